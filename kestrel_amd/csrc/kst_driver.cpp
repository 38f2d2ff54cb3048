// kst_driver.cpp -- the caller of the backend: the driver's block loop
// (src/driver.rs:240-1192) with the note-off gate table and the per-tile
// controller rows it publishes (src/voice.rs:212-658), and the C ABI of the
// bank, the MIDI reader and the driver.
//
// Per block: drain the events that land in it (tempo changes apply at their own
// tick), turn each note-on into admission candidates from the bank's preview
// tables, publish the gate and controller rows, lay the LFOs over them, admit
// what fits (evenly thinned or ranked by loudness within 64 time strata), build
// only the admitted voices, and hand them to the backend. After the render:
// the limiter, the hard clamp, and the end-of-file tail.
//
// The event side of a block (everything up to admission) reads nothing that
// admission, the render or the post-render step of the block before it
// writes, so it runs one block ahead on a thread of its own and hands each
// block over in one swap of buffers.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>

#include "kst_host.h"
#include "kst_stage.h"

namespace kst {
namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

constexpr uint32_t GATE_SLOTS = 16 * 128;  // src/voice.rs:17
constexpr uint32_t CHANNELS = 16, CHAN_FIELDS = 4;  // src/voice.rs:460-471
constexpr uint32_t CHAN_BEND = 0, CHAN_GAIN_L = 1, CHAN_GAIN_R = 2, CHAN_VARIANT = 3;
constexpr uint32_t BEND_ONE = 1u << 24;     // src/fixed.rs:97
constexpr size_t ADMIT_STRATA = 64;         // src/driver.rs:21
constexpr uint16_t RPN_NONE = 0x4000;       // src/driver.rs:28
constexpr uint32_t SOUND_CC_SHIFT = 1;      // src/driver.rs:37
constexpr double MAX_TAIL_SECONDS = 60.0;   // src/driver.rs:139
constexpr uint32_t DEFERRED = UINT32_MAX;   // src/driver.rs:233
constexpr float FRAC_PI_2_F32 = 1.57079632679489661923f;
constexpr float SQRT_2_F32 = 1.41421356237309504880f;
constexpr double TAU = 6.28318530717958647692;

uint32_t f2u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
double fract(double x) { return x - std::trunc(x); }

// bend_factor, src/fixed.rs:99-112: semitones to an 8.24 step multiplier.
uint32_t bend_factor(double semitones) {
    if (!std::isfinite(semitones)) return BEND_ONE;
    const double f = exp2(semitones / 12.0) * (double)BEND_ONE;
    if (f <= 0.0) return 1;
    if (f >= 4294967295.0) return UINT32_MAX;
    return (uint32_t)std::round(f);
}

// The tiebreak of the ranking key: a bijective avalanche of the candidate's
// index on 48 bits (src/voice.rs:168-177).
uint64_t mix48(uint64_t index) {
    const uint64_t M = 0x0000FFFFFFFFFFFFull;
    uint64_t z = index & M;
    z = (z ^ (z >> 24)) & M;
    z = (z * 0x9E3779B97F4Bull) & M;
    z = (z ^ (z >> 23)) & M;
    z = (z * 0xBF58476D1CE5ull) & M;
    z = (z ^ (z >> 25)) & M;
    return z;
}

// The top 15 bits below the sign of |l| + |r|: a logarithmic loudness step
// (src/voice.rs:203-210).
uint64_t rank_gain_q(float gain_l, float gain_r) {
    const float g = std::fabs(gain_l) + std::fabs(gain_r);
    if (g != g || g <= 0.0f) return 0;
    return (f2u(g) >> 16) & 0x7FFF;
}

uint64_t rank_key(float gain_l, float gain_r, size_t index) {  // src/driver.rs:236-238
    return (rank_gain_q(gain_l, gain_r) << 48) | mix48(index);
}

// Config::admit_take, src/config.rs:120-131.
uint32_t admit_take(const kst_config &c, uint32_t live, uint32_t queued) {
    const uint32_t cap = c.max_voices;
    const uint32_t want = std::min(queued, cap);
    if (live + want <= cap) return want;
    uint32_t need = 0;
    if (c.steal_rule != 1) {
        const uint64_t ms = std::max<uint64_t>((uint64_t)c.max_voices * c.max_steal_percent / 100, 1);
        need = std::min(std::min(live + want - cap, live), (uint32_t)ms);
    }
    return std::min(want, cap - (live - need));
}

// ---------------------------------------------------------------------------
// Note-off gate table, src/voice.rs:212-452. A voice releases once the count of
// published note-offs at its (channel, key) slot reaches its own ordinal; rows
// carry that count at the start of each gate tile.
// ---------------------------------------------------------------------------
class GateTable {
public:
    std::vector<uint32_t> rows;
    uint32_t tiles, tile_frames;

    GateTable(uint32_t block_frames, uint32_t gate_frames)
        : rows((size_t)(block_frames / gate_frames) * GATE_SLOTS, 0),
          tiles(block_frames / gate_frames), tile_frames(gate_frames), on_(GATE_SLOTS, 0),
          off_(GATE_SLOTS, 0), pending_(GATE_SLOTS, 0), held_(GATE_SLOTS, 0) {}

    static uint32_t slot(uint8_t ch, uint8_t key) { return (ch & 15u) * 128u + (key & 127u); }
    void begin_block() { cursor_ = 0; }
    void end_block() { publish_through(tiles); }

    uint32_t note_on(uint8_t ch, uint8_t key, uint32_t frame) {
        advance_to(frame);
        const uint32_t s = slot(ch, key);
        return ++on_[s];
    }
    void note_off(uint8_t ch, uint8_t key, uint32_t frame) {
        advance_to(frame);
        const uint32_t s = slot(ch, key);
        // An off that nothing is sounding for is dropped; a pedal holds it as
        // a count, clamped to what is actually sounding.
        const bool unheld = off_[s] + pending_[s] != on_[s];
        if ((sostenuto_ >> (ch & 15) & 1) && held_[s] > 0) {
            if (unheld) {
                ++pending_[s];
                --held_[s];
            }
        } else if (sustained(ch)) {
            if (unheld) ++pending_[s];
        } else if (off_[s] != on_[s]) {
            ++off_[s];
        }
    }
    bool sustained(uint8_t ch) const { return sustain_ >> (ch & 15) & 1; }
    void set_sustain(uint8_t ch, bool down, uint32_t frame) {
        if (down == sustained(ch)) return;
        advance_to(frame);
        const uint16_t bit = (uint16_t)(1u << (ch & 15));
        if (down) {
            sustain_ |= bit;
            return;
        }
        sustain_ &= (uint16_t)~bit;
        if (!(sostenuto_ & bit)) flush(ch);
    }
    void set_sostenuto(uint8_t ch, bool down, uint32_t frame) {
        const uint16_t bit = (uint16_t)(1u << (ch & 15));
        if (down == ((sostenuto_ & bit) != 0)) return;
        advance_to(frame);
        const uint32_t base = (ch & 15u) * 128u;
        if (down) {
            sostenuto_ |= bit;
            for (uint32_t s = base; s < base + 128; ++s) held_[s] = on_[s] - off_[s] - pending_[s];
            return;
        }
        sostenuto_ &= (uint16_t)~bit;
        std::fill(held_.begin() + base, held_.begin() + base + 128, 0u);
        if (!sustained(ch)) flush(ch);
    }
    void all_notes_off(uint8_t ch, uint32_t frame) {
        advance_to(frame);
        const uint32_t base = (ch & 15u) * 128u;
        for (uint32_t s = base; s < base + 128; ++s) {
            if (sustained(ch))
                pending_[s] = on_[s] - off_[s];
            else
                off_[s] = on_[s];
        }
    }
    void all_sound_off(uint8_t ch, uint32_t frame) {
        advance_to(frame);
        const uint32_t base = (ch & 15u) * 128u;
        for (uint32_t s = base; s < base + 128; ++s) {
            off_[s] = on_[s];
            pending_[s] = held_[s] = 0;
        }
    }
    void reset_controllers(uint8_t ch, uint32_t frame) {
        set_sostenuto(ch, false, frame);
        set_sustain(ch, false, frame);
    }

private:
    std::vector<uint32_t> on_, off_, pending_, held_;
    uint16_t sustain_ = 0, sostenuto_ = 0;
    uint32_t cursor_ = 0;

    void publish_through(uint32_t end) {
        for (; cursor_ < end && cursor_ < tiles; ++cursor_)
            std::copy(off_.begin(), off_.end(), rows.begin() + (size_t)cursor_ * GATE_SLOTS);
    }
    // Rows up to and including the tile of `frame` take the state before it.
    void advance_to(uint32_t frame) { publish_through(frame / tile_frames + 1); }
    void flush(uint8_t ch) {
        const uint32_t base = (ch & 15u) * 128u;
        for (uint32_t s = base; s < base + 128; ++s) {
            off_[s] += pending_[s];
            pending_[s] = 0;
        }
    }
};

// ---------------------------------------------------------------------------
// Controller rows, src/voice.rs:454-658: per tile and channel the 8.24 bend
// factor, the two channel gains (f32 bits) and the params variant.
// ---------------------------------------------------------------------------
class ChannelTable {
public:
    std::vector<uint32_t> rows;
    uint32_t tiles, tile_frames;
    bool bend_active = false, gain_active = false, variant_active = false;

    ChannelTable(uint32_t block_frames, uint32_t gate_frames)
        : tiles(block_frames / gate_frames), tile_frames(gate_frames) {
        for (uint32_t c = 0; c < CHANNELS; ++c) {
            now_[c * CHAN_FIELDS + CHAN_BEND] = BEND_ONE;
            now_[c * CHAN_FIELDS + CHAN_GAIN_L] = f2u(1.0f);
            now_[c * CHAN_FIELDS + CHAN_GAIN_R] = f2u(1.0f);
            now_[c * CHAN_FIELDS + CHAN_VARIANT] = 0;
        }
        rows.resize((size_t)tiles * W);
        for (uint32_t t = 0; t < tiles; ++t) std::copy(now_, now_ + W, rows.begin() + (size_t)t * W);
    }
    void begin_block() { cursor_ = 0; }
    void end_block() { publish_through(tiles); }
    void set_bend(uint8_t ch, uint32_t factor, uint32_t frame) {
        set(ch, CHAN_BEND, factor, frame);
    }
    void set_variant(uint8_t ch, uint32_t v, uint32_t frame) { set(ch, CHAN_VARIANT, v, frame); }
    void set_gain(uint8_t ch, float l, float r, uint32_t frame) {
        const uint32_t base = (ch & 15u) * CHAN_FIELDS;
        if (now_[base + CHAN_GAIN_L] == f2u(l) && now_[base + CHAN_GAIN_R] == f2u(r)) return;
        advance_to(frame);
        now_[base + CHAN_GAIN_L] = f2u(l);
        now_[base + CHAN_GAIN_R] = f2u(r);
    }
    // LFOs, laid over what the rows already publish.
    void modulate_bend(uint8_t ch, uint32_t tile, uint32_t factor) {
        uint32_t &e = rows[((size_t)tile * CHANNELS + (ch & 15u)) * CHAN_FIELDS + CHAN_BEND];
        const uint32_t scaled = (uint32_t)(((uint64_t)e * factor) >> 24);
        e = std::max(scaled, 1u);
    }
    void modulate_gain(uint8_t ch, uint32_t tile, float factor) {
        const size_t base = ((size_t)tile * CHANNELS + (ch & 15u)) * CHAN_FIELDS;
        rows[base + CHAN_GAIN_L] = f2u(u2f(rows[base + CHAN_GAIN_L]) * factor);
        rows[base + CHAN_GAIN_R] = f2u(u2f(rows[base + CHAN_GAIN_R]) * factor);
    }
    void refresh_active() {
        bend_active = gain_active = variant_active = false;
        const uint32_t one = f2u(1.0f);
        for (size_t e = 0; e < rows.size(); e += CHAN_FIELDS) {
            bend_active |= rows[e + CHAN_BEND] != BEND_ONE;
            gain_active |= rows[e + CHAN_GAIN_L] != one || rows[e + CHAN_GAIN_R] != one;
            variant_active |= rows[e + CHAN_VARIANT] != 0;
        }
    }

private:
    static constexpr uint32_t W = CHANNELS * CHAN_FIELDS;
    uint32_t now_[W];
    uint32_t cursor_ = 0;
    void set(uint8_t ch, uint32_t field, uint32_t v, uint32_t frame) {
        const uint32_t i = (ch & 15u) * CHAN_FIELDS + field;
        if (now_[i] == v) return;
        advance_to(frame);
        now_[i] = v;
    }
    void publish_through(uint32_t end) {
        for (; cursor_ < end && cursor_ < tiles; ++cursor_)
            std::copy(now_, now_ + W, rows.begin() + (size_t)cursor_ * W);
    }
    void advance_to(uint32_t frame) { publish_through(frame / tile_frames + 1); }
};

}  // namespace

// ---------------------------------------------------------------------------
// The driver.
// ---------------------------------------------------------------------------
class Driver {
public:
    Driver(const kst_config &cfg, const kst_driver_config &dcfg, const Bank &bank,
           std::unique_ptr<EventSource> src)
        : cfg_(cfg), dcfg_(dcfg), bank_(bank), src_(std::move(src)),
          clock_(src_->division, cfg.sample_rate), gates_(cfg.block_frames, cfg.gate_frames),
          chan_(cfg.block_frames, cfg.gate_frames) {
        for (int c = 0; c < 16; ++c) {
            bend_range_[c] = 2.0;
            cc_volume_[c] = (uint8_t)std::min<uint32_t>(dcfg.default_channel_volume, 127);
            cc_expression_[c] = 127;
            cc_pan_[c] = 64;
            for (int w = 0; w < 5; ++w) cc_sound_[c][w] = 64;
            cc_vib_rate_[c] = cc_vib_depth_[c] = 64;
        }
        variants_.push_back(neutral_state());
        seen_states_.push_back(neutral_state());
        variant_seen_.push_back(0);
        tail_blocks_left_ = (uint64_t)(MAX_TAIL_SECONDS * (double)cfg.sample_rate / (double)cfg.block_frames);
        for (int c = 0; c < 16; ++c) refresh_preset(c, 0);
        if (dcfg.limiter && dcfg.limiter_cfg.mode != KST_LIMITER_OFF) {
            kst_limiter_config lc = dcfg.limiter_cfg;
            lc.sample_rate = cfg.sample_rate;
            if (kst_limiter_create(&lc, &limiter_) != KST_OK)
                throw std::invalid_argument(kst_post_last_error());
        }
        // The front starts as full-size rows: each swap hands the event side
        // buffers of the size its tables write through.
        f_.gate_rows = gates_.rows;
        f_.chan_rows = chan_.rows;
        events_ = std::thread([this] { event_thread(); });
    }
    ~Driver() {
        {
            std::unique_lock<std::mutex> g(m_);
            cv_.wait(g, [&] { return !job_; });
            stop_ = true;
        }
        cv_.notify_all();
        events_.join();
        if (limiter_) kst_limiter_destroy(limiter_);
    }

    // Everything before admission (src/driver.rs:450-555). Returns how many
    // params variants wait to be uploaded. The block was usually begun by the
    // event thread while the caller finished the one before; its results are
    // taken over here and that thread starts on the block after.
    // `start_next`: begin the block after on the event thread now (next_block
    // passes false when it will begin that block itself, inline).
    size_t begin_block(bool start_next = true) {
        if (ahead_)
            join_ahead();  // rethrows what the event side threw
        else if (ready_)
            ready_ = false;  // begun inline during the last block's render
        else
            begin_at();
        publish();
        if (start_next) start_ahead();
        return f_.uploads.size();
    }

    // The event side of the next block into the back buffers. Runs on the
    // event thread or on the caller's; it touches only state that admission,
    // the render and the post-render step never do.
    void begin_at() {
        const auto t0 = Clock::now();
        const uint64_t F = cfg_.block_frames;
        const uint64_t block_start = begin_index_ * F, block_end = block_start + F;
        gates_.begin_block();
        chan_.begin_block();
        variant_used_ = 1;
        for (uint32_t v : cur_variant_) variant_used_ |= 1ull << v;
        notes_.clear();
        cands_.clear();
        deferred_now_.clear();
        block_first_id_ = next_note_id_;
        spawn_variants_ = false;
        uploads_.clear();

        // Deferred voices whose start lands in this block, visited in the
        // order a swap-remove walk over the queue visits them.
        for (size_t i = 0; i < deferred_.size();) {
            if (deferred_[i].first < block_end) {
                const uint64_t frame = deferred_[i].first;
                kst_spawn_cmd cmd = deferred_[i].second;
                deferred_[i] = deferred_.back();
                deferred_.pop_back();
                cmd.start_rel = (uint32_t)(frame > block_start ? frame - block_start : 0);
                if (cmd.variant != 0) {
                    variant_used_ |= 1ull << cmd.variant;
                    spawn_variants_ = true;
                }
                cands_.push_back(Cand{rank_key(cmd.gain_l, cmd.gain_r, cands_.size()), DEFERRED,
                                      (uint32_t)deferred_now_.size(), 0});
                deferred_now_.push_back(cmd);
            } else {
                ++i;
            }
        }

        // Events of this block; a tempo change applies from its own tick.
        // The source hands out runs of events; an event past the block stays
        // at the cursor for the block it lands in.
        for (;;) {
            if (sp_n_ == 0 && (sp_n_ = src_->span(&sp_)) == 0) {
                stream_done_ = true;
                break;
            }
            const kst_midi_event &ev = *sp_;
            if (ev.kind == KST_EV_TEMPO) {
                clock_.set_tempo(ev.tick, (uint32_t)ev.value);
                ++counts_.events;
                ++sp_;
                --sp_n_;
                continue;
            }
            const uint64_t frame = sat_u64(clock_.frame_at(ev.tick));
            if (frame >= block_end) break;
            ++sp_;
            --sp_n_;
            const uint32_t rel = std::min<uint32_t>(
                (uint32_t)(frame > block_start ? frame - block_start : 0), cfg_.block_frames - 1);
            ++counts_.events;
            handle_event(ev, rel, block_start);
        }
        // Past the end of the file everything is released, pedals or not.
        if (stream_done_ && !end_sent_) {
            for (uint8_t ch = 0; ch < 16; ++ch) gates_.all_sound_off(ch, 0);
            end_sent_ = true;
        }
        for (const auto &pv : pending_variants_) {
            uploads_.emplace_back(pv.first, std::vector<kst_region_params>());
            bank_.build_variant(pv.second, uploads_.back().second);
        }
        pending_variants_.clear();
        gates_.end_block();
        chan_.end_block();
        apply_modulation();
        chan_.refresh_active();
        ++begin_index_;
        ev_time_ += Clock::now() - t0;
    }

    // Hands the back buffers over to admission (the front) in swaps; the
    // event side overwrites every row of what it gets back before reading it.
    void publish() {
        f_.gate_rows.swap(gates_.rows);
        f_.chan_rows.swap(chan_.rows);
        f_.bend = chan_.bend_active;
        f_.gain = chan_.gain_active;
        f_.variant = chan_.variant_active || spawn_variants_;
        f_.uploads.swap(uploads_);
        f_.cands.swap(cands_);
        f_.notes.swap(notes_);
        f_.deferred_now.swap(deferred_now_);
        f_.first_id = block_first_id_;
        f_.src_more = !stream_done_ || !deferred_.empty();
        stats_.events = counts_.events;
        stats_.notes = counts_.notes;
        stats_.variant_states = counts_.variant_states;
        stats_.param_variants = counts_.param_variants;
        stats_.variant_rebuilds = counts_.variant_rebuilds;
        stats_.variant_fallbacks = counts_.variant_fallbacks;
        t_events_ = ev_time_;
    }

    void start_ahead() {
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = true;
        }
        cv_.notify_all();
        ahead_ = true;
    }

    void join_ahead() {
        const auto t0 = Clock::now();
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return !job_; });
        ahead_ = false;
        t_join_ += Clock::now() - t0;
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }

    void event_thread() {
        std::unique_lock<std::mutex> g(m_);
        for (;;) {
            cv_.wait(g, [&] { return job_ || stop_; });
            if (stop_) return;
            g.unlock();
            std::exception_ptr e;
            try {
                begin_at();
            } catch (...) {
                e = std::current_exception();
            }
            g.lock();
            err_ = e;
            job_ = false;
            cv_.notify_all();
        }
    }

    // Admission and the build of what it admits (src/driver.rs:556-594).
    void admit(uint64_t live64, uint64_t stolen_before) {
        const auto t0 = Clock::now();
        const size_t take = admit_plan(live64, stolen_before);
        if (spawn_buf_.size() < take) spawn_buf_.resize(take);  // grows only: no refill
        spawn_n_ = materialise(spawn_buf_.data(), take);
        t_admit_ += Clock::now() - t0;
    }

    // Admission proper: which candidates, in which order, and their counters.
    // Returns how many of the (reordered) candidates to build.
    size_t admit_plan(uint64_t live64, uint64_t stolen_before) {
        const uint32_t live = (uint32_t)live64;
        const size_t want = f_.cands.size();
        const size_t take = admit_take(cfg_, live, (uint32_t)want);
        stats_.dropped += want - take;
        stolen_before_ = stolen_before;
        stats_.last_live = live;
        stats_.last_want = (uint32_t)want;
        stats_.last_take = (uint32_t)take;
        if (take < want && take > 0) {
            if (cfg_.admit_rule == 1) {  // Even: thin by position
                for (size_t i = 0; i < take; ++i) f_.cands[i] = f_.cands[(size_t)((uint64_t)i * want / take)];
            } else {
                rank_candidates(want, take);
            }
        }
        uint64_t want_e = 0, take_e = 0;
        for (size_t i = 0; i < f_.cands.size(); ++i) {
            const uint64_t e = (f_.cands[i].key >> 48) & 0x7FFF;
            want_e += e;
            if (i < take) take_e += e;
        }
        stats_.last_want_energy = want_e;
        stats_.last_take_energy = take_e;
        stats_.voices_spawned += want;
        return std::min(take, f_.cands.size());
    }

    // After the render (src/driver.rs:596-641). Returns whether there is more.
    int end_block(float *out, size_t n, uint64_t stolen_after_spawn, const kst_block_stats &st) {
        return end_at(out, n, stolen_after_spawn, st, f_.src_more);
    }

    // `src_more`: the file or the deferred queue still had something after
    // this block's own events were drained.
    int end_at(float *out, size_t n, uint64_t stolen_after_spawn, const kst_block_stats &st,
               bool src_more) {
        const auto t0 = Clock::now();
        struct Tally {
            Clock::time_point t0;
            Clock::duration &acc;
            ~Tally() { acc += Clock::now() - t0; }
        } tally{t0, t_post_};
        stats_.last_stolen = stolen_after_spawn - stolen_before_;
        if (limiter_ && kst_limiter_process(limiter_, out, n) != KST_OK)
            return fail(KST_EINVAL, kst_post_last_error());
        stats_.clipped += kst_clamp_block(out, n);
        if (dcfg_.nan_guard)
            for (size_t i = 0; i < n; ++i)
                if (!std::isfinite(out[i]))
                    return fail(KST_EINVAL, "block " + std::to_string(block_index_) + " sample " +
                                                std::to_string(i) + " is " + std::to_string(out[i]) +
                                                "; the synth produced a non-finite value");
        stats_.peak = std::fmax(stats_.peak, st.peak);
        stats_.frames += cfg_.block_frames;
        stats_.blocks += 1;
        ++block_index_;
        if (src_more) return 1;
        if (st.active_voices > 0) {
            tail_blocks_left_ = tail_blocks_left_ > 0 ? tail_blocks_left_ - 1 : 0;
            return tail_blocks_left_ == 0 ? 0 : 1;
        }
        return 0;
    }

    // One block against this library's backend. The event thread drains the
    // events of block b+1 and builds its rows while this thread admits block
    // b, uploads it, waits for the device and runs the post-render step: none
    // of that reads what the other side writes, and the counters of b+1 are
    // held on the event side until begin_block takes the block over, so what
    // each call returns and reports is what the sequential loop would show.
    int next_block(kst_backend *b, float *out, size_t n) {
        if (n != (size_t)cfg_.block_frames * 2)
            return fail(KST_EINVAL, "output block is " + std::to_string(n) + " samples, expected " +
                                        std::to_string(cfg_.block_frames * 2));
        const auto t_block = Clock::now();
        const bool inl = pace_inline_;
        try {
            begin_block(!inl);
        } catch (const std::exception &e) {
            return fail(KST_EINVAL, e.what());
        }
        auto be = [](int rc) {
            if (rc != KST_OK) throw std::runtime_error(kst_last_error());
        };
        try {
            auto tb = Clock::now();
            for (const auto &u : f_.uploads)
                be(kst_set_params_variant(b, u.first, u.second.data(), u.second.size()));
            be(kst_set_gates(b, f_.gate_rows.data(), f_.gate_rows.size()));
            be(kst_set_channels(b, f_.chan_rows.data(), f_.chan_rows.size(), f_.bend, f_.gain,
                                f_.variant));
            kst_block_stats st{};
            be(kst_stats(b, &st));
            t_backend_ += Clock::now() - tb;
            // Admission builds the admitted layers straight into the backend's
            // pinned command staging (what kst_spawn would copy them into).
            const auto ta = Clock::now();
            const size_t take = admit_plan(st.active_voices, st.stolen);
            kst_spawn_cmd *cmds = nullptr;
            be(kst::spawn_reserve(b, take, &cmds));
            const size_t built = materialise(cmds, take);
            t_admit_ += Clock::now() - ta;
            tb = Clock::now();
            be(kst::spawn_commit(b, built));
            be(kst_stats(b, &st));
            const uint64_t stolen_after = st.stolen;
            be(kst_render_submit(b));
            t_backend_ += Clock::now() - tb;
            if (inl) {  // the next block's event side while the device renders
                try {
                    begin_at();
                } catch (const std::exception &e) {
                    (void)kst_render_collect(b, out, n);  // leave the backend usable
                    return fail(KST_EINVAL, e.what());
                }
                ready_ = true;
            }
            tb = Clock::now();
            be(kst_render_collect(b, out, n));
            be(kst_stats(b, &st));
            t_wait_ += Clock::now() - tb;
            ++timed_blocks_;
            const int more = end_at(out, n, stolen_after, st, f_.src_more);
            pace(inl, Clock::now() - t_block);
            return more;
        } catch (const std::runtime_error &e) {
            return fail(KST_EDEVICE, e.what());
        }
    }

    // Where next_block runs the event side: on the event thread (beside
    // admission, the device and the post-render step) or inline on the
    // caller's thread while the device renders. Which is faster depends on the
    // machine -- how the two threads' shared buffers travel between the cores
    // the scheduler picked -- so next_block times both: a 16-block trial of
    // each, then the faster for 480 blocks, then a trial of the other against
    // the chosen one's latest blocks, and so on. The first blocks of a stint
    // are left out of its mean (the switch itself costs one block's overlap).
    void pace(bool inl, std::chrono::steady_clock::duration took) {
        ++inline_blocks_[inl];
        const uint32_t k = PACE_TRIAL - stint_left_;  // position in the last PACE_TRIAL
        if (stint_left_ <= PACE_TRIAL && k >= PACE_SKIP) {
            stint_sum_ += std::chrono::duration<double>(took).count();
            ++stint_n_;
        }
        if (--stint_left_ > 0) return;
        pace_mean_[inl] = stint_n_ ? stint_sum_ / stint_n_ : 0.0;
        stint_sum_ = 0.0;
        stint_n_ = 0;
        if (pace_trial_ && pace_mean_[!inl] == 0.0) {  // the other has no trial yet
            pace_inline_ = !inl;
            stint_left_ = PACE_TRIAL;
        } else if (pace_trial_) {  // a trial ended: keep the faster
            pace_inline_ = pace_mean_[1] < pace_mean_[0];
            pace_trial_ = false;
            stint_left_ = PACE_STINT;
        } else {  // a stint ended: try the other
            pace_inline_ = !inl;
            pace_trial_ = true;
            stint_left_ = PACE_TRIAL;
        }
    }

    size_t timings(kst_timing *out, size_t cap) const {
        if (timed_blocks_ == 0) return 0;
        const double k = 1e-6 / (double)timed_blocks_;
        const kst_timing t[7] = {
            {"events_rows", std::chrono::duration<double, std::nano>(t_events_).count() * k},
            {"events_wait", std::chrono::duration<double, std::nano>(t_join_).count() * k},
            {"admission_build", std::chrono::duration<double, std::nano>(t_admit_).count() * k},
            {"post_render", std::chrono::duration<double, std::nano>(t_post_).count() * k},
            {"backend_calls", std::chrono::duration<double, std::nano>(t_backend_).count() * k},
            {"device_wait", std::chrono::duration<double, std::nano>(t_wait_).count() * k},
            {"inline_share", (double)inline_blocks_[1] / (double)timed_blocks_}};
        for (size_t i = 0; i < 7 && i < cap; ++i) out[i] = t[i];
        return 7;
    }

    // The block begin_block took over: valid until the next begin_block.
    const std::vector<std::pair<uint32_t, std::vector<kst_region_params>>> &uploads() const {
        return f_.uploads;
    }
    const std::vector<uint32_t> &gate_rows() const { return f_.gate_rows; }
    const std::vector<uint32_t> &chan_rows() const { return f_.chan_rows; }
    void chan_flags(uint32_t flags[3]) const {
        flags[0] = f_.bend;
        flags[1] = f_.gain;
        flags[2] = f_.variant;
    }
    const kst_spawn_cmd *spawned() const { return spawn_buf_.data(); }
    size_t spawned_count() const { return spawn_n_; }
    const kst_driver_stats &stats() const { return stats_; }
    uint32_t block_frames() const { return cfg_.block_frames; }

private:
    // One layer a block might admit (src/driver.rs:210-228).
    struct Cand {
        uint64_t key;
        uint32_t note;    // index into notes_, or DEFERRED
        uint32_t region;  // or the index into deferred_now_
        uint32_t id_off;  // note id minus block_first_id_
    };
    struct Note {
        uint8_t ch, key, vel, variant;
        uint32_t rel, ordinal;
    };
    typedef std::array<uint8_t, 5> State;

    // Read-only after construction.
    kst_config cfg_;
    kst_driver_config dcfg_;
    const Bank &bank_;

    // Event side: the event thread's own state (or the caller's before the
    // first block), kept off the cache lines the caller's side writes.
    alignas(128) std::unique_ptr<EventSource> src_;
    TempoClock clock_;
    GateTable gates_;
    ChannelTable chan_;
    uint8_t bank_msb_[16] = {}, bank_lsb_[16] = {};
    uint32_t preset_[16] = {};
    int16_t bend_val_[16] = {};
    double bend_range_[16];
    uint16_t rpn_sel_[16] = {};
    uint8_t cc_volume_[16], cc_expression_[16], cc_pan_[16], cc_sound_[16][5];
    uint8_t cc_mod_[16] = {}, cc_vib_rate_[16], cc_vib_depth_[16], cc_tremolo_[16] = {},
            cc_soft_[16] = {};
    uint8_t lsb_mod_[16] = {}, lsb_volume_[16] = {}, lsb_pan_[16] = {}, lsb_expression_[16] = {};
    double lfo_phase_[16] = {};
    std::vector<State> variants_, seen_states_;
    uint32_t cur_variant_[16] = {};
    uint64_t variant_used_ = 1;
    std::vector<uint64_t> variant_seen_;
    uint64_t variant_clock_ = 0;
    std::vector<std::pair<uint32_t, ParamMod>> pending_variants_;
    uint64_t next_note_id_ = 1;
    const kst_midi_event *sp_ = nullptr;  // cursor into the source's current run
    size_t sp_n_ = 0;
    bool stream_done_ = false, end_sent_ = false;
    std::vector<Preview> preview_;
    std::vector<std::pair<uint64_t, kst_spawn_cmd>> deferred_;
    uint64_t begin_index_ = 0;  // the block begin_at drains next
    // The back buffers begin_at fills, handed over by publish().
    std::vector<std::pair<uint32_t, std::vector<kst_region_params>>> uploads_;
    bool spawn_variants_ = false;
    std::vector<kst_spawn_cmd> deferred_now_;
    std::vector<Note> notes_;
    std::vector<Cand> cands_;
    uint64_t block_first_id_ = 0;
    // Event-side counters (events, notes, variant bookkeeping), copied into
    // stats_ as each block is taken over.
    kst_driver_stats counts_{};
    typedef std::chrono::steady_clock Clock;
    Clock::duration ev_time_{};  // event side's own total, copied on take-over

    // The caller's side: admission, the backend calls and the post-render step.
    // One block's event-side results as admission and the backend read them.
    struct Front {
        std::vector<uint32_t> gate_rows, chan_rows;
        bool bend = false, gain = false, variant = false;
        std::vector<std::pair<uint32_t, std::vector<kst_region_params>>> uploads;
        std::vector<Cand> cands;
        std::vector<Note> notes;
        std::vector<kst_spawn_cmd> deferred_now;
        uint64_t first_id = 0;
        bool src_more = true;  // the file or the deferred queue had more after it
    };
    alignas(128) Front f_;
    std::vector<kst_spawn_cmd> spawn_buf_;  // the split entry points' spawn list
    size_t spawn_n_ = 0;
    kst_driver_stats stats_{};
    kst_limiter *limiter_ = nullptr;
    uint64_t block_index_ = 0, tail_blocks_left_ = 0, stolen_before_ = 0;
    Clock::duration t_events_{}, t_admit_{}, t_post_{}, t_backend_{}, t_wait_{}, t_join_{};
    uint64_t timed_blocks_ = 0;
    // next_block's choice of where the event side runs (see pace()).
    static constexpr uint32_t PACE_TRIAL = 16, PACE_SKIP = 2, PACE_STINT = 480;
    bool pace_inline_ = false, pace_trial_ = true, ready_ = false;
    uint32_t stint_left_ = PACE_TRIAL, stint_n_ = 0;
    double stint_sum_ = 0.0, pace_mean_[2] = {0.0, 0.0};
    uint64_t inline_blocks_[2] = {0, 0};

    // The event thread and its one-job handshake: job_ is set while it
    // begins a block, err_ holds what that threw.
    alignas(128) std::mutex m_;
    std::condition_variable cv_;
    bool job_ = false, stop_ = false, ahead_ = false;
    std::exception_ptr err_;
    std::thread events_;

    static State neutral_state() {
        State s;
        s.fill((uint8_t)(64 >> SOUND_CC_SHIFT));
        return s;
    }

    void refresh_preset(int c, uint8_t program) {  // src/driver.rs:419-432
        const uint32_t bank_num = c == 9 ? 128u : bank_msb_[c];  // GM drums on channel 10
        const int64_t p = bank_.find_preset(bank_num, program);
        preset_[c] = p < 0 ? 0u : (uint32_t)p;
    }

    static kst_spawn_cmd make_cmd(const Layer &v, uint32_t variant, uint32_t gate_slot,
                                  uint32_t ordinal, uint32_t start_rel, uint64_t id) {
        kst_spawn_cmd c{};  // src/driver.rs:646-678
        c.phase_lo = (uint32_t)v.phase;
        c.phase_hi = (uint32_t)(v.phase >> 32);
        c.step_lo = (uint32_t)v.step;
        c.step_hi = (uint32_t)(v.step >> 32);
        c.smp_base = v.smp_base;
        c.smp_len = v.smp_len;
        c.loop_start = v.loop_start;
        c.loop_end = v.loop_end;
        c.flags = v.flags;
        c.params = v.params;
        c.variant = variant;
        c.region = v.region;
        c.gate_slot = gate_slot;
        c.ordinal = ordinal;
        c.start_rel = start_rel;
        c.note_id_lo = (uint32_t)id;
        c.note_id_hi = (uint32_t)(id >> 32);
        c.gain_l = v.gain_l;
        c.gain_r = v.gain_r;
        return c;
    }

    // Loudest admission (src/driver.rs:680-749): stratum s of the want
    // candidates keeps its proportional share of take, the highest keys, and
    // the winners slide down to the front in stratum order. Within a stratum
    // the winners keep their arrival order.
    void rank_candidates(size_t want, size_t take) {
        const size_t strata = std::min(take, ADMIT_STRATA);
        std::vector<uint64_t> keys;
        std::vector<Cand> tmp;
        for (size_t s = 0; s < strata; ++s) {
            const size_t lo = (size_t)((uint64_t)s * want / strata);
            const size_t hi = (size_t)((uint64_t)(s + 1) * want / strata);
            const size_t olo = (size_t)((uint64_t)s * take / strata);
            const size_t ohi = (size_t)((uint64_t)(s + 1) * take / strata);
            const size_t q = ohi - olo, n = hi - lo;
            if (q == 0) continue;
            if (q < n) {
                keys.resize(n);
                for (size_t i = 0; i < n; ++i) keys[i] = f_.cands[lo + i].key;
                std::nth_element(keys.begin(), keys.begin() + (q - 1), keys.end(),
                                 std::greater<uint64_t>());
                const uint64_t kth = keys[q - 1];  // keys are distinct: exactly q at or above
                tmp.clear();
                for (size_t i = lo; i < hi; ++i)
                    if (f_.cands[i].key >= kth) tmp.push_back(f_.cands[i]);
                for (size_t i = lo; i < hi; ++i)
                    if (f_.cands[i].key < kth) tmp.push_back(f_.cands[i]);
                std::copy(tmp.begin(), tmp.end(), f_.cands.begin() + lo);
            }
            for (size_t j = 0; j < q; ++j) std::swap(f_.cands[olo + j], f_.cands[lo + j]);
        }
    }

    // Builds the first `take` candidates into out[0..], skipping a layer the
    // bank cannot build (src/driver.rs:751-783); returns how many it wrote.
    size_t materialise(kst_spawn_cmd *out, size_t take) {
        size_t n = 0;
        for (size_t i = 0; i < take; ++i) {
            const Cand &c = f_.cands[i];
            if (c.note == DEFERRED) {
                out[n++] = f_.deferred_now[c.region];
                continue;
            }
            const Note &nt = f_.notes[c.note];
            Layer v;
            if (!bank_.build_layer(c.region, nt.key, nt.vel, v)) continue;
            const uint32_t start_rel = v.delay_frames == 0 ? nt.rel : nt.rel + v.delay_frames;
            out[n++] = make_cmd(v, nt.variant, GateTable::slot(nt.ch, nt.key), nt.ordinal, start_rel,
                                f_.first_id + c.id_off);
        }
        return n;
    }

    void handle_event(const kst_midi_event &ev, uint32_t rel, uint64_t block_start) {
        const uint8_t ch = ev.ch;
        const int c = ch & 15;
        switch (ev.kind) {
        case KST_EV_NOTE_ON: {  // src/driver.rs:787-872
            const uint8_t key = ev.a, vel = ev.b;
            const uint32_t ordinal = gates_.note_on(ch, key, rel);
            ++counts_.notes;
            const uint32_t variant = cur_variant_[c];
            preview_.clear();
            bank_.preview_note_on(preset_[c], key, vel, dcfg_.max_layers, preview_);
            const uint32_t note = (uint32_t)notes_.size();
            bool recorded = false;
            for (const Preview &p : preview_) {
                const uint64_t id = next_note_id_++;  // one id per layer, delayed ones too
                if (p.delay_frames != 0) {
                    const uint64_t start = block_start + rel + p.delay_frames;
                    if (start >= block_start + cfg_.block_frames) {
                        Layer v;
                        if (bank_.build_layer(p.region, key, vel, v))
                            deferred_.emplace_back(
                                start, make_cmd(v, variant, GateTable::slot(ch, key), ordinal, rel, id));
                        continue;
                    }
                }
                if (!recorded) {
                    notes_.push_back(Note{(uint8_t)(ch & 0xF), key, vel, (uint8_t)variant, rel, ordinal});
                    recorded = true;
                }
                cands_.push_back(Cand{rank_key(p.gain_l, p.gain_r, cands_.size()), note, p.region,
                                      (uint32_t)(id - block_first_id_)});
                if (variant != 0) {
                    variant_used_ |= 1ull << variant;
                    spawn_variants_ = true;
                }
            }
            break;
        }
        case KST_EV_NOTE_OFF:
            gates_.note_off(ch, ev.a, rel);
            break;
        case KST_EV_CC:
            control_change(ch, ev.a, ev.b, rel);
            break;
        case KST_EV_PROGRAM:
            refresh_preset(c, ev.a);
            break;
        case KST_EV_PITCH_BEND:
            bend_val_[c] = (int16_t)ev.value;
            refresh_bend(c, rel);
            break;
        default:
            break;
        }
    }

    void control_change(uint8_t ch, uint8_t num, uint8_t val, uint32_t rel) {  // :874-978
        const int c = ch & 15;
        auto select_kept = [&]() -> uint16_t { return rpn_sel_[c] == RPN_NONE ? 0 : rpn_sel_[c]; };
        switch (num) {
        case 0: bank_msb_[c] = val; break;
        case 1: cc_mod_[c] = val; lsb_mod_[c] = 0; break;
        case 33: lsb_mod_[c] = val; break;
        case 39: lsb_volume_[c] = val; refresh_gain(c, rel); break;
        case 42: lsb_pan_[c] = val; refresh_gain(c, rel); break;
        case 43: lsb_expression_[c] = val; refresh_gain(c, rel); break;
        case 32: bank_lsb_[c] = val; break;
        // Data entry and increment/decrement act on RPN 0 (bend range) only.
        case 6:
            if (rpn_sel_[c] == 0) { bend_range_[c] = (double)val; refresh_bend(c, rel); }
            break;
        case 38:
            if (rpn_sel_[c] == 0) {
                bend_range_[c] = std::trunc(bend_range_[c]) + (double)val / 100.0;
                refresh_bend(c, rel);
            }
            break;
        case 96:
            if (rpn_sel_[c] == 0) { bend_range_[c] = std::fmin(bend_range_[c] + 1.0, 127.0); refresh_bend(c, rel); }
            break;
        case 97:
            if (rpn_sel_[c] == 0) { bend_range_[c] = std::fmax(bend_range_[c] - 1.0, 0.0); refresh_bend(c, rel); }
            break;
        case 98: case 99: rpn_sel_[c] = RPN_NONE; break;
        case 100: rpn_sel_[c] = (uint16_t)((select_kept() & 0x3F80) | val); break;
        case 101: rpn_sel_[c] = (uint16_t)((select_kept() & 0x7F) | ((uint16_t)val << 7)); break;
        case 7: cc_volume_[c] = val; lsb_volume_[c] = 0; refresh_gain(c, rel); break;
        case 10: cc_pan_[c] = val; lsb_pan_[c] = 0; refresh_gain(c, rel); break;
        case 11: cc_expression_[c] = val; lsb_expression_[c] = 0; refresh_gain(c, rel); break;
        case 64: gates_.set_sustain(ch, val >= 64, rel); break;
        case 66: gates_.set_sostenuto(ch, val >= 64, rel); break;
        case 67: cc_soft_[c] = val; refresh_gain(c, rel); break;
        case 71: case 72: case 73: case 74: case 75: set_sound_cc(ch, num - 71, val, rel); break;
        case 76: cc_vib_rate_[c] = val; break;
        case 77: cc_vib_depth_[c] = val; break;
        case 92: cc_tremolo_[c] = val; break;
        case 120: gates_.all_sound_off(ch, rel); break;
        case 121: reset_controllers(ch, rel); break;
        case 123: case 124: case 125: case 126: case 127: gates_.all_notes_off(ch, rel); break;
        default: break;
        }
    }

    // CC71-CC75 select (and if new, queue the build of) a params variant for
    // the channel's quantised controller state (src/driver.rs:980-1056).
    void set_sound_cc(uint8_t ch, int which, uint8_t val, uint32_t rel) {
        const int c = ch & 15;
        if (cc_sound_[c][which] == val) return;
        cc_sound_[c][which] = val;
        State want;
        for (int i = 0; i < 5; ++i) want[i] = (uint8_t)(cc_sound_[c][i] >> SOUND_CC_SHIFT);
        if (std::find(seen_states_.begin(), seen_states_.end(), want) == seen_states_.end()) {
            seen_states_.push_back(want);
            counts_.variant_states = (uint32_t)seen_states_.size();
        }
        const uint32_t cap = std::min<uint32_t>(std::max<uint32_t>(cfg_.max_param_variants, 1), 63);
        uint32_t idx;
        auto it = std::find(variants_.begin(), variants_.end(), want);
        if (it != variants_.end()) {
            idx = (uint32_t)(it - variants_.begin());
        } else if (variants_.size() < cap) {
            idx = (uint32_t)variants_.size();
            variants_.push_back(want);
            variant_seen_.push_back(0);
            counts_.param_variants = idx + 1;
            build_variant_at(idx, c);
        } else {
            // Reuse the stalest slot no row of this block names yet.
            int64_t victim = -1;
            for (size_t i = 1; i < variants_.size(); ++i)
                if (!(variant_used_ >> i & 1) &&
                    (victim < 0 || variant_seen_[i] < variant_seen_[(size_t)victim]))
                    victim = (int64_t)i;
            if (victim >= 0) {
                idx = (uint32_t)victim;
                variants_[idx] = want;
                build_variant_at(idx, c);
                ++counts_.variant_rebuilds;
            } else {  // every slot is spoken for: the nearest state
                ++counts_.variant_fallbacks;
                uint32_t best = 0, best_d = UINT32_MAX;
                for (size_t i = 0; i < variants_.size(); ++i) {
                    uint32_t d = 0;
                    for (int k = 0; k < 5; ++k) d += (uint32_t)std::abs((int)variants_[i][k] - (int)want[k]);
                    if (d < best_d) {
                        best_d = d;
                        best = (uint32_t)i;
                    }
                }
                idx = best;
            }
        }
        variant_seen_[idx] = ++variant_clock_;
        variant_used_ |= 1ull << idx;
        cur_variant_[c] = idx;
        chan_.set_variant(ch, idx, rel);
    }

    void build_variant_at(uint32_t i, int c) {  // src/driver.rs:1058-1071
        const ParamMod m = ParamMod::from_controllers(cc_sound_[c][0], cc_sound_[c][1], cc_sound_[c][2],
                                                      cc_sound_[c][3], cc_sound_[c][4]);
        pending_variants_.erase(std::remove_if(pending_variants_.begin(), pending_variants_.end(),
                                               [&](const auto &p) { return p.first == i; }),
                                pending_variants_.end());
        pending_variants_.emplace_back(i, m);
    }

    void reset_controllers(uint8_t ch, uint32_t rel) {  // src/driver.rs:1073-1108
        const int c = ch & 15;
        bank_msb_[c] = bank_lsb_[c] = 0;
        bend_range_[c] = 2.0;
        bend_val_[c] = 0;
        rpn_sel_[c] = RPN_NONE;
        cc_volume_[c] = (uint8_t)std::min<uint32_t>(dcfg_.default_channel_volume, 127);
        cc_expression_[c] = 127;
        cc_pan_[c] = 64;
        cc_soft_[c] = cc_mod_[c] = lsb_mod_[c] = lsb_volume_[c] = lsb_pan_[c] = lsb_expression_[c] = 0;
        cc_tremolo_[c] = 0;
        cc_vib_rate_[c] = cc_vib_depth_[c] = 64;
        for (int w = 0; w < 5; ++w) set_sound_cc(ch, w, 64, rel);
        refresh_bend(c, rel);
        refresh_gain(c, rel);
        gates_.reset_controllers(ch, rel);
    }

    // Vibrato and tremolo, one sample per channel per gate tile
    // (src/driver.rs:1110-1144).
    void apply_modulation() {
        const uint32_t tiles = chan_.tiles;
        const double tile_seconds = (double)cfg_.gate_frames / (double)cfg_.sample_rate;
        for (int c = 0; c < 16; ++c) {
            const double vib = ((double)cc_mod_[c] * 128.0 + (double)lsb_mod_[c]) / 16383.0 *
                               ((double)cc_vib_depth_[c] / 64.0) * 50.0;
            const double trem = (double)cc_tremolo_[c] / 127.0 * 0.25;
            const double rate = 5.0 * exp2(((double)cc_vib_rate_[c] - 64.0) / 32.0);
            if (vib <= 0.0 && trem <= 0.0) {
                lfo_phase_[c] = fract(lfo_phase_[c] + rate * tile_seconds * (double)tiles);
                continue;
            }
            double phase = lfo_phase_[c];
            for (uint32_t t = 0; t < tiles; ++t) {
                const double s = std::sin(phase * TAU);
                if (vib > 0.0) chan_.modulate_bend((uint8_t)c, t, bend_factor(vib * s / 100.0));
                if (trem > 0.0) chan_.modulate_gain((uint8_t)c, t, (float)(1.0 - trem + trem * s));
                phase = fract(phase + rate * tile_seconds);
            }
            lfo_phase_[c] = phase;
        }
    }

    void refresh_bend(int c, uint32_t rel) {  // src/driver.rs:1146-1159
        const double semitones = ((double)bend_val_[c] / 8192.0) * bend_range_[c];
        chan_.set_bend((uint8_t)c, bend_factor(semitones), rel);
    }

    // Volume and expression squared (GM), soft pedal up to -6 dB, constant
    // power pan with an exact unity centre (src/driver.rs:1161-1188).
    void refresh_gain(int c, uint32_t rel) {
        auto fine = [](uint8_t msb, uint8_t lsb) {
            return lsb == 0 ? (float)msb / 127.0f : ((float)msb * 128.0f + (float)lsb) / 16383.0f;
        };
        const float v = fine(cc_volume_[c], lsb_volume_[c]);
        const float e = fine(cc_expression_[c], lsb_expression_[c]);
        const float soft = 1.0f - 0.5f * ((float)cc_soft_[c] / 127.0f);
        const float amp = (v * v) * (e * e) * soft;
        const float theta = fine(cc_pan_[c], lsb_pan_[c]) * FRAC_PI_2_F32;
        float l = 1.0f, r = 1.0f;
        if (!(cc_pan_[c] == 64 && lsb_pan_[c] == 0)) {
            l = cosf(theta) * SQRT_2_F32;
            r = sinf(theta) * SQRT_2_F32;
        }
        chan_.set_gain((uint8_t)c, amp * l, amp * r, rel);
    }
};

}  // namespace kst

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct kst_bank {
    kst::Bank bank;
};
struct kst_midi {
    kst::MidiFile file;
};
struct kst_driver {
    std::unique_ptr<kst::Driver> d;
};

namespace {

using kst::fail;

int check_driver_config(const kst_config *cfg, const kst_driver_config *dcfg) {
    if (!cfg || !dcfg) return fail(KST_EINVAL, "null config");
    if (cfg->sample_rate < 8000 || cfg->sample_rate > 768000)
        return fail(KST_EINVAL, "sample_rate " + std::to_string(cfg->sample_rate) + " out of range");
    if (cfg->gate_frames == 0 || cfg->block_frames == 0 || cfg->block_frames % cfg->gate_frames != 0)
        return fail(KST_EINVAL, "gate_frames (" + std::to_string(cfg->gate_frames) +
                                    ") must divide block_frames (" +
                                    std::to_string(cfg->block_frames) + ")");
    if (cfg->block_frames > 65536)
        return fail(KST_EINVAL, "block_frames " + std::to_string(cfg->block_frames) +
                                    " exceeds the 16-bit frame offset a queued note carries");
    if (dcfg->max_layers == 0) return fail(KST_EINVAL, "max_layers must be non-zero");
    if (cfg->max_voices == 0) return fail(KST_EINVAL, "max_voices must be non-zero");
    return KST_OK;
}

kst::BankSettings bank_settings(const kst_config &c, const kst_driver_config &d) {
    kst::BankSettings s;
    s.sample_rate = c.sample_rate;
    s.decay_curve = c.decay_curve;
    s.release_curve = c.release_curve;
    s.filter_enabled = d.filter_enabled;
    s.env_floor = c.env_floor;
    s.master_volume = d.master_volume;
    return s;
}

int make_driver(const kst_config *cfg, const kst_driver_config *dcfg, const kst_bank *bank,
                std::unique_ptr<kst::EventSource> src, kst_driver **out) {
    try {
        auto *d = new kst_driver;
        d->d = std::make_unique<kst::Driver>(*cfg, *dcfg, bank->bank, std::move(src));
        *out = d;
        return KST_OK;
    } catch (const std::exception &e) {
        return fail(KST_EINVAL, e.what());
    }
}

}  // namespace

extern "C" {

void kst_default_driver_config(kst_driver_config *c) {
    if (!c) return;
    *c = kst_driver_config{};
    c->max_layers = 16;
    c->master_volume = 1.0f;
    c->default_channel_volume = 127;
    c->filter_enabled = 1;
    c->limiter = 1;
    c->nan_guard = 0;
    kst_default_limiter_config(&c->limiter_cfg);
}

int kst_bank_create(const kst_config *cfg, const kst_driver_config *dcfg, const int16_t *pool,
                    size_t pool_len, uint32_t pool_rate, const kst_sample_info *samples,
                    size_t n_samples, const kst_region *regions, size_t n_regions,
                    const kst_preset *presets, size_t n_presets, const uint32_t *preset_regions,
                    size_t n_preset_regions, kst_bank **out) {
    if (!cfg || !dcfg || !out) return fail(KST_EINVAL, "null argument");
    if ((!pool && pool_len) || (!samples && n_samples) || (!regions && n_regions) ||
        (!presets && n_presets) || (!preset_regions && n_preset_regions))
        return fail(KST_EINVAL, "null array with a non-zero length");
    if (cfg->sample_rate == 0) return fail(KST_EINVAL, "sample_rate must be non-zero");
    try {
        *out = new kst_bank{kst::Bank(bank_settings(*cfg, *dcfg), pool, pool_len, pool_rate, samples,
                                      n_samples, regions, n_regions, presets, n_presets,
                                      preset_regions, n_preset_regions)};
        return KST_OK;
    } catch (const std::exception &e) {
        return fail(KST_EINVAL, e.what());
    }
}

int kst_bank_info(const kst_bank *b, size_t *pool_len, size_t *n_params, uint32_t *n_regions,
                  uint32_t *max_sample_len) {
    if (!b) return fail(KST_EINVAL, "null bank");
    if (pool_len) *pool_len = b->bank.pool.size();
    if (n_params) *n_params = b->bank.params.size();
    if (n_regions) *n_regions = (uint32_t)b->bank.regions.size();
    if (max_sample_len) *max_sample_len = b->bank.max_sample_len;
    return KST_OK;
}

const int16_t *kst_bank_pool(const kst_bank *b) { return b ? b->bank.pool.data() : nullptr; }
const kst_region_params *kst_bank_params(const kst_bank *b) {
    return b ? b->bank.params.data() : nullptr;
}

int kst_bank_variant(const kst_bank *b, const uint8_t cc[5], kst_region_params *out, size_t n) {
    if (!b || !cc || (!out && n)) return fail(KST_EINVAL, "null argument");
    if (n != b->bank.params.size())
        return fail(KST_EINVAL, "variant holds " + std::to_string(b->bank.params.size()) +
                                    " entries, got room for " + std::to_string(n));
    std::vector<kst_region_params> v;
    b->bank.build_variant(kst::ParamMod::from_controllers(cc[0], cc[1], cc[2], cc[3], cc[4]), v);
    std::copy(v.begin(), v.end(), out);
    return KST_OK;
}

int64_t kst_bank_find_preset(const kst_bank *b, uint32_t bank, uint32_t program) {
    return b ? b->bank.find_preset(bank, program) : -1;
}

void kst_bank_destroy(kst_bank *b) { delete b; }

int kst_bank_contents(const kst_bank *b, const kst_sample_info **samples, size_t *n_samples,
                      const kst_region **regions, size_t *n_regions, const kst_preset **presets,
                      size_t *n_presets, const uint32_t **preset_regions,
                      size_t *n_preset_regions, uint32_t *pool_rate) {
    if (!b) return fail(KST_EINVAL, "null bank");
    const kst::Bank &k = b->bank;
    if (samples) *samples = k.samples.data();
    if (n_samples) *n_samples = k.samples.size();
    if (regions) *regions = k.regions.data();
    if (n_regions) *n_regions = k.regions.size();
    if (presets) *presets = k.preset_desc.data();
    if (n_presets) *n_presets = k.preset_desc.size();
    if (preset_regions) *preset_regions = k.preset_regions.data();
    if (n_preset_regions) *n_preset_regions = k.preset_regions.size();
    if (pool_rate) *pool_rate = k.pool_rate;
    return KST_OK;
}

void kst_default_load_config(kst_load_config *c) {
    if (!c) return;
    c->resample_pool = 1;
    c->pad = 0;
    c->sample_pool_budget = 2ull << 30;
}

int kst_bank_load_sf2(const kst_config *cfg, const kst_driver_config *dcfg,
                      const kst_load_config *lcfg, const char *path, kst_bank **out) {
    if (!cfg || !dcfg || !lcfg || !path || !out) return fail(KST_EINVAL, "null argument");
    if (cfg->sample_rate == 0) return fail(KST_EINVAL, "sample_rate must be non-zero");
    try {
        std::unique_ptr<kst::Bank> bank(kst::load_sf2(bank_settings(*cfg, *dcfg), path,
                                                      lcfg->resample_pool != 0,
                                                      lcfg->sample_pool_budget));
        *out = new kst_bank{std::move(*bank)};
        return KST_OK;
    } catch (const std::exception &e) {
        return fail(KST_EINVAL, e.what());
    }
}

int kst_bank_load_sfz(const kst_config *cfg, const kst_driver_config *dcfg,
                      const kst_load_config *lcfg, const char *path, kst_bank **out) {
    if (!cfg || !dcfg || !lcfg || !path || !out) return fail(KST_EINVAL, "null argument");
    if (cfg->sample_rate == 0) return fail(KST_EINVAL, "sample_rate must be non-zero");
    try {
        std::unique_ptr<kst::Bank> bank(
            kst::load_sfz(bank_settings(*cfg, *dcfg), path, lcfg->resample_pool != 0));
        *out = new kst_bank{std::move(*bank)};
        return KST_OK;
    } catch (const std::exception &e) {
        return fail(KST_EINVAL, e.what());
    }
}

int64_t kst_resample_i16(const int16_t *src, size_t n, double ratio, int16_t *out, size_t cap) {
    if ((!src && n) || (!out && cap)) return fail(KST_EINVAL, "null argument");
    if (!(ratio > 0.0) || !std::isfinite(ratio)) return fail(KST_EINVAL, "ratio must be positive");
    const std::vector<int16_t> v = kst::resample_i16(src, n, ratio);
    std::copy(v.begin(), v.begin() + (ptrdiff_t)std::min(cap, v.size()), out);
    return (int64_t)v.size();
}

int kst_midi_open(const char *path, kst_midi **out) {
    if (!path || !out) return fail(KST_EINVAL, "null argument");
    try {
        *out = new kst_midi{kst::MidiFile(path)};
        return KST_OK;
    } catch (const std::exception &e) {
        return fail(KST_EINVAL, e.what());
    }
}

int kst_midi_info(const kst_midi *m, kst_division *div, uint32_t *format, uint32_t *tracks) {
    if (!m) return fail(KST_EINVAL, "null stream");
    if (div) *div = m->file.division;
    if (format) *format = m->file.format;
    if (tracks) *tracks = m->file.track_count;
    return KST_OK;
}

int64_t kst_midi_read(kst_midi *m, kst_midi_event *out, size_t cap) {
    if (!m || (!out && cap)) return fail(KST_EINVAL, "null argument");
    size_t n = 0;
    while (n < cap && m->file.next(out[n])) ++n;
    return (int64_t)n;
}

void kst_midi_close(kst_midi *m) { delete m; }

int kst_driver_open(const kst_config *cfg, const kst_driver_config *dcfg, const kst_bank *bank,
                    const char *midi_path, kst_driver **out) {
    if (!bank || !midi_path || !out) return fail(KST_EINVAL, "null argument");
    if (int rc = check_driver_config(cfg, dcfg)) return rc;
    std::unique_ptr<kst::EventSource> src;
    try {
        src = std::make_unique<kst::Prefetch>(std::make_unique<kst::MidiFile>(midi_path));
    } catch (const std::exception &e) {
        return fail(KST_EINVAL, e.what());
    }
    return make_driver(cfg, dcfg, bank, std::move(src), out);
}

int kst_driver_create(const kst_config *cfg, const kst_driver_config *dcfg,
                      const kst_bank *bank, const kst_division *div,
                      const kst_midi_event *events, size_t n, kst_driver **out) {
    if (!bank || !div || !out || (!events && n)) return fail(KST_EINVAL, "null argument");
    if (int rc = check_driver_config(cfg, dcfg)) return rc;
    for (size_t i = 1; i < n; ++i)
        if (events[i].tick < events[i - 1].tick)
            return fail(KST_EINVAL, "event " + std::to_string(i) + " is earlier than the one before it");
    return make_driver(cfg, dcfg, bank, std::make_unique<kst::EventList>(events, n, *div), out);
}

int kst_driver_next_block(kst_driver *d, kst_backend *b, float *out, size_t n) {
    if (!d || !b || !out) return fail(KST_EINVAL, "null argument");
    return d->d->next_block(b, out, n);
}

int kst_driver_begin_block(kst_driver *d) {
    if (!d) return fail(KST_EINVAL, "null driver");
    try {
        return (int)d->d->begin_block();
    } catch (const std::exception &e) {
        return fail(KST_EINVAL, e.what());
    }
}

int kst_driver_pending_variant(const kst_driver *d, size_t i, uint32_t *index,
                               const kst_region_params **data, size_t *n) {
    if (!d || !index || !data || !n) return fail(KST_EINVAL, "null argument");
    const auto &u = d->d->uploads();
    if (i >= u.size()) return fail(KST_EINVAL, "variant " + std::to_string(i) + " of " + std::to_string(u.size()));
    *index = u[i].first;
    *data = u[i].second.data();
    *n = u[i].second.size();
    return KST_OK;
}

int kst_driver_rows(const kst_driver *d, const uint32_t **gates, size_t *n_gates,
                    const uint32_t **chans, size_t *n_chans, uint32_t flags[3]) {
    if (!d || !gates || !n_gates || !chans || !n_chans || !flags) return fail(KST_EINVAL, "null argument");
    *gates = d->d->gate_rows().data();
    *n_gates = d->d->gate_rows().size();
    *chans = d->d->chan_rows().data();
    *n_chans = d->d->chan_rows().size();
    d->d->chan_flags(flags);
    return KST_OK;
}

int kst_driver_admit(kst_driver *d, uint64_t live, uint64_t stolen_before,
                     const kst_spawn_cmd **cmds, size_t *n) {
    if (!d || !cmds || !n) return fail(KST_EINVAL, "null argument");
    d->d->admit(live, stolen_before);
    *cmds = d->d->spawned();
    *n = d->d->spawned_count();
    return KST_OK;
}

int kst_driver_end_block(kst_driver *d, float *out, size_t n, uint64_t stolen_after_spawn,
                         const kst_block_stats *after_render) {
    if (!d || !out || !after_render) return fail(KST_EINVAL, "null argument");
    if (n != (size_t)d->d->block_frames() * 2)
        return fail(KST_EINVAL, "output block is " + std::to_string(n) + " samples, expected " +
                                    std::to_string(d->d->block_frames() * 2));
    return d->d->end_block(out, n, stolen_after_spawn, *after_render);
}

int kst_driver_get_stats(const kst_driver *d, kst_driver_stats *out) {
    if (!d || !out) return fail(KST_EINVAL, "null argument");
    *out = d->d->stats();
    return KST_OK;
}

size_t kst_driver_timings(const kst_driver *d, kst_timing *out, size_t cap) {
    if (!d || (!out && cap)) return 0;
    return d->d->timings(out, cap);
}

void kst_driver_destroy(kst_driver *d) { delete d; }

const char *kst_host_last_error(void) { return kst::g_error.c_str(); }

// What the driver does with each controller number (src/driver.rs:58-131):
// the role of every number is the reference's; names are the MIDI 1.0 ones.
int kst_cc_role(uint32_t num, const char **name, const char **note) {
    static const char *NO_DEFAULT = "no General MIDI destination; only a soundfont could bind it";
    static const char *UNDEF = "undefined in the MIDI specification";
    static const char *GENERAL = "general purpose, no defined destination";
    static const char *LSB = "fine half of a controller that is not implemented";
    static const char *FX = "effects are not implemented";
    static const char *PORTA = "needs per-voice glide state";
    const char *n = "", *t = "";
    int role = KST_CC_INERT;
    auto applied = [&](const char *nm) { n = nm; t = ""; role = KST_CC_APPLIED; };
    auto inert = [&](const char *nm, const char *why) { n = nm; t = why; role = KST_CC_INERT; };
    auto missing = [&](const char *nm, const char *why) { n = nm; t = why; role = KST_CC_MISSING; };
    switch (num) {
    case 0: applied("bank select msb"); break;
    case 1: applied("modulation"); break;
    case 2: inert("breath", NO_DEFAULT); break;
    case 4: inert("foot controller", NO_DEFAULT); break;
    case 5: missing("portamento time", PORTA); break;
    case 6: applied("data entry msb"); break;
    case 7: applied("channel volume"); break;
    case 8: inert("balance", "for two-source channels; a sampler voice has one"); break;
    case 10: applied("pan"); break;
    case 11: applied("expression"); break;
    case 12: case 13: missing("effect control", FX); break;
    case 16: case 17: case 18: case 19: inert("general purpose", GENERAL); break;
    case 32: applied("bank select lsb"); break;
    case 33: applied("modulation lsb"); break;
    case 38: applied("data entry lsb"); break;
    case 39: applied("channel volume lsb"); break;
    case 42: applied("pan lsb"); break;
    case 43: applied("expression lsb"); break;
    case 64: applied("sustain pedal"); break;
    case 65: missing("portamento", PORTA); break;
    case 66: applied("sostenuto"); break;
    case 67: applied("soft pedal"); break;
    case 68: missing("legato footswitch", "needs mono voice allocation"); break;
    case 69: missing("hold 2", "a second hold that lengthens the release"); break;
    case 70: inert("sound variation", NO_DEFAULT); break;
    case 71: applied("resonance"); break;
    case 72: applied("release time"); break;
    case 73: applied("attack time"); break;
    case 74: applied("brightness"); break;
    case 75: applied("decay time"); break;
    case 76: applied("vibrato rate"); break;
    case 77: applied("vibrato depth"); break;
    case 78: missing("vibrato delay", "the LFO starts at once"); break;
    case 79: inert("sound controller 10", NO_DEFAULT); break;
    case 80: case 81: case 82: case 83: inert("general purpose", GENERAL); break;
    case 84: missing("portamento control", PORTA); break;
    case 88: missing("high resolution velocity", "velocity is read as 7 bits"); break;
    case 91: missing("reverb send", FX); break;
    case 92: applied("tremolo depth"); break;
    case 93: missing("chorus send", FX); break;
    case 94: missing("celeste depth", FX); break;
    case 95: missing("phaser depth", FX); break;
    case 96: applied("data increment"); break;
    case 97: applied("data decrement"); break;
    case 98: case 99: applied("nrpn select"); break;
    case 100: case 101: applied("rpn select"); break;
    case 120: applied("all sound off"); break;
    case 121: applied("reset controllers"); break;
    case 122: inert("local control", "an offline render has no keyboard"); break;
    case 123: applied("all notes off"); break;
    case 124: applied("omni off"); break;
    case 125: applied("omni on"); break;
    case 126: applied("mono mode on"); break;
    case 127: applied("poly mode on"); break;
    default:
        if (num >= 34 && num <= 63) inert("", LSB);
        else inert("", UNDEF);
        break;
    }
    if (name) *name = n;
    if (note) *note = t;
    return role;
}

}  // extern "C"
