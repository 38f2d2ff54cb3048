// kst_host.h -- host-side classes of the driver's block loop: the flattened
// instrument bank, the SMF event stream and tempo clock, and the driver.
// The C ABI over them is in include/kestrel_hip.h; nothing here touches the GPU.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <memory>
#include <mutex>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "kestrel_hip.h"

namespace kst {

// ---------------------------------------------------------------------------
// Saturating conversions with Rust's `as` semantics (NaN -> 0, clamp to range),
// which several reference formulas rely on.
// ---------------------------------------------------------------------------
inline uint32_t sat_u32(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 4294967296.0f) return UINT32_MAX;
    return (uint32_t)v;
}
inline uint64_t sat_u64(double v) {
    if (!(v > 0.0)) return 0;
    if (v >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)v;
}
inline int64_t sat_i64(float v) {
    if (v != v) return 0;
    if (v >= 9223372036854775808.0f) return INT64_MAX;
    if (v < -9223372036854775808.0f) return INT64_MIN;
    return (int64_t)v;
}

// ---------------------------------------------------------------------------
// Bank: src/bank.rs. Regions arrive with every generator resolved; the bank
// lays out the params table, designs variants and answers note-on previews.
// ---------------------------------------------------------------------------
struct ParamMod {  // src/bank.rs:206-217
    float cutoff_cents = 0.0f, q_cb = 0.0f;
    float attack_scale = 1.0f, attack_add = 0.0f;
    float decay_scale = 1.0f, decay_add = 0.0f;
    float release_scale = 1.0f, release_add = 0.0f;
    static ParamMod from_controllers(uint8_t cc71, uint8_t cc72, uint8_t cc73, uint8_t cc74,
                                     uint8_t cc75);
};

struct Preview {  // PreviewLayer, src/bank.rs:421-431
    uint32_t region;
    float gain_l, gain_r;
    uint32_t delay_frames;
};

struct Layer {  // VoiceSpawn, src/bank.rs:433-450
    uint64_t phase, step;
    uint32_t smp_base, smp_len, loop_start, loop_end, flags, params, region;
    float gain_l, gain_r;
    uint32_t delay_frames;
    uint32_t exclusive_class;
};

struct BankSettings {
    uint32_t sample_rate = 48000, decay_curve = 0, release_curve = 0, filter_enabled = 1;
    float env_floor = 1.0e-5f, master_volume = 1.0f;
};

class Bank {
public:
    std::vector<int16_t> pool;
    uint32_t pool_rate = 0;
    std::vector<kst_sample_info> samples;
    std::vector<kst_region> regions;
    BankSettings set;
    // Per region: params layout (src/bank.rs:689-721).
    std::vector<uint32_t> params_base, params_stride, params_vel_span;
    std::vector<kst_region_params> params;  // variant 0
    uint32_t max_sample_len = 0;
    std::vector<kst_preset> preset_desc;      // as given: regions are
    std::vector<uint32_t> preset_regions;     // preset_regions[first .. first+count]

    // Builds the layout, variant 0 and the admission tables. Throws
    // std::invalid_argument naming what is wrong with the input.
    Bank(const BankSettings &s, const int16_t *pool, size_t pool_len, uint32_t pool_rate,
         const kst_sample_info *samples, size_t n_samples, const kst_region *regions,
         size_t n_regions, const kst_preset *presets, size_t n_presets,
         const uint32_t *preset_regions, size_t n_preset_regions);

    int64_t find_preset(uint32_t bank, uint32_t program) const;  // -1: no presets
    void preview_note_on(uint32_t preset, uint8_t key, uint8_t vel, size_t max_layers,
                         std::vector<Preview> &out) const;
    bool build_layer(uint32_t region, uint8_t key, uint8_t vel, Layer &out) const;
    void build_variant(const ParamMod &m, std::vector<kst_region_params> &out) const;

private:
    struct PresetEntry {
        uint32_t bank, program;
        std::vector<uint32_t> key_index;    // 129 offsets into key_regions
        std::vector<uint32_t> key_regions;
    };
    std::vector<PresetEntry> presets_;
    std::vector<std::pair<uint64_t, uint32_t>> index_;  // ((bank<<16)|program, preset), sorted
    std::vector<float> gain_table_;                     // region*128+vel -> (l, r)
    std::vector<uint32_t> delay_frames_;
    std::vector<uint64_t> key_ok_;
    // What build_voice computes from the region alone, and the phase step of
    // each (region, key): a layer build is then table lookups.
    struct RegionPlan {
        uint32_t start_offset, smp_len, loop_start, loop_end, flags;
    };
    std::vector<RegionPlan> plan_;
    std::vector<uint64_t> step_;                        // region*128+key
    bool key_ok(uint32_t region, uint8_t key) const {
        const size_t bit = (size_t)region * 128 + key;
        return (key_ok_[bit / 64] >> (bit % 64)) & 1;
    }
    bool build_voice(const kst_region &r, uint32_t ri, uint8_t key, uint8_t vel, Layer &out) const;
    kst_region_params make_params(const kst_region &r, uint8_t key, uint8_t vel,
                                  const ParamMod &m) const;
};

// ---------------------------------------------------------------------------
// Event sources: the streaming SMF reader (src/midi.rs:38-358) or a list.
// ---------------------------------------------------------------------------
class EventSource {
public:
    virtual ~EventSource() = default;
    // Next event in tick order; false at the end.
    virtual bool next(kst_midi_event &ev) = 0;
    // The next run of events in tick order, [*first, *first + n), valid until
    // the next call; 0 at the end. By default one event at a time.
    virtual size_t span(const kst_midi_event **first) {
        if (!next(one_)) return 0;
        *first = &one_;
        return 1;
    }
    kst_division division{};

private:
    kst_midi_event one_{};
};

class MidiFile : public EventSource {
public:
    // Throws std::runtime_error with the reference's message on a bad file.
    explicit MidiFile(const std::string &path);
    ~MidiFile() override;
    bool next(kst_midi_event &ev) override;
    uint32_t format = 0, track_count = 0;

private:
    struct Track {
        int fd = -1;
        uint64_t off = 0;  // file offset of the next unread chunk byte
        std::vector<uint8_t> buf;
        size_t pos = 0, filled = 0;
        uint64_t remaining = 0, tick = 0;
        uint8_t running = 0;
        bool ended = false;
        bool fill();
        bool byte(uint8_t &b);
        bool varlen(uint64_t &v);
        bool skip(uint64_t n);
        bool next_event(kst_midi_event &ev);
        // Decoded events not yet merged, [at, n) of ev.
        std::vector<kst_midi_event> ev;
        size_t at = 0, n = 0;
        bool refill();
    };
    int fd_ = -1;
    std::vector<Track> tracks_;
    // Winner tree over the tracks' head events keyed (tick << 20 | track), so
    // one integer minimum orders by tick and then by track; an exhausted
    // track's key is all ones. node_[1] is the next event; leaves sit at
    // [leaves_, 2 * leaves_). A tick too wide for the key (2^44 and up)
    // switches the merge to a scan over (tick, track) pairs.
    uint32_t leaves_ = 1;
    std::vector<uint64_t> node_;
    bool wide_ = false;
    // A run: track run_w_ keeps winning while its head key stays below the
    // best of the other tracks, run_limit_ -- dense files give runs of many
    // events at one tick, emitted without touching the tree.
    bool run_ = false;
    uint32_t run_w_ = 0;
    uint64_t run_limit_ = 0;
    uint64_t leaf_key(uint32_t t) const;
    void update(uint32_t t);
    uint32_t scan_min() const;
};

class EventList : public EventSource {
public:
    EventList(const kst_midi_event *ev, size_t n, const kst_division &div)
        : events_(ev, ev + n) {
        division = div;
    }
    bool next(kst_midi_event &ev) override {
        if (at_ >= events_.size()) return false;
        ev = events_[at_++];
        return true;
    }
    size_t span(const kst_midi_event **first) override {
        *first = events_.data() + at_;
        const size_t n = events_.size() - at_;
        at_ = events_.size();
        return n;
    }

private:
    std::vector<kst_midi_event> events_;
    size_t at_ = 0;
};

// src/sf2.rs: load a SoundFont 2 file into a bank (throws std::runtime_error
// with the reference's message). src/resample.rs: the pool's rate converter.
Bank *load_sf2(const BankSettings &set, const std::string &path, bool resample_pool,
               uint64_t budget);
std::vector<int16_t> resample_i16(const int16_t *src, size_t n, double ratio);
// src/sfz.rs: load an SFZ instrument (and the WAV files it names) into a bank.
Bank *load_sfz(const BankSettings &set, const std::string &path, bool resample_pool);

// Decodes another source ahead on a thread of its own, in chunks through a
// bounded queue: the consumer sees exactly the inner sequence, and the file's
// decoding leaves the block loop's critical path. Consumed chunks go back to
// the decoder for reuse, so a long file streams through a fixed set of
// buffers (no allocation, no fresh pages) once the queue has filled.
class Prefetch : public EventSource {
public:
    explicit Prefetch(std::unique_ptr<EventSource> inner, size_t chunk = 1 << 15, size_t depth = 32);
    ~Prefetch() override;
    bool next(kst_midi_event &ev) override;
    size_t span(const kst_midi_event **first) override;  // the rest of the current chunk

private:
    void run();
    bool refill();  // the next chunk into cur_; false at the end
    std::unique_ptr<EventSource> inner_;
    size_t chunk_, depth_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::vector<kst_midi_event>> q_;
    std::vector<std::vector<kst_midi_event>> spare_;  // consumed chunks to refill
    bool done_ = false, stop_ = false;
    std::vector<kst_midi_event> cur_;
    size_t at_ = 0;
    std::thread th_;
};

// src/midi.rs:363-418
class TempoClock {
public:
    TempoClock(const kst_division &d, uint32_t sample_rate);
    void set_tempo(uint64_t tick, uint32_t us_per_qn);
    double frame_at(uint64_t tick) const {
        return base_frame_ + (double)(tick - base_tick_) * frames_per_tick_;
    }

private:
    void set_raw(uint32_t us_per_qn);
    kst_division div_;
    double sample_rate_;
    uint64_t base_tick_ = 0;
    double base_frame_ = 0.0, frames_per_tick_ = 0.0;
};

}  // namespace kst
