// kestrel_oracle.cpp -- CPU oracle for the Kestrel render hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker the HIP backend is compared against.
// It restates the reference's CPU backend statement for statement:
//
//   Pool, push, stable compaction ........ src/cpu.rs:24-142
//   spawn (steal / drop / thinning) ...... src/cpu.rs:330-390
//   render (per voice, per frame) ........ src/cpu.rs:392-639
//   fetch / neighbour index / interp ..... src/cpu.rs:187-302
//   steal key ............................ src/cpu.rs:208-221
//   32.32 phase, 8.24 bend ............... src/fixed.rs:75-171
//   derived config arithmetic ............ src/config.rs:120-131,532-552
//   spawn_pick ........................... src/voice.rs:72-75
//   sort key layout ...................... src/gpu/mod.rs:76-105, shaders/sort.wgsl:46-72
//   coefficient design ................... src/bank.rs:876-979
//
// It is compiled with -ffp-contract=off so no multiply-add is fused: every f32
// operation rounds exactly where the Rust source rounds. Rust's float->int `as`
// casts saturate and map NaN to zero; the helpers below do the same, where C++
// would be undefined.

#include "kestrel_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

namespace {
// Runs fn(lo, hi) over [0, n) split into `threads` contiguous ranges. Only
// independent per-index work goes through here, so the result does not depend
// on the thread count.
template <typename Fn>
void par_for(size_t n, int threads, Fn fn) {
    if (threads <= 1 || n < 65536) {
        fn((size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int k = 0; k < threads; ++k)
        th.emplace_back([=]() { fn(n * k / threads, n * (k + 1) / threads); });
    for (auto &t : th) t.join();
}
}  // namespace

namespace {

constexpr uint32_t ENV_ATTACK = 0, ENV_DECAY = 1, ENV_SUSTAIN = 2, ENV_RELEASE = 3,
                   ENV_DEAD = 4;
constexpr uint32_t VF_LOOP = 1, VF_LOOP_UNTIL_RELEASE = 2;
constexpr uint32_t RP_FILTER = 1;
constexpr size_t GATE_SLOTS = 2048;
constexpr size_t BEND_CHANNELS = 16, CHAN_FIELDS = 4;
constexpr size_t CHAN_BEND = 0, CHAN_GAIN_L = 1, CHAN_GAIN_R = 2, CHAN_VARIANT = 3;
constexpr float FRAC_SCALE_F32 = 1.0f / 4294967296.0f;  // src/fixed.rs:73
constexpr uint32_t BEND_FRAC_BITS = 24;                  // src/fixed.rs:154
constexpr uint32_t BEND_ONE = 1u << BEND_FRAC_BITS;

// Rust `f as u32`: saturating, NaN -> 0.
inline uint32_t rs_f32_to_u32(float x) {
    if (std::isnan(x) || x <= 0.0f) return 0;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

inline float bits_f32(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
inline uint32_t f32_bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

// Rust f32::clamp (NaN passes through).
template <typename R>
inline R rs_clamp(R x, R lo, R hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

inline uint64_t fixed_scale(uint64_t phase, uint32_t factor) {
    // src/fixed.rs:141-148
    unsigned __int128 p = ((unsigned __int128)phase * (unsigned __int128)factor) >> BEND_FRAC_BITS;
    if (p > (unsigned __int128)UINT64_MAX) return UINT64_MAX;
    return (uint64_t)p;
}

uint32_t max_steal(const kor_config &c) {
    // src/config.rs:549-552
    uint64_t n = (uint64_t)c.max_voices * (uint64_t)c.max_steal_percent / 100;
    uint32_t v = (uint32_t)n;
    return v > 1 ? v : 1;
}

uint32_t steal_span(const kor_config &c) {
    // src/config.rs:534-536
    uint32_t s = c.block_frames > c.steal_fade_frames ? c.block_frames - c.steal_fade_frames : 0;
    return s > 1 ? s : 1;
}

inline uint64_t spawn_pick(uint64_t i, uint64_t want, uint64_t take) {
    // src/voice.rs:72-75
    return i * want / take;
}

// Per-precision copy of one region's DSP constants.
template <typename R>
struct Params {
    R attack_rate, attack_end, decay_coef, decay_target, sustain, release_coef;
    R b0, b1, a1, a2;
    uint32_t flags;
};

template <typename R>
Params<R> widen(const kor_region_params &p) {
    return Params<R>{(R)p.attack_rate, (R)p.attack_end, (R)p.decay_coef, (R)p.decay_target,
                     (R)p.sustain,     (R)p.release_coef, (R)p.b0,        (R)p.b1,
                     (R)p.a1,          (R)p.a2,           p.flags};
}
Params<double> widen64(const kor_region_params64 &p) {
    return Params<double>{p.attack_rate, p.attack_end, p.decay_coef, p.decay_target,
                          p.sustain,     p.release_coef, p.b0,        p.b1,
                          p.a1,          p.a2,           p.flags};
}

// src/cpu.rs:24-52. Float state is kept at the render precision R; with R =
// float it is bit for bit the reference's.
template <typename R>
struct Pool {
    std::vector<uint64_t> phase, step;
    std::vector<uint32_t> smp_base, smp_len, loop_start, loop_end, flags, env_stage;
    std::vector<R> env_level;
    std::vector<float> gain_l, gain_r;
    std::vector<R> filt_z1, filt_z2;
    std::vector<uint32_t> params, region, gate_slot, ordinal, start_rel, born_variant, stop_rel;
    std::vector<uint64_t> note_id;

    size_t len() const { return phase.size(); }

    void push(const kor_spawn_cmd &c) {
        // src/cpu.rs:59-82
        phase.push_back(((uint64_t)c.phase_hi << 32) | c.phase_lo);
        step.push_back(((uint64_t)c.step_hi << 32) | c.step_lo);
        smp_base.push_back(c.smp_base);
        smp_len.push_back(c.smp_len);
        loop_start.push_back(c.loop_start);
        loop_end.push_back(c.loop_end);
        flags.push_back(c.flags);
        env_stage.push_back(ENV_ATTACK);
        env_level.push_back((R)0);
        gain_l.push_back(c.gain_l);
        gain_r.push_back(c.gain_r);
        filt_z1.push_back((R)0);
        filt_z2.push_back((R)0);
        params.push_back(c.params);
        region.push_back(c.region);
        gate_slot.push_back(c.gate_slot);
        ordinal.push_back(c.ordinal);
        start_rel.push_back(c.start_rel);
        born_variant.push_back(c.variant + 1);
        stop_rel.push_back(0);
        note_id.push_back(((uint64_t)c.note_id_hi << 32) | (uint64_t)c.note_id_lo);
    }

    template <typename F>
    void for_each_field(F &&f) {
        f(phase); f(step); f(smp_base); f(smp_len); f(loop_start); f(loop_end); f(flags);
        f(env_stage); f(env_level); f(gain_l); f(gain_r); f(filt_z1); f(filt_z2); f(params);
        f(region); f(gate_slot); f(ordinal); f(start_rel); f(born_variant); f(stop_rel);
        f(note_id);
    }

    // src/cpu.rs:84-117: stable compaction, order preserved.
    void compact(const std::vector<uint8_t> &keep) {
        size_t w = 0;
        for (size_t r = 0; r < keep.size() && r < len(); ++r) {
            if (keep[r]) {
                if (w != r) for_each_field([&](auto &v) { v[w] = v[r]; });
                ++w;
            }
        }
        for_each_field([&](auto &v) { v.resize(w); });
    }

    // Reorder by a permutation (sorted mode): new[j] = old[perm[j]].
    void permute(const std::vector<uint32_t> &perm, int threads) {
        std::vector<std::thread> th;
        for_each_field([&](auto &v) {
            auto job = [&v, &perm]() {
                auto old = v;
                v.resize(perm.size());
                for (size_t j = 0; j < perm.size(); ++j) v[j] = old[perm[j]];
            };
            if (threads > 1 && perm.size() >= 65536)
                th.emplace_back(job);
            else
                job();
        });
        for (auto &t : th) t.join();
    }
};

}  // namespace

struct kor_synth {
    kor_config cfg;
    int precision;  // 0 f32, 1 f64
    int state_only;
    std::vector<int16_t> bank_pool;
    std::vector<kor_region_params> bank_params;
    std::vector<kor_region_params64> bank_params64;
    std::vector<std::vector<kor_region_params>> variants;
    Pool<float> pool32;
    Pool<double> pool64;
    std::vector<double> mix;
    std::vector<uint32_t> gate_rows, chan_rows;
    bool bend_active = false, gain_active = false, variant_active = false;
    size_t tiles, tile_frames;
    uint64_t stolen = 0, dropped = 0;
    float peak = 0.0f;
    bool sorted = false;
    uint32_t sort_layout[6] = {0, 0, 0, 0, 0, 0};
    int threads = 1;
};

namespace {

// ---------------------------------------------------------------------------
// render, templated on the arithmetic precision
// ---------------------------------------------------------------------------

// The slot indices in stable order of their 32-bit keys: an LSD counting sort,
// 16 bits per pass. A stable sort by key has one result, so this is the order
// std::stable_sort gives, in linear time (pools of 67M voices).
std::vector<uint32_t> stable_order(const std::vector<uint32_t> &key) {
    const size_t n = key.size();
    std::vector<uint32_t> a(n), b(n);
    std::iota(a.begin(), a.end(), 0u);
    std::vector<size_t> cnt(65537);
    for (int shift = 0; shift < 32; shift += 16) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (size_t i = 0; i < n; ++i) ++cnt[((key[a[i]] >> shift) & 0xFFFFu) + 1];
        for (size_t d = 0; d < 65536; ++d) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; ++i) b[cnt[(key[a[i]] >> shift) & 0xFFFFu]++] = a[i];
        a.swap(b);
    }
    return a;
}

template <typename R>
struct Renderer {
    kor_synth &s;
    Pool<R> &pool;

    // src/cpu.rs:190-194: divide by 32767 the way unpack2x16snorm does, clamp at -1.
    R fetch(uint32_t base, uint32_t idx) const {
        uint32_t i = base + idx;  // u32 arithmetic, as in the reference
        int16_t v = i < s.bank_pool.size() ? s.bank_pool[i] : (int16_t)0;
        if constexpr (sizeof(R) == 4) {
            float x = (float)v * (1.0f / 32767.0f);
            return std::fmax(x, -1.0f);
        } else {
            double x = (double)v / 32767.0;
            return std::fmax(x, -1.0);
        }
    }

    // src/cpu.rs:243-259
    static uint32_t advance_index(uint32_t idx, int32_t off, bool looping, uint32_t loop_start,
                                  uint32_t loop_end, uint32_t len) {
        int64_t raw = (int64_t)idx + (int64_t)off;
        if (looping) {
            int64_t ls = loop_start, le = loop_end;
            if (raw >= le) {
                int64_t span = std::max<int64_t>(le - ls, 1);
                return (uint32_t)(ls + (raw - ls) % span);
            } else if (raw < 0) {
                return 0;
            }
            return (uint32_t)raw;
        }
        int64_t hi = (int64_t)len - 1;
        // Rust i64::clamp(0, len-1); len >= 1 for any voice the driver builds.
        if (raw < 0) raw = 0;
        if (raw > hi) raw = hi;
        return (uint32_t)raw;
    }

    // src/cpu.rs:295-302, the shader's Horner form.
    static R catmull_rom(R sm1, R s0, R s1, R s2, R t) {
        R a = (R)-0.5 * sm1 + (R)1.5 * s0 - (R)1.5 * s1 + (R)0.5 * s2;
        R b = sm1 - (R)2.5 * s0 + (R)2.0 * s1 - (R)0.5 * s2;
        R c = (R)-0.5 * sm1 + (R)0.5 * s1;
        return ((a * t + b) * t + c) * t + s0;
    }

    R interpolate(uint32_t base, uint32_t idx, R frac, bool looping, uint32_t ls, uint32_t le,
                  uint32_t len) const {
        switch (s.cfg.interpolation) {
            case 0:
                return fetch(base, idx);
            case 1: {
                uint32_t i1 = advance_index(idx, 1, looping, ls, le, len);
                R s0 = fetch(base, idx);
                R s1 = fetch(base, i1);
                return s0 + (s1 - s0) * frac;
            }
            default: {
                uint32_t im1 = advance_index(idx, -1, looping, ls, le, len);
                uint32_t i1 = advance_index(idx, 1, looping, ls, le, len);
                uint32_t i2 = advance_index(idx, 2, looping, ls, le, len);
                return catmull_rom(fetch(base, im1), fetch(base, idx), fetch(base, i1),
                                   fetch(base, i2), frac);
            }
        }
    }

    // src/cpu.rs:228-238
    Params<R> params_of(uint32_t variant, size_t base) const {
        if constexpr (sizeof(R) == 8) {
            if (variant == 0 && !s.bank_params64.empty()) return widen64(s.bank_params64[base]);
        }
        if (variant == 0) return widen<R>(s.bank_params[base]);
        size_t i = variant - 1;
        if (i < s.variants.size() && base < s.variants[i].size())
            return widen<R>(s.variants[i][base]);
        if constexpr (sizeof(R) == 8) {
            if (!s.bank_params64.empty()) return widen64(s.bank_params64[base]);
        }
        return widen<R>(s.bank_params[base]);
    }

    // One voice, src/cpu.rs:403-613, accumulating into `mix`.
    void render_voice(size_t v, double *mix) {
        const kor_config &cfg = s.cfg;
        const bool exp_decay = cfg.decay_curve == 0;
        const bool exp_release = cfg.release_curve == 0;
        const R floor = (R)cfg.env_floor;
        const size_t fade = cfg.steal_fade_frames;
        const bool state_only = s.state_only != 0;

        uint32_t stage0 = pool.env_stage[v];
        if (stage0 == ENV_DEAD) return;
        size_t params_base = pool.params[v];
        uint32_t born_variant = pool.born_variant[v];
        uint32_t variant = born_variant > 0 ? born_variant - 1 : 0;  // saturating_sub
        Params<R> p = params_of(variant, params_base);
        bool use_filter = (p.flags & RP_FILTER) != 0;

        uint32_t base = pool.smp_base[v];
        uint32_t len = pool.smp_len[v];
        uint32_t loop_start = pool.loop_start[v];
        uint32_t loop_end = pool.loop_end[v];
        uint32_t vflags = pool.flags[v];
        bool loop_enabled = (vflags & VF_LOOP) != 0;
        bool loop_until_release = (vflags & VF_LOOP_UNTIL_RELEASE) != 0;
        size_t gate_slot = pool.gate_slot[v];
        uint32_t ordinal = pool.ordinal[v];
        float base_gain_l = pool.gain_l[v];
        float base_gain_r = pool.gain_r[v];
        float gain_l = base_gain_l;
        float gain_r = base_gain_r;
        uint64_t base_step = pool.step[v];
        size_t channel = (size_t)((uint32_t)gate_slot >> 7) & 15;  // ChannelTable::channel_of
        uint64_t step = base_step;

        uint64_t phase = pool.phase[v];
        uint32_t stage = stage0;
        R level = pool.env_level[v];
        R z1 = pool.filt_z1[v];
        R z2 = pool.filt_z2[v];
        size_t start_rel = pool.start_rel[v];
        size_t stop_rel = pool.stop_rel[v];
        size_t born_tile = start_rel / s.tile_frames;

        for (size_t tile = 0; tile < s.tiles; ++tile) {
            // Note-off gate, sampled once per gate tile.
            if (stage < ENV_RELEASE && s.gate_rows[tile * GATE_SLOTS + gate_slot] >= ordinal) {
                stage = ENV_RELEASE;
                level = std::fmin(level, (R)1);
            }
            if (s.bend_active || s.gain_active || s.variant_active) {
                size_t ci = (tile * BEND_CHANNELS + channel) * CHAN_FIELDS;
                if (s.bend_active) step = fixed_scale(base_step, s.chan_rows[ci + CHAN_BEND]);
                if (s.gain_active) {
                    gain_l = base_gain_l * bits_f32(s.chan_rows[ci + CHAN_GAIN_L]);
                    gain_r = base_gain_r * bits_f32(s.chan_rows[ci + CHAN_GAIN_R]);
                }
                if (s.variant_active) {
                    uint32_t want = s.chan_rows[ci + CHAN_VARIANT];
                    bool born_here = born_variant != 0 && tile <= born_tile;
                    if (want != variant && !born_here) {
                        variant = want;
                        p = params_of(variant, params_base);
                        use_filter = (p.flags & RP_FILTER) != 0;
                    }
                }
            }

            size_t f0 = tile * s.tile_frames;
            for (size_t i = 0; i < s.tile_frames; ++i) {
                size_t f = f0 + i;
                if (f < start_rel) continue;
                if (stage == ENV_DEAD) goto done;

                bool looping = loop_enabled && !(loop_until_release && stage >= ENV_RELEASE);
                uint32_t idx = (uint32_t)(phase >> 32);
                if (!looping && idx >= len) {
                    stage = ENV_DEAD;
                    goto done;
                }

                switch (stage) {
                    case ENV_ATTACK:
                        level += p.attack_rate;
                        if (level >= p.attack_end) {
                            stage = ENV_DECAY;
                            level = (R)1;
                        }
                        break;
                    case ENV_DECAY:
                        level = exp_decay ? level * p.decay_coef : level - p.decay_coef;
                        if (level <= p.decay_target) {
                            if (p.sustain <= floor) {
                                stage = ENV_DEAD;
                                level = (R)0;
                            } else {
                                stage = ENV_SUSTAIN;
                                level = p.sustain;
                            }
                        }
                        break;
                    case ENV_SUSTAIN:
                        break;
                    case ENV_RELEASE:
                        level = exp_release ? level * p.release_coef : level - p.release_coef;
                        if (level <= floor) {
                            stage = ENV_DEAD;
                            level = (R)0;
                        }
                        break;
                    default:
                        break;
                }

                if (!state_only) {
                    R frac;
                    if constexpr (sizeof(R) == 4)
                        frac = (float)(uint32_t)phase * FRAC_SCALE_F32;
                    else
                        frac = (double)(uint32_t)phase / 4294967296.0;
                    R smp = interpolate(base, idx, frac, looping, loop_start, loop_end, len);
                    R g = std::fmin(level, (R)1);
                    R x = smp * g;
                    R y;
                    if (use_filter) {
                        // Transposed direct form II, b2 == b0.
                        y = p.b0 * x + z1;
                        z1 = p.b1 * x - p.a1 * y + z2;
                        z2 = p.b0 * x - p.a2 * y;
                    } else {
                        y = x;
                    }
                    if (stop_rel != 0 && f + 1 >= stop_rel) {
                        size_t d = f + 1 - stop_rel;
                        if (d >= fade) {
                            stage = ENV_DEAD;
                            y = (R)0;
                        } else {
                            y = y * ((R)1 - (R)d / (R)fade);
                        }
                    }
                    mix[f * 2] += (double)(y * (R)gain_l);
                    mix[f * 2 + 1] += (double)(y * (R)gain_r);
                } else if (stop_rel != 0 && f + 1 >= stop_rel) {
                    if (f + 1 - stop_rel >= fade) stage = ENV_DEAD;
                }

                // ---- advance the phase ----
                phase += step;  // wrapping_add
                if (looping) {
                    uint32_t hi = (uint32_t)(phase >> 32);
                    if (hi >= loop_end) {
                        uint32_t span = loop_end - loop_start;
                        if (span < 1) span = 1;
                        uint32_t wrapped = loop_start + (hi - loop_start) % span;
                        phase = ((uint64_t)wrapped << 32) | (phase & 0xFFFFFFFFull);
                    }
                }
            }
        }
    done:
        pool.phase[v] = phase;
        pool.env_stage[v] = stage;
        pool.env_level[v] = level;
        pool.filt_z1[v] = z1;
        pool.filt_z2[v] = z2;
        pool.start_rel[v] = 0;
        pool.born_variant[v] = 0;
        pool.stop_rel[v] = 0;
    }

    void render_all() {
        size_t n = pool.len();
        std::fill(s.mix.begin(), s.mix.end(), 0.0);
        int t = s.threads;
        if (t <= 1 || n < 4096) {
            for (size_t v = 0; v < n; ++v) render_voice(v, s.mix.data());
            return;
        }
        // Voice-partitioned threads, each with its own f64 mix, merged in
        // thread order so the result does not depend on scheduling.
        std::vector<std::vector<double>> mixes(t, std::vector<double>(s.mix.size(), 0.0));
        std::vector<std::thread> th;
        for (int k = 0; k < t; ++k) {
            size_t lo = n * k / t, hi = n * (k + 1) / t;
            th.emplace_back([this, lo, hi, k, &mixes]() {
                for (size_t v = lo; v < hi; ++v) render_voice(v, mixes[k].data());
            });
        }
        for (auto &x : th) x.join();
        for (int k = 0; k < t; ++k)
            for (size_t j = 0; j < s.mix.size(); ++j) s.mix[j] += mixes[k][j];
    }

    uint64_t steal_key(size_t i) const {
        // src/cpu.rs:213-221
        uint64_t id = pool.note_id[i];
        if (s.cfg.steal_rule != 2) return id;
        float level = rs_clamp((float)pool.env_level[i], 0.0f, 1.0f);
        uint64_t q = (uint64_t)rs_f32_to_u32(level * 65535.0f);
        return (q << 48) | (id & 0x0000FFFFFFFFFFFFull);
    }

    int spawn(const kor_spawn_cmd *cmds, size_t want) {
        // src/cpu.rs:330-390
        const kor_config &cfg = s.cfg;
        size_t cap = cfg.max_voices;
        size_t live = pool.len();
        auto pick = [&](size_t i, size_t w, size_t t) -> size_t {
            return cfg.admit_rule == 0 ? i : (size_t)spawn_pick(i, w, t);
        };
        if (live + want > cap) {
            if (cfg.steal_rule == 1) {  // DropNew
                size_t room = cap > live ? cap - live : 0;
                size_t take = std::min(want, room);
                s.dropped += want - take;
                for (size_t i = 0; i < take; ++i) pool.push(cmds[pick(i, want, take)]);
                return 0;
            }
            size_t need = live + want > cap ? live + want - cap : 0;
            need = std::min(need, live);
            need = std::min(need, (size_t)max_steal(cfg));
            if (need > 0) {
                // The victim set is the `need` smallest keys; keys are a total
                // order with no ties, so any selection algorithm names the same
                // set as select_nth_unstable_by_key.
                // The threshold is the need-th smallest key; exactly `need`
                // keys are at or below it.
                std::vector<uint64_t> keys(live);
                par_for(live, s.threads, [&](size_t lo, size_t hi) {
                    for (size_t i = lo; i < hi; ++i) keys[i] = steal_key(i);
                });
                std::vector<uint64_t> tmp(keys);
                std::nth_element(tmp.begin(), tmp.begin() + (need - 1), tmp.end());
                const uint64_t threshold = tmp[need - 1];
                uint32_t span = steal_span(cfg);
                for (size_t i = 0; i < live; ++i) {
                    if (keys[i] > threshold) continue;
                    uint32_t id = (uint32_t)pool.note_id[i];
                    pool.stop_rel[i] = id % span + 1;
                }
                s.stolen += need;
            }
            size_t take = std::min(want, cap - (live - need));
            s.dropped += want - take;
            for (size_t i = 0; i < take; ++i) pool.push(cmds[pick(i, want, take)]);
            return 0;
        }
        for (size_t i = 0; i < want; ++i) pool.push(cmds[i]);
        return 0;
    }

    void sort_and_compact() {
        size_t n = pool.len();
        if (!s.sorted) {
            std::vector<uint8_t> keep(n);
            bool any_dead = false;
            for (size_t i = 0; i < n; ++i) {
                keep[i] = pool.env_stage[i] != ENV_DEAD;
                any_dead |= !keep[i];
            }
            if (any_dead) pool.compact(keep);
            return;
        }
        // shaders/sort.wgsl:46-72 keys, then a stable sort; dead voices carry
        // dead_region and so sort past every survivor.
        const uint32_t *L = s.sort_layout;
        uint32_t region_shift = L[1], stage_shift = L[2], phase_shift = L[3], phase_mask = L[4],
                 dead_region = L[5];
        std::vector<uint32_t> key(n);
        par_for(n, s.threads, [&](size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; ++i) {
                uint32_t stage = pool.env_stage[i];
                uint32_t region = dead_region;
                if (stage != ENV_DEAD) region = std::min(pool.region[i], dead_region - 1);
                uint32_t ph = ((uint32_t)(pool.phase[i] >> 32) >> phase_shift) & phase_mask;
                key[i] = (region << region_shift) | (std::min(stage, 7u) << stage_shift) | ph;
            }
        });
        size_t live_new = 0;
        for (size_t i = 0; i < n; ++i) live_new += pool.env_stage[i] != ENV_DEAD;
        std::vector<uint32_t> perm = stable_order(key);
        perm.resize(live_new);
        pool.permute(perm, s.threads);
    }
};

template <typename R>
Pool<R> &pool_of(kor_synth *s);
template <>
Pool<float> &pool_of<float>(kor_synth *s) { return s->pool32; }
template <>
Pool<double> &pool_of<double>(kor_synth *s) { return s->pool64; }

template <typename R>
int render_block(kor_synth *s, float *out32, double *out64, size_t n) {
    size_t block = s->cfg.block_frames;
    if (n != block * 2) return -1;
    Renderer<R> r{*s, pool_of<R>(s)};
    r.render_all();
    // src/cpu.rs:615-625
    float peak = 0.0f;
    for (size_t j = 0; j < block * 2; ++j) {
        float v = (float)s->mix[j];
        if (out32) out32[j] = v;
        if (out64) out64[j] = s->mix[j];
        float a = std::fabs(v);
        if (a > peak) peak = a;
    }
    s->peak = peak;
    r.sort_and_compact();
    return 0;
}

}  // namespace

template <typename R>
static void export_pool(const Pool<R> &p, uint32_t *o, size_t stride, int threads) {
    const size_t n = p.len();
    // One field per task, each written front to back.
    auto field = [&](int f) {
        uint32_t *d = o + (size_t)f * stride;
        for (size_t i = 0; i < n; ++i) {
            switch (f) {
                case 0: d[i] = (uint32_t)p.phase[i]; break;
                case 1: d[i] = (uint32_t)(p.phase[i] >> 32); break;
                case 2: d[i] = (uint32_t)p.step[i]; break;
                case 3: d[i] = (uint32_t)(p.step[i] >> 32); break;
                case 4: d[i] = p.smp_base[i]; break;
                case 5: d[i] = p.smp_len[i]; break;
                case 6: d[i] = p.loop_start[i]; break;
                case 7: d[i] = p.loop_end[i]; break;
                case 8: d[i] = p.flags[i]; break;
                case 9: d[i] = p.env_stage[i]; break;
                case 10: d[i] = f32_bits((float)p.env_level[i]); break;
                case 11: d[i] = f32_bits(p.gain_l[i]); break;
                case 12: d[i] = f32_bits(p.gain_r[i]); break;
                case 13: d[i] = f32_bits((float)p.filt_z1[i]); break;
                case 14: d[i] = f32_bits((float)p.filt_z2[i]); break;
                case 15: d[i] = p.params[i]; break;
                case 16: d[i] = p.region[i]; break;
                case 17: d[i] = p.gate_slot[i]; break;
                case 18: d[i] = p.ordinal[i]; break;
                case 19: d[i] = p.start_rel[i]; break;
                case 20: d[i] = (uint32_t)p.note_id[i]; break;
                case 21: d[i] = (uint32_t)(p.note_id[i] >> 32); break;
                case 22: d[i] = p.born_variant[i]; break;
                default: d[i] = p.stop_rel[i]; break;
            }
        }
    };
    par_for(24, threads > 1 && n >= 65536 ? 24 : 1, [&](size_t lo, size_t hi) {
        for (size_t f = lo; f < hi; ++f) field((int)f);
    });
}

template <typename R>
static void import_pool(Pool<R> &p, const uint32_t *o, size_t n, size_t stride, int threads) {
    p.for_each_field([&](auto &v) { v.resize(n); });
    auto w = [&](int f) { return o + (size_t)f * stride; };
    par_for(n, threads, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            p.phase[i] = ((uint64_t)w(1)[i] << 32) | w(0)[i];
            p.step[i] = ((uint64_t)w(3)[i] << 32) | w(2)[i];
            p.smp_base[i] = w(4)[i];
            p.smp_len[i] = w(5)[i];
            p.loop_start[i] = w(6)[i];
            p.loop_end[i] = w(7)[i];
            p.flags[i] = w(8)[i];
            p.env_stage[i] = w(9)[i];
            p.env_level[i] = (R)bits_f32(w(10)[i]);
            p.gain_l[i] = bits_f32(w(11)[i]);
            p.gain_r[i] = bits_f32(w(12)[i]);
            p.filt_z1[i] = (R)bits_f32(w(13)[i]);
            p.filt_z2[i] = (R)bits_f32(w(14)[i]);
            p.params[i] = w(15)[i];
            p.region[i] = w(16)[i];
            p.gate_slot[i] = w(17)[i];
            p.ordinal[i] = w(18)[i];
            p.start_rel[i] = w(19)[i];
            p.note_id[i] = ((uint64_t)w(21)[i] << 32) | w(20)[i];
            p.born_variant[i] = w(22)[i];
            p.stop_rel[i] = w(23)[i];
        }
    });
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

uint64_t kor_fixed_from_f64(double v) {
    // src/fixed.rs:84-96
    double scaled = v * 4294967296.0;
    if (std::isnan(scaled) || scaled <= 0.0) return 0;
    if (scaled >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)scaled;
}

uint64_t kor_fixed_scale(uint64_t phase, uint32_t factor) { return fixed_scale(phase, factor); }

float kor_fixed_frac_f32(uint64_t phase) { return (float)(uint32_t)phase * FRAC_SCALE_F32; }

uint32_t kor_bend_factor(double semitones) {
    // src/fixed.rs:159-171
    if (!std::isfinite(semitones)) return BEND_ONE;
    double f = std::exp2(semitones / 12.0) * (double)BEND_ONE;
    if (f <= 0.0) return 1;
    if (f >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)std::round(f);
}

uint32_t kor_max_steal(const kor_config *c) { return max_steal(*c); }
uint32_t kor_pool_slots(const kor_config *c) {
    uint64_t v = (uint64_t)c->max_voices + max_steal(*c);  // saturating_add
    return v > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)v;
}
uint32_t kor_steal_span(const kor_config *c) { return steal_span(*c); }

uint32_t kor_admit_take(const kor_config *c, uint32_t live, uint32_t queued) {
    // src/config.rs:120-131
    uint32_t cap = c->max_voices;
    uint32_t want = std::min(queued, cap);
    if (live + want <= cap) return want;
    uint32_t need = 0;
    if (c->steal_rule != 1) need = std::min(std::min(live + want - cap, live), max_steal(*c));
    return std::min(want, cap - (live - need));
}

uint64_t kor_spawn_pick(uint64_t i, uint64_t want, uint64_t take) {
    return spawn_pick(i, want, take);
}

uint64_t kor_steal_key(uint64_t note_id, float level, int quietest) {
    if (!quietest) return note_id;
    float l = rs_clamp(level, 0.0f, 1.0f);
    uint64_t q = (uint64_t)rs_f32_to_u32(l * 65535.0f);
    return (q << 48) | (note_id & 0x0000FFFFFFFFFFFFull);
}

void kor_sort_key_plan(uint32_t n_regions, uint32_t max_len, uint32_t out[6]) {
    // src/gpu/mod.rs:76-105
    auto bitlen = [](uint32_t x) -> uint32_t { return x == 0 ? 0 : 32 - __builtin_clz(x); };
    uint32_t dead_region = n_regions > 1 ? n_regions : 1;
    uint32_t region_bits = bitlen(dead_region);
    uint32_t stage_bits = 3;
    uint32_t ml = max_len > 1 ? max_len : 1;
    uint32_t len_bits = bitlen(ml);
    uint32_t avail = 32 > region_bits + stage_bits ? 32 - (region_bits + stage_bits) : 0;
    uint32_t phase_bits = std::min(avail, len_bits);
    uint32_t phase_shift = len_bits > phase_bits ? len_bits - phase_bits : 0;
    uint32_t phase_mask = phase_bits >= 32 ? 0xFFFFFFFFu : ((1u << phase_bits) - 1);
    out[0] = region_bits + stage_bits + phase_bits;
    out[1] = phase_bits + stage_bits;
    out[2] = phase_bits;
    out[3] = phase_shift;
    out[4] = phase_mask;
    out[5] = dead_region;
}

void kor_biquad_lowpass(float fc, float q_cb, float sr, float out[4]) {
    // src/bank.rs:963-979
    const float Q_DEFAULT = 0.707106769f;  // std::f32::consts::FRAC_1_SQRT_2
    const float PI32 = 3.14159274f;        // std::f32::consts::PI
    float q_db = q_cb / 10.0f - 3.01f;
    float q = std::fmax(std::pow(10.0f, q_db / 20.0f), 0.001f);
    float gain = std::sqrt(Q_DEFAULT / q);
    float w0 = 2.0f * PI32 * rs_clamp(fc / sr, 1.0e-5f, 0.49f);
    float sin_w0 = std::sin(w0), cos_w0 = std::cos(w0);
    float alpha = sin_w0 / (2.0f * q);
    float a0 = 1.0f + alpha;
    out[0] = gain * (1.0f - cos_w0) * 0.5f / a0;
    out[1] = gain * (1.0f - cos_w0) / a0;
    out[2] = -2.0f * cos_w0 / a0;
    out[3] = (1.0f - alpha) / a0;
}

void kor_biquad_lowpass_f64(double fc, double q_cb, double sr, double out[4]) {
    // The same RBJ design carried out in f64 (precision form).
    const double Q_DEFAULT = 0.70710678118654752440;
    const double PI64 = 3.14159265358979323846;
    double q_db = q_cb / 10.0 - 3.01;
    double q = std::fmax(std::pow(10.0, q_db / 20.0), 0.001);
    double gain = std::sqrt(Q_DEFAULT / q);
    double w0 = 2.0 * PI64 * rs_clamp(fc / sr, 1.0e-5, 0.49);
    double sin_w0 = std::sin(w0), cos_w0 = std::cos(w0);
    double alpha = sin_w0 / (2.0 * q);
    double a0 = 1.0 + alpha;
    out[0] = gain * (1.0 - cos_w0) * 0.5 / a0;
    out[1] = gain * (1.0 - cos_w0) / a0;
    out[2] = -2.0 * cos_w0 / a0;
    out[3] = (1.0 - alpha) / a0;
}

void kor_make_params(float attack_s, float hold_s, float decay_s, float sustain_in,
                     float release_s, float fc_cents, float q_cb, float sr, int exp_decay,
                     int exp_release, float env_floor, int filter_enabled,
                     kor_region_params *out) {
    // src/bank.rs:876-954 with key scaling and controller modifiers neutral.
    float attack_frames = std::fmax(attack_s * sr, 0.0f);
    float hold_frames = std::fmax(hold_s * sr, 0.0f);
    float decay_frames = std::fmax(decay_s * sr, 1.0f);
    float release_frames = std::fmax(release_s * sr, 1.0f);
    float attack_rate = attack_frames < 1.0f ? 1.0f : 1.0f / attack_frames;
    float attack_end = 1.0f + hold_frames * attack_rate;
    float sustain = rs_clamp(sustain_in, 0.0f, 1.0f);
    float decay_coef = exp_decay ? std::pow(10.0f, -100.0f / (20.0f * decay_frames))
                                 : 1.0f / decay_frames;
    float release_coef = exp_release ? std::pow(10.0f, -100.0f / (20.0f * release_frames))
                                     : 1.0f / release_frames;
    float fc = 8.176f * std::pow(2.0f, fc_cents / 1200.0f);  // cents_to_hz
    float nyq_guard = sr * 0.49f;
    bool use_filter = filter_enabled && fc_cents < 13500.0f && fc < nyq_guard && fc > 20.0f;
    float bq[4] = {1.0f, 0.0f, 0.0f, 0.0f};
    if (use_filter) kor_biquad_lowpass(fc, q_cb, sr, bq);
    out->attack_rate = attack_rate;
    out->attack_end = attack_end;
    out->decay_coef = decay_coef;
    out->decay_target = std::fmax(sustain, env_floor);
    out->sustain = sustain;
    out->release_coef = release_coef;
    out->b0 = bq[0];
    out->b1 = bq[1];
    out->a1 = bq[2];
    out->a2 = bq[3];
    out->flags = use_filter ? RP_FILTER : 0;
    out->pad = 0;
}

void kor_make_params_f64(double attack_s, double hold_s, double decay_s, double sustain_in,
                         double release_s, double fc_cents, double q_cb, double sr,
                         int exp_decay, int exp_release, double env_floor, int filter_enabled,
                         kor_region_params64 *out) {
    double attack_frames = std::fmax(attack_s * sr, 0.0);
    double hold_frames = std::fmax(hold_s * sr, 0.0);
    double decay_frames = std::fmax(decay_s * sr, 1.0);
    double release_frames = std::fmax(release_s * sr, 1.0);
    double attack_rate = attack_frames < 1.0 ? 1.0 : 1.0 / attack_frames;
    double attack_end = 1.0 + hold_frames * attack_rate;
    double sustain = rs_clamp(sustain_in, 0.0, 1.0);
    double decay_coef = exp_decay ? std::pow(10.0, -100.0 / (20.0 * decay_frames))
                                  : 1.0 / decay_frames;
    double release_coef = exp_release ? std::pow(10.0, -100.0 / (20.0 * release_frames))
                                      : 1.0 / release_frames;
    double fc = 8.176 * std::pow(2.0, fc_cents / 1200.0);
    double nyq_guard = sr * 0.49;
    bool use_filter = filter_enabled && fc_cents < 13500.0 && fc < nyq_guard && fc > 20.0;
    double bq[4] = {1.0, 0.0, 0.0, 0.0};
    if (use_filter) kor_biquad_lowpass_f64(fc, q_cb, sr, bq);
    out->attack_rate = attack_rate;
    out->attack_end = attack_end;
    out->decay_coef = decay_coef;
    out->decay_target = std::fmax(sustain, env_floor);
    out->sustain = sustain;
    out->release_coef = release_coef;
    out->b0 = bq[0];
    out->b1 = bq[1];
    out->a1 = bq[2];
    out->a2 = bq[3];
    out->flags = use_filter ? RP_FILTER : 0;
    out->pad = 0;
}

kor_synth *kor_create(const kor_config *cfg, const int16_t *pool, size_t pool_len,
                      const kor_region_params *params, size_t n_params, int precision,
                      int state_only) {
    if (!cfg || cfg->gate_frames == 0 || cfg->block_frames % cfg->gate_frames != 0) return nullptr;
    kor_synth *s = new kor_synth();
    s->cfg = *cfg;
    s->precision = precision;
    s->state_only = state_only;
    s->bank_pool.assign(pool, pool + pool_len);
    s->bank_params.assign(params, params + n_params);
    s->tiles = cfg->block_frames / cfg->gate_frames;
    s->tile_frames = cfg->gate_frames;
    s->mix.assign((size_t)cfg->block_frames * 2, 0.0);
    s->gate_rows.assign(s->tiles * GATE_SLOTS, 0);
    s->chan_rows.assign(s->tiles * BEND_CHANNELS * CHAN_FIELDS, 0);
    return s;
}

void kor_destroy(kor_synth *s) { delete s; }

void kor_set_params64(kor_synth *s, const kor_region_params64 *params, size_t n) {
    s->bank_params64.assign(params, params + n);
}

int kor_set_gates(kor_synth *s, const uint32_t *rows, size_t n) {
    if (n != s->gate_rows.size()) return -1;
    std::memcpy(s->gate_rows.data(), rows, n * 4);
    return 0;
}

int kor_set_channels(kor_synth *s, const uint32_t *rows, size_t n, int bend, int gain,
                     int variant) {
    if (n != s->chan_rows.size()) return -1;
    std::memcpy(s->chan_rows.data(), rows, n * 4);
    s->bend_active = bend != 0;
    s->gain_active = gain != 0;
    s->variant_active = variant != 0;
    return 0;
}

int kor_set_params_variant(kor_synth *s, uint32_t index, const kor_region_params *data,
                           size_t n) {
    // src/cpu.rs:313-323
    if (index == 0) return 0;
    size_t i = index - 1;
    if (s->variants.size() <= i) s->variants.resize(i + 1);
    s->variants[i].assign(data, data + n);
    return 0;
}

int kor_spawn(kor_synth *s, const kor_spawn_cmd *cmds, size_t n) {
    if (s->precision) return Renderer<double>{*s, s->pool64}.spawn(cmds, n);
    return Renderer<float>{*s, s->pool32}.spawn(cmds, n);
}

int kor_render(kor_synth *s, float *out, size_t n) {
    if (s->precision) return render_block<double>(s, out, nullptr, n);
    return render_block<float>(s, out, nullptr, n);
}

int kor_render_f64(kor_synth *s, double *out, size_t n) {
    if (s->precision) return render_block<double>(s, nullptr, out, n);
    return render_block<float>(s, nullptr, out, n);
}

void kor_stats(const kor_synth *s, kor_block_stats *out) {
    out->active_voices = s->precision ? s->pool64.len() : s->pool32.len();
    out->stolen = s->stolen;
    out->dropped = s->dropped;
    out->peak = s->peak;
    out->pad = 0;
}

void kor_set_sorted(kor_synth *s, uint32_t n_regions, uint32_t max_len, int on) {
    s->sorted = on != 0;
    kor_sort_key_plan(n_regions, max_len, s->sort_layout);
}

void kor_set_threads(kor_synth *s, int n) { s->threads = n < 1 ? 1 : n; }

size_t kor_pool_len(const kor_synth *s) {
    return s->precision ? s->pool64.len() : s->pool32.len();
}

void kor_pool_export(const kor_synth *s, uint32_t *out, size_t stride) {
    if (s->precision)
        export_pool(s->pool64, out, stride, s->threads);
    else
        export_pool(s->pool32, out, stride, s->threads);
}

void kor_pool_import(kor_synth *s, const uint32_t *soa, size_t n, size_t stride) {
    if (s->precision)
        import_pool(s->pool64, soa, n, stride, s->threads);
    else
        import_pool(s->pool32, soa, n, stride, s->threads);
}

void kor_set_counters(kor_synth *s, uint64_t stolen, uint64_t dropped) {
    s->stolen = stolen;
    s->dropped = dropped;
}

}  // extern "C"
