// kst_render.hip -- per-voice block render and the mixdown.
//
// Per-voice semantics follow src/cpu.rs:403-613 (= shaders/render.wgsl:146-421)
// statement for statement, compiled with -ffp-contract=off so every f32
// operation rounds where the reference rounds; the 32.32 phase is a native
// uint64_t. Two forms:
//
//  * wave64 form (k_render_fast): one lane per voice, a persistent grid of
//    waves walking the pool in 64-voice batches. Each 16-frame sub-tile the
//    wave writes its 64 x 32 stereo contributions to LDS (row-major by output
//    sample) and reads them back transposed, so a lane sums one output
//    sample over 32 voices with plain adds and one cross-half exchange --
//    about two VALU ops per voice-frame for the whole mixdown. Sums are
//    accumulated per wave in LDS across batches (optionally as compensated
//    hi/lo pairs), combined per workgroup, and written once as one partial
//    row per workgroup.
//  * plain form (k_render_plain + k_reduce_plain): the reference's structure
//    -- reduce_tile frames per tile, an LDS staging array and a two-level tree
//    per tile, partials[j * nwg + wg] accumulated across grid-stride batches,
//    then one workgroup per output sample (shaders/render.wgsl:121-144,
//    shaders/reduce.wgsl).
//
// No float atomics anywhere; every summation order is fixed, so two runs are
// bitwise identical.
#include <hip/hip_ext.h>

#include "kst_device.h"

// The structured buffer load (LLVM's llvm.amdgcn.struct.buffer.load.i32, which
// this clang has no builtin for): the 4-byte word at descriptor base + vindex *
// stride + voffset + soffset, with idxen addressing. Declared at file scope so
// the name binds to the intrinsic.
typedef int kst_rsrc_t __attribute__((ext_vector_type(4)));
__device__ int kst_struct_load_i32(kst_rsrc_t rsrc, int vindex, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.struct.buffer.load.i32");

namespace kst {
namespace {

constexpr float INV_32767 = 1.0f / 32767.0f;       // unpack2x16snorm scale, src/cpu.rs:193
constexpr float FRAC_SCALE = 1.0f / 4294967296.0f;  // src/fixed.rs:73

struct Voice {
    uint64_t phase, step, base_step;
    uint32_t base, len, ls, le, vflags, stage, gate_slot, ordinal;
    uint32_t start_rel, stop_rel, params_base, variant, born_variant, region;
    float level, z1, z2, gl, gr, bgl, bgr;
    kst_region_params p;
    bool live;
};

__device__ __forceinline__ kst_region_params load_params(const Args &a, uint32_t variant,
                                                          uint32_t base) {
    // src/cpu.rs:228-238: a variant outside the table reads the bank's own.
    // Every variant slot is initialised with the bank's table, which covers a
    // variant that was never uploaded.
    // A params index past the table (the CPU reference would panic) reads
    // entry 0 rather than out of bounds.
    uint32_t v = variant < a.max_variants ? variant : 0;
    uint32_t i = base < a.ppv ? base : 0;
    return a.params[(size_t)v * a.ppv + i];
}

__device__ __forceinline__ void load_voice(const Args &a, const uint32_t *pool, uint32_t i,
                                           bool live, Voice &v) {
    const uint32_t c = a.cap;
    v.live = live;
    if (!live) {
        // A neutral lane: silent, and every address it could form is in bounds.
        v.phase = v.step = v.base_step = 0;
        v.base = 0;
        v.len = v.ls = v.le = 1;
        v.vflags = 0;
        v.stage = ENV_DEAD;
        v.gate_slot = v.ordinal = v.start_rel = v.stop_rel = 0;
        v.params_base = v.variant = v.born_variant = v.region = 0;
        v.level = v.z1 = v.z2 = v.gl = v.gr = v.bgl = v.bgr = 0.0f;
        v.p = kst_region_params{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        return;
    }
    v.phase = ((uint64_t)pool[F_PHASE_HI * c + i] << 32) | pool[F_PHASE_LO * c + i];
    v.base_step = ((uint64_t)pool[F_STEP_HI * c + i] << 32) | pool[F_STEP_LO * c + i];
    v.step = v.base_step;
    v.base = pool[F_SMP_BASE * c + i];
    v.len = pool[F_SMP_LEN * c + i];
    v.ls = pool[F_LOOP_START * c + i];
    v.le = pool[F_LOOP_END * c + i];
    v.vflags = pool[F_FLAGS * c + i];
    v.stage = pool[F_ENV_STAGE * c + i];
    v.level = __uint_as_float(pool[F_ENV_LEVEL * c + i]);
    v.bgl = __uint_as_float(pool[F_GAIN_L * c + i]);
    v.bgr = __uint_as_float(pool[F_GAIN_R * c + i]);
    v.gl = v.bgl;
    v.gr = v.bgr;
    v.z1 = __uint_as_float(pool[F_FILT_Z1 * c + i]);
    v.z2 = __uint_as_float(pool[F_FILT_Z2 * c + i]);
    v.params_base = pool[F_PARAMS * c + i];
    v.region = pool[F_REGION * c + i];
    v.gate_slot = pool[F_GATE_SLOT * c + i];
    v.ordinal = pool[F_ORDINAL * c + i];
    v.start_rel = pool[F_START_REL * c + i];
    v.born_variant = pool[F_BORN_VARIANT * c + i];
    v.stop_rel = pool[F_STOP_REL * c + i];
    v.variant = v.born_variant > 0 ? v.born_variant - 1 : 0;
    v.p = load_params(a, v.variant, v.params_base);
}

// Fused compaction: the fields a block never changes move with the voice to
// its slot in the other buffer (store_voice writes the rest).
__device__ __forceinline__ void copy_static(const Args &a, const uint32_t *src, uint32_t s,
                                            uint32_t *dst, uint32_t d, const Voice &v) {
    const uint32_t c = a.cap;
    dst[F_STEP_LO * c + d] = (uint32_t)v.base_step;
    dst[F_STEP_HI * c + d] = (uint32_t)(v.base_step >> 32);
    dst[F_SMP_BASE * c + d] = v.base;
    dst[F_SMP_LEN * c + d] = v.len;
    dst[F_LOOP_START * c + d] = v.ls;
    dst[F_LOOP_END * c + d] = v.le;
    dst[F_FLAGS * c + d] = v.vflags;
    dst[F_GAIN_L * c + d] = __float_as_uint(v.bgl);
    dst[F_GAIN_R * c + d] = __float_as_uint(v.bgr);
    dst[F_PARAMS * c + d] = v.params_base;
    dst[F_REGION * c + d] = v.region;
    dst[F_GATE_SLOT * c + d] = v.gate_slot;
    dst[F_ORDINAL * c + d] = v.ordinal;
    dst[F_NOTE_LO * c + d] = src[F_NOTE_LO * c + s];
    dst[F_NOTE_HI * c + d] = src[F_NOTE_HI * c + s];
}

// Write back what the block changed, clear the per-block fields
// (src/cpu.rs:605-612), and emit this voice's sort key (shaders/sort.wgsl:
// 246-262) or its dead bit when sorting is off.
__device__ __forceinline__ void store_voice(const Args &a, uint32_t *pool, uint32_t i,
                                            const Voice &v) {
    const uint32_t c = a.cap;
    pool[F_PHASE_LO * c + i] = (uint32_t)v.phase;
    pool[F_PHASE_HI * c + i] = (uint32_t)(v.phase >> 32);
    pool[F_ENV_STAGE * c + i] = v.stage;
    pool[F_ENV_LEVEL * c + i] = __float_as_uint(v.level);
    pool[F_FILT_Z1 * c + i] = __float_as_uint(v.z1);
    pool[F_FILT_Z2 * c + i] = __float_as_uint(v.z2);
    pool[F_START_REL * c + i] = 0;
    pool[F_BORN_VARIANT * c + i] = 0;
    pool[F_STOP_REL * c + i] = 0;
    uint32_t key;
    if (a.sort_on) {
        uint32_t region = a.dead_region;
        if (v.stage != ENV_DEAD) region = v.region < a.dead_region - 1 ? v.region : a.dead_region - 1;
        uint32_t ph = ((uint32_t)(v.phase >> 32) >> a.phase_shift) & a.phase_mask;
        uint32_t st = v.stage < 7 ? v.stage : 7;
        key = (region << a.region_shift) | (st << a.stage_shift) | ph;
    } else {
        key = v.stage == ENV_DEAD ? 1u : 0u;
    }
    a.keys[0][i] = key;
}

// 8.24 bend of a 32.32 step, truncating and saturating (src/fixed.rs:141-148).
__device__ __forceinline__ uint64_t fixed_scale(uint64_t phase, uint32_t factor) {
    uint64_t lo = phase * (uint64_t)factor;          // bits 0..63 of the product
    uint64_t hi = __umul64hi(phase, (uint64_t)factor);  // bits 64..127
    if (hi >> 24) return ~0ull;
    return (hi << 40) | (lo >> 24);
}

// Note-off gate and controller row for one gate tile (src/cpu.rs:449-487).
// Returns true when it changed the stage or the step (what the straight
// path's phase limit depends on).
template <bool CHAN>
__device__ __forceinline__ bool gate_tile(const Args &a, const uint32_t *gates,
                                          const uint32_t *chan, uint32_t chan_flags,
                                          uint32_t gt, Voice &v) {
    if (!v.live) return false;
    bool changed = false;
    // The slot is masked to the row (the driver never builds one past 2047).
    if (v.stage < ENV_RELEASE && gates[gt * GATE_SLOTS + (v.gate_slot & (GATE_SLOTS - 1))] >= v.ordinal) {
        v.stage = ENV_RELEASE;
        v.level = fminf(v.level, 1.0f);
        changed = true;
    }
    if (CHAN && chan_flags) {
        const uint32_t channel = (v.gate_slot >> 7) & 15;
        const uint32_t ci = (gt * BEND_CHANNELS + channel) * CHAN_FIELDS;
        if (chan_flags & CHAN_ACTIVE_BEND) {
            v.step = fixed_scale(v.base_step, chan[ci + CHAN_BEND]);
            changed = true;
        }
        if (chan_flags & CHAN_ACTIVE_GAIN) {
            v.gl = v.bgl * __uint_as_float(chan[ci + CHAN_GAIN_L]);
            v.gr = v.bgr * __uint_as_float(chan[ci + CHAN_GAIN_R]);
        }
        if (chan_flags & CHAN_ACTIVE_VARIANT) {
            const uint32_t want = chan[ci + CHAN_VARIANT];
            const uint32_t born_tile = v.start_rel / a.gate_frames;
            const bool born_here = v.born_variant != 0 && gt <= born_tile;
            if (want != v.variant && !born_here) {
                v.variant = want;
                v.p = load_params(a, v.variant, v.params_base);
            }
        }
    }
    return changed;
}

// Gate tiles whose rows one load round covers (the wave64 render).
constexpr uint32_t GATE_WINDOW = 16;

// The first gate tile in [g0, g0 + GATE_WINDOW) whose row releases this
// lane's note (gates[tile][slot] >= ordinal, src/cpu.rs:449-455), or
// g0 + GATE_WINDOW. Stages only advance, so a note is released at the first
// such tile if at all: once the stage reaches RELEASE later rows change
// nothing. The window's rows are loaded together, so a block of up to
// GATE_WINDOW gate tiles waits on its gate rows once.
__device__ __forceinline__ uint32_t gate_window(const uint32_t *gates, uint32_t g0,
                                               uint32_t tiles, const Voice &v) {
    const uint32_t slot = v.gate_slot & (GATE_SLOTS - 1);
    uint32_t row[GATE_WINDOW];
#pragma unroll
    for (uint32_t k = 0; k < GATE_WINDOW; ++k)
        row[k] = g0 + k < tiles ? gates[(g0 + k) * GATE_SLOTS + slot] : 0u;
    uint32_t first = g0 + GATE_WINDOW;
#pragma unroll
    for (uint32_t k = GATE_WINDOW; k-- > 0;)
        if (g0 + k < tiles && row[k] >= v.ordinal) first = g0 + k;
    return first;
}

// gate_tile with the note-off decided by gate_window: `release` is this
// lane's first releasing tile.
template <bool CHAN>
__device__ __forceinline__ bool gate_tile_at(const Args &a, const uint32_t *chan,
                                             uint32_t chan_flags, uint32_t gt, uint32_t release,
                                             Voice &v) {
    if (!v.live) return false;
    bool changed = false;
    if (v.stage < ENV_RELEASE && gt == release) {
        v.stage = ENV_RELEASE;
        v.level = fminf(v.level, 1.0f);
        changed = true;
    }
    if (CHAN && chan_flags) {
        const uint32_t channel = (v.gate_slot >> 7) & 15;
        const uint32_t ci = (gt * BEND_CHANNELS + channel) * CHAN_FIELDS;
        if (chan_flags & CHAN_ACTIVE_BEND) {
            v.step = fixed_scale(v.base_step, chan[ci + CHAN_BEND]);
            changed = true;
        }
        if (chan_flags & CHAN_ACTIVE_GAIN) {
            v.gl = v.bgl * __uint_as_float(chan[ci + CHAN_GAIN_L]);
            v.gr = v.bgr * __uint_as_float(chan[ci + CHAN_GAIN_R]);
        }
        if (chan_flags & CHAN_ACTIVE_VARIANT) {
            const uint32_t want = chan[ci + CHAN_VARIANT];
            const uint32_t born_tile = v.start_rel / a.gate_frames;
            const bool born_here = v.born_variant != 0 && gt <= born_tile;
            if (want != v.variant && !born_here) {
                v.variant = want;
                v.p = load_params(a, v.variant, v.params_base);
            }
        }
    }
    return changed;
}

__device__ __forceinline__ float fetch(const Args &a, uint32_t base, uint32_t idx) {
    // src/cpu.rs:190-194. u32 index arithmetic as in the reference; out of the
    // pool reads zero.
    const uint32_t i = base + idx;
    const int v = i < a.bank_len ? (int)a.bank[i] : 0;
    return fmaxf((float)v * INV_32767, -1.0f);
}

// src/cpu.rs:243-259
__device__ __forceinline__ uint32_t advance_index(uint32_t idx, int32_t off, bool looping,
                                                  uint32_t ls, uint32_t le, uint32_t len) {
    const int64_t raw = (int64_t)idx + off;
    if (looping) {
        const int64_t l0 = ls, l1 = le;
        if (raw >= l1) {
            int64_t span = l1 - l0;
            if (span < 1) span = 1;
            return (uint32_t)(l0 + (raw - l0) % span);
        }
        if (raw < 0) return 0;
        return (uint32_t)raw;
    }
    int64_t r = raw;
    const int64_t hi = (int64_t)len - 1;
    if (r < 0) r = 0;
    if (r > hi) r = hi;
    return (uint32_t)r;
}

// The Catmull-Rom coefficients of src/cpu.rs:297-302,
//   ca = -0.5 sm1 + 1.5 s0 - 1.5 s1 + 0.5 s2,
//   cb = sm1 - 2.5 s0 + 2 s1 - 0.5 s2,
//   cc = -0.5 sm1 + 0.5 s1,
// each summed left to right. A term scaled by 0.5 or 2 is exact (every sample
// is 0 or at least 1/32767 in magnitude and at most 1), so adding it is one
// fused multiply-add with the same single rounding: the same bits in 11
// operations instead of 16.
__device__ __forceinline__ void cubic_coefs(float sm1, float s0, float s1, float s2, float &ca,
                                            float &cb, float &cc) {
    ca = fmaf(0.5f, s2, fmaf(-0.5f, sm1, 1.5f * s0) - 1.5f * s1);
    cb = fmaf(-0.5f, s2, fmaf(2.0f, s1, sm1 - 2.5f * s0));
    cc = fmaf(-0.5f, sm1, 0.5f * s1);
}

template <int INTERP>
__device__ __forceinline__ float interpolate(const Args &a, const Voice &v, uint32_t idx,
                                             float frac, bool looping) {
    if (INTERP == 0) return fetch(a, v.base, idx);
    if (INTERP == 1) {
        const uint32_t i1 = advance_index(idx, 1, looping, v.ls, v.le, v.len);
        const float s0 = fetch(a, v.base, idx);
        const float s1 = fetch(a, v.base, i1);
        return s0 + (s1 - s0) * frac;
    }
    const uint32_t im1 = advance_index(idx, -1, looping, v.ls, v.le, v.len);
    const uint32_t i1 = advance_index(idx, 1, looping, v.ls, v.le, v.len);
    const uint32_t i2 = advance_index(idx, 2, looping, v.ls, v.le, v.len);
    const float sm1 = fetch(a, v.base, im1);
    const float s0 = fetch(a, v.base, idx);
    const float s1 = fetch(a, v.base, i1);
    const float s2 = fetch(a, v.base, i2);
    // src/cpu.rs:297-302, Horner form.
    float ca, cb, cc;
    cubic_coefs(sm1, s0, s1, s2, ca, cb, cc);
    return ((ca * frac + cb) * frac + cc) * frac + s0;
}

// One frame of one voice, src/cpu.rs:491-601. Returns the (y*gain_l,
// y*gain_r) contribution, or zeros when the voice is silent this frame.
template <int INTERP>
__device__ __forceinline__ void frame(const Args &a, uint32_t f, Voice &v, float &ol,
                                      float &orr) {
    ol = 0.0f;
    orr = 0.0f;
    if (!v.live || v.stage == ENV_DEAD || f < v.start_rel) return;
    const bool loop_en = (v.vflags & VF_LOOP) != 0;
    const bool lur = (v.vflags & VF_LOOP_UNTIL_RELEASE) != 0;
    const bool looping = loop_en && !(lur && v.stage >= ENV_RELEASE);
    const uint32_t idx = (uint32_t)(v.phase >> 32);
    if (!looping && idx >= v.len) {
        v.stage = ENV_DEAD;
        return;
    }
    // Envelope first, so an instant attack is at full level on frame 0.
    switch (v.stage) {
        case ENV_ATTACK:
            v.level = v.level + v.p.attack_rate;
            if (v.level >= v.p.attack_end) {
                v.stage = ENV_DECAY;
                v.level = 1.0f;
            }
            break;
        case ENV_DECAY:
            v.level = a.exp_decay ? v.level * v.p.decay_coef : v.level - v.p.decay_coef;
            if (v.level <= v.p.decay_target) {
                if (v.p.sustain <= a.env_floor) {
                    v.stage = ENV_DEAD;
                    v.level = 0.0f;
                } else {
                    v.stage = ENV_SUSTAIN;
                    v.level = v.p.sustain;
                }
            }
            break;
        case ENV_RELEASE:
            v.level = a.exp_release ? v.level * v.p.release_coef : v.level - v.p.release_coef;
            if (v.level <= a.env_floor) {
                v.stage = ENV_DEAD;
                v.level = 0.0f;
            }
            break;
        default:
            break;
    }
    const float frac = (float)(uint32_t)v.phase * FRAC_SCALE;
    const float s = interpolate<INTERP>(a, v, idx, frac, looping);
    const float g = fminf(v.level, 1.0f);
    const float x = s * g;
    float y = x;
    if (v.p.flags & RP_FILTER) {
        // Transposed direct form II, b2 == b0.
        y = v.p.b0 * x + v.z1;
        v.z1 = v.p.b1 * x - v.p.a1 * y + v.z2;
        v.z2 = v.p.b0 * x - v.p.a2 * y;
    }
    // A stolen voice fades over steal_fade_frames from its own stop frame,
    // after the filter.
    if (v.stop_rel != 0 && f + 1 >= v.stop_rel) {
        const uint32_t d = f + 1 - v.stop_rel;
        if (d >= a.fade) {
            v.stage = ENV_DEAD;
            y = 0.0f;
        } else {
            y = y * (1.0f - (float)d / (float)a.fade);
        }
    }
    ol = y * v.gl;
    orr = y * v.gr;
    v.phase += v.step;
    if (looping) {
        const uint32_t hi = (uint32_t)(v.phase >> 32);
        if (hi >= v.le) {
            uint32_t span = v.le - v.ls;
            if (span < 1) span = 1;
            const uint32_t wrapped = v.ls + (hi - v.ls) % span;
            v.phase = ((uint64_t)wrapped << 32) | (v.phase & 0xFFFFFFFFull);
        }
    }
}


// ---------------------------------------------------------------------------
// wave64 fast path: a 16-frame sub-tile in which no lane can reach a loop end,
// a sample end, its start offset or a steal fade. Then the per-frame work is
// straight-line: the envelope is one fused multiply-add per frame -- exact,
// because in every stage one of its two operations is exact (x*1 or +(-0)) --
// with a compare that sends the rare stage change to a uniform fix-up branch,
// and the sample taps are plain neighbours.
// ---------------------------------------------------------------------------

constexpr float F_INF = __builtin_huge_valf();

// level' = fmaf(level, m, c), stage change when level' >= thi or level' <= tlo
// (src/cpu.rs:512-549).
__device__ __forceinline__ void env_coefs(const Args &a, const Voice &v, float &m, float &c,
                                          float &thi, float &tlo) {
    m = 1.0f;
    c = -0.0f;
    thi = F_INF;
    tlo = -F_INF;
    switch (v.stage) {
        case ENV_ATTACK:
            c = v.p.attack_rate;
            thi = v.p.attack_end;
            break;
        case ENV_DECAY:
            if (a.exp_decay) m = v.p.decay_coef; else c = -v.p.decay_coef;
            tlo = v.p.decay_target;
            break;
        case ENV_RELEASE:
            if (a.exp_release) m = v.p.release_coef; else c = -v.p.release_coef;
            tlo = a.env_floor;
            break;
        default:
            break;
    }
}

// The stage change of src/cpu.rs:512-549 for a lane whose compare fired.
__device__ __forceinline__ void env_transition(const Args &a, Voice &v) {
    if (v.stage == ENV_ATTACK) {
        v.stage = ENV_DECAY;
        v.level = 1.0f;
    } else if (v.stage == ENV_DECAY) {
        if (v.p.sustain <= a.env_floor) {
            v.stage = ENV_DEAD;
            v.level = 0.0f;
        } else {
            v.stage = ENV_SUSTAIN;
            v.level = v.p.sustain;
        }
    } else if (v.stage == ENV_RELEASE) {
        v.stage = ENV_DEAD;
        v.level = 0.0f;
    }
}

// The straight path for frames [f0, f0 + SUBTILE) needs every tap of the
// sub-tile inside the sample (before the loop end when looping) and inside the
// bank: phase + SUBTILE * step < limit. The limit depends on the stage (loop
// until release), the step (overflow guard) and fixed voice fields, so it is
// recomputed only when a gate tile changes one of those.
template <int INTERP>
__device__ __forceinline__ uint64_t fast_limit(const Args &a, const Voice &v) {
    if ((v.step >> 32) >= (1u << 26)) return 0;  // SUBTILE steps could overflow
    const bool loop_en = (v.vflags & VF_LOOP) != 0;
    const bool lur = (v.vflags & VF_LOOP_UNTIL_RELEASE) != 0;
    const bool looping = loop_en && !(lur && v.stage >= ENV_RELEASE);
    const uint64_t reach = INTERP == 2 ? 2u : (INTERP == 1 ? 1u : 0u);
    // A branch-free pick: a plain select of the two fields let the compiler
    // address the voice struct in scratch memory.
    const uint32_t pick = 0u - (uint32_t)looping;
    const uint64_t limit = v.len ^ ((v.le ^ v.len) & pick);
    // The in-sample index also stays below 2^24, so the tap offset is one
    // 24-bit multiply-add (longer samples take the edge path there).
    const uint64_t lc = limit < (1u << 24) ? limit : (1u << 24);
    const uint64_t l1 = lc > reach ? lc - reach : 0;
    const uint64_t room = (uint64_t)a.bank_len;
    const uint64_t l2 = room > (uint64_t)v.base + reach ? room - v.base - reach : 0;
    return (l1 < l2 ? l1 : l2) << 32;  // < 2^63: a phase at or past 2^63 never passes
}

// Can this lane take the straight-line path for frames [f0, f0 + SUBTILE)?
// Silent lanes (no voice, or DEAD) always can: they are neutralised.
// A steal fade in the sub-tile is checked apart (fade_in): the straight path
// has a variant for it.
template <int INTERP>
__device__ __forceinline__ bool fast_ok(const Voice &v, uint64_t lim, uint32_t f0) {
    if (!v.live || v.stage == ENV_DEAD) return true;
    if (v.start_rel > f0) return false;
    if (INTERP == 2 && (v.phase >> 32) == 0) return false;  // tap idx-1
    // phase + SUBTILE * step < lim, without wrapping (step < 2^58).
    return v.phase < lim && v.step * SUBTILE < lim - v.phase;
}

// The bank as a raw buffer resource: a tap load takes its 32-bit byte offset
// straight from a VGPR (no 64-bit address arithmetic per load), and a load
// past the padded bank reads 0 instead of faulting. The host keeps the pool
// under 2^31 frames, so every byte offset fits 32 bits. Dword 3 = 0x00020000:
// 32-bit data format, the CDNA3/CDNA4 untyped-buffer setting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bank_rsrc(const Args &a) {
    const uint64_t nb = ((uint64_t)a.bank_len + 8) * 2;  // the pool plus its guard frames
    const uint32_t n = nb > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)nb;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<int16_t *>(a.bank), (short)0, (int)n,
                                             0x00020000);
}

// The bank as a structured buffer of 2-byte records: a tap load's index is the
// frame number itself, so a phase chain that carries the sample's base in its
// high word gives each tap's address with no arithmetic of its own. Dword 1
// holds the address's high bits and the stride (bits 16-29); dword 2 the record
// count (the pool and its guard frames). readfirstlane keeps the descriptor in
// scalar registers (a descriptor the compiler places in vector registers turns
// every load into a loop over lanes).
__device__ __forceinline__ kst_rsrc_t bank_records(const Args &a) {
    const uint64_t nr = (uint64_t)a.bank_len + 8;
    const uint64_t ba = (uint64_t)a.bank;
    kst_rsrc_t r;
    r.x = __builtin_amdgcn_readfirstlane((int)(uint32_t)ba);
    r.y = __builtin_amdgcn_readfirstlane((int)(((uint32_t)(ba >> 32) & 0xFFFFu) | (2u << 16)));
    r.z = __builtin_amdgcn_readfirstlane((int)(nr > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)nr));
    r.w = 0x00020000;
    return r;
}

// One 32-bit word holding the int16 frames at byte offsets off and off + 2
// (gfx950 buffer loads take 2-byte-aligned offsets).
__device__ __forceinline__ uint32_t tap_pair(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0);
}

// int16 -> f32 as src/cpu.rs:190-194 does it: times 1/32767, clamped at -1.
// Without a -32768 in the bank the clamp cannot change anything.
template <bool CLAMP>
__device__ __forceinline__ float norm_lo(uint32_t w) {
    const float f = (float)(int)(int16_t)(w & 0xFFFFu) * INV_32767;
    return CLAMP ? fmaxf(f, -1.0f) : f;
}
template <bool CLAMP>
__device__ __forceinline__ float norm_hi(uint32_t w) {
    const float f = (float)((int)w >> 16) * INV_32767;
    return CLAMP ? fmaxf(f, -1.0f) : f;
}

// Does this lane's steal fade (src/cpu.rs:577-586) reach frames [f0, f0 + SUBTILE)?
__device__ __forceinline__ bool fade_in(const Voice &v, uint32_t f0) {
    return v.live && v.stage != ENV_DEAD && v.stop_rel != 0 && v.stop_rel <= f0 + SUBTILE;
}

// FILT: 0 every sounding lane unfiltered, 1 every sounding lane filtered,
// 2 mixed. ENV_STATIC: every sounding lane sustains through the sub-tile.
// FADE: some lane's steal fade reaches the sub-tile. Frame f of a lane with
// d = f + 1 - stop_rel >= 0 is scaled by the fade table's 1 - d / fade, and
// from d == fade on it is silent and the voice dies -- what the edge path and
// the per-frame path do, so all three give the same bits.
// WRAP (linear and nearest): some looping lane reaches its loop end in the
// sub-tile. Every frame's phase wraps back by one loop span once its index
// reaches the loop end, and a linear tap after the loop end is the loop start
// (src/cpu.rs:243-259), as the edge path does it.
// A lane with no voice, or a dead one, is neutralised for the straight path:
// silent, and every tap it loads is the bank's first.
template <int INTERP>
__device__ __forceinline__ void neutralise(Voice &v) {
    if (!v.live || v.stage == ENV_DEAD) {
        v.gl = v.gr = 0.0f;
        v.step = 0;
        v.phase = (uint64_t)(INTERP == 2 ? 1u : 0u) << 32;
        v.base = 0;
    }
}

// Every frame's taps first: the phase chain is integer arithmetic, so all of
// the sub-tile's loads are in flight before the first is needed. One copy of
// this loop serves every straight-path variant (the dispatch loads the taps
// and then picks the variant), so a sub-tile issues its loads once.
template <int INTERP>
__device__ __forceinline__ void straight_taps(const Args &a, const Voice &v,
                                              uint32_t (&w0)[SUBTILE],
                                              uint32_t (&w1)[INTERP == 2 ? SUBTILE : 1],
                                              uint32_t (&lo)[SUBTILE]) {
    // The chain runs phase + (base << 32) (cubic: base - 1, its first tap
    // being idx - 1), whose high word is the tap's frame in the pool and whose
    // low word is the phase's own (lo, the frame loop's fraction for nearest
    // and linear). A cubic frame's second word, frames idx + 1 and idx + 2, is
    // 4 bytes on: the instruction's own offset.
    const kst_rsrc_t bank = bank_records(a);
    uint64_t pb = v.phase + ((uint64_t)(v.base - (INTERP == 2 ? 1u : 0u)) << 32);
#pragma unroll
    for (uint32_t i = 0; i < SUBTILE; ++i) {
        const int r = (int)(uint32_t)(pb >> 32);
        w0[i] = (uint32_t)kst_struct_load_i32(bank, r, 0, 0, 0);
        if (INTERP == 2) w1[i] = (uint32_t)kst_struct_load_i32(bank, r, 4, 0, 0);
        lo[i] = (uint32_t)pb;
        pb += v.step;
    }
}

// The WRAP variant's taps: each frame's phase wraps back one loop span once
// its index reaches the loop end, and a tap past the loop end reads from the
// loop start (src/cpu.rs:243-259): for linear the tap after idx = le - 1 is
// ls; for cubic idx + 1 and idx + 2 wrap the same way (ls, then ls + 1, or ls
// again for a one-frame loop) while idx - 1 never does. Returns the phase after
// the sub-tile.
template <int INTERP>
__device__ __forceinline__ uint64_t wrap_taps(const Args &a, const Voice &v,
                                              uint32_t (&w0)[SUBTILE],
                                              uint32_t (&w1)[INTERP == 2 ? SUBTILE : 1]) {
    const __amdgpu_buffer_rsrc_t bank = bank_rsrc(a);
    const bool sounding = v.live && v.stage != ENV_DEAD;
    const bool looping = sounding && (v.vflags & VF_LOOP) != 0 &&
                         !((v.vflags & VF_LOOP_UNTIL_RELEASE) != 0 && v.stage >= ENV_RELEASE);
    const uint32_t le = looping ? v.le : 0xFFFFFFFFu;  // never reached when not looping
    const uint32_t span = v.le - v.ls;
    const uint32_t base2 = v.base * 2u;
    // Frames ls and ls + 1 of a looping lane (ls twice for a one-frame loop).
    uint32_t ls_pair = 0;
    if (looping) {
        const uint32_t p = INTERP == 2 ? tap_pair(bank, base2 + v.ls * 2u)
                                       : (uint32_t)(uint16_t)a.bank[v.base + v.ls];
        ls_pair = INTERP == 2 && span > 1 ? p : (p & 0xFFFFu) * 0x10001u;
    }
    uint64_t ph = v.phase;
#pragma unroll
    for (uint32_t i = 0; i < SUBTILE; ++i) {
        const uint32_t idx = (uint32_t)(ph >> 32);
        const uint32_t off = __umul24(idx, a.tap_bytes) + base2;
        if (INTERP == 2) {
            w0[i] = tap_pair(bank, off - 2);  // frames idx-1, idx (never past the loop end)
            const uint32_t w = tap_pair(bank, off + 2);  // frames idx+1, idx+2
            w1[i] = idx + 2 < le ? w : (idx + 2 == le ? (w & 0xFFFFu) | (ls_pair << 16) : ls_pair);
        } else {
            const uint32_t w = tap_pair(bank, off);
            w0[i] = INTERP == 1 && idx + 1 == le ? (w & 0xFFFFu) | (ls_pair << 16) : w;
        }
        ph += v.step;
        if ((uint32_t)(ph >> 32) >= le) ph -= (uint64_t)span << 32;
    }
    return ph;
}

// The straight path's frame loop over taps already loaded (neutralised lanes).
// lo: nearest and linear take each frame's phase low word from the tap chain.
// WRAP: the phase after the sub-tile is `wrapped_end` (see wrap_taps).
template <int INTERP, bool CLAMP, int FILT, bool ENV_STATIC, bool FADE, bool WRAP = false>
__device__ __forceinline__ void fast_subtile(const Args &a, Voice &v, float *tb,
                                             const float *ftab, uint32_t lane, uint32_t f0,
                                             const uint32_t (&w0)[SUBTILE],
                                             const uint32_t (&w1)[INTERP == 2 ? SUBTILE : 1],
                                             const uint32_t *lo, uint64_t wrapped_end = 0) {
    float m, c, thi, tlo;
    env_coefs(a, v, m, c, thi, tlo);
    const bool filt = FILT == 1 || (FILT == 2 && (v.p.flags & RP_FILTER));
    float g = fminf(v.level, 1.0f);
    bool kill = false;
    // The steal fade: frame f0 + i is scaled by ftab[d0 + i], d0 = f0 + 1 -
    // stop_rel, where the table (FADE_PAD entries of 1 before it, of 0 after
    // it) gives 1 before the fade, 1 - d / fade during it and 0 from d = fade
    // on, the frame that kills the voice. A lane with no fade reads ones.
    const float *gtab = ftab;
    bool died = false;
    if (FADE) {
        const int32_t fade = (int32_t)a.fade;
        const int32_t pad = (int32_t)FADE_PAD;
        const int32_t d0 = v.stop_rel != 0 ? (int32_t)(f0 + 1) - (int32_t)v.stop_rel : -pad;
        gtab = ftab + (d0 < -pad ? -pad : (d0 > fade ? fade : d0));
        died = v.stop_rel != 0 && d0 + (int32_t)SUBTILE - 1 >= fade;
    }
    // Cubic (and the wrap variant) runs its own 32-bit chain of the phase's
    // low word.
    uint32_t lo32 = (uint32_t)v.phase;
#pragma unroll
    for (uint32_t i = 0; i < SUBTILE; ++i) {
        if (!ENV_STATIC) {
            if (kill) {
                // A lane that died on the previous frame is silent from here.
                if (v.stage == ENV_DEAD) v.gl = v.gr = 0.0f;
                kill = false;
            }
            v.level = fmaf(v.level, m, c);
            const bool tr = (v.level >= thi) | (v.level <= tlo);
            if (__builtin_expect(__any(tr), 0)) {
                if (tr) {
                    env_transition(a, v);
                    env_coefs(a, v, m, c, thi, tlo);
                }
                kill = __any(tr && v.stage == ENV_DEAD);
            }
            g = fminf(v.level, 1.0f);
        }
        const float flo = (float)(INTERP == 2 || WRAP ? lo32 : lo[i]);
        float s;
        if (INTERP == 0) {
            s = norm_lo<CLAMP>(w0[i]);
        } else if (INTERP == 1) {
            const float s0 = norm_lo<CLAMP>(w0[i]);
            const float s1 = norm_hi<CLAMP>(w0[i]);
            // s0 + (s1 - s0) * (flo * 2^-32): the power-of-two scale commutes
            // with rounding, so one fused step gives the same bits.
            s = fmaf((s1 - s0) * flo, FRAC_SCALE, s0);
        } else {
            const float sm1 = norm_lo<CLAMP>(w0[i]);
            const float s0 = norm_hi<CLAMP>(w0[i]);
            const float s1 = norm_lo<CLAMP>(w1[i]);
            const float s2 = norm_hi<CLAMP>(w1[i]);
            float ca, cb, cc;
            cubic_coefs(sm1, s0, s1, s2, ca, cb, cc);
            // ((ca * frac + cb) * frac + cc) * frac + s0 with frac = flo * 2^-32:
            // each product by frac is the product by flo scaled by 2^-32, which
            // commutes with rounding, so each step is one fused multiply-add of
            // that product.
            s = fmaf(fmaf(fmaf(ca * flo, FRAC_SCALE, cb) * flo, FRAC_SCALE, cc) * flo, FRAC_SCALE,
                     s0);
        }
        const float x = s * g;
        float y = x;
        if (FILT == 1 || (FILT == 2 && filt)) {
            const float b0x = v.p.b0 * x;
            y = b0x + v.z1;
            v.z1 = v.p.b1 * x - v.p.a1 * y + v.z2;
            v.z2 = b0x - v.p.a2 * y;
        }
        if (FADE) {
            // y * 1 is y itself; a zero factor marks the frame that kills the
            // voice, which sounds nothing (src/cpu.rs:577-586).
            const float fg = gtab[i];
            const bool gone = fg == 0.0f;
            y = y * fg;
            tb[(2 * i) * TROW + lane] = gone ? 0.0f : y * v.gl;
            tb[(2 * i + 1) * TROW + lane] = gone ? 0.0f : y * v.gr;
        } else {
            tb[(2 * i) * TROW + lane] = y * v.gl;
            tb[(2 * i + 1) * TROW + lane] = y * v.gr;
        }
        if (INTERP == 2 || WRAP)
            lo32 += (uint32_t)v.step;
    }
    if (WRAP)
        v.phase = wrapped_end;
    else
        v.phase += v.step * SUBTILE;  // the same sum, mod 2^64
    if (!ENV_STATIC && kill && v.stage == ENV_DEAD) v.gl = v.gr = 0.0f;
    if (FADE && died) {
        v.stage = ENV_DEAD;
        v.level = 0.0f;
        v.gl = v.gr = 0.0f;
    }
}

template <int INTERP, bool CLAMP, bool FADE>
__device__ __forceinline__ void fast_subtile_dispatch(const Args &a, Voice &v, float *tb,
                                                      const float *ftab, uint32_t lane,
                                                      uint32_t f0) {
    const bool sounding = v.live && v.stage != ENV_DEAD;
    const bool env_static = !__any(sounding && v.stage != ENV_SUSTAIN);
    const bool f = (v.p.flags & RP_FILTER) != 0;
    const bool all_filtered = !__any(sounding && !f);
    const bool none_filtered = !__any(sounding && f);
    neutralise<INTERP>(v);
    uint32_t w0[SUBTILE], w1[INTERP == 2 ? SUBTILE : 1], lo[SUBTILE];
    straight_taps<INTERP>(a, v, w0, w1, lo);
    if (all_filtered) {
        if (env_static)
            fast_subtile<INTERP, CLAMP, 1, true, FADE>(a, v, tb, ftab, lane, f0, w0, w1, lo);
        else
            fast_subtile<INTERP, CLAMP, 1, false, FADE>(a, v, tb, ftab, lane, f0, w0, w1, lo);
    } else if (none_filtered) {
        fast_subtile<INTERP, CLAMP, 0, false, FADE>(a, v, tb, ftab, lane, f0, w0, w1, lo);
    } else {
        fast_subtile<INTERP, CLAMP, 2, false, FADE>(a, v, tb, ftab, lane, f0, w0, w1, lo);
    }
}

// ---------------------------------------------------------------------------
// wave64 edge path: a sub-tile in which some lane starts late, fades out after
// a steal, wraps at its loop end or runs off its sample end. Each of these is a
// per-lane predicate on the straight-line frame:
//  * frames before start_rel change nothing and sound nothing;
//  * the steal fade multiplies by 1 - d / fade from a table computed with the
//    same division, and the frame at d == fade is silent and kills the voice;
//  * the phase wraps by one loop span (the step is shorter than the span), and
//    the tap after the loop end is the loop start, as src/cpu.rs:243-259 has it;
//  * a non-looping voice dies on the first frame whose index is past its
//    sample, and its last frame's second tap is the last sample itself.
// The phase chain and the taps are computed first, as in the straight path;
// the per-frame loop only needs the phase's low word.
// ---------------------------------------------------------------------------

template <int INTERP>
__device__ __forceinline__ bool edge_ok(const Args &a, const Voice &v, uint32_t f0) {
    if (!v.live || v.stage == ENV_DEAD) return true;
    const uint32_t hi = (uint32_t)(v.phase >> 32);
    if ((v.step >> 32) >= (1u << 26) || hi >= 0x80000000u) return false;
    const bool loop_en = (v.vflags & VF_LOOP) != 0;
    const bool lur = (v.vflags & VF_LOOP_UNTIL_RELEASE) != 0;
    const bool looping = loop_en && !(lur && v.stage >= ENV_RELEASE);
    if (looping && (v.ls >= v.le || hi >= v.le || (uint32_t)(v.step >> 32) >= v.le - v.ls))
        return false;
    // Taps reach index le (looping) or len (not looping); all inside the bank.
    if ((uint64_t)v.base + (looping ? v.le : v.len) >= a.bank_len) return false;
    // A fade this sub-tile needs the staged table.
    if (v.stop_rel != 0 && v.stop_rel <= f0 + SUBTILE && a.fade_lds == 0) return false;
    return true;
}

// Can this lane take the straight path's WRAP variant for frames [f0, f0 +
// SUBTILE)? A started, looping lane whose loop is longer than its step (one
// wrap per frame at most), with the tap offsets inside the 24-bit range and the
// bank. Other lanes of the wave must pass fast_ok.
template <int INTERP>
__device__ __forceinline__ bool wrap_ok(const Args &a, const Voice &v, uint32_t f0) {
    if (!v.live || v.stage == ENV_DEAD) return true;
    if (v.start_rel > f0) return false;
    const bool looping = (v.vflags & VF_LOOP) != 0 &&
                         !((v.vflags & VF_LOOP_UNTIL_RELEASE) != 0 && v.stage >= ENV_RELEASE);
    if (!looping) return false;
    const uint32_t hi = (uint32_t)(v.phase >> 32);
    if (v.ls >= v.le || hi >= v.le || (uint32_t)(v.step >> 32) >= v.le - v.ls) return false;
    // Cubic's tap idx - 1 stays a plain read: the index never reaches 0.
    if (INTERP == 2 && (hi == 0 || v.ls == 0)) return false;
    if (v.le >= (1u << 24)) return false;
    return (uint64_t)v.base + v.le < a.bank_len;
}

template <int INTERP, bool CLAMP>
__device__ __forceinline__ void fast_subtile_edge(const Args &a, Voice &v, float *tb,
                                                  const float *ftab, uint32_t lane,
                                                  uint32_t f0) {
    const bool sounding = v.live && v.stage != ENV_DEAD;
    if (!sounding) {
        v.gl = v.gr = 0.0f;
        v.step = 0;
        v.phase = 0;
        v.base = 0;
    }
    const bool loop_en = (v.vflags & VF_LOOP) != 0;
    const bool lur = (v.vflags & VF_LOOP_UNTIL_RELEASE) != 0;
    const bool looping = sounding && loop_en && !(lur && v.stage >= ENV_RELEASE);
    const uint64_t span_hi = (uint64_t)(v.le - v.ls) << 32;
    const __amdgpu_buffer_rsrc_t bank = bank_rsrc(a);
    const uint32_t base2 = v.base * 2u;
    // Frames [f0 + i] that sound: started, not past the sample, not past the fade.
    uint32_t amask = 0;
    uint32_t w0[SUBTILE];
    uint64_t ph = v.phase;
#pragma unroll
    for (uint32_t i = 0; i < SUBTILE; ++i) {
        const uint32_t f = f0 + i;
        const uint32_t idx = (uint32_t)(ph >> 32);
        const bool started = f >= v.start_rel;
        const bool past_end = !looping && idx >= v.len;
        const bool faded = v.stop_rel != 0 && f + 1 >= v.stop_rel + a.fade;
        const bool act = sounding && started && !past_end && !faded;
        uint32_t w = tap_pair(bank, act ? base2 + idx * 2u : 0u);
        if (INTERP == 1 && act && (looping ? idx + 1 == v.le : idx + 1 == v.len)) {
            const uint32_t j = looping ? v.ls : idx;
            w = (w & 0xFFFFu) | ((uint32_t)(uint16_t)a.bank[v.base + j] << 16);
        }
        w0[i] = w;
        amask |= act ? 1u << i : 0u;
        if (started) {
            ph += v.step;
            if (looping && (uint32_t)(ph >> 32) >= v.le) ph -= span_hi;
        }
    }
    float m, c, thi, tlo;
    env_coefs(a, v, m, c, thi, tlo);
    const bool filt = (v.p.flags & RP_FILTER) != 0;
    uint32_t lo = (uint32_t)v.phase;
    const uint32_t slo = (uint32_t)v.step;
    bool kill = false;
    bool died = false;
#pragma unroll
    for (uint32_t i = 0; i < SUBTILE; ++i) {
        const uint32_t f = f0 + i;
        const bool act = (amask >> i) & 1u;
        if (kill) {
            if (v.stage == ENV_DEAD) v.gl = v.gr = 0.0f;
            kill = false;
        }
        const float ln = fmaf(v.level, m, c);
        const bool tr = act && ((ln >= thi) | (ln <= tlo));
        v.level = act ? ln : v.level;
        if (__builtin_expect(__any(tr), 0)) {
            if (tr) {
                env_transition(a, v);
                env_coefs(a, v, m, c, thi, tlo);
            }
            kill = __any(tr && v.stage == ENV_DEAD);
        }
        const float g = fminf(v.level, 1.0f);
        const float flo = (float)lo;
        float s;
        if (INTERP == 0) {
            s = norm_lo<CLAMP>(w0[i]);
        } else {
            const float s0 = norm_lo<CLAMP>(w0[i]);
            const float s1 = norm_hi<CLAMP>(w0[i]);
            s = fmaf((s1 - s0) * flo, FRAC_SCALE, s0);
        }
        const float x = s * g;
        const float b0x = v.p.b0 * x;
        const float yf = b0x + v.z1;
        const float z1n = v.p.b1 * x - v.p.a1 * yf + v.z2;
        const float z2n = b0x - v.p.a2 * yf;
        const bool upd = act && filt;
        v.z1 = upd ? z1n : v.z1;
        v.z2 = upd ? z2n : v.z2;
        float y = filt ? yf : x;
        const bool fading = v.stop_rel != 0 && f + 1 >= v.stop_rel;
        if (fading && act) y = y * ftab[f + 1 - v.stop_rel];
        tb[(2 * i) * TROW + lane] = act ? y * v.gl : 0.0f;
        tb[(2 * i + 1) * TROW + lane] = act ? y * v.gr : 0.0f;
        lo += (f >= v.start_rel) ? slo : 0u;
        // Past the sample or through the fade: dead from here on.
        died |= sounding && f >= v.start_rel && !act;
    }
    if (kill && v.stage == ENV_DEAD) v.gl = v.gr = 0.0f;
    v.phase = ph;
    if (died) {
        v.stage = ENV_DEAD;
        v.level = 0.0f;
        v.gl = v.gr = 0.0f;
    }
}

// Error-free transformation of a + b (Knuth's two-sum).
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    const float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

// ---------------------------------------------------------------------------
// wave64 form
// ---------------------------------------------------------------------------

// One wave renders one batch of 64 voices for the whole block. Each 16-frame
// sub-tile it reduces its 64 x 32 contributions through LDS and writes the 32
// sums into its own row of the partial buffer (row = batch), so no two waves
// share an accumulator and the summation order never depends on scheduling,
// nor on which wave renders which batch.
template <int INTERP, bool CHAN, bool CLAMP>
__device__ __forceinline__ void render_batch(const Args &a, uint32_t batch, uint32_t live_total,
                                             float *tb, const float *ftab, uint32_t lane) {
    DevState *st = a.st;
    const uint32_t F = a.F;
    const uint32_t two_f = 2 * F;
    const uint32_t b = st->block;
    const uint32_t chan_flags = CHAN ? st->chan_flags : 0u;
    uint32_t *pool = sel2(a.pools, st->parity);
    const uint32_t *gates = a.gates + (size_t)b * a.gate_words;
    const uint32_t *chan = a.chans + (size_t)b * a.chan_words;
    float *prow = a.partials + (size_t)batch * two_f;

    const uint32_t half = lane >> 5;  // which 32 voices this lane sums over
    const uint32_t o = lane & 31;     // which output of the sub-tile it owns
    const uint32_t vi = batch * 64 + lane;
    const bool in = vi < live_total;
    // Fused: voice vi of this block is survivor perm[vi] of the live pool
    // (vi < live) or this block's spawn, which k_spawn wrote to the other
    // buffer at vi; either way the voice leaves in the other buffer at vi.
    uint32_t *dst = a.fused ? sel2(a.pools, st->parity ^ 1) : pool;
    const bool survivor = !a.fused || vi < st->live;
    const uint32_t *perm = a.fused ? live_perm(a) : nullptr;
    const uint32_t *src = survivor ? pool : dst;
    const uint32_t slot = (survivor && perm && in) ? perm[vi] : vi;
    Voice v;
    load_voice(a, src, slot, in, v);
    if (a.fused && in && survivor) copy_static(a, src, slot, dst, vi, v);
    uint64_t lim = fast_limit<INTERP>(a, v);
    const uint32_t tiles = F / a.gate_frames;
    uint32_t release = gate_window(gates, 0, tiles, v), window_end = GATE_WINDOW;
    for (uint32_t s = 0; s < F / SUBTILE; ++s) {
        const uint32_t f0 = s * SUBTILE;
        if (f0 % a.gate_frames == 0) {
            const uint32_t gt = f0 / a.gate_frames;
            if (gt >= window_end) {
                release = gate_window(gates, gt, tiles, v);
                window_end = gt + GATE_WINDOW;
            }
            if (gate_tile_at<CHAN>(a, chan, chan_flags, gt, release, v))
                lim = fast_limit<INTERP>(a, v);
        }
        const bool fok = fast_ok<INTERP>(v, lim, f0);
        const bool straight = !__any(!fok);
        const bool fading = __any(fade_in(v, f0));
        if (straight && !fading) {
            fast_subtile_dispatch<INTERP, CLAMP, false>(a, v, tb, ftab, lane, f0);
        } else if (straight && a.fade_lds != 0) {
            fast_subtile_dispatch<INTERP, CLAMP, true>(a, v, tb, ftab, lane, f0);
        } else if ((!fading || a.fade_lds != 0) && !__any(!(fok || wrap_ok<INTERP>(a, v, f0)))) {
            neutralise<INTERP>(v);
            uint32_t w0[SUBTILE], w1[INTERP == 2 ? SUBTILE : 1];
            const uint64_t end = wrap_taps<INTERP>(a, v, w0, w1);
            if (fading)
                fast_subtile<INTERP, CLAMP, 2, false, true, true>(a, v, tb, ftab, lane, f0, w0, w1,
                                                                  nullptr, end);
            else
                fast_subtile<INTERP, CLAMP, 2, false, false, true>(a, v, tb, ftab, lane, f0, w0, w1,
                                                                   nullptr, end);
        } else if (INTERP != 2 && !__any(!edge_ok<INTERP>(a, v, f0))) {
            fast_subtile_edge<INTERP, CLAMP>(a, v, tb, ftab, lane, f0);
        } else {
#pragma unroll 1
            for (uint32_t i = 0; i < SUBTILE; ++i) {
                float ol, orr;
                frame<INTERP>(a, f0 + i, v, ol, orr);
                tb[(2 * i) * TROW + lane] = ol;
                tb[(2 * i + 1) * TROW + lane] = orr;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Transposed read: output o over the 32 voices of this half.
        const float4 *row = reinterpret_cast<const float4 *>(tb + o * TROW + 32 * half);
        float4 q[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = row[k];
        float t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = (q[k].x + q[k].y) + (q[k].z + q[k].w);
        const float sum = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
        const float other = __shfl_xor(sum, 32);
        // Same value in both halves: (h0 + h1) == (h1 + h0). Half s%2 stores it.
        const float total = half ? other + sum : sum + other;
        if ((s & 1) == half) prow[s * 32 + o] = total;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (v.live) store_voice(a, dst, vi, v);
}

// The render's prologue: the fade table into LDS, and (one thread) what the
// next block's victim selection reads, the slots this render writes.
__device__ __forceinline__ uint32_t render_prologue(const Args &a, float *ftab) {
    for (uint32_t i = threadIdx.x; i < a.fade_lds; i += blockDim.x) ftab[i] = a.fade_tab[i];
    __syncthreads();
    DevState *st = a.st;
    const uint32_t live_total = st->live_total;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->snap_n = live_total;
        st->snap_parity = a.fused ? st->parity ^ 1 : st->parity;
        st->snap_block = st->block;
    }
    return live_total;
}

// The per-block render. Workgroup w runs on XCD x = w mod 8 (up to a rotation
// that keeps the classes together) and each XCD starts its workgroups in
// order, so workgroup w takes entries (w / 8) * wpg + wave of XCD x's dispatch
// list (k_order): the XCD works through one contiguous run of the sorted pool
// (its L2 holds the samples of a few regions), new voices and steal victims --
// the longest waves -- first. Without the lists (render_order null) the same
// runs are walked from their ends.
// Waves per SIMD: nearest and linear fit in 128 VGPRs and keep four; cubic's
// wider tap window takes more, and three waves without spills render 2.5 %
// faster than four with 4-14 VGPRs spilled.
template <int INTERP, bool CHAN, bool CLAMP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(INTERP == 2 ? 3 : 4)))
k_render_fast(Args a) {
    extern __shared__ float lds[];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t wpg = blockDim.x >> 6;
    float *tb = lds + wave * TBUF_FLOATS;
    float *ftab = lds + wpg * TBUF_FLOATS;
    const uint32_t live_total = render_prologue(a, ftab);
    ftab += FADE_PAD;  // the fade gains, d = 0 .. fade - 1
    uint32_t batch;
    if (a.render_order) {
        const DevState *st = a.st;
        const uint32_t nb = (live_total + 63) / 64;
        const uint32_t nsurv = st->live / 64 < nb ? st->live / 64 : nb;
        const uint32_t ns = nb - nsurv;
        const uint32_t x = blockIdx.x % XCDS;
        const uint32_t k = (blockIdx.x / XCDS) * wpg + wave;
        const uint32_t len = run_len(nsurv, x) + (ns > x ? (ns - 1 - x) / XCDS + 1 : 0u);
        if (k >= len) return;
        batch = a.render_order[(size_t)x * a.order_stride + k];
    } else {
        const uint32_t groups = (live_total + 64 * wpg - 1) / (64 * wpg);
        if (blockIdx.x >= groups) return;
        batch = xcd_group_rev(blockIdx.x, groups) * wpg + wave;
        if (batch * 64 >= live_total) return;
    }
    render_batch<INTERP, CHAN, CLAMP>(a, batch, live_total, tb, ftab, lane);
}

// The same pass over a batch of renders: blockIdx.z picks the render, and
// workgroup w renders the batches w * wpg + wave of its XCD's run, walked
// from the run's end (new voices, the slowest waves, first).
template <int INTERP, bool CHAN, bool CLAMP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kb_render_fast(const Args *__restrict__ aa) {
    extern __shared__ float lds[];
    const Args &a = aa[blockIdx.z];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t wpg = blockDim.x >> 6;
    float *tb = lds + wave * TBUF_FLOATS;
    float *ftab = lds + wpg * TBUF_FLOATS;
    const uint32_t live_total = render_prologue(a, ftab);
    ftab += FADE_PAD;  // the fade gains, d = 0 .. fade - 1
    const uint32_t groups = (live_total + 64 * wpg - 1) / (64 * wpg);
    if (blockIdx.x >= groups) return;
    const uint32_t batch = xcd_group_rev(blockIdx.x, groups) * wpg + wave;
    if (batch * 64 >= live_total) return;
    render_batch<INTERP, CHAN, CLAMP>(a, batch, live_total, tb, ftab, lane);
}

// Stage 1 of the mixdown: column sums of the batch rows in fixed order, one
// chunk of REDUCE_ROWS rows per workgroup (64 columns x 4 row phases), written
// as (hi, lo) pairs. With COMP the sums are compensated (two-sum), so the
// cross-batch sum of a million voices keeps its low-order bits.
template <bool COMP>
__device__ __forceinline__ void reduce_rows_body(const Args &a) {
    DevState *st = a.st;
    // The render has ended: publish it to the selection on the other stream.
    if (a.sig_render && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
        signal_seq(&st->render_seq, a.sig_render);
    const uint32_t two_f = 2 * a.F;
    const uint32_t rows = (st->live_total + 63) / 64;
    const uint32_t chunk = blockIdx.y;
    const uint32_t r0 = chunk * REDUCE_ROWS;
    if (r0 >= rows) return;
    const uint32_t r1 = r0 + REDUCE_ROWS < rows ? r0 + REDUCE_ROWS : rows;
    const uint32_t col = blockIdx.x * 64 + (threadIdx.x & 63);
    const uint32_t ph = threadIdx.x >> 6;
    __shared__ float sh[2][4][64];
    float hi = 0.0f, lo = 0.0f;
    if (col < two_f) {
        // Rows r0+ph, r0+ph+4, ... in order; loads issued RU at a time ahead
        // of the dependent sum.
        constexpr uint32_t RU = 8;
        for (uint32_t r = r0 + ph; r < r1; r += 4 * RU) {
            float x[RU];
#pragma unroll
            for (uint32_t u = 0; u < RU; ++u) {
                const uint32_t rr = r + 4 * u;
                x[u] = rr < r1 ? a.partials[(size_t)rr * two_f + col] : 0.0f;
            }
#pragma unroll
            for (uint32_t u = 0; u < RU; ++u) {
                if (r + 4 * u >= r1) break;
                if (COMP) {
                    float s, e;
                    two_sum(hi, x[u], s, e);
                    hi = s;
                    lo = lo + e;
                } else {
                    hi = hi + x[u];
                }
            }
        }
    }
    sh[0][ph][threadIdx.x & 63] = hi;
    sh[1][ph][threadIdx.x & 63] = lo;
    __syncthreads();
    if (ph == 0 && col < two_f) {
        float h = sh[0][0][threadIdx.x], l = sh[1][0][threadIdx.x];
        for (uint32_t k = 1; k < 4; ++k) {
            if (COMP) {
                float s, e;
                two_sum(h, sh[0][k][threadIdx.x], s, e);
                h = s;
                l = l + (e + sh[1][k][threadIdx.x]);
            } else {
                h = h + sh[0][k][threadIdx.x];
            }
        }
        a.rsum[((size_t)chunk * two_f + col) * 2] = h;
        a.rsum[((size_t)chunk * two_f + col) * 2 + 1] = l;
    }
}
template <bool COMP>
__global__ void __launch_bounds__(256) k_reduce_rows(Args a) { reduce_rows_body<COMP>(a); }
// The same pass over a batch of renders: blockIdx.z picks the render.
template <bool COMP>
__global__ void __launch_bounds__(256) kb_reduce_rows(const Args *__restrict__ aa) {
    reduce_rows_body<COMP>(aa[blockIdx.z]);
}

// Stage 2: the chunks in order, one thread per output sample, then the block
// peak (src/cpu.rs:615-625).
template <bool COMP>
__device__ __forceinline__ void reduce_final_body(const Args &a) {
    DevState *st = a.st;
    if (st->idle) return;
    const uint32_t two_f = 2 * a.F;
    // FQ threads per output sample, each over a contiguous quarter of the
    // chunks; the quarters are joined in order.
    constexpr uint32_t FQ = 4;
    __shared__ float qh[FQ][64], ql[FQ][64];
    const uint32_t q = threadIdx.x / 64, t = threadIdx.x % 64;
    const uint32_t col = blockIdx.x * 64 + t;
    const uint32_t rows = (st->live_total + 63) / 64;
    const uint32_t chunks = (rows + REDUCE_ROWS - 1) / REDUCE_ROWS;
    const uint32_t cq = (chunks + FQ - 1) / FQ;
    const uint32_t cbeg = q * cq, cend = cbeg + cq < chunks ? cbeg + cq : chunks;
    float h = 0.0f, l = 0.0f;
    constexpr uint32_t CU = 8;
    for (uint32_t c0 = cbeg; c0 < cend && col < two_f; c0 += CU) {
        float2 x[CU];
#pragma unroll
        for (uint32_t u = 0; u < CU; ++u) {
            const uint32_t c = c0 + u;
            x[u] = c < cend ? reinterpret_cast<const float2 *>(a.rsum)[(size_t)c * two_f + col]
                            : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (uint32_t u = 0; u < CU; ++u) {
            if (c0 + u >= cend) break;
            if (COMP) {
                float s, e;
                two_sum(h, x[u].x, s, e);
                h = s;
                l = l + (e + x[u].y);
            } else {
                h = h + x[u].x;
            }
        }
    }
    qh[q][t] = h;
    ql[q][t] = l;
    __syncthreads();
    if (q != 0 || col >= two_f) return;
    for (uint32_t k = 1; k < FQ; ++k) {
        if (COMP) {
            float s, e;
            two_sum(h, qh[k][t], s, e);
            h = s;
            l = l + (e + ql[k][t]);
        } else {
            h = h + qh[k][t];
        }
    }
    const float v = COMP ? h + l : h;
    a.out[(size_t)st->block * two_f + col] = v;
    const float m = fabsf(v);
    if (m == m) atomicMax(&st->peak_bits, __float_as_uint(m));
}
template <bool COMP>
__global__ void __launch_bounds__(256) k_reduce_final(Args a) { reduce_final_body<COMP>(a); }
// The same pass over a batch of renders: blockIdx.z picks the render.
template <bool COMP>
__global__ void __launch_bounds__(256) kb_reduce_final(const Args *__restrict__ aa) {
    reduce_final_body<COMP>(aa[blockIdx.z]);
}

// ---------------------------------------------------------------------------
// plain form: the reference's render + reduce structure
// ---------------------------------------------------------------------------

template <int INTERP, bool CHAN>
__global__ void __launch_bounds__(1024) k_render_plain(Args a) {
    extern __shared__ float lds[];
    const uint32_t WG = blockDim.x;
    const uint32_t TILE = a.plain_tile;
    const uint32_t M = TILE * 2;
    const uint32_t PER_LANE = WG / M;
    float *sh = lds;             // [M][WG]
    float *sh2 = lds + WG * M;   // [WG]
    const uint32_t tid = threadIdx.x, wg = blockIdx.x, nwg = gridDim.x;
    DevState *st = a.st;
    const uint32_t live = st->live_total;
    const uint32_t b = st->block;
    const uint32_t chan_flags = CHAN ? st->chan_flags : 0u;
    uint32_t *pool = sel2(a.pools, st->parity);
    const uint32_t *gates = a.gates + (size_t)b * a.gate_words;
    const uint32_t *chan = a.chans + (size_t)b * a.chan_words;
    const uint32_t samples = a.F * 2;

    for (uint32_t j = tid; j < samples; j += WG) a.partials[(size_t)j * nwg + wg] = 0.0f;
    __syncthreads();

    for (uint32_t batch = wg; batch * WG < live; batch += nwg) {
        const uint32_t vi = batch * WG + tid;
        Voice v;
        load_voice(a, pool, vi, vi < live, v);
        const uint32_t tiles = a.F / TILE;
        const uint32_t tiles_per_gate = a.gate_frames / TILE;
        for (uint32_t tile = 0; tile < tiles; ++tile) {
            if (tile % tiles_per_gate == 0)
                gate_tile<CHAN>(a, gates, chan, chan_flags, tile / tiles_per_gate, v);
            const uint32_t f0 = tile * TILE;
            for (uint32_t i = 0; i < TILE; ++i) {
                float ol, orr;
                frame<INTERP>(a, f0 + i, v, ol, orr);
                sh[(i * 2) * WG + tid] = ol;
                sh[(i * 2 + 1) * WG + tid] = orr;
            }
            __syncthreads();
            // shaders/render.wgsl:125-144
            const uint32_t lane = tid / PER_LANE, chunk = tid % PER_LANE;
            float acc = 0.0f;
            for (uint32_t k = 0; k < M; ++k) acc = acc + sh[lane * WG + chunk + k * PER_LANE];
            sh2[lane * PER_LANE + chunk] = acc;
            __syncthreads();
            if (tid < M) {
                float total = 0.0f;
                for (uint32_t k = 0; k < PER_LANE; ++k) total = total + sh2[tid * PER_LANE + k];
                const size_t j = (size_t)(f0 * 2 + tid) * nwg + wg;
                a.partials[j] = a.partials[j] + total;
            }
            __syncthreads();
        }
        if (v.live) store_voice(a, pool, vi, v);
    }
}

// shaders/reduce.wgsl:110-150
__global__ void __launch_bounds__(1024) k_reduce_plain(Args a) {
    extern __shared__ float red[];
    DevState *st = a.st;
    const uint32_t j = blockIdx.x, tid = threadIdx.x, WG = blockDim.x;
    const uint32_t nwg = a.render_groups;
    const size_t base = (size_t)j * nwg;
    float acc = 0.0f;
    if (a.kahan) {
        float comp = 0.0f;
        for (uint32_t w = tid; w < nwg; w += WG) {
            const float y = a.partials[base + w] - comp;
            const float t = acc + y;
            comp = (t - acc) - y;
            acc = t;
        }
    } else {
        for (uint32_t w = tid; w < nwg; w += WG) acc = acc + a.partials[base + w];
    }
    red[tid] = acc;
    __syncthreads();
    for (uint32_t s = WG / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = red[tid] + red[tid + s];
        __syncthreads();
    }
    if (tid == 0) {
        const float v = red[0];
        a.out[(size_t)st->block * 2 * a.F + j] = v;
        const float m = fabsf(v);
        if (m == m) atomicMax(&st->peak_bits, __float_as_uint(m));
    }
}

template <int INTERP, bool CHAN, bool CLAMP>
void launch_fast_t(const Args &a, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
    const uint32_t wpg = a.waves_per_group;
    const uint32_t threads = wpg * 64;
    const size_t lds = fast_render_lds_bytes(wpg, a.fade_lds);
    uint32_t grid = a.render_groups;
    if (a.render_order) {
        // Each XCD's list holds at most its run plus its share of the new
        // voices' batches: ceil(batches / 8) + 9 entries (its chunks round up).
        const uint32_t per_xcd = (a.render_groups * wpg + XCDS - 1) / XCDS + RUN_CHUNKS / XCDS + 2;
        grid = XCDS * ((per_xcd + wpg - 1) / wpg);
    }
    if (start || stop)
        hipExtLaunchKernelGGL(k_render_fast<INTERP, CHAN, CLAMP>, dim3(grid), dim3(threads),
                              (uint32_t)lds, s, start, stop, 0u, a);
    else
        k_render_fast<INTERP, CHAN, CLAMP><<<grid, threads, lds, s>>>(a);
}

template <int INTERP, bool CHAN>
void launch_plain_t(const Args &a, hipStream_t s) {
    const uint32_t WG = a.plain_wg;
    const size_t lds = (size_t)WG * (a.plain_tile * 2 + 1) * sizeof(float);
    k_render_plain<INTERP, CHAN><<<a.render_groups, WG, lds, s>>>(a);
}

}  // namespace

bool fast_render_supported(uint32_t F, uint32_t gate_frames, uint32_t interp) {
    return F % 32 == 0 && gate_frames % SUBTILE == 0 && interp <= 2;
}

size_t fast_render_lds_bytes(uint32_t waves_per_group, uint32_t fade_lds) {
    return ((size_t)waves_per_group * TBUF_FLOATS + ((fade_lds + 3) & ~3u)) * sizeof(float);
}

void launch_render_fast(const Args &a, bool chan, hipStream_t s, hipEvent_t start,
                        hipEvent_t stop) {
#define KST_FAST(I, CL)                                                                    \
    if (a.interp == I && (a.clamp != 0) == CL) {                                           \
        chan ? launch_fast_t<I, true, CL>(a, s, start, stop)                               \
             : launch_fast_t<I, false, CL>(a, s, start, stop);                             \
        return;                                                                            \
    }
    KST_FAST(0, false)
    KST_FAST(1, false)
    KST_FAST(2, false)
    KST_FAST(0, true)
    KST_FAST(1, true)
    KST_FAST(2, true)
#undef KST_FAST
}

void launch_batch_render(const Args *aa, uint32_t nr, uint32_t interp, bool chan, bool clamp,
                         uint32_t waves_per_group, uint32_t groups, uint32_t fade_lds,
                         hipStream_t s) {
    const dim3 grid((groups + 7) & ~7u, 1, nr);  // a multiple of 8: XCD = blockIdx.x % 8
    const uint32_t threads = waves_per_group * 64;
    const size_t lds = fast_render_lds_bytes(waves_per_group, fade_lds);
#define KST_BFAST(I, CH, CL)                                                                \
    if (interp == I && chan == CH && clamp == CL) {                                         \
        kb_render_fast<I, CH, CL><<<grid, threads, lds, s>>>(aa);                          \
        return;                                                                             \
    }
    KST_BFAST(0, false, false) KST_BFAST(1, false, false) KST_BFAST(2, false, false)
    KST_BFAST(0, true, false) KST_BFAST(1, true, false) KST_BFAST(2, true, false)
    KST_BFAST(0, false, true) KST_BFAST(1, false, true) KST_BFAST(2, false, true)
    KST_BFAST(0, true, true) KST_BFAST(1, true, true) KST_BFAST(2, true, true)
#undef KST_BFAST
}

void launch_batch_reduce(const Args *aa, uint32_t nr, bool comp, uint32_t max_two_f,
                         uint32_t max_rows, hipStream_t s) {
    const dim3 g1((max_two_f + 63) / 64, (max_rows + REDUCE_ROWS - 1) / REDUCE_ROWS, nr);
    const dim3 g2((max_two_f + 63) / 64, 1, nr);
    if (comp) {
        kb_reduce_rows<true><<<g1, 256, 0, s>>>(aa);
        kb_reduce_final<true><<<g2, 256, 0, s>>>(aa);
    } else {
        kb_reduce_rows<false><<<g1, 256, 0, s>>>(aa);
        kb_reduce_final<false><<<g2, 256, 0, s>>>(aa);
    }
}

void launch_reduce_fast(const Args &a, bool comp, uint32_t max_rows, hipStream_t s) {
    const uint32_t two_f = 2 * a.F;
    const dim3 g1((two_f + 63) / 64, (max_rows + REDUCE_ROWS - 1) / REDUCE_ROWS);
    const uint32_t g2 = (two_f + 63) / 64;
    if (comp) {
        k_reduce_rows<true><<<g1, 256, 0, s>>>(a);
        k_reduce_final<true><<<g2, 256, 0, s>>>(a);
    } else {
        k_reduce_rows<false><<<g1, 256, 0, s>>>(a);
        k_reduce_final<false><<<g2, 256, 0, s>>>(a);
    }
}

void launch_render_plain(const Args &a, bool chan, hipStream_t s) {
    if (a.interp == 0) chan ? launch_plain_t<0, true>(a, s) : launch_plain_t<0, false>(a, s);
    else if (a.interp == 1) chan ? launch_plain_t<1, true>(a, s) : launch_plain_t<1, false>(a, s);
    else chan ? launch_plain_t<2, true>(a, s) : launch_plain_t<2, false>(a, s);
}

void launch_reduce_plain(const Args &a, hipStream_t s) {
    const uint32_t WG = a.plain_wg;
    k_reduce_plain<<<2 * a.F, WG, WG * sizeof(float), s>>>(a);
}

}  // namespace kst
