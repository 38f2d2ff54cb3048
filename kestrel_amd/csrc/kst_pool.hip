// kst_pool.hip -- the per-block pool passes around the render kernel:
//
//   k_begin      steal / spawn arithmetic of one block   src/gpu/mod.rs:1015-1040
//   k_select     radix select of the k-th smallest key   shaders/select.wgsl
//   k_mark       schedule the victims' stop frames       shaders/compact.wgsl:148-191
//   k_spawn      append the admitted note-ons            shaders/spawn.wgsl
//   k_sort_*     stable LSD radix sort by locality key   shaders/sort.wgsl
//                (or the 1-bit alive key: compaction)    shaders/compact.wgsl:26-146
//   k_gather     compaction + reorder in one copy        shaders/sort.wgsl:373-401
//   k_end        publish the compacted count             shaders/compact.wgsl:181-186
//
// All integer; the only atomics are integer counters whose totals do not
// depend on order, so every pass is deterministic.
#include "kst_device.h"

namespace kst {
namespace {

__device__ __forceinline__ uint64_t steal_key(const Args &a, const uint32_t *pool, uint32_t i) {
    // src/cpu.rs:208-221 / shaders/common.wgsl:170-177
    uint64_t id = ((uint64_t)pool[F_NOTE_HI * a.cap + i] << 32) | pool[F_NOTE_LO * a.cap + i];
    if (a.steal_rule != STEAL_QUIETEST) return id;
    float level = __uint_as_float(pool[F_ENV_LEVEL * a.cap + i]);
    // Rust clamp keeps NaN, and `NaN as u32` is 0.
    uint32_t q = 0;
    if (level == level) {
        level = level < 0.0f ? 0.0f : (level > 1.0f ? 1.0f : level);
        q = (uint32_t)(level * 65535.0f);
    }
    return ((uint64_t)q << 48) | (id & 0x0000FFFFFFFFFFFFull);
}

__global__ void k_begin(Args a) {
    DevState *st = a.st;
    const uint32_t b = st->block;
    // Clear the select histograms and sort totals for this block.
    for (uint32_t i = threadIdx.x; i < SEL_ROUNDS * 256; i += blockDim.x) a.sel_hist[i] = 0;
    const uint32_t radix = 1u << a.digit_bits;
    for (uint32_t i = threadIdx.x; i < a.sort_passes * radix; i += blockDim.x) a.sort_tot[i] = 0;
    if (threadIdx.x != 0) return;
    // src/gpu/mod.rs:1015-1040 (and src/cpu.rs:330-384, which agrees).
    const uint64_t cap = a.max_voices;
    const uint64_t live = st->live;
    const uint64_t raw = a.cmd_counts[b];
    const uint64_t want = raw < cap ? raw : cap;
    uint64_t steal_k = 0;
    if (live + want > cap && a.steal_rule != STEAL_DROP_NEW) {
        steal_k = live + want - cap;
        if (steal_k > live) steal_k = live;
        if (steal_k > a.max_steal) steal_k = a.max_steal;
    }
    uint64_t room = cap - (live - steal_k);
    uint64_t spawn = want < room ? want : room;
    st->want = (uint32_t)raw;
    st->steal_k = (uint32_t)steal_k;
    st->spawn_count = (uint32_t)spawn;
    st->live_total = (uint32_t)(live + spawn);
    st->cmd_off = a.cmd_offs[b];
    st->chan_flags = a.chan_flags[b];
    st->live_new = 0;
    st->peak_bits = 0;
    st->stolen += steal_k;
    st->dropped += raw - spawn;
    st->sel_prefix_hi[0] = 0;
    st->sel_prefix_lo[0] = 0;
    st->sel_k[0] = steal_k > 0 ? (uint32_t)steal_k : 1;
}

// Pick one byte of the answer from the previous round's histogram
// (shaders/select.wgsl:235-262): the first bin whose inclusive count reaches k.
// The whole workgroup scans the 256 bins in LDS; every workgroup computes the
// same answer, and workgroup 0 publishes it for the next round.
__device__ void refine(const Args &a, uint32_t r, uint64_t *s_prefix, uint32_t *s_k,
                       uint32_t *scan) {
    DevState *st = a.st;
    const uint32_t prev = r - 1;
    const uint32_t t = threadIdx.x;  // blockDim.x == 256
    uint32_t h = a.sel_hist[prev * 256 + t];
    scan[t] = h;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        uint32_t x = t >= off ? scan[t - off] : 0;
        __syncthreads();
        scan[t] += x;
        __syncthreads();
    }
    const uint32_t kk = st->sel_k[prev];
    const uint32_t incl = scan[t];
    const uint32_t excl = incl - h;
    if (incl >= kk && excl < kk) {
        uint64_t p = ((uint64_t)st->sel_prefix_hi[prev] << 32) | st->sel_prefix_lo[prev];
        p |= (uint64_t)t << (56 - 8 * prev);
        *s_prefix = p;
        *s_k = kk - excl;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_select(Args a, uint32_t r) {
    DevState *st = a.st;
    if (st->steal_k == 0) return;
    __shared__ uint32_t hist[256];
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_k;
    if (threadIdx.x == 0) {
        // A histogram that never reaches k (it cannot, for unique keys) keeps
        // bin 255, as the reference's refine does.
        s_prefix = r == 0 ? 0ull : ~0ull;
        s_k = st->sel_k[0];
    }
    __syncthreads();
    if (r > 0) refine(a, r, &s_prefix, &s_k, hist);
    if (threadIdx.x == 0 && blockIdx.x == 0 && r > 0) {
        st->sel_prefix_hi[r] = (uint32_t)(s_prefix >> 32);
        st->sel_prefix_lo[r] = (uint32_t)s_prefix;
        st->sel_k[r] = s_k;
    }
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t prefix = s_prefix;
    const uint32_t *pool = sel2(a.pools, st->parity);
    const uint32_t live = st->live;
    const uint32_t shift = 56 - 8 * r;
    const uint32_t hi_shift = 64 - 8 * r;  // bits above this round's byte
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < live;
         i += gridDim.x * blockDim.x) {
        uint64_t key = steal_key(a, pool, i);
        bool match = r == 0 || (key >> hi_shift) == (prefix >> hi_shift);
        if (match) atomicAdd(&hist[(key >> shift) & 255], 1u);
    }
    __syncthreads();
    uint32_t c = hist[threadIdx.x];
    if (c) atomicAdd(&a.sel_hist[r * 256 + threadIdx.x], c);
}

// The threshold is the k-th smallest key; keys are unique, so `key <= T` names
// exactly steal_k voices (shaders/compact.wgsl:151-179, src/cpu.rs:365-370).
__global__ void __launch_bounds__(256) k_mark(Args a) {
    DevState *st = a.st;
    if (st->steal_k == 0) return;
    __shared__ uint64_t s_thresh;
    __shared__ uint32_t s_k;
    __shared__ uint32_t scan[256];
    if (threadIdx.x == 0) s_thresh = ~0ull;
    __syncthreads();
    refine(a, SEL_ROUNDS, &s_thresh, &s_k, scan);
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->thresh_hi = (uint32_t)(s_thresh >> 32);
        st->thresh_lo = (uint32_t)s_thresh;
    }
    const uint64_t t = s_thresh;
    uint32_t *pool = sel2(a.pools, st->parity);
    const uint32_t live = st->live;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < live;
         i += gridDim.x * blockDim.x) {
        if (steal_key(a, pool, i) <= t) {
            uint32_t lo = pool[F_NOTE_LO * a.cap + i];
            pool[F_STOP_REL * a.cap + i] = lo % a.steal_span + 1;
        }
    }
}

// shaders/spawn.wgsl:35-78; thinning per src/voice.rs:72-75 when the block is
// oversubscribed under AdmitRule::Even (src/gpu/mod.rs:1056-1071).
__global__ void __launch_bounds__(256) k_spawn(Args a) {
    DevState *st = a.st;
    const uint32_t n = st->spawn_count;
    if (n == 0) return;
    const uint32_t want = st->want;
    const bool thin = a.admit_rule == ADMIT_EVEN && n < want;
    const uint32_t base = st->live;
    const kst_spawn_cmd *cmds = a.cmds + st->cmd_off;
    uint32_t *pool = sel2(a.pools, st->parity);
    const uint32_t c = a.cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        uint32_t src = thin ? (uint32_t)((uint64_t)i * want / n) : i;
        const kst_spawn_cmd s = cmds[src];
        uint32_t d = base + i;
        pool[F_PHASE_LO * c + d] = s.phase_lo;
        pool[F_PHASE_HI * c + d] = s.phase_hi;
        pool[F_STEP_LO * c + d] = s.step_lo;
        pool[F_STEP_HI * c + d] = s.step_hi;
        pool[F_SMP_BASE * c + d] = s.smp_base;
        pool[F_SMP_LEN * c + d] = s.smp_len;
        pool[F_LOOP_START * c + d] = s.loop_start;
        pool[F_LOOP_END * c + d] = s.loop_end;
        pool[F_FLAGS * c + d] = s.flags;
        pool[F_ENV_STAGE * c + d] = ENV_ATTACK;
        pool[F_ENV_LEVEL * c + d] = 0u;
        pool[F_GAIN_L * c + d] = __float_as_uint(s.gain_l);
        pool[F_GAIN_R * c + d] = __float_as_uint(s.gain_r);
        pool[F_FILT_Z1 * c + d] = 0u;
        pool[F_FILT_Z2 * c + d] = 0u;
        pool[F_PARAMS * c + d] = s.params;
        pool[F_REGION * c + d] = s.region;
        pool[F_GATE_SLOT * c + d] = s.gate_slot;
        pool[F_ORDINAL * c + d] = s.ordinal;
        pool[F_START_REL * c + d] = s.start_rel;
        pool[F_NOTE_LO * c + d] = s.note_id_lo;
        pool[F_NOTE_HI * c + d] = s.note_id_hi;
        pool[F_BORN_VARIANT * c + d] = s.variant + 1u;
        pool[F_STOP_REL * c + d] = 0u;
    }
}

// ---------------------------------------------------------------------------
// Locality sort / compaction: stable LSD radix sort of (key, slot) pairs over
// the render's voices. The render pass wrote each voice's key; dead voices
// carry the dead region (or the bit 1 when sorting is off), so they land past
// every survivor and the gather stops before them.
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool key_alive(const Args &a, uint32_t key) {
    return a.sort_on ? (key >> a.region_shift) < a.dead_region : key == 0;
}

// Per-tile digit counts of pass p over that pass's input order (the previous
// pass's output), plus the digit totals. Pass 0 also counts the survivors.
__global__ void __launch_bounds__(SORT_TPB) k_sort_hist(Args a, uint32_t p) {
    DevState *st = a.st;
    const uint32_t n = st->live_total;
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint32_t radix = 1u << a.digit_bits, mask = radix - 1;
    const uint32_t shift = p * a.digit_bits;
    __shared__ uint32_t sh[256];
    __shared__ uint32_t s_alive;
    const uint32_t *keys = sel2(a.keys, p & 1);
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < radix; i += SORT_TPB) sh[i] = 0;
        if (threadIdx.x == 0) s_alive = 0;
        __syncthreads();
        uint32_t alive = 0;
        for (uint32_t k = 0; k < SORT_IPT; ++k) {
            uint32_t item = t * SORT_TILE + k * SORT_TPB + threadIdx.x;
            if (item >= n) break;
            uint32_t key = keys[item];
            if (p == 0) alive += key_alive(a, key) ? 1u : 0u;
            atomicAdd(&sh[(key >> shift) & mask], 1u);
        }
        if (alive) atomicAdd(&s_alive, alive);
        __syncthreads();
        for (uint32_t d = threadIdx.x; d < radix; d += SORT_TPB) {
            uint32_t c = sh[d];
            a.sort_hist[(size_t)(p * radix + d) * a.tiles_max + t] = c;
            if (c) atomicAdd(&a.sort_tot[p * radix + d], c);
        }
        if (threadIdx.x == 0 && s_alive) atomicAdd(&st->live_new, s_alive);
        __syncthreads();
    }
}

// One workgroup per (pass, digit): the digit's base is the sum of the smaller
// digits' totals; each tile's offset adds the same digit's counts in the tiles
// before it.
__global__ void __launch_bounds__(256) k_sort_scan(Args a, uint32_t p) {
    DevState *st = a.st;
    const uint32_t n = st->live_total;
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint32_t radix = 1u << a.digit_bits;
    const uint32_t d = blockIdx.x;
    __shared__ uint32_t red[256];
    uint32_t s = 0;
    for (uint32_t i = threadIdx.x; i < d; i += 256) s += a.sort_tot[p * radix + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    uint32_t running = red[0];
    __syncthreads();
    uint32_t *row = a.sort_hist + (size_t)(p * radix + d) * a.tiles_max;
    // Exclusive scan of row[0..tiles) in chunks of 256.
    for (uint32_t base = 0; base < tiles; base += 256) {
        uint32_t t = base + threadIdx.x;
        uint32_t v = t < tiles ? row[t] : 0;
        red[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 256; off <<= 1) {
            uint32_t x = threadIdx.x >= off ? red[threadIdx.x - off] : 0;
            __syncthreads();
            red[threadIdx.x] += x;
            __syncthreads();
        }
        if (t < tiles) row[t] = running + red[threadIdx.x] - v;
        uint32_t total = red[255];
        __syncthreads();
        running += total;
    }
}

// Stable scatter of one digit pass. Items of a tile are ranked in index
// order: rounds of 256 consecutive items, waves in order within a round,
// lanes in order within a wave (match masks built from ballots).
__global__ void __launch_bounds__(SORT_TPB) k_sort_scatter(Args a, uint32_t p) {
    DevState *st = a.st;
    const uint32_t n = st->live_total;
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint32_t radix = 1u << a.digit_bits, mask = radix - 1;
    const uint32_t shift = p * a.digit_bits;
    const uint32_t *kin = sel2(a.keys, p & 1);
    const uint32_t *iin = sel2(a.idx, p & 1);
    uint32_t *kout = sel2(a.keys, (p + 1) & 1);
    uint32_t *iout = sel2(a.idx, (p + 1) & 1);
    __shared__ uint32_t cnt[256];
    __shared__ uint32_t wc[SORT_TPB / 64][256];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        for (uint32_t d = threadIdx.x; d < radix; d += SORT_TPB)
            cnt[d] = a.sort_hist[(size_t)(p * radix + d) * a.tiles_max + t];
        for (uint32_t k = 0; k < SORT_IPT; ++k) {
            for (uint32_t i = threadIdx.x; i < (SORT_TPB / 64) * radix; i += SORT_TPB)
                wc[i / radix][i % radix] = 0;
            __syncthreads();
            const uint32_t item = t * SORT_TILE + k * SORT_TPB + threadIdx.x;
            const bool valid = item < n;
            uint32_t key = valid ? kin[item] : 0;
            uint32_t digit = (key >> shift) & mask;
            uint64_t m = __ballot(valid);
            for (uint32_t bit = 0; bit < a.digit_bits; ++bit) {
                uint64_t bb = __ballot(valid && ((digit >> bit) & 1));
                m &= ((digit >> bit) & 1) ? bb : ~bb;
            }
            uint32_t rank = __popcll(m & lt);
            if (valid && rank == 0) wc[wave][digit] = __popcll(m);
            __syncthreads();
            if (valid) {
                uint32_t off = cnt[digit] + rank;
                for (uint32_t w = 0; w < wave; ++w) off += wc[w][digit];
                kout[off] = key;
                iout[off] = p == 0 ? item : iin[item];
            }
            __syncthreads();
            for (uint32_t d = threadIdx.x; d < radix; d += SORT_TPB) {
                uint32_t s = 0;
                for (uint32_t w = 0; w < SORT_TPB / 64; ++w) s += wc[w][d];
                cnt[d] += s;
            }
            __syncthreads();
        }
    }
}

// shaders/sort.wgsl:376-400 / compact.wgsl:116-146: survivors move to the other
// buffer in sorted (or original) order; per-block fields are cleared.
__global__ void __launch_bounds__(256) k_gather(Args a) {
    DevState *st = a.st;
    const uint32_t n = st->live_new;
    const uint32_t *perm = sel2(a.idx, a.sort_passes & 1);
    const uint32_t *src = sel2(a.pools, st->parity);
    uint32_t *dst = sel2(a.pools, st->parity ^ 1);
    const uint32_t c = a.cap;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += gridDim.x * blockDim.x) {
        uint32_t s = perm[j];
#pragma unroll
        for (uint32_t f = 0; f < VOICE_FIELDS; ++f) dst[f * c + j] = src[f * c + s];
        dst[F_START_REL * c + j] = 0;
        dst[F_BORN_VARIANT * c + j] = 0;
        dst[F_STOP_REL * c + j] = 0;
    }
}

__global__ void k_end(Args a) {
    DevState *st = a.st;
    BlockResult r;
    r.active_voices = st->live_new;
    r.stolen = st->stolen;
    r.dropped = st->dropped;
    r.peak_bits = st->peak_bits;
    r.pad = 0;
    a.results[st->block] = r;
    st->live = st->live_new;
    st->parity ^= 1;
    st->block += 1;
}

}  // namespace

void launch_begin(const Args &a, hipStream_t s) { k_begin<<<1, 256, 0, s>>>(a); }

void launch_select(const Args &a, uint32_t grid, hipStream_t s) {
    for (uint32_t r = 0; r < SEL_ROUNDS; ++r) k_select<<<grid, 256, 0, s>>>(a, r);
    k_mark<<<grid, 256, 0, s>>>(a);
}

void launch_spawn(const Args &a, uint32_t grid, hipStream_t s) {
    k_spawn<<<grid, 256, 0, s>>>(a);
}

void launch_sort(const Args &a, uint32_t grid, hipStream_t s) {
    const uint32_t radix = 1u << a.digit_bits;
    for (uint32_t p = 0; p < a.sort_passes; ++p) {
        k_sort_hist<<<grid, SORT_TPB, 0, s>>>(a, p);
        k_sort_scan<<<radix, 256, 0, s>>>(a, p);
        k_sort_scatter<<<grid, SORT_TPB, 0, s>>>(a, p);
    }
    k_gather<<<grid, 256, 0, s>>>(a);
}

void launch_end(const Args &a, hipStream_t s) { k_end<<<1, 1, 0, s>>>(a); }

}  // namespace kst
