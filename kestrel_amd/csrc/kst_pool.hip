// kst_pool.hip -- the per-block pool passes around the render kernel:
//
//   k_begin      steal / spawn arithmetic of one block   src/gpu/mod.rs:1015-1040
//   k_sel_*      radix select of the steal_k smallest    shaders/select.wgsl
//                keys, marking the victims' stop frames  shaders/compact.wgsl:148-191
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

__device__ __forceinline__ void begin_body(const Args &a) {
    DevState *st = a.st;
    const uint32_t b = st->block;
    if (threadIdx.x != 0) return;
    if (b >= a.nblocks) {
        // A batch member whose blocks are done: nothing to select, spawn,
        // render or sort, and the pool and counters stay as they are.
        st->idle = 1;
        st->steal_k = 0;
        st->spawn_count = 0;
        st->live_total = 0;
        return;
    }
    st->idle = 0;
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
}
__global__ void k_begin(Args a) { begin_body(a); }
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void kb_begin(const Args *__restrict__ aa) {
    begin_body(aa[blockIdx.z]);
}

// The selection's set-up (src/cpu.rs:345-371 picks victims among the voices
// live at the block's start). Candidates are the sounding voices among a
// range of slots: after a fused block, the slots its render wrote (survivors
// and the dying, in render order -- the victim set does not depend on order);
// otherwise the compacted live pool. With sel_ahead the block is the one after
// the last render, selected while that render's block still reduces and
// sorts; the steal count repeats k_begin's arithmetic from the same inputs.
// src/gpu/mod.rs:1015-1040: how many of `live` voices a block with `raw`
// note-ons steals.
__device__ __forceinline__ uint32_t steal_count(const Args &a, uint64_t live, uint64_t raw) {
    const uint64_t cap = a.max_voices;
    const uint64_t want = raw < cap ? raw : cap;
    uint64_t steal_k = 0;
    if (live + want > cap && a.steal_rule != STEAL_DROP_NEW) {
        steal_k = live + want - cap;
        if (steal_k > live) steal_k = live;
        if (steal_k > a.max_steal) steal_k = a.max_steal;
    }
    return (uint32_t)steal_k;
}

// The steal count once the number of sounding candidates is known (the first
// refine's bin total, or the collected count when no round ran).
__device__ __forceinline__ uint32_t resolve_steal(const Args &a, DevState *st, uint32_t live) {
    const uint32_t k = steal_count(a, live, st->sel_want);
    st->sel_steal = k;
    st->sel_k = k > 0 ? k : 1;
    return k;
}

__device__ __forceinline__ void sel_begin_body(const Args &a) {
    if (threadIdx.x != 0) return;
    DevState *st = a.st;
    const bool from_render = a.sel_ahead || (a.fused && st->perm_valid);
    uint32_t blk, n, par;
    if (from_render) {
        blk = a.sel_ahead ? st->snap_block + 1 : st->block;
        n = st->snap_n;
        par = a.sel_ahead ? st->snap_parity : st->parity;
    } else {
        blk = st->block;
        n = st->live;
        par = st->parity;
    }
    st->sel_n = n;
    st->sel_parity = par;
    st->sel_want = blk < a.nblocks ? a.cmd_counts[blk] : 0u;
    st->sel_prefix = 0;
    st->sel_top = 64;
    st->cand_n = 0;
    if (blk >= a.nblocks) {
        st->sel_steal = 0;
    } else if (from_render) {
        // How many of the render's voices still sound is counted by the first
        // histogram round (or by the collection); the steal count waits for it.
        st->sel_steal = SEL_PENDING;
        st->sel_k = 1;
        st->sel_count = n;
        st->sel_done = n <= SEL_CAP;
        return;
    } else {
        resolve_steal(a, st, n);
    }
    st->sel_count = n;
    st->sel_done = n <= SEL_CAP;  // few enough for the final workgroup alone
}
// In a staged run the selection of the block after block b first waits until
// b's render has ended (its mixdown has started).
__global__ void k_sel_begin(Args a) {
    if (a.wait_render) wait_seq(a.st, &a.st->render_seq, a.wait_render);
    sel_begin_body(a);
}
__global__ void kb_sel_begin(const Args *__restrict__ aa) { sel_begin_body(aa[blockIdx.z]); }

__device__ __forceinline__ bool sounding_at(const Args &a, const uint32_t *pool, uint32_t i) {
    return pool[F_ENV_STAGE * a.cap + i] != ENV_DEAD;
}

// Slot i's steal key and whether it is a candidate (sounding), with every load
// issued unconditionally so a thread's slots are fetched together.
__device__ __forceinline__ bool cand_at(const Args &a, const uint32_t *pool, uint32_t i,
                                        uint32_t n, unsigned long long &key) {
    if (i >= n) {
        key = 0;
        return false;
    }
    const bool sounding = pool[F_ENV_STAGE * a.cap + i] != ENV_DEAD;
    key = steal_key(a, pool, i);
    return sounding;
}

// Slots per thread fetched before any is used by the selection's passes.
constexpr uint32_t SEL_UNROLL = 4;

// ---------------------------------------------------------------------------
// Steal selection: the steal_k smallest keys, which are unique (the note id is
// in every key), so "key <= the k-th smallest" names exactly steal_k voices --
// the set src/cpu.rs:351-370 takes with select_nth_unstable. Should a caller
// repeat note ids, every voice whose key is at or below the k-th smallest is a
// victim (the oracle's rule), so a group tied at the threshold goes whole.
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long y = __shfl_xor(x, off);
        x = y < x ? y : x;
    }
    return x;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long x) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long y = __shfl_xor(x, off);
        x = y > x ? y : x;
    }
    return x;
}

__device__ __forceinline__ bool sel_match(unsigned long long key, unsigned long long prefix,
                                          uint32_t top) {
    return top >= 64 || (key >> top) == (prefix >> top);
}

// The window of the next digit: bits [lo, top).
__device__ __forceinline__ uint32_t sel_lo(uint32_t top) {
    return top > SEL_BITS ? top - SEL_BITS : 0;
}

// Count, smallest and largest key per bin of the candidates' next digit, into
// per-bin LDS slots (cnt[], mn[], mx[]). Lanes sharing the wave's first digit
// are folded into one update (neighbouring voices mostly do).
__device__ __forceinline__ void sel_accumulate(bool match, unsigned long long key, uint32_t d,
                                               uint32_t *cnt, unsigned long long *mn,
                                               unsigned long long *mx) {
    const uint64_t mm = __ballot(match);
    if (!mm) return;
    const uint32_t first = __ffsll((long long)mm) - 1;
    const uint32_t d0 = __shfl(d, first);
    const bool same = match && d == d0;
    const uint64_t ms = __ballot(same);
    const unsigned long long lo = wave_min_u64(same ? key : ~0ull);
    const unsigned long long hi = wave_max_u64(same ? key : 0ull);
    if ((threadIdx.x & 63) == first) {
        atomicAdd(&cnt[d0], (uint32_t)__popcll(ms));
        atomicMin(&mn[d0], lo);
        atomicMax(&mx[d0], hi);
    }
    if (match && !same) {
        atomicAdd(&cnt[d], 1u);
        atomicMin(&mn[d], key);
        atomicMax(&mx[d], key);
    }
}

// Pick the bin holding the k-th key from `cnt/mn/mx` (256 entries, one per
// thread of the calling workgroup) and narrow the state to it.
__device__ void sel_choose(DevState *st, const uint32_t *cnt,
                           const unsigned long long *mn, const unsigned long long *mx,
                           uint32_t *scan) {
    const uint32_t t = threadIdx.x;
    const uint32_t c = t < SEL_BINS ? cnt[t] : 0;
    if (t < SEL_BINS) scan[t] = c;
    __syncthreads();
    for (uint32_t off = 1; off < SEL_BINS; off <<= 1) {
        uint32_t x = (t < SEL_BINS && t >= off) ? scan[t - off] : 0;
        __syncthreads();
        if (t < SEL_BINS) scan[t] += x;
        __syncthreads();
    }
    const uint32_t k = st->sel_k;
    if (t < SEL_BINS) {
        const uint32_t excl = scan[t] - c;
        if (excl < k && excl + c >= k) {
            // Keys of the bin lie in [mn, mx]: they share every bit above the
            // highest one in which those two differ.
            const unsigned long long diff = mn[t] ^ mx[t];
            const uint32_t top = diff ? 64 - __clzll(diff) : 0;
            st->sel_prefix = top >= 64 ? 0ull : (mn[t] >> top) << top;
            st->sel_top = top;
            st->sel_k = k - excl;
            st->sel_count = c;
            st->sel_done = c <= SEL_CAP || top == 0;
        }
    }
    __syncthreads();
}

// One MSB-first round over the pool: per-workgroup bin statistics of the
// candidates' next digit, stored without atomics.
__device__ __forceinline__ void sel_hist_body(const Args &a) {
    DevState *st = a.st;
    if (st->sel_steal == 0 || st->sel_done) return;
    __shared__ uint32_t cnt[SEL_BINS];
    __shared__ unsigned long long mn[SEL_BINS], mx[SEL_BINS];
    const uint32_t top = st->sel_top;
    const unsigned long long prefix = st->sel_prefix;
    const uint32_t lo = sel_lo(top);
    const uint32_t dmask = (1u << (top - lo)) - 1;
    for (uint32_t i = threadIdx.x; i < SEL_BINS; i += blockDim.x) {
        cnt[i] = 0;
        mn[i] = ~0ull;
        mx[i] = 0;
    }
    __syncthreads();
    const uint32_t *pool = sel2(a.pools, st->sel_parity);
    const uint32_t n = st->sel_n;
    const uint32_t lim = (n + 63) & ~63u;  // whole waves, so ballots see every lane
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < lim; i0 += SEL_UNROLL * stride) {
        unsigned long long key[SEL_UNROLL];
        bool match[SEL_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < SEL_UNROLL; ++u)
            match[u] = cand_at(a, pool, i0 + u * stride, n, key[u]);
#pragma unroll
        for (uint32_t u = 0; u < SEL_UNROLL; ++u) {
            // Lanes of a wave agree on i < lim (lim and stride are whole waves).
            if (i0 + u * stride >= lim) break;
            const bool m = match[u] && sel_match(key[u], prefix, top);
            sel_accumulate(m, key[u], (uint32_t)(key[u] >> lo) & dmask, cnt, mn, mx);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < SEL_BINS; i += blockDim.x) {
        const size_t o = (size_t)blockIdx.x * SEL_BINS + i;
        a.sel_cnt[o] = cnt[i];
        a.sel_min[o] = mn[i];
        a.sel_max[o] = mx[i];
    }
}
__global__ void __launch_bounds__(SEL_THREADS) k_sel_hist(Args a) { sel_hist_body(a); }
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void __launch_bounds__(SEL_THREADS) kb_sel_hist(const Args *__restrict__ aa) {
    sel_hist_body(aa[blockIdx.z]);
}

// One workgroup sums the per-workgroup counts (four threads per bin, each over
// a quarter of the rows) and picks the bin; only that bin's smallest and
// largest keys are then read back from every row.
__device__ __forceinline__ void sel_refine_body(const Args &a, uint32_t blocks) {
    DevState *st = a.st;
    if (st->sel_steal == 0 || st->sel_done) return;
    constexpr uint32_t QUADS = SEL_BINS / 4;          // 16-byte loads per row
    constexpr uint32_t G = SEL_THREADS / QUADS;       // rows read at once
    constexpr uint32_t WAVES = SEL_THREADS / 64;
    __shared__ uint32_t part[G][SEL_BINS], cnt[SEL_BINS];
    __shared__ unsigned long long red_lo[WAVES], red_hi[WAVES];
    __shared__ uint32_t s_bin, s_excl, s_k;
    const uint32_t t = threadIdx.x;
    // Thread t sums four neighbouring bins over every G-th row: all of its
    // loads are independent, so the counts arrive in one memory round trip.
    {
        const uint32_t b4 = (t % QUADS) * 4, r0 = t / QUADS;
        uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma unroll 8
        for (uint32_t r = r0; r < blocks; r += G) {
            const uint4 x = *reinterpret_cast<const uint4 *>(a.sel_cnt + (size_t)r * SEL_BINS + b4);
            c0 += x.x;
            c1 += x.y;
            c2 += x.z;
            c3 += x.w;
        }
        part[r0][b4] = c0;
        part[r0][b4 + 1] = c1;
        part[r0][b4 + 2] = c2;
        part[r0][b4 + 3] = c3;
    }
    __syncthreads();
    if (t < SEL_BINS) {
        uint32_t c = 0;
#pragma unroll
        for (uint32_t g = 0; g < G; ++g) c += part[g][t];
        cnt[t] = c;
    }
    __syncthreads();
    if (t < 64) {
        // One wave scans the bins (four per lane) and picks the k-th key's.
        uint32_t c[4], sum = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            c[b] = cnt[4 * t + b];
            sum += c[b];
        }
        uint32_t incl = sum;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if (t >= off) incl += y;
        }
        uint32_t k = st->sel_k;
        if (st->sel_steal == SEL_PENDING) {
            // The first round over the render's voices: the bins hold every
            // sounding one, so their total is the live count.
            const uint32_t total = __shfl(incl, 63);
            k = steal_count(a, total, st->sel_want);
            if (t == 0) {
                st->sel_steal = k;
                st->sel_k = k > 0 ? k : 1;
            }
        }
        if (t == 0) {
            s_k = k;
            s_bin = SEL_BINS;
        }
        uint32_t run = incl - sum;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (k > 0 && run < k && run + c[b] >= k) {
                s_bin = 4 * t + b;
                s_excl = run;
            }
            run += c[b];
        }
    }
    __syncthreads();
    const uint32_t chosen = s_bin;
    if (chosen >= SEL_BINS) return;  // nothing to steal (or k outside the counts: cannot happen)
    unsigned long long lo = ~0ull, hi = 0;
    for (uint32_t b = t; b < blocks; b += SEL_THREADS) {
        const unsigned long long x = a.sel_min[(size_t)b * SEL_BINS + chosen];
        const unsigned long long y = a.sel_max[(size_t)b * SEL_BINS + chosen];
        lo = x < lo ? x : lo;
        hi = y > hi ? y : hi;
    }
    lo = wave_min_u64(lo);
    hi = wave_max_u64(hi);
    if ((t & 63) == 0) {
        red_lo[t >> 6] = lo;
        red_hi[t >> 6] = hi;
    }
    __syncthreads();
    if (t == 0) {
        for (uint32_t w = 1; w < WAVES; ++w) {
            lo = red_lo[w] < lo ? red_lo[w] : lo;
            hi = red_hi[w] > hi ? red_hi[w] : hi;
        }
        // The bin's keys lie in [lo, hi]: they share every bit above the
        // highest one in which those two differ.
        const unsigned long long diff = lo ^ hi;
        const uint32_t top = diff ? 64 - __clzll(diff) : 0;
        const uint32_t n = cnt[chosen];
        st->sel_prefix = top >= 64 ? 0ull : (lo >> top) << top;
        st->sel_top = top;
        st->sel_k = s_k - s_excl;
        st->sel_count = n;
        st->sel_done = n <= SEL_CAP || top == 0;
    }
}

__global__ void __launch_bounds__(SEL_THREADS) k_sel_refine(Args a, uint32_t blocks) { sel_refine_body(a, blocks); }
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void __launch_bounds__(SEL_THREADS) kb_sel_refine(const Args *__restrict__ aa, uint32_t blocks) {
    sel_refine_body(aa[blockIdx.z], blocks);
}

// Slot i's voice is a victim: its stop frame from the low word of its note id
// (src/cpu.rs:365-369 / shaders/compact.wgsl:174-175). Under either rule that
// word is the steal key's low word, so callers pass it from the key rather
// than loading it again.
__device__ __forceinline__ void mark_victim(const Args &a, uint32_t *pool, uint32_t i,
                                            uint32_t note_lo) {
    pool[F_STOP_REL * a.cap + i] = note_lo % a.steal_span + 1;
}

// Every key below the candidates is a victim; the candidates are collected
// (one global reservation per workgroup).
__device__ __forceinline__ void sel_collect_body(const Args &a) {
    DevState *st = a.st;
    if (st->sel_steal == 0) return;
    __shared__ uint32_t s_n, s_base;
    uint32_t *pool = sel2(a.pools, st->sel_parity);
    const uint32_t n = st->sel_n;
    const uint32_t top = st->sel_top;
    const unsigned long long prefix = st->sel_prefix;
    const unsigned long long pb = top >= 64 ? 0ull : prefix >> top;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += SEL_UNROLL * stride) {
        unsigned long long key[SEL_UNROLL];
        bool ok[SEL_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < SEL_UNROLL; ++u) ok[u] = cand_at(a, pool, i0 + u * stride, n, key[u]);
#pragma unroll
        for (uint32_t u = 0; u < SEL_UNROLL; ++u) {
            if (!ok[u]) continue;
            const unsigned long long kb = top >= 64 ? 0ull : key[u] >> top;
            if (kb < pb) mark_victim(a, pool, i0 + u * stride, (uint32_t)key[u]);
            else if (kb == pb) atomicAdd(&s_n, 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_n ? atomicAdd(&st->cand_n, s_n) : 0;
    __syncthreads();
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += SEL_UNROLL * stride) {
        unsigned long long key[SEL_UNROLL];
        bool ok[SEL_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < SEL_UNROLL; ++u) ok[u] = cand_at(a, pool, i0 + u * stride, n, key[u]);
#pragma unroll
        for (uint32_t u = 0; u < SEL_UNROLL; ++u) {
            if (!ok[u] || (top >= 64 ? 0ull : key[u] >> top) != pb) continue;
            const uint32_t at = s_base + atomicAdd(&s_n, 1u);
            a.cand_key[at] = key[u];
            a.cand_slot[at] = i0 + u * stride;
        }
    }
}
__global__ void __launch_bounds__(SEL_THREADS) k_sel_collect(Args a) { sel_collect_body(a); }
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void __launch_bounds__(SEL_THREADS) kb_sel_collect(const Args *__restrict__ aa) {
    sel_collect_body(aa[blockIdx.z]);
}

// The final rounds when the candidates fit in registers: at most SEL_CAP keys
// (16 per thread) that agree on every bit from 32 up, so only their low words
// are held. Each round histograms the active keys' next digit in LDS, one wave
// scans the bins and picks the one holding the k-th key, keys in lower bins
// become victims and keys in higher bins drop out -- the choices of the
// general rounds below, without a trip through global memory per round.
__device__ void sel_final_regs(const Args &a, uint32_t n) {
    DevState *st = a.st;
    constexpr uint32_t PER = SEL_CAP / SEL_THREADS;
    __shared__ uint32_t cnt[SEL_BINS], mn[SEL_BINS], mx[SEL_BINS];
    __shared__ uint32_t s_bin, s_excl;
    uint32_t *pool = sel2(a.pools, st->sel_parity);
    const uint32_t t = threadIdx.x;
    // Key j of this thread is candidate j * SEL_THREADS + t; its slot is read
    // back only when it turns out a victim.
    uint32_t key[PER];
    uint32_t act = 0;  // bit j: key[j] is still a candidate
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t i = j * SEL_THREADS + t;
        const bool in = i < n;
        act |= in ? 1u << j : 0u;
        key[j] = in ? (uint32_t)a.cand_key[i] : 0u;
    }
    uint32_t top = st->sel_top, k = st->sel_k, count = n;
    // Each round lowers top by at least SEL_BITS; unique keys leave one, keys
    // that tie leave their group (same: every active key is the same value).
    bool same = top == 0;
    for (uint32_t round = 0; round <= 32 / SEL_BITS && count > 1 && !same; ++round) {
        const uint32_t lo = sel_lo(top);
        const uint32_t dmask = (1u << (top - lo)) - 1;
        if (t < SEL_BINS) {
            cnt[t] = 0;
            mn[t] = ~0u;
            mx[t] = 0;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            if (!((act >> j) & 1u)) continue;
            const uint32_t d = (key[j] >> lo) & dmask;
            atomicAdd(&cnt[d], 1u);
            atomicMin(&mn[d], key[j]);
            atomicMax(&mx[d], key[j]);
        }
        __syncthreads();
        if (t < 64) {
            // Four bins per lane, an inclusive scan across the wave.
            uint32_t c[4], sum = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                c[b] = cnt[4 * t + b];
                sum += c[b];
            }
            uint32_t incl = sum;
            for (uint32_t off = 1; off < 64; off <<= 1) {
                const uint32_t y = __shfl_up(incl, off);
                if (t >= off) incl += y;
            }
            uint32_t run = incl - sum;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if (run < k && run + c[b] >= k) {
                    s_bin = 4 * t + b;
                    s_excl = run;
                }
                run += c[b];
            }
        }
        __syncthreads();
        const uint32_t bin = s_bin;
        k -= s_excl;
        count = cnt[bin];
        const uint32_t diff = mn[bin] ^ mx[bin];
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            if (!((act >> j) & 1u)) continue;
            const uint32_t d = (key[j] >> lo) & dmask;
            if (d < bin) mark_victim(a, pool, a.cand_slot[j * SEL_THREADS + t], key[j]);
            if (d != bin) act &= ~(1u << j);
        }
        top = diff ? 32 - __clz(diff) : 0;
        same = diff == 0;
        __syncthreads();  // every read of this round's bins before the next clears them
    }
    // The keys left are the k-th smallest: one, or a group of equal keys (note
    // ids the caller repeated), every one of which is at the threshold and a
    // victim, as every key at or below the threshold is in the oracle.
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j)
        if (((act >> j) & 1u) && (count == 1 || same))
            mark_victim(a, pool, a.cand_slot[j * SEL_THREADS + t], key[j]);
}

// Finish on the candidates with more rounds in one workgroup until the bin
// holds the k-th key alone; keys below each chosen bin become victims as the
// candidate list is narrowed, and the last one standing is the k-th.
__device__ __forceinline__ void sel_final_body(const Args &a) {
    DevState *st = a.st;
    if (st->sel_steal == 0) return;
    if (st->sel_steal == SEL_PENDING) {
        // No round ran: every sounding voice was collected.
        __syncthreads();  // every thread has read sel_steal before it changes
        if (threadIdx.x == 0) {
            resolve_steal(a, st, st->cand_n);
            st->sel_count = st->cand_n;
        }
        __syncthreads();
        if (st->sel_steal == 0) return;
    }
    if (st->cand_n <= SEL_CAP && st->sel_top <= 32 && blockDim.x == SEL_THREADS) {
        sel_final_regs(a, st->cand_n);
        return;
    }
    __shared__ uint32_t cnt[SEL_BINS], scan[SEL_BINS];
    __shared__ unsigned long long mn[SEL_BINS], mx[SEL_BINS];
    __shared__ uint32_t s_n;
    uint32_t *pool = sel2(a.pools, st->sel_parity);
    uint32_t n = st->cand_n;
    // Each round lowers sel_top by at least SEL_BITS (the chosen bin's keys
    // agree on every bit above the digit), and unique keys leave one
    // candidate when it reaches 0 (equal keys leave their group); the bound
    // only guards a broken state.
    for (uint32_t round = 0; round <= 64 / SEL_BITS && st->sel_count > 1 && st->sel_top > 0;
         ++round) {
        const uint32_t top = st->sel_top;
        const uint32_t lo = sel_lo(top);
        const uint32_t dmask = (1u << (top - lo)) - 1;
        for (uint32_t i = threadIdx.x; i < SEL_BINS; i += blockDim.x) {
            cnt[i] = 0;
            mn[i] = ~0ull;
            mx[i] = 0;
        }
        __syncthreads();
        const uint32_t lim = (n + 63) & ~63u;
        for (uint32_t i = threadIdx.x; i < lim; i += blockDim.x) {
            const bool in = i < n;
            const unsigned long long k = in ? a.cand_key[i] : 0ull;
            sel_accumulate(in, k, in ? (uint32_t)(k >> lo) & dmask : 0u, cnt, mn, mx);
        }
        __syncthreads();
        sel_choose(st, cnt, mn, mx, scan);
        // Keep the chosen bin's candidates; those below it are victims.
        const uint32_t ntop = st->sel_top;
        const unsigned long long pb = ntop >= 64 ? 0ull : st->sel_prefix >> ntop;
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        for (uint32_t base = 0; base < n; base += blockDim.x) {
            const uint32_t i = base + threadIdx.x;
            unsigned long long k = 0;
            uint32_t sl = 0;
            bool keep = false;
            if (i < n) {
                k = a.cand_key[i];
                sl = a.cand_slot[i];
                const unsigned long long kb = ntop >= 64 ? 0ull : k >> ntop;
                if (kb < pb) mark_victim(a, pool, sl, (uint32_t)k);
                keep = kb == pb;
            }
            __syncthreads();  // every read of this chunk before any write
            if (keep) {
                const uint32_t at = atomicAdd(&s_n, 1u);
                a.cand_key[at] = k;
                a.cand_slot[at] = sl;
            }
            __syncthreads();
        }
        n = s_n;
        __syncthreads();
    }
    // The keys left are the k-th smallest: one, or a group of equal keys, all
    // at the threshold and victims (every key at or below it is one).
    if (st->sel_count <= 1 || st->sel_top == 0)
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
            mark_victim(a, pool, a.cand_slot[i], (uint32_t)a.cand_key[i]);
}
// The selection's last kernel: in a staged run it then publishes that the
// victims are marked (every thread's stores are in before thread 0's release).
__global__ void __launch_bounds__(SEL_THREADS) k_sel_final(Args a) {
    sel_final_body(a);
    if (a.sig_sel) {
        __syncthreads();
        if (threadIdx.x == 0) signal_seq(&a.st->sel_seq, a.sig_sel);
    }
}
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void __launch_bounds__(SEL_THREADS) kb_sel_final(const Args *__restrict__ aa) {
    sel_final_body(aa[blockIdx.z]);
}

// shaders/spawn.wgsl:35-78; thinning per src/voice.rs:72-75 when the block is
// oversubscribed under AdmitRule::Even (src/gpu/mod.rs:1056-1071).
// Which batches of the live pool hold a steal victim, for the render's
// dispatch order: batch j (voices 64j .. 64j + 63, all survivors) when some
// voice's stop_rel is set, read where the render reads it (slot perm[vi] of
// the live pool). One wave per batch, grid-stride.
__device__ __forceinline__ void victim_flags(const Args &a, const DevState *st) {
    const uint32_t nsurv = st->live / 64;
    const uint32_t *pool = sel2(a.pools, st->parity);
    const uint32_t *perm = live_perm(a);
    const uint32_t stride = gridDim.x * blockDim.x;  // a multiple of 64
    for (uint32_t vi = blockIdx.x * blockDim.x + threadIdx.x; vi < nsurv * 64; vi += stride) {
        const uint32_t slot = perm ? perm[vi] : vi;
        const bool victim = __any(pool[F_STOP_REL * a.cap + slot] != 0);
        if ((threadIdx.x & 63) == 0) a.batch_victim[vi / 64] = victim ? 1u : 0u;
    }
}

__device__ __forceinline__ void spawn_body(const Args &a) {
    DevState *st = a.st;
    if (a.batch_victim) victim_flags(a, st);
    const uint32_t n = st->spawn_count;
    if (n == 0) return;
    const uint32_t want = st->want;
    const bool thin = a.admit_rule == ADMIT_EVEN && n < want;
    const uint32_t base = st->live;
    const kst_spawn_cmd *cmds = a.cmds + st->cmd_off;
    // A fused block renders into the other buffer; its spawns go there directly.
    uint32_t *pool = sel2(a.pools, st->parity ^ a.fused);
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
__global__ void __launch_bounds__(256) k_spawn(Args a) { spawn_body(a); }
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void __launch_bounds__(256) kb_spawn(const Args *__restrict__ aa) {
    spawn_body(aa[blockIdx.z]);
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
// pass's output). Each tile's keys are loaded up front; counts go to LDS and
// then to the tile's row of sort_hist (tile-major: a tile's digits are
// adjacent), with no global atomics.
__device__ __forceinline__ void sort_hist_body(const Args &a, uint32_t p) {
    if (p >= a.sort_passes) return;  // a batch runs the most passes any member needs
    DevState *st = a.st;
    const uint32_t n = st->live_total;
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint32_t radix = 1u << a.digit_bits, mask = radix - 1;
    const uint32_t shift = p * a.digit_bits;
    __shared__ uint32_t sh[SORT_MAX_RADIX];
    const uint32_t *keys = sel2(a.keys, p & 1);
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        uint32_t kk[SORT_IPT];
#pragma unroll
        for (uint32_t k = 0; k < SORT_IPT; ++k) {
            const uint32_t item = t * SORT_TILE + k * SORT_TPB + threadIdx.x;
            kk[k] = item < n ? keys[item] : 0u;
        }
        for (uint32_t i = threadIdx.x; i < radix; i += SORT_TPB) sh[i] = 0;
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < SORT_IPT; ++k) {
            const uint32_t item = t * SORT_TILE + k * SORT_TPB + threadIdx.x;
            if (item < n) atomicAdd(&sh[(kk[k] >> shift) & mask], 1u);
        }
        __syncthreads();
        for (uint32_t d = threadIdx.x; d < radix; d += SORT_TPB)
            a.sort_hist[((size_t)p * a.tiles_max + t) * radix + d] = sh[d];
        __syncthreads();
    }
}
__global__ void __launch_bounds__(SORT_TPB) k_sort_hist(Args a, uint32_t p) { sort_hist_body(a, p); }
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void __launch_bounds__(SORT_TPB) kb_sort_hist(const Args *__restrict__ aa, uint32_t p) {
    sort_hist_body(aa[blockIdx.z], p);
}

// Exclusive scan of `v` (one value per thread of a 256-thread workgroup):
// each wave scans its 64 values with lane shuffles, the four wave totals meet
// in `wtot` behind one barrier. Returns the exclusive prefix and leaves the
// total in *total. `wtot` may be reused only after another barrier.
__device__ __forceinline__ uint32_t block_exscan256(uint32_t v, uint32_t *wtot, uint32_t *total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < SORT_TPB / 64; ++w) {
        const uint32_t x = wtot[w];
        before += w < wave ? x : 0u;
        all += x;
    }
    *total = all;
    return before + incl - v;
}

// Each tile's offset within a digit is the same digit's count in the tiles
// before it; the digit's total goes to sort_tot for the scatter to place the
// digits. A workgroup takes SCAN_DIGITS adjacent digits (a half-wave reads a
// tile's 32 adjacent counts at once) and splits its tiles into SCAN_LANES
// contiguous ranges, one per half-wave: each sums its range, the ranges' sums
// are scanned in LDS, and each range's offsets are written from its start.
// Its tiles are all of them, or (k_sort_scan_grouped, pools of many tiles)
// group blockIdx.y of SCAN_GROUP_TILES, whose digit counts k_sort_scan_sums
// has summed: the groups before it give its start.
constexpr uint32_t SCAN_DIGITS = 32, SCAN_LANES = 32;

// Tiles [g0, g1) of this workgroup: all, or its group.
__device__ __forceinline__ void scan_group_range(bool grouped, uint32_t tiles, uint32_t &g0,
                                                 uint32_t &g1) {
    if (!grouped) {
        g0 = 0;
        g1 = tiles;
        return;
    }
    g0 = blockIdx.y * SCAN_GROUP_TILES < tiles ? blockIdx.y * SCAN_GROUP_TILES : tiles;
    g1 = g0 + SCAN_GROUP_TILES < tiles ? g0 + SCAN_GROUP_TILES : tiles;
}

// Groups holding a tile (at least one, so some group writes the totals).
__device__ __forceinline__ uint32_t scan_groups_used(uint32_t tiles) {
    const uint32_t g = (tiles + SCAN_GROUP_TILES - 1) / SCAN_GROUP_TILES;
    return g > 0 ? g : 1u;
}

// Grouped scan, first kernel: each group's count of each digit.
__global__ void __launch_bounds__(1024) k_sort_scan_sums(Args a, uint32_t p) {
    const uint32_t radix = 1u << a.digit_bits;
    if (p >= a.sort_passes || blockIdx.x * SCAN_DIGITS >= radix) return;
    const uint32_t n = a.st->live_total;
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    if (blockIdx.y >= scan_groups_used(tiles)) return;
    uint32_t g0, g1;
    scan_group_range(true, tiles, g0, g1);
    const uint32_t col = threadIdx.x % SCAN_DIGITS, w = threadIdx.x / SCAN_DIGITS;
    const uint32_t d = blockIdx.x * SCAN_DIGITS + col;
    const uint32_t per = (g1 - g0 + SCAN_LANES - 1) / SCAN_LANES;
    const uint32_t t0 = g0 + w * per < g1 ? g0 + w * per : g1;
    const uint32_t t1 = t0 + per < g1 ? t0 + per : g1;
    const uint32_t *h = a.sort_hist + (size_t)p * a.tiles_max * radix;
    __shared__ uint32_t part[SCAN_LANES][SCAN_DIGITS];
    uint32_t sum = 0;
    if (d < radix) {
#pragma unroll 8
        for (uint32_t t = t0; t < t1; ++t) sum += h[(size_t)t * radix + d];
    }
    part[w][col] = sum;
    __syncthreads();
    if (w == 0 && d < radix) {
        uint32_t tot = 0;
        for (uint32_t l = 0; l < SCAN_LANES; ++l) tot += part[l][col];
        a.scan_grp[(size_t)blockIdx.y * radix + d] = tot;
    }
}

template <bool GROUPED>
__device__ __forceinline__ void sort_scan_body(const Args &a, uint32_t p) {
    const uint32_t radix = 1u << a.digit_bits;
    if (p >= a.sort_passes || blockIdx.x * SCAN_DIGITS >= radix) return;
    DevState *st = a.st;
    const uint32_t n = st->live_total;
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint32_t used = GROUPED ? scan_groups_used(tiles) : 1u;
    if (blockIdx.y >= used) return;
    uint32_t g0, g1;
    scan_group_range(GROUPED, tiles, g0, g1);
    const uint32_t col = threadIdx.x % SCAN_DIGITS, w = threadIdx.x / SCAN_DIGITS;
    const uint32_t d = blockIdx.x * SCAN_DIGITS + col;
    const uint32_t per = (g1 - g0 + SCAN_LANES - 1) / SCAN_LANES;
    const uint32_t t0 = g0 + w * per < g1 ? g0 + w * per : g1;
    const uint32_t t1 = t0 + per < g1 ? t0 + per : g1;
    uint32_t *h = a.sort_hist + (size_t)p * a.tiles_max * radix;
    __shared__ uint32_t part[SCAN_LANES][SCAN_DIGITS];
    __shared__ uint32_t before[GROUPED ? SCAN_LANES : 1][SCAN_DIGITS];
    if (GROUPED) {
        // The groups before this one: lane w sums groups w, w + 32, ...
        uint32_t pre = 0;
        if (d < radix)
            for (uint32_t q = w; q < blockIdx.y; q += SCAN_LANES)
                pre += a.scan_grp[(size_t)q * radix + d];
        before[GROUPED ? w : 0][col] = pre;
    }
    uint32_t sum = 0;
    if (d < radix) {
#pragma unroll 8
        for (uint32_t t = t0; t < t1; ++t) sum += h[(size_t)t * radix + d];
    }
    part[w][col] = sum;
    __syncthreads();
    uint32_t run = 0;
    if (GROUPED)
        for (uint32_t l = 0; l < SCAN_LANES; ++l) run += before[GROUPED ? l : 0][col];
    for (uint32_t l = 0; l < w; ++l) run += part[l][col];
    if (d < radix) {
#pragma unroll 8
        for (uint32_t t = t0; t < t1; ++t) {
            const uint32_t v = h[(size_t)t * radix + d];
            h[(size_t)t * radix + d] = run;
            run += v;
        }
        if (w == SCAN_LANES - 1 && blockIdx.y == used - 1) a.sort_tot[p * radix + d] = run;
    }
}
__global__ void __launch_bounds__(1024) k_sort_scan(Args a, uint32_t p) {
    sort_scan_body<false>(a, p);
}
__global__ void __launch_bounds__(1024) k_sort_scan_grouped(Args a, uint32_t p) {
    sort_scan_body<true>(a, p);
}
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void __launch_bounds__(1024) kb_sort_scan(const Args *__restrict__ aa, uint32_t p) {
    sort_scan_body<false>(aa[blockIdx.z], p);
}

// Stable scatter of one digit pass. Each wave ranks its own contiguous quarter
// of the tile (512 items, 8 rounds of 64 in index order) against a per-wave
// digit table, with match masks built from ballots and no workgroup barrier
// between rounds. The four tables then become per-wave offsets within each
// digit, the tile's digit counts are scanned into local bases, the ranked tile
// is placed in LDS in digit order (a stable local sort) and written out by
// consecutive threads, so each digit's run within the tile leaves in adjacent
// addresses. Each XCD takes a contiguous range of tiles: the runs of one digit
// from neighbouring tiles are adjacent in the output, so their cache lines
// fill within one L2.
__device__ __forceinline__ void sort_scatter_body(const Args &a, uint32_t p) {
    if (p >= a.sort_passes) return;  // a batch runs the most passes any member needs
    DevState *st = a.st;
    const uint32_t n = st->live_total;
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint32_t radix = 1u << a.digit_bits, mask = radix - 1;
    const uint32_t shift = p * a.digit_bits;
    const uint32_t *kin = sel2(a.keys, p & 1);
    const uint32_t *iin = sel2(a.idx, p & 1);
    uint32_t *kout = sel2(a.keys, (p + 1) & 1);
    uint32_t *iout = sel2(a.idx, (p + 1) & 1);
    constexpr uint32_t WAVES = SORT_TPB / 64;
    constexpr uint32_t PER = SORT_MAX_RADIX / SORT_TPB;  // digits per thread
    constexpr uint32_t CHUNK = SORT_TILE / WAVES;        // items per wave
    static_assert(CHUNK / 64 == SORT_IPT, "a wave ranks SORT_IPT rounds of 64 items");
    __shared__ uint32_t wcnt[WAVES][SORT_MAX_RADIX];  // per-wave counts, then offsets
    __shared__ uint32_t wtot[WAVES];
    __shared__ uint32_t gbase[SORT_MAX_RADIX];  // the tile's first output slot per digit
    __shared__ uint32_t lbase[SORT_MAX_RADIX];  // the digit's first slot in the local order
    __shared__ uint32_t lkey[SORT_TILE], lidx[SORT_TILE];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    // Digit bases: exclusive scan of this pass's digit totals; thread i holds
    // digits i * PER .. i * PER + PER - 1.
    uint32_t dbase[PER];
    {
        uint32_t v[PER], sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint32_t d = threadIdx.x * PER + j;
            v[j] = d < radix ? a.sort_tot[p * radix + d] : 0u;
            sum += v[j];
        }
        uint32_t total;
        uint32_t run = block_exscan256(sum, wtot, &total);
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            dbase[j] = run;
            run += v[j];
        }
    }
    // Fused: survivors sort before every dead key, so when the dead keys begin
    // at a digit boundary of the last pass, their count is that digit's base.
    if (a.fused && p + 1 == a.sort_passes && a.dead_digit != ~0u && blockIdx.x == 0) {
        if (a.dead_digit >= radix) {
            if (threadIdx.x == 0) st->live_new = n;
        } else if (a.dead_digit / PER == threadIdx.x) {
#pragma unroll
            for (uint32_t j = 0; j < PER; ++j)
                if (a.dead_digit % PER == j) st->live_new = dbase[j];
        }
    }
    for (uint32_t id = blockIdx.x; id < tiles; id += gridDim.x) {
        const uint32_t t = xcd_group(id, tiles);
        const uint32_t t0 = t * SORT_TILE;
        __syncthreads();  // the previous tile's readers of the LDS tables are done
        uint32_t kk[SORT_IPT], ii[SORT_IPT], rk[SORT_IPT], hv[PER];
        const uint32_t i0 = t0 + wave * CHUNK + lane;
#pragma unroll
        for (uint32_t k = 0; k < SORT_IPT; ++k) {
            const uint32_t item = i0 + k * 64;
            const bool valid = item < n;
            kk[k] = valid ? kin[item] : 0u;
            ii[k] = valid ? (p == 0 ? item : iin[item]) : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint32_t d = threadIdx.x * PER + j;
            hv[j] = d < radix ? a.sort_hist[((size_t)p * a.tiles_max + t) * radix + d] : 0u;
        }
        for (uint32_t d = lane; d < SORT_MAX_RADIX; d += 64) wcnt[wave][d] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Ranks within the wave's quarter: the lanes of one digit read its
        // running count, the lowest of them advances it by their number.
#pragma unroll
        for (uint32_t k = 0; k < SORT_IPT; ++k) {
            const bool valid = i0 + k * 64 < n;
            const uint32_t digit = (kk[k] >> shift) & mask;
            uint64_t m = __ballot(valid);
            if (a.digit_bits == SORT_MAX_BITS) {  // the usual width, unrolled
#pragma unroll
                for (uint32_t bit = 0; bit < SORT_MAX_BITS; ++bit) {
                    const uint64_t bb = __ballot(valid && ((digit >> bit) & 1));
                    m &= ((digit >> bit) & 1) ? bb : ~bb;
                }
            } else {
                for (uint32_t bit = 0; bit < a.digit_bits; ++bit) {
                    const uint64_t bb = __ballot(valid && ((digit >> bit) & 1));
                    m &= ((digit >> bit) & 1) ? bb : ~bb;
                }
            }
            const uint32_t below = __popcll(m & lt);
            const uint32_t pre = valid ? wcnt[wave][digit] : 0u;
            rk[k] = pre + below;
            if (valid && below == 0) wcnt[wave][digit] = pre + (uint32_t)__popcll(m);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        __syncthreads();
        // Per digit: each wave's offset (the earlier waves' counts), the
        // tile's count, and its local and global bases.
        uint32_t c[PER], sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint32_t d = threadIdx.x * PER + j;
            uint32_t run = 0;
#pragma unroll
            for (uint32_t w = 0; w < WAVES; ++w) {
                const uint32_t x = wcnt[w][d];
                wcnt[w][d] = run;
                run += x;
            }
            c[j] = run;
            sum += run;
        }
        uint32_t total;
        uint32_t run = block_exscan256(sum, wtot, &total);
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint32_t d = threadIdx.x * PER + j;
            lbase[d] = run;
            gbase[d] = dbase[j] + hv[j];
            run += c[j];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < SORT_IPT; ++k) {
            if (i0 + k * 64 < n) {
                const uint32_t d = (kk[k] >> shift) & mask;
                const uint32_t lpos = lbase[d] + wcnt[wave][d] + rk[k];
                lkey[lpos] = kk[k];
                lidx[lpos] = ii[k];
            }
        }
        __syncthreads();
        // Out in local order: slot j holds the (j - lbase[d])-th item of digit d.
        const uint32_t items = n - t0 < SORT_TILE ? n - t0 : SORT_TILE;
        for (uint32_t j = threadIdx.x; j < items; j += SORT_TPB) {
            const uint32_t key = lkey[j];
            const uint32_t d = (key >> shift) & mask;
            const uint32_t off = gbase[d] + (j - lbase[d]);
            kout[off] = key;
            iout[off] = lidx[j];
        }
    }
}
__global__ void __launch_bounds__(SORT_TPB) k_sort_scatter(Args a, uint32_t p) { sort_scatter_body(a, p); }
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void __launch_bounds__(SORT_TPB) kb_sort_scatter(const Args *__restrict__ aa, uint32_t p) {
    sort_scatter_body(aa[blockIdx.z], p);
}

// shaders/sort.wgsl:376-400 / compact.wgsl:116-146: survivors move to the other
// buffer in sorted (or original) order; per-block fields are cleared. Dead
// voices sort after every survivor, so the last survivor's thread records the
// new live count.
__device__ __forceinline__ void gather_body(const Args &a) {
    DevState *st = a.st;
    const uint32_t n = st->live_total;
    const uint32_t *perm = sel2(a.idx, a.sort_passes & 1);
    const uint32_t *skey = sel2(a.keys, a.sort_passes & 1);
    const uint32_t *src = sel2(a.pools, st->parity);
    uint32_t *dst = sel2(a.pools, st->parity ^ 1);
    const uint32_t c = a.cap;
    // Each workgroup copies one contiguous run of the sorted order, and each
    // XCD a contiguous set of runs: the source slots a run reads are the
    // neighbourhood the previous block's order left them in, so they share
    // that XCD's L2.
    const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t j0 = xcd_group(blockIdx.x, gridDim.x) * per;
    const uint32_t j1 = j0 + per < n ? j0 + per : n;
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        const bool alive = key_alive(a, skey[j]);
        const bool next_alive = j + 1 < n && key_alive(a, skey[j + 1]);
        if (alive && !next_alive) st->live_new = j + 1;
        if (!alive) continue;
        const uint32_t s = perm[j];
        uint32_t v[VOICE_FIELDS];
#pragma unroll
        for (uint32_t f = 0; f < VOICE_FIELDS; ++f) v[f] = src[f * c + s];
        v[F_START_REL] = 0;
        v[F_BORN_VARIANT] = 0;
        v[F_STOP_REL] = 0;
#pragma unroll
        for (uint32_t f = 0; f < VOICE_FIELDS; ++f) dst[f * c + j] = v[f];
    }
}
__global__ void __launch_bounds__(256) k_gather(Args a) { gather_body(a); }
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void __launch_bounds__(256) kb_gather(const Args *__restrict__ aa) {
    gather_body(aa[blockIdx.z]);
}

// After a fused block the survivors are a prefix of the sorted keys (dead
// voices sort after every survivor), so their count is the first dead
// position: one workgroup finds the chunk holding it, then counts the
// survivors inside that chunk. Two rounds of reads, no global atomics.
__device__ uint32_t count_survivors(const Args &a, uint32_t n) {
    __shared__ uint32_t s_chunk, s_cnt;
    const uint32_t *skey = sel2(a.keys, a.sort_passes & 1);
    const uint32_t t = threadIdx.x, T = blockDim.x;
    if (t == 0) {
        s_chunk = T;
        s_cnt = 0;
    }
    __syncthreads();
    const uint32_t b0 = (uint32_t)((uint64_t)t * n / T);
    const uint32_t e0 = (uint32_t)((uint64_t)(t + 1) * n / T);
    if (e0 > b0 && !key_alive(a, skey[e0 - 1])) atomicMin(&s_chunk, t);
    __syncthreads();
    const uint32_t c = s_chunk;
    if (c == T) return n;  // every key alive (the same branch in every thread)
    const uint32_t lo = (uint32_t)((uint64_t)c * n / T);
    const uint32_t hi = (uint32_t)((uint64_t)(c + 1) * n / T);
    uint32_t k = 0;
    for (uint32_t p = lo + t; p < hi; p += T) k += key_alive(a, skey[p]) ? 1u : 0u;
    if (k) atomicAdd(&s_cnt, k);
    __syncthreads();
    return lo + s_cnt;
}

__device__ __forceinline__ void end_body(const Args &a) {
    DevState *st = a.st;
    if (st->idle) return;
    const uint32_t live_new = a.fused && a.dead_digit == ~0u ? count_survivors(a, st->live_total)
                                                           : st->live_new;
    if (threadIdx.x != 0) return;
    BlockResult r;
    r.active_voices = live_new;
    r.stolen = st->stolen;
    r.dropped = st->dropped;
    r.peak_bits = st->peak_bits;
    r.pad = 0;
    a.results[st->block] = r;
    st->live_new = live_new;
    st->live = live_new;
    st->parity ^= 1;
    st->perm_valid = a.fused;
    st->block += 1;
    // Begin the next block here when there is one (after the last, the state
    // keeps this block's counts for a settle's gather).
    if (a.end_begins && st->block < a.nblocks) begin_body(a);  // thread 0, as begin_body needs
}
// In a staged run the block ends once the next block's victims are marked.
__global__ void k_end(Args a) {
    if (a.wait_sel) {
        if (threadIdx.x == 0) wait_seq(a.st, &a.st->sel_seq, a.wait_sel);
        __syncthreads();
    }
    end_body(a);
}

// The per-block render's dispatch lists (one workgroup per XCD). The live
// pool's survivor batches [0, nsurv) are dealt to the XCDs in interleaved
// chunks (run_batch); the batches [nsurv, nb) that hold new voices are dealt
// to the XCDs in turn. XCD x renders, in order: its new-voice batches, its
// run's batches holding a steal victim, then its run's other batches, each
// class walked from the run's end. New voices start late and victims fade, so
// those waves run longest; started first, they no longer trail the launch.
// Every batch appears in exactly one list once.
__global__ void __launch_bounds__(1024) k_order(Args a) {
    const DevState *st = a.st;
    if (st->idle) return;
    const uint32_t live_total = st->live_total;
    const uint32_t nb = (live_total + 63) / 64;
    const uint32_t nsurv = st->live / 64 < nb ? st->live / 64 : nb;
    const uint32_t ns = nb - nsurv;
    const uint32_t x = blockIdx.x;
    const uint32_t n = run_len(nsurv, x);  // its run: run_batch(nsurv, x, 0 .. n - 1)
    const uint32_t nsp = ns > x ? (ns - 1 - x) / XCDS + 1 : 0;
    uint32_t *out = a.render_order + (size_t)x * a.order_stride;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = blockDim.x >> 6;
    for (uint32_t k = t; k < nsp; k += blockDim.x) out[k] = nsurv + x + k * XCDS;
    __shared__ uint32_t wsum[16], total;
    // How many of the run's batches hold a victim.
    uint32_t cnt = 0;
    for (uint32_t j = t; j < n; j += blockDim.x) cnt += a.batch_victim[run_batch(nsurv, x, j)] ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if (lane == 0) wsum[wave] = cnt;
    __syncthreads();
    if (t == 0) {
        uint32_t e = 0;
        for (uint32_t w = 0; w < nw; ++w) e += wsum[w];
        total = e;
    }
    __syncthreads();
    const uint32_t nvict = total;
    // Stable placement, chunk by chunk of the run walked from its end.
    uint32_t before_e = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += blockDim.x) {
        const uint32_t j = c0 + t;
        const bool in = j < n;
        const uint32_t bj = in ? run_batch(nsurv, x, n - 1 - j) : 0u;
        const bool e = in && a.batch_victim[bj] != 0;
        const uint64_t m = __ballot(e);
        const uint32_t below = __popcll(m & ((1ull << lane) - 1ull));
        __syncthreads();  // wsum of the previous chunk is consumed
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        uint32_t wbase = 0, ctotal = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            wbase += w < wave ? wsum[w] : 0u;
            ctotal += wsum[w];
        }
        if (in) {
            const uint32_t pe = before_e + wbase + below;  // victims before this one
            const uint32_t pos = e ? nsp + pe : nsp + nvict + (j - pe);
            out[pos] = bj;
        }
        before_e += ctotal;
    }
}

// A single submitted block's per-block inputs, from kernel arguments.
__global__ void k_stage_one(Args a, uint32_t chan_flags, uint32_t ncmd) {
    const_cast<uint32_t *>(a.chan_flags)[0] = chan_flags;
    const_cast<uint32_t *>(a.cmd_counts)[0] = ncmd;
    const_cast<uint32_t *>(a.cmd_offs)[0] = 0;
    a.st->block = 0;
}

// After the gather of a settle: the gathered buffer becomes the live pool, in
// order, with no permutation pending.
__global__ void k_settle(Args a) {
    DevState *st = a.st;
    st->live = st->live_new;
    st->parity ^= 1;
    st->perm_valid = 0;
}
// The same pass over a batch of renders: blockIdx.z picks the render.
__global__ void kb_end(const Args *__restrict__ aa) {
    end_body(aa[blockIdx.z]);
}

}  // namespace

void launch_begin(const Args &a, hipStream_t s) { k_begin<<<1, 256, 0, s>>>(a); }

void launch_select(const Args &a, uint32_t grid, uint32_t rounds, hipStream_t s) {
    k_sel_begin<<<1, 1, 0, s>>>(a);
    for (uint32_t r = 0; r < rounds; ++r) {
        k_sel_hist<<<grid, SEL_THREADS, 0, s>>>(a);
        k_sel_refine<<<1, SEL_THREADS, 0, s>>>(a, grid);
    }
    k_sel_collect<<<grid, SEL_THREADS, 0, s>>>(a);
    k_sel_final<<<1, SEL_THREADS, 0, s>>>(a);
}

void launch_spawn(const Args &a, uint32_t grid, hipStream_t s) {
    k_spawn<<<grid, 256, 0, s>>>(a);
}

void launch_order(const Args &a, hipStream_t s) { k_order<<<XCDS, 1024, 0, s>>>(a); }

void launch_sort(const Args &a, uint32_t grid, uint32_t gather_grid, hipStream_t s) {
    const uint32_t radix = 1u << a.digit_bits;
    for (uint32_t p = 0; p < a.sort_passes; ++p) {
        k_sort_hist<<<grid, SORT_TPB, 0, s>>>(a, p);
        const dim3 sg((radix + SCAN_DIGITS - 1) / SCAN_DIGITS, a.scan_groups ? a.scan_groups : 1);
        if (a.scan_groups) {
            k_sort_scan_sums<<<sg, SCAN_DIGITS * SCAN_LANES, 0, s>>>(a, p);
            k_sort_scan_grouped<<<sg, SCAN_DIGITS * SCAN_LANES, 0, s>>>(a, p);
        } else {
            k_sort_scan<<<sg, SCAN_DIGITS * SCAN_LANES, 0, s>>>(a, p);
        }
        k_sort_scatter<<<grid, SORT_TPB, 0, s>>>(a, p);
    }
    if (!a.fused) k_gather<<<gather_grid, 256, 0, s>>>(a);
}

void launch_end(const Args &a, hipStream_t s) { k_end<<<1, a.fused ? 256 : 1, 0, s>>>(a); }
void launch_stage_one(const Args &a, uint32_t chan_flags, uint32_t ncmd, hipStream_t s) {
    k_stage_one<<<1, 1, 0, s>>>(a, chan_flags, ncmd);
}

void launch_settle(const Args &a, uint32_t gather_grid, hipStream_t s) {
    k_gather<<<gather_grid, 256, 0, s>>>(a);
    k_settle<<<1, 1, 0, s>>>(a);
}

void launch_batch_begin(const Args *aa, uint32_t nr, hipStream_t s) {
    kb_begin<<<dim3(1, 1, nr), 256, 0, s>>>(aa);
}

void launch_batch_select(const Args *aa, uint32_t nr, uint32_t grid, uint32_t rounds,
                         hipStream_t s) {
    kb_sel_begin<<<dim3(1, 1, nr), 1, 0, s>>>(aa);
    for (uint32_t r = 0; r < rounds; ++r) {
        kb_sel_hist<<<dim3(grid, 1, nr), SEL_THREADS, 0, s>>>(aa);
        kb_sel_refine<<<dim3(1, 1, nr), SEL_THREADS, 0, s>>>(aa, grid);
    }
    kb_sel_collect<<<dim3(grid, 1, nr), SEL_THREADS, 0, s>>>(aa);
    kb_sel_final<<<dim3(1, 1, nr), SEL_THREADS, 0, s>>>(aa);
}

void launch_batch_spawn(const Args *aa, uint32_t nr, uint32_t grid, hipStream_t s) {
    kb_spawn<<<dim3(grid, 1, nr), 256, 0, s>>>(aa);
}

void launch_batch_sort(const Args *aa, uint32_t nr, uint32_t grid, uint32_t gather_grid,
                       uint32_t passes, uint32_t max_radix, hipStream_t s) {
    for (uint32_t p = 0; p < passes; ++p) {
        kb_sort_hist<<<dim3(grid, 1, nr), SORT_TPB, 0, s>>>(aa, p);
        kb_sort_scan<<<dim3((max_radix + SCAN_DIGITS - 1) / SCAN_DIGITS, 1, nr),
                       SCAN_DIGITS * SCAN_LANES, 0, s>>>(aa, p);
        kb_sort_scatter<<<dim3(grid, 1, nr), SORT_TPB, 0, s>>>(aa, p);
    }
    kb_gather<<<dim3(gather_grid, 1, nr), 256, 0, s>>>(aa);
}

void launch_batch_end(const Args *aa, uint32_t nr, hipStream_t s) {
    kb_end<<<dim3(1, 1, nr), 1, 0, s>>>(aa);
}

}  // namespace kst
