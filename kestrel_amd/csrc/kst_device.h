// kst_device.h -- layouts and kernel arguments shared by the gfx950 kernels and
// the host orchestration of the Kestrel render backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/kestrel_hip.h"

namespace kst {

// ---- voice pool: 24 u32 fields, structure of arrays, field f of voice i at
// f * capacity + i (shaders/common.wgsl:10-44) ----
constexpr uint32_t VOICE_FIELDS = 24;
enum : uint32_t {
    F_PHASE_LO = 0, F_PHASE_HI = 1, F_STEP_LO = 2, F_STEP_HI = 3,
    F_SMP_BASE = 4, F_SMP_LEN = 5, F_LOOP_START = 6, F_LOOP_END = 7,
    F_FLAGS = 8, F_ENV_STAGE = 9, F_ENV_LEVEL = 10, F_GAIN_L = 11,
    F_GAIN_R = 12, F_FILT_Z1 = 13, F_FILT_Z2 = 14, F_PARAMS = 15,
    F_REGION = 16, F_GATE_SLOT = 17, F_ORDINAL = 18, F_START_REL = 19,
    F_NOTE_LO = 20, F_NOTE_HI = 21, F_BORN_VARIANT = 22, F_STOP_REL = 23,
};

// Envelope stages (src/voice.rs:11-15), voice flags (src/bank.rs:12-14),
// region-params flags (src/bank.rs:192).
constexpr uint32_t ENV_ATTACK = 0, ENV_DECAY = 1, ENV_SUSTAIN = 2, ENV_RELEASE = 3,
                   ENV_DEAD = 4;
constexpr uint32_t VF_LOOP = 1, VF_LOOP_UNTIL_RELEASE = 2;
constexpr uint32_t RP_FILTER = 1;

// Gate and controller rows (src/voice.rs:18, 454-474).
constexpr uint32_t GATE_SLOTS = 2048;
constexpr uint32_t BEND_CHANNELS = 16, CHAN_FIELDS = 4;
constexpr uint32_t CHAN_BEND = 0, CHAN_GAIN_L = 1, CHAN_GAIN_R = 2, CHAN_VARIANT = 3;
constexpr uint32_t CHAN_ACTIVE_BEND = 1, CHAN_ACTIVE_GAIN = 2, CHAN_ACTIVE_VARIANT = 4;

constexpr uint32_t STEAL_DROP_NEW = 1, STEAL_QUIETEST = 2;
constexpr uint32_t ADMIT_EVEN = 1;

// Radix select of the k-th smallest 64-bit steal key (shaders/select.wgsl):
// MSB-first rounds of SEL_BITS-wide digits. Each round also keeps the smallest
// and largest key of every bin, so the next digit starts at the highest bit
// that still varies inside the chosen bin (note-id keys skip their zero high
// bits). Once the bin holds at most SEL_CAP keys, one workgroup finishes the
// rounds on the collected candidates, held in registers (16 per thread).
// Per-workgroup histograms are stored without atomics and summed by one
// workgroup: the same global address hit from every CU is what costs.
constexpr uint32_t SEL_BITS = 8;
constexpr uint32_t SEL_BINS = 1u << SEL_BITS;
constexpr uint32_t SEL_MAX_BLOCKS = 1024;
constexpr uint32_t SEL_THREADS = 1024;
constexpr uint32_t SEL_CAP = 16 * SEL_THREADS;
// DevState::sel_steal before the selection has counted its candidates.
constexpr uint32_t SEL_PENDING = 0xFFFFFFFFu;

// Locality sort tiles: 256 threads x 8 items, processed in index order.
constexpr uint32_t SORT_TPB = 256;
constexpr uint32_t SORT_IPT = 8;
constexpr uint32_t SORT_TILE = SORT_TPB * SORT_IPT;
// The sort's scan over groups of tiles, from pools of this many tiles on.
constexpr uint32_t SCAN_GROUP_TILES = 512, SCAN_GROUPED_MIN_TILES = 4096;
constexpr uint32_t MAX_SORT_PASSES = 32;
constexpr uint32_t SORT_MAX_BITS = 9;
constexpr uint32_t SORT_MAX_RADIX = 1u << SORT_MAX_BITS;

// Wave64 render form: 16-frame sub-tiles, transposed through LDS with a row
// stride of 68 words (68 = 4 mod 32 keeps the 128-bit column reads
// bank-conflict free).
constexpr uint32_t SUBTILE = 16;
constexpr uint32_t TROW = 68;
constexpr uint32_t FADE_LDS_MAX = 2048;
// The staged fade table: FADE_PAD gains of 1 before the fade's and FADE_PAD of
// 0 after them, so any sub-tile reads it at 16 consecutive entries.
constexpr uint32_t FADE_PAD = 16;
constexpr uint32_t TBUF_FLOATS = 2 * SUBTILE * TROW;  // 32 rows
// Batch rows summed per workgroup in the first mixdown stage.
constexpr uint32_t REDUCE_ROWS = 256;
// Workgroups are dealt to the XCDs in turn (workgroup w runs on XCD w mod 8,
// up to a rotation), and each XCD starts its workgroups in order.
constexpr uint32_t XCDS = 8;

// Device-resident state of one backend, carried across blocks. The live count
// lives here rather than in a kernel argument so a whole block of passes, or a
// run of blocks, can be launched without the host in between
// (src/gpu/mod.rs:1-7 made the same move for wgpu).
// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD
// i % 8). Give each XCD one contiguous run of the sorted pool instead, so its
// L2 holds the samples of a few regions rather than of all of them. A
// bijection on [0, n): XCD x takes groups [x*q + min(x, r), ... + q + (x < r)).
__device__ __forceinline__ uint32_t xcd_group(uint32_t id, uint32_t n) {
    const uint32_t x = id % XCDS, k = id / XCDS;
    const uint32_t q = n / XCDS, r = n % XCDS;
    return x * q + (x < r ? x : r) + k;
}

// The same runs, each walked from its end: the first workgroups an XCD
// launches take the last batches of its run. The pool's tail holds the block's
// new voices (late starts, the slowest waves), so they start first instead of
// trailing the launch.
__device__ __forceinline__ uint32_t xcd_group_rev(uint32_t id, uint32_t n) {
    const uint32_t x = id % XCDS, k = id / XCDS;
    const uint32_t q = n / XCDS, r = n % XCDS;
    const uint32_t cnt = q + (x < r ? 1u : 0u);
    return x * q + (x < r ? x : r) + (cnt - 1 - k);
}

// The render's dispatch lists deal the sorted pool's survivor batches to the
// XCDs in RUN_CHUNKS interleaved chunks (chunk c to XCD c mod XCDS): each XCD
// renders several stretches of the pool, so no XCD draws a region whose voices
// cost more than the rest. run_len: how many survivor batches XCD x renders;
// run_batch: its i-th in pool order.
constexpr uint32_t RUN_CHUNKS = 8 * XCDS;
__device__ __forceinline__ uint32_t run_chunk(uint32_t nsurv) {
    const uint32_t ch = (nsurv + RUN_CHUNKS - 1) / RUN_CHUNKS;
    return ch > 0 ? ch : 1u;
}
__device__ __forceinline__ uint32_t run_len(uint32_t nsurv, uint32_t x) {
    const uint32_t ch = run_chunk(nsurv);
    const uint32_t chunks = (nsurv + ch - 1) / ch;  // the last may be short
    if (x >= chunks) return 0;
    const uint32_t mine = (chunks - 1 - x) / XCDS + 1;  // chunks x, x + XCDS, ...
    const uint32_t lastc = x + (mine - 1) * XCDS;
    return (mine - 1) * ch + (nsurv - lastc * ch < ch ? nsurv - lastc * ch : ch);
}
__device__ __forceinline__ uint32_t run_batch(uint32_t nsurv, uint32_t x, uint32_t i) {
    const uint32_t ch = run_chunk(nsurv);
    return (x + (i / ch) * XCDS) * ch + i % ch;
}

struct DevState {
    uint32_t live;    // compacted live count at the start of the block
    uint32_t parity;  // which pool buffer holds the live pool
    uint32_t block;   // index of the block within the staged inputs
    uint32_t idle;    // no staged input left for this block: every pass is a no-op
    unsigned long long stolen, dropped;  // cumulative
    // Derived per block by k_begin (src/gpu/mod.rs:1015-1040).
    uint32_t want, steal_k, spawn_count, live_total;
    uint32_t cmd_off, chan_flags, live_new, peak_bits;
    // Radix select state.
    unsigned long long sel_prefix;  // bits >= sel_top of every remaining candidate
    uint32_t sel_top;               // bits below this are still unresolved
    uint32_t sel_k;                 // rank of the k-th key among the candidates
    uint32_t sel_count;             // remaining candidates
    uint32_t sel_done;              // candidates few enough to rank in LDS
    uint32_t cand_n;                // candidates collected
    uint32_t perm_valid;            // the live pool is read through the last sort's permutation
    // What the last render left, for the next block's victim selection: the
    // voices it ran (slots [0, snap_n) of pool snap_parity) and its block
    // index. How many still sound is counted by the selection's first round.
    uint32_t snap_n, snap_parity, snap_block;
    // The selection's own view (k_sel_begin): candidates are the sounding
    // voices among slots [0, sel_n) of pool sel_parity; sel_steal become
    // victims (SEL_PENDING until the candidates are counted), of a block
    // with sel_want note-ons.
    uint32_t sel_n, sel_parity, sel_steal, sel_want;
    // Staged runs order the two streams on the device instead of through
    // events: the first mixdown kernel after block b's render stores b's
    // sequence number in render_seq (the render has ended, so the pool it
    // wrote is visible), the selection waits for it; the selection's last
    // kernel stores it in sel_seq, k_end waits for it. sync_timeout: a wait
    // gave up (the host reports an error).
    uint32_t render_seq, sel_seq, sync_timeout, sync_pad;
};

// Per-block result row, what stats() reads after each render.
struct BlockResult {
    unsigned long long active_voices, stolen, dropped;
    uint32_t peak_bits, pad;
};

struct Args {
    DevState *st;
    uint32_t *pools[2];
    uint32_t cap;  // allocated slots = SoA stride (Config::pool_slots)
    uint32_t bank_len;
    uint32_t clamp;  // the bank holds -32768, so taps need the -1 clamp
    uint32_t tap_bytes;  // 2: bytes per int16 frame (a runtime value, so tap offsets use one mad_u24)
    const int16_t *bank;
    const kst_region_params *params;
    uint32_t ppv;           // params per variant
    uint32_t max_variants;
    // Staged per-block inputs, indexed by st->block.
    const uint32_t *gates;       // [nblocks][gate_words]
    const uint32_t *chans;       // [nblocks][chan_words]
    const uint32_t *chan_flags;  // [nblocks]
    const kst_spawn_cmd *cmds;   // concatenated
    const uint32_t *cmd_counts;  // [nblocks]
    const uint32_t *cmd_offs;    // [nblocks]
    uint32_t gate_words, chan_words;
    float *out;               // [nblocks][2F]
    BlockResult *results;     // [nblocks]
    // Config.
    uint32_t nblocks;  // block slots holding inputs; a block at or past this idles
    uint32_t F, gate_frames, max_voices, max_steal, steal_span, fade;
    // [FADE_PAD + fade + FADE_PAD]: 1 ..., then 1 - d / fade for d = 0 .. fade - 1
    // as the per-frame fade computes it, then 0 ...
    const float *fade_tab;
    uint32_t fade_lds;      // entries of fade_tab staged in LDS (0: too long to stage)
    uint32_t steal_rule, admit_rule, interp, exp_decay, exp_release;
    float env_floor;
    // Render / reduce.
    float *partials;         // wave64: [batch][2F]; plain: [2F][nwg]
    float *rsum;             // wave64 mixdown stage 1: [chunk][2F][hi, lo]
    uint32_t render_groups;  // render workgroups launched this block
    uint32_t waves_per_group;
    uint32_t plain_tile, plain_wg, kahan;
    // Locality sort and compaction.
    uint32_t *keys[2];
    uint32_t *idx[2];
    uint32_t *sort_hist;  // [passes][radix][tiles_max]
    uint32_t *sort_tot;   // [passes][radix] digit totals
    uint32_t sort_on, sort_passes, digit_bits, tiles_max;
    // Pools of many tiles scan each digit's counts in groups of SCAN_GROUP_TILES
    // tiles (scan_groups = the most groups, 0: one scan kernel); scan_grp holds
    // each group's count per digit, [groups][radix].
    uint32_t *scan_grp;
    uint32_t scan_groups, scan_pad;
    uint32_t region_shift, stage_shift, phase_shift, phase_mask, dead_region;
    // Fused compaction (wave64 per-block and staged runs): the render reads
    // the survivors through the previous sort's permutation and writes every
    // field of every voice to the other pool buffer, so no gather pass runs.
    uint32_t fused;
    // Select the victims of the block after the last render (staged runs,
    // on the side stream) rather than of the block about to run.
    uint32_t sel_ahead;
    // k_end also begins the next block (staged runs with the selection ahead,
    // where nothing else runs between the two): one launch fewer per block.
    uint32_t end_begins;
    // Fused: the top digit of the last sort pass at which dead keys begin, when
    // that boundary falls on a digit (the survivor count is then that digit's
    // base); ~0u when it does not.
    uint32_t dead_digit;
    uint32_t *sel_cnt;             // [SEL_MAX_BLOCKS][SEL_BINS] per-workgroup counts
    unsigned long long *sel_min;   // [SEL_MAX_BLOCKS][SEL_BINS]
    unsigned long long *sel_max;   // [SEL_MAX_BLOCKS][SEL_BINS]
    unsigned long long *cand_key;  // [cap]
    uint32_t *cand_slot;           // [cap]
    // Per-block render dispatch order (wave64, fused; null: the plain XCD
    // runs). batch_victim[j]: survivor batch j holds a steal victim (written
    // by k_spawn). render_order[x * order_stride + k]: the k-th batch XCD x
    // renders (written by k_order): its share of the new voices' batches, its
    // run's batches holding a victim, then the run's other batches.
    uint32_t *batch_victim;  // [cap / 64 + 1]
    uint32_t *render_order;  // [XCDS][order_stride]
    uint32_t order_stride;
    // Device-side ordering of a staged run's two streams (0: none): the
    // sequence number the mixdown publishes after the render and the
    // selection waits for, and the one the selection publishes and k_end
    // waits for.
    uint32_t sig_render, wait_render, sig_sel, wait_sel;
};

// A wait gives up after this many ticks of the 100 MHz real-time counter
// (20 s), records sync_timeout and lets the chain run on, so no kernel can
// spin forever; every later wait then returns at once.
constexpr uint64_t SYNC_TIMEOUT_TICKS = 2000000000ull;

// Spin (one thread) until *flag has reached `want` (wrap-safe), with acquire
// ordering at device scope: what was written before the flag's release is
// visible afterwards.
__device__ __forceinline__ void wait_seq(DevState *st, uint32_t *flag, uint32_t want) {
    if (__hip_atomic_load(&st->sync_timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const uint64_t t0 = wall_clock64();
    while ((int32_t)(__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) - want) < 0) {
        if (wall_clock64() - t0 > SYNC_TIMEOUT_TICKS) {
            __hip_atomic_store(&st->sync_timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Publish `seq` in *flag with release ordering at device scope.
__device__ __forceinline__ void signal_seq(uint32_t *flag, uint32_t seq) {
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// One of a pair of buffers, chosen without indexing the kernel-argument array
// (a dynamic index would force the pair out to scratch memory).
template <typename T>
__device__ __forceinline__ T *sel2(T *const (&pair)[2], uint32_t i) {
    return i ? pair[1] : pair[0];
}

// The live pool's voice i sits at slot perm[i] after a fused block (the last
// sort's permutation, never applied as a copy), at slot i otherwise.
__device__ __forceinline__ const uint32_t *live_perm(const Args &a) {
    return a.st->perm_valid ? sel2(a.idx, a.sort_passes & 1) : nullptr;
}

// ---- launchers (kst_pool.hip) ----
void launch_begin(const Args &a, hipStream_t s);
void launch_select(const Args &a, uint32_t grid, uint32_t rounds, hipStream_t s);
void launch_spawn(const Args &a, uint32_t grid, hipStream_t s);
void launch_sort(const Args &a, uint32_t grid, uint32_t gather_grid, hipStream_t s);
void launch_end(const Args &a, hipStream_t s);
// The per-block render's dispatch lists from k_spawn's victim flags.
void launch_order(const Args &a, hipStream_t s);
// The per-block inputs of a single submitted block (channel flags, command
// count, command offset 0, block index 0), written from kernel arguments.
void launch_stage_one(const Args &a, uint32_t chan_flags, uint32_t ncmd, hipStream_t s);
// Apply a fused block's permutation as a copy (before the pool is read or
// written in place): gather, then flip to the gathered buffer.
void launch_settle(const Args &a, uint32_t gather_grid, hipStream_t s);

// ---- launchers (kst_render.hip) ----
// Returns false when the wave64 form does not cover this configuration.
bool fast_render_supported(uint32_t F, uint32_t gate_frames, uint32_t interp);
size_t fast_render_lds_bytes(uint32_t waves_per_group, uint32_t fade_lds);
// With `start`/`stop`, the launch itself records the kernel's start and end
// (hipExtLaunchKernel): no event packets of their own in the stream.
void launch_render_fast(const Args &a, bool chan, hipStream_t s, hipEvent_t start = nullptr,
                        hipEvent_t stop = nullptr);

// A batch of renders in one launch per pass: `aa` is a device array of `nr`
// Args (the render's own buffers and config), and blockIdx.z picks the
// render. Grids are sized for the largest render; smaller ones return early.
void launch_batch_begin(const Args *aa, uint32_t nr, hipStream_t s);
void launch_batch_select(const Args *aa, uint32_t nr, uint32_t grid, uint32_t rounds,
                         hipStream_t s);
void launch_batch_spawn(const Args *aa, uint32_t nr, uint32_t grid, hipStream_t s);
void launch_batch_render(const Args *aa, uint32_t nr, uint32_t interp, bool chan, bool clamp,
                         uint32_t waves_per_group, uint32_t groups, uint32_t fade_lds,
                         hipStream_t s);
void launch_batch_reduce(const Args *aa, uint32_t nr, bool comp, uint32_t max_two_f,
                         uint32_t max_rows, hipStream_t s);
void launch_batch_sort(const Args *aa, uint32_t nr, uint32_t grid, uint32_t gather_grid,
                       uint32_t passes, uint32_t max_radix, hipStream_t s);
void launch_batch_end(const Args *aa, uint32_t nr, hipStream_t s);
// Rows of the allocation-free design entry point's device scratch.
constexpr size_t DESIGN_SCRATCH_ROWS = 65536;
hipError_t design_scratch(int device, kst_region_design **out);
void launch_design(const kst_region_design *in, uint32_t n, kst_region_params *out, double sr,
                   bool exp_decay, bool exp_release, float env_floor, hipStream_t s);
void launch_reduce_fast(const Args &a, bool comp, uint32_t max_rows, hipStream_t s);
void launch_render_plain(const Args &a, bool chan, hipStream_t s);
void launch_reduce_plain(const Args &a, hipStream_t s);

}  // namespace kst
