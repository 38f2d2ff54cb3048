// kst_backend.hip -- host orchestration of the gfx950 render backend behind the
// C ABI of include/kestrel_hip.h.
//
// One backend = one pool on one device, one HIP stream, and the per-block
// launch chain of src/gpu/mod.rs:1014-1289 rebuilt around a device-resident
// block state: k_begin evaluates the block's steal/spawn arithmetic on the
// device from the live count the previous block left there, so the same chain
// serves the synchronous Backend::render and runs of staged blocks (optionally
// replayed from a captured hipGraph) without the host in between.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kst_device.h"
#include "kst_stage.h"

using namespace kst;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(KST_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

// The reference's Config::max_steal / pool_slots / steal_span
// (src/config.rs:534-552); both sides must use exactly this arithmetic.
uint32_t cfg_max_steal(const kst_config &c) {
    uint64_t n = (uint64_t)c.max_voices * c.max_steal_percent / 100;
    uint32_t v = (uint32_t)n;
    return v > 1 ? v : 1;
}
uint32_t cfg_pool_slots(const kst_config &c) {
    uint64_t v = (uint64_t)c.max_voices + cfg_max_steal(c);
    return v > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)v;
}
uint32_t cfg_steal_span(const kst_config &c) {
    uint32_t s = c.block_frames > c.steal_fade_frames ? c.block_frames - c.steal_fade_frames : 0;
    return s > 1 ? s : 1;
}

bool is_pow2(uint32_t x) { return x && !(x & (x - 1)); }

// src/config.rs:412-525, the render-path subset, plus the device's own limits.
int validate(const kst_config &c) {
    if (c.sample_rate < 8000 || c.sample_rate > 768000)
        return fail(KST_EINVAL, "sample_rate %u out of range", c.sample_rate);
    if (c.block_frames == 0 || c.reduce_tile == 0 || c.block_frames % c.reduce_tile != 0)
        return fail(KST_EINVAL, "block_frames (%u) must be a non-zero multiple of reduce_tile (%u)",
                    c.block_frames, c.reduce_tile);
    if (!is_pow2(c.workgroup_size) || c.workgroup_size < 32 || c.workgroup_size > 1024)
        return fail(KST_EINVAL, "workgroup_size %u must be a power of two in [32, 1024]",
                    c.workgroup_size);
    if (!is_pow2(c.reduce_tile))
        return fail(KST_EINVAL, "reduce_tile %u must be a power of two", c.reduce_tile);
    if (c.gate_frames == 0 || c.gate_frames % c.reduce_tile != 0 ||
        c.block_frames % c.gate_frames != 0)
        return fail(KST_EINVAL,
                    "gate_frames (%u) must be a multiple of reduce_tile (%u) and divide "
                    "block_frames (%u)",
                    c.gate_frames, c.reduce_tile, c.block_frames);
    uint64_t shared = (uint64_t)c.workgroup_size * (c.reduce_tile * 2ull + 1) * 4;
    if (shared > 65536)
        return fail(KST_EINVAL,
                    "workgroup_size %u with reduce_tile %u needs %llu bytes of workgroup "
                    "storage, over the 64 KiB a workgroup can have",
                    c.workgroup_size, c.reduce_tile, (unsigned long long)shared);
    if (c.workgroup_size < c.reduce_tile * 2)
        return fail(KST_EINVAL, "workgroup_size %u must be at least twice reduce_tile %u",
                    c.workgroup_size, c.reduce_tile);
    if (c.max_voices == 0) return fail(KST_EINVAL, "max_voices must be non-zero");
    if (c.steal_fade_frames == 0 || c.steal_fade_frames >= c.block_frames)
        return fail(KST_EINVAL, "steal_fade_frames %u must be in 1..block_frames (%u)",
                    c.steal_fade_frames, c.block_frames);
    if (c.max_steal_percent == 0 || c.max_steal_percent > 100)
        return fail(KST_EINVAL, "max_steal_percent %u must be in 1..=100", c.max_steal_percent);
    if (c.block_frames > 65535)
        return fail(KST_EINVAL, "block_frames %u must be at most 65535", c.block_frames);
    if (c.steal_rule > 2) return fail(KST_EINVAL, "steal_rule %u unknown", c.steal_rule);
    if (c.admit_rule > 1) return fail(KST_EINVAL, "admit_rule %u unknown", c.admit_rule);
    if (c.interpolation > 2) return fail(KST_EINVAL, "interpolation %u unknown", c.interpolation);
    if (c.decay_curve > 1 || c.release_curve > 1)
        return fail(KST_EINVAL, "envelope curve must be 0 (exponential) or 1 (linear)");
    if (c.max_param_variants == 0) return fail(KST_EINVAL, "max_param_variants must be non-zero");
    if (c.waves_per_group == 0 || c.waves_per_group > 4)
        return fail(KST_EINVAL, "waves_per_group %u must be in 1..=4", c.waves_per_group);
    if (c.sort_digit_bits == 0 || c.sort_digit_bits > SORT_MAX_BITS)
        return fail(KST_EINVAL, "sort_digit_bits %u must be in 1..=%u", c.sort_digit_bits,
                    SORT_MAX_BITS);
    // 32-bit SoA word offsets: field 23 of the last slot must stay below 2^32.
    if ((uint64_t)cfg_pool_slots(c) * VOICE_FIELDS >= (1ull << 32))
        return fail(KST_EINVAL,
                    "max_voices %u allocates %u pool slots; the 24-field pool is addressed "
                    "with 32-bit word offsets, so cap --max-voices at %llu",
                    c.max_voices, cfg_pool_slots(c),
                    (unsigned long long)(((1ull << 32) / VOICE_FIELDS - 1) * 100 /
                                         (100 + c.max_steal_percent)));
    return KST_OK;
}

// Pools from this size on select the next block's victims on the side stream
// during a staged run.
constexpr uint32_t PIPELINE_MIN_VOICES = 1u << 15;

struct SortPlan {
    uint32_t bits, region_shift, stage_shift, phase_shift, phase_mask, dead_region;
};

// src/gpu/mod.rs:76-105
SortPlan plan_sort(uint32_t n_regions, uint32_t max_len) {
    auto bitlen = [](uint32_t x) -> uint32_t { return x == 0 ? 0 : 32 - __builtin_clz(x); };
    SortPlan p;
    p.dead_region = n_regions > 1 ? n_regions : 1;
    uint32_t region_bits = bitlen(p.dead_region);
    uint32_t stage_bits = 3;
    uint32_t len_bits = bitlen(max_len > 1 ? max_len : 1);
    uint32_t avail = 32 > region_bits + stage_bits ? 32 - (region_bits + stage_bits) : 0;
    uint32_t phase_bits = std::min(avail, len_bits);
    p.bits = region_bits + stage_bits + phase_bits;
    p.region_shift = phase_bits + stage_bits;
    p.stage_shift = phase_bits;
    p.phase_shift = len_bits > phase_bits ? len_bits - phase_bits : 0;
    p.phase_mask = phase_bits >= 32 ? 0xFFFFFFFFu : ((1u << phase_bits) - 1);
    return p;
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        return hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

const char *PASS_NAMES[5] = {"steal", "spawn", "render", "reduce", "compact"};

}  // namespace

struct kst_backend {
    kst_config cfg;
    int device = 0;
    uint32_t cus = 256;
    hipStream_t stream = nullptr;
    bool fast = true;  // wave64 form in use

    uint32_t slots = 0, ppv = 0, tiles = 0, gate_words = 0, chan_words = 0;
    uint32_t max_steal = 0;
    SortPlan sort{};
    uint32_t sort_passes = 1, digit_bits = 1, tiles_max = 1, scan_groups = 0;
    uint32_t dead_digit = ~0u;  // Args::dead_digit
    uint32_t max_groups = 1;  // render workgroups at capacity (plain: partial-buffer width)
    uint32_t max_rows = 1;    // wave64 partial rows at capacity (one per 64-voice batch)

    DevBuf<uint32_t> pools[2];
    DevBuf<int16_t> bank;
    uint32_t bank_len = 0;
    uint32_t bank_clamp = 1;
    DevBuf<kst_region_params> params;
    DevBuf<float> partials, rsum, fade_tab;
    uint32_t fade_lds = 0;
    DevBuf<uint32_t> keys[2], idx[2], sort_hist, sort_tot, scan_grp, sel_cnt, cand_slot;
    // The per-block render's dispatch lists (wave64 form): per-batch victim
    // flags and one list per XCD of order_stride entries.
    DevBuf<uint32_t> batch_victim, render_order;
    uint32_t order_stride = 0;
    DevBuf<unsigned long long> cand_key, sel_min, sel_max;
    DevBuf<DevState> st;

    // Staged inputs (one slot for the per-block API, nblocks for staged runs).
    uint32_t stage_cap = 0, staged = 0;
    DevBuf<uint32_t> g_gates, g_chans, g_chan_flags, g_cmd_counts, g_cmd_offs;
    DevBuf<kst_spawn_cmd> g_cmds;
    DevBuf<float> g_out;
    DevBuf<BlockResult> g_results;

    // Host side of the per-block API.
    uint32_t *h_gates = nullptr, *h_chans = nullptr;  // pinned: uploads run without staging
    uint32_t h_chan_flags = 0;
    bool staged_chan = true, graph_chan = true;
    size_t n_pending = 0;  // this block's spawn commands, staged in pin_cmds
    uint32_t *pin_small = nullptr;  // pinned staging for per-block uploads
    size_t pin_cmds_cap = 0;
    kst_spawn_cmd *pin_cmds = nullptr;
    float *pin_out = nullptr;
    BlockResult *pin_result = nullptr;

    // Host mirror, exact: every change is host-decided or read back.
    uint32_t live = 0, parity = 0;
    // A fused block left the live pool behind its sort's permutation (the
    // device's DevState::perm_valid, mirrored here).
    bool perm_valid = false;
    uint64_t stolen = 0, dropped = 0;
    float peak = 0.0f;

    // 0 off, 1 every pass, 2 the render kernel only (KST_PROFILE_*).
    uint32_t profile = 0;
    bool inflight = false;  // a render submitted and not yet collected
    // Select rounds launched per block. A round whose bin is already small
    // enough returns at once; rounds still needed after these are finished by
    // the final single-workgroup pass, so the count only affects speed.
    uint32_t sel_rounds = 3;
    hipEvent_t ev[6] = {};
    std::vector<hipEvent_t> stage_ev;
    std::vector<kst_timing> timings;

    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // The captured chain holds its kernels' arguments by value, the staged
    // block count among them: a run of another count captures again.
    uint32_t graph_staged = 0;
    // Staged runs select the next block's victims on a side stream while the
    // current block reduces and sorts (default priority, like the main stream).
    hipStream_t side = nullptr;
    hipEvent_t ev_rendered = nullptr, ev_selected = nullptr;
    uint32_t sync_seq = 0;  // last sequence number of a device-ordered block
    // Staged runs order their two streams on the device (see launch_block)
    // unless kernels may be serialized across streams, where a device-side
    // wait could not be met: KST_DEVICE_SYNC=0, AMD_SERIALIZE_KERNEL or
    // HIP_LAUNCH_BLOCKING set (profilers that collect counters serialize too;
    // run them with KST_DEVICE_SYNC=0). Events order the streams then.
    bool device_sync = true;

    ~kst_backend() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : stage_ev) (void)hipEventDestroy(e);
        for (auto &p : pools) p.release();
        bank.release();
        params.release();
        partials.release();
        rsum.release();
        fade_tab.release();
        for (auto &k : keys) k.release();
        for (auto &k : idx) k.release();
        sort_hist.release();
        sort_tot.release();
        scan_grp.release();
        sel_cnt.release();
        sel_min.release();
        sel_max.release();
        cand_slot.release();
        cand_key.release();
        st.release();
        g_gates.release();
        g_chans.release();
        g_chan_flags.release();
        g_cmd_counts.release();
        g_cmd_offs.release();
        g_cmds.release();
        g_out.release();
        g_results.release();
        if (pin_small) (void)hipHostFree(pin_small);
        if (h_gates) (void)hipHostFree(h_gates);
        if (h_chans) (void)hipHostFree(h_chans);
        if (pin_cmds) (void)hipHostFree(pin_cmds);
        if (pin_out) (void)hipHostFree(pin_out);
        if (pin_result) (void)hipHostFree(pin_result);
        if (ev_rendered) (void)hipEventDestroy(ev_rendered);
        if (ev_selected) (void)hipEventDestroy(ev_selected);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int ensure_stage(uint32_t nblocks, size_t total_cmds) {
        if (nblocks > stage_cap) {
            HIPCHK(g_gates.alloc((size_t)nblocks * gate_words));
            HIPCHK(g_chans.alloc((size_t)nblocks * chan_words));
            HIPCHK(g_chan_flags.alloc(nblocks));
            HIPCHK(g_cmd_counts.alloc(nblocks));
            HIPCHK(g_cmd_offs.alloc(nblocks));
            HIPCHK(g_out.alloc((size_t)nblocks * 2 * cfg.block_frames));
            HIPCHK(g_results.alloc(nblocks));
            stage_cap = nblocks;
            drop_graph();
        }
        if (total_cmds > g_cmds.n || !g_cmds.p) {
            size_t n = std::max<size_t>(total_cmds, 1024);
            HIPCHK(g_cmds.alloc(n));
            drop_graph();
        }
        return KST_OK;
    }

    // Apply a pending permutation as a copy, so the live pool is in order in
    // one buffer: before the host reads or writes it, and before a batch (which
    // compacts with its own gather) runs this render.
    int settle() {
        if (!perm_valid) return KST_OK;
        Args a = args(1);
        launch_settle(a, per_voice_grid(slots), stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
        parity ^= 1;
        perm_valid = false;
        return KST_OK;
    }

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }

    uint32_t per_voice_grid(uint32_t n) const {
        uint32_t g = (n + 255) / 256;
        return std::max(1u, std::min(g, cus * 8));
    }
    uint32_t sort_grid(uint32_t n) const {
        uint32_t t = (n + SORT_TILE - 1) / SORT_TILE;
        return std::max(1u, std::min(t, cus * 4));
    }
    uint32_t render_groups_for(uint32_t n) const {
        if (fast) {
            uint32_t waves = (n + 63) / 64;
            return std::max(1u, (waves + cfg.waves_per_group - 1) / cfg.waves_per_group);
        }
        uint32_t cap = std::min(std::max(cfg.max_render_workgroups, 1u), 65535u);
        uint32_t g = (n + cfg.workgroup_size - 1) / cfg.workgroup_size;
        return std::max(1u, std::min(g, std::min(cap, max_groups)));
    }

    Args args(uint32_t render_groups) const {
        Args a{};
        a.st = st.p;
        a.pools[0] = pools[0].p;
        a.pools[1] = pools[1].p;
        a.cap = slots;
        a.bank_len = bank_len;
        a.tap_bytes = 2;
        a.clamp = bank_clamp;
        a.bank = bank.p;
        a.params = params.p;
        a.ppv = ppv;
        a.max_variants = cfg.max_param_variants;
        a.gates = g_gates.p;
        a.chans = g_chans.p;
        a.chan_flags = g_chan_flags.p;
        a.cmds = g_cmds.p;
        a.cmd_counts = g_cmd_counts.p;
        a.cmd_offs = g_cmd_offs.p;
        a.gate_words = gate_words;
        a.chan_words = chan_words;
        a.out = g_out.p;
        a.results = g_results.p;
        a.F = cfg.block_frames;
        a.gate_frames = cfg.gate_frames;
        a.max_voices = cfg.max_voices;
        a.max_steal = max_steal;
        a.steal_span = cfg_steal_span(cfg);
        a.fade = cfg.steal_fade_frames;
        a.steal_rule = cfg.steal_rule;
        a.admit_rule = cfg.admit_rule;
        a.interp = cfg.interpolation;
        a.exp_decay = cfg.decay_curve == 0;
        a.exp_release = cfg.release_curve == 0;
        a.env_floor = cfg.env_floor;
        a.partials = partials.p;
        a.fade_tab = fade_tab.p;
        a.fade_lds = fade_lds;
        a.nblocks = std::max(staged, 1u);
        a.rsum = rsum.p;
        a.render_groups = render_groups;
        a.waves_per_group = cfg.waves_per_group;
        a.plain_tile = cfg.reduce_tile;
        a.plain_wg = cfg.workgroup_size;
        a.kahan = cfg.kahan_reduce;
        for (int k = 0; k < 2; ++k) {
            a.keys[k] = keys[k].p;
            a.idx[k] = idx[k].p;
        }
        a.sort_hist = sort_hist.p;
        a.sort_tot = sort_tot.p;
        a.sort_on = cfg.sort_voices;
        a.sort_passes = sort_passes;
        a.digit_bits = digit_bits;
        a.tiles_max = tiles_max;
        a.scan_grp = scan_grp.p;
        a.scan_groups = scan_groups;
        a.region_shift = sort.region_shift;
        a.stage_shift = sort.stage_shift;
        a.phase_shift = sort.phase_shift;
        a.phase_mask = sort.phase_mask;
        a.dead_region = sort.dead_region;
        a.fused = fast ? 1u : 0u;
        a.dead_digit = dead_digit;
        a.sel_cnt = sel_cnt.p;
        a.sel_min = sel_min.p;
        a.sel_max = sel_max.p;
        a.cand_key = cand_key.p;
        a.cand_slot = cand_slot.p;
        a.batch_victim = batch_victim.p;
        a.render_order = render_order.p;
        a.order_stride = order_stride;
        return a;
    }

    // Victim selection for the block about to run, or (ahead) for the block
    // after the last render.
    // `sync`: this selection's sequence number in a device-ordered staged run
    // (it waits for the render's and publishes its own), 0 otherwise.
    void launch_sel(uint32_t n_live, bool ahead, hipStream_t s, uint32_t sync = 0) {
        Args a = args(1);
        a.sel_ahead = ahead ? 1u : 0u;
        a.wait_render = sync;
        a.sig_sel = sync;
        launch_select(a, std::max(1u, std::min(std::min(cus, SEL_MAX_BLOCKS),
                                               (n_live + SEL_THREADS * 4 - 1) / (SEL_THREADS * 4))),
                      sel_rounds, s);
    }

    // The per-block chain. `n_live`/`n_total` size the grids; the kernels read
    // the true counts from the device state, so capacity-sized grids (staged
    // runs) are as correct as exact ones (per-block calls). Pipelined (staged
    // runs, wave64 form): this block's victims were selected during the last
    // block; after the render the side stream selects the next block's while
    // this one reduces, sorts and compacts, and the block ends by joining it.
    //
    // device_sync (pipelined, not captured): the two streams are ordered on
    // the device rather than by events, whose packets cost the chain ~20 us
    // per fork and join. The mixdown after the render publishes the block's
    // sequence number, the side stream's selection waits for it and publishes
    // its own at its end, and k_end waits for that. Every wait is launched
    // after the kernel it waits for, so the chain cannot deadlock even where
    // both streams share one hardware queue, and every wait is bounded.
    void launch_block(uint32_t n_live, uint32_t n_total, bool may_steal, bool spawn, bool chan,
                      hipEvent_t *ev, bool pipelined = false, bool device_sync = false) {
        // Events around every pass, or (KST_PROFILE_RENDER) only around the
        // render kernel: each event in the stream costs the chain a few
        // microseconds.
        const bool timed_render = ev != nullptr;
        const bool timed = timed_render && !render_only();
        const uint32_t rg = render_groups_for(n_total);
        Args a = args(rg);
        hipStream_t s = stream;
        if (timed) (void)hipEventRecord(ev[0], s);
        if (may_steal && !pipelined) launch_sel(n_live, false, s);
        // Pipelined, the last block's k_end began this one (the run's first
        // block is begun before its loop).
        if (!pipelined) launch_begin(a, s);
        if (timed) (void)hipEventRecord(ev[1], s);
        // The wave64 form's k_spawn also flags the batches holding a victim,
        // from which k_order builds the render's dispatch lists.
        if (spawn || fast) launch_spawn(a, per_voice_grid(n_total), s);
        if (fast) launch_order(a, s);
        if (fast && timed_render && !timed) {
            // The render kernel alone: its launch records its own start and end.
            launch_render_fast(a, chan, s, ev[2], ev[3]);
        } else {
            if (timed_render) (void)hipEventRecord(ev[2], s);
            if (fast)
                launch_render_fast(a, chan, s);
            else
                launch_render_plain(a, chan, s);
            if (timed_render) (void)hipEventRecord(ev[3], s);
        }
        const bool dsync = pipelined && device_sync;
        uint32_t seq = 0;
        if (dsync) {
            seq = ++sync_seq;
            if (seq == 0) seq = ++sync_seq;  // 0 means "no wait"
        } else if (pipelined) {
            (void)hipEventRecord(ev_rendered, s);
            (void)hipStreamWaitEvent(side, ev_rendered, 0);
            launch_sel(n_live, true, side);
            (void)hipEventRecord(ev_selected, side);
        }
        Args ar = a;
        ar.sig_render = seq;
        if (fast)
            launch_reduce_fast(ar, cfg.compensated != 0, (n_total + 63) / 64, s);
        else
            launch_reduce_plain(a, s);
        if (dsync) launch_sel(n_live, true, side, seq);
        if (timed) (void)hipEventRecord(ev[4], s);
        launch_sort(a, sort_grid(n_total), per_voice_grid(n_total), s);
        if (pipelined) a.end_begins = 1;
        a.wait_sel = seq;
        launch_end(a, s);
        if (pipelined && !dsync) (void)hipStreamWaitEvent(s, ev_selected, 0);
        if (timed) (void)hipEventRecord(ev[5], s);
    }

    bool render_only() const {
        return profile == KST_PROFILE_RENDER || profile == KST_PROFILE_RENDER_SAMPLED;
    }

    // Per-pass milliseconds, averaged over the event sets of 6 of blocks 0,
    // stride, 2 * stride, ... below `nblocks` (only the render's pair under
    // the render-only levels).
    void read_timings(const hipEvent_t *evs, uint32_t nblocks, uint32_t stride = 1) {
        timings.clear();
        for (int i = 0; i < 5; ++i) {
            if (render_only() && i != 2) continue;
            double sum = 0.0;
            uint32_t cnt = 0;
            for (uint32_t k = 0; k < nblocks; k += stride) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, evs[6 * k + i], evs[6 * k + i + 1]) == hipSuccess)
                    sum += ms;
                ++cnt;
            }
            timings.push_back({PASS_NAMES[i], cnt ? sum / cnt : 0.0});
        }
    }

    int ensure_stage_events(uint32_t nblocks) {
        while (stage_ev.size() < 6ull * nblocks) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            stage_ev.push_back(e);
        }
        return KST_OK;
    }
};

namespace kst {

// Room for n commands in the pinned staging the render uploads from.
int spawn_reserve(kst_backend *b, size_t n, kst_spawn_cmd **out) {
    if (!b || !out) return fail(KST_EINVAL, "null argument");
    if (b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (n > 0xFFFFFFFFull) return fail(KST_EINVAL, "%zu spawn commands in one block", n);
    if (n > b->pin_cmds_cap || !b->pin_cmds) {
        if (b->pin_cmds) (void)hipHostFree(b->pin_cmds);
        b->pin_cmds = nullptr;
        b->pin_cmds_cap = 0;
        const size_t cap = std::max<size_t>(n + n / 4, 1024);
        HIPCHK(hipHostMalloc(&b->pin_cmds, cap * sizeof(kst_spawn_cmd)));
        b->pin_cmds_cap = cap;
    }
    *out = b->pin_cmds;
    return KST_OK;
}

// The first n staged commands are this block's; the eager accounting of
// src/cpu.rs:330-390, which the device repeats in k_begin from the same inputs.
int spawn_commit(kst_backend *b, size_t n) {
    if (!b) return fail(KST_EINVAL, "null argument");
    if (b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (n > b->pin_cmds_cap) return fail(KST_EINVAL, "%zu spawn commands past the %zu reserved", n,
                                         b->pin_cmds_cap);
    b->n_pending = n;
    const uint64_t cap = b->cfg.max_voices, live = b->live, raw = n;
    const uint64_t want = std::min<uint64_t>(raw, cap);
    uint64_t need = 0;
    if (live + want > cap && b->cfg.steal_rule != STEAL_DROP_NEW)
        need = std::min<uint64_t>(std::min<uint64_t>(live + want - cap, live), b->max_steal);
    const uint64_t take = std::min<uint64_t>(want, cap - (live - need));
    b->stolen += need;
    b->dropped += raw - take;
    return KST_OK;
}

}  // namespace kst

extern "C" {

void kst_default_config(kst_config *c) {
    std::memset(c, 0, sizeof *c);
    // src/config.rs:345-409
    c->sample_rate = 48000;
    c->block_frames = 4096;
    c->gate_frames = 32;
    c->reduce_tile = 4;
    c->workgroup_size = 256;
    c->max_render_workgroups = 2048;
    c->max_voices = 1u << 20;
    c->steal_rule = 2;  // Quietest
    c->admit_rule = 0;  // Loudest
    c->max_steal_percent = 25;
    c->steal_fade_frames = 96;
    c->sort_voices = 1;
    c->interpolation = 1;
    c->decay_curve = 0;
    c->release_curve = 0;
    c->env_floor = 1.0e-5f;
    c->max_param_variants = 32;
    c->kahan_reduce = 0;
    c->form = 1;
    c->compensated = 0;
    c->waves_per_group = 1;
    c->sort_digit_bits = 9;
    c->device = 0;
    c->profile = 0;
}

const char *kst_last_error(void) { return g_last_error.c_str(); }

int kst_create(const kst_config *cfg, const int16_t *pool, size_t pool_len,
               const kst_region_params *params_in, size_t n_params, uint32_t n_regions,
               uint32_t max_sample_len, kst_backend **out) {
    if (!cfg || !out) return fail(KST_EINVAL, "null argument");
    *out = nullptr;
    int rc = validate(*cfg);
    if (rc) return rc;
    auto *b = new kst_backend();
    b->cfg = *cfg;
    b->device = cfg->device;
    auto cleanup = [&](int code) {
        delete b;
        return code;
    };
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess)
        return cleanup(fail(KST_EDEVICE, "hipSetDevice(%d): %s", b->device, hipGetErrorString(e)));
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess &&
        cus > 0)
        b->cus = (uint32_t)cus;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->ev_rendered, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->ev_selected, hipEventDisableTiming)) != hipSuccess)
        return cleanup(fail(KST_EDEVICE, "creating streams and events: %s", hipGetErrorString(e)));

    const kst_config &c = b->cfg;
    b->fast = c.form == 1 && fast_render_supported(c.block_frames, c.gate_frames, c.interpolation);
    b->slots = cfg_pool_slots(c);
    b->max_steal = cfg_max_steal(c);
    b->tiles = c.block_frames / c.gate_frames;
    b->gate_words = b->tiles * GATE_SLOTS;
    b->chan_words = b->tiles * BEND_CHANNELS * CHAN_FIELDS;
    b->sort = plan_sort(n_regions, max_sample_len);
    if (c.sort_voices) {
        b->digit_bits = c.sort_digit_bits;
        b->sort_passes = (b->sort.bits + b->digit_bits - 1) / b->digit_bits;
        if (b->sort_passes == 0) b->sort_passes = 1;
    } else {
        b->digit_bits = 1;
        b->sort_passes = 1;
    }
    b->tiles_max = (b->slots + SORT_TILE - 1) / SORT_TILE;
    {
        // Dead keys are those at or above D (the dead region in the key's top
        // field, or the dead bit); the last pass sees key >> s.
        const uint64_t D = c.sort_voices ? (uint64_t)b->sort.dead_region << b->sort.region_shift : 1;
        const uint32_t sl = (b->sort_passes - 1) * b->digit_bits;
        b->dead_digit = (D & ((1ull << sl) - 1)) == 0 ? (uint32_t)(D >> sl) : ~0u;
    }
    // A fourth global round keeps the final single-workgroup pass short once
    // the pool is large enough that three leave tens of thousands of keys.
    b->sel_rounds = c.max_voices > (1u << 22) ? 4 : 3;
    if (b->fast) {
        b->max_rows = (b->slots + 63) / 64;
        b->max_groups = (b->max_rows + c.waves_per_group - 1) / c.waves_per_group;
    } else {
        uint32_t cap = std::min(std::max(c.max_render_workgroups, 1u), 65535u);
        b->max_groups = std::max(1u, std::min(cap, (b->slots + c.workgroup_size - 1) / c.workgroup_size));
    }

    // ---- static data: the bank's pool, padded with zero guard frames so
    // every neighbour read of the last frame stays inside the buffer.
    std::vector<int16_t> padded(pool, pool + pool_len);
    padded.resize(pool_len + 8 + (pool_len & 1), 0);
    if (pool_len >= (1ull << 31))
        return cleanup(fail(KST_EINVAL, "sample pool of %zu frames exceeds the 2^31-frame limit "
                                        "of 32-bit byte offsets", pool_len));
    b->bank_len = (uint32_t)pool_len;
    // -32768 is the one value the -1 clamp of src/cpu.rs:193 changes; a bank
    // without it lets the taps skip the clamp.
    b->bank_clamp = std::find(pool, pool + pool_len, (int16_t)-32768) != pool + pool_len;
    if (b->bank.alloc(padded.size()) != hipSuccess)
        return cleanup(fail(KST_EDEVICE, "allocating the %zu-sample pool", padded.size()));
    if (hipMemcpy(b->bank.p, padded.data(), padded.size() * 2, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(KST_EDEVICE, "uploading the sample pool"));

    // src/gpu/mod.rs:357-382: an empty table gets one neutral entry.
    std::vector<kst_region_params> ptab(params_in, params_in + n_params);
    if (ptab.empty()) ptab.push_back({1, 1, 1, 1, 1, 0, 1, 0, 0, 0, 0, 0});
    b->ppv = (uint32_t)ptab.size();
    // Every variant slot starts as the bank's own table (what src/cpu.rs:228-238
    // falls back to for a variant never uploaded).
    std::vector<kst_region_params> all((size_t)b->ppv * c.max_param_variants);
    for (uint32_t v = 0; v < c.max_param_variants; ++v)
        std::copy(ptab.begin(), ptab.end(), all.begin() + (size_t)v * b->ppv);
    if (b->params.alloc(all.size()) != hipSuccess ||
        hipMemcpy(b->params.p, all.data(), all.size() * sizeof(kst_region_params),
                  hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(KST_EDEVICE, "uploading the params table"));

    // ---- per-block buffers ----
    const size_t pool_words = (size_t)b->slots * VOICE_FIELDS;
    const size_t two_f = 2 * (size_t)c.block_frames;
    for (int k = 0; k < 2; ++k) {
        if (b->pools[k].alloc(pool_words) != hipSuccess)
            return cleanup(fail(KST_EDEVICE,
                                "allocating a %.2f GiB voice pool buffer for %u slots",
                                pool_words * 4.0 / (1 << 30), b->slots));
        if (b->keys[k].alloc(b->slots) != hipSuccess || b->idx[k].alloc(b->slots) != hipSuccess)
            return cleanup(fail(KST_EDEVICE, "allocating the sort buffers"));
    }
    const size_t prows = b->fast ? b->max_rows : b->max_groups;
    const size_t chunks = (b->max_rows + REDUCE_ROWS - 1) / REDUCE_ROWS;
    if (b->partials.alloc(two_f * prows) != hipSuccess ||
        b->rsum.alloc(two_f * 2 * std::max<size_t>(chunks, 1)) != hipSuccess)
        return cleanup(fail(KST_EDEVICE, "allocating the partial sums"));
    const uint32_t radix = 1u << b->digit_bits;
    if (b->sort_hist.alloc((size_t)b->sort_passes * radix * b->tiles_max) != hipSuccess ||
        b->sort_tot.alloc((size_t)b->sort_passes * radix) != hipSuccess ||
        b->sel_cnt.alloc(SEL_MAX_BLOCKS * SEL_BINS) != hipSuccess ||
        b->sel_min.alloc(SEL_MAX_BLOCKS * SEL_BINS) != hipSuccess ||
        b->sel_max.alloc(SEL_MAX_BLOCKS * SEL_BINS) != hipSuccess ||
        b->cand_key.alloc(b->slots) != hipSuccess || b->cand_slot.alloc(b->slots) != hipSuccess ||
        b->st.alloc(1) != hipSuccess)
        return cleanup(fail(KST_EDEVICE, "allocating the select/sort scratch"));
    // Many tiles: the scan runs over groups of tiles in parallel (one scan
    // workgroup per 32 digits would walk thousands of tiles in sequence).
    if (b->tiles_max >= SCAN_GROUPED_MIN_TILES) {
        b->scan_groups = (b->tiles_max + SCAN_GROUP_TILES - 1) / SCAN_GROUP_TILES;
        if (b->scan_grp.alloc((size_t)b->scan_groups * radix) != hipSuccess)
            return cleanup(fail(KST_EDEVICE, "allocating the grouped scan's sums"));
    }
    if (b->fast) {
        // An XCD's list: at most ceil(rows / XCDS) + 7 survivor batches (its
        // chunks, rounded up) and its share of the new voices' batches.
        b->order_stride = (b->max_rows + XCDS - 1) / XCDS + RUN_CHUNKS / XCDS + 2;
        if (b->batch_victim.alloc(b->max_rows + 1) != hipSuccess ||
            b->render_order.alloc((size_t)XCDS * b->order_stride) != hipSuccess)
            return cleanup(fail(KST_EDEVICE, "allocating the render's dispatch lists"));
    }
    // The steal fade's gain for d = 0 .. fade-1 (src/cpu.rs:577-586): the same
    // correctly rounded f32 division and subtraction the per-frame path does.
    {
        // Padded with FADE_PAD gains of 1 before and of 0 after, the frames
        // before a steal's fade and from its last on.
        const uint32_t fade = c.steal_fade_frames;
        std::vector<float> tab(fade + 2 * FADE_PAD, 0.0f);
        for (uint32_t d = 0; d < FADE_PAD; ++d) tab[d] = 1.0f;
        for (uint32_t d = 0; d < fade; ++d) tab[FADE_PAD + d] = 1.0f - (float)d / (float)fade;
        if (b->fade_tab.alloc(tab.size()) != hipSuccess ||
            hipMemcpy(b->fade_tab.p, tab.data(), tab.size() * sizeof(float),
                      hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(KST_EDEVICE, "uploading the fade table"));
        b->fade_lds = fade <= FADE_LDS_MAX ? (uint32_t)tab.size() : 0;
    }
    DevState zero{};
    if (hipMemcpy(b->st.p, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(KST_EDEVICE, "initialising the device state"));
    if (b->ensure_stage(1, 1024) != KST_OK) return cleanup(KST_EDEVICE);

    if (hipHostMalloc(&b->h_gates, (size_t)b->gate_words * 4) != hipSuccess ||
        hipHostMalloc(&b->h_chans, (size_t)b->chan_words * 4) != hipSuccess ||
        hipHostMalloc(&b->pin_small, 4096 * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc(&b->pin_out, two_f * sizeof(float)) != hipSuccess ||
        hipHostMalloc(&b->pin_result, sizeof(BlockResult)) != hipSuccess)
        return cleanup(fail(KST_EDEVICE, "allocating pinned staging"));
    std::memset(b->h_gates, 0, (size_t)b->gate_words * 4);
    std::memset(b->h_chans, 0, (size_t)b->chan_words * 4);
    for (auto &x : b->ev)
        if (hipEventCreate(&x) != hipSuccess)
            return cleanup(fail(KST_EDEVICE, "creating timing events"));
    b->profile = c.profile <= KST_PROFILE_RENDER_SAMPLED ? c.profile : KST_PROFILE_ALL;
    if (b->profile == 0 && getenv("KST_PROFILE") != nullptr) b->profile = KST_PROFILE_ALL;
    {
        auto set = [](const char *name) {
            const char *v = getenv(name);
            return v != nullptr && *v != '\0' && std::strcmp(v, "0") != 0;
        };
        const char *ds = getenv("KST_DEVICE_SYNC");
        b->device_sync = !(ds != nullptr && std::strcmp(ds, "0") == 0) &&
                         !set("AMD_SERIALIZE_KERNEL") && !set("HIP_LAUNCH_BLOCKING");
    }
    *out = b;
    return KST_OK;
}

void kst_destroy(kst_backend *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    delete b;
}

int kst_set_gates(kst_backend *b, const uint32_t *rows, size_t n) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b || (!rows && n)) return fail(KST_EINVAL, "null argument");
    if (n != b->gate_words)
        return fail(KST_EINVAL, "gate rows are %zu words, expected %u", n, b->gate_words);
    std::memcpy(b->h_gates, rows, n * 4);
    return KST_OK;
}

int kst_set_channels(kst_backend *b, const uint32_t *rows, size_t n, int bend, int gain,
                     int variant) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b || (!rows && n)) return fail(KST_EINVAL, "null argument");
    if (n != b->chan_words)
        return fail(KST_EINVAL, "channel rows are %zu words, expected %u", n, b->chan_words);
    std::memcpy(b->h_chans, rows, n * 4);
    b->h_chan_flags = (bend ? CHAN_ACTIVE_BEND : 0) | (gain ? CHAN_ACTIVE_GAIN : 0) |
                      (variant ? CHAN_ACTIVE_VARIANT : 0);
    return KST_OK;
}

int kst_set_params_variant(kst_backend *b, uint32_t index, const kst_region_params *data,
                           size_t n) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b || (!data && n)) return fail(KST_EINVAL, "null argument");
    // src/gpu/mod.rs:972-987
    if (n != b->ppv)
        return fail(KST_EINVAL, "params variant %u has %zu entries, expected %u", index, n, b->ppv);
    if (index >= b->cfg.max_param_variants)
        return fail(KST_EINVAL, "params variant %u is past the configured maximum", index);
    (void)hipSetDevice(b->device);
    HIPCHK(hipMemcpy(b->params.p + (size_t)index * b->ppv, data, n * sizeof(kst_region_params),
                     hipMemcpyHostToDevice));
    return KST_OK;
}

int kst_design_params(kst_backend *b, uint32_t index, const kst_region_design *in, size_t n) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b || (!in && n)) return fail(KST_EINVAL, "null argument");
    if (n != b->ppv)
        return fail(KST_EINVAL, "params variant %u has %zu entries, expected %u", index, n, b->ppv);
    if (index >= b->cfg.max_param_variants)
        return fail(KST_EINVAL, "params variant %u is past the configured maximum", index);
    HIPCHK(hipSetDevice(b->device));
    kst_region_design *d = nullptr;
    HIPCHK(hipMalloc(&d, n * sizeof(kst_region_design)));
    hipError_t e = hipMemcpy(d, in, n * sizeof(kst_region_design), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_design(d, (uint32_t)n, b->params.p + (size_t)index * b->ppv,
                      (double)b->cfg.sample_rate, b->cfg.decay_curve == 0,
                      b->cfg.release_curve == 0, b->cfg.env_floor, b->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(KST_EDEVICE, "designing params: %s", hipGetErrorString(e));
    return KST_OK;
}

int kst_design_params_noalloc(kst_backend *b, uint32_t index, const kst_region_design *in,
                              size_t n) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b || (!in && n)) return fail(KST_EINVAL, "null argument");
    if (n != b->ppv)
        return fail(KST_EINVAL, "params variant %u has %zu entries, expected %u", index, n, b->ppv);
    if (index >= b->cfg.max_param_variants)
        return fail(KST_EINVAL, "params variant %u is past the configured maximum", index);
    if (n > kst::DESIGN_SCRATCH_ROWS)
        return fail(KST_EINVAL, "%zu design rows exceed the allocation-free entry point's "
                                "%zu-row scratch", n, kst::DESIGN_SCRATCH_ROWS);
    HIPCHK(hipSetDevice(b->device));
    kst_region_design *d = nullptr;
    hipError_t e = kst::design_scratch(b->device, &d);
    if (e != hipSuccess) return fail(KST_EDEVICE, "design scratch address: %s", hipGetErrorString(e));
    // Every row the kernel reads is written here first, so nothing of an
    // earlier call's rows is read.
    e = hipMemcpyAsync(d, in, n * sizeof(kst_region_design), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) {
        launch_design(d, (uint32_t)n, b->params.p + (size_t)index * b->ppv,
                      (double)b->cfg.sample_rate, b->cfg.decay_curve == 0,
                      b->cfg.release_curve == 0, b->cfg.env_floor, b->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail(KST_EDEVICE, "designing params: %s", hipGetErrorString(e));
    return KST_OK;
}

int kst_params_export(kst_backend *b, uint32_t index, kst_region_params *out, size_t n) {
    if (!b || (!out && n)) return fail(KST_EINVAL, "null argument");
    if (n != b->ppv)
        return fail(KST_EINVAL, "params variant %u has %zu entries, expected %u", index, n, b->ppv);
    if (index >= b->cfg.max_param_variants)
        return fail(KST_EINVAL, "params variant %u is past the configured maximum", index);
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(out, b->params.p + (size_t)index * b->ppv, n * sizeof(kst_region_params),
                     hipMemcpyDeviceToHost));
    return KST_OK;
}

int kst_spawn(kst_backend *b, const kst_spawn_cmd *cmds, size_t n) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b || (!cmds && n)) return fail(KST_EINVAL, "null argument");
    kst_spawn_cmd *dst = nullptr;
    if (int rc = kst::spawn_reserve(b, n, &dst)) return rc;
    if (n) std::memcpy(dst, cmds, n * sizeof(kst_spawn_cmd));
    return kst::spawn_commit(b, n);
}

int kst_render_submit(kst_backend *b) {
    if (!b) return fail(KST_EINVAL, "null argument");
    if (b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    const size_t two_f = 2 * (size_t)b->cfg.block_frames;
    HIPCHK(hipSetDevice(b->device));
    const uint32_t ncmd = (uint32_t)b->n_pending;
    int rc = b->ensure_stage(1, ncmd);
    if (rc) return rc;
    hipStream_t s = b->stream;
    // Block slot 0: gates, channels, flags, commands, and the block index.
    HIPCHK(hipMemcpyAsync(b->g_gates.p, b->h_gates, b->gate_words * 4,
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->g_chans.p, b->h_chans, b->chan_words * 4,
                          hipMemcpyHostToDevice, s));
    if (ncmd)
        HIPCHK(hipMemcpyAsync(b->g_cmds.p, b->pin_cmds, (size_t)ncmd * sizeof(kst_spawn_cmd),
                              hipMemcpyHostToDevice, s));
    // Flags, command count, command offset 0 and block index 0, as kernel
    // arguments rather than four copies.
    launch_stage_one(b->args(1), b->h_chan_flags, ncmd, s);

    // Host copy of k_begin's arithmetic, for grid sizes only.
    const uint64_t cap = b->cfg.max_voices, live = b->live;
    const uint64_t want = std::min<uint64_t>(ncmd, cap);
    uint64_t need = 0;
    if (live + want > cap && b->cfg.steal_rule != STEAL_DROP_NEW)
        need = std::min<uint64_t>(std::min<uint64_t>(live + want - cap, live), b->max_steal);
    const uint64_t take = std::min<uint64_t>(want, cap - (live - need));
    const uint32_t total = (uint32_t)(live + take);
    b->launch_block((uint32_t)live, total, need > 0, take > 0, b->h_chan_flags != 0,
                    b->profile ? b->ev : nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->pin_out, b->g_out.p, two_f * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(b->pin_result, b->g_results.p, sizeof(BlockResult),
                          hipMemcpyDeviceToHost, s));
    b->inflight = true;
    return KST_OK;
}

int kst_render_collect(kst_backend *b, float *out, size_t n) {
    if (!b || !out) return fail(KST_EINVAL, "null argument");
    const size_t two_f = 2 * (size_t)b->cfg.block_frames;
    if (n != two_f) return fail(KST_EINVAL, "output block is %zu samples, expected %zu", n, two_f);
    if (!b->inflight) return fail(KST_ESTATE, "no render was submitted");
    HIPCHK(hipSetDevice(b->device));
    b->inflight = false;
    HIPCHK(hipStreamSynchronize(b->stream));
    std::memcpy(out, b->pin_out, two_f * 4);
    const BlockResult &r = *b->pin_result;
    b->live = (uint32_t)r.active_voices;
    b->parity ^= 1;
    b->perm_valid = b->fast;
    float pk;
    std::memcpy(&pk, &r.peak_bits, 4);
    b->peak = pk;
    if (r.stolen != b->stolen || r.dropped != b->dropped)
        return fail(KST_ESTATE, "device steal/drop counters (%llu/%llu) disagree with the host's "
                                "(%llu/%llu)",
                    (unsigned long long)r.stolen, (unsigned long long)r.dropped,
                    (unsigned long long)b->stolen, (unsigned long long)b->dropped);
    if (b->profile) b->read_timings(b->ev, 1);
    b->n_pending = 0;
    return KST_OK;
}

int kst_render(kst_backend *b, float *out, size_t n) {
    if (!b || !out) return fail(KST_EINVAL, "null argument");
    const size_t two_f = 2 * (size_t)b->cfg.block_frames;
    if (n != two_f) return fail(KST_EINVAL, "output block is %zu samples, expected %zu", n, two_f);
    if (int rc = kst_render_submit(b)) return rc;
    return kst_render_collect(b, out, n);
}

int kst_stats(const kst_backend *b, kst_block_stats *out) {
    if (!b || !out) return fail(KST_EINVAL, "null argument");
    // After spawn and before render the CPU reference counts the appended
    // voices (src/cpu.rs:641-648); so does this.
    uint64_t active = b->live;
    if (b->n_pending) {
        const uint64_t cap = b->cfg.max_voices, live = b->live;
        const uint64_t want = std::min<uint64_t>(b->n_pending, cap);
        uint64_t need = 0;
        if (live + want > cap && b->cfg.steal_rule != STEAL_DROP_NEW)
            need = std::min<uint64_t>(std::min<uint64_t>(live + want - cap, live), b->max_steal);
        active = live + std::min<uint64_t>(want, cap - (live - need));
    }
    out->active_voices = active;
    out->stolen = b->stolen;
    out->dropped = b->dropped;
    out->peak = b->peak;
    out->pad = 0;
    return KST_OK;
}

const char *kst_name(const kst_backend *b) {
    if (!b) return "";
    return b->fast ? (b->cfg.compensated ? "hip-wave64-compensated" : "hip-wave64") : "hip-plain";
}

int kst_set_profile(kst_backend *b, uint32_t level) {
    if (!b) return fail(KST_EINVAL, "null backend");
    if (level > KST_PROFILE_RENDER_SAMPLED)
        return fail(KST_EINVAL,
                    "profile level %u must be 0, 1 (every pass), 2 (render only) or 3 (render "
                    "only, every %u-th staged block)", level, KST_RENDER_SAMPLE);
    if (b->inflight) return fail(KST_ESTATE, "a render is in flight");
    b->profile = level;
    b->cfg.profile = level;
    b->timings.clear();
    return KST_OK;
}

size_t kst_timings(const kst_backend *b, kst_timing *out, size_t cap) {
    if (!b) return 0;
    size_t n = b->timings.size();
    for (size_t i = 0; i < n && i < cap; ++i) out[i] = b->timings[i];
    return n;
}

int kst_stage_blocks(kst_backend *b, uint32_t nblocks, const uint32_t *gates,
                     const uint32_t *chans, const uint32_t *chan_flags,
                     const kst_spawn_cmd *cmds, const uint32_t *cmd_counts) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b || !gates || !chans || !chan_flags || !cmd_counts || nblocks == 0)
        return fail(KST_EINVAL, "null argument or zero blocks");
    HIPCHK(hipSetDevice(b->device));
    std::vector<uint32_t> offs(nblocks);
    size_t total = 0;
    for (uint32_t i = 0; i < nblocks; ++i) {
        offs[i] = (uint32_t)total;
        total += cmd_counts[i];
    }
    if (total && !cmds) return fail(KST_EINVAL, "null spawn commands");
    if (total > 0xFFFFFFFFull) return fail(KST_EINVAL, "too many staged spawn commands");
    int rc = b->ensure_stage(nblocks, total);
    if (rc) return rc;
    hipStream_t s = b->stream;
    HIPCHK(hipMemcpyAsync(b->g_gates.p, gates, (size_t)nblocks * b->gate_words * 4,
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->g_chans.p, chans, (size_t)nblocks * b->chan_words * 4,
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->g_chan_flags.p, chan_flags, nblocks * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->g_cmd_counts.p, cmd_counts, nblocks * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->g_cmd_offs.p, offs.data(), nblocks * 4, hipMemcpyHostToDevice, s));
    if (total)
        HIPCHK(hipMemcpyAsync(b->g_cmds.p, cmds, total * sizeof(kst_spawn_cmd),
                              hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    b->staged = nblocks;
    // The controller-row render variant only when some staged block has rows.
    b->staged_chan = std::any_of(chan_flags, chan_flags + nblocks, [](uint32_t f) { return f != 0; });
    return KST_OK;
}

int kst_run_staged(kst_backend *b, int use_graph) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b) return fail(KST_EINVAL, "null argument");
    if (!b->staged) return fail(KST_ESTATE, "no staged blocks");
    HIPCHK(hipSetDevice(b->device));
    hipStream_t s = b->stream;
    uint32_t zero = 0;
    HIPCHK(hipMemcpyAsync(&b->st.p->block, &zero, 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(&b->st.p->sync_timeout, &zero, 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    // Capacity-sized grids: the live count is on the device only.
    const uint32_t n = b->slots;
    const bool chan = b->staged_chan;
    // The wave64 form selects each block's victims during the block before;
    // the first block's are selected here. Small pools keep the plain chain:
    // their selection is a few launches, cheaper than the fork and join.
    const bool pipe = b->fast && b->cfg.max_voices >= PIPELINE_MIN_VOICES;
    if (pipe) {
        b->launch_sel(b->cfg.max_voices, false, s);
        launch_begin(b->args(1), s);  // each block's k_end begins the next
        HIPCHK(hipGetLastError());
    }
    if (use_graph) {
        if (b->graph_exec && (b->graph_chan != chan || b->graph_staged != b->staged))
            b->drop_graph();
        if (!b->graph_exec) {
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            b->launch_block(b->cfg.max_voices, n, true, true, chan, nullptr, pipe);
            b->graph_chan = chan;
            b->graph_staged = b->staged;
            hipError_t e = hipStreamEndCapture(s, &b->graph);
            if (e != hipSuccess) return fail(KST_EDEVICE, "graph capture: %s", hipGetErrorString(e));
            HIPCHK(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
        }
        for (uint32_t i = 0; i < b->staged; ++i) HIPCHK(hipGraphLaunch(b->graph_exec, s));
    } else {
        if (b->profile) {
            int rc = b->ensure_stage_events(b->staged);
            if (rc) return rc;
        }
        const uint32_t stride = b->profile == KST_PROFILE_RENDER_SAMPLED ? KST_RENDER_SAMPLE : 1;
        for (uint32_t i = 0; i < b->staged; ++i)
            b->launch_block(b->cfg.max_voices, n, true, true, chan,
                            b->profile && i % stride == 0 ? &b->stage_ev[6 * i] : nullptr, pipe,
                            b->device_sync);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    if (pipe && !use_graph) {
        HIPCHK(hipStreamSynchronize(b->side));
        uint32_t timeout = 0;
        HIPCHK(hipMemcpy(&timeout, &b->st.p->sync_timeout, 4, hipMemcpyDeviceToHost));
        if (timeout)
            return fail(KST_EDEVICE, "a device-side wait between the render and selection "
                                     "streams timed out; the staged results are not valid");
    }
    if (b->profile && !use_graph)
        b->read_timings(b->stage_ev.data(), b->staged,
                        b->profile == KST_PROFILE_RENDER_SAMPLED ? KST_RENDER_SAMPLE : 1);
    // Bring the host mirror up to date from the last block's result row.
    BlockResult r;
    HIPCHK(hipMemcpy(&r, b->g_results.p + (b->staged - 1), sizeof r, hipMemcpyDeviceToHost));
    b->live = (uint32_t)r.active_voices;
    if (b->staged & 1) b->parity ^= 1;
    b->perm_valid = b->fast;
    b->stolen = r.stolen;
    b->dropped = r.dropped;
    std::memcpy(&b->peak, &r.peak_bits, 4);
    return KST_OK;
}

int kst_staged_results(kst_backend *b, float *out, kst_block_stats *stats) {
    if (!b) return fail(KST_EINVAL, "null argument");
    if (!b->staged) return fail(KST_ESTATE, "no staged blocks");
    HIPCHK(hipSetDevice(b->device));
    const size_t two_f = 2 * (size_t)b->cfg.block_frames;
    if (out)
        HIPCHK(hipMemcpy(out, b->g_out.p, b->staged * two_f * 4, hipMemcpyDeviceToHost));
    if (stats) {
        std::vector<BlockResult> r(b->staged);
        HIPCHK(hipMemcpy(r.data(), b->g_results.p, r.size() * sizeof(BlockResult),
                         hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < b->staged; ++i) {
            stats[i].active_voices = r[i].active_voices;
            stats[i].stolen = r[i].stolen;
            stats[i].dropped = r[i].dropped;
            std::memcpy(&stats[i].peak, &r[i].peak_bits, 4);
            stats[i].pad = 0;
        }
    }
    return KST_OK;
}

// ---- batches: many renders, one launch per pass ----

struct kst_batch {
    struct Group {
        uint32_t first = 0, count = 0, interp = 0, wpg = 1, groups = 1, fade_lds = 0;
        bool chan = false, clamp = false;
    };
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<kst_backend *> r;      // launch order: grouped by render variant
    std::vector<Group> groups;
    DevBuf<Args> args;
    uint32_t nr = 0, sel_grid = 1, spawn_grid = 1, sort_grid = 1, gather_grid = 1;
    uint32_t passes = 1, max_radix = 2, max_two_f = 2, max_rows = 1;
    bool comp = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~kst_batch() {
        (void)hipSetDevice(device);
        drop_graph();
        args.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    // One block of every member: the chain of kst_backend::launch_block with
    // each pass launched once for all members.
    void launch_step() {
        const Args *aa = args.p;
        hipStream_t s = stream;
        launch_batch_begin(aa, nr, s);
        launch_batch_select(aa, nr, sel_grid, 3, s);
        launch_batch_spawn(aa, nr, spawn_grid, s);
        for (const Group &g : groups)
            launch_batch_render(aa + g.first, g.count, g.interp, g.chan, g.clamp, g.wpg, g.groups,
                                g.fade_lds, s);
        launch_batch_reduce(aa, nr, comp, max_two_f, max_rows, s);
        launch_batch_sort(aa, nr, sort_grid, gather_grid, passes, max_radix, s);
        launch_batch_end(aa, nr, s);
    }
};

int kst_batch_create(kst_backend *const *renders, size_t n, kst_batch **out) {
    if (!renders || !out || n == 0) return fail(KST_EINVAL, "a batch needs at least one render");
    if (n > 65535) return fail(KST_EINVAL, "%zu renders in one batch (at most 65535)", n);
    *out = nullptr;
    const kst_backend *r0 = renders[0];
    for (size_t i = 0; i < n; ++i) {
        const kst_backend *b = renders[i];
        if (!b) return fail(KST_EINVAL, "render %zu is null", i);
        if (!b->fast) return fail(KST_EINVAL, "render %zu is not the wave64 form", i);
        if (b->device != r0->device)
            return fail(KST_EINVAL, "render %zu is on device %d, the batch on %d", i, b->device,
                        r0->device);
        if ((b->cfg.compensated != 0) != (r0->cfg.compensated != 0))
            return fail(KST_EINVAL, "render %zu differs in compensated", i);
        for (size_t j = 0; j < i; ++j)
            if (renders[j] == b) return fail(KST_EINVAL, "render %zu appears twice", i);
    }
    auto *bt = new kst_batch();
    bt->device = r0->device;
    bt->comp = r0->cfg.compensated != 0;
    bt->nr = (uint32_t)n;
    bt->r.assign(renders, renders + n);
    if (hipSetDevice(bt->device) != hipSuccess ||
        hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking) != hipSuccess ||
        bt->args.alloc(n) != hipSuccess) {
        delete bt;
        return fail(KST_EDEVICE, "creating the batch stream and table");
    }
    *out = bt;
    return KST_OK;
}

void kst_batch_destroy(kst_batch *bt) { delete bt; }

int kst_batch_run(kst_batch *bt, int use_graph) {
    if (!bt) return fail(KST_EINVAL, "null argument");
    HIPCHK(hipSetDevice(bt->device));
    // Group members by render variant (the render kernel's template), largest
    // first inside a group is not needed: every grid covers the largest.
    auto key = [](const kst_backend *b) {
        return (b->cfg.interpolation << 3) | ((b->staged_chan ? 1u : 0u) << 2) |
               ((b->bank_clamp ? 1u : 0u) << 1);
    };
    std::vector<kst_backend *> order(bt->r);
    std::stable_sort(order.begin(), order.end(), [&](const kst_backend *x, const kst_backend *y) {
        return std::make_pair(key(x), x->cfg.waves_per_group) <
               std::make_pair(key(y), y->cfg.waves_per_group);
    });
    std::vector<kst_batch::Group> groups;
    uint32_t steps = 0;
    uint32_t sel = 1, spawn = 1, sort = 1, gather = 1, passes = 1, radix = 2, two_f = 2, rows = 1;
    std::vector<Args> host(order.size());
    for (size_t i = 0; i < order.size(); ++i) {
        kst_backend *b = order[i];
        if (!b->staged) return fail(KST_ESTATE, "a batch member has no staged blocks");
        HIPCHK(hipStreamSynchronize(b->stream));
        if (int rc = b->settle()) return rc;
        const uint32_t n = b->slots;
        host[i] = b->args(b->render_groups_for(n));
        host[i].fused = 0;  // a batch compacts with its own gather pass
        host[i].batch_victim = nullptr;  // and maps its render workgroups statically
        host[i].render_order = nullptr;
        steps = std::max(steps, b->staged);
        sel = std::max(sel, std::max(1u, std::min(std::min(b->cus, SEL_MAX_BLOCKS),
                                                  (b->cfg.max_voices + SEL_THREADS * 4 - 1) /
                                                      (SEL_THREADS * 4))));
        spawn = std::max(spawn, b->per_voice_grid(n));
        sort = std::max(sort, b->sort_grid(n));
        gather = std::max(gather, b->per_voice_grid(n));
        passes = std::max(passes, b->sort_passes);
        radix = std::max(radix, 1u << b->digit_bits);
        two_f = std::max(two_f, 2 * b->cfg.block_frames);
        rows = std::max(rows, (n + 63) / 64);
        const bool chan = b->staged_chan;
        if (groups.empty() || groups.back().interp != b->cfg.interpolation ||
            groups.back().chan != chan || groups.back().clamp != (b->bank_clamp != 0) ||
            groups.back().wpg != b->cfg.waves_per_group) {
            kst_batch::Group g;
            g.first = (uint32_t)i;
            g.interp = b->cfg.interpolation;
            g.chan = chan;
            g.clamp = b->bank_clamp != 0;
            g.wpg = b->cfg.waves_per_group;
            groups.push_back(g);
        }
        kst_batch::Group &g = groups.back();
        g.count += 1;
        g.groups = std::max(g.groups, b->render_groups_for(n));
        g.fade_lds = std::max(g.fade_lds, b->fade_lds);
    }
    const bool same_shape =
        bt->exec && order == bt->r && groups.size() == bt->groups.size() &&
        std::equal(groups.begin(), groups.end(), bt->groups.begin(),
                   [](const kst_batch::Group &x, const kst_batch::Group &y) {
                       return x.first == y.first && x.count == y.count && x.interp == y.interp &&
                              x.chan == y.chan && x.clamp == y.clamp && x.wpg == y.wpg &&
                              x.groups == y.groups && x.fade_lds == y.fade_lds;
                   }) &&
        sel == bt->sel_grid && spawn == bt->spawn_grid && sort == bt->sort_grid &&
        gather == bt->gather_grid && passes == bt->passes && radix == bt->max_radix &&
        two_f == bt->max_two_f && rows == bt->max_rows;
    if (!same_shape) bt->drop_graph();
    bt->r = order;
    bt->groups = groups;
    bt->sel_grid = sel;
    bt->spawn_grid = spawn;
    bt->sort_grid = sort;
    bt->gather_grid = gather;
    bt->passes = passes;
    bt->max_radix = radix;
    bt->max_two_f = two_f;
    bt->max_rows = rows;
    hipStream_t s = bt->stream;
    HIPCHK(hipMemcpyAsync(bt->args.p, host.data(), host.size() * sizeof(Args),
                          hipMemcpyHostToDevice, s));
    uint32_t zero = 0;
    for (kst_backend *b : order)
        HIPCHK(hipMemcpyAsync(&b->st.p->block, &zero, 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    if (use_graph) {
        if (!bt->exec) {
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            bt->launch_step();
            hipError_t e = hipStreamEndCapture(s, &bt->graph);
            if (e != hipSuccess) return fail(KST_EDEVICE, "batch graph capture: %s", hipGetErrorString(e));
            HIPCHK(hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0));
        }
        for (uint32_t i = 0; i < steps; ++i) HIPCHK(hipGraphLaunch(bt->exec, s));
    } else {
        for (uint32_t i = 0; i < steps; ++i) bt->launch_step();
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    // Every member's host mirror from its last block's result row.
    for (kst_backend *b : order) {
        BlockResult res;
        HIPCHK(hipMemcpy(&res, b->g_results.p + (b->staged - 1), sizeof res, hipMemcpyDeviceToHost));
        b->live = (uint32_t)res.active_voices;
        if (b->staged & 1) b->parity ^= 1;
        b->stolen = res.stolen;
        b->dropped = res.dropped;
        std::memcpy(&b->peak, &res.peak_bits, 4);
    }
    return KST_OK;
}

int64_t kst_pool_export(kst_backend *b, uint32_t *soa, size_t stride) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b || !soa) return fail(KST_EINVAL, "null argument");
    if (stride < b->live) return fail(KST_EINVAL, "stride %zu below the live count %u", stride, b->live);
    if (hipSetDevice(b->device) != hipSuccess) return fail(KST_EDEVICE, "hipSetDevice");
    if (int rc = b->settle()) return rc;
    for (uint32_t f = 0; f < VOICE_FIELDS; ++f) {
        if (b->live == 0) break;
        hipError_t e = hipMemcpy(soa + (size_t)f * stride, b->pools[b->parity].p + (size_t)f * b->slots,
                                 (size_t)b->live * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(KST_EDEVICE, "pool export: %s", hipGetErrorString(e));
    }
    return (int64_t)b->live;
}

int kst_pool_import(kst_backend *b, const uint32_t *soa, size_t n, size_t stride) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b || (!soa && n)) return fail(KST_EINVAL, "null argument");
    if (n > b->cfg.max_voices)
        return fail(KST_EINVAL, "%zu voices exceed max_voices %u", n, b->cfg.max_voices);
    if (stride < n) return fail(KST_EINVAL, "stride below n");
    HIPCHK(hipSetDevice(b->device));
    for (uint32_t f = 0; f < VOICE_FIELDS && n; ++f)
        HIPCHK(hipMemcpy(b->pools[b->parity].p + (size_t)f * b->slots, soa + (size_t)f * stride,
                         n * 4, hipMemcpyHostToDevice));
    DevState s;
    HIPCHK(hipMemcpy(&s, b->st.p, sizeof s, hipMemcpyDeviceToHost));
    s.live = (uint32_t)n;
    s.parity = b->parity;
    s.perm_valid = 0;  // the imported pool is in order
    HIPCHK(hipMemcpy(b->st.p, &s, sizeof s, hipMemcpyHostToDevice));
    b->live = (uint32_t)n;
    b->perm_valid = false;
    return KST_OK;
}

int kst_set_counters(kst_backend *b, uint64_t stolen, uint64_t dropped) {
    if (b && b->inflight) return fail(KST_ESTATE, "a submitted render has not been collected");
    if (!b) return fail(KST_EINVAL, "null argument");
    HIPCHK(hipSetDevice(b->device));
    DevState s;
    HIPCHK(hipMemcpy(&s, b->st.p, sizeof s, hipMemcpyDeviceToHost));
    s.stolen = stolen;
    s.dropped = dropped;
    HIPCHK(hipMemcpy(b->st.p, &s, sizeof s, hipMemcpyHostToDevice));
    b->stolen = stolen;
    b->dropped = dropped;
    return KST_OK;
}

}  // extern "C"
