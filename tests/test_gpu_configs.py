"""Parity of the gfx950 backend with the CPU oracle at each configuration's own
size: the whole C1 file, a 2-second window of the 1M-voice C2 pool (state
only: selection, steal marks and compaction), one full C3 block at 2,097,152
voices, batched C4 members, and three blocks of the 67,108,864-voice C5 pool.

Bar, as in test_gpu_parity: integer and float voice state bit-exact, compared
slot by slot against the oracle in sorted mode (state-only windows leave out
the filter memories, which state-only rendering does not compute); the mix
nulls to -100 dB or better relative to the block's reference peak.
"""
import math
import os

import numpy as np
import pytest

from kestrel_amd import workload as W

pytestmark = pytest.mark.gpu

NULL_DB = -100.0
THREADS = max(1, min(16, os.cpu_count() or 1))
FILTER_FIELDS = (13, 14)  # filt_z1, filt_z2: not computed in state-only mode


def null_db(a, ref):
    d = float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(ref, np.float64))))
    peak = float(np.max(np.abs(ref)))
    return -math.inf if d == 0.0 else 20 * math.log10(d / max(peak, 1e-30))


def stage(g, blocks):
    g.stage_blocks(np.concatenate([b.gates for b in blocks]),
                   np.concatenate([b.chans for b in blocks]),
                   np.array([b.chan_flag_word for b in blocks], np.uint32),
                   np.concatenate([b.cmds for b in blocks]),
                   np.array([len(b.cmds) for b in blocks], np.uint32))


def fill_until_stealing(g, bb, max_chunks=10, per=64):
    """Staged runs of `per` blocks until the pool is full and each of a chunk's
    last 16 blocks steals."""
    for _ in range(max_chunks):
        before = g.stats()["stolen"]
        stage(g, [bb.next_block() for _ in range(per)])
        g.run_staged(False)
        _, stats = g.staged_results()
        tail = [s["stolen"] for s in stats[-17:]]
        if stats[-1]["stolen"] > before and all(b > a for a, b in zip(tail, tail[1:])):
            return g.stats()
    raise AssertionError("the pool never saturated")


def oracle_from(cfg, bank, table, snap, st, state_only=False):
    from oracle import oracle
    o = oracle.Oracle(cfg, bank.pool, table, state_only=state_only)
    o.set_sorted(bank.n_regions, bank.max_sample_len, True)
    o.set_threads(THREADS)
    o.pool_import(snap)
    o.set_counters(st["stolen"], st["dropped"])
    return o


def feed(be, blk):
    be.set_gates(blk.gates)
    be.set_channels(blk.chans, *blk.flags)
    be.spawn(blk.cmds)


def counters(s):
    return s["active_voices"], s["stolen"], s["dropped"]


def same_pool(pg, pc, skip=()):
    assert pg.shape == pc.shape, f"live {pg.shape[1]} vs {pc.shape[1]}"
    keep = [f for f in range(24) if f not in skip]
    bad = [f for f in keep if not np.array_equal(pg[f], pc[f])]
    assert not bad, f"fields {bad} differ"


def test_c1_whole_file():
    """All 1,875 blocks of the 10-second C1 file: pool, counters and the mix
    every block, and the whole file through the default post-render step
    (brickwall, clamp) on both sides, nulled as `kestrel null` does."""
    from kestrel_amd import post
    from kestrel_amd.backend import HipSynth
    from oracle import oracle
    cfg, bank, notes = W.workload_c1()
    cfg = cfg.replace(compensated=True)
    nblocks = math.ceil(10.0 * cfg.sample_rate / cfg.block_frames)
    assert nblocks == 1875
    gpu = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    cpu = oracle.Oracle(cfg, bank.pool, bank.params)
    cpu.set_sorted(bank.n_regions, bank.max_sample_len, True)
    bb = W.BlockBuilder(cfg, bank, notes)
    outs_g = np.empty((nblocks, cfg.block_frames * 2), np.float32)
    outs_c = np.empty((nblocks, cfg.block_frames * 2), np.float64)
    worst, live_max = -math.inf, 0
    for b in range(nblocks):
        blk = bb.next_block()
        feed(gpu, blk)
        feed(cpu, blk)
        assert counters(gpu.stats()) == counters(cpu.stats()), b
        outs_g[b] = gpu.render()
        outs_c[b] = cpu.render_f64()
        assert counters(gpu.stats()) == counters(cpu.stats()), b
        same_pool(gpu.pool_export(), cpu.pool_export())
        if np.abs(outs_c[b]).max() > 0:
            worst = max(worst, null_db(outs_g[b], outs_c[b]))
        live_max = max(live_max, gpu.stats()["active_voices"])
    assert bb.on_ptr == len(notes)  # every note of the file was spawned
    la = post.Limiter(sample_rate=cfg.sample_rate, mode=post.BRICKWALL)
    lb = post.Limiter(sample_rate=cfg.sample_rate, mode=post.BRICKWALL)
    a, c = outs_g.copy(), outs_c.astype(np.float32)
    for i in range(nblocks):
        la.process(a[i])
        post.clamp_block(a[i])
        lb.process(c[i])
        post.clamp_block(c[i])
    r = post.null(c.ravel(), a.ravel())
    print(f"c1 whole file: {nblocks} blocks, max live {live_max}, worst block null "
          f"{worst:.1f} dB, WAV null {r['peak_diff_db']:.1f} dBFS", flush=True)
    assert live_max > 200
    assert worst <= NULL_DB
    assert r["identical"] or r["peak_diff_db"] <= NULL_DB
    gpu.close()


def test_c2_two_second_window_state_only():
    """188 consecutive blocks (2 s) of the saturated 1,048,576-voice pool from a
    device-filled snapshot against the state-only oracle: counters every block,
    the pool (every field but the filter memories) every 8th block and the
    last. Every block steals, so this is selection and compaction over 2 s."""
    from kestrel_amd.backend import HipSynth
    cfg = W.config_c2(compensated=True)
    window = 188
    _, bank, notes = W.workload_c2(seconds=(640 + window + 4) * 512 / 48000, cfg=cfg,
                                   base_rate_hz=500_000.0)
    bb = W.BlockBuilder(cfg, bank, notes)
    g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    st = fill_until_stealing(g, bb)
    snap = g.pool_export()
    assert snap.shape[1] >= cfg.max_voices - cfg.max_steal()
    o = oracle_from(cfg, bank, bank.params, snap, st, state_only=True)
    del snap
    steals = 0
    for b in range(window):
        blk = bb.next_block()
        before = g.stats()["stolen"]  # spawn counts its steals at once, as src/cpu.rs does
        feed(g, blk)
        feed(o, blk)
        assert counters(g.stats()) == counters(o.stats()), b
        g.render()
        o.render()
        assert counters(g.stats()) == counters(o.stats()), b
        steals += g.stats()["stolen"] > before
        if b % 8 == 0 or b == window - 1:
            same_pool(g.pool_export(), o.pool_export(), FILTER_FIELDS)
    print(f"c2 2 s window: {window} blocks, {steals} with steals, live "
          f"{g.stats()['active_voices']}", flush=True)
    assert steals >= window * 9 // 10
    g.close()


def test_c3_full_size_block():
    """One block of the precision form at its own size (2,097,152-voice pool,
    stereo samples as two voices, cubic, bends, device-designed resonant
    filters) from a device-filled snapshot: every field bit-exact, mix nulled."""
    from kestrel_amd.backend import HipSynth
    cfg = W.config_c3()
    _, bank, notes = W.workload_c3(seconds=(640 + 4) * 512 / 48000, cfg=cfg,
                                   base_rate_hz=500_000.0)
    bb = W.BlockBuilder(cfg, bank, notes, bends=W.random_walk_bends(0x4B53, 2.0, 0.05))
    g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    g.design_params(0, bank.design)
    table = g.params_export(0)
    st = fill_until_stealing(g, bb)
    snap = g.pool_export()
    assert snap.shape[1] >= cfg.max_voices - cfg.max_steal()
    o = oracle_from(cfg, bank, table, snap, st)
    del snap
    blk = bb.next_block()
    assert blk.flags[0]  # the bend rows are live
    feed(g, blk)
    feed(o, blk)
    assert counters(g.stats()) == counters(o.stats())
    og = g.render()
    oc = o.render_f64()
    assert counters(g.stats()) == counters(o.stats())
    assert g.stats()["stolen"] > st["stolen"]
    same_pool(g.pool_export(), o.pool_export())
    n = null_db(og, oc)
    print(f"c3 full-size block: {g.stats()['active_voices']} voices, null {n:.1f} dB",
          flush=True)
    assert n <= NULL_DB
    g.close()


C4_MEMBERS = (18, 51, 38, 36, 58, 34, 23, 52)  # largest pool, cubic, saturated, smallest


def test_c4_batched_members_against_the_oracle():
    """Eight renders of the batching workload (the largest and the smallest
    pool, cubic and linear, saturated members that steal every block), 512- and
    1024-frame blocks, one and two waves per workgroup, warmed up in the batch
    through its graph; then the next two batched blocks of each against the
    oracle from the member's own snapshot: counters per block, mix nulled,
    pool bit-exact."""
    from kestrel_amd.backend import HipBatch, HipSynth
    specs = {s["index"]: s for s in W.c4_specs()}
    ms = []
    for k, idx in enumerate(C4_MEMBERS):
        sp = specs[idx]
        f = 1024 if k % 2 else 512
        cfg, bank, notes = W.c4_render(sp, 1.2, block_frames=f, waves_per_group=1 + k % 2)
        g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
        ms.append(dict(idx=idx, cfg=cfg, bank=bank, bb=W.BlockBuilder(cfg, bank, notes), g=g))
    assert any(m["cfg"].interpolation == 2 for m in ms)
    assert max(m["cfg"].max_voices for m in ms) == max(s["max_voices"] for s in specs.values())
    batch = HipBatch([m["g"] for m in ms])
    for m in ms:
        stage(m["g"], [m["bb"].next_block() for _ in range(48)])
    batch.run(True)
    for m in ms:
        m["st"] = m["g"].stats()
        m["o"] = oracle_from(m["cfg"], m["bank"], m["bank"].params, m["g"].pool_export(),
                             m["st"])
        m["blocks"] = [m["bb"].next_block() for _ in range(2)]
        stage(m["g"], m["blocks"])
    batch.run(True)
    stealing = 0
    for m in ms:
        outs, stats = m["g"].staged_results()
        o = m["o"]
        for b, blk in enumerate(m["blocks"]):
            feed(o, blk)
            oc = o.render_f64()
            assert counters(stats[b]) == counters(o.stats()), (m["idx"], b)
            if np.abs(oc).max() > 0:
                assert null_db(outs[b], oc) <= NULL_DB, (m["idx"], b)
        same_pool(m["g"].pool_export(), o.pool_export())
        stealing += stats[-1]["stolen"] > stats[0]["stolen"] or \
            stats[0]["stolen"] > m["st"]["stolen"]
        print(f"c4 member {m['idx']}: pool {m['cfg'].max_voices}, F {m['cfg'].block_frames}, "
              f"interp {m['cfg'].interpolation}, live {stats[-1]['active_voices']}", flush=True)
    assert stealing >= 2
    batch.close()
    for m in ms:
        m["g"].close()


def test_c5_three_blocks_state_only():
    """The 67,108,864-voice pool (cubic, steal 5 %, stereo bank of 128
    instruments): a synthetic full snapshot imported on the device, one device
    block, then three consecutive blocks on both sides from the device's pool
    against the state-only oracle: counters every block (steals every block),
    pool bit-exact but the filter memories after the first and the third."""
    from kestrel_amd.backend import HipSynth
    cfg = W.config_c5()
    seed = 0x4B530005
    cfg, bank, notes = W.workload_c5(seed, seconds=8 * 512 / 48000, cfg=cfg)
    bb = W.BlockBuilder(cfg, bank, notes)
    bb.next_id = cfg.max_voices  # the snapshot holds ids 0 .. max_voices-1
    g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    pool = W.synthetic_pool(cfg, bank, cfg.max_voices, seed + 2)
    g.pool_import(pool)
    del pool
    print("c5: snapshot imported", flush=True)
    stage(g, [bb.next_block()])
    g.run_staged(False)
    st = g.stats()
    snap = g.pool_export()
    assert snap.shape[1] >= cfg.max_voices - cfg.max_steal()
    o = oracle_from(cfg, bank, bank.params, snap, st, state_only=True)
    del snap
    for b in range(3):
        blk = bb.next_block()
        before = g.stats()["stolen"]  # spawn counts its steals at once
        feed(g, blk)
        feed(o, blk)
        assert counters(g.stats()) == counters(o.stats()), b
        g.render()
        o.render()
        assert counters(g.stats()) == counters(o.stats()), b
        assert g.stats()["stolen"] > before, b
        print(f"c5 block {b}: live {g.stats()['active_voices']}, stolen "
              f"{g.stats()['stolen'] - before}", flush=True)
        if b in (0, 2):
            same_pool(g.pool_export(), o.pool_export(), FILTER_FIELDS)
    g.close()
