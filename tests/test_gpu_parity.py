"""Parity of the gfx950 backend with the CPU oracle on identical inputs.

Bar (stated per check below):
* integer state -- phase accumulators, envelope stages, voice order, steal
  victims (stop frames), compaction order -- bit-exact, compared slot by slot
  (the oracle runs in sorted mode when the device sorts);
* float voice state (level, filter memory) bit-exact too: each voice's f32
  arithmetic is the reference's, operation for operation;
* the mix: the device sums in f32 in a fixed tree order, the oracle in f64, so
  the mix nulls to NULL_DB or better relative to the block's reference peak.
"""
import math

import numpy as np
import pytest

from kestrel_amd import workload as W
from kestrel_amd.config import Config

pytestmark = pytest.mark.gpu

NULL_DB = -100.0


def make_pair(cfg, bank, sorted_mode=None):
    from kestrel_amd.backend import HipSynth
    from oracle import oracle
    gpu = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    cpu = oracle.Oracle(cfg, bank.pool, bank.params)
    if sorted_mode if sorted_mode is not None else cfg.sort_voices:
        cpu.set_sorted(bank.n_regions, bank.max_sample_len, True)
    return gpu, cpu


def null_db(a, ref):
    d = float(np.max(np.abs(a.astype(np.float64) - ref.astype(np.float64))))
    peak = float(np.max(np.abs(ref)))
    if d == 0.0:
        return -math.inf
    return 20 * math.log10(d / max(peak, 1e-30))


def step_both(gpu, cpu, blk):
    for be in (gpu, cpu):
        be.set_gates(blk.gates)
        be.set_channels(blk.chans, *blk.flags)
        be.spawn(blk.cmds)
    sg, sc = gpu.stats(), cpu.stats()
    assert (sg["active_voices"], sg["stolen"], sg["dropped"]) == \
        (sc["active_voices"], sc["stolen"], sc["dropped"])
    og = gpu.render()
    oc = cpu.render()
    return og, oc


def compare_pools(gpu, cpu, where):
    pg, pc = gpu.pool_export(), cpu.pool_export()
    assert pg.shape == pc.shape, f"{where}: live {pg.shape[1]} vs {pc.shape[1]}"
    bad = np.flatnonzero((pg != pc).any(axis=1))
    assert bad.size == 0, f"{where}: fields {bad.tolist()} differ"


def run_stream(cfg, bank, builder, blocks, check_pool_every=1, min_live=1):
    gpu, cpu = make_pair(cfg, bank)
    worst = -math.inf
    max_live = 0
    for b in range(blocks):
        blk = builder.next_block()
        og, oc = step_both(gpu, cpu, blk)
        worst = max(worst, null_db(og, oc))
        sg, sc = gpu.stats(), cpu.stats()
        assert (sg["active_voices"], sg["stolen"], sg["dropped"]) == \
            (sc["active_voices"], sc["stolen"], sc["dropped"]), b
        max_live = max(max_live, sg["active_voices"])
        if check_pool_every and b % check_pool_every == 0:
            compare_pools(gpu, cpu, f"block {b}")
    assert max_live >= min_live
    assert worst <= NULL_DB, worst
    gpu.close()
    return worst


@pytest.mark.parametrize("form,sort", [(1, True), (1, False), (0, True), (0, False)])
def test_c1_stream_bit_exact_state_and_null(form, sort):
    cfg, bank, notes = W.workload_c1()
    cfg = cfg.replace(form=form, sort_voices=sort)
    print(f"c1 form={form} sort={sort}", flush=True)
    run_stream(cfg, bank, W.BlockBuilder(cfg, bank, notes), 120, min_live=100)


@pytest.mark.parametrize("interp", [0, 2])
def test_c1_other_interpolations(interp):
    cfg, bank, notes = W.workload_c1()
    cfg = cfg.replace(interpolation=interp)
    run_stream(cfg, bank, W.BlockBuilder(cfg, bank, notes), 60, min_live=100)


@pytest.mark.parametrize("interp", [0, 1, 2])
def test_bank_holding_minus_32768_takes_the_clamped_taps(interp):
    """-32768 is the one sample value the -1 clamp of src/cpu.rs:190-194
    changes; a bank holding it runs the render's clamped tap variants."""
    import dataclasses
    cfg, bank, notes = W.workload_c1()
    pool = bank.pool.copy()
    for s0, n in zip(bank.smp_start, bank.smp_len):
        pool[int(s0) + 5:int(s0) + int(n):97] = -32768
    assert (pool == -32768).sum() > 100
    bank = dataclasses.replace(bank, pool=pool)
    cfg = cfg.replace(interpolation=interp)
    print(f"clamped taps interp={interp}", flush=True)
    run_stream(cfg, bank, W.BlockBuilder(cfg, bank, notes), 40, min_live=100)


@pytest.mark.parametrize("rule,admit", [(2, 0), (0, 0), (1, 1), (2, 1)])
def test_saturated_small_pool_steals_and_drops(rule, admit):
    """Pool of 512 under ~3000 note-ons a second: steal/drop every block."""
    cfg = Config(block_frames=256, max_voices=512, steal_rule=rule, admit_rule=admit,
                 max_steal_percent=10)
    bank = W.make_bank(5, instruments=8, zones=4, key_lo=21, key_hi=108)
    notes = W.poisson_notes(6, 12_000, 4.0, key_lo=21, key_hi=108, mean_dur_s=1.0)
    print(f"saturated rule={rule} admit={admit}", flush=True)
    run_stream(cfg, bank, W.BlockBuilder(cfg, bank, notes), 150, min_live=500)


def test_bends_drive_the_controller_path():
    cfg, bank, notes = W.workload_c1()
    cfg = cfg.replace(interpolation=2)
    bb = W.BlockBuilder(cfg, bank, notes, bends=W.random_walk_bends(9, 2.0, 0.05))
    run_stream(cfg, bank, bb, 60, min_live=100)


@pytest.mark.parametrize("stream", ["c1", "40k"])
def test_compensated_form_matches_and_nulls_deeper(stream):
    """The compensated mixdown keeps the voice state bit-exact and nulls at
    least as deep as the plain one; over 40k voices, where the plain f32 tree
    sum drops low-order bits, strictly deeper."""
    if stream == "c1":
        cfg, bank, notes = W.workload_c1()
        blocks, min_live = 80, 100
        mk = lambda c: W.BlockBuilder(c, bank, notes)
    else:
        cfg = W.config_c2(max_voices=40_000)
        bank = W.make_bank(11, instruments=16, zones=4, key_lo=0, key_hi=127)
        notes = W.burst_notes(12, 0.4, base_rate_hz=150_000.0, burst_notes_n=20_000,
                              burst_every_s=0.25)
        blocks, min_live = 30, 30_000
        mk = lambda c: W.BlockBuilder(c, bank, notes)
    print(f"compensated vs plain, {stream}", flush=True)
    w_plain = run_stream(cfg, bank, mk(cfg), blocks, check_pool_every=5, min_live=min_live)
    cfgc = cfg.replace(compensated=True)
    w_comp = run_stream(cfgc, bank, mk(cfgc), blocks, check_pool_every=5, min_live=min_live)
    print(f"null plain {w_plain:.1f} dB, compensated {w_comp:.1f} dB", flush=True)
    assert w_comp <= w_plain
    if stream == "40k":
        assert w_comp < w_plain


def test_gain_and_variant_rows():
    """Channel gains and a params variant switch mid-stream, as CC7/CC74 would."""
    cfg, bank, notes = W.workload_c1()
    gpu, cpu = make_pair(cfg, bank)
    variant = bank.params.copy()
    variant["b0"] *= np.float32(0.5)
    variant["b1"] *= np.float32(0.5)
    for be in (gpu, cpu):
        be.set_params_variant(1, variant)
    bb = W.BlockBuilder(cfg, bank, notes)
    worst = -math.inf
    for b in range(60):
        blk = bb.next_block()
        ch = blk.chans.reshape(cfg.tiles(), 16, 4).copy()
        if b >= 20:
            ch[:, 3, 1] = np.float32(0.5).view(np.uint32)
            ch[:, 3, 2] = np.float32(0.25).view(np.uint32)
        if b >= 30:
            ch[cfg.tiles() // 2:, 5, 3] = 1
        blk.chans = ch.ravel()
        blk.flags = (False, b >= 20, b >= 30)
        og, oc = step_both(gpu, cpu, blk)
        worst = max(worst, null_db(og, oc))
        compare_pools(gpu, cpu, f"block {b}")
    assert worst <= NULL_DB


def test_staged_run_equals_per_block_calls():
    cfg, bank, notes = W.workload_c1()
    from kestrel_amd.backend import HipSynth
    blocks = [blk for blk in (W.BlockBuilder(cfg, bank, notes).next_block() for _ in range(40))]
    a = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    outs = []
    for blk in blocks:
        a.set_gates(blk.gates)
        a.set_channels(blk.chans, *blk.flags)
        a.spawn(blk.cmds)
        outs.append(a.render().copy())
    pool_a = a.pool_export()
    for use_graph in (False, True):
        b = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
        b.stage_blocks(np.concatenate([x.gates for x in blocks]),
                       np.concatenate([x.chans for x in blocks]),
                       np.array([x.chan_flag_word for x in blocks], np.uint32),
                       np.concatenate([x.cmds for x in blocks]),
                       np.array([len(x.cmds) for x in blocks], np.uint32))
        b.run_staged(use_graph)
        out, stats = b.staged_results()
        assert np.array_equal(out, np.stack(outs)), use_graph
        assert np.array_equal(b.pool_export(), pool_a)
        assert stats[-1]["active_voices"] == a.stats()["active_voices"]
        b.close()


@pytest.mark.parametrize("device_sync", ["1", "0"])
def test_pipelined_staged_run_equals_per_block_calls_while_stealing(device_sync, monkeypatch):
    """A staged run of a pool large enough to select each block's victims on
    the side stream during the block before (with and without the graph; the
    two streams ordered on the device, or by events with KST_DEVICE_SYNC=0)
    gives the per-block calls' outputs, stats and pool bit for bit, over
    blocks that steal."""
    from kestrel_amd.backend import HipSynth
    monkeypatch.setenv("KST_DEVICE_SYNC", device_sync)
    cfg = W.config_c2(max_voices=40_000)
    bank = W.make_bank(11, instruments=16, zones=4, key_lo=0, key_hi=127)
    notes = W.burst_notes(12, 0.6, base_rate_hz=150_000.0, burst_notes_n=20_000,
                          burst_every_s=0.25)
    blocks = [blk for blk in (W.BlockBuilder(cfg, bank, notes).next_block() for _ in range(50))]
    print("pipelined staged run, 40k voices, 50 blocks", flush=True)
    a = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    outs, stats_a = [], []
    for blk in blocks:
        a.set_gates(blk.gates)
        a.set_channels(blk.chans, *blk.flags)
        a.spawn(blk.cmds)
        outs.append(a.render().copy())
        stats_a.append(a.stats())
    assert stats_a[-1]["stolen"] > 0 and sum(
        x["stolen"] != y["stolen"] for x, y in zip(stats_a, stats_a[1:])) >= 10
    pool_a = a.pool_export()
    for use_graph in (False, True):
        b = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
        b.stage_blocks(np.concatenate([x.gates for x in blocks]),
                       np.concatenate([x.chans for x in blocks]),
                       np.array([x.chan_flag_word for x in blocks], np.uint32),
                       np.concatenate([x.cmds for x in blocks]),
                       np.array([len(x.cmds) for x in blocks], np.uint32))
        b.run_staged(use_graph)
        out, stats = b.staged_results()
        assert np.array_equal(out, np.stack(outs)), use_graph
        assert [s["stolen"] for s in stats] == [s["stolen"] for s in stats_a], use_graph
        assert [s["active_voices"] for s in stats] == [s["active_voices"] for s in stats_a]
        assert np.array_equal(b.pool_export(), pool_a), use_graph
        b.close()
    a.close()


def test_repeated_note_ids_steal_every_voice_at_the_threshold():
    """Every block re-spawns the same notes with the same note ids, so steal
    keys tie: the oracle marks every voice whose key is at or below the
    steal_k-th smallest, and so must the device, per block and in a staged run
    whose selection runs ahead on the side stream (pool bit-exact, counters,
    null)."""
    cfg = W.config_c2(max_voices=40_000)
    bank = W.make_bank(11, instruments=16, zones=4, key_lo=0, key_hi=127)
    notes = W.burst_notes(12, 0.6, base_rate_hz=150_000.0, burst_notes_n=20_000,
                          burst_every_s=0.25)
    # A fresh builder per block: block 0's notes (ids from 0) every time.
    blocks = [W.BlockBuilder(cfg, bank, notes).next_block() for _ in range(30)]
    print("repeated note ids, 40k voices, 30 blocks", flush=True)
    gpu, cpu = make_pair(cfg, bank)
    worst = -math.inf
    for b, blk in enumerate(blocks):
        og, oc = step_both(gpu, cpu, blk)
        worst = max(worst, null_db(og, oc))
        compare_pools(gpu, cpu, f"block {b}")
    assert gpu.stats()["stolen"] > 0
    assert worst <= NULL_DB, worst
    gpu.close()
    from kestrel_amd.backend import HipSynth
    from oracle import oracle
    g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    c = oracle.Oracle(cfg, bank.pool, bank.params)
    c.set_sorted(bank.n_regions, bank.max_sample_len, True)
    stage_all(g, blocks)
    g.run_staged(False)
    outs, stats = g.staged_results()
    for b, blk in enumerate(blocks):
        c.set_gates(blk.gates)
        c.set_channels(blk.chans, *blk.flags)
        c.spawn(blk.cmds)
        oc = c.render()
        sc = c.stats()
        assert (stats[b]["active_voices"], stats[b]["stolen"], stats[b]["dropped"]) == \
            (sc["active_voices"], sc["stolen"], sc["dropped"]), b
        assert null_db(outs[b], oc) <= NULL_DB, b
    compare_pools(g, c, "staged run")
    g.close()


def stage_all(g, blocks):
    g.stage_blocks(np.concatenate([x.gates for x in blocks]),
                   np.concatenate([x.chans for x in blocks]),
                   np.array([x.chan_flag_word for x in blocks], np.uint32),
                   np.concatenate([x.cmds for x in blocks]),
                   np.array([len(x.cmds) for x in blocks], np.uint32))


@pytest.mark.parametrize("size", ["pipelined-40k", "c1-4096"])
def test_graph_replays_restaged_with_other_block_counts(size):
    """One backend replays its captured chain over staged runs of 40, then 24,
    then 40 blocks. The chain's kernels hold the staged block count by value,
    so a run of another count must be captured again: every run equals the
    per-block calls over the same blocks (outputs, stats, pool), at a size
    whose selection runs a block ahead on the side stream and at one that keeps
    the plain chain."""
    from kestrel_amd.backend import HipSynth
    if size == "pipelined-40k":
        cfg = W.config_c2(max_voices=40_000)
        bank = W.make_bank(11, instruments=16, zones=4, key_lo=0, key_hi=127)
        notes = W.burst_notes(12, 1.3, base_rate_hz=150_000.0, burst_notes_n=20_000,
                              burst_every_s=0.25)
    else:
        cfg, bank, notes = W.workload_c1()
    bb = W.BlockBuilder(cfg, bank, notes)
    runs = [[bb.next_block() for _ in range(n)] for n in (40, 24, 40)]
    print(f"graph re-staged, {size}", flush=True)
    a = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    b = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    steals = 0
    for k, blocks in enumerate(runs):
        outs, stats_a = [], []
        for blk in blocks:
            a.set_gates(blk.gates)
            a.set_channels(blk.chans, *blk.flags)
            a.spawn(blk.cmds)
            outs.append(a.render().copy())
            stats_a.append(a.stats())
        stage_all(b, blocks)
        b.run_staged(True)
        out, stats = b.staged_results()
        assert np.array_equal(out, np.stack(outs)), k
        for key in ("active_voices", "stolen", "dropped"):
            assert [s[key] for s in stats] == [s[key] for s in stats_a], (k, key)
        assert b.stats()["stolen"] == a.stats()["stolen"], k
        assert np.array_equal(b.pool_export(), a.pool_export()), k
        steals += stats_a[-1]["stolen"] - (stats_a[0]["stolen"])
    if size == "pipelined-40k":
        assert steals > 0
    a.close()
    b.close()


def test_reruns_are_byte_identical():
    cfg, bank, notes = W.workload_c1()
    from kestrel_amd.backend import HipSynth
    res = []
    for _ in range(2):
        g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
        bb = W.BlockBuilder(cfg, bank, notes)
        outs = []
        for _ in range(40):
            blk = bb.next_block()
            g.set_gates(blk.gates)
            g.set_channels(blk.chans, *blk.flags)
            g.spawn(blk.cmds)
            outs.append(g.render().copy())
        res.append(np.concatenate(outs))
        g.close()
    assert res[0].tobytes() == res[1].tobytes()


def test_validation_errors_name_the_flag():
    from kestrel_amd.backend import HipSynth
    from kestrel_amd.abi import BackendError
    cfg, bank, _ = W.workload_c1()
    g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    with pytest.raises(BackendError, match="gate rows"):
        g.set_gates(np.zeros(5, np.uint32))
    with pytest.raises(BackendError, match="past the configured maximum"):
        g.set_params_variant(cfg.max_param_variants, bank.params)
    with pytest.raises(BackendError, match="entries, expected"):
        g.set_params_variant(1, bank.params[:3])
    with pytest.raises(BackendError, match="output block"):
        g.render(np.zeros(7, np.float32))
    g.close()


def test_profile_levels_time_the_passes_they_name():
    """PROFILE_RENDER times the render kernel only, PROFILE_RENDER_SAMPLED the
    same on every 8th staged block, PROFILE_ALL every pass; either way the
    staged blocks' outputs are the unprofiled ones."""
    from kestrel_amd.abi import BackendError
    from kestrel_amd.backend import HipSynth
    from kestrel_amd.config import (PROFILE_ALL, PROFILE_OFF, PROFILE_RENDER,
                                    PROFILE_RENDER_SAMPLED)
    cfg, bank, notes = W.workload_c1()
    outs = {}
    for level in (PROFILE_OFF, PROFILE_RENDER, PROFILE_RENDER_SAMPLED, PROFILE_ALL):
        g = HipSynth(cfg.replace(profile=level), bank.pool, bank.params, bank.n_regions,
                     bank.max_sample_len)
        bb = W.BlockBuilder(cfg, bank, notes)
        blocks = [bb.next_block() for _ in range(12)]
        g.stage_blocks(np.concatenate([b.gates for b in blocks]),
                       np.concatenate([b.chans for b in blocks]),
                       np.array([b.chan_flag_word for b in blocks], np.uint32),
                       np.concatenate([b.cmds for b in blocks]),
                       np.array([len(b.cmds) for b in blocks], np.uint32))
        g.run_staged(False)
        outs[level], _ = g.staged_results()
        names = [n for n, _ in g.timings()]
        assert names == {PROFILE_OFF: [], PROFILE_RENDER: ["render"],
                         PROFILE_RENDER_SAMPLED: ["render"],
                         PROFILE_ALL: ["steal", "spawn", "render", "reduce", "compact"]}[level]
        assert all(ms > 0 for _, ms in g.timings())
        with pytest.raises(BackendError, match="profile level"):
            g.set_profile(4)
        g.close()
    for level in (PROFILE_RENDER, PROFILE_RENDER_SAMPLED, PROFILE_ALL):
        assert np.array_equal(outs[PROFILE_OFF], outs[level])


@pytest.mark.parametrize("rule,digit_bits", [(2, 8), (0, 9), (2, 5)])
def test_multi_tile_sort_and_steal_at_40k_voices(rule, digit_bits):
    """A pool large enough for many sort tiles and select workgroups, saturated
    so every block steals; the oracle runs in sorted mode. Oldest keys (note
    ids) and quietest keys (level over id) take different select rounds; the
    radix width changes the number of sort passes."""
    cfg = W.config_c2(max_voices=40_000).replace(steal_rule=rule, sort_digit_bits=digit_bits)
    bank = W.make_bank(11, instruments=16, zones=4, key_lo=0, key_hi=127)
    notes = W.burst_notes(12, 0.8, base_rate_hz=150_000.0, burst_notes_n=20_000,
                          burst_every_s=0.25)
    print(f"40k voices rule={rule} digit_bits={digit_bits}", flush=True)
    run_stream(cfg, bank, W.BlockBuilder(cfg, bank, notes), 60, check_pool_every=5,
               min_live=39_000)


def test_one_block_at_full_size_matches_the_oracle():
    """The bench workload at ~1M voices: fill the pool on the device, then one
    block from the same snapshot on both sides -- pool bit-exact, mix nulled."""
    from kestrel_amd.backend import HipSynth
    from oracle import oracle
    cfg = W.config_c2()
    _, bank, notes = W.workload_c2(seconds=3.6, cfg=cfg, base_rate_hz=500_000.0)
    bb = W.BlockBuilder(cfg, bank, notes)
    g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    for _ in range(5):
        chunk = [bb.next_block() for _ in range(64)]
        g.stage_blocks(np.concatenate([b.gates for b in chunk]),
                       np.concatenate([b.chans for b in chunk]),
                       np.array([b.chan_flag_word for b in chunk], np.uint32),
                       np.concatenate([b.cmds for b in chunk]),
                       np.array([len(b.cmds) for b in chunk], np.uint32))
        g.run_staged(False)
    st = g.stats()
    print("live after warmup", st, flush=True)
    assert st["active_voices"] > 1_000_000
    snap = g.pool_export()
    o = oracle.Oracle(cfg, bank.pool, bank.params)
    o.set_sorted(bank.n_regions, bank.max_sample_len, True)
    o.pool_import(snap)
    o.set_counters(st["stolen"], st["dropped"])
    blk = bb.next_block()
    og, oc = step_both(g, o, blk)
    assert g.stats()["stolen"] > st["stolen"]   # this block steals
    compare_pools(g, o, "full-size block")
    assert null_db(og, oc) <= NULL_DB
    g.close()
