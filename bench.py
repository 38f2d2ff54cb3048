#!/usr/bin/env python3
"""Headline benchmark: voice-samples/s and x-realtime at ~1M voices.

Default workload (c2, the plain method at saturation): 48 kHz stereo, 512-frame
blocks, a 1,048,576-voice pool with steal-percent 10, linear interpolation, a
16 x 4-zone bank of 32768-frame int16 samples, and dense bursts of note-ons
(calibrated so the pool stays saturated and stealing fires every block).
`--workload c1` runs the small CPU-runnable configuration over its whole
10-second file; `--workload c3` the precision form (2,097,152 voices, stereo
samples as two voices, cubic, +-2 semitone bends, resonant filters designed in
FP64 on the device).

A "step" is one block of the whole hot path on the device: steal select and
marking, spawn, render, mixdown, locality sort and compaction. The saturated
workloads (c2, c3) first run real blocks of their note stream until the pool is
full and every block steals (whatever `--warmup` says), then `--warmup` more
blocks; then the inputs of `--steps` blocks are staged in HBM and run back to
back (the timed region). The run exits non-zero when the timed blocks were not
at saturation (mean voices below the pool size, or a block without steals).
Reported beside it:

* plain: the same blocks through the reference-structured render and reduce,
  inputs staged in HBM before the clock like the chosen form's; plain_sync: the
  reference's call pattern, one synchronous Backend::render call per block with
  its uploads;
* parity: the CPU oracle (a restatement of src/cpu.rs) from the same pool
  snapshot over the first 16 timed blocks (c1: the whole file) -- counters
  every block, the mix nulled, the pool after the window bit-exact;
* cpu_baseline: that oracle window timed with every host thread, and the first
  block on one thread (the reference's CPU backend is single-threaded);
* roofline: the render kernel's algorithmic bytes per launch over its mean
  duration (HIP events on the backend's own stream), against 8 TB/s, with the
  HBM traffic of a counter profile of the same workload when one is committed,
  and from its SQ counter profile the VALU instructions per voice-frame and the
  share of the chip's VALU issue slots they fill in this run's render time;
* precision: the device's and the CPU f32 render's error against an FP64 render
  of the same algorithm (f64 design, f64 voice arithmetic, f64 mix).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voice-samples/s & x-realtime at 1M voices; vs plain GPU path at <=-100dB null"
HBM_PEAK_GBPS = 8000.0
# The reference's own figure for this metric (README.md:13-14): 1,048,576
# concurrent voices at 1.02x realtime, 48 kHz, on an RTX 5060 Laptop GPU through
# wgpu/Vulkan. Only the 1M-voice workload's line is compared with it.
REFERENCE_1M = {"value": round(1_048_576 * 48_000 * 1.02, 1), "unit": "voice-samples/s",
                "x_realtime": 1.02, "voices": 1_048_576,
                "hardware": "RTX 5060 Laptop (wgpu/Vulkan)", "source": "README.md:13-14"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def null_db(a, ref):
    d = float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(ref, np.float64))))
    peak = float(np.max(np.abs(ref)))
    return -math.inf if d == 0.0 else 20.0 * math.log10(d / max(peak, 1e-30))


def stage(backend, blocks):
    backend.stage_blocks(np.concatenate([b.gates for b in blocks]),
                         np.concatenate([b.chans for b in blocks]),
                         np.array([b.chan_flag_word for b in blocks], np.uint32),
                         np.concatenate([b.cmds for b in blocks]),
                         np.array([len(b.cmds) for b in blocks], np.uint32))


def wav_null(dev_blocks, cpu_blocks, sample_rate) -> dict:
    """The three null figures: raw mix absolute (dBFS), and what `kestrel null`
    reports on the two renders after the reference's default post-render step
    (true-peak brickwall, then the hard clamp), each side through its own
    limiter, block by block."""
    from kestrel_amd import post
    dev = np.asarray(dev_blocks, np.float32).reshape(len(dev_blocks), -1)
    cpu = np.asarray(cpu_blocks, np.float64).astype(np.float32).reshape(len(cpu_blocks), -1)
    raw = float(np.max(np.abs(dev.astype(np.float64) - np.asarray(cpu_blocks, np.float64))))
    la = post.Limiter(sample_rate=sample_rate, mode=post.BRICKWALL)
    lb = post.Limiter(sample_rate=sample_rate, mode=post.BRICKWALL)
    a, b = cpu.copy(), dev.copy()
    for i in range(len(a)):
        la.process(a[i])
        post.clamp_block(a[i])
        lb.process(b[i])
        post.clamp_block(b[i])
    r = post.null(a.ravel(), b.ravel())
    return {"raw_null_dbfs": round(20 * math.log10(raw), 1) if raw > 0 else None,
            "wav_null_dbfs": round(r["peak_diff_db"], 1) if r["peak_diff"] > 0 else None,
            "wav_null_rms_dbfs": round(r["rms_diff_db"], 1) if r["rms_diff"] > 0 else None,
            "wav_identical": r["identical"]}


def render_voices(cfg, live0, blocks, stats):
    """Voices entering render per block: live + spawned (src/gpu/mod.rs:1015-1040)."""
    out = []
    live = live0
    for b, st in zip(blocks, stats):
        out.append(live + cfg.admit_take(live, len(b.cmds)))
        live = st["active_voices"]
    return np.array(out, np.float64)


WORKLOADS = {
    "c1": "plain method, small: 48 kHz stereo, 256-frame blocks, 4096-voice pool, steal-percent "
          "25, linear interpolation, 16x4-zone bank of 32768-frame int16 samples, 10,000 "
          "Poisson notes over 10 s (the whole file)",
    "c2": "plain method at saturation: 48 kHz stereo, 512-frame blocks, 1,048,576-voice pool "
          "filled by real blocks until every block steals, steal-percent 10 (Quietest), linear "
          "interpolation, 16x4-zone bank (keys 0-127) of 32768-frame int16 samples; note-ons: "
          "a 500,000 notes/s Poisson background plus 20 ms bursts of 50,000 every 0.5 s, "
          "durations log-uniform 0.5-8 s (the rate that sustains ~1.3M concurrent notes)",
    "c4": "batching form: 64 independent renders (pools 4,096-524,288 log-uniform, 1e3-5e6 "
          "notes over 5-60 s, 4-64 instruments, 4k-128k-frame samples, linear or cubic, block "
          "length and waves per workgroup from the autotuner), one launch per pass for all, "
          "hipGraph-replayed block step",
    "c5": "extreme pool: 48 kHz stereo, 512-frame blocks, 67,108,864-voice pool (6.4 GB per "
          "buffer) started full from a synthetic snapshot, steal-percent 5, cubic, 128x4 stereo "
          "65536-frame regions, 4M notes/s Poisson with 50 ms bursts, exponential 1 s durations, "
          "compensated mixdown",
    "midi": "end to end: a dense format-1 SMF (16 tracks, 2-60 ms notes, 4M note-ons/s by "
            "default) read by the native driver (src/driver.rs block loop: events, gate and "
            "controller rows, ranked admission, voice build) into the 1,048,576-voice backend "
            "(512-frame blocks, steal-percent 10, linear, compensated), then the brickwall "
            "limiter and clamp, one synchronous block at a time as the reference runs",
    "c3": "precision form: 48 kHz stereo, 512-frame blocks, 2,097,152-voice pool filled by "
          "real blocks until every block steals (steal Oldest 10 %), stereo 65536-frame "
          "saw+noise samples as two hard-panned voices, cubic interpolation, +-2 semitone bend "
          "random walks, lowpass Q log-uniform 0.5-20 designed in FP64 on the device, "
          "compensated mixdown; note-ons: 500,000 notes/s Poisson plus 20 ms bursts of 50,000 "
          "every 0.5 s",
}


def timed_profile(args) -> int:
    """The timed run records HIP events around the render kernel of every
    RENDER_SAMPLE-th block (blocks 0, 8, 16, ...): their mean duration is the
    roofline's denominator. Each event packet costs the chain a few
    microseconds, so the other blocks run with none; events between every pass
    would add them to every pass. The per-pass breakdown comes from a replay of
    the same blocks afterwards."""
    from kestrel_amd.config import PROFILE_OFF, PROFILE_RENDER_SAMPLED
    return PROFILE_OFF if args.graph else PROFILE_RENDER_SAMPLED


def setup(name: str, args, seconds: float):
    """(cfg, bank, builder, design_on_device) for a named workload."""
    from kestrel_amd import workload as W
    if name == "c1":
        cfg = W.config_c1(compensated=bool(args.compensated), profile=timed_profile(args))
        cfg, bank, notes = W.workload_c1(cfg=cfg)
        return cfg, bank, W.BlockBuilder(cfg, bank, notes), False
    if name == "c2":
        cfg = W.config_c2(compensated=bool(args.compensated), profile=timed_profile(args))
        _, bank, notes = W.workload_c2(args.seed or 0x4B530002, seconds=seconds, cfg=cfg,
                                       base_rate_hz=500_000.0)
        return cfg, bank, W.BlockBuilder(cfg, bank, notes), False
    if name == "c3":
        cfg = W.config_c3(profile=timed_profile(args))
        _, bank, notes = W.workload_c3(args.seed or 0x4B530003, seconds=seconds, cfg=cfg,
                                       base_rate_hz=500_000.0)
        return cfg, bank, W.BlockBuilder(cfg, bank, notes,
                                         bends=W.random_walk_bends(0x4B53, 2.0, 0.05)), True
    raise SystemExit(f"unknown workload {name!r}; one of {sorted(WORKLOADS)}")


def calibrate_copy():
    import torch
    n = 1 << 28
    src = torch.full((n,), 1.0, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}),
          flush=True)


def run_c4(args):
    """The batching form: every member's window in one batch vs the plain path
    (each member alone, one synchronous reference-structured call per block)."""
    from kestrel_amd import autotune, workload as W
    from kestrel_amd.backend import HipBatch, HipSynth
    warm_s = (args.warmup if args.warmup is not None else 96) * 512 / 48000
    timed_s = (args.steps if args.steps is not None else 96) * 512 / 48000
    specs = W.c4_specs(args.seed or 0x4B530004)
    t = time.time()
    tuned = {} if args.no_tune else autotune.tune(specs, log=log)
    log(f"autotune: {len(tuned)} classes in {time.time() - t:.1f} s")
    members = []
    for sp in specs:
        f, w, bits = tuned.get((autotune.size_class(sp["max_voices"]), sp["interpolation"]),
                               (512, 1, 9))
        cfg, bank, notes = W.c4_render(sp, warm_s + timed_s + 0.1, block_frames=f,
                                       waves_per_group=w, sort_digit_bits=bits)
        bb = W.BlockBuilder(cfg, bank, notes)
        g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
        members.append(dict(spec=sp, cfg=cfg, bank=bank, bb=bb, g=g,
                            warm=max(1, int(round(warm_s * 48000 / f))),
                            timed=max(1, int(round(timed_s * 48000 / f)))))
    log(f"c4: {len(members)} renders built in {time.time() - t:.1f} s, pools "
        f"{sum(m['cfg'].max_voices for m in members)} voices in all")
    batch = HipBatch([m["g"] for m in members])
    t = time.time()
    for m in members:
        stage(m["g"], [m["bb"].next_block() for _ in range(m["warm"])])
    batch.run(True)
    log(f"warmup in {time.time() - t:.1f} s, live {sum(m['g'].stats()['active_voices'] for m in members)}")
    for m in members:
        m["st0"] = m["g"].stats()
        m["snap"] = m["g"].pool_export()
        m["blocks"] = [m["bb"].next_block() for _ in range(m["timed"])]
        stage(m["g"], m["blocks"])
    t0 = time.perf_counter()
    batch.run(True)
    wall = time.perf_counter() - t0
    vs_total, audio_total = 0.0, 0.0
    for m in members:
        m["outs"], m["stats"] = m["g"].staged_results()
        V = render_voices(m["cfg"], m["snap"].shape[1], m["blocks"], m["stats"])
        m["vs"] = float(V.sum()) * m["cfg"].block_frames
        vs_total += m["vs"]
        audio_total += m["timed"] * m["cfg"].block_frames / 48000
    value = vs_total / wall
    log(f"timed: {len(members)} renders, {audio_total:.1f} s of audio in {wall * 1e3:.1f} ms: "
        f"{value:.3e} voice-samples/s, {audio_total / wall:.1f}x realtime in aggregate")
    plain = plain_sync = None
    parity = {}
    if not args.no_plain:
        tp = tq = 0.0
        worst = -math.inf
        equal = same = True
        for m in members:
            pc = m["cfg"].replace(form=0, compensated=False, profile=False)
            b = m["bank"]
            p = HipSynth(pc, b.pool, b.params, b.n_regions, b.max_sample_len)
            # (1) inputs staged in HBM, blocks back to back, one render at a time
            p.pool_import(m["snap"])
            p.set_counters(m["st0"]["stolen"], m["st0"]["dropped"])
            stage(p, m["blocks"])
            t1 = time.perf_counter()
            p.run_staged(False)
            tp += time.perf_counter() - t1
            outs, _ = p.staged_results()
            if np.abs(m["outs"]).max() > 0:
                worst = max(worst, null_db(outs, m["outs"]))
            equal &= bool(np.array_equal(p.pool_export(), m["g"].pool_export()))
            # (2) the reference's call pattern: per-block uploads and a host sync
            p.pool_import(m["snap"])
            p.set_counters(m["st0"]["stolen"], m["st0"]["dropped"])
            outs2 = np.empty_like(m["outs"])
            t1 = time.perf_counter()
            for i, blk in enumerate(m["blocks"]):
                p.set_gates(blk.gates)
                p.set_channels(blk.chans, *blk.flags)
                p.spawn(blk.cmds)
                p.render(outs2[i])
            tq += time.perf_counter() - t1
            same &= bool(np.array_equal(outs2, outs))
            p.close()
        plain = {"value": round(vs_total / tp, 1), "unit": "voice-samples/s",
                 "x_realtime": round(audio_total / tp, 2),
                 "form": "each render alone, reference-structured render+reduce, its inputs "
                         "staged in HBM before the clock, blocks back to back"}
        plain_sync = {"value": round(vs_total / tq, 1), "unit": "voice-samples/s",
                      "x_realtime": round(audio_total / tq, 2),
                      "form": "each render alone, reference-structured render+reduce, one "
                              "synchronous Backend::render call per block with its uploads"}
        parity["plain_vs_batch_null_db_worst"] = round(worst, 1)
        parity["pools_equal_plain_vs_batch"] = equal
        parity["plain_sync_equals_plain"] = same
        log(f"plain: {plain}; plain_sync: {plain_sync}")
    if not args.no_cpu:
        from oracle import oracle
        by_size = sorted(members, key=lambda m: m["snap"].shape[1])
        checked = []
        for m in by_size[:2] + by_size[-2:]:  # the two smallest and the two largest
            b = m["bank"]
            o = oracle.Oracle(m["cfg"], b.pool, b.params)
            o.set_sorted(b.n_regions, b.max_sample_len, True)
            o.pool_import(m["snap"])
            o.set_counters(m["st0"]["stolen"], m["st0"]["dropped"])
            blk = m["blocks"][0]
            o.set_gates(blk.gates)
            o.set_channels(blk.chans, *blk.flags)
            o.spawn(blk.cmds)
            oc = o.render_f64()
            checked.append(dict(render=m["spec"]["index"], voices=int(m["snap"].shape[1]),
                                null_db=round(null_db(m["outs"][0], oc), 1)
                                if np.abs(oc).max() > 0 else None,
                                stats_match=o.stats()["active_voices"] ==
                                m["stats"][0]["active_voices"]))
        parity["cpu_first_block"] = checked
        log(f"cpu parity: {checked}")
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "voice-samples/s", "n_gpus": 1,
        "steps": max(m["timed"] for m in members), "warmup": max(m["warm"] for m in members),
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS["c4"], "name": "c4", "renders": len(members),
                   "audio_seconds": round(audio_total, 2), "wall_ms": round(wall * 1e3, 2),
                   "tuned": {f"2^{c}/{'linear' if i == 1 else 'cubic'}": list(v)
                             for (c, i), v in sorted(tuned.items())}},
        "x_realtime": round(audio_total / wall, 2),
        "plain": plain, "vs_plain": round(value / plain["value"], 3) if plain else None,
        "plain_sync": plain_sync,
        "vs_plain_sync": round(value / plain_sync["value"], 3) if plain_sync else None,
        "parity": parity,
    }
    batch.close()
    for m in members:
        m["g"].close()
    print(json.dumps(line), flush=True)


def run_midi(args):
    """MIDI file -> native driver -> backend -> limiter, block by block. Reports
    the end-to-end rate, the driver's own host cost per block (the same file
    through a second driver with a saturated pool stand-in, no device calls),
    and the oracle's Python driver on the first blocks as the CPU baseline."""
    import tempfile
    from kestrel_amd import workload as W
    from kestrel_amd.backend import HipSynth
    from kestrel_amd.driver import Bank, Driver
    K = args.steps if args.steps is not None else 200
    WU = args.warmup if args.warmup is not None else 200
    rate = args.notes_per_s
    cfg = W.config_c2(compensated=True, limiter=True, limiter_mode=2)
    F = cfg.block_frames
    seconds = (WU + K + 2) * F / cfg.sample_rate
    seed = args.seed or 0x4B530006
    t = time.time()
    desc = W.driver_bank_desc(seed)
    fd, path = tempfile.mkstemp(suffix=".mid")
    os.close(fd)
    size = W.dense_midi(path, seed + 1, int(rate * seconds), seconds)
    bank = Bank(cfg, *desc)
    log(f"midi: {int(rate * seconds)} notes over {seconds:.2f} s, {size / 1e6:.1f} MB SMF, "
        f"bank {bank.n_regions} regions, built in {time.time() - t:.1f} s")
    gpu = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    drv = Driver.open(cfg, bank, path)
    out = np.zeros(F * 2, np.float32)
    for _ in range(WU):
        drv.next_block(gpu, out)
    s0 = drv.stats()
    t0 = time.perf_counter()
    for _ in range(K):
        drv.next_block(gpu, out)
    wall = time.perf_counter() - t0
    s1 = drv.stats()
    live = gpu.stats()["active_voices"]
    notes = s1["notes"] - s0["notes"]
    x_rt = K * F / cfg.sample_rate / wall
    phases = {k: round(v, 4) for k, v in drv.timings().items()}
    log(f"timed: {K} blocks in {wall * 1e3:.1f} ms ({wall / K * 1e3:.3f} ms/block), {x_rt:.2f}x "
        f"realtime, {notes / wall:.3e} note-ons/s, live {live}, dropped "
        f"{s1['dropped'] - s0['dropped']}, last take {s1['last_take']}; host ms/block by phase "
        f"(warm-up included) {phases}")
    gpu.close()

    # The driver alone: same file, the pool held at capacity, its outputs
    # left in place (no copies to Python).
    import ctypes as C
    from kestrel_amd import abi
    lib = abi.load()
    host = Driver.open(cfg, bank, path)
    zero = np.zeros(F * 2, np.float32)
    st = dict(active_voices=cfg.max_voices, stolen=0, dropped=0, peak=0.0)
    kst = abi.KstBlockStats(cfg.max_voices, 0, 0, 0.0, 0)
    cp, cn = C.c_void_p(), C.c_size_t()

    def host_block():
        lib.kst_driver_begin_block(host._h)
        lib.kst_driver_admit(host._h, cfg.max_voices, 0, C.byref(cp), C.byref(cn))
        lib.kst_driver_end_block(host._h, abi.ptr(zero), zero.size, 0, C.byref(kst))
    for _ in range(WU):
        host_block()
    th = time.perf_counter()
    for _ in range(K):
        host_block()
    host_s = (time.perf_counter() - th) / K
    log(f"driver host cost: {host_s * 1e3:.3f} ms/block ({host_s / (F / cfg.sample_rate):.2f} of "
        f"a block's duration)")

    cpu = None
    if not args.no_cpu:
        from oracle import driver as OD
        settings = dict(sample_rate=cfg.sample_rate, decay_curve=cfg.decay_curve,
                        release_curve=cfg.release_curve, filter_enabled=cfg.filter_enabled,
                        env_floor=np.float32(cfg.env_floor), master_volume=cfg.master_volume)
        ob = OD.Bank(settings, *desc[1:], pool_rate=0)
        div, oev = OD.read_midi(path)
        od = OD.Driver(cfg, ob, div, oev)
        nb, tc = 0, time.perf_counter()
        while nb < 3 or time.perf_counter() - tc < 10.0:
            od.begin_block()
            od.admit(cfg.max_voices, 0)
            od.end_block(zero.copy(), 0, st)
            nb += 1
            if nb >= 40:
                break
        tcpu = (time.perf_counter() - tc) / nb
        cpu = {"value": round(od.stats["notes"] / (tcpu * nb), 1), "unit": "note-ons/s",
               "cores": 1, "kind": "port",
               "sample": f"the oracle's Python driver over the first {nb} blocks of the same file "
                         f"(host only, pool at capacity), {tcpu * 1e3:.1f} ms/block"}
        log(f"cpu baseline: {cpu}")
    os.unlink(path)
    line = {
        "metric": "x-realtime end to end (MIDI file -> driver -> 1M-voice render -> limiter)",
        "value": round(x_rt, 3), "unit": "x realtime", "n_gpus": 1, "steps": K, "warmup": WU,
        "ms_per_step": round(wall / K * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS["midi"], "name": "midi", "block_frames": F,
                   "max_voices": cfg.max_voices, "notes_per_s": rate, "live_voices": live},
        "note_ons_per_s": round(notes / wall, 1),
        "driver_host_ms_per_block": round(host_s * 1e3, 4),
        "driver_phases_ms_per_block": phases,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_c5(args):
    """The extreme pool: throughput over the timed blocks, then the first timed
    block again from the same snapshot against the CPU oracle in state-only
    mode (every pool field but the filter memories, which state-only skips)."""
    from kestrel_amd import workload as W
    from kestrel_amd.abi import FIELD
    from kestrel_amd.backend import HipSynth
    K = args.steps if args.steps is not None else 12
    WU = args.warmup if args.warmup is not None else 3
    cfg = W.config_c5(profile=not args.graph)
    seed = args.seed or 0x4B530005
    t = time.time()
    cfg, bank, notes = W.workload_c5(seed, seconds=(WU + K + 2) * 512 / 48000, cfg=cfg)
    bb = W.BlockBuilder(cfg, bank, notes)
    bb.next_id = cfg.max_voices  # the snapshot holds ids 0 .. max_voices-1
    gpu = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    pool = W.synthetic_pool(cfg, bank, cfg.max_voices, seed + 2)
    gpu.pool_import(pool)
    del pool
    log(f"c5: bank {bank.n_regions} stereo regions, {len(notes)} notes, pool of "
        f"{cfg.max_voices} voices imported in {time.time() - t:.1f} s")
    stage(gpu, [bb.next_block() for _ in range(WU)])
    gpu.run_staged(False)
    st0 = gpu.stats()
    snapshot = gpu.pool_export()
    timed = [bb.next_block() for _ in range(K)]
    stage(gpu, timed)
    t0 = time.perf_counter()
    gpu.run_staged(use_graph=args.graph)
    wall = time.perf_counter() - t0
    outs, stats = gpu.staged_results()
    timings = dict(gpu.timings())
    F = cfg.block_frames
    V = render_voices(cfg, snapshot.shape[1], timed, stats)
    vs = float(V.sum()) * F
    value = vs / wall
    x_rt = K * F / cfg.sample_rate / wall
    log(f"timed: {K} blocks in {wall * 1e3:.1f} ms, mean voices {V.mean():.0f}, {value:.3e} "
        f"voice-samples/s, {x_rt:.3f}x realtime; per-pass ms/block {timings}")
    # The render kernel's algorithmic bytes per launch (cubic: four taps).
    step = (snapshot[3].astype(np.float64) * 2 ** 32 + snapshot[2]) / 2 ** 32
    mean_step = float(step.mean()) if step.size else 1.0
    spawned = np.array([min(len(b.cmds), cfg.max_voices) for b in timed], np.float64)
    bytes_block = V * (192.0 + 2.0 * 3 + 2.0 * F * mean_step) + 76.0 * spawned + 8.0 * F
    roofline = None
    if timings.get("render"):
        achieved = float(bytes_block.mean()) / (timings["render"] * 1e-3) / 1e9
        traffic, source = measured_traffic("c5", float(V.mean()))
        vpf, vfrac, vsource = measured_valu("c5", float(V.mean()), timings["render"])
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                    "traffic": traffic, "traffic_source": source,
                    "valu_per_voice_frame": vpf, "valu_issue_frac": vfrac,
                    "valu_source": vsource, "kernel": "k_render_fast",
                    "algorithmic_bytes_per_launch": round(float(bytes_block.mean())),
                    "bytes_per_voice_sample": round(float(bytes_block.sum() / vs), 4),
                    "render_ms_per_block": round(timings["render"], 4)}
    log(f"roofline: {roofline}")
    parity = {}
    cpu_baseline = None
    if not args.no_cpu:
        from oracle import oracle
        gpu.pool_import(snapshot)
        gpu.set_counters(st0["stolen"], st0["dropped"])
        stage(gpu, timed[:1])
        gpu.run_staged(False)
        _, st1 = gpu.staged_results()
        after = gpu.pool_export()
        o = oracle.Oracle(cfg, bank.pool, bank.params, state_only=True)
        o.set_sorted(bank.n_regions, bank.max_sample_len, True)
        o.set_threads(min(16, os.cpu_count() or 1))
        o.pool_import(snapshot)
        o.set_counters(st0["stolen"], st0["dropped"])
        blk = timed[0]
        o.set_gates(blk.gates)
        o.set_channels(blk.chans, *blk.flags)
        o.spawn(blk.cmds)
        tc = time.perf_counter()
        o.render()
        tcpu = time.perf_counter() - tc
        ref = o.pool_export()
        keep = [i for i in range(24) if i not in (FIELD["filt_z1"], FIELD["filt_z2"])]
        parity = {"blocks_checked": 1, "cpu_state_only_s": round(tcpu, 1),
                  "stats_match_cpu": o.stats()["active_voices"] == st1[0]["active_voices"] and
                  o.stats()["stolen"] == st1[0]["stolen"],
                  "stolen_this_block": int(st1[0]["stolen"] - st0["stolen"]),
                  "pool_bit_exact_vs_cpu_but_filter_state": bool(
                      ref.shape == after.shape and np.array_equal(ref[keep], after[keep]))}
        log(f"cpu parity: {parity}")
        del o
        # Throughput sample of the full CPU render: the oracle over a slice of
        # the snapshot (its first 2,097,152 voices), one block with no note-ons.
        slice_n = min(1 << 21, snapshot.shape[1])
        zero_cmds = np.zeros(0, timed[0].cmds.dtype)
        tc_all, tc_one = None, None
        for threads in (host_threads(), 1):
            o = oracle.Oracle(cfg, bank.pool, bank.params)
            o.set_sorted(bank.n_regions, bank.max_sample_len, True)
            o.set_threads(threads)
            o.pool_import(np.ascontiguousarray(snapshot[:, :slice_n]))
            o.set_gates(blk.gates)
            o.set_channels(blk.chans, *blk.flags)
            o.spawn(zero_cmds)
            tc = time.perf_counter()
            o.render_f64()
            tc = time.perf_counter() - tc
            if threads == 1:
                tc_one = tc
            else:
                tc_all = tc
            del o
        sample = f"one block of {F} frames over the first {slice_n} voices of the pool snapshot"
        cpu_baseline = {"value": round(slice_n * F / tc_all, 1), "unit": "voice-samples/s",
                        "cores": host_threads(), "kind": "port",
                        "sample": sample + f" ({tc_all:.1f} s), every host thread",
                        "cores_1": {"value": round(slice_n * F / tc_one, 1),
                                    "unit": "voice-samples/s", "threads": 1,
                                    "sample": sample + f" ({tc_one:.1f} s)"},
                        "host_cores": os.cpu_count(), "cpu_model": cpu_model()}
        log(f"cpu: {cpu_baseline}")
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "voice-samples/s", "n_gpus": 1,
        "steps": K, "warmup": WU, "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS["c5"], "name": "c5", "block_frames": F,
                   "max_voices": cfg.max_voices, "mean_voices": round(float(V.mean()), 1),
                   "form": gpu.name(), "graph": bool(args.graph)},
        "x_realtime": round(x_rt, 4), "timings_ms": timings, "roofline": roofline,
        "cpu_baseline": cpu_baseline, "parity": parity,
    }
    gpu.close()
    print(json.dumps(line), flush=True)


PREFILL_BLOCKS = 64   # blocks per staged prefill run
PREFILL_CHUNKS = 12   # at most this many runs to reach saturation
PARITY_BLOCKS = 16    # consecutive timed blocks checked against the oracle


def prefill(gpu, builder) -> int:
    """Real blocks of the workload, staged 64 at a time, until the pool is full
    and each of a run's last 16 blocks steals. Returns the blocks run."""
    n = 0
    for _ in range(PREFILL_CHUNKS):
        before = gpu.stats()["stolen"]
        stage(gpu, [builder.next_block() for _ in range(PREFILL_BLOCKS)])
        gpu.run_staged(use_graph=False)
        n += PREFILL_BLOCKS
        _, stats = gpu.staged_results()
        tail = [s["stolen"] for s in stats[-17:]]
        if stats[-1]["stolen"] > before and all(b > a for a, b in zip(tail, tail[1:])):
            return n
    raise SystemExit(f"the pool did not saturate in {n} prefill blocks")


def host_threads() -> int:
    """The CPU threads this process may run on (the box's share, not the
    machine's count)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_one_thread(cfg, bank, table, snapshot, st0, timed, V, wl) -> dict:
    """The oracle on one thread, as the reference's CPU backend runs: the first
    timed block (c1: the first 200 blocks)."""
    from oracle import oracle
    o = oracle.Oracle(cfg, bank.pool, table)
    o.set_sorted(bank.n_regions, bank.max_sample_len, True)
    o.set_threads(1)
    o.pool_import(snapshot)
    o.set_counters(st0["stolen"], st0["dropped"])
    nb = min(200, len(timed)) if wl == "c1" else 1
    tc = 0.0
    for blk in timed[:nb]:
        o.set_gates(blk.gates)
        o.set_channels(blk.chans, *blk.flags)
        t = time.perf_counter()
        o.spawn(blk.cmds)
        o.render_f64()
        tc += time.perf_counter() - t
    F = cfg.block_frames
    return {"value": round(float(V[:nb].sum()) * F / tc, 1), "unit": "voice-samples/s",
            "threads": 1, "x_realtime": round(nb * F / cfg.sample_rate / tc, 4),
            "sample": f"{nb} block(s) of {F} frames at {int(V[:nb].mean())} voices "
                      f"({tc:.1f} s)"}


def precision_leg(cfg, bank, table, snapshot, st0, timed, outs, oc, nb) -> dict:
    """Peak and RMS error of the device's block(s) and of the CPU f32 render
    against an FP64 render of the same algorithm: f64 coefficient design from
    the bank's design inputs, f64 voice arithmetic, f64 mix."""
    from oracle import oracle
    o64 = oracle.Oracle(cfg, bank.pool, table, precision=1)
    o64.set_params64(oracle.design_table_f64(bank.design, cfg))
    o64.set_sorted(bank.n_regions, bank.max_sample_len, True)
    o64.set_threads(host_threads())
    o64.pool_import(snapshot)
    o64.set_counters(st0["stolen"], st0["dropped"])
    ref = np.empty((nb, cfg.block_frames * 2), np.float64)
    for i, blk in enumerate(timed[:nb]):
        o64.set_gates(blk.gates)
        o64.set_channels(blk.chans, *blk.flags)
        o64.spawn(blk.cmds)
        ref[i] = o64.render_f64()
    del o64
    peak = float(np.abs(ref).max())
    dg = outs[:nb].astype(np.float64) - ref
    dc = oc[:nb] - ref
    to_db = lambda x: round(20 * math.log10(max(x, 1e-300) / max(peak, 1e-30)), 1)
    return {"vs": f"oracle FP64 render of the first {nb} timed block(s)",
            "peak_ref": round(peak, 6),
            "device_peak_err_db": to_db(float(np.abs(dg).max())),
            "device_rms_err_db": to_db(float(np.sqrt(np.mean(dg * dg)))),
            "cpu_f32_peak_err_db": to_db(float(np.abs(dc).max())),
            "cpu_f32_rms_err_db": to_db(float(np.sqrt(np.mean(dc * dc))))}


def measured_traffic(wl: str, mean_voices: float):
    """HBM bytes per render launch from the committed counter profile of this
    workload, when that profile's mean voice count is within 2 % of this run's;
    else None (the traffic of another workload says nothing about this one)."""
    path = os.path.join(ROOT, "profiles", f"render_traffic_{wl}.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    mv = float(d.get("mean_voices") or 0)
    if d.get("workload") != wl or not mv or abs(mv - mean_voices) > 0.02 * mean_voices:
        return None, None
    return d.get("bytes_per_launch"), f"profiles/render_traffic_{wl}.json " \
        f"(mean voices {mv:.0f})"


# MI355X VALU issue: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every
# 4 cycles at 2.4 GHz.
VALU_WAVE_INSTS_PER_S = 1024 * 2.4e9 / 4


def measured_valu(wl: str, mean_voices: float, render_ms: float):
    """(VALU instructions per voice-frame, VALU issue fraction, source) of the
    render from the committed SQ counter profile of this workload, when its mean
    voice count is within 2 % of this run's: the profile's instruction count,
    scaled to this run's voices, over this run's render time at one wave64
    instruction per SIMD every 4 cycles. Else Nones."""
    path = os.path.join(ROOT, "profiles", f"render_valu_{wl}.json")
    if not os.path.exists(path) or not render_ms:
        return None, None, None
    with open(path) as f:
        d = json.load(f)
    mv = float(d.get("mean_voices") or 0)
    if d.get("workload") != wl or not mv or abs(mv - mean_voices) > 0.02 * mean_voices:
        return None, None, None
    insts = float(d["sq_insts_valu_per_launch"]) * mean_voices / mv
    frac = insts / (VALU_WAVE_INSTS_PER_S * render_ms * 1e-3)
    return d.get("valu_per_voice_frame"), round(frac, 4), \
        f"profiles/render_valu_{wl}.json (mean voices {mv:.0f})"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=None,
                    help="timed blocks (default: 200 for c2, 100 for c3, the whole file for c1)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed blocks after the prefill (default 16; 0 for c1). c2 and c3 "
                         "first run real blocks until the pool is full and stealing every "
                         "block, whatever this says")
    ap.add_argument("--compensated", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true",
                    help="replay the block chain from a hipGraph (no per-pass events)")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--waves-per-group", type=int, default=None)
    ap.add_argument("--sort-digit-bits", type=int, default=None)
    ap.add_argument("--no-tune", action="store_true", help="c4: skip the autotuner (512/1)")
    ap.add_argument("--notes-per-s", type=float, default=4e6, help="midi: note-on rate")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the render "
              "path runs on one GPU (sharding by voice is a design note, not built)")
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "c4":
        run_c4(args)
        return
    if args.workload == "c5":
        run_c5(args)
        return
    if args.workload == "midi":
        run_midi(args)
        return

    from kestrel_amd.backend import HipSynth

    wl = args.workload
    K = args.steps if args.steps is not None else {"c1": 1875, "c2": 200, "c3": 100}[wl]
    WU = args.warmup if args.warmup is not None else (0 if wl == "c1" else 16)
    block_frames = 256 if wl == "c1" else 512
    saturate = wl != "c1"
    seconds = ((PREFILL_CHUNKS * PREFILL_BLOCKS if saturate else 0) + WU + K + 4) * \
        block_frames / 48000
    t = time.time()
    cfg, bank, builder, design = setup(wl, args, seconds)
    for knob in ("waves_per_group", "sort_digit_bits"):
        if getattr(args, knob) is not None:
            cfg = cfg.replace(**{knob: getattr(args, knob)})
    log(f"workload {wl}: {len(builder.notes)} notes, bank {bank.n_regions} regions"
        f"{' (stereo)' if bank.stereo else ''}, built in {time.time() - t:.1f} s")

    gpu = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    table = bank.params
    if design:
        gpu.design_params(0, bank.design)  # FP64 on the device, rounded once to f32
        table = gpu.params_export(0)
    # ---- prefill (saturated workloads): real blocks until the pool is full and
    # stealing every block, whatever --warmup says; then --warmup more blocks ----
    t = time.time()
    pre = prefill(gpu, builder) if saturate else 0
    done = 0
    while done < WU:
        n = min(64, WU - done)
        stage(gpu, [builder.next_block() for _ in range(n)])
        gpu.run_staged(use_graph=False)
        done += n
    st0 = gpu.stats()
    log(f"prefill {pre} blocks + warmup {WU} blocks in {time.time() - t:.1f} s, live "
        f"{st0['active_voices']}, stolen {st0['stolen']}, dropped {st0['dropped']}")
    snapshot = gpu.pool_export()
    live0 = snapshot.shape[1]

    timed = [builder.next_block() for _ in range(K)]
    stage(gpu, timed)
    # ---- timed region: K blocks, inputs resident in HBM ----
    t0 = time.perf_counter()
    gpu.run_staged(use_graph=args.graph)
    t1 = time.perf_counter()
    wall = t1 - t0
    outs, stats = gpu.staged_results()
    timings = dict(gpu.timings())
    V = render_voices(cfg, live0, timed, stats)
    F = cfg.block_frames
    voice_samples = float(V.sum()) * F
    value = voice_samples / wall
    audio_s = K * F / cfg.sample_rate
    x_rt = audio_s / wall
    steals = [b["stolen"] - a["stolen"] for a, b in zip([st0] + stats[:-1], stats)]
    n_steal = sum(1 for x in steals if x > 0)
    log(f"timed: {K} blocks in {wall * 1e3:.1f} ms ({wall / K * 1e3:.3f} ms/block), "
        f"mean voices {V.mean():.0f}, {value:.3e} voice-samples/s, {x_rt:.1f}x realtime, "
        f"blocks with steals {n_steal}/{K}")
    log(f"render ms/block (events around the render kernel): {timings}")
    pool_final = gpu.pool_export()
    saturated = (not saturate) or (float(V.mean()) >= cfg.max_voices and n_steal == K)
    # ---- per-pass breakdown: the first timed blocks again from the snapshot,
    # with events between every pass (outside the timed region) ----
    passes = None
    if not args.graph:
        from kestrel_amd.config import PROFILE_ALL
        nb_b = min(K, 64)
        gpu.pool_import(snapshot)
        gpu.set_counters(st0["stolen"], st0["dropped"])
        stage(gpu, timed[:nb_b])
        gpu.set_profile(PROFILE_ALL)
        gpu.run_staged(False)
        passes = {k: round(v, 4) for k, v in gpu.timings()}
        gpu.set_profile(timed_profile(args))
        log(f"per-pass ms/block (replay of {nb_b} timed blocks with events between passes): "
            f"{passes}")

    # ---- roofline of the render kernel ----
    step = (snapshot[3].astype(np.float64) * 2 ** 32 + snapshot[2]) / 2 ** 32
    mean_step = float(step.mean()) if step.size else 1.0
    taps = {0: 1, 1: 2, 2: 4}[cfg.interpolation]
    spawned = np.array([min(len(b.cmds), cfg.max_voices) for b in timed], np.float64)
    bytes_block = V * (192.0 + 2.0 * (taps - 1) + 2.0 * F * mean_step) + 76.0 * spawned + 8.0 * F
    per_unit = float(bytes_block.sum() / voice_samples)
    render_ms = timings.get("render")
    from kestrel_amd.config import PROFILE_RENDER_SAMPLED, RENDER_SAMPLE
    roofline = None
    if render_ms:
        achieved = float(bytes_block.mean()) / (render_ms * 1e-3) / 1e9
        traffic, source = measured_traffic(wl, float(V.mean()))
        vpf, vfrac, vsource = measured_valu(wl, float(V.mean()), render_ms)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                    "traffic": traffic, "traffic_source": source, "kernel": "k_render_fast",
                    "valu_per_voice_frame": vpf, "valu_issue_frac": vfrac,
                    "valu_source": vsource,
                    "algorithmic_bytes_per_launch": round(float(bytes_block.mean())),
                    "bytes_per_voice_sample": round(per_unit, 4),
                    "render_ms_per_block": round(render_ms, 4),
                    "render_timed_blocks": (len(range(0, K, RENDER_SAMPLE))
                                            if timed_profile(args) == PROFILE_RENDER_SAMPLED
                                            else K)}
    log(f"roofline: {roofline}")

    # ---- CPU baseline + parity over consecutive timed blocks ----
    cpu_baseline = None
    parity = {}
    precision = None
    if not args.no_cpu:
        from oracle import oracle
        nb = K if wl == "c1" else min(PARITY_BLOCKS, K)
        threads = host_threads()
        o = oracle.Oracle(cfg, bank.pool, table)
        o.set_sorted(bank.n_regions, bank.max_sample_len, True)
        o.set_threads(threads)
        o.pool_import(snapshot)
        o.set_counters(st0["stolen"], st0["dropped"])
        oc = np.empty((nb, cfg.block_frames * 2), np.float64)
        tcpu = 0.0
        stats_ok = True
        for i, blk in enumerate(timed[:nb]):
            o.set_gates(blk.gates)
            o.set_channels(blk.chans, *blk.flags)
            tc = time.perf_counter()
            o.spawn(blk.cmds)
            oc[i] = o.render_f64()
            tcpu += time.perf_counter() - tc
            s_o = o.stats()
            stats_ok &= (s_o["active_voices"], s_o["stolen"], s_o["dropped"]) == \
                (stats[i]["active_voices"], stats[i]["stolen"], stats[i]["dropped"])
        oracle_pool = o.pool_export()
        del o
        if nb == K:
            dev_pool = pool_final
        else:
            # The device again from the same snapshot over the same nb blocks.
            gpu.pool_import(snapshot)
            gpu.set_counters(st0["stolen"], st0["dropped"])
            stage(gpu, timed[:nb])
            gpu.run_staged(False)
            dev_pool = gpu.pool_export()
        parity["blocks_checked"] = nb
        parity["null_db_vs_cpu"] = round(null_db(outs[:nb], oc), 1)
        parity.update(wav_null(outs[:nb], oc, cfg.sample_rate))
        parity["stats_match_cpu_every_block"] = bool(stats_ok)
        parity["pool_bit_exact_vs_cpu"] = bool(dev_pool.shape == oracle_pool.shape and
                                               np.array_equal(dev_pool, oracle_pool))
        del dev_pool, oracle_pool
        vs_all = float(V[:nb].sum()) * F / tcpu
        cores_all = {"value": round(vs_all, 1), "unit": "voice-samples/s", "threads": threads,
                     "x_realtime": round(nb * F / cfg.sample_rate / tcpu, 4),
                     "sample": f"{nb} consecutive block(s) of {F} frames at "
                               f"{int(V[:nb].mean())} voices ({tcpu:.1f} s)"}
        cores_1 = cpu_one_thread(cfg, bank, table, snapshot, st0, timed, V, wl)
        cpu_baseline = {"value": cores_all["value"], "unit": "voice-samples/s",
                        "cores": threads, "kind": "port",
                        "sample": cores_all["sample"] + f", {threads} threads: the oracle (a "
                                  "restatement of src/cpu.rs) with voice-partitioned threads "
                                  "and a fixed-order merge of their f64 mixes",
                        "x_realtime": cores_all["x_realtime"],
                        "cores_1": cores_1, "cores_all": cores_all,
                        "host_cores": os.cpu_count(), "cpu_model": cpu_model()}
        log(f"cpu: {cpu_baseline}")
        log(f"parity: {parity}")
        precision = precision_leg(cfg, bank, table, snapshot, st0, timed, outs, oc,
                                  nb if wl == "c1" else 1)
        log(f"precision: {precision}")

    # ---- plain form: reference-structured kernels over the same blocks ----
    plain = plain_sync = None
    if not args.no_plain:
        pcfg = cfg.replace(form=0, compensated=False, profile=False)
        p = HipSynth(pcfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
        if design:
            p.set_params_variant(0, table)
        # (1) inputs staged in HBM before the clock, blocks back to back: the
        # same timed work as the chosen form, only the kernels differ.
        p.pool_import(snapshot)
        p.set_counters(st0["stolen"], st0["dropped"])
        stage(p, timed)
        tp = time.perf_counter()
        p.run_staged(False)
        tp = time.perf_counter() - tp
        pouts, _ = p.staged_results()
        parity["plain_vs_chosen_null_db"] = round(null_db(pouts, outs), 1)
        parity["pools_equal_plain_vs_chosen"] = bool(np.array_equal(p.pool_export(), pool_final))
        plain = {"value": round(voice_samples / tp, 1), "unit": "voice-samples/s",
                 "ms_per_step": round(tp / K * 1e3, 4), "x_realtime": round(audio_s / tp, 2),
                 "form": "reference-structured render+reduce, inputs staged in HBM before the "
                         "clock, blocks back to back (the chosen form's timed work)"}
        # (2) the reference's call pattern: per-block uploads and a host sync.
        p.pool_import(snapshot)
        p.set_counters(st0["stolen"], st0["dropped"])
        pouts2 = np.empty_like(outs)
        tq = time.perf_counter()
        for i, blk in enumerate(timed):
            p.set_gates(blk.gates)
            p.set_channels(blk.chans, *blk.flags)
            p.spawn(blk.cmds)
            p.render(pouts2[i])
        tq = time.perf_counter() - tq
        parity["plain_sync_equals_plain"] = bool(np.array_equal(pouts2, pouts))
        plain_sync = {"value": round(voice_samples / tq, 1), "unit": "voice-samples/s",
                      "ms_per_step": round(tq / K * 1e3, 4),
                      "x_realtime": round(audio_s / tq, 2),
                      "form": "reference-structured render+reduce, one synchronous "
                              "Backend::render call per block with its uploads"}
        log(f"plain: {plain}; plain_sync: {plain_sync}")
        p.close()

    at_1m = cfg.max_voices == REFERENCE_1M["voices"]
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "voice-samples/s",
        "n_gpus": 1, "steps": K, "warmup": WU, "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(value / REFERENCE_1M["value"], 2) if at_1m else None,
        "baseline": REFERENCE_1M if at_1m else None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[wl], "name": wl,
                   "block_frames": F, "max_voices": cfg.max_voices,
                   "mean_voices": round(float(V.mean()), 1),
                   "blocks_with_steals": n_steal, "prefill_blocks": pre,
                   "form": gpu.name(), "graph": bool(args.graph)},
        "x_realtime": round(x_rt, 2),
        "roofline": roofline,
        "passes_ms": passes,
        "cpu_baseline": cpu_baseline,
        "plain": plain,
        "vs_plain": round(value / plain["value"], 3) if plain else None,
        "plain_sync": plain_sync,
        "vs_plain_sync": round(value / plain_sync["value"], 3) if plain_sync else None,
        "parity": parity,
        "precision": precision,
    }
    gpu.close()
    print(json.dumps(line), flush=True)
    if not saturated:
        log(f"error: the timed blocks were not at saturation (mean voices {V.mean():.0f}, "
            f"max_voices {cfg.max_voices}, blocks with steals {n_steal}/{K})")
        sys.exit(3)


if __name__ == "__main__":
    main()
