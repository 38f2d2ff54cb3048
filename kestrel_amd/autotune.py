"""Autotuner for the batching form: block length and waves per render
workgroup, per class of render (pool size by powers of two, interpolation).

Each candidate is timed on the device on a short synthetic run of its class
(a pool filled by bursts of note-ons, then a few staged blocks), and the one
with the most voice-samples per second wins. Block length changes the output
(gate cadence, steal span), so a render's block length is fixed by the tuner
before its input stream is built; waves per workgroup and the sort radix only
change timing.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import workload as W
from .config import Config

CANDIDATES = [(f, w) for f in (256, 512, 1024) for w in (1, 2)]
MIXED_WPG_MARGIN = 1.03


def size_class(max_voices: int) -> int:
    return int(round(math.log2(max(max_voices, 1))))


def _stage(g, blocks):
    g.stage_blocks(np.concatenate([b.gates for b in blocks]),
                   np.concatenate([b.chans for b in blocks]),
                   np.array([b.chan_flag_word for b in blocks], np.uint32),
                   np.concatenate([b.cmds for b in blocks]),
                   np.array([len(b.cmds) for b in blocks], np.uint32))


def measure(max_voices: int, interpolation: int, block_frames: int, waves_per_group: int,
            warm_s: float = 0.25, timed_s: float = 0.1, seed: int = 0x4B53A7) -> float:
    """Voice-samples/s of one candidate on a saturated synthetic pool."""
    from .backend import HipSynth
    cfg = Config(block_frames=block_frames, max_voices=max_voices, interpolation=interpolation,
                 max_steal_percent=10, compensated=True, waves_per_group=waves_per_group)
    bank = W.make_bank(seed, instruments=8, zones=4, key_lo=0, key_hi=127, frames=16384,
                       loop=(4096, 15000))
    seconds = warm_s + timed_s + 0.05
    rate = 1.5 * max_voices / warm_s  # with the bursts, fills the pool inside the warm-up
    notes = W.burst_notes(seed + 1, seconds, cfg.sample_rate, base_rate_hz=rate,
                          burst_notes_n=max(1, max_voices // 8), burst_every_s=0.1,
                          dur_range_s=(0.5, 4.0))
    bb = W.BlockBuilder(cfg, bank, notes)
    g = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
    try:
        warm = [bb.next_block() for _ in range(max(1, int(warm_s * 48000 / block_frames)))]
        _stage(g, warm)
        g.run_staged(True)
        live0 = g.stats()["active_voices"]
        timed = [bb.next_block() for _ in range(max(1, int(timed_s * 48000 / block_frames)))]
        _stage(g, timed)
        t0 = time.perf_counter()
        g.run_staged(True)
        wall = time.perf_counter() - t0
        _, stats = g.staged_results()
        live, vs = live0, 0.0
        for b, st in zip(timed, stats):
            vs += (live + cfg.admit_take(live, len(b.cmds))) * block_frames
            live = st["active_voices"]
        return vs / wall
    finally:
        g.close()


def tune(specs: list[dict], log=print) -> dict:
    """{(size class, interpolation): (block_frames, waves_per_group)} for the
    classes the specs use."""
    classes = sorted({(size_class(s["max_voices"]), s["interpolation"]) for s in specs})
    best = {}
    for cls, interp in classes:
        pool = 1 << cls
        scores = {c: measure(pool, interp, *c) for c in CANDIDATES}
        # In a batch, members that differ in waves per workgroup need render
        # launches of their own, run one after the other; two waves per
        # workgroup has to win clearly to be worth that.
        eff = {(f, w): v / (MIXED_WPG_MARGIN if w != 1 else 1.0) for (f, w), v in scores.items()}
        pick = max(eff, key=eff.get)
        best[(cls, interp)] = pick
        log(f"autotune pool 2^{cls} interp {interp}: " +
            ", ".join(f"{f}/{w}: {v:.2e}" for (f, w), v in scores.items()) + f" -> {pick}")
    return best
