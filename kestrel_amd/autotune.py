"""Autotuner for the batching form: block length, waves per render workgroup
and the locality sort's radix width, per class of render (pool size by powers
of two, interpolation).

Each candidate is timed on the device on a short synthetic run of its class
(a pool filled by bursts of note-ons, then a few staged blocks), and the one
with the most voice-samples per second wins. The search is coordinate-wise:
block length over 256-8192 frames (the reference's default is 4096) with one
wave per workgroup and 9-bit digits, then waves per workgroup at the chosen
length, then the radix width at both. Block length changes the output (gate
cadence, steal span), so a render's block length is fixed by the tuner before
its input stream is built; waves per workgroup and the radix width only change
timing. Radix widths stop at 9 bits: the sort's scatter keeps per-digit
tables for every wave of its workgroup in LDS.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import workload as W
from .config import Config

BLOCK_FRAMES = (256, 512, 1024, 2048, 4096, 8192)
WAVES = (1, 2)
DIGIT_BITS = (6, 8, 9)
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
            digit_bits: int = 9, warm_s: float = 0.25, timed_s: float = 0.1,
            seed: int = 0x4B53A7) -> float:
    """Voice-samples/s of one candidate on a saturated synthetic pool (at least
    four timed blocks)."""
    from .backend import HipSynth
    cfg = Config(block_frames=block_frames, max_voices=max_voices, interpolation=interpolation,
                 max_steal_percent=10, compensated=True, waves_per_group=waves_per_group,
                 sort_digit_bits=digit_bits)
    n_timed = max(4, int(timed_s * 48000 / block_frames))
    timed_s = n_timed * block_frames / 48000
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
        timed = [bb.next_block() for _ in range(n_timed)]
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
    """{(size class, interpolation): (block_frames, waves_per_group, digit_bits)}
    for the classes the specs use."""
    classes = sorted({(size_class(s["max_voices"]), s["interpolation"]) for s in specs})
    best = {}
    for cls, interp in classes:
        pool = 1 << cls
        scores = {}

        def score(c):
            if c not in scores:
                scores[c] = measure(pool, interp, *c)
            return scores[c]

        f = max(BLOCK_FRAMES, key=lambda f: score((f, 1, 9)))
        # In a batch, members that differ in waves per workgroup need render
        # launches of their own, run one after the other; two waves per
        # workgroup has to win clearly to be worth that.
        w = max(WAVES, key=lambda w: score((f, w, 9)) / (MIXED_WPG_MARGIN if w != 1 else 1.0))
        b = max(DIGIT_BITS, key=lambda b: score((f, w, b)))
        best[(cls, interp)] = (f, w, b)
        log(f"autotune pool 2^{cls} interp {interp}: " +
            ", ".join(f"{c[0]}/{c[1]}/{c[2]}: {v:.2e}" for c, v in scores.items()) +
            f" -> {(f, w, b)}")
    return best
