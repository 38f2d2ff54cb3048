"""Batching form: several independent renders run as one batch (one launch per
pass for all members) give, member by member, exactly what each gives alone.

Bar: bitwise-identical output blocks, per-block stats and pool state (every
pass is deterministic and each member's summation order does not depend on
the grid that covers it).
"""
import numpy as np
import pytest

from kestrel_amd import workload as W
from kestrel_amd.config import Config

pytestmark = pytest.mark.gpu


def members():
    """Mixed sizes, block lengths, interpolations, controller rows and sort modes."""
    out = []
    cfg, bank, notes = W.workload_c1()
    out.append((cfg, bank, W.BlockBuilder(cfg, bank, notes), 50))
    cfg = W.config_c3(max_voices=20_000)
    _, bank, notes = W.workload_c3(seconds=1.0, cfg=cfg, base_rate_hz=40_000.0,
                                   burst_notes_n=4_000, burst_every_s=0.25)
    out.append((cfg, bank, W.BlockBuilder(cfg, bank, notes,
                                          bends=W.random_walk_bends(5, 2.0, 0.05)), 35))
    cfg = Config(block_frames=512, max_voices=8192, interpolation=1, max_steal_percent=10,
                 compensated=True)
    bank = W.make_bank(21, instruments=8, zones=4, key_lo=0, key_hi=127)
    notes = W.burst_notes(22, 1.0, base_rate_hz=20_000.0, burst_notes_n=3_000,
                          burst_every_s=0.3)
    out.append((cfg, bank, W.BlockBuilder(cfg, bank, notes), 45))
    cfg = Config(block_frames=256, max_voices=3000, interpolation=0, sort_voices=False,
                 compensated=True)
    bank = W.make_bank(31, instruments=4, zones=4)
    notes = W.poisson_notes(32, 4_000, 2.0, key_lo=21, key_hi=108, mean_dur_s=0.5)
    out.append((cfg, bank, W.BlockBuilder(cfg, bank, notes), 30))
    return out


def stage(g, blocks):
    g.stage_blocks(np.concatenate([b.gates for b in blocks]),
                   np.concatenate([b.chans for b in blocks]),
                   np.array([b.chan_flag_word for b in blocks], np.uint32),
                   np.concatenate([b.cmds for b in blocks]),
                   np.array([len(b.cmds) for b in blocks], np.uint32))


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_each_member_alone(use_graph):
    from kestrel_amd.backend import HipBatch, HipSynth
    ms = members()
    for cfg, *_ in ms:
        cfg.compensated = True
    alone, batched = [], []
    for cfg, bank, bb, nblocks in ms:
        chunks = [[bb.next_block() for _ in range(nblocks)] for _ in range(2)]
        a = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
        b = HipSynth(cfg, bank.pool, bank.params, bank.n_regions, bank.max_sample_len)
        alone.append((a, chunks))
        batched.append(b)
    batch = HipBatch(batched)
    for rnd in range(2):
        for (a, chunks), b in zip(alone, batched):
            stage(a, chunks[rnd])
            a.run_staged(False)
            stage(b, chunks[rnd])
        print(f"round {rnd}: batch of {len(batched)} (graph={use_graph})", flush=True)
        batch.run(use_graph)
        for i, ((a, _), b) in enumerate(zip(alone, batched)):
            oa, sa = a.staged_results()
            ob, sb = b.staged_results()
            assert np.array_equal(oa, ob), (rnd, i)
            assert sa == sb, (rnd, i)
            assert np.array_equal(a.pool_export(), b.pool_export()), (rnd, i)
            assert a.stats() == b.stats(), (rnd, i)
    assert max(b.stats()["active_voices"] for b in batched) > 1000
    batch.close()
    for (a, _), b in zip(alone, batched):
        a.close()
        b.close()
