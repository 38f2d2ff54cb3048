"""Host-side logic: the GateTable port against the reference's own gate tests
(src/voice.rs:673-765), the vectorised block builder against the port, the
velocity curve (src/bank.rs:989-997), and config validation."""
import json
import os

import numpy as np
import pytest

from kestrel_amd import workload as W
from kestrel_amd.config import Config, ConfigError

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_unit_tests.json")))


def gate_cfg():
    g = GOLD["gate_config"]
    return Config(block_frames=g["block_frames"], reduce_tile=g["reduce_tile"],
                  gate_frames=g["gate_frames"])


def play(g, events):
    ords = []
    for ev in events:
        kind = ev[0]
        if kind == "on":
            ords.append(g.note_on(ev[1], ev[2], ev[3]))
        elif kind == "off":
            g.note_off(ev[1], ev[2], ev[3])
        elif kind == "sustain":
            g.set_sustain(ev[1], ev[2], ev[3])
    return ords


def test_gate_rows_reflect_state_at_tile_start():
    t = GOLD["gate_rows_reflect_state_at_tile_start"]
    g = W.GateTable(gate_cfg())
    assert g.tiles == t["expect_tiles"]
    g.begin_block()
    assert play(g, t["events"]) == [t["expect_ordinal"]]
    g.end_block()
    s = W.GateTable.slot(0, 60)
    assert [int(g.row(i)[s]) for i in range(4)] == t["expect_rows"]


def test_the_pedal_defers_note_offs_and_releases_them_together():
    t = GOLD["the_pedal_defers_note_offs_and_releases_them_together"]
    g = W.GateTable(gate_cfg())
    s = W.GateTable.slot(0, 60)
    g.begin_block()
    play(g, t["block1"])
    g.end_block()
    assert g.row(3)[s] == t["block1_row3"] and g.sounding() == t["block1_sounding"]
    g.begin_block()
    play(g, t["block2"])
    g.end_block()
    assert g.row(1)[s] == t["block2_row1"] and g.row(2)[s] == t["block2_row2"]
    assert g.sounding() == t["block2_sounding"]


def test_a_restrike_under_the_pedal_releases_only_what_was_lifted():
    t = GOLD["a_restrike_under_the_pedal_releases_only_what_was_lifted"]
    g = W.GateTable(gate_cfg())
    s = W.GateTable.slot(0, 60)
    g.begin_block()
    play(g, t["block1"])
    g.end_block()
    assert g.sounding() == t["block1_sounding"]
    g.begin_block()
    play(g, t["block2"])
    g.end_block()
    assert g.row(3)[s] == t["block2_row3"] and g.sounding() == t["block2_sounding"]


def test_unmatched_note_off_does_not_pre_release():
    t = GOLD["unmatched_note_off_does_not_pre_release"]
    g = W.GateTable(gate_cfg())
    g.begin_block()
    assert play(g, t["events"]) == [t["expect_ordinal"]]
    g.end_block()
    assert g.row(3)[W.GateTable.slot(0, 60)] == t["expect_row3"]


def test_retrigger_releases_in_order():
    t = GOLD["retrigger_releases_in_order"]
    g = W.GateTable(gate_cfg())
    g.begin_block()
    a, b = play(g, t["events"])
    g.end_block()
    assert [a, b] == t["expect_ordinals"]
    r = g.row(3)[W.GateTable.slot(0, 60)]
    assert a <= r < b


def test_block_builder_matches_gate_table_port():
    """The vectorised builder publishes the rows and ordinals the port does."""
    cfg = Config(block_frames=256, gate_frames=32, max_voices=4096)
    bank = W.make_bank(1, instruments=4, zones=2, frames=4096, loop=(1000, 3000))
    notes = W.poisson_notes(2, 3000, 2.0, key_lo=60, key_hi=64, channels=2,
                            mean_dur_s=0.01)
    bb = W.BlockBuilder(cfg, bank, notes)
    g = W.GateTable(cfg)
    evs = sorted([(int(f), 1, i) for i, f in enumerate(notes.on)]
                 + [(int(f), 0, i) for i, f in enumerate(notes.off)])
    k = 0
    for blk in range(40):
        inp = bb.next_block()
        g.begin_block()
        ords = []
        end = (blk + 1) * cfg.block_frames
        while k < len(evs) and evs[k][0] < end:
            f, kind, i = evs[k]
            rel = f - blk * cfg.block_frames
            if kind == 1:
                ords.append(g.note_on(int(notes.ch[i]), int(notes.key[i]), rel))
            else:
                g.note_off(int(notes.ch[i]), int(notes.key[i]), rel)
            k += 1
        g.end_block()
        assert np.array_equal(inp.gates, g.rows), blk
        assert list(inp.cmds["ordinal"]) == ords, blk


def test_velocity_curve_is_sane():
    t = GOLD["velocity_curve_is_sane"]
    cb = lambda v: -200.0 * np.log10(W.velocity_gain(np.array([v]))[0].astype(np.float64))
    assert cb(127) < t["vel127_max_cb"]
    assert t["vel64_cb_range"][0] <= cb(64) < t["vel64_cb_range"][1]
    assert cb(1) >= t["vel1_min_cb"]


def test_spawn_commands_follow_build_voice():
    cfg, bank, notes = W.workload_c1()
    bb = W.BlockBuilder(cfg, bank, notes)
    cmds = np.concatenate([bb.next_block().cmds for _ in range(20)])
    assert np.array_equal(cmds["note_id_lo"], np.arange(len(cmds)))
    region = cmds["region"]
    key = cmds["gate_slot"] & 127
    ratio = ((cmds["step_hi"].astype(np.float64) * 2 ** 32 + cmds["step_lo"]) / 2 ** 32)
    assert np.allclose(ratio, 2.0 ** ((key - bank.root_key[region]) / 12.0), rtol=1e-9)
    assert np.all(cmds["start_rel"] < cfg.block_frames)
    assert np.all(cmds["gain_l"] == cmds["gain_r"])      # centre pan


@pytest.mark.parametrize("bad", [dict(block_frames=0), dict(reduce_tile=3),
                                 dict(workgroup_size=100), dict(gate_frames=48),
                                 dict(steal_fade_frames=0), dict(max_steal_percent=101),
                                 dict(max_voices=0), dict(block_frames=131072)])
def test_config_validation_names_the_flag(bad):
    with pytest.raises(ConfigError):
        Config(**bad).validate()


def test_default_config_is_valid():
    Config().validate()
    W.config_c1().validate()
    W.config_c2().validate()
