"""The driver's block loop (src/driver.rs, src/voice.rs:212-658, src/bank.rs,
src/midi.rs) in the library against the oracle's restatement, on the CPU.

Pins of the oracle itself: the reference's own unit tests of the gate table
(src/voice.rs:673-765) and of the bank's curves (src/bank.rs:985-1034), run on
the oracle; closed forms and invariants (mix48 is a bijection, ranked admission
keeps exactly each stratum's loudest, even admission picks i*want/take, the
tempo clock's frame positions); the controller role table of
src/driver.rs:58-131. Then the library: bank tables and variants bit for bit,
the SMF reader event for event (including running status, truncation and
garbage), and whole random streams block by block -- variants, gate rows,
controller rows and flags, spawn commands, limiter output, counters.
"""
import math
import os

import numpy as np
import pytest

import driver_cases as DC
from kestrel_amd.abi import EVENT_DTYPE, PARAMS_DTYPE

OD = pytest.importorskip("oracle.driver")


def lib_bank(cfg, desc):
    from kestrel_amd.driver import Bank
    pool, samples, regions, presets, pre, rate = desc
    return Bank(cfg, pool, samples, regions, presets, pre, pool_rate=rate)


def ora_bank(cfg, desc):
    _, samples, regions, presets, pre, rate = desc
    return OD.Bank(DC.oracle_settings(cfg), samples, regions, presets, pre, pool_rate=rate)


# ---------------------------------------------------------------------------
# The oracle against the reference's own unit tests and closed forms
# ---------------------------------------------------------------------------
def test_oracle_gate_table_reference_tests():
    """src/voice.rs:673-765, on the oracle's gate table (block 64, gate 16)."""
    G = OD.Gates
    s = G.slot(0, 60)
    g = G(4, 16)
    assert g.note_on(0, 60, 0) == 1
    g.note_off(0, 60, 40)
    g.end_block()
    assert list(g.rows[:, s]) == [0, 0, 0, 1]  # the off lands inside tile 2

    g = G(4, 16)
    g.note_on(0, 60, 0)
    g.set_sustain(0, True, 0)
    g.note_off(0, 60, 16)
    g.end_block()
    assert g.rows[3, s] == 0 and g.sounding() == 1
    g.cursor = 0
    g.set_sustain(0, False, 16)
    g.end_block()
    assert g.rows[1, s] == 0 and g.rows[2, s] == 1 and g.sounding() == 0

    g = G(4, 16)  # a restrike under the pedal releases only what was lifted
    g.set_sustain(0, True, 0)
    g.note_on(0, 60, 0)
    g.note_off(0, 60, 0)
    g.note_on(0, 60, 16)
    g.note_off(0, 60, 16)
    g.note_off(0, 60, 16)
    g.end_block()
    assert g.sounding() == 2
    g.cursor = 0
    g.set_sustain(0, False, 0)
    g.end_block()
    assert g.rows[3, s] == 2 and g.sounding() == 0

    g = G(4, 16)  # an unmatched note-off does not pre-release
    g.note_off(0, 60, 0)
    assert g.note_on(0, 60, 16) == 1
    g.end_block()
    assert g.rows[3, s] == 0

    g = G(4, 16)  # retrigger releases in order
    a, b = g.note_on(0, 60, 0), g.note_on(0, 60, 0)
    g.note_off(0, 60, 16)
    g.end_block()
    assert (a, b) == (1, 2) and a <= g.rows[3, s] < b


def test_oracle_bank_curves_reference_tests():
    """src/bank.rs:985-1034 on the oracle's f32 curves."""
    va = OD.velocity_atten_cb
    assert va(127, 100.0) < 0.01
    assert 100.0 <= va(64, 100.0) < 140.0
    assert va(1, 100.0) >= 800.0
    for vel in (1, 40, 64, 100, 127):
        assert va(vel, 0.0) == 0.0
    for vel in (1, 40, 64, 100):
        assert 0.0 < va(vel, 50.0) < va(vel, 100.0)
    assert va(127, 50.0) < 0.01
    assert va(127, -100.0) >= 960.0 and va(1, -100.0) < 0.01
    F = np.float32
    b0, b1, a1, a2 = OD.biquad_lowpass(F(1000), F(0), F(48000))
    h = (F(2) * b0 + b1) / (F(1) + a1 + a2)
    assert abs(h - 1.0) < 0.02
    b0, b1, a1, a2 = OD.biquad_lowpass(F(1000), F(120), F(48000))
    assert (F(2) * b0 + b1) / (F(1) + a1 + a2) < h
    # timecents: 2^(tc/1200) through the same exp2f the bank uses
    assert abs(OD.exp2f(F(0) / F(1200)) - 1.0) < 1e-6
    assert abs(OD.exp2f(F(1200) / F(1200)) - 2.0) < 1e-5


def test_oracle_closed_forms():
    # mix48 is a bijection on 48 bits: distinct on a run of consecutive indices,
    # and neighbouring indices land far apart (channel parity carries no rank).
    xs = [OD.mix48(i) for i in range(1 << 16)]
    assert len(set(xs)) == len(xs) and all(x < 1 << 48 for x in xs)
    assert sum(xs[i] > xs[i + 1] for i in range(0, 1 << 16, 2)) in range(14000, 18800)
    # rank_gain_q is monotonic in |l| + |r| and 0 for silence or NaN.
    gs = np.float32([1e-6, 1e-3, 0.1, 0.5, 1.0, 1.5, 8.0])
    q = [OD.rank_gain_q(g, np.float32(0)) for g in gs]
    assert q == sorted(q) and len(set(q)) == len(q)
    assert OD.rank_gain_q(np.float32(0), np.float32(0)) == 0
    assert OD.rank_gain_q(np.float32("nan"), np.float32(1)) == 0
    # bend_factor: exact octaves, unity, saturation
    assert OD.bend_factor(0.0) == 1 << 24
    assert OD.bend_factor(12.0) == 1 << 25 and OD.bend_factor(-12.0) == 1 << 23
    assert OD.bend_factor(math.nan) == 1 << 24 and OD.bend_factor(1e9) == 0xFFFFFFFF
    # tempo clock: 120 bpm at 480 ppq and 48 kHz is 50 frames a tick; a tempo
    # change applies from its own tick.
    c = OD.TempoClock(("ppq", 480), 48000)
    assert c.frame_at(480) == 24000.0
    c.set_tempo(480, 1_000_000)
    assert c.frame_at(960) == 24000.0 + 48000.0
    s = OD.TempoClock(("smpte", 25, 40), 48000)
    s.set_tempo(10, 1)
    assert s.frame_at(1000) == 48000.0


def test_oracle_admission_keeps_each_strata_loudest():
    """Ranked admission (src/driver.rs:680-749): the admitted set is, stratum by
    stratum, the top share of keys; even admission picks i*want//take."""
    cfg = DC.driver_config(max_voices=50, max_layers=4)
    desc = DC.make_bank_desc(3)
    for rule in (0, 1):
        cfg.admit_rule = rule
        ob = ora_bank(cfg, desc)
        ev = DC.random_events(4, 600, 200, cc_rate=0.0, channels=[0, 1])
        d = OD.Driver(cfg, ob, ("ppq", 480), DC.events_to_oracle(ev))
        checked = 0
        for _ in range(30):
            d.begin_block()
            before = [tuple(c) for c in d.cands]
            live = 40
            d.admit(live, 0)
            want, take = len(before), cfg.admit_take(live, len(before))
            if 0 < take < want:
                got = [tuple(c) for c in d.cands[:take]]
                if rule == 1:
                    assert got == [before[i * want // take] for i in range(take)]
                else:
                    strata = min(take, 64)
                    exp = []
                    for s in range(strata):
                        lo, hi = s * want // strata, (s + 1) * want // strata
                        q = (s + 1) * take // strata - s * take // strata
                        top = sorted(before[lo:hi], key=lambda c: -c[0])[:q]
                        exp.extend(sorted(top, key=lambda c: before.index(c)))
                    assert got == exp
                checked += 1
            out = np.zeros(cfg.block_frames * 2, np.float32)
            d.end_block(out, 0, dict(active_voices=live, stolen=0, dropped=0, peak=0.0))
        assert checked >= 5


# ---------------------------------------------------------------------------
# The library against the oracle
# ---------------------------------------------------------------------------
def test_cc_role_table():
    """Every controller number's role, src/driver.rs:58-131."""
    from kestrel_amd.driver import cc_role
    applied = {0, 1, 6, 7, 10, 11, 32, 33, 38, 39, 42, 43, 64, 66, 67, 71, 72, 73, 74, 75, 76, 77,
               92, 96, 97, 98, 99, 100, 101, 120, 121, 123, 124, 125, 126, 127}
    missing = {5, 12, 13, 65, 68, 69, 78, 84, 88, 91, 93, 94, 95}
    for n in range(128):
        role = cc_role(n)[1]
        want = "applied" if n in applied else "missing" if n in missing else "inert"
        assert role == want, (n, role)


@pytest.mark.parametrize("seed,pool_rate", [(11, 0), (12, 48000)])
def test_bank_tables_match_the_oracle(seed, pool_rate):
    cfg = DC.driver_config(master_volume=0.8)
    desc = DC.make_bank_desc(seed, pool_rate=pool_rate)
    lb, ob = lib_bank(cfg, desc), ora_bank(cfg, desc)
    want = OD.params_array(ob.params, PARAMS_DTYPE)
    assert lb.params.size == want.size > 48
    assert np.array_equal(DC.u32(lb.params), DC.u32(want))
    for cc in ((64, 64, 64, 64, 64), (0, 127, 0, 127, 0), (127, 0, 127, 0, 127),
               (90, 30, 64, 10, 100)):
        got = lb.variant(*cc)
        assert np.array_equal(DC.u32(got), DC.u32(OD.params_array(
            ob.build_variant(OD.param_mod(*cc)), PARAMS_DTYPE))), cc
    for bank in (0, 1, 5, 128, 129):
        for prog in list(range(12)) + [99, 127]:
            assert lb.find_preset(bank, prog) == ob.find_preset(bank, prog), (bank, prog)
    assert lb.n_regions == len(desc[2]) and lb.max_sample_len == int(desc[1]["len"].max())


def test_bank_rejects_what_it_cannot_build():
    from kestrel_amd.abi import BackendError
    cfg = DC.driver_config()
    pool, samples, regions, presets, pre, rate = DC.make_bank_desc(5)
    bad = samples.copy()
    bad[0]["len"] = pool.size + 1
    with pytest.raises(BackendError, match="runs past the pool"):
        lib_bank(cfg, (pool, bad, regions, presets, pre, rate))
    badp = pre.copy()
    badp[0] = len(regions)
    with pytest.raises(BackendError, match="names region"):
        lib_bank(cfg, (pool, samples, regions, presets, badp, rate))


def smf_cases(tmp_path):
    """(name, path) of SMF files covering the reader's branches."""
    ev = DC.random_events(21, 300, 4000)
    cases = []
    p = str(tmp_path / "multi.mid")
    DC.write_smf(p, [DC.track_bytes(ev[0::3]), DC.track_bytes(ev[1::3], use_running=False),
                     DC.track_bytes(ev[2::3])], ppq=96)
    cases.append(("multi", p))
    # sysex, an unknown meta, poly and channel pressure, a system-common byte,
    # a note-on with velocity 0, and a track cut short in the middle of an event
    t = bytearray()
    t += DC.varlen(0) + bytes([0xF0, 3, 1, 2, 0xF7])
    t += DC.varlen(5) + bytes([0xFF, 0x03, 4]) + b"name"
    t += DC.varlen(1) + bytes([0xA2, 60, 10]) + DC.varlen(1) + bytes([0xD3, 50])
    t += DC.varlen(2) + bytes([0x91, 64, 0]) + DC.varlen(0) + bytes([70, 90])  # running status
    t += DC.varlen(3) + bytes([0xF2, 5]) + DC.varlen(1) + bytes([0xE0, 0, 64])
    t += DC.varlen(1) + bytes([0xFF, 0x51, 3, 0x07, 0xA1, 0x20])
    t += DC.varlen(4) + bytes([0xB5, 7])  # truncated
    p = str(tmp_path / "branches.mid")
    DC.write_smf(p, [bytes(t), DC.track_bytes(ev[:40])], ppq=480,
                 extra=b"XTRA" + (4).to_bytes(4, "big") + b"abcd" + bytes(64))
    cases.append(("branches", p))
    rng = np.random.default_rng(7)
    for i in range(6):  # garbage after a valid header
        p = str(tmp_path / f"fuzz{i}.mid")
        DC.write_smf(p, [rng.integers(0, 256, int(rng.integers(1, 400))).astype(np.uint8).tobytes()
                         for _ in range(int(rng.integers(1, 4)))], ppq=int(rng.integers(1, 960)))
        cases.append((f"fuzz{i}", p))
    p = str(tmp_path / "smpte.mid")
    DC.write_smf(p, [DC.track_bytes(ev[:50])], ppq=(0x10000 - (25 << 8)) | 40)  # -25 fps, 40 tpf
    cases.append(("smpte", p))
    return cases


def test_midi_reader_matches_the_oracle(tmp_path):
    from kestrel_amd.driver import MidiFile
    for name, path in smf_cases(tmp_path):
        m = MidiFile(path)
        got = m.events()
        div, want = OD.read_midi(path)
        assert len(got) == len(want), name
        for e, (tick, w) in zip(got, want):
            assert (int(e["tick"]), int(e["kind"]), int(e["ch"]), int(e["a"]), int(e["b"]),
                    int(e["value"])) == (tick, *w), name
        d = m.division
        assert (("smpte", d.fps, d.ticks_per_frame) if d.smpte else ("ppq", d.ppq)) == div, name
    assert len(MidiFile(smf_cases(tmp_path)[0][1]).events()) == 2 * 300 + int(300 * 0.4) + 3


def test_midi_reader_errors(tmp_path):
    from kestrel_amd.abi import BackendError
    from kestrel_amd.driver import MidiFile
    p = tmp_path / "short.mid"
    p.write_bytes(b"MThd\0\0")
    with pytest.raises(BackendError, match="too short for an SMF header"):
        MidiFile(str(p))
    p = tmp_path / "riff.mid"
    p.write_bytes(b"RIFF" + bytes(20))
    with pytest.raises(BackendError, match="not a standard MIDI file"):
        MidiFile(str(p))
    p = tmp_path / "empty.mid"
    DC.write_smf(str(p), [])
    with pytest.raises(BackendError, match="no MTrk chunks"):
        MidiFile(str(p))


CASES = [
    dict(max_voices=40),
    dict(admit_rule=1, limiter_mode=1, max_voices=12),
    dict(steal_rule=1, limiter=False, block_frames=512, gate_frames=64),
    dict(max_param_variants=2, max_layers=16, limiter_mode=0, max_voices=2000),
    dict(decay_curve=1, release_curve=1, filter_enabled=False, default_channel_volume=100),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_driver_blocks_match_the_oracle(case):
    from kestrel_amd.driver import Division, Driver
    cfg = DC.driver_config(**CASES[case])
    desc = DC.make_bank_desc(30 + case)
    ev = DC.random_events(40 + case, 900, 1500)
    lb, ob = lib_bank(cfg, desc), ora_bank(cfg, desc)
    ld = Driver(cfg, lb, division=Division.ppq_of(480), events=ev)
    od = OD.Driver(cfg, ob, ("ppq", 480), DC.events_to_oracle(ev))
    busy = DC.run_pair(cfg, desc, ld, od, blocks=160, seed=case)
    assert busy > 20
    st = ld.stats()
    assert st["notes"] > 100 and st["events"] > 500
    if case in (0, 1):
        assert st["dropped"] > 0  # admission had to choose
    if case == 3:
        assert st["variant_rebuilds"] + st["variant_fallbacks"] > 0


def test_driver_from_a_file_matches_the_oracle(tmp_path):
    """kst_driver_open over a multi-track file with tempo changes, to its end
    and through the tail."""
    from kestrel_amd.driver import Driver
    cfg = DC.driver_config(block_frames=512, gate_frames=32)
    desc = DC.make_bank_desc(77)
    ev = DC.random_events(78, 400, 2000)
    p = str(tmp_path / "song.mid")
    DC.write_smf(p, [DC.track_bytes(ev[0::2]), DC.track_bytes(ev[1::2])], ppq=240)
    lb, ob = lib_bank(cfg, desc), ora_bank(cfg, desc)
    div, oev = OD.read_midi(p)
    ld = Driver.open(cfg, lb, p)
    od = OD.Driver(cfg, ob, div, oev)
    DC.run_pair(cfg, desc, ld, od, blocks=100_000, seed=9)
    st = ld.stats()
    assert st["notes"] == 400 and st["events"] == len(oev)


def test_driver_block_stays_put_while_the_next_one_begins():
    """begin_block starts the next block's event side on the driver's thread:
    what the caller holds of the current block (rows, flags, variants, spawn
    list) must not move while that runs, and pacing the calls differently
    must not change a single output or counter."""
    import time
    from kestrel_amd.driver import Division, Driver
    cfg = DC.driver_config(max_voices=60, max_param_variants=3)
    desc = DC.make_bank_desc(55)
    ev = DC.random_events(56, 1500, 1200, cc_rate=0.3)
    lb = lib_bank(cfg, desc)
    runs = []
    for pause in (0.0, 0.003):
        ld = Driver(cfg, lb, division=Division.ppq_of(480), events=ev)
        seen = []
        for b in range(120):
            live = cfg.max_voices - 10 if b % 2 else 0  # every other block admission chooses
            ups = ld.begin_block()
            g1, c1, f1 = ld.rows()
            cmds = ld.admit(live, 0)
            if pause:
                time.sleep(pause)  # the event thread finishes the next block meanwhile
            g2, c2, f2 = ld.rows()
            assert np.array_equal(g1, g2) and np.array_equal(c1, c2) and f1 == f2, b
            out = np.full(cfg.block_frames * 2, 0.25, np.float32)
            more = ld.end_block(out, 0, dict(active_voices=live, stolen=0, dropped=0, peak=0.25))
            seen.append((g1.tobytes(), c1.tobytes(), f1, [(i, t.tobytes()) for i, t in ups],
                         cmds.tobytes(), more))
        runs.append((seen, ld.stats()))
        ld.close()
    assert runs[0][0] == runs[1][0]
    assert runs[0][1] == runs[1][1]
    assert runs[0][1]["dropped"] > 0 and runs[0][1]["notes"] > 100


def test_driver_validates_its_inputs():
    from kestrel_amd.abi import BackendError
    from kestrel_amd.driver import Division, Driver
    cfg = DC.driver_config()
    lb = lib_bank(cfg, DC.make_bank_desc(1))
    ev = np.zeros(2, EVENT_DTYPE)
    ev["tick"] = [5, 3]
    with pytest.raises(BackendError, match="earlier than the one before"):
        Driver(cfg, lb, division=Division.ppq_of(96), events=ev)
    with pytest.raises(BackendError, match="gate_frames"):
        Driver(DC.driver_config(gate_frames=48), lb, division=Division.ppq_of(96), events=ev[:0])
    with pytest.raises(BackendError, match="No such file"):
        Driver.open(cfg, lb, os.path.join("/nonexistent", "x.mid"))
