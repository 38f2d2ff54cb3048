"""Pins the CPU oracle (oracle/kestrel_oracle.cpp) before anything is compared
with it: the reference's own golden vectors for this path (src/fixed.rs:173-225,
src/bank.rs:985-1034), closed forms, invariants, and brute force on tiny pools.
Each check is chosen so a plausible restatement mistake (a dropped term, a wrong
index or sign, an operation fused or reordered) fails it."""
import json
import math
import os

import numpy as np
import pytest

from kestrel_amd.config import Config

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_unit_tests.json")))
F32 = np.float32


# ---------------------------------------------------------------------------
# 32.32 phase and 8.24 bend (src/fixed.rs)
# ---------------------------------------------------------------------------

def test_round_trip_unity(oracle_mod):
    g = GOLD["fixed_round_trip_unity"]
    f = oracle_mod.fixed_from_f64(g["input"])
    assert f >> 32 == g["hi"] and f & 0xFFFFFFFF == g["lo"]
    assert oracle_mod.frac_f32(f) == g["frac"]


def test_accumulates_without_drift(oracle_mod):
    g = GOLD["fixed_accumulates_without_drift"]
    ratio = g["ratio_num"] / g["ratio_den"]
    step = oracle_mod.fixed_from_f64(ratio)
    p = (step * g["short_steps"]) % (1 << 64)
    assert abs(p / 2 ** 32 - ratio * g["short_steps"]) < g["short_tolerance"]
    n = g["long_steps"]
    err = (step * n) / 2 ** 32 - ratio * n
    assert abs(err) < g["long_tolerance_samples"]


def test_frac_matches_shader_expression(oracle_mod):
    g = GOLD["fixed_frac_matches_shader_expression"]
    for lo in g["lo"]:
        f = (g["hi"] << 32) | lo
        # lo as f32 (round to nearest even) times 2^-32, which is exact
        assert oracle_mod.frac_f32(f) == F32(lo) * F32(2.0 ** -32)


def test_f32_phase_would_have_failed():
    g = GOLD["fixed_f32_phase_would_have_failed"]
    step = F32(44100.0 / 48000.0)
    # f32 accumulation stalls: once p reaches 2^24 * ulp the sum stops moving;
    # summing in float32 by repeated doubling of chunks is exact enough to show it.
    p = np.cumsum(np.full(g["steps"], step, F32), dtype=F32)[-1]
    exact = g["steps"] * (44100.0 / 48000.0)
    assert abs(float(p) - exact) > g["min_error_samples"]


@pytest.mark.parametrize("seed", range(4))
def test_fixed_scale_brute_force(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    for _ in range(2000):
        phase = int(rng.integers(0, 2 ** 63)) * int(rng.integers(1, 3)) % (1 << 64)
        factor = int(rng.integers(0, 2 ** 32))
        want = (phase * factor) >> 24
        want = min(want, (1 << 64) - 1)
        assert oracle_mod.fixed_scale(phase, factor) == want


def test_fixed_scale_saturates_and_identity(oracle_mod):
    assert oracle_mod.fixed_scale((1 << 64) - 1, 1 << 25) == (1 << 64) - 1
    for p in (0, 1, 12345678901234, (1 << 64) - 1):
        assert oracle_mod.fixed_scale(p, 1 << 24) == p


def test_bend_factor(oracle_mod):
    assert oracle_mod.bend_factor(0.0) == 1 << 24
    assert oracle_mod.bend_factor(12.0) == 1 << 25
    assert oracle_mod.bend_factor(-12.0) == 1 << 23
    assert oracle_mod.bend_factor(float("nan")) == 1 << 24
    assert oracle_mod.bend_factor(1e9) == 0xFFFFFFFF
    # a semitone: round(2^(1/12) * 2^24)
    assert oracle_mod.bend_factor(1.0) == round(2 ** (1 / 12) * (1 << 24))


def test_from_f64_saturates(oracle_mod):
    assert oracle_mod.fixed_from_f64(-1.0) == 0
    assert oracle_mod.fixed_from_f64(float("nan")) == 0
    assert oracle_mod.fixed_from_f64(1e30) == (1 << 64) - 1
    assert oracle_mod.fixed_from_f64(0.5) == 1 << 31


# ---------------------------------------------------------------------------
# derived config arithmetic (src/config.rs:120-131, 532-552)
# ---------------------------------------------------------------------------

def test_derived_arithmetic_matches_host(oracle_mod):
    rng = np.random.default_rng(7)
    for _ in range(300):
        cfg = Config(max_voices=int(rng.integers(1, 5_000_000)),
                     max_steal_percent=int(rng.integers(1, 101)),
                     block_frames=int(rng.choice([64, 256, 512, 4096])),
                     steal_fade_frames=int(rng.integers(1, 63)),
                     steal_rule=int(rng.integers(0, 3)))
        assert oracle_mod.max_steal(cfg) == cfg.max_steal()
        assert oracle_mod.pool_slots(cfg) == cfg.pool_slots()
        assert oracle_mod.steal_span(cfg) == cfg.steal_span()
        for _ in range(5):
            live = int(rng.integers(0, cfg.max_voices + 1))
            queued = int(rng.integers(0, 2 * cfg.max_voices + 2))
            assert oracle_mod.admit_take(cfg, live, queued) == cfg.admit_take(live, queued)


def test_max_steal_closed_form(oracle_mod):
    cfg = Config(max_voices=1 << 20, max_steal_percent=10)
    assert oracle_mod.max_steal(cfg) == 104857
    assert oracle_mod.pool_slots(cfg) == 1153433
    assert oracle_mod.max_steal(Config(max_voices=5, max_steal_percent=1)) == 1


def test_spawn_pick_is_even_thinning(oracle_mod):
    # floor(i * want / take): strictly increasing, starts at 0, stays below want
    for want, take in ((10, 3), (1000, 999), (7, 7), (100000, 1)):
        picks = [oracle_mod.spawn_pick(i, want, take) for i in range(take)]
        assert picks[0] == 0 and picks[-1] < want
        assert all(b > a for a, b in zip(picks, picks[1:]))


# ---------------------------------------------------------------------------
# steal key and sort key layout
# ---------------------------------------------------------------------------

def test_steal_key(oracle_mod):
    assert oracle_mod.steal_key(12345, 0.7, False) == 12345
    # quietest: level quantised to 16 bits above the low 48 bits of the id
    assert oracle_mod.steal_key(5, 1.0, True) == (65535 << 48) | 5
    assert oracle_mod.steal_key(5, 0.0, True) == 5
    assert oracle_mod.steal_key(5, 2.0, True) == (65535 << 48) | 5
    assert oracle_mod.steal_key(5, -1.0, True) == 5
    assert oracle_mod.steal_key(5, float("nan"), True) == 5
    assert oracle_mod.steal_key((1 << 50) + 3, 0.5, True) == (32767 << 48) | 3
    # order: quieter sorts first regardless of id
    assert oracle_mod.steal_key(999, 0.1, True) < oracle_mod.steal_key(1, 0.2, True)


def test_sort_key_plan(oracle_mod):
    # 64 regions, 32768-frame samples: 7 region bits + 3 stage bits + 16 phase bits
    p = oracle_mod.sort_key_plan(64, 32768)
    assert p == {"bits": 26, "region_shift": 19, "stage_shift": 16, "phase_shift": 0,
                 "phase_mask": 0xFFFF, "dead_region": 64}
    # 2048 regions: 12 + 3 + 16 = 31 bits
    assert oracle_mod.sort_key_plan(2048, 32768)["bits"] == 31
    # 2^20 regions leave 9 phase bits of a 17-bit length
    p = oracle_mod.sort_key_plan(1 << 20, 65536)
    assert p["bits"] == 32 and p["phase_shift"] == 17 - 8 and p["phase_mask"] == 0xFF


# ---------------------------------------------------------------------------
# coefficient design (src/bank.rs:963-979, tests 1015-1027)
# ---------------------------------------------------------------------------

def dc_gain(b):
    b0, b1, a1, a2 = (float(x) for x in b)
    return (2 * b0 + b1) / (1 + a1 + a2)


def response(b, w):
    b0, b1, a1, a2 = (float(x) for x in b)
    z = np.exp(-1j * w)
    return abs((b0 + b1 * z + b0 * z * z) / (1 + a1 * z + a2 * z * z))


def test_lowpass_is_unity_at_dc(oracle_mod):
    g = GOLD["lowpass_is_unity_at_dc"]
    h = dc_gain(oracle_mod.biquad_lowpass(g["fc"], g["q_cb"], g["sr"]))
    assert abs(h - 1.0) < g["dc_tolerance"]
    h_res = dc_gain(oracle_mod.biquad_lowpass(g["fc"], g["resonant_q_cb"], g["sr"]))
    assert h_res < h


@pytest.mark.parametrize("fc,q", [(1000.0, 0.7071), (5000.0, 2.0), (200.0, 20.0)])
def test_lowpass_matches_rbj_response(oracle_mod, fc, q):
    # RBJ cookbook low-pass scaled by sqrt(Q_default / Q): |H(w0)| = Q * gain
    q_cb = 10.0 * (20.0 * math.log10(q) + 3.01)
    sr = 48000.0
    for design in (oracle_mod.biquad_lowpass, oracle_mod.biquad_lowpass_f64):
        b = design(fc, q_cb, sr)
        w0 = 2 * math.pi * fc / sr
        gain = math.sqrt(0.70710678 / q)
        tol = 2e-3 if design is oracle_mod.biquad_lowpass_f64 else 2e-2
        assert abs(response(b, w0) / (q * gain) - 1.0) < tol
        assert abs(response(b, 1e-6) / gain - 1.0) < tol
        assert response(b, 0.99 * math.pi) < 0.05 * gain


def test_make_params_envelope_closed_forms(oracle_mod):
    sr = 48000.0
    p = oracle_mod.make_params(0.005, 0.0, 0.100, 0.6, 0.150, 13500.0, 0.0, sr)
    assert p["attack_rate"] == F32(1.0) / (F32(0.005) * F32(sr))
    assert p["attack_end"] == 1.0
    # exponential segment: falls 100 dB over the quoted time
    frames = 0.100 * sr
    assert abs(float(p["decay_coef"]) ** frames - 1e-5) < 1e-8
    assert p["decay_target"] == F32(0.6) and p["sustain"] == F32(0.6)
    assert p["flags"] == 0          # 13500 cents: no filter
    p = oracle_mod.make_params(0.0, 0.0, 0.1, 0.6, 0.15, 6000.0, 0.0, sr)
    assert p["attack_rate"] == 1.0  # instant attack
    assert p["flags"] == 1
    p64 = oracle_mod.make_params_f64(0.0, 0.0, 0.1, 0.6, 0.15, 6000.0, 0.0, sr)
    for k in ("decay_coef", "release_coef", "b0", "b1", "a1", "a2"):
        assert abs(float(p64[k]) - float(p[k])) <= 4e-7 * max(1.0, abs(float(p64[k])))


# ---------------------------------------------------------------------------
# render: closed forms
# ---------------------------------------------------------------------------

def one_param(oracle_mod, **kw):
    p = np.zeros(1, oracle_mod.PARAMS_DTYPE)
    base = dict(attack_rate=1.0, attack_end=1.0, decay_coef=1.0, decay_target=1.0, sustain=1.0,
                release_coef=0.5, b0=1.0, b1=0.0, a1=0.0, a2=0.0, flags=0)
    base.update(kw)
    for k, v in base.items():
        p[k] = v
    return p


def one_cmd(oracle_mod, **kw):
    c = np.zeros(1, oracle_mod.SPAWN_DTYPE)
    base = dict(step_hi=1, smp_len=4096, loop_end=4096, gate_slot=60, ordinal=1,
                gain_l=1.0, gain_r=0.5)
    base.update(kw)
    for k, v in base.items():
        c[k] = v
    return c


def small_cfg(**kw):
    base = dict(block_frames=64, gate_frames=16, reduce_tile=8, max_voices=64,
                steal_fade_frames=8, interpolation=1)
    base.update(kw)
    return Config(**base)


def test_instant_envelope_renders_the_sample(oracle_mod):
    # src/testkit.rs:319-348: "a single note renders to exactly the sample repeated"
    rng = np.random.default_rng(1)
    smp = rng.integers(-32767, 32768, 4096).astype(np.int16)
    cfg = small_cfg()
    o = oracle_mod.Oracle(cfg, smp, one_param(oracle_mod))
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    o.spawn(one_cmd(oracle_mod, flags=1))
    outs = []
    for _ in range(3):
        outs.append(o.render().reshape(-1, 2))
    out = np.concatenate(outs)
    s = (smp[:192].astype(F32) * F32(1.0 / 32767.0))
    assert np.array_equal(out[:, 0], s)
    assert np.array_equal(out[:, 1], s * F32(0.5))


def test_half_step_interpolates_midpoints_of_a_ramp(oracle_mod):
    n = 2048
    ramp = np.round((np.arange(n) / (n - 1) * 2 - 1) * 32767).astype(np.int16)  # testkit ramp
    cfg = small_cfg()
    o = oracle_mod.Oracle(cfg, ramp, one_param(oracle_mod))
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    o.spawn(one_cmd(oracle_mod, step_hi=0, step_lo=1 << 31, smp_len=n, loop_end=0))
    out = o.render().reshape(-1, 2)[:, 0].astype(np.float64)
    f = np.arange(64)
    lo = ramp[f // 2].astype(np.float64) / 32767
    hi = ramp[f // 2 + 1].astype(np.float64) / 32767
    want = np.where(f % 2 == 0, lo, (lo + hi) / 2)
    assert np.max(np.abs(out - want)) < 2e-7


@pytest.mark.parametrize("step", [0.918750, 1.0, 1.7, 3.25])
def test_phase_accumulator_closed_form(oracle_mod, step):
    cfg = small_cfg()
    ls, le = 100, 1100
    stepfx = oracle_mod.fixed_from_f64(step)
    o = oracle_mod.Oracle(cfg, np.zeros(4096, np.int16), one_param(oracle_mod))
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    o.spawn(one_cmd(oracle_mod, step_hi=stepfx >> 32, step_lo=stepfx & 0xFFFFFFFF,
                    loop_start=ls, loop_end=le, flags=1))
    blocks = 40
    for _ in range(blocks):
        o.render()
    pool = o.pool_export()
    phase = (int(pool[1, 0]) << 32) | int(pool[0, 0])
    total = stepfx * blocks * cfg.block_frames
    hi, lo = total >> 32, total & 0xFFFFFFFF
    if hi >= le:
        hi = ls + (hi - ls) % (le - ls)
    assert phase == (hi << 32) | lo


def test_sample_end_kills_without_advancing(oracle_mod):
    cfg = small_cfg()
    o = oracle_mod.Oracle(cfg, np.ones(4096, np.int16), one_param(oracle_mod))
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    o.spawn(one_cmd(oracle_mod, smp_len=40))
    out = o.render().reshape(-1, 2)[:, 0]
    assert np.all(out[:40] != 0) and np.all(out[40:] == 0)
    assert o.stats()["active_voices"] == 0


def test_envelope_transitions_closed_form(oracle_mod):
    # attack 0.1 per frame: full level after ceil(1/0.1) = 10 frames, then an
    # exponential decay by 0.9 per frame reaches 0.5 after
    # ceil(ln 0.5 / ln 0.9) = 7 more frames.
    cfg = small_cfg(block_frames=32, gate_frames=32, steal_fade_frames=4)
    p = one_param(oracle_mod, attack_rate=0.1, attack_end=1.0, decay_coef=0.9,
                  decay_target=0.5, sustain=0.5)
    o = oracle_mod.Oracle(cfg, np.ones(4096, np.int16), p)
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    o.spawn(one_cmd(oracle_mod))
    out = o.render().reshape(-1, 2)[:, 0].astype(np.float64)
    lvl = out / (1.0 / 32767.0)
    n_att = math.ceil(1 / 0.1)
    n_dec = math.ceil(math.log(0.5) / math.log(0.9))
    # level during attack ~ 0.1 * (f+1); allow the f32 accumulation its ulps
    assert np.allclose(lvl[:n_att - 1], 0.1 * np.arange(1, n_att), rtol=1e-5)
    stage_at_end = o.pool_export()[9, 0]
    assert stage_at_end == 2  # sustain
    # the frame the level first equals sustain
    first_sus = n_att + int(np.argmax(np.isclose(lvl[n_att:], 0.5, rtol=1e-6)))
    assert abs(first_sus - (n_att - 1 + n_dec)) <= 1


def test_filter_steady_state_matches_analytic_gain(oracle_mod):
    sr, fc, q = 48000.0, 3000.0, 2.0
    q_cb = 10.0 * (20.0 * math.log10(q) + 3.01)
    b = oracle_mod.biquad_lowpass(fc, q_cb, sr)
    freq = 1500.0
    per = 32
    n = 4096
    t = np.arange(n)
    smp = np.round(0.5 * np.sin(2 * np.pi * t / per) * 32767).astype(np.int16)
    cfg = small_cfg(block_frames=256, gate_frames=32)
    p = one_param(oracle_mod, b0=b[0], b1=b[1], a1=b[2], a2=b[3], flags=1)
    o = oracle_mod.Oracle(cfg, smp, p)
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    o.spawn(one_cmd(oracle_mod, loop_start=0, loop_end=n, smp_len=n, flags=1))
    out = np.concatenate([o.render().reshape(-1, 2)[:, 0] for _ in range(8)])
    amp = np.max(np.abs(out[-512:]))
    w = 2 * math.pi * (sr / per) / sr
    assert abs(amp / 0.5 - response(b, w)) < 0.01 * response(b, w)
    assert freq > 0


# ---------------------------------------------------------------------------
# selection and compaction: invariants and brute force on tiny pools
# ---------------------------------------------------------------------------

def spawn_many(oracle_mod, n, ids, levels_attack=None, smp_len=4096):
    c = np.zeros(n, oracle_mod.SPAWN_DTYPE)
    c["step_hi"] = 1
    c["smp_len"] = smp_len
    c["gate_slot"] = np.arange(n) % 2048
    c["ordinal"] = 1
    c["note_id_lo"] = np.asarray(ids, np.uint64) & 0xFFFFFFFF
    c["note_id_hi"] = np.asarray(ids, np.uint64) >> 32
    c["gain_l"] = 1.0
    c["gain_r"] = 1.0
    if levels_attack is not None:
        c["params"] = levels_attack
    return c


@pytest.mark.parametrize("rule", [0, 2])
def test_victims_are_the_smallest_keys_independent_of_order(oracle_mod, rule):
    rng = np.random.default_rng(rule + 10)
    cap = 64
    cfg = small_cfg(max_voices=cap, steal_rule=rule, max_steal_percent=25,
                    block_frames=64, gate_frames=16)
    # Distinct attack rates give every voice a distinct level after one block.
    rates = (np.arange(cap) + 1) / 1000.0
    params = np.zeros(cap, oracle_mod.PARAMS_DTYPE)
    params["attack_rate"] = rates
    params["attack_end"] = 1e9
    params["sustain"] = 1.0
    params["decay_target"] = 1.0
    params["decay_coef"] = 1.0
    perm = rng.permutation(cap)
    ids = rng.permutation(10_000)[:cap] + 100
    results = []
    for order in (np.arange(cap), perm):
        o = oracle_mod.Oracle(cfg, np.ones(8192, np.int16), params)
        o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
        o.spawn(spawn_many(oracle_mod, cap, ids[order], levels_attack=order))
        o.render()
        need = 10
        new = spawn_many(oracle_mod, need, np.arange(need) + 50_000)
        o.spawn(new)
        pool = o.pool_export()
        stop = pool[23, :cap]
        victims = set((pool[20, :cap][stop != 0]).tolist())
        assert len(victims) == need
        assert o.stats()["stolen"] == need
        results.append(victims)
        note = pool[20, :cap]
        level = pool[10, :cap].view(np.float32)
        if rule == 0:
            assert victims == set(sorted(note.tolist())[:need])
        else:
            assert victims == set(note[np.argsort(level, kind="stable")[:need]].tolist())
        # stop frame = id % (F - fade) + 1
        for i in np.flatnonzero(stop):
            assert stop[i] == note[i] % (64 - 8) + 1
    assert results[0] == results[1]


def test_compaction_is_a_stable_partition(oracle_mod):
    cfg = small_cfg(max_voices=256)
    n = 200
    rng = np.random.default_rng(3)
    lens = rng.integers(5, 300, n)
    c = spawn_many(oracle_mod, n, np.arange(n))
    c["smp_len"] = lens
    o = oracle_mod.Oracle(cfg, np.ones(8192, np.int16), one_param(oracle_mod))
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    o.spawn(c)
    o.render()
    ids = o.pool_export()[20]
    assert np.array_equal(ids, np.flatnonzero(lens > 64))


def test_drop_new_and_even_thinning(oracle_mod):
    cfg = small_cfg(max_voices=16, steal_rule=1, admit_rule=1)
    o = oracle_mod.Oracle(cfg, np.ones(8192, np.int16), one_param(oracle_mod))
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    o.spawn(spawn_many(oracle_mod, 40, np.arange(40)))
    st = o.stats()
    assert st["active_voices"] == 16 and st["dropped"] == 24 and st["stolen"] == 0
    ids = o.pool_export()[20]
    assert list(ids) == [i * 40 // 16 for i in range(16)]


def test_f64_render_tracks_f32_render(oracle_mod):
    """The f64 render is the same algorithm in double precision. It parts from
    the f32 one only through f32 rounding inside each voice -- envelope
    multiplies accumulated over thousands of frames and the frame a stage
    boundary lands on -- so the two agree to about -60 dB of the signal."""
    from kestrel_amd import workload as W
    cfg, bank, notes = W.workload_c1()
    b32 = W.BlockBuilder(cfg, bank, notes)
    o32 = oracle_mod.Oracle(cfg, bank.pool, bank.params)
    o64 = oracle_mod.Oracle(cfg, bank.pool, bank.params, precision=1)
    errs = []
    for _ in range(60):
        blk = b32.next_block()
        for o in (o32, o64):
            o.set_gates(blk.gates)
            o.set_channels(blk.chans, *blk.flags)
            o.spawn(blk.cmds)
        errs.append(o32.render().astype(np.float64) - o64.render_f64())
        # phase accumulators are integer and must not drift at all
        assert np.array_equal(o32.pool_export()[:2], o64.pool_export()[:2])
    e = np.concatenate(errs)
    assert 0 < np.max(np.abs(e)) < 5e-3
    assert np.sqrt(np.mean(e ** 2)) < 1e-3


def test_sorted_mode_is_a_stable_sort_of_the_survivors(oracle_mod):
    """Sorted mode (the order the device's locality sort leaves): the survivors
    of a block, in their insertion order, stably sorted by the documented key
    (region, stage, phase bucket; src/gpu/mod.rs:76-105). Many equal keys, so a
    sort that is not stable, or a radix pass that drops a digit, fails."""
    rng = np.random.default_rng(21)
    n, regions = 3000, 5
    cfg = small_cfg(max_voices=4096, block_frames=64, gate_frames=16)
    c = spawn_many(oracle_mod, n, np.arange(n))
    c["smp_len"] = rng.integers(5, 300, n)
    c["region"] = rng.integers(0, regions, n)
    c["phase_hi"] = rng.integers(0, 3, n)
    c["step_hi"] = rng.integers(0, 3, n)
    pools = []
    for sort in (False, True):
        o = oracle_mod.Oracle(cfg, np.ones(8192, np.int16), one_param(oracle_mod))
        o.set_sorted(regions, 4096, sort)
        o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
        o.spawn(c)
        o.render()
        pools.append(o.pool_export())
    plain, srt = pools
    L = oracle_mod.sort_key_plan(regions, 4096)
    key = ((np.minimum(plain[16], L["dead_region"] - 1) << L["region_shift"]) |
           (np.minimum(plain[9], 7) << L["stage_shift"]) |
           ((plain[1] >> L["phase_shift"]) & L["phase_mask"]))
    order = np.argsort(key, kind="stable")
    assert len(set(key.tolist())) < len(key) // 4  # plenty of ties
    assert np.array_equal(srt, plain[:, order])
