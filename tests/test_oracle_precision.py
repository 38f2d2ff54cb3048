"""Pins the oracle's render in both of its precisions: the f32 restatement of
src/cpu.rs:392-625 and the same algorithm carried out in f64 (`precision=1`),
which is the reference every precision-form error figure is measured against.

Each test takes a special case whose answer the mathematics fixes
independently of the code: dyadic envelope parameters whose levels are exact
in both precisions, a sample that is a quadratic polynomial (which Catmull-Rom
reproduces exactly), an integer ramp under linear interpolation at arbitrary
phase fractions, the steady-state response of a biquad to a periodic input
(from the DFT of one period and the transfer function), and a steal fade with a
power-of-two length. The f64 tolerances are tight enough that a 1/32768 scale,
a frac taken in f32, a wrong filter coefficient or a params table picked from
the wrong precision fails.
"""
import math

import numpy as np
import pytest

from kestrel_amd.config import Config

F32 = np.float32
PRECISIONS = [0, 1]


def params(oracle_mod, n=1, **kw):
    p = np.zeros(n, oracle_mod.PARAMS_DTYPE)
    base = dict(attack_rate=1.0, attack_end=1.0, decay_coef=1.0, decay_target=1.0, sustain=1.0,
                release_coef=0.5, b0=1.0, b1=0.0, a1=0.0, a2=0.0, flags=0)
    base.update(kw)
    for k, v in base.items():
        p[k] = v
    return p


def cmd(oracle_mod, n=1, **kw):
    c = np.zeros(n, oracle_mod.SPAWN_DTYPE)
    base = dict(step_hi=1, smp_len=4096, loop_end=4096, gate_slot=60, ordinal=1,
                gain_l=1.0, gain_r=0.5)
    base.update(kw)
    for k, v in base.items():
        c[k] = v
    return c


def cfg_of(**kw):
    base = dict(block_frames=64, gate_frames=16, reduce_tile=8, max_voices=64,
                steal_fade_frames=8, interpolation=1)
    base.update(kw)
    return Config(**base)


def make(oracle_mod, cfg, bank, table, prec):
    o = oracle_mod.Oracle(cfg, bank, table, precision=prec)
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    return o


def block(o, prec):
    """One block as float64 (interleaved L, R): the f64 mix itself at
    precision 1, the f32 output block at precision 0."""
    return (o.render_f64() if prec else o.render().astype(np.float64)).reshape(-1, 2)


def norm(v, prec):
    """int16 -> float the way src/cpu.rs:190-194 scales it, at each precision."""
    v = np.asarray(v)
    if prec:
        return np.maximum(v.astype(np.float64) / 32767.0, -1.0)
    return np.maximum(v.astype(F32) * F32(1.0 / 32767.0), F32(-1.0)).astype(np.float64)


# ---------------------------------------------------------------------------
# fetch scale: an instant-envelope looped sample renders exactly the sample
# (src/testkit.rs:319-348)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("prec", PRECISIONS)
def test_instant_envelope_renders_the_sample_exactly(oracle_mod, prec):
    rng = np.random.default_rng(11)
    smp = rng.integers(-32768, 32768, 4096).astype(np.int16)
    smp[5] = -32768  # the clamp at -1
    cfg = cfg_of()
    o = make(oracle_mod, cfg, smp, params(oracle_mod), prec)
    o.spawn(cmd(oracle_mod, flags=1))
    out = np.concatenate([block(o, prec) for _ in range(3)])
    if prec:
        want = np.maximum(smp[:192].astype(np.float64) / 32767.0, -1.0)
    else:
        want = np.maximum(smp[:192].astype(F32) * F32(1.0 / 32767.0), F32(-1.0))
    assert np.array_equal(out[:, 0], want)
    assert np.array_equal(out[:, 1], want * 0.5)
    assert out[5, 0] == -1.0


# ---------------------------------------------------------------------------
# interpolation at arbitrary phases: closed forms
# ---------------------------------------------------------------------------

def phase_walk(phase0, step, n):
    """(index, frac) of n frames of a 32.32 accumulator, exact."""
    p = (phase0 + step * np.arange(n, dtype=object))
    idx = np.array([int(x) >> 32 for x in p], np.int64)
    frac = np.array([(int(x) & 0xFFFFFFFF) / 2.0 ** 32 for x in p], np.float64)
    return idx, frac


@pytest.mark.parametrize("prec", PRECISIONS)
def test_linear_interpolation_of_an_integer_ramp(oracle_mod, prec):
    # s[k] = k - 2048: the interpolant at index + frac is exactly
    # (index + frac - 2048) / 32767.
    n = 4096
    smp = (np.arange(n) - 2048).astype(np.int16)
    cfg = cfg_of(interpolation=1, block_frames=128, gate_frames=16)
    step = oracle_mod.fixed_from_f64(1.3183098861837907)
    phase0 = (700 << 32) | 0x9E3779B9
    o = make(oracle_mod, cfg, smp, params(oracle_mod), prec)
    o.spawn(cmd(oracle_mod, phase_hi=phase0 >> 32, phase_lo=phase0 & 0xFFFFFFFF,
                step_hi=step >> 32, step_lo=step & 0xFFFFFFFF, smp_len=n, loop_end=0))
    out = np.concatenate([block(o, prec)[:, 0] for _ in range(2)])
    idx, frac = phase_walk(phase0, step, len(out))
    want = (idx - 2048 + frac) / 32767.0
    err = np.max(np.abs(out - want))
    # f64: a few ulps of values below 0.1 (an f32 frac alone would be ~1e-12 off);
    # f32: the reference's own rounding, a few ulps of f32.
    assert err < (1e-15 if prec else 2e-8), err


@pytest.mark.parametrize("prec", PRECISIONS)
def test_cubic_interpolation_reproduces_a_quadratic(oracle_mod, prec):
    # Catmull-Rom takes central differences as tangents, which are exact for a
    # quadratic, so the Hermite cubic equals the quadratic everywhere between
    # the samples: s(x) = (x - 150)^2 - 11000.
    n = 300
    k = np.arange(n + 8)
    smp = ((k - 150) ** 2 - 11000).astype(np.int16)
    smp[n:] = 0
    cfg = cfg_of(interpolation=2, block_frames=64, gate_frames=16)
    step = oracle_mod.fixed_from_f64(0.7071067811865476)
    phase0 = (20 << 32) | 0x12345678
    o = make(oracle_mod, cfg, smp, params(oracle_mod), prec)
    o.spawn(cmd(oracle_mod, phase_hi=phase0 >> 32, phase_lo=phase0 & 0xFFFFFFFF,
                step_hi=step >> 32, step_lo=step & 0xFFFFFFFF, smp_len=n, loop_end=0))
    out = np.concatenate([block(o, prec)[:, 0] for _ in range(3)])
    idx, frac = phase_walk(phase0, step, len(out))
    x = idx + frac
    want = ((x - 150.0) ** 2 - 11000.0) / 32767.0
    err = np.max(np.abs(out - want))
    assert err < (1e-13 if prec else 2e-6), err


# ---------------------------------------------------------------------------
# envelope: dyadic parameters, every level exact in both precisions
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("curve", ["exponential", "linear"])
def test_envelope_levels_and_transition_frames_exact(oracle_mod, prec, curve):
    exp = curve == "exponential"
    cfg = cfg_of(block_frames=64, gate_frames=16, decay_curve=0 if exp else 1,
                 release_curve=0 if exp else 1)
    if exp:
        p = params(oracle_mod, attack_rate=0.125, attack_end=1.0, decay_coef=0.5,
                   decay_target=0.0625, sustain=0.0625, release_coef=0.5)
    else:
        p = params(oracle_mod, attack_rate=0.125, attack_end=1.0, decay_coef=0.125,
                   decay_target=0.5, sustain=0.5, release_coef=0.125)
    smp = np.full(4096, 16384, np.int16)
    o = oracle_mod.Oracle(cfg, smp, p, precision=prec)
    gates = np.zeros((cfg.tiles(), 2048), np.uint32)
    gates[2:, 60] = 1  # released from gate tile 2 (frame 32) on
    o.set_gates(gates.ravel())
    o.spawn(cmd(oracle_mod, flags=1))
    out = block(o, prec)
    # The levels the reference's stage machine (src/cpu.rs:508-549) must produce.
    lv = []
    level, stage = 0.0, "attack"
    for f in range(64):
        if f == 32 and stage in ("attack", "decay", "sustain"):
            stage, level = "release", min(level, 1.0)
        if stage == "attack":
            level += 0.125
            if level >= 1.0:
                stage, level = "decay", 1.0
        elif stage == "decay":
            level = level * 0.5 if exp else level - 0.125
            if level <= (0.0625 if exp else 0.5):
                stage, level = "sustain", (0.0625 if exp else 0.5)
        elif stage == "release":
            level = level * 0.5 if exp else level - 0.125
            if level <= 1e-5:
                stage, level = "dead", 0.0
        lv.append(level if stage != "dead" else 0.0)
        if stage == "dead":
            break
    lv = np.array(lv + [0.0] * (64 - len(lv)))
    # Closed forms the loop above must agree with: full level after 8 attack
    # frames; the sustain level from frame 11 (exp: 1 -> 1/16 in four halvings)
    # or frame 11 (linear: 1 -> 0.5 in four steps of 1/8); death in release when
    # the level first reaches the 1e-5 floor: exp 0.0625 * 2^-13 = 7.6e-6 at
    # the 13th release frame, linear 0.5 - 4/8 = 0 at the 4th.
    assert lv[7] == 1.0 and lv[11] == (0.0625 if exp else 0.5)
    death = 32 + (12 if exp else 3)
    assert lv[death] == 0.0 and lv[death - 1] > 0.0
    dt = np.float64 if prec else F32
    want = dt(norm(16384, prec)) * lv.astype(dt)  # x = s * level, rounded once
    assert np.array_equal(out[:, 0], want)
    assert np.array_equal(out[:, 1], want * dt(0.5))
    assert o.stats()["active_voices"] == 0


# ---------------------------------------------------------------------------
# biquad: steady state of a periodic input, from the transfer function
# ---------------------------------------------------------------------------

def steady_state(x_period, b):
    """Steady-state output of H(z) = (b0 + b1 z^-1 + b0 z^-2) / (1 + a1 z^-1 +
    a2 z^-2) to the periodic input x_period, one period."""
    b0, b1, a1, a2 = (float(v) for v in b)
    n = len(x_period)
    z = np.exp(-2j * np.pi * np.arange(n) / n)
    H = (b0 + b1 * z + b0 * z * z) / (1 + a1 * z + a2 * z * z)
    return np.real(np.fft.ifft(np.fft.fft(x_period) * H))


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("fc,q,period", [(3000.0, 2.0, 32), (500.0, 20.0, 96)])
def test_filter_steady_state_matches_the_transfer_function(oracle_mod, prec, fc, q, period):
    # A period of 96 frames puts the fundamental (500 Hz) on the resonance.
    sr = 48000.0
    q_cb = 10.0 * (20.0 * math.log10(q) + 3.01)
    n = period * (4096 // period)
    rng = np.random.default_rng(5)
    one = np.round(rng.uniform(-0.5, 0.5, period) * 32767).astype(np.int16)
    smp = np.tile(one, n // period)
    b32 = oracle_mod.biquad_lowpass(fc, q_cb, sr)
    cfg = cfg_of(block_frames=512, gate_frames=32, max_voices=4)
    table = params(oracle_mod, b0=b32[0], b1=b32[1], a1=b32[2], a2=b32[3], flags=1)
    o = oracle_mod.Oracle(cfg, smp, table, precision=prec)
    o.set_gates(np.zeros(cfg.tiles() * 2048, np.uint32))
    if prec:
        b = oracle_mod.biquad_lowpass_f64(fc, q_cb, sr)
        t64 = np.zeros(1, oracle_mod.PARAMS64_DTYPE)
        for k in table.dtype.names:
            t64[k] = table[k]
        t64["b0"], t64["b1"], t64["a1"], t64["a2"] = b
        o.set_params64(t64)
    else:
        b = b32
    o.spawn(cmd(oracle_mod, loop_start=0, loop_end=n, smp_len=n, flags=1))
    # the transient decays as |pole|^f: ~1e-16 of it is left after 48 blocks at Q=20
    blocks = 48
    for _ in range(blocks - 1):
        block(o, prec)
    out = block(o, prec)[:, 0]
    x = norm(one, prec)
    y = steady_state(x, b)
    f = (blocks - 1) * 512 + np.arange(512)
    amp = np.max(np.abs(y))
    err = np.max(np.abs(out - y[f % period])) / amp
    # f32: the reference's own rounding, amplified by the resonance (1e-4 at Q 20)
    assert err < (1e-12 if prec else 5e-4), err
    assert amp > 0.05


# ---------------------------------------------------------------------------
# which params table the f64 render reads (src/cpu.rs:228-238 at f64)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("variant", [0, 3, 5])
def test_params_table_pick(oracle_mod, prec, variant):
    # Bank table (f32): no filter. Bank table in f64: the FIR y[n] = (x[n] +
    # x[n-2]) / 2 (b0 = 1/2, b1 = a1 = a2 = 0). Variant 3 (f32, uploaded): no
    # filter. Variant 5 was never uploaded: it reads the bank table. So at f64
    # voices born in variants 0 and 5 are filtered and the variant-3 voice is
    # not; at f32 none is.
    cfg = cfg_of(block_frames=64, gate_frames=16, max_voices=8)
    rng = np.random.default_rng(9)
    smp = rng.integers(-20000, 20000, 4096).astype(np.int16)
    table = params(oracle_mod)
    o = make(oracle_mod, cfg, smp, table, prec)
    t64 = np.zeros(1, oracle_mod.PARAMS64_DTYPE)
    for k in table.dtype.names:
        t64[k] = table[k]
    t64["b0"], t64["flags"] = 0.5, 1
    o.set_params64(t64)
    o.set_params_variant(3, params(oracle_mod))
    o.spawn(cmd(oracle_mod, variant=variant, flags=1))
    out = block(o, prec)[:, 0]
    x = norm(smp[:64], prec)
    fir = 0.5 * x + 0.5 * np.r_[0.0, 0.0, x[:-2]]
    assert np.array_equal(out, fir if prec and variant != 3 else x)


# ---------------------------------------------------------------------------
# steal fade: a power-of-two fade, exact in both precisions
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("prec", PRECISIONS)
def test_steal_fade_closed_form(oracle_mod, prec):
    # Two voices fill a two-voice pool; a third note steals the oldest (id 0),
    # which then fades as y * (1 - d / 8) from its stop frame and is silent and
    # dead from d = 8 on (src/cpu.rs:577-586).
    F, fade = 64, 8
    cfg = cfg_of(block_frames=F, gate_frames=16, max_voices=2, steal_rule=0,
                 max_steal_percent=50, steal_fade_frames=fade)
    smp = np.full(4096, 8192, np.int16)
    o = make(oracle_mod, cfg, smp, params(oracle_mod), prec)
    c = cmd(oracle_mod, n=2, flags=1, gain_r=0.0)
    c["note_id_lo"] = [20, 21]
    c["gate_slot"] = [60, 61]
    c["gain_l"] = [1.0, 0.0]  # voice 21 sounds in neither channel
    o.spawn(c)
    block(o, prec)
    o.spawn(cmd(oracle_mod, flags=1, gain_l=0.0, gain_r=1.0, note_id_lo=30, gate_slot=62))
    assert o.stats()["stolen"] == 1
    stop = 20 % (F - fade) + 1
    assert o.pool_export()[23].tolist() == [stop, 0, 0]
    out = block(o, prec)
    dt = np.float64 if prec else F32
    s = dt(norm(8192, prec))
    d = np.arange(F) + 1 - stop
    g = np.where(d < 0, 1.0, np.where(d < fade, 1.0 - d / fade, 0.0)).astype(dt)
    assert np.array_equal(out[:, 0], s * g)          # the victim, faded once
    assert np.array_equal(out[:, 1], np.full(F, s))  # the new voice alone
    assert o.stats()["active_voices"] == 2
    ids = o.pool_export()[20].tolist()
    assert ids == [21, 30]
