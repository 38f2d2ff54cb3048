"""The post-render step on the host (no GPU needed): the library's limiters,
clamp, WAV writer and null comparison against a pure-Python restatement of
src/limiter.rs and `kestrel null` (bit for bit, short signals), and the
reference's own limiter tests (src/limiter.rs:473-705) run on the library.
"""
import math
import os

import numpy as np
import pytest

from kestrel_amd import post
from oracle import post as ref

SR = 48000


@pytest.fixture(scope="module", autouse=True)
def lib():
    from kestrel_amd import abi, build
    if not os.path.exists(abi.LIB_PATH):
        build.build()
    return abi.load()


def signal(seed, frames, loud=6.0):
    rng = np.random.default_rng(seed)
    t = np.arange(frames)
    base = np.sin(t * 0.03) * loud + rng.normal(0.0, 0.3, frames)
    spikes = (rng.random(frames) < 0.002) * rng.uniform(-80, 80, frames)
    l = base + spikes
    r = np.cos(t * 0.021) * loud * 0.7 + rng.normal(0.0, 0.3, frames)
    return np.stack([l, r], axis=1).astype(np.float32).ravel()


@pytest.mark.parametrize("mode,kw", [
    (post.OMNI, {}),
    (post.BRICKWALL, dict(true_peak=1)),
    (post.BRICKWALL, dict(true_peak=0)),
    (post.BRICKWALL, dict(true_peak=1, sustain_ms=400.0, ceiling_db=-6.0)),
    (post.BRICKWALL, dict(true_peak=1, lookahead_ms=0.1, release_ms=5.0)),
])
def test_limiters_match_the_restatement_bit_for_bit(mode, kw):
    x = signal(1, 6000)
    lim = post.Limiter(mode=mode, **kw)
    cfg = lim.cfg
    if mode == post.OMNI:
        r = ref.Omni(SR)
    else:
        r = ref.Brickwall(SR, 10 ** (cfg.ceiling_db / 20.0), cfg.lookahead_ms, cfg.release_ms,
                          cfg.sustain_ms, cfg.true_peak)
        assert lim.latency == r.delay_frames
    got, want = x.copy(), x.copy()
    for b0 in range(0, got.size, 1024):  # state carries across blocks
        lim.process(got[b0:b0 + 1024])
        r.process(want[b0:b0 + 1024])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_clamp_and_null_match_the_restatement():
    x = signal(2, 3000, loud=1.5)
    a, b = x.copy(), x.copy()
    assert post.clamp_block(a) == ref.clamp_block(b) > 0
    assert np.array_equal(a, b) and np.abs(a).max() <= 1.0
    y = x + np.float32(1e-3) * np.sin(np.arange(x.size)).astype(np.float32)
    got, want = post.null(x, y), ref.null(x, y)
    for k in ("compared", "worst", "identical"):
        assert got[k] == want[k]
    for k in ("peak_ref", "peak_diff", "rms_diff"):
        assert got[k] == pytest.approx(want[k], rel=1e-12)
    same = post.null(x, x.copy())
    assert same["identical"] and same["peak_diff_db"] == -math.inf


def test_brickwall_latency_is_lookahead_plus_detector_delay():
    assert post.Limiter(mode=post.BRICKWALL).latency == 96 + 3   # 2 ms at 48 kHz + 3
    assert post.Limiter(mode=post.BRICKWALL, true_peak=0).latency == 96
    assert post.Limiter(mode=post.OMNI).latency == 0


# ---- the reference's own limiter tests (src/limiter.rs:473-705) ----

def test_quiet_signal_passes_through_scaled():
    buf = np.full(SR * 6, 0.1, np.float32)
    post.Limiter(mode=post.OMNI).process(buf)
    assert np.all(np.abs(buf[-1000:] - 0.125) < 0.005)


def test_loud_signal_is_pulled_down():
    buf = np.full(SR * 2, 8.0, np.float32)
    post.Limiter(mode=post.OMNI).process(buf)
    assert np.all(np.abs(buf[-100:]) <= 1.0)


def test_brickwall_never_exceeds_the_ceiling():
    for ceiling in (1.0, 0.5):
        t = np.arange(SR * 2) // 2
        buf = (np.sin(t * 0.05) * 40.0 + np.where(t % 977 == 0, 300.0, 0.0)).astype(np.float32)
        post.Limiter(mode=post.BRICKWALL, ceiling_db=20 * math.log10(ceiling),
                     sustain_ms=400.0).process(buf)
        assert float(np.abs(buf).max()) <= ceiling + 1e-4
        assert post.clamp_block(buf) == 0


def brickwall_after_transient(sustain_ms):
    quiet = 0.5
    t = np.arange(SR * 2) // 2
    buf = np.where((t >= 4800) & (t < 4848), 60.0, quiet).astype(np.float32)
    post.Limiter(mode=post.BRICKWALL, sustain_ms=sustain_ms).process(buf)
    level = lambda sec: float(np.abs(buf[int(sec * SR) * 2:int(sec * SR) * 2 + 4800]).max())
    return quiet, level


def test_brickwall_recovers_quickly_after_a_transient():
    quiet, level = brickwall_after_transient(400.0)
    assert abs(level(0.02) - quiet) < 0.02
    assert abs(level(0.35) - quiet) < 0.02


def test_true_peak_detection_catches_intersample_overshoot():
    t = (np.arange(SR * 2) // 2).astype(np.float64)
    src = (np.sin(t * math.pi / 2 + math.pi / 4) * math.sqrt(2)).astype(np.float32)
    tp, sp = src.copy(), src.copy()
    post.Limiter(mode=post.BRICKWALL, sustain_ms=400.0, true_peak=1).process(tp)
    post.Limiter(mode=post.BRICKWALL, sustain_ms=400.0, true_peak=0).process(sp)
    assert np.abs(tp).max() < np.abs(sp).max() - 0.05


def test_limiters_are_deterministic():
    for mode in (post.OMNI, post.BRICKWALL):
        runs = []
        for _ in range(2):
            b = (np.sin(np.arange(8192, dtype=np.float32) * 0.017) * 9.0).astype(np.float32)
            post.Limiter(mode=mode, sustain_ms=400.0).process(b)
            runs.append(b)
        assert np.array_equal(runs[0], runs[1])


def test_the_sustained_stage_holds_the_gain_steady_on_a_loud_passage():
    def wobble(sustain_ms):
        t = (np.arange(SR * 2 * 2) // 2) / SR
        src = ((np.sin(t * 1000 * math.tau) + np.sin(t * 1050 * math.tau)) * 20.0).astype(
            np.float32)
        buf = src.copy()
        lim = post.Limiter(mode=post.BRICKWALL, release_ms=5.0, sustain_ms=sustain_ms)
        d = lim.latency * 2
        lim.process(buf)
        k = np.arange(src.size // 2, src.size - d)
        k = k[np.abs(src[k]) > 4.0]
        assert k.size > 10000
        g = buf[k + d].astype(np.float64) / src[k].astype(np.float64)
        return g.std() / g.mean()
    assert wobble(400.0) < wobble(0.0) * 0.85


def test_the_sustained_stage_does_not_reintroduce_pumping():
    quiet, level = brickwall_after_transient(400.0)
    assert abs(level(0.35) - quiet) < 0.02


# ---- WAV ----

def test_wav_writer_layout_and_round_trip(tmp_path):
    x = signal(3, 1000, loud=0.8)
    p = tmp_path / "a.wav"
    w = post.WavWriter(str(p), SR, 2, float32=True)
    w.write_block(x[:800])
    w.write_block(x[800:])
    w.finish()
    raw = open(p, "rb").read()
    assert raw[:4] == b"RIFF" and int.from_bytes(raw[4:8], "little") == 36 + 4 * x.size
    assert raw[8:16] == b"WAVEfmt " and int.from_bytes(raw[16:20], "little") == 16
    assert [int.from_bytes(raw[a:b], "little") for a, b in ((20, 22), (22, 24), (24, 28),
                                                              (28, 32), (32, 34), (34, 36))] == \
        [3, 2, SR, SR * 8, 8, 32]
    assert raw[36:40] == b"data" and int.from_bytes(raw[40:44], "little") == 4 * x.size
    y, sr, ch = post.read_wav(str(p))
    assert (sr, ch) == (SR, 2) and np.array_equal(x, y)
    q = tmp_path / "b.wav"
    w = post.WavWriter(str(q), SR, 2, float32=False)
    w.write_block(np.array([0.0, 1.5, -1.0, 0.5, np.nan, -0.00002], np.float32))
    w.finish()
    pcm = np.frombuffer(open(q, "rb").read()[44:], "<i2")
    assert pcm.tolist() == [0, 32767, -32767, 16384, 0, -1]
