"""The SoundFont 2 loader and the pool's rate converter (src/sf2.rs,
src/resample.rs) in the library against the oracle's restatement.

Pins of the oracle: the reference's own converter tests (src/resample.rs:
113-171) run on it, and a known-answer soundfont (one looped sine zone with
every envelope generator at its default) whose loaded region is fixed by the
format: instant envelope, unity sustain, continuous loop over the whole
sample, no filter. Then the library: random soundfonts (mixed rates, ROM and
out-of-range samples, global zones, reversed and disjoint ranges, preset
offsets, absolute-only generators at preset level), with and without pool
resampling and under a budget that forces the pool rate down -- samples,
regions, presets and the pool bit for bit; the converter sample for sample at
several ratios; the loader's errors.
"""
import math

import numpy as np
import pytest

import sf2_cases as S
from kestrel_amd.abi import REGION_DTYPE, SAMPLE_DTYPE
from kestrel_amd.config import Config

OS = pytest.importorskip("oracle.sf2")


def test_oracle_converter_reference_tests():
    """src/resample.rs:113-171 on the oracle's converter."""
    src = np.full(4000, 10000, np.int16)
    out = OS.resample_i16(src, 48000.0 / 44100.0)
    assert out.size == round(4000 * 48000.0 / 44100.0)
    assert np.all(np.abs(out[200:-200].astype(int) - 10000) < 60)
    n = 44100
    t = np.arange(n) / 44100.0
    src = (16000.0 * np.sin(2 * np.pi * 1000.0 * t)).astype(np.int16)
    out = OS.resample_i16(src, 48000.0 / 44100.0)
    w = out[1000:47000]
    crossings = int(np.sum((w[:-1] < 0) != (w[1:] < 0)))
    assert abs(crossings / 2.0 / (46000.0 / 48000.0) - 1000.0) < 5.0
    src = (26000.0 * np.sin(2 * np.pi * 440.0 * np.arange(2000) / 44100.0)).astype(np.int16)
    assert np.all(np.abs(OS.resample_i16(src, 2.0).astype(int)) < 32000)
    src = np.where(np.arange(500) % 2 == 0, 32767, -32768).astype(np.int16)
    out = OS.resample_i16(src, 2.0)
    assert out.dtype == np.int16 and out.size == 1000


def test_oracle_loader_known_answer(tmp_path):
    p = str(tmp_path / "simple.sf2")
    data = S.simple_sf2(p)
    pool, samples, regions, presets, pre, rate = OS.load(p, 48000)
    assert rate == 48000 and len(samples) == 1 and len(regions) == 1 and presets == [(0, 0, 0, 1)]
    s, r = samples[0], regions[0]
    assert (s["start"], s["len"], s["loop_start"], s["loop_end"]) == (0, data.size, 0, data.size)
    assert np.array_equal(pool[:data.size], data) and not pool[data.size:].any()
    assert r["attack"] == 0 and r["hold"] == 0 and r["decay"] == 0 and r["release"] == 0
    assert r["delay"] == 0 and r["sustain"] == 1.0 and r["loop_mode"] == 1
    assert (r["key_lo"], r["key_hi"], r["vel_lo"], r["vel_hi"]) == (0, 127, 0, 127)
    assert r["filter_fc_cents"] == 13500 and r["pan"] == 0 and r["scale_tuning"] == 100
    assert r["root_key_override"] == -1 and r["fixed_key"] == -1 and r["amp_veltrack"] == 100
    # timecents: 1200 is two seconds, -12000 and below is instant
    assert OS.timecents(1200) == 2.0 and OS.timecents(-12000) == 0 and OS.timecents(0) == 1.0
    assert abs(float(OS.cb_to_gain(200)) - 0.1) < 1e-7


@pytest.mark.parametrize("ratio", [48000 / 44100, 2.0, 0.5, 22050 / 48000, 48000 / 32000, 0.37, 1.0])
def test_converter_matches_the_oracle(ratio):
    from kestrel_amd.driver import resample_i16
    rng = np.random.default_rng(int(ratio * 1000))
    for x in ((rng.standard_normal(2500) * 12000).clip(-32768, 32767).astype(np.int16),
              np.where(np.arange(300) % 2 == 0, 32767, -32768).astype(np.int16),
              np.array([5], np.int16), np.zeros(0, np.int16)):
        a, b = resample_i16(x, ratio), OS.resample_i16(x, ratio)
        assert a.dtype == np.int16 and np.array_equal(a, b), (ratio, x.size)


def compare(lib_bank, ora):
    pool, samples, regions, presets, pre, rate = ora
    c = lib_bank.contents()
    assert c["pool_rate"] == rate
    assert np.array_equal(c["pool"], pool)
    assert c["samples"].size == len(samples) and c["regions"].size == len(regions)
    for got, want in zip(c["samples"], samples):
        for k, v in want.items():
            assert np.asarray(got[k]).tobytes() == np.asarray(v, SAMPLE_DTYPE[k]).tobytes(), k
    for i, (got, want) in enumerate(zip(c["regions"], regions)):
        for k, v in want.items():
            assert np.asarray(got[k]).tobytes() == np.asarray(v, REGION_DTYPE[k]).tobytes(), (i, k)
    assert [tuple(int(x) for x in p) for p in c["presets"]] == presets
    assert list(c["preset_regions"]) == pre


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("resample,budget", [(True, 2 << 30), (False, 2 << 30), (True, 3_000)])
def test_loader_matches_the_oracle(tmp_path, seed, resample, budget):
    from kestrel_amd.driver import Bank
    p = str(tmp_path / f"r{seed}.sf2")
    S.random_sf2(p, seed)
    cfg = Config(sample_rate=48000)
    try:
        ora = OS.load(p, 48000, resample, budget)
    except ValueError as e:
        with pytest.raises(Exception, match="no usable presets"):
            Bank.load_sf2(cfg, p, resample, budget)
        assert "no usable presets" in str(e)
        return
    lib = Bank.load_sf2(cfg, p, resample, budget)
    compare(lib, ora)
    if budget < 1 << 20:
        assert ora[5] < 48000  # the budget forced the pool rate down


def test_loader_feeds_the_bank(tmp_path):
    """A loaded soundfont builds the same params table as kst_bank_create given
    the loaded contents."""
    from kestrel_amd.driver import Bank
    p = str(tmp_path / "r.sf2")
    S.random_sf2(p, 11)
    cfg = Config(sample_rate=48000)
    a = Bank.load_sf2(cfg, p)
    c = a.contents()
    b = Bank(cfg, c["pool"], c["samples"], c["regions"], c["presets"], c["preset_regions"],
             pool_rate=c["pool_rate"])
    assert np.array_equal(a.params.view(np.uint32), b.params.view(np.uint32))
    assert a.n_regions == b.n_regions and a.max_sample_len == b.max_sample_len


def test_loader_errors(tmp_path):
    from kestrel_amd.abi import BackendError
    from kestrel_amd.driver import Bank
    cfg = Config(sample_rate=48000)
    s = [dict(name="s", data=S.sine(440, 48000, 100), rate=48000, root=60)]
    i = [dict(name="i", zones=[dict(gens=[], sample=0)])]
    pr = [dict(name="p", bank=0, program=0, zones=[dict(gens=[], instrument=0)])]
    cases = [("riff.sf2", dict(form=b"WAVE"), "expected sfbk"),
             ("nosmpl.sf2", dict(drop=(b"smpl",)), "no smpl chunk"),
             ("noshdr.sf2", dict(drop=(b"shdr",)), "pdta is missing shdr"),
             ("nopre.sf2", dict(), "no presets")]
    for name, kw, msg in cases:
        p = str(tmp_path / name)
        S.write_sf2(p, s, i, pr if name != "nopre.sf2" else [], **kw)
        with pytest.raises(BackendError, match=msg):
            Bank.load_sf2(cfg, p)
    p = str(tmp_path / "rom.sf2")
    s[0]["type"] = 0x8001
    S.write_sf2(p, s, i, pr)
    with pytest.raises(BackendError, match="no usable presets"):
        Bank.load_sf2(cfg, p)
    p = tmp_path / "text.sf2"
    p.write_bytes(b"hello world, not a soundfont")
    with pytest.raises(BackendError, match="not a RIFF file"):
        Bank.load_sf2(cfg, str(p))
