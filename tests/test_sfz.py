"""The SFZ loader and its WAV reader (src/sfz.rs, src/wav.rs:166-365) in the
library against the oracle's restatement.

Pins of the oracle: the reference's own SFZ tests (note names, values with
spaces, a value running to the end; src/sfz.rs:678-698) and closed forms of
the WAV decode (every sample format reproduces its integer codes within one
16-bit step; the smpl chunk's root key, fine tune from the 0.32 pitch fraction
and first loop). Then the library: random instrument directories -- every
sample format, stereo files, inheritance through the five headers, an
#include, aliases, note names, comments, Windows paths and case, disabled and
missing samples -- with and without pool resampling: samples, regions, the
preset and the pool bit for bit; and the loader's errors.
"""
import os

import numpy as np
import pytest

import sfz_cases as Z
from kestrel_amd.abi import REGION_DTYPE, SAMPLE_DTYPE
from kestrel_amd.config import Config

OZ = pytest.importorskip("oracle.sfz")


def test_oracle_reference_sfz_tests():
    assert OZ.note_name("c4") == 60 and OZ.note_name("a4") == 69
    assert OZ.note_name("c#4") == 61 and OZ.note_name("c-1") == 0
    assert OZ.note_name("bb3") == 58 and OZ.note_name("x4") is None
    s = "My Sample Name.wav lokey=30"
    assert s[:OZ.value_end(s)].strip() == "My Sample Name.wav"
    assert OZ.value_end("60") == 2
    assert OZ.value_end("a.wav<region>") == 5


@pytest.mark.parametrize("fmt", Z.FORMATS)
def test_oracle_wav_decode(tmp_path, fmt):
    rng = np.random.default_rng(4)
    codes = rng.integers(-30000, 30000, 400)
    x = codes / 32767.0
    p = str(tmp_path / f"{fmt}.wav")
    Z.write_wav(p, np.stack([x, -x], 1), 44100, fmt, smpl=(62, 1 << 31, [(10, 300)]))
    w = OZ.read_wav(p)
    assert w["rate"] == 44100 and w["channels"] == 2
    # root key and fine tune from the fields the reference reads (offsets 20
    # and 24 of smpl), the first loop from the first loop record
    assert w["root"] == 62 and w["loop"] == (10, 300) and w["fine"] == np.float32(50.0)
    left, right = OZ.channel_i16(w, 0), OZ.channel_i16(w, 1)
    tol = 400 if fmt == "pcm8" else 1  # 8 bits: a 256-code step and the 127 vs 128 scale
    assert np.all(np.abs(left.astype(int) - codes) <= tol)
    assert np.all(np.abs(right.astype(int) + codes) <= tol + 1)


def compare(lib_bank, ora):
    pool, samples, regions, presets, pre, rate = ora
    c = lib_bank.contents()
    assert c["pool_rate"] == rate
    assert c["samples"].size == len(samples) and c["regions"].size == len(regions)
    assert np.array_equal(c["pool"], pool)
    for i, (got, want) in enumerate(zip(c["samples"], samples)):
        for k, v in want.items():
            assert np.asarray(got[k]).tobytes() == np.asarray(v, SAMPLE_DTYPE[k]).tobytes(), (i, k)
    for i, (got, want) in enumerate(zip(c["regions"], regions)):
        for k, v in want.items():
            assert np.asarray(got[k]).tobytes() == np.asarray(v, REGION_DTYPE[k]).tobytes(), (i, k)
    assert [tuple(int(x) for x in p) for p in c["presets"]] == presets
    assert list(c["preset_regions"]) == pre


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("resample", [True, False])
def test_loader_matches_the_oracle(tmp_path, seed, resample):
    from kestrel_amd.driver import Bank
    path = Z.random_sfz(str(tmp_path / f"inst{seed}"), seed)
    cfg = Config(sample_rate=48000)
    ora = OZ.load(path, 48000, resample)
    lib = Bank.load_sfz(cfg, path, resample)
    compare(lib, ora)
    assert len(ora[2]) >= 10  # regions that loaded
    assert lib.params.size > 0


def test_loader_errors(tmp_path):
    from kestrel_amd.abi import BackendError
    from kestrel_amd.driver import Bank
    cfg = Config(sample_rate=48000)
    p = tmp_path / "empty.sfz"
    p.write_text("<group> volume=-3\n")
    with pytest.raises(BackendError, match="no <region> sections"):
        Bank.load_sfz(cfg, str(p))
    p = tmp_path / "missing.sfz"
    p.write_text("<region> sample=nowhere.wav\n<region> key=60\n")
    with pytest.raises(BackendError, match="every region was unusable"):
        Bank.load_sfz(cfg, str(p))
    w = tmp_path / "adpcm.wav"
    Z.write_wav(str(w), np.zeros(100), 48000, "pcm16")
    raw = bytearray(w.read_bytes())
    raw[20:22] = (2).to_bytes(2, "little")  # format tag 2: not a format the reader takes
    w.write_bytes(bytes(raw))
    p = tmp_path / "adpcm.sfz"
    p.write_text("<region> sample=adpcm.wav\n")
    with pytest.raises(BackendError, match="every region was unusable"):
        Bank.load_sfz(cfg, str(p))
    with pytest.raises(BackendError, match="reading"):
        Bank.load_sfz(cfg, str(tmp_path / "absent.sfz"))
