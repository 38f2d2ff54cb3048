"""The caller's side of `render()`: the reference's limiters, the hard clamp,
the WAV writer and the `kestrel null` comparison, over the C ABI
(kst_limiter_*, kst_clamp_block, kst_wav_*, kst_null).

Mirrors src/limiter.rs (`Limiter`, `Brickwall`, `clamp_block`), src/wav.rs
(`WavWriter`) and src/bin/kestrel.rs:597-646 (`null`).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import abi

OFF, OMNI, BRICKWALL = 0, 1, 2


class LimiterConfig(C.Structure):
    _fields_ = [("sample_rate", C.c_uint32), ("mode", C.c_uint32), ("ceiling_db", C.c_double),
                ("lookahead_ms", C.c_double), ("release_ms", C.c_double),
                ("sustain_ms", C.c_double), ("true_peak", C.c_uint32), ("pad", C.c_uint32)]

    @classmethod
    def default(cls, **kw) -> "LimiterConfig":
        c = cls()
        abi.load().kst_default_limiter_config(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class NullResult(C.Structure):
    _fields_ = [("compared", C.c_uint64), ("worst", C.c_uint64), ("peak_ref", C.c_double),
                ("peak_diff", C.c_double), ("rms_diff", C.c_double), ("identical", C.c_uint32),
                ("pad", C.c_uint32)]


def _check(rc: int) -> None:
    if rc != 0:
        raise abi.BackendError(abi.load().kst_post_last_error().decode())


def _f32(buf: np.ndarray) -> np.ndarray:
    if buf.dtype != np.float32 or not buf.flags.c_contiguous:
        raise ValueError("expected a contiguous float32 array")
    return buf


class Limiter:
    """LimiterMode::Omni / Brickwall (src/limiter.rs), state carried across blocks."""

    def __init__(self, cfg: LimiterConfig | None = None, **kw):
        self._lib = abi.load()
        self.cfg = cfg or LimiterConfig.default(**kw)
        h = C.c_void_p()
        _check(self._lib.kst_limiter_create(C.byref(self.cfg), C.byref(h)))
        self._h = h

    def process(self, buf: np.ndarray) -> np.ndarray:
        """In place on an interleaved stereo float32 block; returns it."""
        b = _f32(buf)
        _check(self._lib.kst_limiter_process(self._h, abi.ptr(b), b.size))
        return b

    @property
    def latency(self) -> int:
        return int(self._lib.kst_limiter_latency(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.kst_limiter_destroy(self._h)
            self._h = None

    __del__ = close


def clamp_block(buf: np.ndarray) -> int:
    b = _f32(buf)
    return int(abi.load().kst_clamp_block(abi.ptr(b), b.size))


class WavWriter:
    """src/wav.rs:50-159: f32 (or PCM16) RIFF/WAVE, sizes patched on finish."""

    def __init__(self, path: str, sample_rate: int, channels: int = 2, float32: bool = True):
        self._lib = abi.load()
        h = C.c_void_p()
        _check(self._lib.kst_wav_create(str(path).encode(), sample_rate, channels, int(float32),
                                        C.byref(h)))
        self._h = h

    def write_block(self, samples: np.ndarray) -> None:
        s = _f32(samples)
        _check(self._lib.kst_wav_write(self._h, abi.ptr(s), s.size))

    def finish(self) -> None:
        if self._h:
            h, self._h = self._h, None
            _check(self._lib.kst_wav_finish(h))


def read_wav(path: str) -> tuple[np.ndarray, int, int]:
    """(interleaved float32 samples, sample rate, channels) of a file WavWriter wrote."""
    data = open(path, "rb").read()
    if data[:4] != b"RIFF" or data[8:16] != b"WAVEfmt ":
        raise ValueError("not a RIFF/WAVE file")
    tag, ch, sr = int.from_bytes(data[20:22], "little"), int.from_bytes(data[22:24], "little"), \
        int.from_bytes(data[24:28], "little")
    n = int.from_bytes(data[40:44], "little")
    raw = data[44:44 + n]
    if tag == 3:
        return np.frombuffer(raw, "<f4").copy(), sr, ch
    return (np.frombuffer(raw, "<i2").astype(np.float32) / np.float32(32767.0)), sr, ch


def null(a: np.ndarray, b: np.ndarray) -> dict:
    """`kestrel null a b`: peak and RMS difference in dB, `a` the reference."""
    r = NullResult()
    a, b = _f32(a), _f32(b)
    _check(abi.load().kst_null(abi.ptr(a), a.size, abi.ptr(b), b.size, C.byref(r)))
    db = lambda v: -math.inf if v <= 0 else 20 * math.log10(v)
    return {"compared": r.compared, "identical": bool(r.identical), "peak_ref": r.peak_ref,
            "peak_diff": r.peak_diff, "peak_diff_db": db(r.peak_diff), "rms_diff": r.rms_diff,
            "rms_diff_db": db(r.rms_diff), "worst": r.worst}
