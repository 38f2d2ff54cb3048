"""ctypes binding of the CPU oracle (oracle/libkestrel_oracle.so).

TEST INFRASTRUCTURE ONLY. The oracle restates the reference's single-threaded
CPU backend (src/cpu.rs) and its integer helpers; tests/, the package smoke()
and bench.py's cpu_baseline leg use it as the checker the HIP backend is
compared against. Nothing in the product package imports this module.

`Oracle` mirrors the reference's `Backend` trait (src/backend.rs:22-58):
set_gates / set_channels / set_params_variant / spawn / render / stats / name.
Configs are duck-typed: any object with the reference's Config field names
(src/config.rs:164-343) works.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkestrel_oracle.so")

# src/voice.rs:22-57 (76 B, #[repr(C)])
SPAWN_DTYPE = np.dtype([
    ("phase_lo", "<u4"), ("phase_hi", "<u4"), ("step_lo", "<u4"), ("step_hi", "<u4"),
    ("smp_base", "<u4"), ("smp_len", "<u4"), ("loop_start", "<u4"), ("loop_end", "<u4"),
    ("flags", "<u4"), ("params", "<u4"), ("variant", "<u4"), ("region", "<u4"),
    ("gate_slot", "<u4"), ("ordinal", "<u4"), ("start_rel", "<u4"),
    ("note_id_lo", "<u4"), ("note_id_hi", "<u4"), ("gain_l", "<f4"), ("gain_r", "<f4"),
])
assert SPAWN_DTYPE.itemsize == 76

# src/bank.rs:166-190 (48 B)
PARAMS_DTYPE = np.dtype([
    ("attack_rate", "<f4"), ("attack_end", "<f4"), ("decay_coef", "<f4"),
    ("decay_target", "<f4"), ("sustain", "<f4"), ("release_coef", "<f4"),
    ("b0", "<f4"), ("b1", "<f4"), ("a1", "<f4"), ("a2", "<f4"),
    ("flags", "<u4"), ("pad", "<u4"),
])
assert PARAMS_DTYPE.itemsize == 48

PARAMS64_DTYPE = np.dtype([
    ("attack_rate", "<f8"), ("attack_end", "<f8"), ("decay_coef", "<f8"),
    ("decay_target", "<f8"), ("sustain", "<f8"), ("release_coef", "<f8"),
    ("b0", "<f8"), ("b1", "<f8"), ("a1", "<f8"), ("a2", "<f8"),
    ("flags", "<u4"), ("pad", "<u4"),
])
assert PARAMS64_DTYPE.itemsize == 88

VOICE_FIELDS = 24


class KorConfig(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "sample_rate", "block_frames", "gate_frames", "max_voices", "steal_rule",
        "admit_rule", "max_steal_percent", "steal_fade_frames", "interpolation",
        "decay_curve", "release_curve")] + [("env_floor", C.c_float)]


class KorStats(C.Structure):
    _fields_ = [("active_voices", C.c_uint64), ("stolen", C.c_uint64),
                ("dropped", C.c_uint64), ("peak", C.c_float), ("pad", C.c_uint32)]


def build(force: bool = False) -> str:
    """Compile the oracle with its own Makefile (g++ only)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "-s"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        u32, u64, f32, f64, sz, i = C.c_uint32, C.c_uint64, C.c_float, C.c_double, C.c_size_t, C.c_int
        sigs = {
            "kor_fixed_from_f64": (u64, [f64]),
            "kor_fixed_scale": (u64, [u64, u32]),
            "kor_fixed_frac_f32": (f32, [u64]),
            "kor_bend_factor": (u32, [f64]),
            "kor_max_steal": (u32, [P]),
            "kor_pool_slots": (u32, [P]),
            "kor_steal_span": (u32, [P]),
            "kor_admit_take": (u32, [P, u32, u32]),
            "kor_spawn_pick": (u64, [u64, u64, u64]),
            "kor_steal_key": (u64, [u64, f32, i]),
            "kor_sort_key_plan": (None, [u32, u32, P]),
            "kor_biquad_lowpass": (None, [f32, f32, f32, P]),
            "kor_biquad_lowpass_f64": (None, [f64, f64, f64, P]),
            "kor_make_params": (None, [f32, f32, f32, f32, f32, f32, f32, f32, i, i, f32, i, P]),
            "kor_make_params_f64": (None, [f64, f64, f64, f64, f64, f64, f64, f64, i, i, f64, i, P]),
            "kor_create": (P, [P, P, sz, P, sz, i, i]),
            "kor_destroy": (None, [P]),
            "kor_set_params64": (None, [P, P, sz]),
            "kor_set_gates": (i, [P, P, sz]),
            "kor_set_channels": (i, [P, P, sz, i, i, i]),
            "kor_set_params_variant": (i, [P, u32, P, sz]),
            "kor_spawn": (i, [P, P, sz]),
            "kor_render": (i, [P, P, sz]),
            "kor_render_f64": (i, [P, P, sz]),
            "kor_stats": (None, [P, P]),
            "kor_set_sorted": (None, [P, u32, u32, i]),
            "kor_set_threads": (None, [P, i]),
            "kor_pool_len": (sz, [P]),
            "kor_pool_export": (None, [P, P, sz]),
            "kor_pool_import": (None, [P, P, sz, sz]),
            "kor_set_counters": (None, [P, u64, u64]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def kor_config(cfg) -> KorConfig:
    return KorConfig(
        sample_rate=cfg.sample_rate, block_frames=cfg.block_frames,
        gate_frames=cfg.gate_frames, max_voices=cfg.max_voices,
        steal_rule=int(cfg.steal_rule), admit_rule=int(cfg.admit_rule),
        max_steal_percent=cfg.max_steal_percent, steal_fade_frames=cfg.steal_fade_frames,
        interpolation=int(cfg.interpolation), decay_curve=int(cfg.decay_curve),
        release_curve=int(cfg.release_curve), env_floor=cfg.env_floor)


# ---- free functions -------------------------------------------------------

def fixed_from_f64(v: float) -> int:
    return lib().kor_fixed_from_f64(v)


def fixed_scale(phase: int, factor: int) -> int:
    return lib().kor_fixed_scale(phase, factor)


def frac_f32(phase: int) -> float:
    return lib().kor_fixed_frac_f32(phase)


def bend_factor(semitones: float) -> int:
    return lib().kor_bend_factor(semitones)


def max_steal(cfg) -> int:
    k = kor_config(cfg)
    return lib().kor_max_steal(C.byref(k))


def pool_slots(cfg) -> int:
    k = kor_config(cfg)
    return lib().kor_pool_slots(C.byref(k))


def steal_span(cfg) -> int:
    k = kor_config(cfg)
    return lib().kor_steal_span(C.byref(k))


def admit_take(cfg, live: int, queued: int) -> int:
    k = kor_config(cfg)
    return lib().kor_admit_take(C.byref(k), live, queued)


def spawn_pick(i: int, want: int, take: int) -> int:
    return lib().kor_spawn_pick(i, want, take)


def steal_key(note_id: int, level: float, quietest: bool) -> int:
    return lib().kor_steal_key(note_id, level, int(quietest))


def sort_key_plan(n_regions: int, max_len: int) -> dict:
    out = np.zeros(6, np.uint32)
    lib().kor_sort_key_plan(n_regions, max_len, _ptr(out))
    return dict(zip(("bits", "region_shift", "stage_shift", "phase_shift", "phase_mask",
                     "dead_region"), (int(x) for x in out)))


def biquad_lowpass(fc: float, q_cb: float, sr: float):
    out = np.zeros(4, np.float32)
    lib().kor_biquad_lowpass(fc, q_cb, sr, _ptr(out))
    return tuple(out)


def biquad_lowpass_f64(fc: float, q_cb: float, sr: float):
    out = np.zeros(4, np.float64)
    lib().kor_biquad_lowpass_f64(fc, q_cb, sr, _ptr(out))
    return tuple(out)


def make_params(attack_s, hold_s, decay_s, sustain, release_s, fc_cents, q_cb, sr,
                exp_decay=True, exp_release=True, env_floor=1e-5, filter_enabled=True):
    out = np.zeros(1, PARAMS_DTYPE)
    lib().kor_make_params(attack_s, hold_s, decay_s, sustain, release_s, fc_cents, q_cb, sr,
                          int(exp_decay), int(exp_release), env_floor, int(filter_enabled),
                          _ptr(out))
    return out[0]


def make_params_f64(attack_s, hold_s, decay_s, sustain, release_s, fc_cents, q_cb, sr,
                    exp_decay=True, exp_release=True, env_floor=1e-5, filter_enabled=True):
    out = np.zeros(1, PARAMS64_DTYPE)
    lib().kor_make_params_f64(attack_s, hold_s, decay_s, sustain, release_s, fc_cents, q_cb,
                              sr, int(exp_decay), int(exp_release), env_floor,
                              int(filter_enabled), _ptr(out))
    return out[0]


def design_table_f64(design: np.ndarray, cfg) -> np.ndarray:
    """A bank's params table designed in f64 (`make_params_f64` per region), from
    the design inputs the device's FP64 design takes (precision form)."""
    out = np.zeros(len(design), PARAMS64_DTYPE)
    for i, d in enumerate(design):
        out[i] = make_params_f64(
            d["attack_s"], d["hold_s"], d["decay_s"], d["sustain"], d["release_s"],
            d["fc_cents"], d["q_cb"], float(cfg.sample_rate), cfg.decay_curve == 0,
            cfg.release_curve == 0, float(np.float32(cfg.env_floor)), bool(d["filter"]))
    return out


# ---- the CPU backend --------------------------------------------------------

class Oracle:
    """CpuSynth restated (src/cpu.rs). precision=1 renders the same algorithm in f64."""

    def __init__(self, cfg, pool: np.ndarray, params: np.ndarray, precision: int = 0,
                 state_only: bool = False):
        self.cfg = cfg
        self._kcfg = kor_config(cfg)
        pool = np.ascontiguousarray(pool, dtype=np.int16)
        params = np.ascontiguousarray(params, dtype=PARAMS_DTYPE)
        self._h = lib().kor_create(C.byref(self._kcfg), _ptr(pool), pool.size, _ptr(params),
                                   params.size, int(precision), int(state_only))
        if not self._h:
            raise ValueError("oracle rejected the config")
        self.block_samples = cfg.block_frames * 2

    def close(self):
        if getattr(self, "_h", None):
            lib().kor_destroy(self._h)
            self._h = None

    __del__ = close

    def set_params64(self, params64: np.ndarray):
        p = np.ascontiguousarray(params64, dtype=PARAMS64_DTYPE)
        lib().kor_set_params64(self._h, _ptr(p), p.size)

    def set_gates(self, rows: np.ndarray):
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        if lib().kor_set_gates(self._h, _ptr(r), r.size) != 0:
            raise ValueError("gate rows have the wrong length")

    def set_channels(self, rows: np.ndarray, bend: bool, gain: bool, variant: bool):
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        if lib().kor_set_channels(self._h, _ptr(r), r.size, int(bend), int(gain),
                                  int(variant)) != 0:
            raise ValueError("channel rows have the wrong length")

    def set_params_variant(self, index: int, data: np.ndarray):
        d = np.ascontiguousarray(data, dtype=PARAMS_DTYPE)
        lib().kor_set_params_variant(self._h, index, _ptr(d), d.size)

    def spawn(self, cmds: np.ndarray):
        c = np.ascontiguousarray(cmds, dtype=SPAWN_DTYPE)
        lib().kor_spawn(self._h, _ptr(c), c.size)

    def render(self, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.block_samples, np.float32)
        if lib().kor_render(self._h, _ptr(out), out.size) != 0:
            raise ValueError("output block has the wrong length")
        return out

    def render_f64(self) -> np.ndarray:
        out = np.empty(self.block_samples, np.float64)
        if lib().kor_render_f64(self._h, _ptr(out), out.size) != 0:
            raise ValueError("output block has the wrong length")
        return out

    def stats(self) -> dict:
        s = KorStats()
        lib().kor_stats(self._h, C.byref(s))
        return {"active_voices": s.active_voices, "stolen": s.stolen, "dropped": s.dropped,
                "peak": s.peak}

    def name(self) -> str:
        return "cpu-reference"

    def set_sorted(self, n_regions: int, max_len: int, on: bool = True):
        lib().kor_set_sorted(self._h, n_regions, max_len, int(on))

    def set_threads(self, n: int):
        lib().kor_set_threads(self._h, n)

    def pool_len(self) -> int:
        return lib().kor_pool_len(self._h)

    def pool_export(self) -> np.ndarray:
        """The pool as a (24, n) u32 array in the device's SoA field order."""
        n = self.pool_len()
        out = np.zeros((VOICE_FIELDS, max(n, 1)), np.uint32)
        lib().kor_pool_export(self._h, _ptr(out), out.shape[1])
        return out[:, :n]

    def pool_import(self, soa: np.ndarray):
        a = np.ascontiguousarray(soa, dtype=np.uint32)
        lib().kor_pool_import(self._h, _ptr(a), a.shape[1], a.shape[1])

    def set_counters(self, stolen: int, dropped: int):
        lib().kor_set_counters(self._h, stolen, dropped)
