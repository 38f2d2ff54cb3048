"""The caller of the backend: the driver's block loop over the C ABI.

Mirrors the reference's `Driver` (src/driver.rs:240-1192): `Driver.open(cfg,
bank, midi)` then `next_block(backend, out)` until it returns False. The work
is native (kst_driver.cpp, kst_bank.cpp, kst_midi.cpp); this module only
marshals. `Bank` is the flattened instrument set of src/bank.rs, `MidiFile`
the streaming reader of src/midi.rs.

The block can also be driven step by step -- `begin_block`, `rows`,
`pending_variants`, `admit`, `end_block` -- for a caller with its own backend;
`next_block` is exactly those steps against `HipSynth`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import EVENT_DTYPE, PARAMS_DTYPE, PRESET_DTYPE, REGION_DTYPE, SAMPLE_DTYPE, SPAWN_DTYPE
from .config import Config, KstConfig
from .post import LimiterConfig


class DriverConfig(C.Structure):
    """`kst_driver_config`: the driver-side settings of src/config.rs."""
    _fields_ = [("max_layers", C.c_uint32), ("master_volume", C.c_float),
                ("default_channel_volume", C.c_uint32), ("filter_enabled", C.c_uint32),
                ("limiter", C.c_uint32), ("nan_guard", C.c_uint32),
                ("limiter_cfg", LimiterConfig)]

    @classmethod
    def from_config(cls, cfg: Config) -> "DriverConfig":
        d = cls()
        abi.load().kst_default_driver_config(C.byref(d))
        d.max_layers = int(cfg.max_layers)
        d.master_volume = float(cfg.master_volume)
        d.default_channel_volume = int(cfg.default_channel_volume)
        d.filter_enabled = int(cfg.filter_enabled)
        d.limiter = int(cfg.limiter)
        d.nan_guard = int(cfg.nan_guard)
        lc = d.limiter_cfg
        lc.sample_rate = int(cfg.sample_rate)
        lc.mode = int(cfg.limiter_mode)
        lc.ceiling_db = float(cfg.limiter_ceiling_db)
        lc.lookahead_ms = float(cfg.limiter_lookahead_ms)
        lc.release_ms = float(cfg.limiter_release_ms)
        lc.sustain_ms = float(cfg.limiter_sustain_ms)
        lc.true_peak = int(cfg.limiter_true_peak)
        return d


class LoadConfig(C.Structure):
    """`kst_load_config`: resample_pool and sample_pool_budget, src/config.rs:319-322."""
    _fields_ = [("resample_pool", C.c_uint32), ("pad", C.c_uint32),
                ("sample_pool_budget", C.c_uint64)]


def resample_i16(src: np.ndarray, ratio: float) -> np.ndarray:
    """resample_i16, src/resample.rs:102-108."""
    lib = abi.load()
    s = _arr(src, np.int16)
    n = _check(lib.kst_resample_i16(abi.ptr(s), s.size, float(ratio), None, 0))
    out = np.zeros(n, np.int16)
    _check(lib.kst_resample_i16(abi.ptr(s), s.size, float(ratio), abi.ptr(out), out.size))
    return out


class Division(C.Structure):
    _fields_ = [("smpte", C.c_uint32), ("ppq", C.c_uint32), ("fps", C.c_uint32),
                ("ticks_per_frame", C.c_uint32)]

    @classmethod
    def ppq_of(cls, ppq: int) -> "Division":
        return cls(0, int(ppq), 0, 0)


class DriverStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "variant_fallbacks", "variant_rebuilds", "frames", "blocks", "events", "dropped", "notes",
        "voices_spawned", "clipped", "last_stolen", "last_want_energy", "last_take_energy")] + [
        (n, C.c_uint32) for n in ("param_variants", "variant_states", "last_live", "last_want",
                                  "last_take")] + [("peak", C.c_float)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


def _check(rc: int) -> int:
    if rc < 0:
        raise abi.BackendError(abi.load().kst_host_last_error().decode(errors="replace"))
    return rc


def _arr(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


class Bank:
    """Bank (src/bank.rs) built from resolved regions: params layout, variant 0
    and the admission preview tables, with `cfg`'s rate, curves and volume."""

    def __init__(self, cfg: Config, pool: np.ndarray, samples: np.ndarray, regions: np.ndarray,
                 presets: np.ndarray, preset_regions: np.ndarray, pool_rate: int = 0, *,
                 _handle=None):
        self._lib = abi.load()
        self.cfg = cfg
        h = C.c_void_p()
        if _handle is not None:
            h = _handle
        else:
            pool_a = _arr(pool, np.int16)
            smp, reg = _arr(samples, SAMPLE_DTYPE), _arr(regions, REGION_DTYPE)
            pre, pre_r = _arr(presets, PRESET_DTYPE), _arr(preset_regions, np.uint32)
            k, d = KstConfig.from_config(cfg), DriverConfig.from_config(cfg)
            _check(self._lib.kst_bank_create(
                C.byref(k), C.byref(d), abi.ptr(pool_a), pool_a.size, int(pool_rate),
                abi.ptr(smp), smp.size, abi.ptr(reg), reg.size, abi.ptr(pre), pre.size,
                abi.ptr(pre_r), pre_r.size, C.byref(h)))
        self._h = h
        pl, npar, nreg, msl = C.c_size_t(), C.c_size_t(), C.c_uint32(), C.c_uint32()
        _check(self._lib.kst_bank_info(h, C.byref(pl), C.byref(npar), C.byref(nreg), C.byref(msl)))
        self.n_params, self.n_regions, self.max_sample_len = npar.value, nreg.value, msl.value
        self.pool = _copy(C.c_void_p(self._lib.kst_bank_pool(h)), pl.value, np.int16)
        self.params = _copy(C.c_void_p(self._lib.kst_bank_params(h)), self.n_params, PARAMS_DTYPE)

    @classmethod
    def load_sf2(cls, cfg: Config, path: str, resample_pool: bool = True,
                 sample_pool_budget: int = 2 << 30) -> "Bank":
        """sf2::load (src/sf2.rs:221-506) with the pool settings of src/config.rs:319-322."""
        lib = abi.load()
        k, d = KstConfig.from_config(cfg), DriverConfig.from_config(cfg)
        lc = LoadConfig(int(resample_pool), 0, int(sample_pool_budget))
        h = C.c_void_p()
        _check(lib.kst_bank_load_sf2(C.byref(k), C.byref(d), C.byref(lc), str(path).encode(),
                                     C.byref(h)))
        return cls(cfg, None, None, None, None, None, _handle=h)

    @classmethod
    def load_sfz(cls, cfg: Config, path: str, resample_pool: bool = True) -> "Bank":
        """sfz::load (src/sfz.rs:254-425)."""
        lib = abi.load()
        k, d = KstConfig.from_config(cfg), DriverConfig.from_config(cfg)
        lc = LoadConfig(int(resample_pool), 0, 0)
        h = C.c_void_p()
        _check(lib.kst_bank_load_sfz(C.byref(k), C.byref(d), C.byref(lc), str(path).encode(),
                                     C.byref(h)))
        return cls(cfg, None, None, None, None, None, _handle=h)

    def contents(self) -> dict:
        """The flattened samples, regions, presets and their region lists, and the pool rate."""
        ptrs = [C.c_void_p() for _ in range(4)]
        ns = [C.c_size_t() for _ in range(4)]
        rate = C.c_uint32()
        _check(self._lib.kst_bank_contents(self._h, C.byref(ptrs[0]), C.byref(ns[0]),
                                           C.byref(ptrs[1]), C.byref(ns[1]), C.byref(ptrs[2]),
                                           C.byref(ns[2]), C.byref(ptrs[3]), C.byref(ns[3]),
                                           C.byref(rate)))
        return {"samples": _copy(ptrs[0], ns[0].value, SAMPLE_DTYPE),
                "regions": _copy(ptrs[1], ns[1].value, REGION_DTYPE),
                "presets": _copy(ptrs[2], ns[2].value, PRESET_DTYPE),
                "preset_regions": _copy(ptrs[3], ns[3].value, np.uint32),
                "pool": self.pool, "pool_rate": rate.value}

    def variant(self, cc71: int, cc72: int, cc73: int, cc74: int, cc75: int) -> np.ndarray:
        """Bank::build_variant for ParamMod::from_controllers (src/bank.rs:336-350, 836-853)."""
        cc = (C.c_uint8 * 5)(cc71, cc72, cc73, cc74, cc75)
        out = np.zeros(self.n_params, PARAMS_DTYPE)
        _check(self._lib.kst_bank_variant(self._h, cc, abi.ptr(out), out.size))
        return out

    def find_preset(self, bank: int, program: int) -> int:
        return int(self._lib.kst_bank_find_preset(self._h, int(bank), int(program)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.kst_bank_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MidiFile:
    """MidiStream (src/midi.rs:249-361): the tick-ordered merge of every track."""

    def __init__(self, path: str):
        self._lib = abi.load()
        h = C.c_void_p()
        _check(self._lib.kst_midi_open(str(path).encode(), C.byref(h)))
        self._h = h
        div, fmt, tr = Division(), C.c_uint32(), C.c_uint32()
        _check(self._lib.kst_midi_info(h, C.byref(div), C.byref(fmt), C.byref(tr)))
        self.division, self.format, self.track_count = div, fmt.value, tr.value

    def read(self, cap: int = 1 << 16) -> np.ndarray:
        out = np.zeros(cap, EVENT_DTYPE)
        n = _check(self._lib.kst_midi_read(self._h, abi.ptr(out), cap))
        return out[:n]

    def events(self) -> np.ndarray:
        parts = []
        while True:
            e = self.read()
            if e.size == 0:
                break
            parts.append(e)
        return np.concatenate(parts) if parts else np.zeros(0, EVENT_DTYPE)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.kst_midi_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Driver:
    """Driver (src/driver.rs): MIDI in, one block of backend calls per step."""

    def __init__(self, cfg: Config, bank: Bank, midi_path: str | None = None, *,
                 division: Division | None = None, events: np.ndarray | None = None):
        self._lib = abi.load()
        self.cfg, self.bank = cfg, bank  # the bank must outlive the driver
        self._k, self._d = KstConfig.from_config(cfg), DriverConfig.from_config(cfg)
        h = C.c_void_p()
        if midi_path is not None:
            _check(self._lib.kst_driver_open(C.byref(self._k), C.byref(self._d), bank._h,
                                             str(midi_path).encode(), C.byref(h)))
        else:
            ev = _arr(events if events is not None else np.zeros(0, EVENT_DTYPE), EVENT_DTYPE)
            _check(self._lib.kst_driver_create(C.byref(self._k), C.byref(self._d), bank._h,
                                               C.byref(division or Division.ppq_of(480)),
                                               abi.ptr(ev), ev.size, C.byref(h)))
        self._h = h
        self.block_samples = cfg.block_frames * 2

    @classmethod
    def open(cls, cfg: Config, bank: Bank, midi_path: str) -> "Driver":
        return cls(cfg, bank, midi_path)

    def next_block(self, backend, out: np.ndarray) -> bool:
        """Driver::next_block against a HipSynth; False once the file and its tail are done."""
        if out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous float32 array")
        return bool(_check(self._lib.kst_driver_next_block(self._h, backend._h, abi.ptr(out),
                                                           out.size)))

    # ---- the same block, step by step ----
    def begin_block(self) -> list[tuple[int, np.ndarray]]:
        """Events, rows and LFOs; returns the params variants to upload (index, table)."""
        n = _check(self._lib.kst_driver_begin_block(self._h))
        out = []
        for i in range(n):
            idx, data, cnt = C.c_uint32(), C.c_void_p(), C.c_size_t()
            _check(self._lib.kst_driver_pending_variant(self._h, i, C.byref(idx), C.byref(data),
                                                        C.byref(cnt)))
            out.append((idx.value, _copy(data, cnt.value, PARAMS_DTYPE)))
        return out

    def rows(self):
        """(gate rows, channel rows, (bend, gain, variant) flags) of this block."""
        g, ng, c, nc = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
        flags = (C.c_uint32 * 3)()
        _check(self._lib.kst_driver_rows(self._h, C.byref(g), C.byref(ng), C.byref(c),
                                         C.byref(nc), flags))
        return (_copy(g, ng.value, np.uint32), _copy(c, nc.value, np.uint32),
                tuple(bool(x) for x in flags))

    def admit(self, live: int, stolen_before: int) -> np.ndarray:
        """Admission and the build of the admitted voices: this block's spawn commands."""
        p, n = C.c_void_p(), C.c_size_t()
        _check(self._lib.kst_driver_admit(self._h, int(live), int(stolen_before), C.byref(p),
                                          C.byref(n)))
        return _copy(p, n.value, SPAWN_DTYPE)

    def end_block(self, out: np.ndarray, stolen_after_spawn: int, after_render: dict) -> bool:
        """Limiter, clamp and counters on the rendered block (in place)."""
        st = abi.KstBlockStats(int(after_render["active_voices"]), int(after_render["stolen"]),
                               int(after_render["dropped"]), float(after_render["peak"]), 0)
        return bool(_check(self._lib.kst_driver_end_block(self._h, abi.ptr(out), out.size,
                                                          int(stolen_after_spawn), C.byref(st))))

    def stats(self) -> dict:
        s = DriverStats()
        _check(self._lib.kst_driver_get_stats(self._h, C.byref(s)))
        return s.as_dict()

    def timings(self) -> dict:
        """Host milliseconds per block of next_block, by phase."""
        out = (abi.KstTiming * 8)()
        n = self._lib.kst_driver_timings(self._h, out, 8)
        return {out[i].name.decode(): out[i].ms for i in range(min(n, 8))}

    def seconds_rendered(self) -> float:
        return self.stats()["frames"] / self.cfg.sample_rate

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.kst_driver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _copy(p: C.c_void_p, n: int, dtype) -> np.ndarray:
    dtype = np.dtype(dtype)
    if n == 0 or not p.value:
        return np.zeros(0, dtype)
    raw = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n * dtype.itemsize,))
    return raw.view(dtype).copy()


def cc_role(num: int) -> tuple[str, str, str]:
    """cc_role (src/driver.rs:58-131): (name, role, note)."""
    name, note = C.c_char_p(), C.c_char_p()
    lib = abi.load()
    role = lib.kst_cc_role(int(num), C.byref(name), C.byref(note))
    return (name.value.decode(), ("applied", "inert", "missing")[role], note.value.decode())
