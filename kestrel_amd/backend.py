"""`HipSynth`: the reference's `Backend` trait (src/backend.rs:22-58) over the
gfx950 library's C ABI.

Method names, argument meaning and error behaviour follow the trait and
`GpuSynth` (src/gpu/mod.rs): a wrong-length slice or an out-of-range params
variant raises with the reference's message; device work happens in `render`,
which is synchronous. There is no CPU path here: without the built library and
a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import PARAMS_DTYPE, SPAWN_DTYPE, VOICE_FIELDS, check, ptr
from .config import Config, KstConfig


class HipSynth:
    """GpuSynth::new(&Config, Arc<Bank>) for the MI355X backend."""

    def __init__(self, cfg: Config, pool: np.ndarray, params: np.ndarray, n_regions: int,
                 max_sample_len: int):
        cfg.validate()
        self.cfg = cfg
        lib = abi.load()
        self._lib = lib
        self._k = KstConfig.from_config(cfg)
        pool = np.ascontiguousarray(pool, dtype=np.int16)
        params = np.ascontiguousarray(params, dtype=PARAMS_DTYPE)
        h = C.c_void_p()
        check(lib.kst_create(C.byref(self._k), ptr(pool), pool.size, ptr(params), params.size,
                             int(n_regions), int(max_sample_len), C.byref(h)))
        self._h = h
        self._n_params = max(int(params.size), 1)  # an empty table gets one neutral entry
        self.block_samples = cfg.block_frames * 2
        self.gate_words = cfg.tiles() * 2048
        self.chan_words = cfg.tiles() * 64

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.kst_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- the Backend trait ----
    def set_gates(self, rows: np.ndarray) -> None:
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        check(self._lib.kst_set_gates(self._h, ptr(r), r.size))

    def set_channels(self, rows: np.ndarray, bend: bool, gain: bool, variant: bool) -> None:
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        check(self._lib.kst_set_channels(self._h, ptr(r), r.size, int(bend), int(gain),
                                         int(variant)))

    def set_params_variant(self, index: int, data: np.ndarray) -> None:
        d = np.ascontiguousarray(data, dtype=PARAMS_DTYPE)
        check(self._lib.kst_set_params_variant(self._h, int(index), ptr(d), d.size))

    def design_params(self, index: int, designs: np.ndarray) -> None:
        """Design params variant `index` on the device in FP64 (precision form)."""
        d = np.ascontiguousarray(designs, dtype=abi.DESIGN_DTYPE)
        check(self._lib.kst_design_params(self._h, int(index), ptr(d), d.size))

    def params_export(self, index: int) -> np.ndarray:
        out = np.zeros(self._n_params, abi.PARAMS_DTYPE)
        check(self._lib.kst_params_export(self._h, int(index), ptr(out), out.size))
        return out

    def spawn(self, cmds: np.ndarray) -> None:
        c = np.ascontiguousarray(cmds, dtype=SPAWN_DTYPE)
        check(self._lib.kst_spawn(self._h, ptr(c), c.size))

    def render(self, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.block_samples, np.float32)
        if out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous float32 array")
        check(self._lib.kst_render(self._h, ptr(out), out.size))
        return out

    def stats(self) -> dict:
        s = abi.KstBlockStats()
        check(self._lib.kst_stats(self._h, C.byref(s)))
        return {"active_voices": s.active_voices, "stolen": s.stolen, "dropped": s.dropped,
                "peak": s.peak}

    def name(self) -> str:
        return self._lib.kst_name(self._h).decode()

    def set_profile(self, level: int) -> None:
        """PROFILE_OFF, PROFILE_ALL (events between every pass), PROFILE_RENDER
        (events around the render kernel only) or PROFILE_RENDER_SAMPLED (the
        same around every RENDER_SAMPLE-th block of a staged run)."""
        abi.check(self._lib.kst_set_profile(self._h, int(level)))

    def timings(self) -> list[tuple[str, float]]:
        buf = (abi.KstTiming * 16)()
        n = self._lib.kst_timings(self._h, buf, 16)
        return [(buf[i].name.decode(), buf[i].ms) for i in range(min(n, 16))]

    # ---- device-resident block program ----
    def stage_blocks(self, gates: np.ndarray, chans: np.ndarray, chan_flags: np.ndarray,
                     cmds: np.ndarray, cmd_counts: np.ndarray) -> None:
        nb = len(cmd_counts)
        g = np.ascontiguousarray(gates, dtype=np.uint32).reshape(nb, self.gate_words)
        ch = np.ascontiguousarray(chans, dtype=np.uint32).reshape(nb, self.chan_words)
        fl = np.ascontiguousarray(chan_flags, dtype=np.uint32)
        cm = np.ascontiguousarray(cmds, dtype=SPAWN_DTYPE)
        cc = np.ascontiguousarray(cmd_counts, dtype=np.uint32)
        if int(cc.sum()) != cm.size:
            raise ValueError("cmd_counts do not add up to the number of commands")
        self._staged = nb
        check(self._lib.kst_stage_blocks(self._h, nb, ptr(g), ptr(ch), ptr(fl),
                                         ptr(cm) if cm.size else None, ptr(cc)))

    def run_staged(self, use_graph: bool = True) -> None:
        check(self._lib.kst_run_staged(self._h, int(use_graph)))

    def staged_results(self) -> tuple[np.ndarray, list[dict]]:
        nb = self._staged
        out = np.empty((nb, self.block_samples), np.float32)
        st = (abi.KstBlockStats * nb)()
        check(self._lib.kst_staged_results(self._h, ptr(out), st))
        stats = [{"active_voices": s.active_voices, "stolen": s.stolen, "dropped": s.dropped,
                  "peak": s.peak} for s in st]
        return out, stats

    # ---- snapshot hooks ----
    def pool_export(self) -> np.ndarray:
        n = self.stats()["active_voices"]
        out = np.zeros((VOICE_FIELDS, max(n, 1)), np.uint32)
        got = self._lib.kst_pool_export(self._h, ptr(out), out.shape[1])
        if got < 0:
            check(int(got))
        return out[:, :got]

    def pool_import(self, soa: np.ndarray) -> None:
        a = np.ascontiguousarray(soa, dtype=np.uint32)
        check(self._lib.kst_pool_import(self._h, ptr(a), a.shape[1], a.shape[1]))

    def set_counters(self, stolen: int, dropped: int) -> None:
        check(self._lib.kst_set_counters(self._h, int(stolen), int(dropped)))


class HipBatch:
    """Several HipSynth renders run together, block step by block step, with
    every pass of the chain one launch for all of them (kst_batch_*). Stage
    each member's blocks first (HipSynth.stage_blocks); read results per member
    with HipSynth.staged_results."""

    def __init__(self, renders: list):
        self._lib = abi.load()
        self.renders = list(renders)
        arr = (C.c_void_p * len(self.renders))(*[r._h.value for r in self.renders])
        h = C.c_void_p()
        check(self._lib.kst_batch_create(arr, len(self.renders), C.byref(h)))
        self._h = h

    def run(self, use_graph: bool = True) -> None:
        check(self._lib.kst_batch_run(self._h, int(use_graph)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.kst_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
