"""ctypes binding of the C ABI in include/kestrel_hip.h (libkestrel_hip.so).

The library is built in-tree by kestrel_amd/build.py. Loading it needs the HIP
runtime; calling anything that touches the device needs a GPU. There is no
fallback: if the library is missing or cannot load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .config import KstConfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkestrel_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "kestrel_hip.h")

# kst_spawn_cmd, src/voice.rs:22-57 (76 B)
SPAWN_DTYPE = np.dtype([
    ("phase_lo", "<u4"), ("phase_hi", "<u4"), ("step_lo", "<u4"), ("step_hi", "<u4"),
    ("smp_base", "<u4"), ("smp_len", "<u4"), ("loop_start", "<u4"), ("loop_end", "<u4"),
    ("flags", "<u4"), ("params", "<u4"), ("variant", "<u4"), ("region", "<u4"),
    ("gate_slot", "<u4"), ("ordinal", "<u4"), ("start_rel", "<u4"),
    ("note_id_lo", "<u4"), ("note_id_hi", "<u4"), ("gain_l", "<f4"), ("gain_r", "<f4"),
])
assert SPAWN_DTYPE.itemsize == 76

# kst_region_params, src/bank.rs:166-190 (48 B)
PARAMS_DTYPE = np.dtype([
    ("attack_rate", "<f4"), ("attack_end", "<f4"), ("decay_coef", "<f4"),
    ("decay_target", "<f4"), ("sustain", "<f4"), ("release_coef", "<f4"),
    ("b0", "<f4"), ("b1", "<f4"), ("a1", "<f4"), ("a2", "<f4"),
    ("flags", "<u4"), ("pad", "<u4"),
])
assert PARAMS_DTYPE.itemsize == 48

# kst_region_design: FP64 inputs of the device coefficient design (64 B)
DESIGN_DTYPE = np.dtype([
    ("attack_s", "<f8"), ("hold_s", "<f8"), ("decay_s", "<f8"), ("sustain", "<f8"),
    ("release_s", "<f8"), ("fc_cents", "<f8"), ("q_cb", "<f8"), ("filter", "<u4"),
    ("pad", "<u4"),
])
assert DESIGN_DTYPE.itemsize == 64

# kst_sample_info, src/bank.rs:35-54 (32 B)
SAMPLE_DTYPE = np.dtype([
    ("start", "<u4"), ("len", "<u4"), ("loop_start", "<u4"), ("loop_end", "<u4"),
    ("rate", "<u4"), ("correction_cents", "<f4"), ("resample_ratio", "<f4"), ("root_key", "<u4"),
])
assert SAMPLE_DTYPE.itemsize == 32

# kst_region, src/bank.rs:57-120 (96 B)
REGION_DTYPE = np.dtype([
    ("sample", "<u4"), ("key_lo", "u1"), ("key_hi", "u1"), ("vel_lo", "u1"), ("vel_hi", "u1"),
    ("root_key_override", "<i2"), ("fixed_key", "<i2"), ("fixed_vel", "<i2"),
    ("coarse_tune", "<i2"), ("fine_tune", "<i2"), ("scale_tuning", "<i2"),
    ("keynum_to_hold", "<i2"), ("keynum_to_decay", "<i2"),
    ("attenuation_cb", "<f4"), ("amp_veltrack", "<f4"), ("pan", "<f4"), ("loop_mode", "<u4"),
    ("addr_start", "<i4"), ("addr_end", "<i4"), ("addr_loop_start", "<i4"),
    ("addr_loop_end", "<i4"), ("exclusive_class", "<u4"),
    ("delay", "<f4"), ("attack", "<f4"), ("hold", "<f4"), ("decay", "<f4"), ("sustain", "<f4"),
    ("release", "<f4"), ("filter_fc_cents", "<f4"), ("filter_q_cb", "<f4"),
    ("filter_veltrack_cents", "<f4"),
])
assert REGION_DTYPE.itemsize == 96

# kst_preset, src/bank.rs:352-362 (16 B)
PRESET_DTYPE = np.dtype([("bank", "<u4"), ("program", "<u4"), ("first", "<u4"), ("count", "<u4")])

# kst_midi_event: (tick, Event), src/midi.rs:17-29 (24 B)
EVENT_DTYPE = np.dtype([
    ("tick", "<u8"), ("kind", "<u4"), ("ch", "u1"), ("a", "u1"), ("b", "u1"), ("pad", "u1"),
    ("value", "<i4"), ("pad2", "<u4"),
])
assert EVENT_DTYPE.itemsize == 24
EV_NOTE_ON, EV_NOTE_OFF, EV_CC, EV_PROGRAM, EV_PITCH_BEND, EV_TEMPO, EV_OTHER = range(7)

VOICE_FIELDS = 24
# shaders/common.wgsl:15-43
FIELDS = ("phase_lo", "phase_hi", "step_lo", "step_hi", "smp_base", "smp_len", "loop_start",
          "loop_end", "flags", "env_stage", "env_level", "gain_l", "gain_r", "filt_z1",
          "filt_z2", "params", "region", "gate_slot", "ordinal", "start_rel", "note_lo",
          "note_hi", "born_variant", "stop_rel")
FIELD = {n: i for i, n in enumerate(FIELDS)}


class KstBlockStats(C.Structure):
    _fields_ = [("active_voices", C.c_uint64), ("stolen", C.c_uint64),
                ("dropped", C.c_uint64), ("peak", C.c_float), ("pad", C.c_uint32)]


class KstTiming(C.Structure):
    _fields_ = [("name", C.c_char_p), ("ms", C.c_double)]


# Every symbol the header declares, with its ctypes signature.
P, u32, u64, i32, i64, sz = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_int64, C.c_size_t
SIGNATURES = {
    "kst_default_config": (None, [P]),
    "kst_create": (i32, [P, P, sz, P, sz, u32, u32, P]),
    "kst_destroy": (None, [P]),
    "kst_set_gates": (i32, [P, P, sz]),
    "kst_set_channels": (i32, [P, P, sz, i32, i32, i32]),
    "kst_set_params_variant": (i32, [P, u32, P, sz]),
    "kst_spawn": (i32, [P, P, sz]),
    "kst_render": (i32, [P, P, sz]),
    "kst_render_submit": (i32, [P]),
    "kst_render_collect": (i32, [P, P, sz]),
    "kst_stats": (i32, [P, P]),
    "kst_name": (C.c_char_p, [P]),
    "kst_timings": (sz, [P, P, sz]),
    "kst_set_profile": (i32, [P, u32]),
    "kst_last_error": (C.c_char_p, []),
    "kst_stage_blocks": (i32, [P, u32, P, P, P, P, P]),
    "kst_run_staged": (i32, [P, i32]),
    "kst_staged_results": (i32, [P, P, P]),
    "kst_batch_create": (i32, [P, sz, P]),
    "kst_batch_run": (i32, [P, i32]),
    "kst_batch_destroy": (None, [P]),
    "kst_design_params": (i32, [P, u32, P, sz]),
    "kst_design_params_noalloc": (i32, [P, u32, P, sz]),
    "kst_default_limiter_config": (None, [P]),
    "kst_limiter_create": (i32, [P, P]),
    "kst_limiter_process": (i32, [P, P, sz]),
    "kst_limiter_latency": (sz, [P]),
    "kst_limiter_destroy": (None, [P]),
    "kst_clamp_block": (u64, [P, sz]),
    "kst_wav_create": (i32, [C.c_char_p, u32, u32, u32, P]),
    "kst_wav_write": (i32, [P, P, sz]),
    "kst_wav_finish": (i32, [P]),
    "kst_null": (i32, [P, sz, P, sz, P]),
    "kst_post_last_error": (C.c_char_p, []),
    "kst_params_export": (i32, [P, u32, P, sz]),
    "kst_pool_export": (i64, [P, P, sz]),
    "kst_pool_import": (i32, [P, P, sz, sz]),
    "kst_set_counters": (i32, [P, u64, u64]),
    "kst_default_driver_config": (None, [P]),
    "kst_bank_create": (i32, [P, P, P, sz, u32, P, sz, P, sz, P, sz, P, sz, P]),
    "kst_bank_info": (i32, [P, P, P, P, P]),
    "kst_bank_pool": (P, [P]),
    "kst_bank_params": (P, [P]),
    "kst_bank_variant": (i32, [P, P, P, sz]),
    "kst_bank_find_preset": (i64, [P, u32, u32]),
    "kst_bank_destroy": (None, [P]),
    "kst_bank_contents": (i32, [P, P, P, P, P, P, P, P, P, P]),
    "kst_default_load_config": (None, [P]),
    "kst_bank_load_sf2": (i32, [P, P, P, C.c_char_p, P]),
    "kst_bank_load_sfz": (i32, [P, P, P, C.c_char_p, P]),
    "kst_resample_i16": (i64, [P, sz, C.c_double, P, sz]),
    "kst_midi_open": (i32, [C.c_char_p, P]),
    "kst_midi_info": (i32, [P, P, P, P]),
    "kst_midi_read": (i64, [P, P, sz]),
    "kst_midi_close": (None, [P]),
    "kst_driver_open": (i32, [P, P, P, C.c_char_p, P]),
    "kst_driver_create": (i32, [P, P, P, P, P, sz, P]),
    "kst_driver_next_block": (i32, [P, P, P, sz]),
    "kst_driver_begin_block": (i32, [P]),
    "kst_driver_pending_variant": (i32, [P, sz, P, P, P]),
    "kst_driver_rows": (i32, [P, P, P, P, P, P]),
    "kst_driver_admit": (i32, [P, u64, u64, P, P]),
    "kst_driver_end_block": (i32, [P, P, sz, u64, P]),
    "kst_driver_get_stats": (i32, [P, P]),
    "kst_driver_timings": (sz, [P, P, sz]),
    "kst_driver_destroy": (None, [P]),
    "kst_cc_role": (i32, [u32, P, P]),
    "kst_host_last_error": (C.c_char_p, []),
}

_lib = None


class BackendError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load the in-tree library. Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BackendError(f"{LIB_PATH} is not built; run python kestrel_amd/build.py")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().kst_last_error().decode(errors="replace")
        raise BackendError(f"kestrel_hip error {rc}: {msg}")


def ptr(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def default_config() -> KstConfig:
    k = KstConfig()
    load().kst_default_config(C.byref(k))
    return k
