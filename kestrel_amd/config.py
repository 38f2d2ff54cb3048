"""Render configuration: the render-path subset of the reference's `Config`
(src/config.rs:164-409) plus the MI355X knobs of the wave64 backend.

The derived arithmetic (`max_steal`, `pool_slots`, `steal_span`, `admit_take`)
restates src/config.rs:120-131,532-552 exactly; the driver and both backends
must agree on it, so it lives in one place per language.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

# src/config.rs:13-20, 45-52, 70-85, 95-109 (#[repr(u32)] discriminants)
INTERP_NEAREST, INTERP_LINEAR, INTERP_CUBIC = 0, 1, 2
CURVE_EXPONENTIAL, CURVE_LINEAR = 0, 1
STEAL_OLDEST, STEAL_DROP_NEW, STEAL_QUIETEST = 0, 1, 2
ADMIT_LOUDEST, ADMIT_EVEN = 0, 1

FORM_PLAIN, FORM_WAVE64 = 0, 1
# kst_config.profile: off, HIP events between every pass, or around the render
# kernel only (the rest of the block chain runs as unprofiled), or that on every
# RENDER_SAMPLE-th block of a staged run (the others run with no events).
PROFILE_OFF, PROFILE_ALL, PROFILE_RENDER, PROFILE_RENDER_SAMPLED = 0, 1, 2, 3
RENDER_SAMPLE = 8  # KST_RENDER_SAMPLE: PROFILE_RENDER_SAMPLED times blocks 0, 8, 16, ...


class ConfigError(ValueError):
    """A validation failure; the message names the flag (src/config.rs:412-525)."""


@dataclasses.dataclass
class Config:
    # ---- stream / block scheduling (src/config.rs:166-200) ----
    sample_rate: int = 48_000
    block_frames: int = 4096
    reduce_tile: int = 4
    gate_frames: int = 32
    workgroup_size: int = 256
    max_render_workgroups: int = 2048
    # ---- voice pool (src/config.rs:202-247) ----
    max_voices: int = 1 << 20
    steal_rule: int = STEAL_QUIETEST
    admit_rule: int = ADMIT_LOUDEST
    max_steal_percent: int = 25
    steal_fade_frames: int = 96
    sort_voices: bool = True
    # ---- dsp (src/config.rs:249-271) ----
    interpolation: int = INTERP_LINEAR
    decay_curve: int = CURVE_EXPONENTIAL
    release_curve: int = CURVE_EXPONENTIAL
    env_floor: float = 1.0e-5
    filter_enabled: bool = True
    max_param_variants: int = 32
    # ---- numerics (src/config.rs:324-334) ----
    kahan_reduce: bool = False
    nan_guard: bool = False
    # ---- driver and output (src/config.rs:207, 258-313) ----
    max_layers: int = 16
    master_volume: float = 1.0
    default_channel_volume: int = 127
    limiter: bool = True
    limiter_mode: int = 2  # 0 Off, 1 Omni, 2 Brickwall
    limiter_ceiling_db: float = 0.0
    limiter_lookahead_ms: float = 2.0
    limiter_release_ms: float = 60.0
    limiter_sustain_ms: float = 0.0
    limiter_true_peak: bool = True
    # ---- MI355X knobs (no counterpart in the reference) ----
    form: int = FORM_WAVE64
    compensated: bool = False
    waves_per_group: int = 1
    sort_digit_bits: int = 9
    device: int = 0
    profile: int = 0  # PROFILE_OFF, PROFILE_ALL, PROFILE_RENDER or PROFILE_RENDER_SAMPLED

    def validate(self) -> None:
        """src/config.rs:412-525, the render-path subset."""
        if not 8_000 <= self.sample_rate <= 768_000:
            raise ConfigError(f"sample_rate {self.sample_rate} out of range")
        if self.block_frames == 0 or self.block_frames % self.reduce_tile:
            raise ConfigError(f"block_frames ({self.block_frames}) must be a non-zero multiple "
                              f"of reduce_tile ({self.reduce_tile})")
        ws = self.workgroup_size
        if ws & (ws - 1) or not 32 <= ws <= 1024:
            raise ConfigError(f"workgroup_size {ws} must be a power of two in [32, 1024]")
        if self.reduce_tile & (self.reduce_tile - 1):
            raise ConfigError(f"reduce_tile {self.reduce_tile} must be a power of two")
        if (self.gate_frames == 0 or self.gate_frames % self.reduce_tile
                or self.block_frames % self.gate_frames):
            raise ConfigError(f"gate_frames ({self.gate_frames}) must be a multiple of "
                              f"reduce_tile ({self.reduce_tile}) and divide block_frames "
                              f"({self.block_frames})")
        if self.max_voices == 0:
            raise ConfigError("max_voices must be non-zero")
        if not 0 < self.steal_fade_frames < self.block_frames:
            raise ConfigError(f"steal_fade_frames {self.steal_fade_frames} must be in "
                              f"1..block_frames ({self.block_frames})")
        if not 1 <= self.max_steal_percent <= 100:
            raise ConfigError(f"max_steal_percent {self.max_steal_percent} must be in 1..=100")
        if self.block_frames > 65535:
            raise ConfigError(f"block_frames {self.block_frames} must be at most 65535")

    # ---- derived arithmetic, src/config.rs:120-131, 532-552 ----
    def max_steal(self) -> int:
        n = self.max_voices * self.max_steal_percent // 100
        return max(n & 0xFFFFFFFF, 1)

    def pool_slots(self) -> int:
        return min(self.max_voices + self.max_steal(), 0xFFFFFFFF)

    def steal_span(self) -> int:
        return max(self.block_frames - self.steal_fade_frames, 0) or 1

    def admit_take(self, live: int, queued: int) -> int:
        cap = self.max_voices
        want = min(queued, cap)
        if live + want <= cap:
            return want
        need = 0 if self.steal_rule == STEAL_DROP_NEW else min(live + want - cap, live,
                                                               self.max_steal())
        return min(want, cap - (live - need))

    def block_samples(self) -> int:
        return self.block_frames * 2

    def tiles(self) -> int:
        return self.block_frames // self.gate_frames

    def replace(self, **kw) -> "Config":
        return dataclasses.replace(self, **kw)


class KstConfig(C.Structure):
    """`kst_config` of include/kestrel_hip.h."""
    _fields_ = [(n, C.c_uint32) for n in (
        "sample_rate", "block_frames", "gate_frames", "reduce_tile", "workgroup_size",
        "max_render_workgroups", "max_voices", "steal_rule", "admit_rule",
        "max_steal_percent", "steal_fade_frames", "sort_voices", "interpolation",
        "decay_curve", "release_curve")] + [("env_floor", C.c_float)] + [
        (n, C.c_uint32) for n in ("max_param_variants", "kahan_reduce", "form",
                                  "compensated", "waves_per_group",
                                  "sort_digit_bits")] + [("device", C.c_int32),
                                                         ("profile", C.c_uint32)]

    @classmethod
    def from_config(cls, cfg: Config) -> "KstConfig":
        k = cls()
        for name, _ in cls._fields_:
            v = getattr(cfg, name)
            setattr(k, name, float(v) if name == "env_floor" else int(v))
        return k
