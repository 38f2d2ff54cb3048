"""MI355X (gfx950) render backend for the Kestrel black-MIDI synthesizer.

`HipSynth` is the reference's `Backend` trait (src/backend.rs) over the HIP
library's C ABI (include/kestrel_hip.h); `Config` is the render-path subset of
the reference's `Config`; `workload` builds synthetic banks and per-block
inputs. Build the library with `kestrel_amd.build.build()`.
"""
from .config import Config, ConfigError  # noqa: F401

__all__ = ["Config", "ConfigError", "HipSynth"]


def __getattr__(name):
    if name == "HipSynth":
        from .backend import HipSynth
        return HipSynth
    raise AttributeError(name)
