"""Build the gfx950 backend library in-tree: kestrel_amd/libkestrel_hip.so.

hipcc cross-compiles for gfx950 without a GPU, so this runs on any machine with
ROCm. Flags that matter for parity:

* ``-ffp-contract=off``: no fused multiply-add the reference does not do; every
  f32 operation rounds where src/cpu.rs rounds. The kernels use an explicit
  fmaf only where the fused and unfused results are provably equal.
* ``-fno-gpu-flush-denormals-to-zero``: keep f32 denormals, as the CPU does.
* ``-target-feature -packed-fp32-ops``: no packed FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libkestrel_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ["kst_backend.hip", "kst_pool.hip", "kst_render.hip", "kst_design.hip", "kst_post.cpp",
           "kst_bank.cpp", "kst_midi.cpp", "kst_driver.cpp", "kst_sf2.cpp",
           "kst_sfz.cpp"]
ARCH = os.environ.get("KST_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 backend needs ROCm's hipcc to build")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        f"-I{os.path.join(ROOT, 'include')}",
    ]


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library. Always rebuilds."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc, *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        objs.append(obj)
    errors = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            errors.append(f"--- {src} ---\n{text}")
        elif verbose and text.strip():
            print(text)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stdout.decode(errors="replace"))
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
