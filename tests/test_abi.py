"""The C ABI library loads and exports every symbol include/kestrel_hip.h
declares, with struct layouts that match the reference's #[repr(C)] types.
No device calls: these run without a GPU."""
import ctypes as C
import os
import re

import pytest

from kestrel_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(abi.HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(kst_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(abi.LIB_PATH):
        from kestrel_amd import build
        build.build()
    return abi.load()


def test_header_declares_the_backend_trait():
    syms = declared_symbols()
    for name in ("kst_create", "kst_destroy", "kst_set_gates", "kst_set_channels",
                 "kst_set_params_variant", "kst_spawn", "kst_render", "kst_stats", "kst_name",
                 "kst_timings", "kst_last_error"):
        assert name in syms


def test_library_exports_every_declared_symbol(lib):
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert set(declared_symbols()) == set(abi.SIGNATURES)


def test_struct_layouts():
    assert abi.SPAWN_DTYPE.itemsize == 76
    assert abi.PARAMS_DTYPE.itemsize == 48
    assert C.sizeof(abi.KstBlockStats) == 32
    from kestrel_amd.config import KstConfig
    assert C.sizeof(KstConfig) == 24 * 4


def test_default_config_mirrors_the_reference(lib):
    k = abi.default_config()
    assert (k.sample_rate, k.block_frames, k.gate_frames, k.reduce_tile) == (48000, 4096, 32, 4)
    assert (k.workgroup_size, k.max_render_workgroups, k.max_voices) == (256, 2048, 1 << 20)
    assert (k.steal_rule, k.admit_rule, k.max_steal_percent, k.steal_fade_frames) == (2, 0, 25, 96)
    assert abs(k.env_floor - 1e-5) < 1e-12 and k.max_param_variants == 32


def test_oracle_is_not_linked_into_the_product(lib):
    from oracle import oracle
    assert "oracle" not in abi.LIB_PATH
    with open(abi.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"kor_create" not in blob
    assert os.path.exists(oracle.LIB_PATH)


def test_product_package_never_imports_the_oracle():
    """The oracle is test infrastructure: nothing under kestrel_amd/ imports it."""
    pkg = os.path.join(ROOT, "kestrel_amd")
    pat = re.compile(r"^\s*(from\s+oracle\b|import\s+oracle\b|from\s+\.\.\s*oracle)", re.M)
    for name in sorted(os.listdir(pkg)):
        if name.endswith(".py"):
            with open(os.path.join(pkg, name)) as f:
                assert not pat.search(f.read()), name
