"""The allocation-free design entry point (kst_design_params_noalloc) against
kst_design_params, through the C ABI: the same coefficients bit for bit, on two
calls in a row whose second stages fewer rows into the shared device scratch
than the first (so rows the first call left behind are never read)."""
import ctypes as C

import numpy as np
import pytest

from kestrel_amd import abi
from kestrel_amd import workload as W
from kestrel_amd.abi import DESIGN_DTYPE, PARAMS_DTYPE
from kestrel_amd.config import KstConfig

pytestmark = pytest.mark.gpu


def grid(n, seed):
    """n design rows over resonances, cutoffs and envelope times."""
    rng = np.random.default_rng(seed)
    rows = np.zeros(n, DESIGN_DTYPE)
    rows["attack_s"] = rng.choice([0.0, 0.001, 0.05, 0.5], n)
    rows["hold_s"] = rng.choice([0.0, 0.01, 0.2], n)
    rows["decay_s"] = rng.uniform(0.0005, 3.0, n)
    rows["sustain"] = rng.uniform(0.0, 1.0, n)
    rows["release_s"] = rng.uniform(0.0005, 8.0, n)
    rows["fc_cents"] = [W.hz_to_cents(f) for f in rng.uniform(25.0, 23000.0, n)]
    rows["q_cb"] = [W.q_to_cb(q) for q in np.exp(rng.uniform(np.log(0.5), np.log(60.0), n))]
    rows["filter"] = rng.random(n) < 0.9
    return rows


def backend(n_regions, seed, frames=256):
    cfg = W.config_c1()
    bank = W.make_bank(seed, instruments=n_regions, zones=1, frames=frames, loop=(0, 0))
    k = KstConfig.from_config(cfg)
    lib = abi.load()
    pool = np.ascontiguousarray(bank.pool, np.int16)
    params = np.ascontiguousarray(bank.params, PARAMS_DTYPE)
    h = C.c_void_p()
    abi.check(lib.kst_create(C.byref(k), abi.ptr(pool), pool.size, abi.ptr(params), params.size,
                             bank.n_regions, bank.max_sample_len, C.byref(h)))
    return lib, h, bank.n_regions


def export(lib, h, index, n):
    out = np.zeros(n, PARAMS_DTYPE)
    abi.check(lib.kst_params_export(h, index, abi.ptr(out), n))
    return out


def test_noalloc_design_equals_design_on_two_calls_the_second_smaller():
    made = []
    for n_regions, seed in ((120, 3), (40, 4)):  # the second call stages fewer rows
        lib, h, n = backend(n_regions, seed)
        made.append((lib, h))
        rows = np.ascontiguousarray(grid(n, seed + 10))
        abi.check(lib.kst_design_params_noalloc(h, 1, abi.ptr(rows), n))
        abi.check(lib.kst_design_params(h, 2, abi.ptr(rows), n))
        got, want = export(lib, h, 1, n), export(lib, h, 2, n)
        print(f"{n} rows: {int((got['flags'] != 0).sum())} filtered", flush=True)
        assert got.tobytes() == want.tobytes(), n
        assert (got["flags"] != 0).any() and (got["flags"] == 0).any()
    for lib, h in made:
        lib.kst_destroy(h)


def test_noalloc_design_validates_and_rejects_a_size_past_its_scratch():
    """A table of another length than the backend's is refused as
    kst_design_params refuses it; a table longer than the 65,536-row scratch is
    refused before any device work, where kst_design_params (which allocates)
    designs it."""
    lib, h, n = backend(8, 5)
    rows = np.ascontiguousarray(grid(n + 1, 6))
    assert lib.kst_design_params_noalloc(h, 1, abi.ptr(rows), n + 1) != 0
    assert b"expected" in lib.kst_last_error()
    lib.kst_destroy(h)
    lib, h, n = backend(65_537, 7, frames=64)
    rows = np.ascontiguousarray(grid(n, 8))
    assert lib.kst_design_params_noalloc(h, 1, abi.ptr(rows), n) != 0
    assert b"scratch" in lib.kst_last_error()
    abi.check(lib.kst_design_params(h, 1, abi.ptr(rows), n))
    lib.kst_destroy(h)
