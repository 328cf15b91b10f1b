"""C ABI without a GPU: the library loads, exports every function that
include/radis_ldm.h declares, and its host-only planning agrees with NumPy."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from radis_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "radis_ldm.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(radis_ldm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (radis_ldm_\w+)", out))
    assert set(names) <= exported
    assert set(_native.EXPORTED) <= exported


def test_abi_version():
    assert _native.load().radis_ldm_abi_version() == _native.ABI_VERSION == 6


@pytest.mark.parametrize("wmin,wmax,wstep", [(1900.0, 2300.0, 0.002), (2000.0, 2500.0, 0.001),
                                             (1000.0, 4000.0, 0.005), (2012.3456, 2063.1, 0.001),
                                             (2100.0, 2100.5, 0.002)])
def test_plan_grid_matches_numpy(wmin, wmax, wstep):
    from radis_amd.engine import plan_grid

    nv, L = plan_grid(wmin, wmax, wstep)
    assert nv == np.arange(wmin, wmax + wstep, wstep).size
    assert L >= 2 * nv and L % 256 == 0
    x = L // 256
    for p in (2, 3, 5, 7):
        while x % p == 0:
            x //= p
    assert x == 1
    # the smallest admissible run-time-sized length: 2 N_v itself, or at
    # least the image correction's floor; a split with compile-time-sized
    # passes is taken up to 1.35x that length
    from radis_amd import autotune

    gen = min(Lc for Lc, _, g in autotune.fft_candidates(nv) if g)
    need = 2 * nv
    assert L == need or L >= autotune.wrap_floor(need)
    assert gen <= L <= 1.35 * gen


def test_random_windows_grid_lengths():
    from radis_amd.engine import plan_grid

    rng = np.random.default_rng(7)
    for _ in range(200):
        a = round(rng.uniform(500, 4000), 6)
        ws = float(rng.choice([0.001, 0.002, 0.005, 0.01]))
        b = round(a + rng.uniform(1, 300), 6)
        nv, _ = plan_grid(a, b, ws)
        assert nv == np.arange(a, b + ws, ws).size


def test_errors_are_reported_not_raised_across_the_abi():
    lib = _native.load()
    nv, L = ctypes.c_int64(), ctypes.c_int64()
    rc = lib.radis_ldm_plan_grid(10.0, 5.0, 0.1, ctypes.byref(nv), ctypes.byref(L))
    assert rc == _native.EINVAL
    assert b"invalid grid" in lib.radis_ldm_last_error()
    rc = lib.radis_ldm_plan_grid(0.0, 1e6, 1e-4, ctypes.byref(nv), ctypes.byref(L))
    assert rc in (_native.EINVAL, _native.ERANGE)
    with pytest.raises(_native.LdmError):
        _native.check(rc)


def test_null_arguments_rejected():
    lib = _native.load()
    assert lib.radis_ldm_ctx_create(0, None) == _native.EINVAL
    assert lib.radis_ldm_db_upload(None, None, None) == _native.EINVAL
    assert lib.radis_ldm_batch_create(None, None, 0, None, None) == _native.EINVAL
    assert lib.radis_ldm_batch_run(None, None, None) == _native.EINVAL
    assert lib.radis_ldm_run_noalloc(None, None, None, None, None, None, None) == _native.EINVAL
