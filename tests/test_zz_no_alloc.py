"""hq_morton_keys_noalloc, the allocation-free device form of hq_morton_keys,
against hq_morton_keys on the same device-resident inputs, through the C ABI.

Keys, permutation and depth are integer results: the two entry points must
agree bit for bit.  Two calls in a row, the second on a smaller input, check
that the fixed workspace keeps nothing from the previous call that matters;
sizes above the workspace are refused with HQ_ERR_ARGS before any device work.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from hierarquicos_amd import _lib as L
from hierarquicos_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 20


def _fn():
    lib = L.lib()
    f = lib.hq_morton_keys_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p,
                  C.c_void_p, C.POINTER(C.c_int)]
    return lib, f


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 1.5, n)
    return m, q


def _run(entry, lib, m, q, depth, ssa=1.2):
    n = len(m)
    dm = DeviceArray.from_numpy(m)
    dq = DeviceArray.from_numpy(np.ascontiguousarray(q))  # (3,N) Fortran = (N,3) C order
    dk = DeviceArray(n, np.int64)
    dp = DeviceArray(n, np.int32)
    Lo = C.c_int(-1)
    if entry == "noalloc":
        rc = lib.hq_morton_keys_noalloc(n, dm.ptr, dq.ptr, depth, ssa, dk.ptr, dp.ptr, C.byref(Lo))
    else:
        rc = lib.hq_morton_keys(n, dm.ptr, dq.ptr, depth, ssa, dk.ptr, dp.ptr, C.byref(Lo),
                                L.HQ_DEVICE)
    assert rc == L.HQ_OK, lib.hq_last_error()
    return dk.numpy(), dp.numpy(), Lo.value


@pytest.mark.parametrize("depth", [0, 5])
def test_noalloc_keys_equal_allocating_entry(depth):
    lib, _ = _fn()
    for n, seed in ((NOALLOC_MAX, 1), (70001, 2)):  # second call smaller
        print(f"noalloc keys n={n} seed={seed} depth={depth}", flush=True)
        m, q = _inputs(n, seed)
        k1, p1, L1 = _run("noalloc", lib, m, q, depth)
        k0, p0, L0 = _run("alloc", lib, m, q, depth)
        assert L1 == L0
        assert np.array_equal(k1, k0)
        assert np.array_equal(p1, p0)
        if n < 100000:  # and both agree with the oracle
            ko, Lk = O.morton_keys(m, q, depth)
            assert Lk == L1 and np.array_equal(k1, ko)
            assert np.array_equal(p1, np.argsort(ko, kind="stable"))


def test_noalloc_keys_reference_fixture(golden):
    lib, _ = _fn()
    d = golden("plummer_n1000_seed3.npz")
    q = np.stack([d["x"], d["y"], d["z"]], 1)
    k, p, Lv = _run("noalloc", lib, d["m"], q, 0)
    assert Lv == int(d["morton_L"])
    assert np.array_equal(p, d["morton_perm"])
    assert np.array_equal(k[p], d["morton_keys_sorted"])


def test_noalloc_rejects_sizes_beyond_workspace():
    lib, _ = _fn()
    Lo = C.c_int(-1)
    dummy = C.c_void_p(1)  # never dereferenced: the size check comes first
    rc = lib.hq_morton_keys_noalloc(NOALLOC_MAX + 1, dummy, dummy, 0, 1.2, dummy, dummy,
                                    C.byref(Lo))
    assert rc == L.HQ_ERR_ARGS
    assert b"workspace" in lib.hq_last_error()
