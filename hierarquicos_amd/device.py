"""Device-resident arrays and device-level entry points.

DeviceArray holds a float64 / int64 buffer in HBM allocated through libhq's
own HIP runtime (hq_dev_alloc), so callers need no second GPU runtime in the
process.  OctreeType, direct_forces and the benchmark take DeviceArrays to
keep inputs and forces resident on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class DeviceArray:
    """A contiguous 1-D (or flattened) numeric buffer in device memory."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = (shape,) if isinstance(shape, (int, np.integer)) else tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        L.check(L.lib().hq_dev_alloc(C.byref(self.ptr), self.nbytes))

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype)
        L.check(L.lib().hq_dev_copy(d.ptr, C.c_void_p(a.ctypes.data), d.nbytes, 0))
        return d

    def numpy(self):
        a = np.empty(self.shape, self.dtype)
        L.check(L.lib().hq_dev_copy(C.c_void_p(a.ctypes.data), self.ptr, self.nbytes, 1))
        return a

    def __len__(self):
        return self.shape[0]

    def free(self):
        if self.ptr:
            L.lib().hq_dev_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def is_device(a):
    return isinstance(a, DeviceArray)


def sync():
    L.check(L.lib().hq_dev_sync())


def direct_forces(m, x, y, z, G=1.0, eps2=0.0, targets=None, precision="fp64", out=None):
    """compute_forces_direct (fortran/direct_forces.f90:8) in gather form on the
    GPU.  Inputs are DeviceArrays or numpy arrays (copied to the device);
    returns the (nt, 3) forces on the targets (all particles by default) as a
    numpy array, or writes them into the DeviceArray ``out``."""
    keep = []

    def dev(a, dt=np.float64):
        if is_device(a):
            return a
        d = DeviceArray.from_numpy(np.asarray(a, dt))
        keep.append(d)
        return d
    dm, dx, dy, dz = (dev(a) for a in (m, x, y, z))
    n = len(dm)
    tg = dev(targets, np.int64) if targets is not None else None
    nt = len(tg) if tg is not None else n
    F = out if out is not None else DeviceArray((nt, 3))
    prec = {"fp64": L.HQ_FP64, "mixed": L.HQ_MIXED}[precision]
    L.check(L.lib().hq_direct_forces_device(n, dm.ptr, dx.ptr, dy.ptr, dz.ptr, nt,
                                            tg.ptr if tg is not None else None, float(G),
                                            float(eps2), prec, F.ptr))
    return F if out is not None else F.numpy()


def morton_keys(m, q, L_depth=0, ssa=1.2):
    """Keys and stable sort permutation of the reference's Morton octree
    (morton_octree.f90:357-417) for host arrays m (N,) and q (N, 3).  Returns
    (keys in original order, perm with perm[i] = original index of sorted
    slot i, depth L)."""
    m = L.f64(m)
    q = L.f64(q)
    n = len(m)
    keys = np.zeros(n, np.int64)
    perm = np.zeros(n, np.int32)
    Lo = C.c_int(0)
    L.check(L.lib().hq_morton_keys(n, C.c_void_p(m.ctypes.data), C.c_void_p(q.ctypes.data),
                                   int(L_depth), float(ssa), C.c_void_p(keys.ctypes.data),
                                   C.c_void_p(perm.ctypes.data), C.byref(Lo), L.HQ_HOST))
    return keys, perm, Lo.value
