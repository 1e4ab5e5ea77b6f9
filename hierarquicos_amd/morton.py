"""Host-side mirror of the reference's morton_tree_type
(fortran/morton/morton_octree.f90:49-180).

    tree = MortonTreeType().init(N, ms, qs, L=0, ssa=1.2, use_quadrupole=True)
    F = tree.forces_par(G, eps2, theta2)       # (3, N), original particle order

qs is (3, N) like the reference (an (N, 3) array is accepted and transposed).
Keys, depth and permutation are bit-exact to the reference; the tree, its
quadrupoles and the stackless walk run in libhq.so on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class MortonTreeType:
    def __init__(self):
        self._h = C.c_void_p()
        self.N = 0

    def init(self, N, ms, qs, L_depth=0, ssa=1.2, use_quadrupole=True):
        ms = L.f64(ms)
        qs = np.asarray(qs, dtype=np.float64)
        if qs.shape == (N, 3) and N != 3:
            qs = qs.T
        q = np.ascontiguousarray(qs.T)  # (N,3) C order == (3,N) Fortran order
        self.close()
        self.N = int(N)
        self.use_quadrupole = bool(use_quadrupole)
        lib = L.lib()
        L.check(lib.hq_morton_create(C.byref(self._h), self.N, int(L_depth), float(ssa),
                                     int(self.use_quadrupole)))
        L.check(lib.hq_morton_build(self._h, C.c_void_p(ms.ctypes.data),
                                    C.c_void_p(q.ctypes.data), L.HQ_HOST))
        dep = C.c_int(0)
        nn = C.c_int64(0)
        L.check(lib.hq_morton_info(self._h, C.byref(dep), C.byref(nn)))
        self.depth = dep.value
        self.nnodes = nn.value
        return self

    def _forces(self, G, eps2, theta2, mult):
        mult = self.use_quadrupole if mult is None else mult
        buf = np.zeros(3 * self.N)
        L.check(L.lib().hq_morton_forces(self._h, float(G), float(eps2), float(theta2),
                                         int(bool(mult)), C.c_void_p(buf.ctypes.data),
                                         L.HQ_LAYOUT_3N, L.HQ_HOST))
        return buf.reshape(self.N, 3).T

    def forces_seq(self, G, eps2, theta2, mult_par=None):
        """forces_seq (:131-148); (3, N)."""
        return self._forces(G, eps2, theta2, mult_par)

    def forces_par(self, G, eps2, theta2, mult_par=None, num_threads=None):
        """forces_par (:150-180); the GPU ignores num_threads."""
        return self._forces(G, eps2, theta2, mult_par)

    def forces_over_p(self, p, G, eps2, theta2, mult_par=None):
        """Force on sorted particle p (0-based), as forces_individual (:113-129)."""
        F = self._forces(G, eps2, theta2, mult_par)
        return F[:, self.export()["perm"][p]].copy()

    def export(self):
        nn, n = self.nnodes, self.N
        o = dict(perm=np.zeros(n, np.int32), prefix=np.zeros(nn, np.int64),
                 level=np.zeros(nn, np.int32), first=np.zeros(nn, np.int32),
                 last=np.zeros(nn, np.int32), parent=np.zeros(nn, np.int32),
                 nchild=np.zeros(nn, np.int32), first_child=np.zeros(nn, np.int32),
                 next=np.zeros(nn, np.int32), mass=np.zeros(nn), qcm=np.zeros((nn, 3)),
                 quad=np.zeros((nn, 6)))
        L.check(L.lib().hq_morton_export(
            self._h, L.ptr(o["perm"], L._i), L.ptr(o["prefix"], L._i64),
            *[L.ptr(o[k], L._i) for k in ("level", "first", "last", "parent", "nchild",
                                          "first_child", "next")],
            L.ptr(o["mass"]), L.ptr(o["qcm"]), L.ptr(o["quad"])))
        return o

    def close(self):
        if self._h:
            L.lib().hq_morton_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
