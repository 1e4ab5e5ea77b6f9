"""ctypes front end of the CPU oracle (hq_oracle.c) -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module, and only as the checker.  Arrays follow the reference's
conventions: forces are returned as (N, 3) float64 (the transpose of the
reference's (3, N) buffers), node and particle indices are 0-based.

Functions and the reference routine each restates:

* direct_forces            fortran/direct_forces.f90:8-64 (serial pair loop)
* direct_targets           the same pair force, gather form over a target subset
* Tree (insertion octree)  fortran/octree_default.f90 pre_init/init/add
                           (:178-580), incorporate_particles (:582-628),
                           evaluate_multipole (:744-882), bh_forces_over_p
                           (:895-1078), prepare_groups/group_worker
                           (:1722-2095), Dehnen (:1095-1712)
* Morton                   fortran/morton/morton_octree.f90 (:184-766)
* morton_keys              morton_key + choose_depth (:184-244)
* errors_ma                benchmarks/test_forces.f90:66-127
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_i64 = C.POINTER(C.c_int64)
_u64 = C.POINTER(C.c_uint64)
_u32 = C.POINTER(C.c_uint32)


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_direct_forces.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_double, C.c_double, _d]
        L.orc_direct_targets.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_int64, _i64,
                                         C.c_double, C.c_double, _d]
        L.orc_tree_new.restype = C.c_void_p
        L.orc_tree_new.argtypes = [C.c_int64, _d, C.c_int, C.c_int]
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_init.argtypes = [C.c_void_p, _d, _d, _d]
        L.orc_tree_nnodes.argtypes = [C.c_void_p]
        L.orc_tree_ngroups.argtypes = [C.c_void_p]
        L.orc_tree_export.argtypes = [C.c_void_p] + [_d] * 9 + [_i] * 4 + [_u64, _u32, _d, _d]
        L.orc_tree_bh_forces.argtypes = [C.c_void_p, C.c_int64, _i64, C.c_double, C.c_double,
                                         C.c_double, _d, _i64]
        L.orc_tree_forces_groups.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                             _d, _i64]
        L.orc_tree_groups_export.argtypes = [C.c_void_p, _i, _i, _i, _i, _i]
        L.orc_tree_dehnen.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _d, _i64]
        L.orc_morton_key.restype = C.c_int64
        L.orc_morton_key.argtypes = [_d, C.c_int]
        L.orc_morton_keys.argtypes = [C.c_int64, _d, _d, C.c_int, C.c_double, _i64]
        L.orc_argsort.argtypes = [C.c_int64, _i64, _i]
        L.orc_morton_build.restype = C.c_void_p
        L.orc_morton_build.argtypes = [C.c_int64, _d, _d, C.c_int, C.c_double, C.c_int]
        L.orc_morton_free.argtypes = [C.c_void_p]
        L.orc_morton_depth.argtypes = [C.c_void_p]
        L.orc_morton_nnodes.argtypes = [C.c_void_p]
        L.orc_morton_export.argtypes = [C.c_void_p, _i, _i64, _i64, _i, _i, _i, _i, _i, _i, _i,
                                        _d, _d, _d, _d]
        L.orc_morton_forces.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int,
                                        _d]
        L.orc_errors_ma.argtypes = [C.c_int64, _d, _d, _d, _d, _d]
        L.orc_integrate.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, _d, C.c_int, C.c_double,
                                    C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double]
        L.orc_energy.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, _d, C.c_double, C.c_double,
                                 _d]
        L.orc_parareal.argtypes = [C.c_int64, _d, _d, C.c_double, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                                   C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _d, _d, _i]
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def direct_forces(m, x, y, z, G=1.0, eps2=0.0):
    """Reference direct sum (direct_forces.f90:19-38); returns (N,3)."""
    m, x, y, z = map(_f64, (m, x, y, z))
    F = np.zeros((len(m), 3))
    lib().orc_direct_forces(len(m), _p(m), _p(x), _p(y), _p(z), G, eps2, _p(F))
    return F


def direct_targets(m, x, y, z, targets, G=1.0, eps2=0.0):
    """FP64 direct forces on the given target indices only; returns (nt,3)."""
    m, x, y, z = map(_f64, (m, x, y, z))
    t = np.ascontiguousarray(targets, dtype=np.int64)
    F = np.zeros((len(t), 3))
    lib().orc_direct_targets(len(m), _p(m), _p(x), _p(y), _p(z), len(t), _p(t, _i64), G, eps2,
                             _p(F))
    return F


class Tree:
    """The reference OctreeType: pre_init(m, method, group_size) + init(x,y,z)."""

    def __init__(self, m, method=12, group_size=64):
        self.m = _f64(m)
        self.N = len(self.m)
        self.method = method
        self._h = lib().orc_tree_new(self.N, _p(self.m), method, group_size)
        if not self._h:
            raise MemoryError("orc_tree_new failed")

    def init(self, x, y, z):
        self.x, self.y, self.z = map(_f64, (x, y, z))
        st = lib().orc_tree_init(self._h, _p(self.x), _p(self.y), _p(self.z))
        if st != 0:
            raise RuntimeError(f"oracle tree init failed (status {st})")
        return self

    @property
    def nnodes(self):
        return lib().orc_tree_nnodes(self._h)

    def export(self):
        """Node store as a dict of numpy arrays (reference node numbering)."""
        n = self.nnodes
        out = {k: np.zeros(n) for k in ("cx", "cy", "cz", "half", "L2", "mass", "qx", "qy", "qz")}
        for k in ("particle", "type", "depth"):
            out[k] = np.zeros(n, dtype=np.int32)
        out["child"] = np.zeros((n, 8), dtype=np.int32)
        out["khi"] = np.zeros(n, dtype=np.uint64)
        out["klo"] = np.zeros(n, dtype=np.uint32)
        out["quad"] = np.zeros((n, 6))
        out["oct"] = np.zeros((n, 14))
        lib().orc_tree_export(
            self._h, *[_p(out[k]) for k in ("cx", "cy", "cz", "half", "L2", "mass", "qx", "qy",
                                             "qz")],
            *[_p(out[k], _i) for k in ("particle", "type", "depth", "child")],
            _p(out["khi"], _u64), _p(out["klo"], _u32), _p(out["quad"]), _p(out["oct"]))
        return out

    def bh_forces(self, theta2, G=1.0, eps2=0.0, targets=None, counts=False):
        """bh_forces_over_p for all (or selected) targets; (nt,3) [+ (nt,2) counts]."""
        if targets is None:
            nt, tp = self.N, None
        else:
            t = np.ascontiguousarray(targets, dtype=np.int64)
            nt, tp = len(t), _p(t, _i64)
        F = np.zeros((nt, 3))
        c = np.zeros((nt, 2), dtype=np.int64) if counts else None
        lib().orc_tree_bh_forces(self._h, nt, tp, theta2, G, eps2, _p(F),
                                 _p(c, _i64) if counts else None)
        return (F, c) if counts else F

    def forces_groups(self, theta2, G=1.0, eps2=0.0, counts=False):
        F = np.zeros((self.N, 3))
        ng = lib().orc_tree_ngroups(self._h)
        c = np.zeros((ng, 2), dtype=np.int64)
        st = lib().orc_tree_forces_groups(self._h, theta2, G, eps2, _p(F), _p(c, _i64))
        if st != 0:
            raise RuntimeError("forces_groups needs a grouped method (30/31/32)")
        return (F, c) if counts else F

    def groups(self):
        ng = lib().orc_tree_ngroups(self._h)
        n = self.nnodes
        g = np.zeros(ng, dtype=np.int32)
        first = np.zeros(n, dtype=np.int32)
        last = np.zeros(n, dtype=np.int32)
        count = np.zeros(n, dtype=np.int32)
        order = np.zeros(self.N, dtype=np.int32)
        lib().orc_tree_groups_export(self._h, _p(g, _i), _p(first, _i), _p(last, _i),
                                     _p(count, _i), _p(order, _i))
        return dict(groups=g, first=first, last=last, count=count, order=order)

    def dehnen(self, theta2, eps2=0.0, G=1.0):
        F = np.zeros((self.N, 3))
        npairs = np.zeros(1, dtype=np.int64)
        st = lib().orc_tree_dehnen(self._h, theta2, eps2, G, _p(F), _p(npairs, _i64))
        if st != 0:
            raise RuntimeError(f"Dehnen needs method 20/21/22 (status {st})")
        return F

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_tree_free(self._h)
            self._h = None


def morton_key(q3, L):
    q = _f64(q3)
    return int(lib().orc_morton_key(_p(q), L))


def morton_keys(m, q, L=0, ssa=1.2):
    """Keys of morton_octree.f90 (normalisation :357-384 + morton_key) in the
    original particle order; q is (N,3).  Returns (keys int64, L)."""
    m = _f64(m)
    q = _f64(q)
    keys = np.zeros(len(m), dtype=np.int64)
    L = lib().orc_morton_keys(len(m), _p(m), _p(q), L, ssa, _p(keys, _i64))
    if L < 0:
        raise RuntimeError("keys not unique at 21 levels (64-bit key limit)")
    return keys, L


def argsort_stable(keys):
    k = np.ascontiguousarray(keys, dtype=np.int64)
    perm = np.zeros(len(k), dtype=np.int32)
    lib().orc_argsort(len(k), _p(k, _i64), _p(perm, _i))
    return perm


class Morton:
    """morton_tree_type (morton_octree.f90:73-180); q is (N,3)."""

    def __init__(self, m, q, L=0, ssa=1.2, use_quad=True):
        self.m = _f64(m)
        self.q = _f64(q)
        self.N = len(self.m)
        self._h = lib().orc_morton_build(self.N, _p(self.m), _p(self.q), L, ssa, int(use_quad))
        self.L = lib().orc_morton_depth(self._h)
        self.nnodes = lib().orc_morton_nnodes(self._h)
        self.use_quad = use_quad

    def export(self):
        n, nn = self.N, self.nnodes
        o = dict(perm=np.zeros(n, np.int32), keys=np.zeros(n, np.int64),
                 prefix=np.zeros(nn, np.int64), level=np.zeros(nn, np.int32),
                 first=np.zeros(nn, np.int32), last=np.zeros(nn, np.int32),
                 parent=np.zeros(nn, np.int32), nchild=np.zeros(nn, np.int32),
                 child=np.zeros((nn, 8), np.int32), next=np.zeros(nn, np.int32),
                 size2=np.zeros(nn), mass=np.zeros(nn), qcm=np.zeros((nn, 3)),
                 quad=np.zeros((nn, 6)))
        lib().orc_morton_export(self._h, _p(o["perm"], _i), _p(o["keys"], _i64),
                                _p(o["prefix"], _i64), *[_p(o[k], _i) for k in (
                                    "level", "first", "last", "parent", "nchild", "child",
                                    "next")],
                                _p(o["size2"]), _p(o["mass"]), _p(o["qcm"]), _p(o["quad"]))
        return o

    def forces(self, G=1.0, eps2=0.0, theta2=0.25, mult=None):
        F = np.zeros((self.N, 3))
        mult = self.use_quad if mult is None else mult
        lib().orc_morton_forces(self._h, G, eps2, theta2, int(mult), _p(F))
        return F

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_morton_free(self._h)
            self._h = None


def errors_ma(m, q, F, Fd):
    """McMillan-Aarseth normalised RMS + sum F + torque (test_forces.f90:66-127)."""
    m, q, F, Fd = map(_f64, (m, q, F, Fd))
    e = np.zeros(7)
    lib().orc_errors_ma(len(m), _p(m), _p(q), _p(F), _p(Fd), _p(e))
    return e


def rel_errors(F, Fd, m=None):
    """RMS and max of |a - a_dir| / |a_dir| per particle (accelerations or forces
    give the same ratio)."""
    F = np.asarray(F, dtype=np.float64)
    Fd = np.asarray(Fd, dtype=np.float64)
    num = np.linalg.norm(F - Fd, axis=1)
    den = np.linalg.norm(Fd, axis=1)
    r = num / np.where(den > 0, den, 1.0)
    return float(np.sqrt(np.mean(r * r))), float(np.max(r))


FORCE_TREE, FORCE_DIRECT = 0, 2
VERLET, EULER = 0, 1


def integrate(m, q, p, scheme, dt, nsteps, force=FORCE_DIRECT, method=11, theta2=0.25, G=1.0,
              eps2=0.0):
    """Verlet (scheme 0, integration.f90:153-171) or symplectic Euler (scheme 1,
    parareal.f90:30-36) from positions q and momenta p, both (N,3); returns the
    new (q, p)."""
    m = _f64(m)
    q = [_f64(np.asarray(q)[:, c]).copy() for c in range(3)]
    p = [_f64(np.asarray(p)[:, c]).copy() for c in range(3)]
    st = lib().orc_integrate(len(m), _p(m), *[_p(a) for a in q], *[_p(a) for a in p], scheme,
                             dt, int(nsteps), force, method, theta2, G, eps2)
    if st != 0:
        raise RuntimeError(f"oracle integration failed (status {st})")
    return np.stack(q, axis=1), np.stack(p, axis=1)


def energy(m, q, p, G=1.0, eps2=0.0):
    """[E, P(3), J(3)] of one system (integration.f90:236-262)."""
    m = _f64(m)
    q, p = _f64(q), _f64(p)
    out = np.zeros(7)
    cols = [np.ascontiguousarray(a[:, c]) for a in (q, p) for c in range(3)]
    lib().orc_energy(len(m), _p(m), *[_p(a) for a in cols], G, eps2, _p(out))
    return out


def parareal(m, q0, p0, tf, nw, itermax, maxerror, G, eps2, coarse=(FORCE_TREE, 11, 0.25, 1),
             fine=(FORCE_DIRECT, 11, 0.25, 16)):
    """integrate_parareal (parareal.f90:97-186); coarse/fine = (force, method,
    theta2, steps per window).  Returns (U (nw, 6, N), errors, iterations)."""
    m = _f64(m)
    n = len(m)
    u0 = np.concatenate([_f64(q0).T, _f64(p0).T]).reshape(-1).copy()
    U = np.zeros(nw * 6 * n)
    err = np.zeros(max(itermax - 1, 1))
    it = np.zeros(1, dtype=np.int32)
    st = lib().orc_parareal(n, _p(m), _p(u0), tf, nw, itermax, maxerror, G, eps2,
                            coarse[0], coarse[1], coarse[2], coarse[3],
                            fine[0], fine[1], fine[2], fine[3], _p(U), _p(err), _p(it, _i))
    if st != 0:
        raise RuntimeError(f"oracle parareal failed (status {st})")
    k = int(it[0])
    return U.reshape(nw, 6, n), err[:k - 1], k
