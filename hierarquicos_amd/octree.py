"""Host-side mirror of the reference's OctreeType (fortran/octree_default.f90).

    tree = OctreeType()
    tree.pre_init(m, method=12)              # pre_init (:178-282)
    tree.init(x, y, z)                       # init (:320-387): GPU build
    F = tree.forces_all(theta2, G, eps2)     # bh_forces_over_p for every p (:895)
    F = tree.forces_groups(theta2, G, eps2)  # forces_groups (:1811)
    F = tree.dehnen_eval(theta2, eps2, G)    # dehnen_eval (:1680)

Forces come back as (3, N) float64 arrays, the reference's ``forces(:,p)``
layout, or are written into a caller-provided device buffer.  Arrays may be
numpy (host) or DeviceArrays (HBM-resident; no host round trip).  All work
runs in libhq.so on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .device import DeviceArray

METHODS = (10, 11, 12, 20, 21, 22, 30, 31, 32)


def _is_device(a):
    return isinstance(a, DeviceArray)


def _addr(a):
    """(pointer, where) of a numpy array or a DeviceArray of float64."""
    if _is_device(a):
        if a.dtype != np.float64:
            raise TypeError("device arrays must be float64")
        return a.ptr, L.HQ_DEVICE
    return C.c_void_p(a.ctypes.data), L.HQ_HOST


class OctreeType:
    """The OctreeType handle: one GPU tree, one HIP stream."""

    def __init__(self):
        self._h = C.c_void_p()
        self.N = 0
        self.method = None
        self.is_allocated = False

    # -- pre_init / init ------------------------------------------------------
    def pre_init(self, m, method=10, group_size=64, precision="fp64", counts=None):
        """counts: sizes of the independent systems of a batch (consecutive
        runs of m); each system gets its own tree, as if pre_init had been
        called on it alone.  None is one system."""
        if method not in METHODS:
            raise ValueError(f"method not identified! ({method})")
        prec = {"fp64": L.HQ_FP64, "mixed": L.HQ_MIXED}[precision]
        self.close()
        self._m = m if _is_device(m) else L.f64(m)
        self.N = int(self._m.shape[0])
        self.method = method
        self.group_size = group_size
        lib = L.lib()
        if counts is None:
            self.counts = None
            L.check(lib.hq_tree_create(C.byref(self._h), self.N, method, group_size, prec))
        else:
            cnt = np.ascontiguousarray(counts, dtype=np.int64)
            if cnt.ndim != 1 or int(cnt.sum()) != self.N:
                raise ValueError("counts must sum to len(m)")
            self.counts = cnt
            L.check(lib.hq_tree_create_batch(C.byref(self._h), len(cnt),
                                              cnt.ctypes.data_as(C.POINTER(C.c_int64)),
                                              method, group_size, prec))
        p, where = _addr(self._m)
        L.check(lib.hq_tree_set_masses(self._h, p, where))
        self.is_allocated = True
        return self

    def init(self, x, y, z):
        if not self.is_allocated:
            raise RuntimeError("Not allocated! Run pre_init first!")
        xs = [a if _is_device(a) else L.f64(a) for a in (x, y, z)]
        if any(int(a.shape[0]) != self.N for a in xs):
            raise ValueError("x, y, z must have the length of m")
        (px, w), (py, _), (pz, _) = (_addr(a) for a in xs)
        L.check(L.lib().hq_tree_build(self._h, px, py, pz, w))
        self._cache = None
        return self

    # -- forces ---------------------------------------------------------------
    def _run(self, fn, out, layout, *args):
        if out is None:
            buf = np.zeros(3 * self.N)
            L.check(fn(self._h, *args, C.c_void_p(buf.ctypes.data), layout, L.HQ_HOST))
            if layout == L.HQ_LAYOUT_3N:       # F[3*p + c] -> (3, N) view
                return buf.reshape(self.N, 3).T
            return buf.reshape(3, self.N).T    # F[c*N + p] -> (N, 3) Fortran order
        p, where = _addr(out)
        L.check(fn(self._h, *args, p, layout, where))
        return out

    def forces_all(self, theta2, G=1.0, eps2=0.0, out=None):
        """bh_forces_over_p (:895-1078) for p = 1..N, as the harness loop does
        (benchmarks/test_forces.f90:199-213); (3, N)."""
        return self._run(L.lib().hq_tree_forces, out, L.HQ_LAYOUT_3N, theta2, G, eps2)

    def forces(self, p, theta2=0.0, G=1.0, eps2=0.0):
        """Force on particle p (0-based), the reference's per-particle call.  The
        GPU evaluates all targets at once; repeated calls with the same
        parameters reuse that result."""
        key = (theta2, G, eps2)
        if getattr(self, "_cache", None) is None or self._cache[0] != key:
            self._cache = (key, self.forces_all(theta2, G, eps2))
        return self._cache[1][:, p].copy()

    def forces_groups(self, theta2, G=1.0, eps2=0.0, out=None):
        if self.method not in (30, 31, 32):
            raise ValueError("forces_groups needs method 30, 31 or 32")
        return self._run(L.lib().hq_tree_forces, out, L.HQ_LAYOUT_3N, theta2, G, eps2)

    def dehnen_eval(self, theta2, eps2=0.0, G=1.0, out=None):
        if self.method not in (20, 21, 22):
            raise ValueError("dehnen_eval needs method 20, 21 or 22")
        return self._run(L.lib().hq_tree_forces, out, L.HQ_LAYOUT_3N, theta2, G, eps2)

    def forces_bunched(self, theta2, G=1.0, eps2=0.0, precision="mixed", out=None):
        """Barnes (1990) shared-list walk over 64-particle bunches in tree order
        (FP64 or mixed precision), by default with the offset acceptance test
        (tune(mac=...)); (3, N)."""
        prec = {"fp64": L.HQ_FP64, "mixed": L.HQ_MIXED}[precision]
        lib = L.lib()

        def fn(h, th2, g, e2, p, layout, where):
            return lib.hq_tree_forces_bunched(h, th2, g, e2, p, layout, where, prec)
        return self._run(fn, out, L.HQ_LAYOUT_3N, theta2, G, eps2)

    def fmm(self, theta2, G=1.0, eps2=0.0, order=4, mode="mfma", out=None):
        """Cartesian FMM of order 4 or 6 on the built tree; (3, N).  mode
        "mfma": same-level M2L as FP32 matrix-core products and an FP32 near
        field; "fp64": everything in scalar FP64."""
        md = {"mfma": L.HQ_FMM_MFMA, "fp64": L.HQ_FMM_FP64}[mode]
        lib = L.lib()

        def fn(h, th2, g, e2, p, layout, where):
            return lib.hq_tree_fmm(h, int(order), th2, g, e2, md, p, layout, where)
        return self._run(fn, out, L.HQ_LAYOUT_3N, theta2, G, eps2)

    def tune(self, bunch=None, super_size=None, fused=None, leaf=None, fmm_leaf=None,
             graph=None, split=None, far=None, far_inner=None, graph_grid=None, mac=None,
             fmm_bin_window=None):
        """Bunch size (<= 64), super-group size, fused/two-kernel mode and leaf
        bucket (cells of <= leaf bodies enter as their bodies) of the bunched
        walk; FMM leaf size; graph: whether integrate() replays one captured
        step (default True); split: 0 never, 1 (default) with few bunches, 2
        always -- the fused walk's batches evaluated one wavefront each (same
        results); far: the mixed walk's far-field factor in bunch radii (default
        5.0; 0 off): farther accepted cells enter the bunch's local expansion;
        far_inner (default 3.0; 0 off): cells between far_inner and far radii
        enter an expansion of higher order; graph_grid: the replayed walk's grid
        in bunches (0, default: twice the first step's bunch count plus 64);
        mac: the bunched walks' acceptance test, "side" (the grouped test of
        octree_default.f90:1927-1937) or "offset" (side + theta |com - centre|);
        fmm_bin_window: target groups whose lattice-pair counts a tile keeps in
        LDS (1..4096, default 4096; fewer send more pairs through global
        atomics, which changes only the order of the FMM's sums)."""
        lib = L.lib()
        if mac is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_MAC, {"side": 0, "offset": 1}[mac]))
        if graph_grid is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_GRAPH_GRID, int(graph_grid)))
        if far is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_FAR, int(round(10 * float(far)))))
        if far_inner is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_FAR_INNER,
                                     int(round(10 * float(far_inner)))))
        if split is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_SPLIT, int(split)))
        if graph is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_GRAPH, int(bool(graph))))
        if fmm_leaf is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_FMM_LEAF, int(fmm_leaf)))
        if fmm_bin_window is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_FMM_BIN_WINDOW, int(fmm_bin_window)))
        if leaf is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_LEAF, int(leaf)))
        if fused is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_FUSED, int(bool(fused))))
        if bunch is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_BUNCH, int(bunch)))
        if super_size is not None:
            L.check(lib.hq_tree_tune(self._h, L.HQ_TUNE_SUPER, int(super_size)))
        return self

    # -- time integration (state held in HBM) ----------------------------------
    def set_state(self, q, p):
        """Positions and momenta, each (N, 3) numpy or a tuple of three float64
        arrays (numpy or DeviceArray)."""
        cols = []
        for a in (q, p):
            if isinstance(a, (tuple, list)):
                cols += [c if _is_device(c) else L.f64(c) for c in a]
            else:
                a = np.asarray(a, dtype=np.float64)
                cols += [np.ascontiguousarray(a[:, c]) for c in range(3)]
        if any(int(c.shape[0]) != self.N for c in cols):
            raise ValueError("state arrays must have the length of m")
        addrs = [_addr(c) for c in cols]
        if len({w for _, w in addrs}) != 1:
            raise ValueError("state arrays must all be host or all device")
        self._state_keep = cols
        L.check(L.lib().hq_tree_set_state(self._h, *[a for a, _ in addrs], addrs[0][1]))
        self._cache = None
        return self

    def get_state(self):
        """(q, p) as (N, 3) numpy arrays."""
        cols = [np.zeros(self.N) for _ in range(6)]
        L.check(L.lib().hq_tree_get_state(self._h, *[C.c_void_p(c.ctypes.data) for c in cols],
                                          L.HQ_HOST))
        return np.stack(cols[:3], axis=1), np.stack(cols[3:], axis=1)

    def integrate(self, dt, nsteps, theta2=0.0, G=1.0, eps2=0.0, scheme="verlet", force="tree"):
        """nsteps steps of dt on the state (rebuilding before every force
        evaluation).  scheme: "verlet" | "euler"; force: "tree" (the handle's
        method), "bunched" (bunched walk at the handle's precision) or
        "direct" (FP64 direct sum within each system)."""
        sch = {"verlet": L.HQ_SCHEME_VERLET, "euler": L.HQ_SCHEME_EULER}[scheme]
        frc = {"tree": L.HQ_FORCE_TREE, "bunched": L.HQ_FORCE_BUNCHED,
               "direct": L.HQ_FORCE_DIRECT}[force]
        L.check(L.lib().hq_tree_integrate(self._h, sch, dt, int(nsteps), theta2, G, eps2, frc))
        self._cache = None
        return self

    def energy(self, G=1.0, eps2=0.0):
        """Per system [E, P(3), J(3)] of the current state; (nsys, 7)."""
        ns = 1 if getattr(self, "counts", None) is None else len(self.counts)
        out = np.zeros((ns, 7))
        L.check(L.lib().hq_tree_energy(self._h, G, eps2, L.ptr(out)))
        return out

    def autotune(self, theta2, G=1.0, eps2=0.0, precision="mixed", reps=2):
        """Searches the bunched walk's bunch size, super-group size, leaf
        bucket and evaluation tiling on the built tree (HIP-event timing, an
        accuracy guard on a sample), keeps the fastest setting and stores it
        in the process's tuning table; returns it with its time in ms."""
        prec = {"fp64": L.HQ_FP64, "mixed": L.HQ_MIXED}[precision]
        b, s_, lf = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        L.check(L.lib().hq_tree_autotune(self._h, theta2, G, eps2, prec, reps, C.byref(b),
                                         C.byref(s_), C.byref(lf), C.byref(ms)))
        return {"bunch": b.value, "super_size": s_.value, "leaf": lf.value,
                "split": self.get_tune("split"), "ms": ms.value}

    _TUNE_KEYS = {"bunch": 1, "super_size": 2, "fused": 3, "leaf": 4, "fmm_leaf": 5, "graph": 6,
                  "split": 7, "far": 8, "far_inner": 9, "graph_grid": 10, "mac": 11,
                  "fmm_bin_window": 12}

    def get_tune(self, key):
        """Current value of a tuning knob (the names of tune(); far and
        far_inner in tenths of a bunch radius, as the library stores them)."""
        v = C.c_int64()
        L.check(L.lib().hq_tree_get_tune(self._h, self._TUNE_KEYS[key], C.byref(v)))
        return v.value

    # -- introspection ----------------------------------------------------------
    def info(self):
        inf = L.TreeInfo()
        L.check(L.lib().hq_tree_info(self._h, C.byref(inf)))
        return inf.as_dict()

    @property
    def number_of_nodes(self):
        return self.info()["nnodes"]

    def export(self):
        """Node store in level order (see include/hq.h hq_tree_export)."""
        nn = self.number_of_nodes
        out = dict(depth=np.zeros(nn, np.int32), type=np.zeros(nn, np.int32),
                   particle=np.zeros(nn, np.int32), child_start=np.zeros(nn + 1, np.int32),
                   key_hi=np.zeros(nn, np.uint64), key_lo=np.zeros(nn, np.uint32),
                   mass=np.zeros(nn), com=np.zeros((nn, 3)), quad=np.zeros((nn, 6)),
                   oct=np.zeros((nn, 14)), center=np.zeros((nn, 3)))
        L.check(L.lib().hq_tree_export(
            self._h, L.ptr(out["depth"], L._i), L.ptr(out["type"], L._i),
            L.ptr(out["particle"], L._i), L.ptr(out["child_start"], L._i),
            L.ptr(out["key_hi"], L._u64), L.ptr(out["key_lo"], L._u32), L.ptr(out["mass"]),
            L.ptr(out["com"]), L.ptr(out["quad"]), L.ptr(out["oct"]), L.ptr(out["center"])))
        return out

    def order(self):
        o = np.zeros(self.N, np.int32)
        L.check(L.lib().hq_tree_order(self._h, L.ptr(o, L._i)))
        return o

    def close(self):
        if self._h:
            L.lib().hq_tree_destroy(self._h)
            self._h = C.c_void_p()
        self.is_allocated = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_table_save(path):
    """Writes the process's tuning table (autotuned bunched-walk settings per
    workload class) to a text file."""
    L.check(L.lib().hq_tune_table_save(str(path).encode()))


def tune_table_load(path):
    """Merges a saved tuning table into the process's table."""
    L.check(L.lib().hq_tune_table_load(str(path).encode()))


def tune_table_clear():
    L.check(L.lib().hq_tune_table_clear())
