"""ctypes binding of libhq.so (the C ABI declared in include/hq.h).

The library is built in-tree by ``python -m hierarquicos_amd.build``.  There
is no CPU fallback: when libhq.so is missing or no GPU is visible, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhq.so")

HQ_OK = 0
HQ_ERR_ARGS, HQ_ERR_CAPACITY, HQ_ERR_LOST, HQ_ERR_HIP = 1, 2, 3, 4
HQ_ERR_NONFINITE, HQ_ERR_DEPTH, HQ_ERR_NOMEM, HQ_ERR_OVERFLOW = 5, 6, 7, 8
HQ_HOST, HQ_DEVICE = 0, 1
HQ_FP64, HQ_MIXED = 0, 1
HQ_LAYOUT_3N, HQ_LAYOUT_N3 = 0, 1
HQ_TUNE_BUNCH, HQ_TUNE_SUPER, HQ_TUNE_FUSED, HQ_TUNE_LEAF, HQ_TUNE_FMM_LEAF = 1, 2, 3, 4, 5
HQ_TUNE_GRAPH, HQ_TUNE_SPLIT, HQ_TUNE_FAR, HQ_TUNE_FAR_INNER = 6, 7, 8, 9
HQ_TUNE_GRAPH_GRID, HQ_TUNE_MAC, HQ_TUNE_FMM_BIN_WINDOW = 10, 11, 12
HQ_FMM_MFMA, HQ_FMM_FP64 = 0, 1
HQ_SCHEME_VERLET, HQ_SCHEME_EULER = 0, 1
HQ_FORCE_TREE, HQ_FORCE_BUNCHED, HQ_FORCE_DIRECT = 0, 1, 2
HQ_CORR_REFERENCE, HQ_CORR_EXACT = 0, 1

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_i64 = C.POINTER(C.c_int64)
_u64 = C.POINTER(C.c_uint64)
_u32 = C.POINTER(C.c_uint32)
_vp = C.c_void_p


class PararealParams(C.Structure):
    """hq_parareal_params of include/hq.h."""
    _fields_ = [("n", C.c_int64), ("num_windows", C.c_int), ("itermax", C.c_int),
                ("tf", C.c_double), ("maxerror", C.c_double), ("G", C.c_double),
                ("eps2", C.c_double),
                ("coarse_force", C.c_int), ("coarse_method", C.c_int),
                ("coarse_precision", C.c_int), ("coarse_steps", C.c_int),
                ("coarse_theta2", C.c_double),
                ("fine_force", C.c_int), ("fine_method", C.c_int), ("fine_precision", C.c_int),
                ("fine_steps", C.c_int), ("fine_theta2", C.c_double),
                ("group_size", C.c_int), ("correction", C.c_int), ("coarse_scheme", C.c_int)]

EXPORTS = (
    "hq_compute_forces_direct", "hq_compute_forces", "hq_generate_tree", "hq_tree_create",
    "hq_tree_create_batch", "hq_tree_set_state", "hq_tree_get_state", "hq_tree_integrate",
    "hq_tree_energy", "hq_parareal", "hq_errors_ma", "hq_tree_autotune", "hq_tree_get_tune", "hq_tune_table_save",
    "hq_tune_table_load", "hq_tune_table_clear", "hq_tree_fmm",
    "hq_tree_set_masses", "hq_tree_build", "hq_tree_forces", "hq_tree_forces_bunched",
    "hq_tree_destroy", "hq_tree_tune", "hq_tree_info", "hq_tree_export", "hq_tree_order",
    "hq_direct_forces_device", "hq_morton_keys", "hq_morton_keys_noalloc", "hq_morton_create", "hq_morton_build",
    "hq_morton_forces", "hq_morton_info", "hq_morton_export", "hq_morton_destroy", "hq_dev_alloc", "hq_dev_free", "hq_dev_copy",
    "hq_dev_sync", "hq_last_error", "hq_version")


class TreeInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnodes", C.c_int64), ("ngroups", C.c_int64),
                ("lost", C.c_int64), ("most_depth", C.c_int), ("multipole", C.c_int),
                ("list_cells", C.c_int64), ("list_parts", C.c_int64), ("n_p2p", C.c_int64),
                ("n_m2p", C.c_int64), ("ms_build", C.c_double), ("ms_keys", C.c_double),
                ("ms_sort", C.c_double), ("ms_nodes", C.c_double), ("ms_upward", C.c_double),
                ("ms_walk", C.c_double), ("ms_lists", C.c_double), ("ms_eval", C.c_double),
                ("exec_cells", C.c_int64), ("exec_parts", C.c_int64), ("max_cells", C.c_int64),
                ("max_parts", C.c_int64), ("max_stack", C.c_int64),
                ("split_walks", C.c_int64), ("far_cells", C.c_int64),
                ("walk_stack", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HQError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[hq status {code}] {msg}")
        self.code = code


_LIB = None


def lib():
    """Load libhq.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hierarquicos_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.hq_last_error.restype = C.c_char_p
    L.hq_version.restype = C.c_char_p
    L.hq_compute_forces_direct.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_double, C.c_double, _d]
    L.hq_compute_forces.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_double, C.c_double,
                                    C.c_double, C.c_int, C.c_int, _d]
    L.hq_generate_tree.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_int]
    L.hq_tree_create.argtypes = [C.POINTER(_vp), C.c_int64, C.c_int, C.c_int, C.c_int]
    L.hq_tree_create_batch.argtypes = [C.POINTER(_vp), C.c_int, _i64, C.c_int, C.c_int, C.c_int]
    L.hq_tree_set_state.argtypes = [_vp] + [_vp] * 6 + [C.c_int]
    L.hq_tree_get_state.argtypes = [_vp] + [_vp] * 6 + [C.c_int]
    L.hq_tree_integrate.argtypes = [_vp, C.c_int, C.c_double, C.c_int64, C.c_double, C.c_double,
                                    C.c_double, C.c_int]
    L.hq_tree_energy.argtypes = [_vp, C.c_double, C.c_double, _d]
    L.hq_tree_get_tune.argtypes = [_vp, C.c_int, _i64]
    L.hq_tune_table_save.argtypes = [C.c_char_p]
    L.hq_tune_table_load.argtypes = [C.c_char_p]
    L.hq_tune_table_clear.argtypes = []
    L.hq_tree_autotune.argtypes = [_vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                   _i, _i, _i, _d]
    L.hq_tree_fmm.argtypes = [_vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _vp,
                              C.c_int, C.c_int]
    L.hq_errors_ma.argtypes = [C.c_int64] + [_vp] * 6 + [C.c_int, C.c_int, _d]
    L.hq_parareal.argtypes = [C.POINTER(PararealParams)] + [_vp] * 7 + [C.c_int, _vp, C.c_int,
                                                                        _d, _i]
    L.hq_tree_set_masses.argtypes = [_vp, _vp, C.c_int]
    L.hq_tree_build.argtypes = [_vp, _vp, _vp, _vp, C.c_int]
    L.hq_tree_forces.argtypes = [_vp, C.c_double, C.c_double, C.c_double, _vp, C.c_int, C.c_int]
    L.hq_tree_forces_bunched.argtypes = [_vp, C.c_double, C.c_double, C.c_double, _vp, C.c_int,
                                         C.c_int, C.c_int]
    L.hq_tree_destroy.argtypes = [_vp]
    L.hq_tree_tune.argtypes = [_vp, C.c_int, C.c_int64]
    L.hq_tree_info.argtypes = [_vp, C.POINTER(TreeInfo)]
    L.hq_tree_export.argtypes = [_vp, _i, _i, _i, _i, _u64, _u32, _d, _d, _d, _d, _d]
    L.hq_tree_order.argtypes = [_vp, _i]
    L.hq_direct_forces_device.argtypes = [C.c_int64, _vp, _vp, _vp, _vp, C.c_int64, _vp,
                                          C.c_double, C.c_double, C.c_int, _vp]
    L.hq_morton_keys.argtypes = [C.c_int64, _vp, _vp, C.c_int, C.c_double, _vp, _vp, _i, C.c_int]
    L.hq_morton_keys_noalloc.argtypes = [C.c_int64, _vp, _vp, C.c_int, C.c_double, _vp, _vp, _i]
    L.hq_morton_create.argtypes = [C.POINTER(_vp), C.c_int64, C.c_int, C.c_double, C.c_int]
    L.hq_morton_build.argtypes = [_vp, _vp, _vp, C.c_int]
    L.hq_morton_forces.argtypes = [_vp, C.c_double, C.c_double, C.c_double, C.c_int, _vp, C.c_int,
                                   C.c_int]
    L.hq_morton_info.argtypes = [_vp, _i, _i64]
    L.hq_morton_export.argtypes = [_vp, _i, _i64, _i, _i, _i, _i, _i, _i, _i, _d, _d, _d]
    L.hq_morton_destroy.argtypes = [_vp]
    L.hq_dev_alloc.argtypes = [C.POINTER(_vp), C.c_int64]
    L.hq_dev_free.argtypes = [_vp]
    L.hq_dev_copy.argtypes = [_vp, _vp, C.c_int64, C.c_int]
    L.hq_dev_sync.argtypes = []
    _LIB = L
    return L


def check(code):
    if code != HQ_OK:
        raise HQError(code, lib().hq_last_error().decode(errors="replace"))


def ptr(a, t=_d):
    """ctypes pointer of a contiguous numpy array."""
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
