"""CPU checks of the HIP library: it is built, loads, and exports every entry
point include/hq.h declares (no compute calls: there is no GPU here)."""
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(REPO, "include", "hq.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(hq_\w+)\s*\(", src, re.M)))


def test_header_declares_the_api():
    names = declared()
    for n in ("hq_compute_forces", "hq_compute_forces_direct", "hq_generate_tree",
              "hq_tree_create", "hq_tree_build", "hq_tree_forces", "hq_morton_keys"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from hierarquicos_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libhq.so not built (python -m hierarquicos_amd.build)")
    L = _lib.lib()
    for n in declared():
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) == set(declared())
    assert L.hq_version().startswith(b"hierarquicos_amd")


def test_empty_inputs_are_refused_before_device_work():
    """N = 0 (and other empty sizes) at every entry point that takes a size:
    HQ_ERR_ARGS, decided from the arguments alone (no device is touched)."""
    import ctypes as C
    from hierarquicos_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libhq.so not built (python -m hierarquicos_amd.build)")
    L = _lib.lib()
    one = np.zeros(1)
    p = one.ctypes.data_as(C.POINTER(C.c_double))
    vp = C.c_void_p(one.ctypes.data)
    h = C.c_void_p()
    Lo = C.c_int()
    calls = {
        "hq_compute_forces_direct": lambda: L.hq_compute_forces_direct(0, p, p, p, p, 1.0, 0.0, p),
        "hq_compute_forces": lambda: L.hq_compute_forces(0, p, p, p, p, 0.5, 1.0, 0.0, 11, 0, p),
        "hq_generate_tree": lambda: L.hq_generate_tree(0, p, p, p, p, 11),
        "hq_tree_create": lambda: L.hq_tree_create(C.byref(h), 0, 11, 64, 0),
        "hq_tree_create_batch": lambda: L.hq_tree_create_batch(
            C.byref(h), 0, np.zeros(1, np.int64).ctypes.data_as(C.POINTER(C.c_int64)), 11, 64, 0),
        "hq_errors_ma": lambda: L.hq_errors_ma(0, p, p, p, p, p, p, 0, 0, p),
        "hq_direct_forces_device": lambda: L.hq_direct_forces_device(0, vp, vp, vp, vp, 0, None,
                                                                     1.0, 0.0, 0, vp),
        "hq_morton_keys": lambda: L.hq_morton_keys(0, vp, vp, 0, 1.2, vp, vp, C.byref(Lo), 0),
        "hq_morton_keys_noalloc": lambda: L.hq_morton_keys_noalloc(0, vp, vp, 0, 1.2, vp, vp,
                                                                   C.byref(Lo)),
        "hq_morton_create": lambda: L.hq_morton_create(C.byref(h), 0, 0, 1.2, 1),
    }
    for name, call in calls.items():
        assert call() == _lib.HQ_ERR_ARGS, name
    # a batch with an empty system
    counts = np.array([5, 0, 5], np.int64)
    assert L.hq_tree_create_batch(C.byref(h), 3, counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                  11, 64, 0) == _lib.HQ_ERR_ARGS
