"""CPU checks of the HIP library: it is built, loads, and exports every entry
point include/hq.h declares (no compute calls: there is no GPU here)."""
import os
import re

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
