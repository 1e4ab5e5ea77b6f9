"""CPU checks of the Fortran boundary (hierarquicos_amd/fortran): the
ISO_C_BINDING module declares every entry point of include/hq.h, its BIND(C)
structs have the C layout, and the Fortran libraries load and export their
symbols.  No compute calls: there is no GPU here."""
import ctypes as C
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(REPO, "hierarquicos_amd", "fortran")
FDRIVER = os.path.join(REPO, "tests", "fortran", "libfdriver.so")


def _declared():
    src = open(os.path.join(REPO, "include", "hq.h")).read()
    return set(re.findall(r"^(?:int|const char \*)\s*(hq_\w+)\s*\(", src, re.M))


def test_iface_binds_every_entry_point():
    src = open(os.path.join(FDIR, "hq_iface.f90")).read()
    bound = set(re.findall(r'BIND\(C,\s*name="(hq_\w+)"\)', src))
    assert _declared() <= bound, sorted(_declared() - bound)


def test_mirror_modules_name_the_reference_procedures():
    src = open(os.path.join(FDIR, "hq_octree.f90")).read().lower()
    for name in ("pre_init", "init", "bh_forces_over_p", "dehnen_eval", "forces_groups",
                 "generate_tree", "compute_forces", "compute_forces_direct",
                 "test_integration_verlet_direct", "test_integration_verlet_bh",
                 "test_integration_verlet_dehnen", "total_energy",
                 "evaluate_error_on_accelerations", "integrate_parareal", "g_til", "f_til",
                 "total_linear_momentum", "total_angular_momentum"):
        assert re.search(r"(subroutine|function)\s+" + name + r"\s*\(", src), name


def _driver():
    if not os.path.exists(FDRIVER):
        pytest.skip("tests/fortran/libfdriver.so not built (amdflang; __graft_entry__.build)")
    return C.CDLL(FDRIVER)


def test_fortran_driver_loads_and_exports():
    import torch  # noqa: F401  (the process the GPU tests run in has torch loaded)
    lib = _driver()
    for n in ("fd_reset", "fd_status", "fd_layout", "fd_api", "fd_octree", "fd_bunched_fmm",
              "fd_direct", "fd_errors", "fd_verlet", "fd_propagators", "fd_parareal"):
        assert hasattr(lib, n), n
    lib.fd_reset()
    assert lib.fd_status() == 0
    hqf = C.CDLL(os.path.join(FDIR, "lib", "libhqf.so"))
    assert hqf is not None


def test_fortran_struct_layout_matches_c(tmp_path):
    lib = _driver()
    out = (C.c_int64 * 6)()
    lib.fd_layout(out)
    probe = tmp_path / "layout.c"
    probe.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "hq.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(hq_parareal_params),\n'
        "         offsetof(hq_parareal_params, fine_theta2),\n"
        "         offsetof(hq_parareal_params, coarse_scheme), sizeof(hq_tree_info_t),\n"
        "         offsetof(hq_tree_info_t, ms_build), offsetof(hq_tree_info_t, max_stack));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(probe), "-o", str(exe)],
                   check=True)
    c_layout = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                               text=True).stdout.split()]
    assert list(out) == c_layout
