"""The force benchmark harness of the reference, on the GPU
(benchmarks/test_forces.f90:129-242 test_forces_time and :66-127
evaluate_error_on_accelerations).

    rows = test_forces_time([1000, 10000], [0.3, 0.5, 0.7], 1e-6, tests=2,
                            file=open("out.txt", "w"), method=12)

Each row is ``N theta time_pre_init time_init time_forces errors(1:7)`` as the
reference writes it (:223, :169 for the direct-sum row ``N -1 0 0 time_df 0..0``).
errors(1) is the McMillan-Aarseth normalised RMS acceleration error against
the FP64 direct sum, errors(2:4) the sum of the forces and errors(5:7) the
total torque.  Inputs are seeded Plummer spheres from hierarquicos_amd.ics
(the reference draws its own with gfortran's RNG).
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from . import ics
from .device import DeviceArray, direct_forces, sync
from .octree import OctreeType


def evaluate_error_on_accelerations(ms, qs, forces, forces_dir):
    """errors(7) of :66-127 for (3,N) qs, forces and forces_dir (host arrays or
    DeviceArrays of 3N in the (3,N) layout)."""
    ms = L.f64(ms)
    n = len(ms)
    qs = np.asarray(qs, dtype=np.float64)
    x, y, z = (np.ascontiguousarray(qs[c]) for c in range(3))
    out = np.zeros(7)
    if isinstance(forces, DeviceArray):
        F, Fd = forces, forces_dir
        dm, dx, dy, dz = (DeviceArray.from_numpy(a) for a in (ms, x, y, z))
        L.check(L.lib().hq_errors_ma(n, dm.ptr, dx.ptr, dy.ptr, dz.ptr, F.ptr, Fd.ptr,
                                     L.HQ_LAYOUT_3N, L.HQ_DEVICE, L.ptr(out)))
        return out
    F = np.ascontiguousarray(np.asarray(forces, dtype=np.float64).T)     # F[3p + c]
    Fd = np.ascontiguousarray(np.asarray(forces_dir, dtype=np.float64).T)
    L.check(L.lib().hq_errors_ma(n, L.ptr(ms), L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(F),
                                 L.ptr(Fd), L.HQ_LAYOUT_3N, L.HQ_HOST, L.ptr(out)))
    return out


def format_row(N, theta, t_pre, t_init, t_forces, errors):
    """One output row in the reference's list-directed layout."""
    vals = [theta, t_pre, t_init, t_forces, *errors]
    return f"{int(N):12d}" + "".join(f" {float(v):25.17E}" for v in vals)


def test_forces_time(N_list, thetas, eps2, tests, file=None, nt=None, method=12, eval_df=True,
                     nt_df=None, gs_par=64, profile="plummer", walk="tree", precision="fp64"):
    """test_forces_time (:129-242).  nt / nt_df (thread counts) have no GPU
    meaning and are accepted for the reference's signature.  walk="bunched"
    times hq_tree_forces_bunched at `precision` instead of the method's own
    walk.  Returns the rows as tuples; writes them to `file` when given."""
    rows = []

    def emit(row):
        rows.append(row)
        if file is not None:
            file.write(format_row(*row) + "\n")

    for N in N_list:
        for i_test in range(1, tests + 1):
            m, q, _ = ics.make(profile, N, seed=i_test)
            dm = DeviceArray.from_numpy(m)
            dq = [DeviceArray.from_numpy(np.ascontiguousarray(q[:, c])) for c in range(3)]
            Fd = DeviceArray((3 * N,))
            if eval_df:
                a = time.perf_counter()
                direct_forces(dm, *dq, 1.0, eps2, out=Fd)
                sync()
                emit((N, -1.0, 0.0, 0.0, time.perf_counter() - a, [0.0] * 7))
            if method == 0:
                continue
            a = time.perf_counter()
            tree = OctreeType().pre_init(dm, method, gs_par, precision=precision)
            sync()
            t_pre = time.perf_counter() - a
            a = time.perf_counter()
            tree.init(*dq)
            sync()
            t_init = time.perf_counter() - a
            F = DeviceArray((3 * N,))
            for theta in thetas:
                th2 = theta ** 2
                a = time.perf_counter()
                if walk == "bunched":
                    tree.forces_bunched(th2, 1.0, eps2, precision, out=F)
                elif method < 20:
                    tree.forces_all(th2, 1.0, eps2, out=F)
                elif method < 30:
                    tree.dehnen_eval(th2, eps2, 1.0, out=F)
                else:
                    tree.forces_groups(th2, 1.0, eps2, out=F)
                sync()
                t_forces = time.perf_counter() - a
                err = (evaluate_error_on_accelerations(m, q.T, F, Fd) if eval_df
                       else np.zeros(7))
                emit((N, theta, t_pre, t_init, t_forces, list(err)))
            tree.close()
    return rows
