"""The reference's f2py module surface, ``api.api_mod`` (fortran/api3d.f90),
backed by libhq.so on the GPU.

    from hierarquicos_amd import api
    F  = api.api_mod.compute_forces(m, x, y, z, theta, G, eps, True)   # (N,3)
    F0 = api.api_mod.compute_forces_direct(m, x, y, z, G, eps)         # (N,3)
    api.api_mod.generate_tree(m, x, y, z, True)

As in the f2py wrapper, inputs are converted to contiguous float64 and the
result is a new (N, 3) Fortran-ordered float64 array owned by the caller.
``quad`` selects the quadrupole (method 11) over the monopole (method 10);
``method`` selects any of 10/11/12, 20/21/22, 30/31/32 directly.  ``theta``
is the opening angle and ``eps`` the softening length; the shipped f2py call
passed theta unsquared to a routine expecting theta^2 (api3d.f90:37), which
``legacy_theta2=True`` reproduces.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L


def _prep(m, x, y, z):
    m, x, y, z = (L.f64(a) for a in (m, x, y, z))
    n = len(m)
    if not (len(x) == len(y) == len(z) == n):
        raise ValueError("m, x, y, z must have the same length")
    return n, m, x, y, z


class api_mod:  # noqa: N801  (the f2py module name)
    @staticmethod
    def generate_tree(m, x, y, z, quad=False, method=None):
        """api3d.f90:9-18: build (and discard) the tree."""
        n, m, x, y, z = _prep(m, x, y, z)
        method = method or (11 if quad else 10)
        L.check(L.lib().hq_generate_tree(n, L.ptr(m), L.ptr(x), L.ptr(y), L.ptr(z), method))

    @staticmethod
    def compute_forces(m, x, y, z, theta, G, eps, quad=False, method=None, legacy_theta2=False):
        """api3d.f90:20-41: Barnes-Hut forces, (N, 3)."""
        n, m, x, y, z = _prep(m, x, y, z)
        method = method or (11 if quad else 10)
        F = np.zeros((n, 3), order="F")
        L.check(L.lib().hq_compute_forces(n, L.ptr(m), L.ptr(x), L.ptr(y), L.ptr(z),
                                          float(theta), float(G), float(eps), int(method),
                                          int(bool(legacy_theta2)), L.ptr(F)))
        return F

    @staticmethod
    def compute_forces_direct(m, x, y, z, G, eps):
        """api3d.f90:43-72: all-pairs forces, (N, 3)."""
        n, m, x, y, z = _prep(m, x, y, z)
        F = np.zeros((n, 3), order="F")
        L.check(L.lib().hq_compute_forces_direct(n, L.ptr(m), L.ptr(x), L.ptr(y), L.ptr(z),
                                                 float(G), float(eps), L.ptr(F)))
        return F
