"""Drivers around the force path, on the GPU: velocity Verlet
(benchmarks/integration_verlet/integration.f90) and Parareal
(parareal/parareal.f90, parareal/api3d.f90).

State stays in HBM for the whole integration; the tree is rebuilt before every
force evaluation, as the reference drivers call ``tree%init`` each step.

    qs, ps = verlet_bh(ms, qs, ps, G, eps2, dt, tf, theta2, mult)      # (3, N) each
    q1, p1, info = integrate_parareal(m, q0, p0, tf, num_windows, N_itermax,
                                      theta, G, eps, maxerror, qntt)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .octree import OctreeType


def steps_until(tf, dt):
    """Step count of ``DO WHILE (t < tf); ...; t = t + ABS(dt)`` (integration.f90:150-171)."""
    t, n = 0.0, 0
    while t < tf:
        t = t + abs(dt)
        n += 1
    return n


def _method(mult, base):
    return base + {1: 0, 4: 1, 8: 2}[mult]


def _run(ms, qs, ps, method, force, G, eps2, dt, tf, theta2, precision="fp64", scheme="verlet"):
    ms = L.f64(ms)
    t = OctreeType().pre_init(ms, method, precision=precision)
    t.set_state(np.asarray(qs, dtype=np.float64).T, np.asarray(ps, dtype=np.float64).T)
    t.integrate(dt, steps_until(tf, dt), theta2, G, eps2, scheme=scheme, force=force)
    q, p = t.get_state()
    t.close()
    return np.ascontiguousarray(q.T), np.ascontiguousarray(p.T)


def verlet_direct(ms, qs, ps, G, eps2, dt, tf):
    """test_integration_verlet_direct (integration.f90:85-108): (3,N) in and out."""
    return _run(ms, qs, ps, 10, "direct", G, eps2, dt, tf, 0.0)


def verlet_bh(ms, qs, ps, G, eps2, dt, tf, theta2, mult=1, walk="tree", precision="fp64"):
    """test_integration_verlet_bh (:110-172): BH forces of order mult (1/4/8)
    with a rebuild each step.  walk="bunched" uses the bunched walk instead of
    the per-target one (precision "fp64" or "mixed")."""
    return _run(ms, qs, ps, _method(mult, 10), walk, G, eps2, dt, tf, theta2, precision)


def verlet_dehnen(ms, qs, ps, G, eps2, dt, tf, theta2, mult=1):
    """test_integration_verlet_dehnen (:174-209)."""
    return _run(ms, qs, ps, _method(mult, 20), "tree", G, eps2, dt, tf, theta2)


def energy_momenta(ms, qs, ps, G=1.0, eps2=0.0):
    """[E, P(3), J(3)] of one system; qs, ps (N,3)."""
    t = OctreeType().pre_init(L.f64(ms), 10)
    t.set_state(qs, ps)
    out = t.energy(G, eps2)[0]
    t.close()
    return out


def total_energy(ms, qs, ps, G, eps2):
    """total_energy (integration.f90:236-262); qs, ps (N,3)."""
    return float(energy_momenta(ms, qs, ps, G, eps2)[0])


def total_linear_momentum(ps):
    return np.asarray(ps, dtype=np.float64).sum(axis=0)


def total_angular_momentum(qs, ps):
    return np.cross(np.asarray(qs, dtype=np.float64), np.asarray(ps, dtype=np.float64)).sum(axis=0)


def _split(u0):
    u0 = np.asarray(u0, dtype=np.float64)
    n = u0.shape[0] // 2
    return n, u0[:n], u0[n:]


def G_til(m, u0, dt, qntt, theta, G, eps):
    """Coarse propagator (parareal.f90:8-44): qntt symplectic Euler steps with
    quadrupole BH forces (theta squared, parareal/api3d.f90:30) or, for
    theta == -1, direct forces.  u0 (2N,3) -> (2N,3)."""
    n, q, p = _split(u0)
    direct = theta == -1
    t = OctreeType().pre_init(L.f64(m), 11)
    t.set_state(q, p)
    t.integrate(dt, qntt, 0.0 if direct else theta * theta, G, eps * eps, scheme="euler",
                force="direct" if direct else "tree")
    q, p = t.get_state()
    return np.concatenate([q, p])


def F_til(m, u0, dt, qntt, theta, G, eps):
    """Fine propagator (parareal.f90:46-69): qntt Verlet steps with direct forces."""
    n, q, p = _split(u0)
    t = OctreeType().pre_init(L.f64(m), 10)
    t.set_state(q, p)
    t.integrate(dt, qntt, 0.0, G, eps * eps, scheme="verlet", force="direct")
    q, p = t.get_state()
    return np.concatenate([q, p])


_FORCE = {"tree": L.HQ_FORCE_TREE, "bunched": L.HQ_FORCE_BUNCHED, "direct": L.HQ_FORCE_DIRECT}
_PREC = {"fp64": L.HQ_FP64, "mixed": L.HQ_MIXED}


def integrate_parareal(m, q0, p0, tf, num_windows, N_itermax, theta, G, eps, maxerror, qntt,
                       coarse=None, fine=None, correction="exact", group_size=64):
    """integrate_parareal (parareal.f90:97-186) on the GPU.

    With coarse/fine left as None this is the reference's configuration: coarse
    = one symplectic Euler step per window with quadrupole BH (theta squared;
    direct when theta == -1), fine = qntt Verlet steps with direct forces.
    coarse / fine may instead be dicts with keys force ("tree" | "bunched" |
    "direct"), method (10-12, 30-32), theta (opening angle), steps (per
    window) and precision ("fp64" | "mixed", bunched only); coarse may also set
    scheme ("euler", the reference's G_til, or "verlet").

    correction: "exact" computes F + (G_new - G_old), so a converged window is
    the sequential fine solution bit for bit; "reference" keeps
    (G_new + F) - G_old as parareal.f90:154-155 orders it.

    Returns q1, p1 with shape (num_windows, N, 3) (the last iterate, as the
    reference's q1(:,:,:)) and info = {"errors", "iterations"}."""
    m = L.f64(m)
    n = len(m)
    q0 = np.asarray(q0, dtype=np.float64)
    p0 = np.asarray(p0, dtype=np.float64)
    if coarse is None:
        coarse = {"force": "direct"} if theta == -1 else {"force": "tree", "method": 11,
                                                          "theta": theta}
        coarse.setdefault("steps", 1)
    if fine is None:
        fine = {"force": "direct", "steps": qntt}
    prm = L.PararealParams()
    prm.n, prm.num_windows, prm.itermax = n, num_windows, N_itermax
    prm.tf, prm.maxerror, prm.G, prm.eps2 = tf, maxerror, G, eps * eps
    for pre, d in (("coarse", coarse), ("fine", fine)):
        setattr(prm, pre + "_force", _FORCE[d.get("force", "tree")])
        setattr(prm, pre + "_method", int(d.get("method", 11)))
        setattr(prm, pre + "_precision", _PREC[d.get("precision", "fp64")])
        setattr(prm, pre + "_steps", int(d.get("steps", 1)))
        th = float(d.get("theta", 0.0))
        setattr(prm, pre + "_theta2", th * th)
    prm.group_size = group_size
    prm.correction = {"exact": L.HQ_CORR_EXACT, "reference": L.HQ_CORR_REFERENCE}[correction]
    prm.coarse_scheme = {"euler": L.HQ_SCHEME_EULER,
                         "verlet": L.HQ_SCHEME_VERLET}[coarse.get("scheme", "euler")]
    cols = [np.ascontiguousarray(a[:, c]) for a in (q0, p0) for c in range(3)]
    U = np.zeros(6 * num_windows * n)
    errors = np.zeros(max(N_itermax - 1, 1))
    it = C.c_int(0)
    L.check(L.lib().hq_parareal(C.byref(prm), L.ptr(m), *[L.ptr(c) for c in cols], L.HQ_HOST,
                                L.ptr(U), L.HQ_HOST, L.ptr(errors), C.byref(it)))
    U = U.reshape(6, num_windows, n)
    q1 = np.ascontiguousarray(U[:3].transpose(1, 2, 0))
    p1 = np.ascontiguousarray(U[3:].transpose(1, 2, 0))
    k = it.value
    return q1, p1, {"errors": errors[:k - 1].copy(), "iterations": k}
