"""CPU pins of the oracle's integrators and Parareal (parareal/parareal.f90,
benchmarks/integration_verlet/integration.f90).

Pins: the closed-form circular two-body orbit fixes the order of each scheme
(Verlet second order: error ratio 4 per halving of dt; symplectic Euler first
order: ratio 2), energy and momenta against a numpy brute force, and
Parareal's defining property: after num_windows - 1 corrections the iterate is
the sequential fine solution (to rounding).
"""
import numpy as np
import pytest

import oracle as O
from hierarquicos_amd.integrate import steps_until


def _circular(dt, T, scheme):
    m = np.array([0.5, 0.5])
    q = np.array([[0.5, 0, 0], [-0.5, 0, 0.0]])
    p = np.array([[0, 0.25, 0], [0, -0.25, 0.0]])   # speed 0.5, omega 1
    q, p = O.integrate(m, q, p, scheme, dt, int(round(T / dt)))
    exact = np.array([0.5 * np.cos(T), 0.5 * np.sin(T), 0.0])
    return np.linalg.norm(q[0] - exact), q, p


@pytest.mark.parametrize("scheme,ratio", [(O.VERLET, 4.0), (O.EULER, 2.0)])
def test_scheme_order_on_circular_orbit(scheme, ratio):
    errs = [_circular(2.0 ** -k, 1.0, scheme)[0] for k in (5, 6, 7, 8)]
    for a, b in zip(errs, errs[1:]):
        assert abs(a / b - ratio) < 0.01 * ratio, errs


def test_verlet_conserves_energy_and_momentum():
    rng = np.random.default_rng(3)
    n = 40
    m = rng.uniform(0.5, 1.5, n) / n
    q = rng.normal(size=(n, 3))
    p = rng.normal(size=(n, 3)) * m[:, None] * 0.3
    e0 = O.energy(m, q, p, 1.0, 0.01)
    q1, p1 = O.integrate(m, q, p, O.VERLET, 1e-3, 200, O.FORCE_DIRECT, G=1.0, eps2=0.01)
    e1 = O.energy(m, q1, p1, 1.0, 0.01)
    assert abs(e1[0] - e0[0]) < 1e-6 * abs(e0[0])
    assert np.allclose(e1[1:4], e0[1:4], atol=1e-15)      # pairwise forces: P exact
    assert np.allclose(e1[4:7], e0[4:7], atol=1e-12)      # central forces: J conserved


def test_energy_matches_brute_force():
    rng = np.random.default_rng(1)
    n = 30
    m = rng.uniform(0.1, 1.0, n)
    q = rng.normal(size=(n, 3))
    p = rng.normal(size=(n, 3))
    eps2 = 0.04
    T = float(np.sum((p * p).sum(1) / (2 * m)))
    V = 0.0
    for a in range(n):
        for b in range(a):
            V -= m[a] * m[b] / np.sqrt(np.sum((q[a] - q[b]) ** 2) + eps2)
    e = O.energy(m, q, p, 1.0, eps2)
    assert abs(e[0] - (T + V)) < 1e-12 * abs(T + V)
    assert np.allclose(e[1:4], p.sum(0), rtol=1e-13, atol=1e-14)
    assert np.allclose(e[4:7], np.cross(q, p).sum(0), rtol=1e-13, atol=1e-14)


def _cube(n, seed=0):
    rng = np.random.default_rng(seed)
    return np.full(n, 1.0 / n), 10.0 * (2.0 * rng.random((n, 3)) - 1.0), np.zeros((n, 3))


def test_parareal_reaches_sequential_fine_solution():
    m, q, p = _cube(60)
    nw, qn, tf, eps2 = 6, 8, 3.0, 0.25
    U, err, k = O.parareal(m, q, p, tf, nw, nw, 0.0, 1.0, eps2,
                           coarse=(O.FORCE_TREE, 11, 1.0, 1), fine=(O.FORCE_DIRECT, 11, 1.0, qn))
    assert k == nw
    qq, pp = q.copy(), p.copy()
    for w in range(nw - 1):
        qq, pp = O.integrate(m, qq, pp, O.VERLET, tf / nw / qn, qn, O.FORCE_DIRECT, eps2=eps2)
        assert np.allclose(U[w + 1, :3].T, qq, rtol=0, atol=1e-12)
        assert np.allclose(U[w + 1, 3:].T, pp, rtol=0, atol=1e-15)


def test_parareal_converges_geometrically():
    m, q, p = _cube(200)
    U, err, k = O.parareal(m, q, p, 10.0, 30, 5, 1e-20, 1.0, 0.25,
                           coarse=(O.FORCE_TREE, 11, 1.0, 1), fine=(O.FORCE_DIRECT, 11, 1.0, 16))
    assert k == 5
    assert np.all(err[1:] < err[:-1] / 30.0), err


def test_steps_until_matches_fortran_loop():
    assert steps_until(1.0, 0.0078125) == 128        # integration.f90:22-23
    assert steps_until(0.0, 0.1) == 0
    assert steps_until(1.0, -0.25) == 4               # t advances by ABS(dt)
