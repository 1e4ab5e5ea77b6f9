"""GPU: the Fortran mirrors of hierarquicos_amd/fortran (OctreeType, api_mod,
compute_forces_direct, the Verlet and Parareal drivers, the harness metric),
driven from tests/fortran/fdriver.f90 the way the reference's Fortran drivers
call theirs, agree bit for bit with the Python mirrors on the same inputs.
Both sit on the same C ABI, so any difference is an argument, layout or
squaring mistake in one of the two host layers; the Python layer itself is
pinned against the oracle and the reference's golden vectors elsewhere.  The
direct sum is also checked against the oracle here, which pins the layouts
of both layers."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from hierarquicos_amd import OctreeType, api, direct_forces, harness, ics, integrate

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDRIVER = os.path.join(REPO, "tests", "fortran", "libfdriver.so")

_d = C.POINTER(C.c_double)
_i64, _i32, _f64 = C.c_int64, C.c_int, C.c_double


@pytest.fixture(scope="module")
def fd():
    assert os.path.exists(FDRIVER), "tests/fortran/libfdriver.so is not built (__graft_entry__.build)"
    lib = C.CDLL(FDRIVER)
    sig = {
        "fd_api": [_i64, _d, _d, _d, _d, _f64, _f64, _f64, _i32, _d, _d],
        "fd_octree": [_i64, _d, _d, _d, _d, _i32, _i32, _f64, _f64, _f64, _i32, _d],
        "fd_bunched_fmm": [_i64, _d, _d, _d, _d, _f64, _f64, _f64, _i32, _i32, _i32, _d, _d],
        "fd_direct": [_i64, _d, _d, _d, _d, _f64, _f64, _d],
        "fd_errors": [_i64, _d, _d, _d, _d, _d],
        "fd_verlet": [_i64, _d, _d, _d, _f64, _f64, _f64, _f64, _f64, _i32, _i32, _d],
        "fd_propagators": [_i64, _d, _d, _f64, _i32, _f64, _f64, _f64, _d, _d],
        "fd_parareal": [_i64, _d, _d, _d, _f64, _i32, _i32, _f64, _f64, _f64, _f64, _i32, _i32,
                        _d, _d, C.POINTER(_i32), _d, _d],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    return lib


def p(a):
    assert a.dtype == np.float64 and (a.flags.c_contiguous or a.flags.f_contiguous)
    return a.ctypes.data_as(_d)


def _sphere(n, seed):
    m, q, v = ics.plummer(n, seed=seed, velocities=True)
    return m, q, v * m[:, None]


def test_api_mod_compute_forces(fd):
    m, q, _ = _sphere(3000, 1)
    x, y, z = (np.ascontiguousarray(q[:, c]) for c in range(3))
    F = np.zeros((3000, 3), order="F")
    Fd = np.zeros((3000, 3), order="F")
    assert fd.fd_api(3000, p(m), p(x), p(y), p(z), 0.5, 1.0, 0.01, 1, p(F), p(Fd)) == 0
    assert np.array_equal(F, api.api_mod.compute_forces(m, x, y, z, 0.5, 1.0, 0.01, True))
    assert np.array_equal(Fd, api.api_mod.compute_forces_direct(m, x, y, z, 1.0, 0.01))
    ref = O.direct_targets(m, x, y, z, np.arange(3000), 1.0, 1e-4)
    assert np.max(np.abs(Fd - ref) / np.abs(ref).max()) < 1e-13


@pytest.mark.parametrize("method,gs,nt", [(12, 64, 4), (11, 64, 1), (21, 64, 1), (31, 32, 1)])
def test_octree_type_methods(fd, method, gs, nt):
    m, q, _ = _sphere(5000, 2)
    x, y, z = (np.ascontiguousarray(q[:, c]) for c in range(3))
    th2, G, e2 = 0.36, 1.0, 1e-4
    F = np.zeros((5000, 3))            # the Fortran (3,N) array
    assert fd.fd_octree(5000, p(m), p(x), p(y), p(z), method, gs, th2, G, e2, nt, p(F)) == 0
    t = OctreeType().pre_init(m, method, group_size=gs).init(x, y, z)
    if method < 20:
        ref = t.forces_all(th2, G, e2)
    elif method < 30:
        ref = t.dehnen_eval(th2, e2, G)
    else:
        ref = t.forces_groups(th2, G, e2)
    if method // 10 == 2:
        # Dehnen's interactions are accumulated with atomics: equal to round-off
        d = np.linalg.norm(F - ref.T, axis=1) / np.linalg.norm(ref.T, axis=1)
        assert d.max() < 1e-12
    else:
        assert np.array_equal(F, ref.T)


def test_bunched_walk_and_fmm(fd):
    m, q, _ = _sphere(6000, 3)
    x, y, z = (np.ascontiguousarray(q[:, c]) for c in range(3))
    Fb = np.zeros((6000, 3))
    Ff = np.zeros((6000, 3))
    assert fd.fd_bunched_fmm(6000, p(m), p(x), p(y), p(z), 0.25, 1.0, 1e-6, 1, 4, 1, p(Fb),
                             p(Ff)) == 0
    t = OctreeType().pre_init(m, 10).init(x, y, z)
    assert np.array_equal(Fb, t.forces_bunched(0.25, 1.0, 1e-6, "mixed").T)
    ref = t.fmm(0.25, 1.0, 1e-6, 4, "fp64").T
    # the FMM's far-field sums are not ordered, so its bits may vary run to run
    assert np.max(np.linalg.norm(Ff - ref, axis=1) / np.linalg.norm(ref, axis=1)) < 1e-12


def test_direct_forces_and_error_metric(fd):
    m, q, _ = _sphere(4000, 4)
    x, y, z = (np.ascontiguousarray(q[:, c]) for c in range(3))
    F = np.zeros((4000, 3))
    assert fd.fd_direct(4000, p(m), p(x), p(y), p(z), 1.0, 1e-4, p(F)) == 0
    assert np.array_equal(F, direct_forces(m, x, y, z, 1.0, 1e-4))
    ref = O.direct_targets(m, x, y, z, np.arange(4000), 1.0, 1e-4)
    assert np.max(np.abs(F - ref) / np.abs(ref).max()) < 1e-13
    Fbh = np.ascontiguousarray(OctreeType().pre_init(m, 11).init(x, y, z).forces_all(0.25, 1.0,
                                                                                    1e-4).T)
    qs = np.ascontiguousarray(q)
    err = np.zeros(7)
    assert fd.fd_errors(4000, p(m), p(qs), p(Fbh), p(F), p(err)) == 0
    ref_err = harness.evaluate_error_on_accelerations(m, q.T, Fbh.T, F.T)
    assert np.array_equal(err, ref_err)
    assert 0 < err[0] < 1e-2


@pytest.mark.parametrize("kind,mult", [(0, 1), (1, 4), (2, 4)])
def test_verlet_drivers(fd, kind, mult):
    m, q, pm = _sphere(2000, 5)
    G, e2, dt, tf, th2 = 1.0, 1e-4, 1e-3, 5e-3, 0.25
    qs, ps = q.copy(), pm.copy()          # advanced in place by the driver
    E = np.zeros(2)
    assert fd.fd_verlet(2000, p(m), p(qs), p(ps), G, e2, dt, tf, th2, mult, kind, p(E)) == 0
    if kind == 0:
        rq, rp = integrate.verlet_direct(m, q.T, pm.T, G, e2, dt, tf)
    elif kind == 1:
        rq, rp = integrate.verlet_bh(m, q.T, pm.T, G, e2, dt, tf, th2, mult)
    else:
        rq, rp = integrate.verlet_dehnen(m, q.T, pm.T, G, e2, dt, tf, th2, mult)
    if kind == 2:
        # Dehnen's forces are accumulated with atomics: equal to round-off
        assert np.max(np.abs(qs - rq.T)) < 1e-12 and np.max(np.abs(ps - rp.T)) < 1e-12
    else:
        assert np.array_equal(qs, rq.T) and np.array_equal(ps, rp.T)
    assert E[0] == integrate.total_energy(m, q, pm, G, e2)
    assert E[1] == integrate.total_energy(m, qs, ps, G, e2)
    assert abs(E[1] - E[0]) < 1e-4 * abs(E[0])


@pytest.mark.parametrize("theta", [0.5, -1.0])
def test_parareal_propagators(fd, theta):
    m, q, pm = _sphere(1500, 6)
    u0 = np.asfortranarray(np.concatenate([q, pm]))
    ug = np.zeros((3000, 3), order="F")
    uf = np.zeros((3000, 3), order="F")
    assert fd.fd_propagators(1500, p(m), p(u0), 2e-3, 3, theta, 1.0, 0.01, p(ug), p(uf)) == 0
    assert np.array_equal(ug, integrate.G_til(m, u0, 2e-3, 3, theta, 1.0, 0.01))
    assert np.array_equal(uf, integrate.F_til(m, u0, 2e-3, 3, theta, 1.0, 0.01))


@pytest.mark.parametrize("correction", [1, 0])
def test_integrate_parareal(fd, correction):
    n, nw, itmax, qntt = 800, 5, 6, 8
    m, q, pm = _sphere(n, 7)
    args = (0.1, nw, itmax, 0.5, 1.0, 0.01, 1e-13, qntt)
    q1 = np.zeros((nw, n, 3), order="F")
    p1 = np.zeros((nw, n, 3), order="F")
    it = C.c_int(0)
    errs = np.zeros(itmax)
    diag = np.zeros(7)
    q0, p0 = np.asfortranarray(q), np.asfortranarray(pm)
    assert fd.fd_parareal(n, p(m), p(q0), p(p0), *args, correction, p(q1), p(p1), C.byref(it),
                          p(errs), p(diag)) == 0
    rq, rp, info = integrate.integrate_parareal(
        m, q, pm, *args, correction={1: "exact", 0: "reference"}[correction])
    assert it.value == info["iterations"] >= 2
    assert np.array_equal(errs[:it.value - 1], info["errors"])
    assert np.array_equal(q1, rq) and np.array_equal(p1, rp)
    E = integrate.energy_momenta(m, rq[-1], rp[-1], 1.0, 0.01 * 0.01)
    assert diag[0] == E[0]
    assert np.allclose(diag[1:4], rp[-1].sum(axis=0), rtol=0, atol=1e-12)
    assert np.allclose(diag[4:7], np.cross(rq[-1], rp[-1]).sum(axis=0), rtol=0, atol=1e-12)


def test_errors_are_reported_not_fatal(fd):
    m, q, _ = _sphere(100, 8)
    x, y, z = (np.ascontiguousarray(q[:, c]) for c in range(3))
    F = np.zeros((100, 3))
    # method 13 does not exist: pre_init fails and the status comes back
    assert fd.fd_octree(100, p(m), p(x), p(y), p(z), 13, 64, 0.25, 1.0, 0.0, 1, p(F)) == 1
    assert fd.fd_octree(100, p(m), p(x), p(y), p(z), 11, 64, 0.25, 1.0, 0.0, 1, p(F)) == 0
