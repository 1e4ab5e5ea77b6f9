"""GPU: Verlet, symplectic Euler, energies and Parareal against the oracle.

Bars: GPU trajectories against the oracle's on the same inputs (forces agree
to round-off for the direct sum and to 1e-10 for the FP64 tree, so the
trajectories agree to 1e-9 over a few dozen steps); a batch of systems
integrated together bit-identical to each system integrated alone; Parareal
with the exact correction bit-identical to the sequential fine propagation
once converged; Parareal with the reference's correction order within 1e-9 of
the oracle's restatement of parareal.f90.
"""
import numpy as np
import pytest

import oracle as O
from hierarquicos_amd import HQError, OctreeType, ics
from hierarquicos_amd import integrate as I

pytestmark = pytest.mark.gpu


def _plummer(n, seed):
    m, q, v = ics.plummer(n, seed=seed, velocities=True)
    return m, q, v * m[:, None]


def _gpu(m, q, p, method, dt, nsteps, scheme, force, theta2=0.25, eps2=1e-4, precision="fp64"):
    t = OctreeType().pre_init(m, method, precision=precision)
    t.set_state(q, p)
    t.integrate(dt, nsteps, theta2, 1.0, eps2, scheme=scheme, force=force)
    return t.get_state()


def _close(a, b, tol):
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max()) <= tol * scale


def test_verlet_direct_matches_oracle():
    m, q, p = _plummer(300, 1)
    qg, pg = _gpu(m, q, p, 10, 1e-2, 20, "verlet", "direct")
    qo, po = O.integrate(m, q, p, O.VERLET, 1e-2, 20, O.FORCE_DIRECT, G=1.0, eps2=1e-4)
    assert _close(qg, qo, 1e-12) and _close(pg, po, 1e-10)


@pytest.mark.parametrize("scheme", ["verlet", "euler"])
def test_direct_graph_replay_matches_launched_steps(scheme):
    """Direct-force integration replays one captured step; the launched steps
    give the same state bit for bit.  Several graph runs of a batch handle, with
    launched work and synchronisations between them, each checked."""
    sizes = [700, 1500, 300]
    sys_ = [_plummer(k, 20 + i) for i, k in enumerate(sizes)]
    m = np.concatenate([s[0] for s in sys_])
    q = np.concatenate([s[1] + 15.0 * i for i, s in enumerate(sys_)])
    p = np.concatenate([s[2] for s in sys_])
    tg = OctreeType().pre_init(m, 10, counts=sizes)
    te = OctreeType().pre_init(m, 10, counts=sizes).tune(graph=False)
    for dt, nsteps in ((1e-3, 7), (2e-3, 3), (5e-4, 12)):
        tg.set_state(q, p).integrate(dt, nsteps, 0.0, 1.0, 1e-4, scheme=scheme, force="direct")
        te.set_state(q, p).integrate(dt, nsteps, 0.0, 1.0, 1e-4, scheme=scheme, force="direct")
        (qg, pg), (qe, pe) = tg.get_state(), te.get_state()
        assert np.array_equal(qg, qe) and np.array_equal(pg, pe), (dt, nsteps)
        assert tg.energy(1.0, 1e-4).shape == (3, 7)   # launched work between graph runs
        q, p = qg, pg


@pytest.mark.parametrize("scheme", ["verlet", "euler"])
def test_bunched_graph_replay_matches_launched_steps(scheme):
    """Bunched mixed-precision integration replays one captured step (node and
    bunch counts on the device, split walk); launched steps with the fused walk
    give the same state bit for bit, over several graph runs of a batch handle
    with launched work between them.  Afterwards the handle describes the last
    step's tree."""
    sizes = [3000, 700, 5000]
    sys_ = [_plummer(k, 40 + i) for i, k in enumerate(sizes)]
    m = np.concatenate([s[0] for s in sys_])
    q = np.concatenate([s[1] + 15.0 * i for i, s in enumerate(sys_)])
    p = np.concatenate([s[2] for s in sys_])
    # graph replays with the split walk (few bunches) against launched fused walks
    tg = OctreeType().pre_init(m, 11, counts=sizes, precision="mixed").tune(leaf=4)
    te = OctreeType().pre_init(m, 11, counts=sizes, precision="mixed").tune(leaf=4, graph=False,
                                                                            split=0)
    for dt, nsteps in ((1e-3, 6), (2e-3, 3), (5e-4, 9)):
        tg.set_state(q, p).integrate(dt, nsteps, 0.25, 1.0, 1e-4, scheme=scheme, force="bunched")
        te.set_state(q, p).integrate(dt, nsteps, 0.25, 1.0, 1e-4, scheme=scheme, force="bunched")
        (qg, pg), (qe, pe) = tg.get_state(), te.get_state()
        assert np.array_equal(qg, qe) and np.array_equal(pg, pe), (dt, nsteps)
        assert tg.info()["nnodes"] == te.info()["nnodes"]
        assert tg.energy(1.0, 1e-4).shape == (3, 7)   # launched work between graph runs
        q, p = qg, pg


@pytest.mark.parametrize("force", ["bunched", "direct"])
def test_graph_cache_and_unrolled_steps(force):
    """A handle keeps its instantiated step graphs: repeated calls with the
    same parameters replay the kept graph (also over a longer run), and a
    changed parameter captures anew while the earlier graph stays kept for
    the next call that needs it.  Every call gives the launched steps' state
    bit for bit."""
    sizes = [2000, 600]
    sys_ = [_plummer(k, 80 + i) for i, k in enumerate(sizes)]
    m = np.concatenate([s[0] for s in sys_])
    q = np.concatenate([s[1] + 15.0 * i for i, s in enumerate(sys_)])
    p = np.concatenate([s[2] for s in sys_])
    method = 11 if force == "bunched" else 10
    prec = "mixed" if force == "bunched" else "fp64"
    tg = OctreeType().pre_init(m, method, counts=sizes, precision=prec)
    te = OctreeType().pre_init(m, method, counts=sizes, precision=prec).tune(graph=False, split=0)
    for dt, nsteps in ((1e-3, 6), (1e-3, 6), (1e-3, 6), (1e-3, 70), (2e-3, 6), (1e-3, 6)):
        tg.set_state(q, p).integrate(dt, nsteps, 0.25, 1.0, 1e-4, force=force)
        te.set_state(q, p).integrate(dt, nsteps, 0.25, 1.0, 1e-4, force=force)
        (qg, pg), (qe, pe) = tg.get_state(), te.get_state()
        assert np.array_equal(qg, qe) and np.array_equal(pg, pe), (dt, nsteps)
        q, p = qg, pg


@pytest.mark.parametrize("split", [0, 1])
def test_bunched_graph_grid_too_small_reruns_launched(split):
    """A replayed step with more bunches than the captured walk's grid (forced
    here: a grid of 8 bunches for about 80) flags the replays; the kernels
    walk only the bunches the grid holds, and the steps are then redone
    launched, which gives the launched path's state bit for bit."""
    n = 5000
    m, q, v = ics.plummer(n, seed=21, velocities=True)
    p = v * m[:, None]
    tg = OctreeType().pre_init(m, 12, precision="mixed").tune(graph_grid=8, split=split)
    te = OctreeType().pre_init(m, 12, precision="mixed").tune(graph=False, split=split)
    for scheme in ("verlet", "euler"):
        tg.set_state(q, p).integrate(1e-3, 5, 0.25, 1.0, 1e-4, scheme=scheme, force="bunched")
        te.set_state(q, p).integrate(1e-3, 5, 0.25, 1.0, 1e-4, scheme=scheme, force="bunched")
        (qg, pg), (qe, pe) = tg.get_state(), te.get_state()
        assert np.array_equal(qg, qe) and np.array_equal(pg, pe), scheme
        assert tg.info()["ngroups"] > 8


def test_bunched_graph_replay_tiny_systems():
    """Graph replay and split walk with systems of 1, 2, 3 and 65 bodies in one
    batch (single-leaf roots, bunches smaller than a wavefront)."""
    sizes = [1, 2, 3, 65, 400]
    sys_ = [_plummer(k, 60 + i) for i, k in enumerate(sizes)]
    m = np.concatenate([s[0] for s in sys_])
    q = np.concatenate([s[1] + 9.0 * i for i, s in enumerate(sys_)])
    p = np.concatenate([s[2] for s in sys_])
    for scheme in ("verlet", "euler"):
        tg = OctreeType().pre_init(m, 12, counts=sizes, precision="mixed")
        te = OctreeType().pre_init(m, 12, counts=sizes, precision="mixed").tune(graph=False,
                                                                                split=0)
        tg.set_state(q, p).integrate(1e-3, 5, 0.25, 1.0, 1e-4, scheme=scheme, force="bunched")
        te.set_state(q, p).integrate(1e-3, 5, 0.25, 1.0, 1e-4, scheme=scheme, force="bunched")
        (qg, pg), (qe, pe) = tg.get_state(), te.get_state()
        assert np.array_equal(qg, qe) and np.array_equal(pg, pe), scheme
        assert np.all(np.isfinite(qg))


def test_bunched_graph_error_reported_like_launched_steps():
    """An error inside the replayed steps (two bodies meeting exactly at step
    3, so lost at depth 25) is flagged on the device; the steps are then rerun
    launched and return the launched path's error."""
    n = 2048
    m, q, v = ics.plummer(n, seed=9, velocities=True)
    m = np.full(n, 1.0 / n)
    q[0], q[1] = (-0.375, 0.1, 0.1), (0.375, 0.1, 0.1)
    v[0], v[1] = (1.0, 0.0, 0.0), (-1.0, 0.0, 0.0)
    p = v * m[:, None]
    codes = []
    for graph in (True, False):
        t = OctreeType().pre_init(m, 11, precision="mixed").tune(graph=graph)
        t.set_state(q, p)
        with pytest.raises(HQError) as e:
            t.integrate(0.125, 6, 0.25, 0.0, 0.0, scheme="verlet", force="bunched")
        codes.append(e.value.code)
    assert codes[0] == codes[1] == 3, codes


@pytest.mark.parametrize("method,scheme", [(10, "verlet"), (11, "verlet"), (12, "euler")])
def test_tree_integration_matches_oracle(method, scheme):
    m, q, p = _plummer(500, 2)
    qg, pg = _gpu(m, q, p, method, 5e-3, 12, scheme, "tree")
    qo, po = O.integrate(m, q, p, O.VERLET if scheme == "verlet" else O.EULER, 5e-3, 12,
                         O.FORCE_TREE, method, 0.25, 1.0, 1e-4)
    assert _close(qg, qo, 1e-9) and _close(pg, po, 1e-8)


def test_verlet_bh_driver_and_step_count():
    """verlet_bh keeps the reference's (3,N) layout and DO WHILE (t < tf) count."""
    m, q, p = _plummer(200, 3)
    qs, ps = I.verlet_bh(m, q.T, p.T, 1.0, 0.05 ** 2, 0.0078125, 0.25, 0.25, mult=1)
    assert qs.shape == (3, 200)
    qo, po = O.integrate(m, q, p, O.VERLET, 0.0078125, 32, O.FORCE_TREE, 10, 0.25, 1.0,
                         0.05 ** 2)
    assert _close(qs.T, qo, 1e-9) and _close(ps.T, po, 1e-8)


def test_energy_matches_oracle():
    m, q, p = _plummer(700, 4)
    e = I.energy_momenta(m, q, p, 1.0, 1e-4)
    eo = O.energy(m, q, p, 1.0, 1e-4)
    assert abs(e[0] - eo[0]) < 1e-12 * abs(eo[0])
    assert np.allclose(e[1:], eo[1:], rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("force,precision,method", [("tree", "fp64", 11),
                                                    ("bunched", "mixed", 12),
                                                    ("direct", "fp64", 10)])
def test_ensemble_integration_bitidentical(force, precision, method):
    rng = np.random.default_rng(11)
    sizes = [int(s) for s in rng.integers(50, 900, 12)]
    sys = [_plummer(n, 20 + k) for k, n in enumerate(sizes)]
    m = np.concatenate([s[0] for s in sys])
    q = np.concatenate([s[1] for s in sys])
    p = np.concatenate([s[2] for s in sys])
    tb = OctreeType().pre_init(m, method, precision=precision, counts=sizes)
    tb.set_state(q, p)
    tb.integrate(5e-3, 6, 0.25, 1.0, 1e-4, force=force)
    qb, pb = tb.get_state()
    eb = tb.energy(1.0, 1e-4)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for k, (mk, qk, pk) in enumerate(sys):
        qa, pa = _gpu(mk, qk, pk, method, 5e-3, 6, "verlet", force, precision=precision)
        assert np.array_equal(qb[off[k]:off[k + 1]], qa), k
        assert np.array_equal(pb[off[k]:off[k + 1]], pa), k
        ea = I.energy_momenta(mk, qa, pa, 1.0, 1e-4)
        assert abs(eb[k, 0] - ea[0]) < 1e-12 * abs(ea[0])


def _cube(n, seed=0):
    rng = np.random.default_rng(seed)
    return np.full(n, 1.0 / n), 10.0 * (2.0 * rng.random((n, 3)) - 1.0), np.zeros((n, 3))


def test_parareal_exact_equals_sequential():
    """Reference propagators (coarse Euler + BH quadrupole, fine Verlet +
    direct); run to num_windows iterates: every window equals the sequential
    fine solution bit for bit."""
    m, q, p = _cube(400)
    nw, qn, tf, eps = 8, 8, 4.0, 0.5
    # window w is exact after w corrections; the error of the last window
    # reads 0 one iterate after that window converged
    q1, p1, info = I.integrate_parareal(m, q, p, tf, nw, nw + 1, 1.0, 1.0, eps, 0.0, qn)
    assert info["iterations"] <= nw + 1
    assert info["errors"][-1] == 0.0
    qq, pp = q.copy(), p.copy()
    for w in range(nw - 1):
        qq, pp = _gpu(m, qq, pp, 10, tf / nw / qn, qn, "verlet", "direct", eps2=eps * eps)
        assert np.array_equal(q1[w + 1], qq), w
        assert np.array_equal(p1[w + 1], pp), w


def test_parareal_batched_bh_fine_exact():
    """Batched form of the Plummer study: coarse monopole BH theta=0.9, fine
    quadrupole bunched mixed-precision walk theta=0.5; converged windows equal
    the sequential fine propagation bit for bit."""
    m, q, p = _plummer(2000, 5)
    nw, tf = 6, 0.6
    coarse = {"force": "tree", "method": 10, "theta": 0.9, "steps": 2}
    fine = {"force": "bunched", "method": 11, "theta": 0.5, "steps": 10, "precision": "mixed"}
    q1, p1, info = I.integrate_parareal(m, q, p, tf, nw, nw + 1, -2, 1.0, 1e-2, 0.0, 0,
                                        coarse=coarse, fine=fine)
    assert info["errors"][-1] == 0.0
    qq, pp = q.copy(), p.copy()
    for w in range(nw - 1):
        qq, pp = _gpu(m, qq, pp, 11, tf / nw / 10, 10, "verlet", "bunched", 0.25, 1e-4,
                      precision="mixed")
        assert np.array_equal(q1[w + 1], qq), w
        assert np.array_equal(p1[w + 1], pp), w


def test_parareal_reference_order_matches_oracle():
    m, q, p = _cube(300, 1)
    nw, qn, tf, eps = 10, 8, 5.0, 0.5
    q1, p1, info = I.integrate_parareal(m, q, p, tf, nw, 4, 1.0, 1.0, eps, 1e-20, qn,
                                        correction="reference")
    U, err, k = O.parareal(m, q, p, tf, nw, 4, 1e-20, 1.0, eps * eps,
                           coarse=(O.FORCE_TREE, 11, 1.0, 1), fine=(O.FORCE_DIRECT, 11, 1.0, qn))
    assert info["iterations"] == k == 4
    assert _close(q1, U[:, :3].transpose(0, 2, 1), 1e-9)
    assert _close(p1, U[:, 3:].transpose(0, 2, 1), 1e-7)
    assert np.allclose(info["errors"], err, rtol=1e-2)
