"""GPU: the benchmark configurations at their full sizes.

* 1M Plummer, octupole, eps 1e-3, theta 0.3 / 0.5 / 0.7: the build (keys,
  permutation, topology, moments) bit-exact against the oracle's, the FP64
  per-target walk within 1e-10 of the oracle's bh_forces_over_p on a 4096
  target sample, and the mixed-precision bunched walk within the 1e-4 RMS
  budget against the FP64 direct sum on seeds 0..4.
* 4M Hernquist, quadrupole, theta 0.6, eps 5e-4, after the autotuner.
* 4M uniform cube and Plummer sphere through the order-4/6 FMM.
* 10^4-body Parareal with 64 windows to 1e-20, and the 256-system ensemble.

The FP64 direct sum that the sample errors are measured against is the GPU's
(k_direct_fp64, pinned to the reference's golden vectors to 1e-13); a part of
each sample is checked against the oracle's direct sum here too.
"""
import numpy as np
import pytest

import oracle as O
from hierarquicos_amd import DeviceArray, OctreeType, direct_forces, ics, tune_table_clear
from hierarquicos_amd import integrate as I
from hierarquicos_amd.device import morton_keys

pytestmark = pytest.mark.gpu

N1M = 1 << 20
N4M = 1 << 22
FP64_TOL = 1e-10
BUDGET = 1e-4


def _sample(n, k=4096, seed=1234):
    return np.sort(np.random.default_rng(seed).choice(n, k, replace=False))


def _rel(F, Fd):
    r = np.linalg.norm(F - Fd, axis=1) / np.linalg.norm(Fd, axis=1)
    return float(np.sqrt(np.mean(r * r))), float(r.max())


def _device(m, q):
    dm = DeviceArray.from_numpy(m)
    dx, dy, dz = (DeviceArray.from_numpy(np.ascontiguousarray(q[:, k])) for k in range(3))
    return dm, dx, dy, dz


def _direct_sample(m, q, d, sample, eps2, check=256):
    """GPU FP64 direct sum on the sample, its first `check` targets against
    the oracle's."""
    Fd = direct_forces(*d, 1.0, eps2, targets=sample)
    x, y, z = (np.ascontiguousarray(q[:, k]) for k in range(3))
    Fo = O.direct_targets(m, x, y, z, sample[:check], 1.0, eps2)
    rel = np.linalg.norm(Fd[:check] - Fo, axis=1) / np.linalg.norm(Fo, axis=1)
    assert float(rel.max()) < 1e-12
    return Fd


def _canonical(depth, khi, klo):
    return np.lexsort((np.asarray(klo), np.asarray(khi), np.asarray(depth)))


# ------------------------------------------------------------------- 1M Plummer
@pytest.fixture(scope="module")
def plummer1m():
    m, q, _ = ics.plummer(N1M, a=1.0, rmax=10.0, seed=0)
    return m, q


def test_c2_morton_keys_and_permutation_bitexact(plummer1m):
    m, q = plummer1m
    keys, perm, L = morton_keys(m, q, 0)
    ko, Lo = O.morton_keys(m, q, 0)
    assert L == Lo and np.array_equal(keys, ko)
    assert np.array_equal(perm, np.argsort(ko, kind="stable"))


def test_c2_build_bitexact(plummer1m):
    """Topology (depth, octant path, type, leaf particle, parent) and moments
    (mass, centre of mass, quadrupole, octupole and the cube centres) of the
    1M tree equal the oracle's insertion build bit for bit."""
    m, q = plummer1m
    x, y, z = (np.ascontiguousarray(q[:, k]) for k in range(3))
    e = OctreeType().pre_init(m, 12).init(x, y, z).export()
    r = O.Tree(m, 12).init(x, y, z).export()
    nn = len(e["depth"])
    assert nn == len(r["depth"])
    cg = _canonical(e["depth"], e["key_hi"], e["key_lo"])
    cr = _canonical(r["depth"], r["khi"], r["klo"])
    for a, b in (("depth", "depth"), ("key_hi", "khi"), ("key_lo", "klo"), ("type", "type"),
                 ("particle", "particle"), ("mass", "mass")):
        assert np.array_equal(np.asarray(e[a])[cg], np.asarray(r[b])[cr]), a
    assert np.array_equal(e["com"][cg], np.stack([r["qx"], r["qy"], r["qz"]], 1)[cr])
    assert np.array_equal(e["quad"][cg], r["quad"][cr])
    assert np.array_equal(e["oct"][cg], r["oct"][cr])
    assert np.array_equal(e["center"][cg], np.stack([r["cx"], r["cy"], r["cz"]], 1)[cr])
    # parents, in canonical numbering
    rank_g = np.empty(nn, np.int64)
    rank_g[cg] = np.arange(nn)
    rank_r = np.empty(nn, np.int64)
    rank_r[cr] = np.arange(nn)
    cs = np.asarray(e["child_start"], np.int64)
    par_g = np.full(nn, -1, np.int64)   # children are consecutive ranges in order
    par_g[cs[0]:cs[nn]] = np.repeat(np.arange(nn), np.diff(cs))
    par_r = np.full(nn, -1, np.int64)
    ch = np.asarray(r["child"], np.int64)
    rows, cols = np.nonzero(ch >= 0)
    par_r[ch[rows, cols]] = rows
    pg = np.where(par_g >= 0, rank_g[np.maximum(par_g, 0)], -1)
    pr = np.where(par_r >= 0, rank_r[np.maximum(par_r, 0)], -1)
    assert np.array_equal(pg[cg], pr[cr])


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_c2_fp64_walk_matches_oracle(plummer1m, theta):
    m, q = plummer1m
    x, y, z = (np.ascontiguousarray(q[:, k]) for k in range(3))
    sample = _sample(N1M)
    th2, e2 = theta * theta, 1e-6
    t = OctreeType().pre_init(m, 12).init(x, y, z)
    F = t.forces_all(th2, 1.0, e2).T[sample]
    Fo = O.Tree(m, 12).init(x, y, z).bh_forces(th2, 1.0, e2, targets=sample)
    rel = np.linalg.norm(F - Fo, axis=1) / np.linalg.norm(Fo, axis=1)
    assert float(rel.max()) < FP64_TOL, float(rel.max())


def _mixed_1m(m, q, d, theta, sample, Fd):
    t = OctreeType().pre_init(d[0], 12, precision="mixed")
    t.tune(super_size=128, far=5.0, far_inner=3.0, leaf=4)   # the benchmark's settings
    dF = DeviceArray((3 * N1M,))
    t.init(*d[1:])
    t.forces_bunched(theta * theta, 1.0, 1e-6, "mixed", out=dF)
    assert t.info()["far_cells"] > 0 and t.info()["lost"] == 0
    return _rel(dF.numpy().reshape(N1M, 3)[sample], Fd)


def test_c2_mixed_walk_budget_every_seed():
    """theta 0.5: the RMS relative error against the FP64 direct sum stays
    below 0.9e-4 (10% under the 1e-4 budget) on each of seeds 0..4."""
    sample = _sample(N1M)
    worst = 0.0
    for seed in range(5):
        m, q, _ = ics.plummer(N1M, a=1.0, rmax=10.0, seed=seed)
        d = _device(m, q)
        Fd = _direct_sample(m, q, d, sample, 1e-6)
        rms, mx = _mixed_1m(m, q, d, 0.5, sample, Fd)
        print(f"seed {seed}: theta 0.5 mixed RMS {rms:.3e} max {mx:.3e}", flush=True)
        worst = max(worst, rms)
    assert worst < 0.9 * BUDGET, worst


@pytest.mark.parametrize("theta", [0.3, 0.7])
def test_c2_mixed_walk_other_angles(plummer1m, theta):
    """theta 0.3 and 0.7 (seed 0): the mixed bunched walk is at least as
    accurate as the reference's per-target FP64 walk at the same angle."""
    m, q = plummer1m
    d = _device(m, q)
    sample = _sample(N1M)
    Fd = _direct_sample(m, q, d, sample, 1e-6)
    rms, mx = _mixed_1m(m, q, d, theta, sample, Fd)
    t = OctreeType().pre_init(d[0], 12).init(*d[1:])
    dF = DeviceArray((3 * N1M,))
    t.forces_all(theta * theta, 1.0, 1e-6, out=dF)
    rms64, _ = _rel(dF.numpy().reshape(N1M, 3)[sample], Fd)
    print(f"theta {theta}: mixed RMS {rms:.3e} (max {mx:.3e}); per-target FP64 RMS {rms64:.3e}")
    assert rms < rms64, (rms, rms64)
    if theta < 0.5:
        assert rms < 0.2 * BUDGET


# ----------------------------------------------------------------- 4M Hernquist
def test_c3_hernquist_autotuned_walk():
    m, q, _ = ics.hernquist(N4M, a=1.0, rmax=50.0, seed=0)
    d = _device(m, q)
    th2, e2 = 0.36, 2.5e-7
    t = OctreeType().pre_init(d[0], 11, precision="mixed").init(*d[1:])
    assert t.info()["lost"] == 0
    tuned = t.autotune(th2, 1.0, e2, "mixed", reps=1)
    dF = DeviceArray((3 * N4M,))
    t.forces_bunched(th2, 1.0, e2, "mixed", out=dF)
    sample = _sample(N4M)
    Fd = _direct_sample(m, q, d, sample, e2, check=128)
    rms, mx = _rel(dF.numpy().reshape(N4M, 3)[sample], Fd)
    t64 = OctreeType().pre_init(d[0], 11).init(*d[1:])
    t64.forces_all(th2, 1.0, e2, out=dF)
    rms64, _ = _rel(dF.numpy().reshape(N4M, 3)[sample], Fd)
    print(f"C3: autotuned {tuned}; mixed RMS {rms:.3e} max {mx:.3e}; per-target FP64 RMS "
          f"{rms64:.3e}")
    assert rms < rms64, (rms, rms64)
    tune_table_clear()   # later handles of this process start from the defaults


def test_c3_theta_zero_is_direct_on_a_subset():
    m, q, _ = ics.hernquist(N4M, a=1.0, rmax=50.0, seed=0)
    sub = np.sort(np.random.default_rng(3).choice(N4M, 20000, replace=False))
    ms, qs = m[sub] * (N4M / 20000), q[sub]
    x, y, z = (np.ascontiguousarray(qs[:, k]) for k in range(3))
    t = OctreeType().pre_init(ms, 11, precision="mixed").init(x, y, z)
    F = t.forces_bunched(0.0, 1.0, 2.5e-7, "fp64").T
    Fd = O.direct_targets(ms, x, y, z, np.arange(20000), 1.0, 2.5e-7)
    rel = np.linalg.norm(F - Fd, axis=1) / np.linalg.norm(Fd, axis=1)
    assert float(rel.max()) < 1e-12
    Fm = t.forces_bunched(0.0, 1.0, 2.5e-7, "mixed").T
    assert _rel(Fm, Fd)[0] < 2e-6


# --------------------------------------------------------------------- 4M FMM
def test_c4_fmm_mfma_matches_fp64_on_plummer():
    m, q, _ = ics.plummer(200000, seed=4)
    x, y, z = q.T.copy()
    t = OctreeType().pre_init(m, 10).init(x, y, z)
    sample = _sample(200000, 2048)
    Fd = O.direct_targets(m, x, y, z, sample, 1.0, 0.0)
    for order in (4, 6):
        F64 = t.fmm(0.25, 1.0, 0.0, order, "fp64").T
        F32 = t.fmm(0.25, 1.0, 0.0, order, "mfma").T
        diff = np.linalg.norm(F32 - F64, axis=1) / np.linalg.norm(F64, axis=1)
        e64, e32 = _rel(F64[sample], Fd)[0], _rel(F32[sample], Fd)[0]
        print(f"Plummer 2e5 p={order}: MFMA-FP64 RMS {np.sqrt(np.mean(diff ** 2)):.2e}; "
              f"errors {e32:.2e} / {e64:.2e}")
        assert float(np.sqrt(np.mean(diff ** 2))) < 1e-5
        assert e32 < 1.1 * e64 + 2e-6, (order, e32, e64)


@pytest.mark.parametrize("profile", ["uniform", "plummer"])
def test_c4_fmm_4m(profile):
    if profile == "uniform":
        m, q, _ = ics.uniform_cube(N4M, half=1.0, seed=0)
    else:
        m, q, _ = ics.plummer(N4M, a=1.0, rmax=10.0, seed=0)
    d = _device(m, q)
    sample = _sample(N4M, 2048)
    Fd = _direct_sample(m, q, d, sample, 0.0, check=64)
    t = OctreeType().pre_init(d[0], 10).init(*d[1:])
    dF = DeviceArray((3 * N4M,))
    err = {}
    for order in (4, 6):
        t.fmm(0.25, 1.0, 0.0, order, "mfma", out=dF)
        err[order] = _rel(dF.numpy().reshape(N4M, 3)[sample], Fd)[0]
    print(f"FMM 4M {profile}: RMS p=4 {err[4]:.3e}, p=6 {err[6]:.3e}")
    assert err[4] < 5e-4 and err[6] < err[4] / 5, err


# ----------------------------------------------------- Parareal and ensemble
def test_c5_parareal_64_windows_to_1e20():
    n, nw, tf, eps = 10000, 64, 4.0, 1e-2
    m, q, v = ics.plummer(n, seed=0, velocities=True)
    p = v * m[:, None]
    coarse = {"force": "bunched", "method": 10, "theta": 0.9, "steps": 6, "scheme": "euler",
              "precision": "mixed"}
    fine = {"force": "bunched", "method": 11, "theta": 0.5, "steps": 64, "precision": "mixed"}
    q1, p1, info = I.integrate_parareal(m, q, p, tf, nw, nw + 1, 1.0, 1.0, eps, 1e-20, 0,
                                        coarse=coarse, fine=fine)
    assert info["errors"][-1] <= 1e-20 and info["iterations"] <= nw + 1, info["iterations"]
    t = OctreeType().pre_init(m, 11, 64, precision="mixed")
    t.set_state(q, p)
    for _ in range(nw - 1):
        t.integrate(tf / nw / 64, 64, 0.25, 1.0, eps * eps, force="bunched")
    qs, ps = t.get_state()
    dq, dp = np.linalg.norm(q1[-1] - qs), np.linalg.norm(p1[-1] - ps)
    err = float(np.sqrt(dq * dq + dp * dp) / (6 * n))
    print(f"Parareal: {info['iterations']} iterates, ||par - seq|| / 6N = {err:.2e}")
    assert err <= 1e-20


def test_c5_ensemble_256_systems():
    rng = np.random.default_rng(0)
    sizes = [int(v) for v in rng.integers(500, 20001, 256)]
    sys_ = [ics.plummer(k, seed=1000 + i, velocities=True) for i, k in enumerate(sizes)]
    m = np.concatenate([a[0] for a in sys_])
    q = np.concatenate([a[1] for a in sys_])
    p = np.concatenate([a[2] * a[0][:, None] for a in sys_])
    t = OctreeType().pre_init(m, 11, 64, precision="mixed", counts=sizes)
    t.set_state(q, p)
    e0 = t.energy(1.0, 1e-4)
    t.integrate(1e-3, 8, 0.25, 1.0, 1e-4, force="bunched")
    e1 = t.energy(1.0, 1e-4)
    drift = float(np.max(np.abs((e1[:, 0] - e0[:, 0]) / e0[:, 0])))
    print(f"ensemble of 256 ({sum(sizes)} bodies): max energy drift {drift:.2e} over 8 steps")
    assert drift < 1e-5
    # two systems against their own handles, bit for bit
    off = np.concatenate([[0], np.cumsum(sizes)])
    qb, pb = t.get_state()
    for k in (0, 255):
        ta = OctreeType().pre_init(sys_[k][0], 11, 64, precision="mixed")
        ta.set_state(sys_[k][1], sys_[k][2] * sys_[k][0][:, None])
        ta.integrate(1e-3, 8, 0.25, 1.0, 1e-4, force="bunched")
        qa, pa = ta.get_state()
        assert np.array_equal(qb[off[k]:off[k + 1]], qa) and np.array_equal(pb[off[k]:off[k + 1]], pa)
