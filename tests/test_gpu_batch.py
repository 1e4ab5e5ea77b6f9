"""GPU: batches of independent systems in one tree handle.

A batch handle holds one tree per system (the ensemble and Parareal drivers
of the reference keep one OctreeType per system, parareal/parareal.f90:28-140).
Bars: every system's FP64 forces bit-identical to a handle built from that
system alone (same tree, same walk, same summation order), and within 1e-10 of
the oracle's reference tree; the mixed bunched walk bit-identical too.  The
Dehnen method accumulates with atomics, so its bar is 1e-12 relative.
"""
import numpy as np
import pytest

import oracle as O
from hierarquicos_amd import HQError, OctreeType, ics

pytestmark = pytest.mark.gpu


def _systems(sizes, seed=0):
    """Plummer spheres of the given sizes, each moved and scaled differently,
    so the root cubes differ from system to system."""
    rng = np.random.default_rng(seed)
    ms, qs = [], []
    for k, n in enumerate(sizes):
        m, q, _ = ics.plummer(n, seed=100 + k)
        q = q * rng.uniform(0.5, 3.0) + rng.uniform(-20, 20, 3)
        ms.append(m * rng.uniform(0.5, 2.0))
        qs.append(q)
    return ms, qs


def _forces(t, method, th2, eps2, how):
    if how == "bunched_mixed":
        return np.ascontiguousarray(t.forces_bunched(th2, 1.0, eps2, "mixed").T)
    if how == "bunched_fp64":
        return np.ascontiguousarray(t.forces_bunched(th2, 1.0, eps2, "fp64").T)
    if method < 20:
        return np.ascontiguousarray(t.forces_all(th2, 1.0, eps2).T)
    if method < 30:
        return np.ascontiguousarray(t.dehnen_eval(th2, eps2, 1.0).T)
    return np.ascontiguousarray(t.forces_groups(th2, 1.0, eps2).T)


def _batch_vs_alone(sizes, method, how="walk", th2=0.36, eps2=1e-6, gs=64):
    ms, qs = _systems(sizes)
    m = np.concatenate(ms)
    q = np.concatenate(qs)
    tb = OctreeType().pre_init(m, method, gs, counts=sizes).init(q[:, 0], q[:, 1], q[:, 2])
    Fb = _forces(tb, method, th2, eps2, how)
    off = np.concatenate([[0], np.cumsum(sizes)])
    alone = []
    for k in range(len(sizes)):
        t = OctreeType().pre_init(ms[k], method, gs).init(qs[k][:, 0], qs[k][:, 1], qs[k][:, 2])
        alone.append(_forces(t, method, th2, eps2, how))
    return ms, qs, off, Fb, alone, tb


SIZES = [1, 2, 7, 300, 1000, 2500, 4096]


@pytest.mark.parametrize("method", [10, 11, 12, 30, 31, 32])
def test_batch_fp64_bitidentical_to_single_systems(method):
    ms, qs, off, Fb, alone, tb = _batch_vs_alone(SIZES, method)
    for k in range(len(SIZES)):
        assert np.array_equal(Fb[off[k]:off[k + 1]], alone[k]), (method, k)
    inf = tb.info()
    assert inf["n"] == sum(SIZES)


@pytest.mark.parametrize("how", ["bunched_fp64", "bunched_mixed"])
def test_batch_bunched_bitidentical(how):
    ms, qs, off, Fb, alone, _ = _batch_vs_alone(SIZES, 12, how)
    for k in range(len(SIZES)):
        assert np.array_equal(Fb[off[k]:off[k + 1]], alone[k]), (how, k)


def test_batch_matches_oracle_per_system():
    sizes = [5, 800, 3000]
    ms, qs, off, Fb, _, _ = _batch_vs_alone(sizes, 12, th2=0.25, eps2=1e-4)
    for k in range(len(sizes)):
        ref = O.Tree(ms[k], 12).init(qs[k][:, 0], qs[k][:, 1], qs[k][:, 2]).bh_forces(
            0.25, 1.0, 1e-4)
        r = np.linalg.norm(Fb[off[k]:off[k + 1]] - ref, axis=1)
        den = np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
        if sizes[k] > 1:
            assert float(np.max(r / den)) < 1e-10, k


@pytest.mark.parametrize("method", [20, 22])
def test_batch_dehnen_matches_single_systems(method):
    ms, qs, off, Fb, alone, _ = _batch_vs_alone([3, 600, 2000], method, th2=0.5, eps2=1e-4)
    for k in range(3):
        F, A = Fb[off[k]:off[k + 1]], alone[k]
        den = np.maximum(np.linalg.norm(A, axis=1), 1e-300)
        assert float(np.max(np.linalg.norm(F - A, axis=1) / den)) < 1e-12, k


def test_ensemble_of_many_systems():
    """An ensemble like the Parareal study's: many Plummer spheres of mixed
    sizes in one batch; every system bit-identical to its own build."""
    rng = np.random.default_rng(7)
    sizes = [int(s) for s in rng.integers(100, 1500, 64)]
    ms, qs, off, Fb, alone, tb = _batch_vs_alone(sizes, 11, th2=0.25, eps2=1e-6)
    for k in range(len(sizes)):
        assert np.array_equal(Fb[off[k]:off[k + 1]], alone[k]), k


def test_batch_rejects_empty_system():
    m = np.ones(10)
    with pytest.raises(HQError):
        OctreeType().pre_init(m, 12, counts=[5, 0, 5])
    with pytest.raises(ValueError):
        OctreeType().pre_init(m, 12, counts=[5, 4])
