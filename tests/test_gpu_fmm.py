"""GPU: the Cartesian FMM of order 4 and 6 (no reference counterpart; the
reference's Dehnen method stops at quadrupole sources).

Pinned by what the mathematics fixes: convergence to the FP64 direct sum as
the order rises and as theta falls, agreement of the matrix-core (FP32) M2L
with the scalar FP64 M2L to FP32 round-off, and agreement of a batch handle
with each system alone.
"""
import numpy as np
import pytest

import oracle as O
from hierarquicos_amd import HQError, OctreeType, direct_forces, ics

pytestmark = pytest.mark.gpu


def _case(profile, n, seed=3):
    m, q, _ = ics.make(profile, n, seed=seed)
    x, y, z = q.T.copy()
    Fd = direct_forces(m, x, y, z, 1.0, 0.0)
    t = OctreeType().pre_init(m, 10).init(x, y, z)
    return m, q, Fd, t


@pytest.mark.parametrize("profile", ["uniform", "plummer"])
def test_fmm_converges_with_order_and_theta(profile):
    m, q, Fd, t = _case(profile, 20000)
    err = {}
    for th2 in (0.25, 0.09):
        for order in (4, 6):
            err[th2, order] = O.rel_errors(t.fmm(th2, 1.0, 0.0, order, "fp64").T, Fd)[0]
    assert err[0.25, 4] < 1e-3, err
    assert err[0.25, 6] < err[0.25, 4] / 5, err
    assert err[0.09, 4] < err[0.25, 4] / 3, err
    assert err[0.09, 6] < err[0.09, 4] / 5, err
    assert err[0.09, 6] < 1e-4, err


@pytest.mark.parametrize("order", [4, 6])
def test_fmm_mfma_matches_fp64(order):
    m, q, Fd, t = _case("uniform", 20000, seed=4)
    F64 = t.fmm(0.25, 1.0, 0.0, order, "fp64").T
    F32 = t.fmm(0.25, 1.0, 0.0, order, "mfma").T
    diff = np.linalg.norm(F32 - F64, axis=1) / np.linalg.norm(F64, axis=1)
    assert float(np.sqrt(np.mean(diff ** 2))) < 5e-6
    e64, e32 = O.rel_errors(F64, Fd)[0], O.rel_errors(F32, Fd)[0]
    assert e32 < 1.1 * e64 + 2e-6, (e32, e64)


def test_fmm_softened_near_field_and_small_n():
    for n in (1, 2, 17, 300):
        m, q, _ = ics.plummer(n, seed=n)
        x, y, z = q.T.copy()
        t = OctreeType().pre_init(m, 11).init(x, y, z)
        F = t.fmm(0.25, 1.0, 1e-4, 6, "fp64").T
        Fd = O.direct_targets(m, x, y, z, np.arange(n), 1.0, 1e-4)
        if n == 1:
            assert np.all(F == 0.0)
        else:
            assert O.rel_errors(F, Fd)[0] < 1e-3


def test_fmm_batch_matches_single_systems():
    sizes = [3000, 50, 8000]
    sys_ = [ics.plummer(k, seed=30 + i)[:2] for i, k in enumerate(sizes)]
    m = np.concatenate([s[0] for s in sys_])
    q = np.concatenate([s[1] + 20.0 * i for i, s in enumerate(sys_)])
    tb = OctreeType().pre_init(m, 10, counts=sizes).init(q[:, 0], q[:, 1], q[:, 2])
    Fb = tb.fmm(0.25, 1.0, 1e-6, 4, "fp64").T
    off = np.concatenate([[0], np.cumsum(sizes)])
    for k, (mk, qk) in enumerate(sys_):
        qk = qk + 20.0 * k
        ta = OctreeType().pre_init(mk, 10).init(qk[:, 0], qk[:, 1], qk[:, 2])
        Fa = ta.fmm(0.25, 1.0, 1e-6, 4, "fp64").T
        d = np.linalg.norm(Fb[off[k]:off[k + 1]] - Fa, axis=1) / np.linalg.norm(Fa, axis=1)
        assert float(d.max()) < 1e-10, k


def test_fmm_rejects_bad_arguments():
    m, q, _ = ics.plummer(100, seed=1)
    t = OctreeType().pre_init(m, 10).init(*q.T.copy())
    with pytest.raises(HQError):
        t.fmm(0.25, 1.0, 0.0, 5)
    with pytest.raises(HQError):
        t.fmm(1.0, 1.0, 0.0, 4)


@pytest.mark.parametrize("window", [1, 3])
def test_fmm_binning_outside_the_lds_window(window):
    """Lattice pairs whose target group lies beyond the tile's LDS window of
    group counts take global atomics; with a window of one or three groups
    most of them do.  The bins then hold the same pairs in another order, so
    the forces equal the default's to FP32 round-off of the M2L sums."""
    m, q, Fd, t = _case("plummer", 20000, seed=9)
    F0 = t.fmm(0.25, 1.0, 0.0, 4, "mfma").T
    assert t.get_tune("fmm_bin_window") == 4096
    t.tune(fmm_bin_window=window)
    assert t.get_tune("fmm_bin_window") == window
    F1 = t.fmm(0.25, 1.0, 0.0, 4, "mfma").T
    diff = np.linalg.norm(F1 - F0, axis=1) / np.linalg.norm(F0, axis=1)
    assert float(np.sqrt(np.mean(diff ** 2))) < 1e-6 and float(diff.max()) < 1e-4, float(diff.max())
    assert abs(O.rel_errors(F1, Fd)[0] - O.rel_errors(F0, Fd)[0]) < 1e-6
    with pytest.raises(HQError):
        t.tune(fmm_bin_window=0)
