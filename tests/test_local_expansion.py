"""CPU checks of the far-field local expansion of the mixed-precision walk
(hierarquicos_amd/csrc/hq_local.h, generated by gen_local.py).

The generator also renders the same straight-line code as Python; here it runs
in float64 on a random cluster of bodies, and its acceleration near a distant
bunch centre is compared with the direct sum over the cluster's bodies -- an
independent check of every sign, index and scale factor.  The remaining
difference is the truncation of the two expansions, whose size and scaling
with distance are asserted too."""
import math
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "hierarquicos_amd", "csrc"))
import gen_local  # noqa: E402

_NS = {}
for _p in gen_local.ORDERS:
    _NS[_p] = {}
    exec(gen_local.python_source(_p), _NS[_p])


def fma(a, b, c):
    return a * b + c


def rsqrt(x):
    return 1.0 / math.sqrt(x)


def record(m, p):
    """A cell's moments about its centre of mass in the FP32 record layout of
    the walk (r[0] M, r[1..6] -3Q, r[7] -1.5 trQ, r[8..10] -1.5 T, r[11..20]
    7.5 O, r[21..23] 15 (O10, O8, O6)), here in float64."""
    M = m.sum()
    com = (m[:, None] * p).sum(0) / M
    x, y, z = (p - com).T
    Q = [(m * x * x).sum(), (m * y * y).sum(), (m * z * z).sum(), (m * x * y).sum(),
         (m * x * z).sum(), (m * y * z).sum()]
    O = [(m * x * x * x).sum(), (m * x * y * y).sum(), (m * x * z * z).sum(),
         (m * y * x * x).sum(), (m * y * y * y).sum(), (m * y * z * z).sum(),
         (m * z * x * x).sum(), (m * z * y * y).sum(), (m * z * z * z).sum(),
         (m * x * y * z).sum()]
    T = [O[0] + O[1] + O[2], O[3] + O[4] + O[5], O[6] + O[7] + O[8]]
    r = [M] + [-3.0 * q for q in Q] + [-1.5 * (Q[0] + Q[1] + Q[2])] + [-1.5 * t for t in T]
    r += [7.5 * o for o in O] + [15.0 * O[9], 15.0 * O[7], 15.0 * O[5]]
    return np.array(r), com


def local_acc(MP, r, com, c, h, eps2, targets, p=5):
    L = [0.0] * len(gen_local.orders(p)[4])
    R = c - com
    _NS[p]["m2l"](MP, R[0], R[1], R[2], eps2, h, r, L, fma, rsqrt)
    out = []
    for t in targets:
        u = (t - c) / h
        acc = [0.0, 0.0, 0.0]
        _NS[p]["l2p"](L, u[0], u[1], u[2], acc, fma)
        out.append([-a / h for a in acc])
    return np.array(out)


def direct(m, p, targets, eps2):
    d = p[None, :, :] - targets[:, None, :]
    s2 = (d * d).sum(2) + eps2
    return (m[None, :, None] * d / s2[:, :, None] ** 1.5).sum(1)


def _case(seed, a, dist, h, eps2, MP, nt=40, p=5):
    rng = np.random.default_rng(seed)
    m = rng.uniform(0.5, 1.5, 25)
    pos = rng.normal(size=(25, 3)) * a
    r, com = record(m, pos)
    u = rng.normal(size=3)
    c = com + dist * u / np.linalg.norm(u)
    v = rng.normal(size=(nt, 3))
    targets = c + h * rng.uniform(0, 1, (nt, 1)) * v / np.linalg.norm(v, axis=1)[:, None]
    F = local_acc(MP, r, com, c, h, eps2, targets, p)
    Fd = direct(m, pos, targets, eps2)
    return np.max(np.linalg.norm(F - Fd, axis=1) / np.linalg.norm(Fd, axis=1))


@pytest.mark.parametrize("p", [5, 6])
@pytest.mark.parametrize("eps2", [0.0, 0.25])
def test_octupole_local_expansion_matches_direct_sum(eps2, p):
    # cluster radius 0.02 and targets within 0.1 of a centre at distance 1:
    # octupole truncation ~ 0.02^4, local truncation ~ 0.1^5 (p5) or 0.1^6 (p6)
    for seed in range(5):
        assert _case(seed, 0.02, 1.0, 0.1, eps2, 8, p=p) < 3e-5


@pytest.mark.parametrize("p", [5, 6])
@pytest.mark.parametrize("MP,bound", [(1, 2e-3), (4, 1e-4), (8, 3e-5)])
def test_lower_orders_truncate_as_expected(MP, bound, p):
    errs = [_case(seed, 0.02, 1.0, 0.1, 1e-3, MP, p=p) for seed in range(5)]
    assert max(errs) < bound, errs


@pytest.mark.parametrize("p", [5, 6])
def test_local_truncation_falls_with_distance(p):
    # a point-like source: only the local expansion truncates; the error falls
    # at least as (h/R)^p (potential order p from the monopole term)
    e1 = _case(7, 1e-6, 1.0, 0.3, 0.0, 8, p=p)
    e2 = _case(7, 1e-6, 2.0, 0.3, 0.0, 8, p=p)
    assert e1 < 1e-2 and e2 < e1 / 2 ** (p - 0.5), (e1, e2)


def test_order_six_is_the_more_accurate():
    # the same far configuration: the p6 expansion truncates less than p5
    e5 = max(_case(s, 0.05, 1.0, 0.3, 1e-3, 8, p=5) for s in range(4))
    e6 = max(_case(s, 0.05, 1.0, 0.3, 1e-3, 8, p=6) for s in range(4))
    assert e6 < 0.5 * e5, (e5, e6)


@pytest.mark.parametrize("order", [5, 6])
def test_exact_at_the_expansion_centre(order):
    # at the bunch centre only the multipole truncation remains: the local
    # expansion adds nothing to the cell's field there
    rng = np.random.default_rng(3)
    m = rng.uniform(0.5, 1.5, 25)
    p = rng.normal(size=(25, 3)) * 0.05
    r, com = record(m, p)
    c = com + np.array([0.6, -0.5, 0.4])
    F = local_acc(8, r, com, c, 0.2, 1e-2, c[None, :], order)
    # the octupole M2P at the same point, from the same moments
    d = com - c
    s2 = d @ d + 1e-2
    M = m.sum()
    x, y, z = (p - com).T
    Qm = np.array([[(m * x * x).sum(), (m * x * y).sum(), (m * x * z).sum()],
                   [(m * x * y).sum(), (m * y * y).sum(), (m * y * z).sum()],
                   [(m * x * z).sum(), (m * y * z).sum(), (m * z * z).sum()]])
    q = p - com
    O3 = np.einsum("j,ja,jb,jc->abc", m, q, q, q)
    T = np.einsum("abb->a", O3)
    v = np.einsum("abc,b,c->a", O3, d, d)
    acc = d * (M / s2 ** 1.5 - 1.5 * np.trace(Qm) / s2 ** 2.5 + 7.5 * (d @ Qm @ d + d @ T) / s2 ** 3.5
               - 17.5 * (d @ v) / s2 ** 4.5) + (-3 * Qm @ d - 1.5 * T) / s2 ** 2.5 + 7.5 * v / s2 ** 3.5
    assert np.allclose(F[0], acc, rtol=1e-12, atol=0)
