"""The seeded initial-condition samplers (hierarquicos_amd/ics.py) against
what the mathematics fixes: closed-form cumulative masses, the truncation,
isotropy, the Plummer speed law as an incomplete beta function, and the
reference's zeroed centre of mass and momentum (initial_values.f90:135-172).

Kolmogorov-Smirnov statistics are compared with the 0.1 % critical value
1.95 / sqrt(n); the inputs are seeded, so each statistic is a fixed number.
"""
import numpy as np
import pytest
from scipy import special

from hierarquicos_amd import ics

N = 200_000
KS = 1.95 / np.sqrt(N)


def ks_stat(sample, cdf):
    x = np.sort(sample)
    F = cdf(x)
    i = np.arange(1, len(x) + 1)
    return float(max(np.max(i / len(x) - F), np.max(F - (i - 1) / len(x))))


def plummer_mass(r, a=1.0):
    """M(r) / M_0 = r^3 / (r^2 + a^2)^(3/2) (density 3 a^2 / (4 pi (r^2 + a^2)^(5/2)))."""
    return r ** 3 / (r * r + a * a) ** 1.5


def hernquist_mass(r, a=1.0):
    """M(r) / M_0 = r^2 / (r + a)^2 (density a / (2 pi r (r + a)^3))."""
    return r * r / (r + a) ** 2


@pytest.mark.parametrize("a", [1.0, 0.58])
def test_inverse_cdfs_invert_the_enclosed_mass(a):
    x = np.linspace(1e-6, 1 - 1e-6, 2001)
    assert np.allclose(plummer_mass(ics.plummer_inverse_cdf(x, a), a), x, rtol=1e-10, atol=1e-13)
    assert np.allclose(hernquist_mass(ics.hernquist_inverse_cdf(x, a), a), x, rtol=1e-10,
                       atol=1e-13)
    # and the radial pdfs of the reference (plummer.f90:50, hernquist.f90:40-51)
    # are their derivatives
    r = np.geomspace(1e-3, 50.0, 400)
    h = 1e-6 * r
    dP = (plummer_mass(r + h, a) - plummer_mass(r - h, a)) / (2 * h)
    assert np.allclose(dP, 3 * a * a * r * r / (r * r + a * a) ** 2.5, rtol=1e-6)
    dH = (hernquist_mass(r + h, a) - hernquist_mass(r - h, a)) / (2 * h)
    assert np.allclose(dH, 2 * r * a / (r + a) ** 3, rtol=1e-6)


@pytest.mark.parametrize("profile,mass,rmax", [("plummer", plummer_mass, 10.0),
                                               ("hernquist", hernquist_mass, 50.0)])
def test_truncated_radii_follow_the_enclosed_mass(profile, mass, rmax):
    inv = ics.plummer_inverse_cdf if profile == "plummer" else ics.hernquist_inverse_cdf
    r = ics._radii(np.random.default_rng(11), N, inv, rmax)
    assert r.max() <= rmax and r.min() >= 0
    # rejection beyond rmax renormalises the cumulative mass by M(rmax)
    d = ks_stat(r, lambda x: mass(x) / mass(rmax))
    assert d < KS, d
    # the untruncated law is visibly different where truncation matters
    if profile == "hernquist":
        assert ks_stat(r, mass) > 5 * KS


@pytest.mark.parametrize("profile", ["plummer", "hernquist"])
def test_public_samplers_radii_and_truncation(profile):
    m, q, _ = ics.make(profile, N, seed=4)
    rmax, mass = (10.0, plummer_mass) if profile == "plummer" else (50.0, hernquist_mass)
    assert np.all(m == 1.0 / N)
    r = np.linalg.norm(q, axis=1)
    # the centre-of-mass shift moves radii by far less than the KS scale
    assert r.max() <= rmax + 0.05
    d = ks_stat(r, lambda x: np.minimum(mass(x) / mass(rmax), 1.0))
    assert d < 3 * KS, d


def test_directions_are_isotropic():
    u = ics._directions(np.random.default_rng(5), N)
    assert np.allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-14)
    # cos(theta) uniform on [-1, 1], phi uniform on [0, 2 pi)
    assert ks_stat(u[:, 2], lambda c: (c + 1) / 2) < KS
    phi = np.mod(np.arctan2(u[:, 1], u[:, 0]), 2 * np.pi)
    assert ks_stat(phi, lambda p: p / (2 * np.pi)) < KS
    # each axis of an isotropic unit vector is uniform on [-1, 1] (Archimedes)
    for c in (0, 1):
        assert ks_stat(u[:, c], lambda c_: (c_ + 1) / 2) < KS
    assert np.all(np.abs(u.mean(axis=0)) < 5 / np.sqrt(N))


def test_plummer_speed_law():
    # q^2 (1 - q^2)^3.5 dq = (1/2) u^(1/2) (1 - u)^(7/2) du with u = q^2, so the
    # cumulative law of q is the regularised incomplete beta I_{q^2}(3/2, 9/2)
    qv = ics._plummer_speed_ratios(np.random.default_rng(6), N)
    assert qv.min() >= 0 and qv.max() <= 1
    d = ks_stat(qv, lambda q: special.betainc(1.5, 4.5, q * q))
    assert d < KS, d
    # the reference's bound (plummer.f90:100-104) is the maximum, at q^2 = 2/9
    peak = (2.0 / 9.0) * (7.0 / 9.0) ** 3.5
    assert abs(ics._PLUMMER_PDF_MAX - peak) < 1e-12
    g = np.linspace(0, 1, 200001)
    assert np.max(g * g * (1 - g * g) ** 3.5) <= ics._PLUMMER_PDF_MAX


def test_plummer_velocities_scale_with_escape_speed():
    m, q, v = ics.plummer(N, seed=7, velocities=True)
    r = np.linalg.norm(q, axis=1)
    ratio = np.linalg.norm(v, axis=1) / np.sqrt(2.0 / np.sqrt(r * r + 1.0))
    # bound by the escape speed up to the tiny momentum and centre shifts
    assert ratio.max() < 1.01
    d = ks_stat(ratio, lambda s: special.betainc(1.5, 4.5, np.minimum(s, 1.0) ** 2))
    assert d < 3 * KS, d
    # virial ratio of the untruncated Plummer sphere, 2K / |W| = 1, to a few %
    K = 0.5 * np.sum(m * np.sum(v * v, axis=1))
    W = -3 * np.pi / 32  # potential energy of the unit Plummer sphere (G = M = a = 1)
    assert abs(2 * K / abs(W) - 1) < 0.05


def test_hernquist_speed_rejection_bound_dominates():
    """The per-radius bound of the rejection sampler is at least the pdf
    q^2 f(E(q)) everywhere (an unbiased von Neumann step needs that)."""
    m, q, v = ics.hernquist(20_000, seed=8, velocities=True)
    assert np.all(np.isfinite(v))
    r = np.linalg.norm(q, axis=1)
    vesc = np.sqrt(2.0 / (1.0 + r))
    assert np.max(np.linalg.norm(v, axis=1) / vesc) < 1.01
    qs = np.linspace(1e-4, 1 - 1e-4, 4001)
    for rr in np.geomspace(1e-3, 50, 40):
        pdf = qs * qs * ics._hernquist_edf(rr, qs, 1.0)
        coarse = np.linspace(1e-3, 1.0 - 1e-3, 512)
        bound = 1.05 * np.max(coarse * coarse * ics._hernquist_edf(rr, coarse, 1.0))
        assert pdf.max() <= bound


@pytest.mark.parametrize("profile", ["plummer", "hernquist", "uniform"])
def test_centre_of_mass_and_momentum_zeroed(profile):
    m, q, v = ics.make(profile, 50_000, seed=9, velocities=True)
    com = (m[:, None] * q).sum(axis=0) / m.sum()
    P = (m[:, None] * v).sum(axis=0)
    scale_q = np.abs(q).max()
    scale_p = np.abs(m[:, None] * v).sum(axis=0).max()
    assert np.all(np.abs(com) < 1e-13 * scale_q), com
    assert np.all(np.abs(P) < 1e-13 * max(scale_p, 1e-300)), P
