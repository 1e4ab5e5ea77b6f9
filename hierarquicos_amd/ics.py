"""Seeded initial conditions (host side, numpy PCG64).

Restates the samplers of initial_values/initial_values.f90 with our own
seeded generator (the reference draws from the Fortran runtime's
RANDOM_NUMBER, whose stream is not reproducible here):

* radius by inverse CDF with rejection beyond ``rmax``
  (initial_values.f90:220-235; plummer.f90:53-64; hernquist.f90:53-66);
* isotropic directions, phi = 2 pi X, cos(theta) = 2X - 1
  (initial_values.f90:355-381);
* Plummer speeds by von Neumann rejection on q^2 (1-q^2)^3.5 with the
  reference's bound 0.0922108113291 (initial_values.f90:237-257,
  plummer.f90:93-104);
* Hernquist speeds by rejection on q^2 f(E(q)) with the isotropic
  distribution function of hernquist.f90:100-111;
* equal masses 1/N, total momentum removed, centre of mass moved to the
  origin (initial_values.f90:135-172).

All positions are returned as an (N, 3) float64 array together with masses
(N,) and, when requested, velocities (N, 3).
"""
from __future__ import annotations

import numpy as np

_PLUMMER_PDF_MAX = 0.0922108113291


def _directions(rng, n):
    phi = 2.0 * np.pi * rng.random(n)
    cth = 2.0 * rng.random(n) - 1.0
    sth = np.sqrt(1.0 - cth * cth)
    return np.stack([sth * np.cos(phi), sth * np.sin(phi), cth], axis=1)


def _radii(rng, n, inverse_cdf, rmax):
    out = np.empty(n)
    filled = 0
    while filled < n:
        k = max(1024, int((n - filled) * 1.2))
        r = inverse_cdf(rng.random(k))
        if rmax is not None and rmax >= 0:
            r = r[r <= rmax]
        r = r[np.isfinite(r)]
        take = min(len(r), n - filled)
        out[filled:filled + take] = r[:take]
        filled += take
    return out


def _finish(m, q, v):
    if v is not None:
        p = v * m[:, None]
        p -= p.sum(axis=0) / len(m)
        v = p / m[:, None]
    com = (m[:, None] * q).sum(axis=0) / m.sum()
    q = q - com
    return m, q, v


def plummer_inverse_cdf(x, a=1.0):
    """r with M(r) / M_0 = x for the Plummer sphere (plummer.f90:53-62)."""
    return a * x ** (1.0 / 3.0) / np.sqrt(1.0 - x ** (2.0 / 3.0))


def hernquist_inverse_cdf(x, a=1.0):
    """r with M(r) / M_0 = x for the Hernquist profile (hernquist.f90:53-66)."""
    s = np.sqrt(x)
    return a * s / (1.0 - s)


def _plummer_speed_ratios(rng, n):
    """q = v / v_esc by von Neumann rejection on q^2 (1 - q^2)^3.5 under the
    reference's bound (initial_values.f90:237-257)."""
    qv = np.empty(n)
    todo = np.arange(n)
    while len(todo):
        x = rng.random(len(todo))
        y = _PLUMMER_PDF_MAX * rng.random(len(todo))
        ok = y < x * x * (1.0 - x * x) ** 3.5
        qv[todo[ok]] = x[ok]
        todo = todo[~ok]
    return qv


def plummer(n, a=1.0, rmax=10.0, seed=0, G=1.0, velocities=False):
    """Plummer sphere of scale radius ``a``, truncated at ``rmax``."""
    rng = np.random.default_rng(seed)
    m = np.full(n, 1.0 / n)
    r = _radii(rng, n, lambda x: plummer_inverse_cdf(x, a), rmax)
    q = r[:, None] * _directions(rng, n)
    v = None
    if velocities:
        vesc = np.sqrt(2.0 * G / np.sqrt(r * r + a * a))
        qv = _plummer_speed_ratios(rng, n)
        v = (qv * vesc)[:, None] * _directions(rng, n)
    return _finish(m, q, v)


def _hernquist_edf(r, x, a):
    """hernquist_edf_norm of hernquist.f90:95-107 as written there (G = M = 1):
    its E is sqrt(-Phi (1 - x^2)), and the second term of the bracket is
    scaled by the first; both differ from Hernquist's (1990) closed form, and
    the sampler follows the reference."""
    e = np.sqrt((1.0 / (a + r)) * (1.0 - x * x))
    f = 1.0 / (8.0 * np.sqrt(2.0) * a * np.pi ** 3)
    f = f / np.sqrt(1.0 - e) ** 5
    f = f * 3.0 * np.arcsin(np.sqrt(e))
    return f + f * np.sqrt(e * (1.0 - e)) * (1.0 - 2.0 * e) * (8.0 * e * e - 8.0 * e - 3.0)


def hernquist(n, a=1.0, rmax=50.0, seed=0, G=1.0, velocities=False):
    """Hernquist profile (cusp rho ~ 1/r), truncated at ``rmax``."""
    rng = np.random.default_rng(seed)
    m = np.full(n, 1.0 / n)
    r = _radii(rng, n, lambda x: hernquist_inverse_cdf(x, a), rmax)
    q = r[:, None] * _directions(rng, n)
    v = None
    if velocities:
        # bound of q^2 f(E) per radius from a tabulated grid (5 % margin)
        qs = np.linspace(1e-3, 1.0 - 1e-3, 512)
        rg = np.geomspace(max(r.min(), 1e-8), r.max(), 256)
        gmax = np.array([np.max(qs * qs * _hernquist_edf(rr, qs, a)) for rr in rg]) * 1.05
        bound = np.interp(np.log(r), np.log(rg), gmax)
        vesc = np.sqrt(2.0 * G / (a + r))
        qv = np.empty(n)
        todo = np.arange(n)
        while len(todo):
            x = rng.random(len(todo))
            y = bound[todo] * rng.random(len(todo))
            ok = y < x * x * _hernquist_edf(r[todo], x, a)
            qv[todo[ok]] = x[ok]
            todo = todo[~ok]
        v = (qv * vesc)[:, None] * _directions(rng, n)
    return _finish(m, q, v)


def uniform_cube(n, half=1.0, seed=0, vmax=0.0, velocities=False):
    """Uniform cube [-half, half]^3 (initial_values.f90:177-198)."""
    rng = np.random.default_rng(seed)
    m = np.full(n, 1.0 / n)
    q = half * (2.0 * rng.random((n, 3)) - 1.0)
    v = vmax * (2.0 * rng.random((n, 3)) - 1.0) if velocities else None
    return _finish(m, q, v)


def homogeneous_sphere(n, R=1.0, seed=0):
    """Homogeneous sphere of radius R (homogeneous.f90:48-61)."""
    rng = np.random.default_rng(seed)
    m = np.full(n, 1.0 / n)
    r = R * rng.random(n) ** (1.0 / 3.0)
    q = r[:, None] * _directions(rng, n)
    return _finish(m, q, None)


def make(profile, n, seed=0, **kw):
    """Dispatch by profile name as generate_initial_values does (:84-132)."""
    fn = {"plummer": plummer, "hernquist": hernquist, "uniform": uniform_cube,
          "homogeneous": homogeneous_sphere}.get(profile)
    if fn is None:
        raise ValueError(f'Unknown density profile: "{profile}"')
    return fn(n, seed=seed, **kw)
