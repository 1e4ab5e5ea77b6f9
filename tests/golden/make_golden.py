"""Write the golden fixtures under tests/golden/ from the reference's own Fortran.

Runs only in the build container (it needs the reference sources): builds
oracle/_ref/libhqref.so (the reference's fortran/octree_default.f90,
fortran/direct_forces.f90 and fortran/morton/morton_octree.f90 compiled with
AMD flang, plus oracle/ref_shim.f90) and stores inputs and outputs as
compressed .npz files with a manifest.  Inputs come from our own seeded
generator (hierarquicos_amd.ics); the fixtures hold the inputs themselves, so
no generator stream has to be reproduced to use them.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
from hierarquicos_amd import ics  # noqa: E402

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_i64 = C.POINTER(C.c_int64)


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def load_ref():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    R = C.CDLL(os.path.join(REPO, "oracle", "_ref", "libhqref.so"))
    R.ref_direct.argtypes = [C.c_int, _d, _d, _d, _d, C.c_double, C.c_double, _d]
    R.ref_tree_build.argtypes = [C.c_int, _d, _d, _d, _d, C.c_int, C.c_int]
    R.ref_tree_build.restype = C.c_int
    R.ref_tree_dump.argtypes = [_d] * 9 + [_i] * 4 + [_d, _d]
    R.ref_tree_bh.argtypes = [C.c_double] * 3 + [_d]
    R.ref_tree_groups.argtypes = [C.c_double] * 3 + [_d, _i]
    R.ref_tree_dehnen.argtypes = [C.c_double] * 3 + [_d]
    R.ref_morton_build.argtypes = [C.c_int, _d, _d, C.c_int, C.c_double, C.c_int, _i]
    R.ref_morton_build.restype = C.c_int
    R.ref_morton_dump.argtypes = [_i, _i64, _i, _i, _i, _i, _i, _i, _i, _d, _d, _d, _d]
    R.ref_morton_forces.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _d]
    return R


def tree_dump(R, nn, with_oct=True):
    f = {k: np.zeros(nn) for k in ("cx", "cy", "cz", "half", "L2", "mass", "qx", "qy", "qz")}
    i = {k: np.zeros(nn, np.int32) for k in ("particle", "type", "depth")}
    child = np.zeros((nn, 8), np.int32)
    quad = np.zeros((nn, 6))
    oct_ = np.zeros((nn, 14))
    R.ref_tree_dump(*[_p(f[k]) for k in ("cx", "cy", "cz", "half", "L2", "mass", "qx", "qy",
                                          "qz")],
                    *[_p(i[k], _i) for k in ("particle", "type", "depth")], _p(child, _i),
                    _p(quad), _p(oct_))
    out = dict(f)
    out.update(i)
    out["child"] = child
    out["quad"] = quad
    out["oct"] = oct_
    return out


def canonical_keys(dump):
    """(depth, khi, klo) per reference node: the octant path from the root with
    digit 4*xb+2*yb+zb, xb = x>=cx, yb = y>=cy, zb = z>=cz.  The reference's
    child slot r (0-based) has bit0 = x>=cx, bit1 = y<cy, bit2 = z<cz
    (octree_default.f90:527-531)."""
    nn = len(dump["depth"])
    khi = np.zeros(nn, np.uint64)
    klo = np.zeros(nn, np.uint32)
    for k in range(nn):
        d = int(dump["depth"][k])
        for r in range(8):
            c = int(dump["child"][k, r])
            if c < 0:
                continue
            digit = (4 if r & 1 else 0) | (0 if r & 2 else 2) | (0 if r & 4 else 1)
            dc = d + 1
            hi, lo = int(khi[k]), int(klo[k])
            if dc <= 21:
                hi |= digit << (3 * (21 - dc))
            else:
                lo |= digit << (3 * (25 - dc))
            khi[c] = np.uint64(hi)
            klo[c] = np.uint32(lo)
    return khi, klo


def run_methods(R, m, x, y, z, theta2, eps2, G=1.0, methods=(10, 11, 12, 20, 21, 22, 30, 31, 32)):
    out = {}
    N = len(m)
    for method in methods:
        nn = R.ref_tree_build(N, _p(m), _p(x), _p(y), _p(z), method, 64)
        F = np.zeros((N, 3))
        if method < 20:
            R.ref_tree_bh(theta2, G, eps2, _p(F))
        elif method < 30:
            R.ref_tree_dehnen(theta2, eps2, G, _p(F))
        else:
            ng = np.zeros(1, np.int32)
            R.ref_tree_groups(theta2, G, eps2, _p(F), _p(ng, _i))
            out[f"ngroups_{method}"] = np.int64(ng[0])
        out[f"F_{method}"] = F
        out[f"nnodes_{method}"] = np.int64(nn)
    return out


def morton_dump(R, m, q, L, use_quad=True):
    N = len(m)
    nn = np.zeros(1, np.int32)
    qq = np.ascontiguousarray(q)
    Lr = R.ref_morton_build(N, _p(m), _p(qq), L, 1.2, int(use_quad), _p(nn, _i))
    nn = int(nn[0])
    o = dict(perm=np.zeros(N, np.int32), prefix=np.zeros(nn, np.int64))
    for k in ("level", "first", "last", "parent", "nchild", "next"):
        o[k] = np.zeros(nn, np.int32)
    o["child"] = np.zeros((nn, 8), np.int32)
    o["size2"] = np.zeros(nn)
    o["mass"] = np.zeros(nn)
    o["qcm"] = np.zeros((nn, 3))
    o["quad"] = np.zeros((nn, 6))
    R.ref_morton_dump(_p(o["perm"], _i), _p(o["prefix"], _i64),
                      *[_p(o[k], _i) for k in ("level", "first", "last", "parent", "nchild")],
                      _p(o["child"], _i), _p(o["next"], _i), _p(o["size2"]), _p(o["mass"]),
                      _p(o["qcm"]), _p(o["quad"]))
    # sorted keys = prefixes of the level-L node holding each sorted particle
    keys = np.zeros(N, np.int64)
    leaf = np.nonzero(o["level"] == Lr)[0]
    for k in leaf:
        keys[o["first"][k]:o["last"][k] + 1] = o["prefix"][k]
    o["keys_sorted"] = keys
    o["L"] = np.int64(Lr)
    return o


def main():
    R = load_ref()
    manifest = {"generator": "numpy.random.default_rng (PCG64) via hierarquicos_amd.ics",
                "reference_build": "AMD flang 22, -O2 -ffp-contract=off (oracle/Makefile ref)",
                "files": {}}

    # 1. C1-like Plummer sphere N=1000, eps=0.01, theta=0.5, all methods.
    N, seed = 1000, 3
    m, q, _ = ics.plummer(N, a=1.0, rmax=10.0, seed=seed)
    x, y, z = (np.ascontiguousarray(q[:, k]) for k in range(3))
    theta2, eps2 = 0.25, 1e-4
    d = dict(m=m, x=x, y=y, z=z, theta2=theta2, eps2=eps2)
    Fd = np.zeros((N, 3))
    R.ref_direct(N, _p(m), _p(x), _p(y), _p(z), 1.0, eps2, _p(Fd))
    d["F_direct"] = Fd
    d.update(run_methods(R, m, x, y, z, theta2, eps2))
    nn = R.ref_tree_build(N, _p(m), _p(x), _p(y), _p(z), 12, 64)
    dump = tree_dump(R, nn)
    khi, klo = canonical_keys(dump)
    for k, v in dump.items():
        d["tree12_" + k] = v
    d["tree12_khi"], d["tree12_klo"] = khi, klo
    mo = morton_dump(R, m, q, 0)
    d["morton_L"] = mo["L"]
    d["morton_perm"] = mo["perm"]
    d["morton_keys_sorted"] = mo["keys_sorted"]
    d["morton_nnodes"] = np.int64(len(mo["level"]))
    Fm = np.zeros((N, 3))
    R.ref_morton_forces(1.0, eps2, theta2, 1, _p(Fm))
    d["F_morton"] = Fm
    fn = "plummer_n1000_seed3.npz"
    np.savez_compressed(os.path.join(HERE, fn), **d)
    manifest["files"][fn] = dict(profile="plummer", N=N, seed=seed, a=1.0, rmax=10.0,
                                 theta2=theta2, eps2=eps2, G=1.0,
                                 contents="direct, methods 10-12/20-22/30-32, method-12 node "
                                          "store, Morton (L=0) permutation/keys/forces")

    # 2. The 10-point set of fortran/morton/test_morton.f90:30-39 (z = 0, m = 1).
    pts = np.array([[0.45931272, 0.22315240], [0.40157369, 0.21211072],
                    [0.70372114, 0.64428738], [0.20444018, 0.68014233],
                    [0.73388286, 0.59264995], [0.68856885, 0.48249009],
                    [0.48843330, 0.28037004], [0.15927911, 0.69084037],
                    [0.30486200, 0.50604824], [0.59676033, 0.80020091]])
    q10 = np.zeros((10, 3))
    q10[:, :2] = pts
    m10 = np.ones(10)
    mo = morton_dump(R, m10, q10, 0)
    Fm = np.zeros((10, 3))
    R.ref_morton_forces(1.0, 1e-4, 0.01, 1, _p(Fm))
    d = {"m": m10, "q": q10, "F_morton": Fm}
    for k, v in mo.items():
        d["morton_" + k] = v
    x, y, z = (np.ascontiguousarray(q10[:, k]) for k in range(3))
    d.update(run_methods(R, m10, x, y, z, 0.01, 1e-4, methods=(11, 12)))
    fn = "points10.npz"
    np.savez_compressed(os.path.join(HERE, fn), **d)
    manifest["files"][fn] = dict(source="fortran/morton/test_morton.f90:30-39", theta2=0.01,
                                 eps2=1e-4, contents="Morton tree (L=0) node store + forces; "
                                                     "methods 11/12")

    # 3. Deeper, uneven tree: Hernquist N=4096 (cusp), theta=0.6, eps=5e-4.
    N, seed = 4096, 5
    m, q, _ = ics.hernquist(N, a=1.0, rmax=50.0, seed=seed)
    x, y, z = (np.ascontiguousarray(q[:, k]) for k in range(3))
    theta2, eps2 = 0.36, 2.5e-7
    d = dict(m=m, x=x, y=y, z=z, theta2=theta2, eps2=eps2)
    Fd = np.zeros((N, 3))
    R.ref_direct(N, _p(m), _p(x), _p(y), _p(z), 1.0, eps2, _p(Fd))
    d["F_direct"] = Fd
    d.update(run_methods(R, m, x, y, z, theta2, eps2, methods=(11, 12, 21, 31, 32)))
    nn = R.ref_tree_build(N, _p(m), _p(x), _p(y), _p(z), 11, 64)
    dump = tree_dump(R, nn)
    khi, klo = canonical_keys(dump)
    d.update(tree11_depth=dump["depth"], tree11_type=dump["type"],
             tree11_particle=dump["particle"], tree11_khi=khi, tree11_klo=klo,
             tree11_mass=dump["mass"])
    fn = "hernquist_n4096_seed5.npz"
    np.savez_compressed(os.path.join(HERE, fn), **d)
    manifest["files"][fn] = dict(profile="hernquist", N=N, seed=seed, a=1.0, rmax=50.0,
                                 theta2=theta2, eps2=eps2,
                                 contents="direct, methods 11/12/21/31/32, method-11 canonical "
                                          "topology")

    # 4. Edge cases: two particles; a coincident pair (reaches max depth 25 and
    #    is dropped as a source, octree_default.f90:570-579).
    d = {}
    m2 = np.array([0.25, 0.75])
    x2, y2, z2 = np.array([0.1, -0.3]), np.array([0.2, 0.4]), np.array([-0.5, 0.7])
    d.update({"two_m": m2, "two_x": x2, "two_y": y2, "two_z": z2})
    for method in (10, 11, 12):
        R.ref_tree_build(2, _p(m2), _p(x2), _p(y2), _p(z2), method, 64)
        F = np.zeros((2, 3))
        R.ref_tree_bh(0.25, 1.0, 1e-4, _p(F))
        d[f"two_F_{method}"] = F
    rng = np.random.default_rng(11)
    mc = np.full(64, 1.0 / 64)
    qc = rng.standard_normal((64, 3))
    qc[17] = qc[40]  # exact duplicate
    xc, yc, zc = (np.ascontiguousarray(qc[:, k]) for k in range(3))
    d.update({"dup_m": mc, "dup_x": xc, "dup_y": yc, "dup_z": zc})
    for method in (10, 11, 12):
        nn = R.ref_tree_build(64, _p(mc), _p(xc), _p(yc), _p(zc), method, 64)
        F = np.zeros((64, 3))
        R.ref_tree_bh(0.25, 1.0, 1e-4, _p(F))
        d[f"dup_F_{method}"] = F
        d[f"dup_nnodes_{method}"] = np.int64(nn)
    fn = "edge_cases.npz"
    np.savez_compressed(os.path.join(HERE, fn), **d)
    manifest["files"][fn] = dict(contents="N=2 BH 10/11/12; N=64 with an exact duplicate "
                                          "(max-depth loss) BH 10/11/12", theta2=0.25,
                                 eps2=1e-4)

    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
    for fn in manifest["files"]:
        print(fn, os.path.getsize(os.path.join(HERE, fn)))


if __name__ == "__main__":
    main()
