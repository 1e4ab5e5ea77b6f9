"""CPU baseline timing for bench.py -- TEST INFRASTRUCTURE ONLY.

Times the reference's own Barnes-Hut (fortran/octree_default.f90 compiled by
oracle/Makefile into oracle/_ref/libhqref.so, kind "reference"), or the C
restatement in hq_oracle.c when that library is absent (kind "port"), on the
host cores, on a bounded sample of the benchmark's workload:

  * pre_init + init (serial insertion build + evaluate_multipole) of all N
    particles, as benchmarks/test_forces.f90:184-191 times them;
  * bh_forces_over_p on `sample` targets spread over the particles, one
    OpenMP thread per target as the harness loop does (:199-213);

and extrapolates the walk to all N targets:
    particles/s = N / (t_build + t_walk(sample) * N / sample).

Run as a separate process (bench.py starts it before touching the GPU):
    python oracle/cpu_baseline.py --n 1048576 --theta 0.5 --eps 1e-3 --sample 65536
It prints one JSON object.
"""
import argparse
import ctypes as C
import json
import os
import resource
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--eps", type=float, default=1e-3)
    ap.add_argument("--method", type=int, default=12)
    ap.add_argument("--profile", choices=("plummer", "hernquist", "cube"), default="plummer")
    ap.add_argument("--sample", type=int, default=65536)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--verlet-steps", type=int, default=0,
                    help="time this many Verlet steps (build + forces of all N each) instead")
    ap.add_argument("--dt", type=float, default=1e-3)
    args = ap.parse_args()
    if args.verlet_steps:
        return verlet(args)

    from hierarquicos_amd import ics
    if args.profile == "hernquist":
        m, q, _ = ics.hernquist(args.n, a=1.0, rmax=50.0, seed=args.seed)
    elif args.profile == "cube":
        m, q, _ = ics.uniform_cube(args.n, half=1.0, seed=args.seed)
    else:
        m, q, _ = ics.plummer(args.n, a=1.0, rmax=10.0, seed=args.seed)
    x, y, z = (np.ascontiguousarray(q[:, k]) for k in range(3))
    th2, e2 = args.theta ** 2, args.eps ** 2
    sample = np.linspace(0, args.n - 1, min(args.sample, args.n)).astype(np.int64)
    threads = args.threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ref_path = os.path.join(HERE, "_ref", "libhqref.so")
    d = C.POINTER(C.c_double)

    def P(a, t=d):
        return a.ctypes.data_as(t)

    if os.path.exists(ref_path):
        kind = "reference"
        R = C.CDLL(ref_path)
        R.ref_tree_build.argtypes = [C.c_int, d, d, d, d, C.c_int, C.c_int]
        R.ref_tree_bh_targets.argtypes = [C.c_int, C.POINTER(C.c_int64), C.c_double, C.c_double,
                                          C.c_double, d]
        t0 = time.perf_counter()
        R.ref_tree_build(args.n, P(m), P(x), P(y), P(z), args.method, 64)
        t1 = time.perf_counter()
        F = np.zeros((len(sample), 3))
        R.ref_tree_bh_targets(len(sample), P(sample, C.POINTER(C.c_int64)), th2, 1.0, e2, P(F))
        t2 = time.perf_counter()
    else:
        kind = "port"
        import oracle as O
        t0 = time.perf_counter()
        tree = O.Tree(m, args.method).init(x, y, z)
        t1 = time.perf_counter()
        F = tree.bh_forces(th2, 1.0, e2, targets=sample)
        t2 = time.perf_counter()
    t_build, t_walk = t1 - t0, t2 - t1
    total = t_build + t_walk * args.n / len(sample)
    print(json.dumps({
        "value": args.n / total, "unit": "particles/s", "cores": threads, "kind": kind,
        "sample": (f"{args.profile} N={args.n} seed {args.seed}, method {args.method}, theta "
                   f"{args.theta}, eps {args.eps}: full insertion build + multipoles "
                   f"({t_build:.2f} s) and bh_forces_over_p on {len(sample)} evenly spaced "
                   f"targets ({t_walk:.2f} s) on {threads} OpenMP threads, walk scaled to N"),
        "t_build_s": t_build, "t_walk_sample_s": t_walk,
        "checksum": float(np.abs(F).sum())}))


def verlet(args):
    """Seconds per Verlet step of the reference (integration.f90:149-171: half
    kick, drift, tree%init, the force loop over all N with OpenMP, half kick)
    on a seeded Plummer system with velocities; the kicks and drift in numpy."""
    from hierarquicos_amd import ics
    m, q, v = ics.plummer(args.n, a=1.0, rmax=10.0, seed=args.seed, velocities=True)
    p = v * m[:, None]
    th2, e2 = args.theta ** 2, args.eps ** 2
    threads = args.threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ref_path = os.path.join(HERE, "_ref", "libhqref.so")
    d = C.POINTER(C.c_double)
    allt = np.arange(args.n, dtype=np.int64)

    def P(a, t=d):
        return a.ctypes.data_as(t)

    if os.path.exists(ref_path):
        kind = "reference"
        R = C.CDLL(ref_path)
        R.ref_tree_build.argtypes = [C.c_int, d, d, d, d, C.c_int, C.c_int]
        R.ref_tree_bh_targets.argtypes = [C.c_int, C.POINTER(C.c_int64), C.c_double, C.c_double,
                                          C.c_double, d]

        def forces(q):
            x, y, z = (np.ascontiguousarray(q[:, k]) for k in range(3))
            R.ref_tree_build(args.n, P(m), P(x), P(y), P(z), args.method, 64)
            F = np.zeros((args.n, 3))
            R.ref_tree_bh_targets(args.n, P(allt, C.POINTER(C.c_int64)), th2, 1.0, e2, P(F))
            return F
    else:
        kind = "port"
        import oracle as O

        def forces(q):
            t = O.Tree(m, args.method).init(q[:, 0], q[:, 1], q[:, 2])
            return t.bh_forces(th2, 1.0, e2)
    h = 0.5 * args.dt
    F = forces(q)
    t0 = time.perf_counter()
    for _ in range(args.verlet_steps):
        p = p + h * F
        q = q + args.dt * p / m[:, None]
        F = forces(q)
        p = p + h * F
    t1 = time.perf_counter()
    print(json.dumps({"s_per_step": (t1 - t0) / args.verlet_steps, "n": args.n, "kind": kind,
                      "cores": threads, "steps": args.verlet_steps, "method": args.method,
                      "checksum": float(np.abs(q).sum())}))


if __name__ == "__main__":
    try:
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    except (ValueError, OSError):
        pass
    main()
