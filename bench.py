"""Benchmark: particles/s per force evaluation at N = 1,048,576 (Plummer,
a = 1, r <= 10a, equal masses, G = 1, seed 0), octupole, theta = 0.5,
eps = 1e-3, mixed precision, with the RMS relative force error against the
FP64 direct sum on a 4096-particle sample.

A step is one force evaluation from device-resident FP64 positions to
device-resident FP64 forces: root cube, octant keys, radix sort, octree from
sorted keys, upward P2M/M2M pass (all of hq_tree_build) and the bunched
Barnes (1990) list walk with FP32 interactions (hq_tree_forces_bunched).

    python bench.py [--steps K] [--warmup W] [--n N] [--no-cpu-baseline]

Prints one JSON line.  Single GPU only: --gpus above 1 exits with a message.
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP32_PEAK_TFLOPS = 157.3   # MI355X dense vector FP32 (published)
FP64_PEAK_TFLOPS = 78.6
# flops per interaction as written in the reference's grouped expressions
# (octree_default.f90 :1978-1987 P2P = monopole, :2009-2027 quadrupole,
# :2032-2082 octupole)
FLOPS_P2P = 20
FLOPS_M2P = {1: 20, 4: 61, 8: 201}


def _run_child(cmd, env, timeout=600):
    """Runs a CPU-baseline child in its own process group and waits for it;
    on a time-out the whole group is killed, so nothing it started outlives
    this call."""
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env,
                         start_new_session=True)
    try:
        out, err = p.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, 9)
        out, err = p.communicate()
        return subprocess.CompletedProcess(cmd, -9, out, err + "\ntimed out")
    return subprocess.CompletedProcess(cmd, p.returncode, out, err)


def cpu_baseline(args):
    """Reference Fortran (or the C port) on the host cores, bounded sample."""
    env = dict(os.environ)
    env.setdefault("OMP_STACKSIZE", "64M")
    env.setdefault("OMP_PROC_BIND", "close")
    threads = int(env.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env["OMP_NUM_THREADS"] = str(threads)
    cmd = [sys.executable, os.path.join(REPO, "oracle", "cpu_baseline.py"), "--n", str(args.n),
           "--seed", str(args.seed), "--theta", str(args.theta), "--eps", str(args.eps),
           "--method", str(args.method), "--profile", args.profile, "--sample",
           str(args.cpu_sample), "--threads", str(threads)]
    r = _run_child(cmd, env)
    if r.returncode != 0:
        return {"value": None, "unit": "particles/s", "cores": threads, "kind": "reference",
                "sample": "failed: " + r.stderr[-300:]}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    for k in ("t_build_s", "t_walk_sample_s", "checksum"):
        out.pop(k, None)
    return out


def cpu_verlet(n, method, theta, eps, steps, seed=0):
    """Seconds per Verlet step of the reference on the host cores (bounded sample)."""
    env = dict(os.environ)
    env.setdefault("OMP_STACKSIZE", "64M")
    env.setdefault("OMP_PROC_BIND", "close")
    threads = int(env.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env["OMP_NUM_THREADS"] = str(threads)
    cmd = [sys.executable, os.path.join(REPO, "oracle", "cpu_baseline.py"), "--n", str(n),
           "--seed", str(seed), "--theta", str(theta), "--eps", str(eps), "--method",
           str(method), "--verlet-steps", str(steps), "--threads", str(threads)]
    r = _run_child(cmd, env)
    if r.returncode != 0:
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def run_ensemble(args):
    """Ensemble of Plummer systems with N uniform in [500, 20000], integrated
    together (one batch handle) with velocity Verlet; particle-steps/s."""
    from hierarquicos_amd import OctreeType, ics
    from hierarquicos_amd.device import sync
    rng = np.random.default_rng(args.seed)
    sizes = [int(v) for v in rng.integers(500, 20001, args.systems)]
    sys_ = [ics.plummer(n, seed=1000 + k, velocities=True) for k, n in enumerate(sizes)]
    m = np.concatenate([a[0] for a in sys_])
    q = np.concatenate([a[1] for a in sys_])
    p = np.concatenate([a[2] * a[0][:, None] for a in sys_])
    ntot = int(sum(sizes))
    th2, e2, dt = args.theta ** 2, args.eps ** 2, args.dt
    base = None
    if not args.no_cpu_baseline:
        samp = sorted(sizes)[:: max(1, len(sizes) // 4)][:4]
        recs = [cpu_verlet(n, 11, args.theta, args.eps, 2) for n in samp]
        if all(recs):
            rate = sum(samp) / sum(r["s_per_step"] for r in recs)
            base = {"value": rate, "unit": "particle-steps/s", "cores": recs[0]["cores"],
                    "kind": recs[0]["kind"],
                    "sample": (f"2 reference Verlet steps (quadrupole BH, theta {args.theta}) "
                               f"of each of the systems N={samp}; rate = sum N / sum s_per_step")}

    def make(precision):
        t = OctreeType().pre_init(m, 11, 64, precision=precision, counts=sizes)
        t.set_state(q, p)
        return t
    t = make("mixed")
    e0 = t.energy(1.0, e2)
    t.integrate(dt, args.warmup, th2, 1.0, e2, force="bunched")
    sync()
    a = time.perf_counter()
    t.integrate(dt, args.steps, th2, 1.0, e2, force="bunched")
    sync()
    el = time.perf_counter() - a
    e1 = t.energy(1.0, e2)
    drift = float(np.max(np.abs((e1[:, 0] - e0[:, 0]) / e0[:, 0])))
    t64 = make("fp64")
    t64.integrate(dt, 1, th2, 1.0, e2, force="tree")
    sync()
    a = time.perf_counter()
    t64.integrate(dt, 2, th2, 1.0, e2, force="tree")
    sync()
    ms64 = (time.perf_counter() - a) * 1e3 / 2
    value = ntot * args.steps / el
    print(json.dumps({
        "metric": "ensemble Verlet particle-steps/s, 256 Plummer systems N in [500, 20000]",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Plummer spheres with velocities)",
        "config": {"workload": f"ensemble_{args.systems}_plummer_500_20000_quadrupole",
                   "systems": args.systems, "N_total": ntot, "theta": args.theta,
                   "eps": args.eps, "dt": dt,
                   "forces": "bunched walk, FP32 interactions, quadrupole, one tree per system"},
        "energy_drift_max_rel": drift,
        "fp64_tree_ms_per_step": ms64,
        "cpu_baseline": base}))


def run_parareal(args):
    """Parareal on one N=10,000 Plummer sphere: 64 windows over t in [0, 4];
    coarse = symplectic Euler, monopole BH theta 0.9, dt ~ 1e-2; fine = Verlet,
    quadrupole BH theta 0.5 (bunched, FP32 interactions), dt ~ 1e-3; all fine
    windows in one batch; iterate until ||U_k+1 - U_k|| / (6N) <= 1e-20."""
    from hierarquicos_amd import OctreeType, ics
    from hierarquicos_amd import integrate as I
    from hierarquicos_amd.device import sync
    n, nw, tf, eps = args.n_parareal, 64, 4.0, args.eps
    m, q, v = ics.plummer(n, seed=args.seed, velocities=True)
    p = v * m[:, None]
    csteps, fsteps = 6, 64
    coarse = {"force": args.coarse_force, "method": 10, "theta": 0.9, "steps": csteps,
              "scheme": args.coarse_scheme, "precision": "mixed"}
    fine = {"force": "bunched", "method": 11, "theta": args.theta, "steps": fsteps,
            "precision": "mixed"}
    dtf = tf / nw / fsteps
    base = None
    if not args.no_cpu_baseline:
        r = cpu_verlet(n, 11, args.theta, eps, 3)
        if r:
            base = {"value": r["s_per_step"] * fsteps * (nw - 1), "unit": "s",
                    "cores": r["cores"], "kind": r["kind"],
                    "sample": (f"3 reference Verlet steps (quadrupole BH theta {args.theta}, "
                               f"N={n}) scaled to the sequential fine solution "
                               f"({nw - 1} windows x {fsteps} steps)")}
    # warm the library (allocations, code objects) on a short run
    I.integrate_parareal(m, q, p, tf / 16, 4, 2, 1.0, 1.0, eps, 1e-20, 0, coarse=coarse, fine=fine)
    sync()
    a = time.perf_counter()
    q1, p1, info = I.integrate_parareal(m, q, p, tf, nw, nw + 1, 1.0, 1.0, eps, 1e-20, 0,
                                        coarse=coarse, fine=fine)
    t_par = time.perf_counter() - a
    # the sequential fine solution it converges to
    t = OctreeType().pre_init(m, 11, 64, precision="mixed")
    t.set_state(q, p)
    sync()
    a = time.perf_counter()
    dq = dp = 0.0
    for w in range(nw - 1):
        t.integrate(dtf, fsteps, args.theta ** 2, 1.0, eps * eps, force="bunched")
    qs, ps = t.get_state()
    t_seq = time.perf_counter() - a
    dq = float(np.linalg.norm(q1[-1] - qs))
    dp = float(np.linalg.norm(p1[-1] - ps))
    err_ps = float(np.sqrt(dq * dq + dp * dp) / (6 * n))
    print(json.dumps({
        "metric": "Parareal wall time to ||U_k+1 - U_k||/(6N) <= 1e-20, N=10k, 64 windows",
        "value": t_par, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_par * 1e3, "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Plummer sphere with velocities)",
        "config": {"workload": f"parareal_plummer_n{n}_w{nw}_t4", "N": n, "windows": nw,
                   "tf": tf, "eps": eps,
                   "coarse": f"{args.coarse_scheme}, monopole BH theta 0.9 ({args.coarse_force} "
                             f"walk), {csteps} steps per window (dt {tf / nw / csteps:.4g})",
                   "fine": f"Verlet, quadrupole BH theta {args.theta} bunched FP32, {fsteps} "
                           f"steps per window (dt {dtf:.4g})",
                   "correction": "F + (G_new - G_old)"},
        "iterations": info["iterations"], "errors": [float(e) for e in info["errors"]],
        "sequential_fine_s": t_seq, "speedup_vs_sequential": t_seq / t_par,
        "par_minus_seq": err_ps,
        "cpu_baseline": base}))


def run_fmm(args):
    """Cartesian FMM at N=4,194,304: uniform cube [-1,1]^3 and Plummer, orders 4
    and 6, theta 0.5; same-level M2L on the matrix cores (FP32) beside the
    scalar FP64 M2L; error against the FP64 direct sum on a sample."""
    from hierarquicos_amd import DeviceArray, OctreeType, direct_forces, ics
    from hierarquicos_amd import _lib as L
    from hierarquicos_amd.device import sync
    N = args.n_fmm
    th2 = args.theta ** 2
    rng = np.random.default_rng(1234)
    sample = np.sort(rng.choice(N, args.sample, replace=False))
    # a plain 256 MiB device-to-device copy: the reference point that scales
    # the FETCH_SIZE / WRITE_SIZE counters of a --pmc profile to bytes
    probe_a, probe_b = DeviceArray((1 << 25,)), DeviceArray((1 << 25,))
    L.check(L.lib().hq_dev_copy(probe_b.ptr, probe_a.ptr, probe_a.nbytes, 2))
    del probe_a, probe_b
    cases = []
    for prof in args.fmm_profiles.split(","):
        if prof == "cube":
            m, q, _ = ics.uniform_cube(N, half=1.0, seed=args.seed)
        else:
            m, q, _ = ics.plummer(N, a=1.0, rmax=10.0, seed=args.seed)
        dm = DeviceArray.from_numpy(m)
        dx, dy, dz = (DeviceArray.from_numpy(np.ascontiguousarray(q[:, k])) for k in range(3))
        Fd = direct_forces(dm, dx, dy, dz, 1.0, 0.0, targets=sample)
        dF = DeviceArray((3 * N,))
        t = OctreeType().pre_init(dm, 10)
        for order in (int(o) for o in args.fmm_orders.split(",")):
            for mode in ("mfma", "fp64"):
                if mode == "fp64" and (order == 6 or args.no_extras):
                    continue
                t.init(dx, dy, dz)
                t.fmm(th2, 1.0, 0.0, order, mode, out=dF)     # warm: tables, buffers
                sync()
                walls = []
                for _ in range(args.steps):
                    a = time.perf_counter()
                    t.init(dx, dy, dz)
                    t.fmm(th2, 1.0, 0.0, order, mode, out=dF)
                    walls.append(time.perf_counter() - a)
                inf = t.info()
                F = dF.numpy().reshape(N, 3)[sample]
                r = np.linalg.norm(F - Fd, axis=1) / np.linalg.norm(Fd, axis=1)
                K = (order + 1) * (order + 2) * (order + 3) // 6
                flops = 2.0 * K * K * inf["exec_cells"]
                ms = float(np.median(walls)) * 1e3
                cases.append({"profile": prof, "order": order, "m2l": mode,
                              "ms_per_eval": ms, "particles_per_s": N / (ms * 1e-3),
                              "rms_rel_err": float(np.sqrt(np.mean(r * r))),
                              "max_rel_err": float(r.max()),
                              "m2l_kernel_ms": inf["ms_lists"], "build_ms": inf["ms_build"],
                              "fmm_ms": inf["ms_walk"], "near_l2p_ms": inf["ms_eval"],
                              "same_level_m2l_pairs": inf["exec_cells"],
                              "other_m2l_pairs": inf["exec_parts"], "near_pairs": inf["n_p2p"],
                              "m2p_pairs": inf["max_cells"],
                              "m2l_tflops": flops / (inf["ms_lists"] * 1e-3) / 1e12})
                print(f"fmm {prof} p={order} {mode}: {ms:.1f} ms", file=sys.stderr, flush=True)
        t.close()
    head = cases[0]
    peak = 157.3 if head["m2l"] == "mfma" else 78.6
    K0 = (head["order"] + 1) * (head["order"] + 2) * (head["order"] + 3) // 6
    workload = f"fmm_{head['profile']}_n{N}_p{head['order']}_theta{args.theta}"
    traffic = None   # HBM bytes per launch from the committed rocprofv3 --pmc summary
    # the matrix-core M2L kernel of this order
    kname = "k_fmm_m2l4" if head["order"] == 4 else "k_fmm_m2l6"
    tf = os.path.join(REPO, "profiles", f"traffic_{kname}_{head['profile']}_n{N}_p{head['order']}.json")
    if os.path.exists(tf) and head["m2l"] == "mfma":
        rec = json.load(open(tf))
        if rec.get("workload") == workload:
            traffic = rec["hbm_bytes_per_launch"]
    print(json.dumps({
        "metric": "FMM particles/s at N=4M uniform cube, p=4, theta=0.5 (matrix-core M2L)",
        "value": head["particles_per_s"], "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": head["ms_per_eval"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded uniform cube and Plummer sphere)",
        "config": {"workload": workload, "N": N, "theta": args.theta, "leaf": 64},
        "rms_rel_err": head["rms_rel_err"], "max_rel_err": head["max_rel_err"],
        "roofline": {"bound": "mfma", "kernel": f"{kname}<{head['order']}>",
                     "achieved": head["m2l_tflops"], "peak": peak, "unit": "TFLOP/s",
                     "frac": head["m2l_tflops"] / peak, "traffic": traffic,
                     "flops_per_launch": 2.0 * K0 * K0 * head["same_level_m2l_pairs"],
                     "flops_note": f"2 K^2 per same-level pair, K = {K0} coefficients (no padding)"},
        "cases": cases, "cpu_baseline": None}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--eps", type=float, default=1e-3)
    ap.add_argument("--method", type=int, default=12)
    ap.add_argument("--profile", choices=("plummer", "hernquist", "cube"), default="plummer")
    ap.add_argument("--bunch", type=int, default=64, help="bunch size (<= 64)")
    ap.add_argument("--super", type=int, default=128, help="super-group size of the bunching")
    ap.add_argument("--unfused", action="store_true", help="lists and evaluation as two kernels")
    ap.add_argument("--leaf", type=int, default=4,
                    help="leaf bucket: cells of <= leaf bodies enter the lists as their bodies")
    ap.add_argument("--far", type=float, default=5.0,
                    help="far-field factor: accepted cells at least this many bunch radii from "
                         "the bunch centre enter its local expansion (0: off)")
    ap.add_argument("--far-inner", type=float, default=3.0,
                    help="inner far field: accepted cells between this many and --far bunch "
                         "radii take a higher-order local expansion (0: off)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick bunch / super-group / leaf bucket with the autotuner first")
    ap.add_argument("--sample", type=int, default=4096)
    ap.add_argument("--seeds", type=int, default=5,
                    help="accuracy and per-seed timing (median of --seed-reps) on seeds 0..seeds-1")
    ap.add_argument("--seed-reps", type=int, default=20)
    ap.add_argument("--thetas", default="0.3,0.7",
                    help="other opening angles timed beside --theta (mixed and FP64 per-target)")
    ap.add_argument("--mac", choices=("side", "offset"), default="offset",
                    help="acceptance test of the bunched walk")
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip FP64 paths and counts")
    ap.add_argument("--workload", choices=("forces", "ensemble", "parareal", "fmm"),
                    default="forces",
                    help="forces: the headline force evaluation (default)")
    ap.add_argument("--systems", type=int, default=256, help="ensemble: number of systems")
    ap.add_argument("--dt", type=float, default=1e-3, help="ensemble: Verlet step")
    ap.add_argument("--n-parareal", type=int, default=10000)
    ap.add_argument("--coarse-scheme", choices=("euler", "verlet"), default="euler",
                    help="parareal: coarse propagator (euler is the reference's G_til)")
    ap.add_argument("--coarse-force", choices=("tree", "bunched"), default="bunched",
                    help="parareal: coarse walk (tree = per-target FP64, bunched = mixed)")
    ap.add_argument("--n-fmm", type=int, default=1 << 22)
    ap.add_argument("--fmm-profiles", default="cube,plummer")
    ap.add_argument("--fmm-orders", default="4,6")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the force path here runs on a single MI355X "
              "and has no multi-GPU mode (--gpus must be 1).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        if args.eps == 1e-3:
            args.eps = 1e-2
        return run_ensemble(args)
    if args.workload == "fmm":
        if args.steps == 20:
            args.steps = 3
        return run_fmm(args)
    if args.workload == "parareal":
        if args.eps == 1e-3:
            args.eps = 1e-2
        return run_parareal(args)

    from hierarquicos_amd import DeviceArray, OctreeType, direct_forces, ics
    from hierarquicos_amd import _lib as L
    from hierarquicos_amd.device import sync

    N = args.n
    if args.profile == "plummer":
        m, q, _ = ics.plummer(N, a=1.0, rmax=10.0, seed=args.seed)
    elif args.profile == "hernquist":
        m, q, _ = ics.hernquist(N, a=1.0, rmax=50.0, seed=args.seed)
    else:
        m, q, _ = ics.uniform_cube(N, half=1.0, seed=args.seed)
    base = None if args.no_cpu_baseline else cpu_baseline(args)   # before the GPU is touched

    th2, e2 = args.theta ** 2, args.eps ** 2
    dm = DeviceArray.from_numpy(m)
    dx, dy, dz = (DeviceArray.from_numpy(np.ascontiguousarray(q[:, k])) for k in range(3))
    dF = DeviceArray((3 * N,))
    tree = OctreeType().pre_init(dm, args.method, args.bunch, precision="mixed")
    tree.tune(bunch=args.bunch, super_size=args.super, fused=not args.unfused, leaf=args.leaf,
              far=args.far, far_inner=args.far_inner, mac=args.mac)
    tuned = None
    if args.autotune:
        tree.init(dx, dy, dz)
        tuned = tree.autotune(th2, 1.0, e2, "mixed", reps=2)
        args.bunch, args.super, args.leaf = tuned["bunch"], tuned["super_size"], tuned["leaf"]

    def step():
        tree.init(dx, dy, dz)
        tree.forces_bunched(th2, 1.0, e2, "mixed", out=dF)

    # a plain 256 MiB device-to-device copy: the reference point that scales
    # the FETCH_SIZE / WRITE_SIZE counters of a --pmc profile to bytes
    probe_a, probe_b = DeviceArray((1 << 25,)), DeviceArray((1 << 25,))
    L.check(L.lib().hq_dev_copy(probe_b.ptr, probe_a.ptr, probe_a.nbytes, 2))
    del probe_a, probe_b

    for _ in range(args.warmup):
        step()
    sync()
    evals, lists, builds, walls = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        step()                     # synchronous: returns when the forces are in HBM
        walls.append((time.perf_counter() - a) * 1e3)
        inf = tree.info()          # device (HIP event) times of this step's kernels
        evals.append(inf["ms_eval"])
        lists.append(inf["ms_lists"])
        builds.append(inf["ms_build"])
    sync()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = N / (ms_step * 1e-3)
    info = tree.info()
    q25, q50, q75 = (float(v) for v in np.percentile(walls, [25, 50, 75]))

    # the same step from host arrays to host forces (PCIe both ways included)
    hx, hy, hz = (np.ascontiguousarray(q[:, k]) for k in range(3))
    a = time.perf_counter()
    tree.init(hx, hy, hz)
    tree.forces_bunched(th2, 1.0, e2, "mixed")
    ms_host = (time.perf_counter() - a) * 1e3

    # accuracy against the FP64 direct sum on a sample of targets
    rng = np.random.default_rng(1234)
    sample = np.sort(rng.choice(N, args.sample, replace=False))
    Fd = direct_forces(dm, dx, dy, dz, 1.0, e2, targets=sample)
    Fm = dF.numpy().reshape(N, 3)[sample]

    def rel(F, ref=None):
        ref = Fd if ref is None else ref
        r = np.linalg.norm(F - ref, axis=1) / np.linalg.norm(ref, axis=1)
        return float(np.sqrt(np.mean(r * r))), float(r.max())
    rms, mx = rel(Fm)
    walk_stack = info["walk_stack"]

    def timed_steps(th2_, reps):
        for _ in range(2):
            tree.init(dx, dy, dz)
            tree.forces_bunched(th2_, 1.0, e2, "mixed", out=dF)
        ts = []
        for _ in range(reps):
            a = time.perf_counter()
            tree.init(dx, dy, dz)
            tree.forces_bunched(th2_, 1.0, e2, "mixed", out=dF)
            ts.append((time.perf_counter() - a) * 1e3)
        return float(np.median(ts))

    # other opening angles on the same input: the mixed walk and, below, the
    # reference's per-target FP64 walk
    sweep = []
    if not args.no_extras:
        for th in (float(v) for v in args.thetas.split(",") if v):
            msm = timed_steps(th * th, 10)
            r_, x_ = rel(dF.numpy().reshape(N, 3)[sample])
            sweep.append({"theta": th, "mixed_ms_per_eval": msm,
                          "mixed_particles_per_s": N / (msm * 1e-3), "mixed_rms_rel_err": r_,
                          "mixed_max_rel_err": x_, "walk_stack": tree.info()["walk_stack"],
                          "far_cells": tree.info()["far_cells"]})

    extras = {}
    flops_alg = None
    if not args.no_extras:
        # reference per-target BH counts (leaf size 1, same theta) from the FP64
        # walk of methods 10-12 = bh_forces_over_p, and the plain FP64 paths
        t64 = OctreeType().pre_init(dm, args.method, args.bunch, precision="fp64")
        t64.tune(bunch=args.bunch, super_size=args.super)
        dF64 = DeviceArray((3 * N,))
        t64.init(dx, dy, dz)
        sync()
        a = time.perf_counter()
        t64.init(dx, dy, dz)
        t64.forces_all(th2, 1.0, e2, out=dF64)
        sync()
        ms_bh64 = (time.perf_counter() - a) * 1e3
        i64 = t64.info()
        flops_alg = FLOPS_P2P * i64["n_p2p"] + FLOPS_M2P[i64["multipole"]] * i64["n_m2p"]
        rms_bh64, mx_bh64 = rel(dF64.numpy().reshape(N, 3)[sample])
        a = time.perf_counter()
        t64.init(dx, dy, dz)
        t64.forces_bunched(th2, 1.0, e2, "fp64", out=dF64)
        sync()
        ms_bun64 = (time.perf_counter() - a) * 1e3
        rms_b64, mx_b64 = rel(dF64.numpy().reshape(N, 3)[sample])
        ib64 = t64.info()
        flops_b64 = (FLOPS_M2P[ib64["multipole"]] * ib64["exec_cells"]
                     + FLOPS_P2P * ib64["exec_parts"])
        for rec in sweep:
            th2_ = rec["theta"] ** 2
            t64.init(dx, dy, dz)
            sync()
            a = time.perf_counter()
            t64.init(dx, dy, dz)
            t64.forces_all(th2_, 1.0, e2, out=dF64)
            sync()
            ms_ = (time.perf_counter() - a) * 1e3
            r_, x_ = rel(dF64.numpy().reshape(N, 3)[sample])
            rec.update({"fp64_reference_bh_ms_per_eval": ms_, "fp64_reference_bh_rms_rel_err": r_,
                        "fp64_reference_bh_max_rel_err": x_,
                        "mixed_speedup_vs_fp64_reference_bh": ms_ / rec["mixed_ms_per_eval"]})
        extras = {
            "fp64_bunched": {"ms_per_eval": ms_bun64, "particles_per_s": N / (ms_bun64 * 1e-3),
                             "rms_rel_err": rms_b64, "max_rel_err": mx_b64,
                             "mixed_speedup": ms_bun64 / ms_step,
                             "walk_ms": ib64["ms_walk"],
                             "executed_tflops": flops_b64 / (ib64["ms_walk"] * 1e-3) / 1e12,
                             "executed_frac_fp64_peak": flops_b64 / (ib64["ms_walk"] * 1e-3)
                             / 1e12 / FP64_PEAK_TFLOPS},
            "fp64_reference_bh": {"ms_per_eval": ms_bh64, "particles_per_s": N / (ms_bh64 * 1e-3),
                                  "rms_rel_err": rms_bh64, "max_rel_err": mx_bh64,
                                  "n_p2p": i64["n_p2p"], "n_m2p": i64["n_m2p"],
                                  "walk_ms": i64["ms_walk"],
                                  "tflops": flops_alg / (i64["ms_walk"] * 1e-3) / 1e12,
                                  "frac_fp64_peak": flops_alg / (i64["ms_walk"] * 1e-3) / 1e12
                                  / FP64_PEAK_TFLOPS,
                                  "mixed_speedup": ms_bh64 / ms_step},
        }

    # every seed of the input: median step time and the error on the sample
    seeds = []
    if not args.no_extras:
        for sd in range(args.seeds):
            if sd != args.seed:
                if args.profile == "plummer":
                    _, qs, _ = ics.plummer(N, a=1.0, rmax=10.0, seed=sd)
                elif args.profile == "hernquist":
                    _, qs, _ = ics.hernquist(N, a=1.0, rmax=50.0, seed=sd)
                else:
                    _, qs, _ = ics.uniform_cube(N, half=1.0, seed=sd)
                for arr, k in ((dx, 0), (dy, 1), (dz, 2)):   # equal masses: only positions change
                    col = np.ascontiguousarray(qs[:, k])   # held until the copy is done
                    L.check(L.lib().hq_dev_copy(arr.ptr, C.c_void_p(col.ctypes.data), arr.nbytes, 0))
            ms_sd = timed_steps(th2, args.seed_reps)
            Fd_sd = direct_forces(dm, dx, dy, dz, 1.0, e2, targets=sample)
            r_, x_ = rel(dF.numpy().reshape(N, 3)[sample], Fd_sd)
            seeds.append({"seed": sd, "ms_per_step_median": ms_sd, "rms_rel_err": r_,
                          "max_rel_err": x_})

    eval_ms = float(np.mean(evals))
    mp = {10: 1, 11: 4, 12: 8}[args.method]
    walk = f"k_walk_mixed<{mp},{walk_stack}>"
    if args.unfused or walk_stack == 0:
        kernel = f"k_eval_mixed<{mp}>"
    elif args.far > 0 and 0 < args.far_inner < args.far:   # the walk and both far-field tiers
        kernel = f"{walk} + k_far<{mp},5> + k_far<{mp},6>"
    elif args.far > 0:   # the walk and its far-field kernel, timed together
        kernel = f"{walk} + k_far<{mp},5>"
    else:
        kernel = walk
    order = {10: "monopole", 11: "quadrupole", 12: "octupole"}[args.method]
    workload = f"{args.profile}_n{N}_{order}_theta{args.theta}_eps{args.eps}"
    traffic = None   # HBM bytes per launch from the committed rocprofv3 --pmc summary
    tf = os.path.join(REPO, "profiles", f"traffic_k_walk_mixed_{workload.split('_theta')[0]}.json")
    if os.path.exists(tf) and not args.unfused:
        rec = json.load(open(tf))
        if (rec.get("workload") == workload and rec.get("leaf", 0) == args.leaf
                and rec.get("far", 0.0) == args.far
                and rec.get("far_inner", 0.0) == args.far_inner
                and rec.get("super", 256) == args.super and rec.get("mac", "side") == args.mac):
            traffic = rec["hbm_bytes_per_launch"]
    roof = None
    if flops_alg is not None:
        achieved = flops_alg / (eval_ms * 1e-3) / 1e12
        executed = (FLOPS_M2P[info["multipole"]] * info["exec_cells"]
                    + FLOPS_P2P * info["exec_parts"]) / (eval_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "compute_unit": "VALU (FP32 FMA/rsqrt; no matrix ops)",
                "kernel": kernel,
                "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                "frac_note": ("achieved prices the reference's per-target interaction counts "
                              "(leaf size 1, same theta); executed_frac prices what the walk "
                              "evaluates one body at a time (far cells excluded)"),
                "executed_tflops": executed, "executed_frac": executed / FP32_PEAK_TFLOPS,
                "flops_per_launch": flops_alg, "kernel_ms": eval_ms,
                "list_entries": {"cells": info["list_cells"], "particles": info["list_parts"]},
                "executed_interactions": {"cells": info["exec_cells"],
                                          "particles": info["exec_parts"],
                                          "far_cells_m2l": info["far_cells"]},
                "bunches": info["ngroups"],
                "longest_lists": {"cells": info["max_cells"], "particles": info["max_parts"],
                                  "stack": info["max_stack"]}}
    if args.far <= 0:
        far_text = "no far field"
    elif 0 < args.far_inner < args.far:
        far_text = (f"far field from {args.far_inner:g} bunch radii (local expansion of order 6, "
                    f"order 5 beyond {args.far:g})")
    else:
        far_text = f"far field beyond {args.far:g} bunch radii (local expansion of order 5)"
    headline = (args.profile == "plummer" and N == 1 << 20 and args.method == 12
                and args.theta == 0.5)
    if headline:
        metric = "particles/s per force eval at N=1M, rms rel err <=1e-4 vs FP64 direct"
    else:
        nl = f"{N >> 20}M" if N % (1 << 20) == 0 else str(N)
        metric = (f"particles/s per force eval at N={nl}, {args.profile}, "
                  f"{ {10: 'monopole', 11: 'quadrupole', 12: 'octupole'}[args.method]}, "
                  f"theta={args.theta:g} (rms rel err vs FP64 direct reported)")
    shape = {"plummer": "Plummer sphere", "hernquist": "Hernquist sphere",
             "cube": "uniform cube"}[args.profile]
    line = {
        "metric": metric,
        "value": value, "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic (seeded {shape})",
        "config": {"workload": workload, "profile": args.profile,
                   "N": N, "seed": args.seed, "theta": args.theta, "eps": args.eps,
                   "order": order,
                   "precision": "mixed (FP32 interactions, FP64 positions/keys/moments)",
                   "walk": (f"bunched Barnes-1990 lists ({args.mac} acceptance test), <= "
                            f"{args.bunch} particles per bunch, "
                            f"cut from sibling subtrees packed to <= {args.super}, leaf bucket "
                            f"{args.leaf}, " + far_text),
                   "autotuned": tuned},
        "rms_rel_err": rms, "max_rel_err": mx, "err_sample": args.sample,
        "rms_rel_err_worst_seed": max([s["rms_rel_err"] for s in seeds], default=rms),
        "seeds": seeds, "theta_sweep": sweep,
        "ms_per_step_median": q50, "ms_per_step_iqr": [q25, q75],
        "host_arrays_ms_per_step": ms_host,
        "host_arrays_particles_per_s": N / (ms_host * 1e-3),
        "phases_ms": {"build": float(np.mean(builds)), "lists": float(np.mean(lists)),
                      "eval": eval_ms, "keys": info["ms_keys"], "sort": info["ms_sort"],
                      "nodes": info["ms_nodes"], "upward": info["ms_upward"]},
        "nnodes": info["nnodes"],
        "walk_stats": {"bunches": info["ngroups"], "list_cells": info["list_cells"],
                       "list_parts": info["list_parts"], "exec_cells": info["exec_cells"],
                       "exec_parts": info["exec_parts"], "max_cells": info["max_cells"],
                       "max_parts": info["max_parts"], "max_stack": info["max_stack"],
                       "far_cells": info["far_cells"]},
        "roofline": roof,
        "cpu_baseline": base,
    }
    line.update(extras)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
