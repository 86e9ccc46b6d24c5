"""Benchmark: seconds for LM to reach 1e-5 relative cost on the 1,500-camera synthetic BA.

Workload ``rings1500`` (seed 0): 1,500 cameras on rings of radius 8/10/12 around the origin,
500,000 points uniform in [-2, 2]^3, track lengths U{3..10} (3,249,262 observations), pixel
noise N(0, 0.5 px), 2% outliers uniform over the image, blockwise Huber delta = 1 px, initial
estimates perturbed by 0.01 rad / 0.05 units.

One "step" is one complete LM solve from the initial estimates (fagra's
LevenbergMarquardt::<Schur>::solve_batch semantics, LM defaults, Schur CG relative tolerance
--cg-tol for every mode). A mode's time-to-target is the wall time from the call until the
first accepted state with cost <= C* (1 + 1e-5), taken in the on_accept observer (as
examples/bal.rs:14-34 does for its 5%/2% targets). C* is the cost of an FP64 solve at CG
tolerance 1e-10 after --cstar-iter accepted steps (or its gradient-tolerance termination), on
the same input. Inputs are resident in HBM before timing; every solve starts from the same
initial estimates. LM is deterministic (fixed-order reductions), so after the warmup solves
have found the step at which a mode reaches the target, the timed solves stop there.

Printed JSON line (one line, a few KB): value = median time-to-target of the headline mode
(FP32 + learned preconditioner when trained weights are present, else FP32 + block-Jacobi),
the FP64 block-Jacobi time beside it, throughput rates, the Schur-product roofline from HIP
events and the CPU oracle baseline. Per-run detail goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s to LM converge @1e-5 rel cost, 1.5k-cam BA: learned+FP32 vs plain f64 PCG"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the Schur-product kernels on rings1500 from the hardware-counter runs
# summarized in profiles/hbm_traffic_rings1500.md (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE,
# over bench.py --profile-products), scaled by what the same counters read for a 256 MiB
# device copy: bytes = 2.0 x FETCH_SIZE KiB x 1024 + 1.0 x WRITE_SIZE KiB x 1024. Counters
# cannot be read in the timed run, so "traffic" carries the figure for the workload it was
# measured on (null elsewhere).
TRAFFIC_SOURCE = "profiles/hbm_traffic_rings1500.md"
MEASURED_TRAFFIC = {
    ("rings1500", "k_points<float, kSchur>"): (2 * 44331.2 + 7812.5) * 1024,
    ("rings1500", "k_schur_cameras<float>"): (2 * 52957.9 + 1829.1) * 1024,
    ("rings1500", "k_points<double, kSchur>"): (2 * 70308.6 + 15625.0) * 1024,
    ("rings1500", "k_schur_cameras<double>"): (2 * 80921.6 + 1514.7) * 1024,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(nc, npts, no, precision):
    """Algorithmic HBM bytes of one launch of each Schur-product kernel, every stream counted
    once (repeated gathers of a camera's table or forward table, or of a point's solve, are
    cache traffic, not algorithmic bytes).

    k_points<kSchur> (one lane per point): per observation the compact record (4 x T) and its
    camera index (4 B); per point the attempt's factor of Hpp + lambda D^2 (6 f64, written once
    per damped solve by the point factorization) in and the solve t (one 4 x T row) out; per
    camera the table row it reads (16 x T) and the forward table of p (12 x T).
    k_schur_cameras (one wave per camera and point-slot range): per observation the record copy
    (4 x T) and the point-slot index (4 B); per point t back in (4 x T); per camera the table
    (32 x T), the forward table (12 x T), the column scales (9 x T) and p in, q out (9 x 8 B
    each).
    """
    t = 4 if precision == "fp32" else 8
    points = no * (4 * t + 4) + npts * (6 * 8 + 4 * t) + nc * (16 + 12) * t
    cameras = no * (4 * t + 4) + npts * 4 * t + nc * ((32 + 12 + 9) * t + 2 * 9 * 8)
    return points, cameras


def stored_jacobian_matvec_bytes(nc, npts, no):
    """The stored-Jacobian two-pass design's bytes per FP32 matvec, for comparison."""
    return no * 2 * (96 + 4) + npts * (24 + 12 + 12) + nc * (180 + 36 + 36)


def make_backend(fa, precision, precond, cg_tol):
    if precond in ("lsmr_block", "lsmr_diagonal"):
        return fa.Lsmr(precision=precision, relative_tolerance=cg_tol, diagonal_preconditioning=True,
                       block_preconditioning=precond == "lsmr_block")
    return fa.Schur(precision=precision, relative_tolerance=cg_tol, preconditioner=precond)


def run_mode(fa, problem_dev, precision, precond, cg_tol, target, max_iter, gradient_tol=0.0):
    """One LM solve from the initial state; precond names the Schur preconditioner or, as
    "lsmr_block" / "lsmr_diagonal", the LSMR backend with that right preconditioning."""
    backend = make_backend(fa, precision, precond, cg_tol)
    lm = fa.LevenbergMarquardt(backend)
    hit = {}
    t0 = [0.0]
    last_log = [0.0]
    tail = []

    def on_accept(st):
        now = time.perf_counter() - t0[0]
        tail.append(st.cost)
        del tail[:-51]
        if now - last_log[0] > 20.0:  # long solves (C* on large workloads) show progress
            last_log[0] = now
            log(f"bench: ... {st.accepted_steps} steps, cost {st.cost!r} at {now:.0f}s")
        if "t" not in hit and st.cost <= target:
            hit["t"] = now
            hit["steps"] = st.accepted_steps
            hit["attempts"] = st.attempts
            hit["linearizations"] = st.linearizations
            hit["cost_evaluations"] = st.cost_evaluations
            hit["solves"], hit["iterations"] = backend.statistics()

    lm.on_accept = on_accept
    problem_dev.set_state(*problem_dev.initial)
    opts = fa.OptimizeOptions(max_iter, gradient_tol, 0.0, 0.0)
    t0[0] = time.perf_counter()
    out = lm.solve_batch(problem_dev, opts)
    total = time.perf_counter() - t0[0]
    backend.close()
    return {"time_to_target": hit.get("t"), "total": total, "status": out.status, "cost": out.cost,
            "accepted": out.accepted_steps, "rejected": out.rejected_steps, "linearizations": out.linearizations,
            "inner_iterations": out.linear_iterations, "solves": out.linear_solves,
            "at_target": {k: v for k, v in hit.items() if k != "t"}, "tail": list(tail)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def oracle_sample(O, problem, cg_tol):
    """One full LM step of the oracle at the benchmark's settings, timed by component:
    linearization, damping preparation, cost of a state, and two damped solves at the initial
    damping 1e-3 (CG tolerance 1e-2 and cg_tol), and three reduced-camera products at that
    damping, whose fastest is the time per PCG iteration (the iteration's vector work is O(cameras)
    and left out). The fixed part of a solve (factorization, right-hand side, recovery,
    verification) is the cg_tol solve less its iterations. Two solves a few iterations apart
    are too close to difference: the first also pays for first-touch memory."""
    ba = O.BA(problem, huber=problem.huber)
    t = time.perf_counter()
    st, s = O.ba_schur(ba)
    t_lin = time.perf_counter() - t
    assert st == O.OK
    t = time.perf_counter()
    assert s.prepare_damping(1e-6) == O.OK
    t_prep = time.perf_counter() - t
    t = time.perf_counter()
    assert ba.cost()[0] == O.OK
    t_cost = time.perf_counter() - t
    solves = []
    for tol in (1e-2, cg_tol):
        s.relative_tolerance = tol
        n0 = s.statistics()[1]
        t = time.perf_counter()
        st, _, _ = s.solve_damped(1e-3)
        solves.append((time.perf_counter() - t, s.statistics()[1] - n0))
    tb, nb = solves[1]
    x = np.random.default_rng(0).standard_normal(9 * problem.n_cameras)
    t_iter = float("inf")
    for _ in range(3):
        t = time.perf_counter()
        assert s.reduced_apply(1e-3, x)[0] == O.OK
        t_iter = min(t_iter, time.perf_counter() - t)
    t_fixed = max(tb - nb * t_iter, 0.0)
    return {"linearize": t_lin, "prepare_damping": t_prep, "cost": t_cost, "solve": tb, "solve_iterations": nb,
            "per_iteration": t_iter, "solve_fixed": t_fixed,
            "first_lm_step": t_lin + t_prep + tb + t_cost}


def cpu_baseline(problem, counts, cg_tol):
    """The CPU oracle (f64 restatement of the reference's Schur-CG path) timed on the host on a
    bounded sample of the same workload -- one LM step, by component -- single-threaded and
    pinned to one core as the reference's runs are (docs/schur-performance.md:189-198), and on
    all host cores; each projected to the target with the FP64 GPU run's work counts."""
    from oracle import oracle as O

    def project(t):
        return (counts["linearizations"] * (t["linearize"] + t["prepare_damping"])
                + counts["solves"] * t["solve_fixed"] + counts["iterations"] * t["per_iteration"]
                + counts["cost_evaluations"] * t["cost"])

    allowed = sorted(os.sched_getaffinity(0))
    all_threads = min(len(allowed), int(os.environ.get("OMP_NUM_THREADS", len(allowed)) or len(allowed)))
    pin = 2 if 2 in allowed else allowed[0]
    O.set_threads(1)
    os.sched_setaffinity(0, {pin})
    try:
        t1 = oracle_sample(O, problem, cg_tol)
    finally:
        os.sched_setaffinity(0, allowed)
    O.set_threads(all_threads)
    tn = oracle_sample(O, problem, cg_tol)
    r = lambda d: {k: (float(f"{v:.6g}") if isinstance(v, float) else v) for k, v in d.items()}
    return {"value": project(t1), "unit": "s", "cores": 1, "kind": "port",
            "sample": ("oracle f64, one LM step of the same workload (linearize, prepare_damping, cost, "
                       f"solve_damped at lambda 1e-3 to CG tol 1e-2 and {cg_tol:g}, 3 reduced-camera products for the "
                       f"per-iteration time), 1 thread pinned to cpu {pin}; "
                       "projected to the target with the FP64 GPU run's counts "
                       f"({counts['linearizations']} linearizations, {counts['solves']} solves, "
                       f"{counts['iterations']} PCG iterations, {counts['cost_evaluations']} costs)"),
            "single_thread": r(t1), "all_cores": {"value": project(tn), "threads": all_threads, **r(tn)},
            "cpu": cpu_model(), "host_threads_allowed": len(allowed)}


def time_linearizations(fa, dev, precision, repeats):
    """Mean wall seconds per linearization (bulk Frame::linearize + accumulate and the gradient
    norm that synchronizes it) at the initial state."""
    dev.set_state(*dev.initial)
    s = fa.Schur(precision=precision)
    assert s.prepare(dev) == fa.OK
    s.linearize()
    s.gradient_norm()
    t = time.perf_counter()
    for _ in range(repeats):
        assert s.linearize() == fa.OK
        assert s.gradient_norm()[0] == fa.OK
    el = (time.perf_counter() - t) / repeats
    s.close()
    return el


def widened_rows(no_cpu):
    """Measurements of the rows beyond the BAL Schur path (python bench.py --widened), each beside
    the oracle timed on the same bounded sample as its CPU baseline: the generic-emission Schur backend (a planar pose / landmark model, one damped solve)
    and the selected covariance (a dense 3,072-coordinate information matrix)."""
    import fagra_amd as fa
    from fagra_amd import generic
    from fagra_amd.planar import Planar
    out = {}
    m = Planar(n_poses=2000, n_landmarks=50000, n_obs=400000, seed=7, noise=0.01)
    g = generic.GenericSchur(3 * m.np_, 3, 2, relative_tolerance=1e-6)
    t = time.perf_counter()
    assert g.prepare(m.dimension) == fa.OK and m.linearize(g) == fa.OK
    st, _ = g.gradient_norm()
    t_acc = time.perf_counter() - t
    t = time.perf_counter()
    assert g.prepare_damping(1e-6) == fa.OK
    t_prep = time.perf_counter() - t
    g.solve_damped(1e-3)  # warm
    reps, t = 5, time.perf_counter()
    for _ in range(reps):
        st, d, pred = g.solve_damped(1e-3)
    t_solve = (time.perf_counter() - t) / reps
    gen = {"model": "planar: 2000 poses (3), 50000 landmarks (2), 400000 observations (2 rows) + prior",
           "dimension": m.dimension, "host_accumulate_s": t_acc, "prepare_damping_s": t_prep,
           "solve_damped_s": t_solve, "pcg_iterations": int(g.diagnostics().iterations), "status": st}
    if not no_cpu:
        from oracle import oracle as O
        O.set_threads(1)  # the single-core baseline
        o = O.Schur(3 * m.np_, 3, 2)
        o.relative_tolerance = 1e-6
        assert o.prepare(m.dimension) == O.OK
        r, pr = m.residuals(m.x)
        jp, jl = m.jacobians(m.x)
        for e in range(m.pi.size):
            o.accumulate(r[e], [(3 * int(m.pi[e]), jp[e]), (3 * m.np_ + 2 * int(m.li[e]), jl[e])])
        o.accumulate(pr, [(0, np.eye(3))])
        t = time.perf_counter()
        o.prepare_damping(1e-6)
        t_oprep = time.perf_counter() - t
        t = time.perf_counter()
        st_o, d_o, _ = o.solve_damped(1e-3)
        t_osolve = time.perf_counter() - t
        gen["cpu_baseline"] = {"prepare_damping_s": t_oprep, "solve_damped_s": t_osolve, "cores": 1,
                             "step_rel_diff": float(np.linalg.norm(d - d_o) / np.linalg.norm(d_o))}
    out["generic_schur"] = gen
    n, k = 3072, 64
    rng = np.random.default_rng(3)
    A = rng.normal(size=(n + 64, n))
    H = np.asfortranarray(A.T @ A)  # column-major, as the ABI takes it
    cols = rng.permutation(n)[:k]
    fa.dense_covariance(H[:128, :128], cols[:4] % 128)  # warm
    t = time.perf_counter()
    st, C = fa.dense_covariance(H, cols)
    t_cov = time.perf_counter() - t
    cov = {"n": n, "selected": k, "seconds": t_cov, "status": st,
           "tflops_cholesky": (n ** 3 / 3) / t_cov / 1e12}
    if not no_cpu:
        from oracle import oracle as O
        t = time.perf_counter()
        st_o, C_o = O.dense_covariance(H, cols)
        cov["cpu_baseline"] = {"seconds": time.perf_counter() - t, "cores": 1,
                             "max_abs_diff": float(np.max(np.abs(C - C_o)))}
    out["covariance"] = cov
    return out


def profile_products(args):
    """Counter-run workload: repeats of each Schur-product kernel at the initial state of the
    workload in both precisions, then a plain 256 MiB device copy whose FETCH_SIZE and
    WRITE_SIZE calibrate the counters."""
    import fagra_amd as fa
    from fagra_amd import synthetic
    import torch
    p = synthetic.generate(args.workload, seed=args.seed)
    out = {}
    for prec, name in ((fa.FP32, "fp32"), (fa.FP64, "fp64")):
        d = fa.BalProblem(p, precision=prec)
        s = fa.Schur(precision=prec)
        assert s.prepare(d) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
        x = np.random.default_rng(1).normal(size=9 * p.n_cameras)
        tp, tc = s.time_products(1e-4, x, repeats=args.kernel_repeats)
        out[name] = {"k_points_us": tp * 1e6, "k_cameras_us": tc * 1e6,
                     "bytes": kernel_bytes(p.n_cameras, p.n_points, p.n_observations, name)}
        log(f"bench: {name} k_points {tp * 1e6:.1f} us, k_cameras {tc * 1e6:.1f} us")
        s.close()
        d.close()
    a = torch.empty(64 * 2**20, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    out["copy_bytes"] = 2 * a.numel() * 4
    print(json.dumps(out))


def median(xs):
    xs = [x for x in xs if x is not None]
    return statistics.median(xs) if xs else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rings1500")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cg-tol", type=float, default=1e-3)
    ap.add_argument("--max-iter", type=int, default=150)
    ap.add_argument("--cstar-iter", type=int, default=400)
    ap.add_argument("--kernel-repeats", type=int, default=50)
    ap.add_argument("--lin-repeats", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="fp64,fp32,learned")
    ap.add_argument("--widened", action="store_true",
                    help="only measure the generic-emission Schur backend and the dense covariance, each "
                         "beside the oracle (one JSON line)")
    ap.add_argument("--profile-products", action="store_true",
                    help="only time the two Schur-product kernels (FP32, FP64) and a 256 MiB device copy, "
                         "for hardware-counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "fagra_amd measures one GPU: --gpus must be 1 (the sharding of the "
                                   "path over several GPUs is described, not measured)"}))
        sys.exit(2)
    if args.profile_products:
        return profile_products(args)
    if args.widened:
        print(json.dumps({"widened": widened_rows(args.no_cpu)}))
        return

    import fagra_amd as fa
    from fagra_amd import learned, synthetic

    t = time.perf_counter()
    p = synthetic.generate(args.workload, seed=args.seed)
    log(f"bench: generated {args.workload}: {p.n_cameras} cameras, {p.n_points} points, "
        f"{p.n_observations} observations in {time.perf_counter() - t:.1f}s")
    dev = {}
    for prec in (fa.FP64, fa.FP32):
        d = fa.BalProblem(p, precision=prec)
        d.initial = (p.cameras, p.points)
        dev[prec] = d

    # C*: FP64 at the reference's default CG tolerance 1e-10, gradient tolerance 1e-8
    t = time.perf_counter()
    ref = run_mode(fa, dev[fa.FP64], fa.FP64, fa.PRECOND_BLOCK_JACOBI, 1e-10, -1.0, args.cstar_iter,
                   gradient_tol=1e-8)
    c_star = ref["cost"]
    target = c_star * (1 + 1e-5)
    tail = ref["tail"]
    cstar = {"cost": c_star, "status": ref["status"], "accepted_steps": ref["accepted"],
             "cg_tolerance": 1e-10, "step_budget": args.cstar_iter,
             "rel_decrease_last_50_steps": (tail[0] - tail[-1]) / c_star if len(tail) > 1 else None,
             "seconds": round(time.perf_counter() - t, 2)}
    log(f"bench: C* {cstar}")

    modes = []
    for m in args.modes.split(","):
        if m == "fp64":
            modes.append(("fp64_block_jacobi", fa.FP64, fa.PRECOND_BLOCK_JACOBI))
        elif m == "fp32":
            modes.append(("fp32_block_jacobi", fa.FP32, fa.PRECOND_BLOCK_JACOBI))
        elif m == "learned" and learned.available():
            modes.append(("fp32_learned", fa.FP32, fa.PRECOND_LEARNED))
        elif m == "lsmr":  # the LSMR backend (block-Jacobi right preconditioning)
            modes.append(("fp64_lsmr_block", fa.FP64, "lsmr_block"))
            modes.append(("fp32_lsmr_block", fa.FP32, "lsmr_block"))
    summary = {}
    for name, prec, pc in modes:
        stop = args.max_iter
        for w in range(args.warmup):
            r = run_mode(fa, dev[prec], prec, pc, args.cg_tol, target, args.max_iter)
            log(f"bench: warmup {name}: time_to_target {r['time_to_target']} at {r['at_target']} "
                f"total {r['total']:.3f}s cost {r['cost']!r}")
            if r["at_target"]:
                stop = r["at_target"]["steps"]
        runs = []
        for k in range(args.steps):
            r = run_mode(fa, dev[prec], prec, pc, args.cg_tol, target, stop)
            log(f"bench: {name}: time_to_target {r['time_to_target']} total {r['total']:.3f}s "
                f"steps {r['accepted']} inner {r['inner_iterations']} at_target {r['at_target']}")
            runs.append(r)
        tt = [r["time_to_target"] for r in runs]
        at = runs[-1]["at_target"] or None
        summary[name] = {"time_to_target_s": median(tt), "min_s": min((x for x in tt if x), default=None),
                         "runs_at_target": sum(x is not None for x in tt), "runs": len(runs), "at_target": at,
                         "lm_iterations_per_s": (at["attempts"] / median(tt)) if at else None}

    head = "fp32_learned" if "fp32_learned" in summary else "fp32_block_jacobi"
    if head not in summary:
        head = next(iter(summary))
    value = summary[head]["time_to_target_s"]
    head_prec = next(m[1] for m in modes if m[0] == head)
    if value is None:
        log(f"bench: the headline mode {head} did not reach the target")
    prec_name = "fp32" if head_prec == fa.FP32 else "fp64"

    # roofline of the dominant kernel and the product rate: back-to-back launches of each
    # Schur-product kernel at the initial state, timed with HIP events on the backend stream
    kern = {}
    for prec, pname in ((fa.FP32, "fp32"), (fa.FP64, "fp64")):
        d = dev[prec]
        d.set_state(*d.initial)
        sch = fa.Schur(precision=prec)
        assert sch.prepare(d) == fa.OK and sch.linearize() == fa.OK and sch.prepare_damping(1e-6) == fa.OK
        x = np.random.default_rng(1).normal(size=9 * p.n_cameras)
        t_pts, t_cams = sch.time_products(1e-4, x, repeats=args.kernel_repeats)
        sch.close()
        kern[pname] = (t_pts, t_cams)
    t_pts, t_cams = kern[prec_name]
    b_pts, b_cams = kernel_bytes(p.n_cameras, p.n_points, p.n_observations, prec_name)
    tname = "float" if prec_name == "fp32" else "double"
    if t_pts >= t_cams:
        kname, kbytes, ksec = f"k_points<{tname}, kSchur>", b_pts, t_pts
    else:
        kname, kbytes, ksec = f"k_schur_cameras<{tname}>", b_cams, t_cams
    achieved = kbytes / ksec / 1e9
    traffic = MEASURED_TRAFFIC.get((args.workload, kname))
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname,
                "bytes_per_launch": kbytes, "avg_launch_us": round(ksec * 1e6, 2),
                "launches_timed": args.kernel_repeats,
                "traffic_source": TRAFFIC_SOURCE if traffic else None,
                "schur_product": {"k_points_us": round(t_pts * 1e6, 2), "k_cameras_us": round(t_cams * 1e6, 2),
                                  "bytes": b_pts + b_cams,
                                  "frac": round((b_pts + b_cams) / (t_pts + t_cams) / 1e9 / HBM_PEAK_GBS, 4),
                                  "stored_jacobian_design_bytes": stored_jacobian_matvec_bytes(
                                      p.n_cameras, p.n_points, p.n_observations)}}
    log(f"bench: roofline {kname}: {ksec * 1e6:.1f} us, {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS:.1%})")

    t_lin = {pname: time_linearizations(fa, dev[prec], prec, args.lin_repeats)
             for prec, pname in ((fa.FP32, "fp32"), (fa.FP64, "fp64"))}
    rates = {"lm_iterations_per_s": summary[head]["lm_iterations_per_s"],
             "schur_matvecs_per_s": {k: round(1.0 / sum(v), 1) for k, v in kern.items()},
             "linearizations_per_s": {k: round(1.0 / v, 1) for k, v in t_lin.items()},
             "lm_iterations": "LM attempts to the target / time to target (headline mode)",
             "schur_matvecs": "1 / (k_points + k_cameras) back-to-back, HIP events",
             "linearizations": "bulk linearization + gradient norm, back-to-back, wall clock"}

    cpu = None
    fp64_at = summary.get("fp64_block_jacobi", {}).get("at_target")
    if not args.no_cpu and fp64_at:
        t = time.perf_counter()
        cpu = cpu_baseline(p, fp64_at, args.cg_tol)
        log(f"bench: cpu baseline in {time.perf_counter() - t:.1f}s: {cpu}")

    workload = (f"{args.workload} seed {args.seed}: {p.n_cameras} cameras, {p.n_points} points, "
                f"{p.n_observations} observations, Huber {p.huber}")
    print(result_line(args, workload, head, prec_name, summary, cstar, rates, roofline, cpu))


def result_line(args, workload, head, prec_name, summary, cstar, rates, roofline, cpu):
    """The one JSON line the benchmark prints: summaries only (no per-run arrays)."""
    value = summary[head]["time_to_target_s"]
    f64_value = summary.get("fp64_block_jacobi", {}).get("time_to_target_s")
    line = {"metric": METRIC, "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1e3 * value, 3) if value else None,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 storage, f64 arithmetic" if prec_name == "fp32" else "f64",
            "data": "synthetic (seeded generator, no dataset)",
            "config": {"workload": workload, "mode": head, "cg_relative_tolerance": args.cg_tol,
                       "target": "C*(1+1e-5)", "step": "one LM solve from the initial estimates to the target"},
            "c_star": cstar, "fp64_block_jacobi_s": f64_value,
            "speedup_vs_fp64": round(f64_value / value, 3) if (value and f64_value) else None,
            "modes": summary, "rates": rates, "roofline": roofline, "cpu_baseline": cpu}
    return json.dumps(line, separators=(",", ":"))


if __name__ == "__main__":
    main()
