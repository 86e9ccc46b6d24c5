"""Benchmark: seconds for LM to reach 1e-5 relative cost on the 1,500-camera synthetic BA.

Workload ``rings1500`` (seed 0): 1,500 cameras on rings of radius 8/10/12 around the origin,
500,000 points uniform in [-2, 2]^3, track lengths U{3..10} (about 3.25M observations),
pixel noise N(0, 0.5 px), 2% outliers uniform over the image, blockwise Huber delta = 1 px,
initial estimates perturbed by 0.01 rad / 0.05 units.

One "step" is one complete LM solve from the initial estimates (fagra's
LevenbergMarquardt::<Schur>::solve_batch semantics, LM defaults, Schur CG relative
tolerance --cg-tol for every mode). A mode's time-to-target is the wall time from the call
until the first accepted state with cost <= C* (1 + 1e-5), taken in the on_accept observer
(as examples/bal.rs:14-34 does for its 5%/2% targets); C* is the converged cost of an FP64
solve at CG tolerance 1e-10 on the same input. Inputs are resident in HBM before timing.

Printed JSON line: value = time-to-target of the headline mode (FP32 + learned preconditioner
when a trained model is present, else FP32 + block-Jacobi), with the FP64 block-Jacobi
time beside it, the Schur-matvec roofline from HIP events, and the CPU oracle baseline.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s to LM converge @1e-5 rel cost, 1.5k-cam BA: learned+FP32 vs plain f64 PCG"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the Schur-product kernels on rings1500 from the hardware-counter
# runs (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over bench.py --profile-products), scaled
# by what the same counters read for a 256 MiB device copy: bytes = 2.0 x FETCH_SIZE KiB x 1024
# + 1.0 x WRITE_SIZE KiB x 1024. Counters cannot be read in the timed run itself, so the
# roofline's "traffic" carries these figures for the workload they were measured on.
MEASURED_TRAFFIC = {
    ("rings1500", "k_points<float, kSchur>"): (2 * 35550.3 + 7812.5) * 1024,
    ("rings1500", "k_cameras<float, cSchur>"): (2 * 71126.3 + 187.5) * 1024,
    ("rings1500", "k_points<double, kSchur>"): (2 * 64437.1 + 15625.0) * 1024,
    ("rings1500", "k_cameras<double, cSchur>"): (2 * 192577.1 + 194.9) * 1024,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(nc, npts, no, precision):
    """Algorithmic HBM bytes of one launch of each Schur-product kernel, every stream counted
    once (repeated gathers of a camera's table or forward table, or of a point's solve, are
    cache traffic, not algorithmic bytes).

    k_points<kSchur> (one lane per point): per observation the compact record (4 x T) and its
    camera index (4 B); per point the column scales (3 x T) in and the solve t (one 4 x T
    row) out; per camera the table row it reads (16 x T) and the forward table of p (12 x T).
    k_cameras<cSchur> (one workgroup per camera): per observation the record copy (4 x T) and
    the point-slot index (4 B); per point t back in (4 x T); per camera the table (32 x T),
    the forward table (12 x T), the column scales (9 x T) and p in, q out (9 x 8 B each).
    """
    t = 4 if precision == "fp32" else 8
    points = no * (4 * t + 4) + npts * (3 * t + 4 * t) + nc * (16 + 12) * t
    cameras = no * (4 * t + 4) + npts * 4 * t + nc * ((32 + 12 + 9) * t + 2 * 9 * 8)
    return points, cameras


def survey_matvec_bytes(nc, npts, no):
    """The stored-Jacobian two-pass design's bytes per FP32 matvec, for comparison."""
    return no * 2 * (96 + 4) + npts * (24 + 12 + 12) + nc * (180 + 36 + 36)


def run_mode(fa, problem_dev, precision, precond, cg_tol, target, max_iter, collect_timings=False):
    """One LM solve from the initial state; precond names the Schur preconditioner or, as
    "lsmr_block" / "lsmr_diagonal", the LSMR backend with that right preconditioning."""
    if precond in ("lsmr_block", "lsmr_diagonal"):
        backend = fa.Lsmr(precision=precision, relative_tolerance=cg_tol, diagonal_preconditioning=True,
                          block_preconditioning=precond == "lsmr_block", collect_timings=collect_timings)
    else:
        backend = fa.Schur(precision=precision, relative_tolerance=cg_tol, preconditioner=precond,
                           collect_timings=collect_timings)
    lm = fa.LevenbergMarquardt(backend)
    hit = {}
    t0 = [0.0]

    trace = []
    last_log = [0.0]

    def on_accept(st):
        now = time.perf_counter() - t0[0]
        trace.append((now, st.cost))
        if now - last_log[0] > 20.0:  # long solves (the converged C* on large workloads) show progress
            last_log[0] = now
            log(f"bench: ... {st.accepted_steps} steps, cost {st.cost!r} at {now:.0f}s")
        if "t" not in hit and st.cost <= target:
            hit["t"] = time.perf_counter() - t0[0]
            hit["steps"] = st.accepted_steps
            hit["attempts"] = st.attempts
            hit["linearizations"] = st.linearizations
            hit["cost_evaluations"] = st.cost_evaluations
            hit["solves"], hit["iterations"] = backend.statistics()

    lm.on_accept = on_accept
    cams0, pts0 = problem_dev.initial
    problem_dev.set_state(cams0, pts0)
    opts = fa.OptimizeOptions(max_iter, 0.0, 0.0, 0.0)
    t0[0] = time.perf_counter()
    out = lm.solve_batch(problem_dev, opts)
    total = time.perf_counter() - t0[0]
    tim = backend.timings()
    res = {"time_to_target": hit.get("t"), "total": total, "status": out.status, "cost": out.cost,
           "accepted": out.accepted_steps, "rejected": out.rejected_steps, "linearizations": out.linearizations,
           "cg_iterations": out.linear_iterations, "solves": out.linear_solves,
           "at_target": {k: v for k, v in hit.items() if k != "t"},
           "matvecs": getattr(tim, "matvecs", None), "matvec_kernel_s": getattr(tim, "matvec_kernel_seconds", None),
           "trace": trace, "timings": {f: getattr(tim, f) for f, _ in tim._fields_}}
    backend.close()
    return res


def cpu_baseline(problem, fp64_counts, budget_s):
    """Time the CPU oracle (single-threaded f64 restatement) on a bounded sample of the same
    workload: one linearization + damping preparation and a few Schur products, then scale by
    the FP64 GPU run's work counts up to the target."""
    from oracle import oracle
    t = time.perf_counter()
    ba = oracle.BA(problem, huber=problem.huber)
    st, s = oracle.ba_schur(ba)
    t_lin = time.perf_counter() - t
    t = time.perf_counter()
    assert s.prepare_damping(1e-6) == oracle.OK
    t_prep = time.perf_counter() - t
    t = time.perf_counter()
    st, c = ba.cost()
    t_cost = time.perf_counter() - t
    x = np.random.default_rng(0).normal(size=9 * problem.n_cameras)
    n_mv, t = 0, time.perf_counter()
    while True:
        s.reduced_apply(1e-3, x)
        n_mv += 1
        el = time.perf_counter() - t
        if el > budget_s or n_mv >= 50:
            break
    t_mv = el / n_mv
    c = fp64_counts
    projected = (c["linearizations"] * (t_lin + t_prep) + c["iterations"] * t_mv
                 + c["solves"] * 3 * t_mv + c["cost_evaluations"] * t_cost)
    return {"value": projected, "unit": "s", "cores": 1, "kind": "port",
            "sample": (f"oracle f64, 1 thread: 1 linearization {t_lin:.2f}s + damping prep {t_prep:.2f}s, "
                       f"1 cost {t_cost:.2f}s, {n_mv} Schur products {t_mv:.3f}s each on rings1500; "
                       f"projected with the FP64 GPU run's counts to target ({c['linearizations']} "
                       f"linearizations, {c['iterations']} CG iterations, {c['solves']} solves, "
                       f"{c['cost_evaluations']} cost evaluations)"),
            "cpu": platform.processor() or platform.machine(), "host_nproc": os.cpu_count()}


def widened_rows(no_cpu):
    """Measurements of the rows beyond the BAL Schur path, each against the oracle on a bounded
    sample: the generic-emission Schur backend (a planar pose / landmark model, one damped solve)
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
        gen["cpu_oracle"] = {"prepare_damping_s": t_oprep, "solve_damped_s": t_osolve, "cores": 1,
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
        cov["cpu_oracle"] = {"seconds": time.perf_counter() - t, "cores": 1,
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rings1500")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cg-tol", type=float, default=1e-3)
    ap.add_argument("--max-iter", type=int, default=150)
    ap.add_argument("--cstar-iter", type=int, default=400)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--kernel-repeats", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-widened", action="store_true",
                    help="skip the generic-emission and covariance measurements")
    ap.add_argument("--modes", default="fp64,fp32,learned,lsmr")
    ap.add_argument("--profile-products", action="store_true",
                    help="only time the two Schur-product kernels (FP32, FP64) and a 256 MiB device copy, "
                         "for hardware-counter runs")
    args = ap.parse_args()
    if args.profile_products:
        return profile_products(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design note only)"}))
        sys.exit(2)

    import fagra_amd as fa
    from fagra_amd import learned, synthetic

    t = time.perf_counter()
    p = synthetic.generate(args.workload, seed=args.seed)
    log(f"bench: generated {args.workload}: {p.n_cameras} cameras, {p.n_points} points, "
        f"{p.n_observations} observations in {time.perf_counter() - t:.1f}s")
    dev = {}
    for prec in (fa.FP64, fa.FP32):
        t = time.perf_counter()
        d = fa.BalProblem(p, precision=prec)
        d.initial = (p.cameras, p.points)
        dev[prec] = d
        log(f"bench: uploaded precision {prec} in {time.perf_counter() - t:.1f}s")

    # C*: the converged FP64 cost on the same input (the reference's default CG tolerance 1e-10)
    t = time.perf_counter()
    ref = run_mode(fa, dev[fa.FP64], fa.FP64, fa.PRECOND_BLOCK_JACOBI, 1e-10, -1.0, args.cstar_iter,
                   collect_timings=False)
    c_star = ref["cost"]
    target = c_star * (1 + 1e-5)
    log(f"bench: C* = {c_star!r} (status {ref['status']}, {ref['accepted']} steps, {ref['cg_iterations']} CG "
        f"iterations, {time.perf_counter() - t:.1f}s)")

    have_model = learned.available()
    modes = []
    for m in args.modes.split(","):
        if m == "fp64":
            modes.append(("fp64_block_jacobi", fa.FP64, fa.PRECOND_BLOCK_JACOBI))
        elif m == "fp32":
            modes.append(("fp32_block_jacobi", fa.FP32, fa.PRECOND_BLOCK_JACOBI))
        elif m == "learned" and have_model:
            modes.append(("fp32_learned", fa.FP32, fa.PRECOND_LEARNED))
        elif m == "lsmr":  # the LSMR backend (block-Jacobi right preconditioning), both precisions
            modes.append(("fp64_lsmr_block", fa.FP64, "lsmr_block"))
            modes.append(("fp32_lsmr_block", fa.FP32, "lsmr_block"))
        elif m == "lsmr_diagonal":
            modes.append(("fp64_lsmr_diagonal", fa.FP64, "lsmr_diagonal"))
    results = {}
    for name, prec, pc in modes:
        for w in range(args.warmup):
            r = run_mode(fa, dev[prec], prec, pc, args.cg_tol, target, args.max_iter)
            log(f"bench: warmup {name}: time_to_target {r['time_to_target']} total {r['total']:.3f}s")
        runs = []
        for k in range(args.steps):
            r = run_mode(fa, dev[prec], prec, pc, args.cg_tol, target, args.max_iter)
            log(f"bench: {name}: time_to_target {r['time_to_target']} total {r['total']:.3f}s "
                f"cost {r['cost']!r} steps {r['accepted']} cg {r['cg_iterations']} matvecs {r['matvecs']}")
            runs.append(r)
        results[name] = runs

    head = "fp32_learned" if "fp32_learned" in results else "fp32_block_jacobi"
    if head not in results:
        head = next((k for k in results if not k.endswith(("lsmr_block", "lsmr_diagonal"))), list(results)[0])
    hr = results[head]
    times = [r["time_to_target"] for r in hr if r["time_to_target"] is not None]
    value = statistics.median(times) if times else None
    mode_s = {}
    for name, runs in results.items():
        ts = [r["time_to_target"] for r in runs if r["time_to_target"] is not None]
        mode_s[name] = statistics.median(ts) if ts else None
    f64_value = mode_s.get("fp64_block_jacobi")

    # roofline of the dominant kernel: back-to-back launches of each Schur-product kernel at the
    # initial state, timed with HIP events on the backend stream
    hname, hprec, hpc = next(m for m in modes if m[0] == head)
    prec_name = "fp32" if head.startswith("fp32") else "fp64"
    d = dev[hprec]
    d.set_state(*d.initial)
    sch = fa.Schur(precision=hprec)
    assert sch.prepare(d) == fa.OK and sch.linearize() == fa.OK and sch.prepare_damping(1e-6) == fa.OK
    x = np.random.default_rng(1).normal(size=9 * p.n_cameras)
    t_pts, t_cams = sch.time_products(1e-4, x, repeats=args.kernel_repeats)
    sch.close()
    b_pts, b_cams = kernel_bytes(p.n_cameras, p.n_points, p.n_observations, prec_name)
    if t_pts >= t_cams:
        kname, kbytes, ksec = "k_points<%s, kSchur>" % ("float" if prec_name == "fp32" else "double"), b_pts, t_pts
    else:
        kname, kbytes, ksec = "k_cameras<%s, cSchur>" % ("float" if prec_name == "fp32" else "double"), b_cams, t_cams
    achieved = kbytes / ksec / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC.get((args.workload, kname)), "kernel": kname,
                "bytes_per_launch": kbytes, "avg_launch_us": ksec * 1e6, "launches_timed": args.kernel_repeats,
                "schur_product": {"k_points_us": t_pts * 1e6, "k_cameras_us": t_cams * 1e6,
                                  "bytes": b_pts + b_cams,
                                  "achieved_GBps": (b_pts + b_cams) / (t_pts + t_cams) / 1e9,
                                  "stored_jacobian_design_bytes": survey_matvec_bytes(p.n_cameras, p.n_points,
                                                                                      p.n_observations)}}
    log(f"bench: roofline {kname}: {ksec * 1e6:.1f} us, {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS:.1%})")

    widened = None
    if not args.no_widened:
        t = time.perf_counter()
        widened = widened_rows(args.no_cpu)
        log(f"bench: widened rows measured in {time.perf_counter() - t:.1f}s: "
            f"generic solve {widened['generic_schur']['solve_damped_s'] * 1e3:.1f} ms, "
            f"covariance {widened['covariance']['seconds'] * 1e3:.1f} ms")

    cpu = None
    f64 = results.get("fp64_block_jacobi", [])
    if not args.no_cpu and f64:
        cnt = f64[-1]["at_target"] or {"linearizations": f64[-1]["linearizations"],
                                        "iterations": f64[-1]["cg_iterations"], "solves": f64[-1]["solves"],
                                        "cost_evaluations": f64[-1]["accepted"] + f64[-1]["rejected"] + 1}
        cpu = cpu_baseline(p, cnt, args.cpu_budget)

    slim = {k: [{kk: vv for kk, vv in r.items() if kk not in ("timings", "trace")} for r in v]
            for k, v in results.items()}
    line = {"metric": METRIC, "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * statistics.mean(r["total"] for r in hr),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": prec_name,
            "data": "synthetic",
            "config": {"workload": f"{args.workload} seed {args.seed}: {p.n_cameras} cameras, {p.n_points} points, "
                                   f"{p.n_observations} observations, Huber {p.huber}", "mode": head,
                       "cg_relative_tolerance": args.cg_tol, "target": "C*(1+1e-5)", "c_star": c_star},
            "fp64_block_jacobi_s": f64_value, "mode_seconds": mode_s,
            "speedup_vs_fp64": (f64_value / value) if (value and f64_value) else None,
            "modes": slim, "roofline": roofline, "cpu_baseline": cpu, "widened": widened}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
