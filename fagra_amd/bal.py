"""BAL problem files and the BAL benchmark runner.

* :func:`parse` / :func:`read` -- ``Problem::parse`` (examples/bal/model.rs:59-113) in the
  native library: the reference's token grammar and validation messages, observations in
  input order. Errors raise :class:`BalError` with the reference's message.
* :func:`format` / :func:`write` -- BAL text that :func:`parse` reads back bit for bit.
* :func:`objective` -- the published BAL formula evaluated independently of the device
  evaluator (examples/bal.rs:202-219), the runner's end check.
* :func:`identify` -- the SHA-256 of a decompressed BAL file against the Ladybug problems
  whose CPU timings the reference publishes (docs/bal-performance.md:145-151).
* ``python -m fagra_amd.bal problem.txt [steps=20] [inner=1000] [repeats=3]
  [scaled|block|identity|schur] [tolerance=1e-6] [--fp32] [--timings]`` -- the reference's
  runner (examples/bal.rs:93-281) on the GPU backends: metadata on stderr, one CSV row per
  run on stdout, the same columns.
"""
import ctypes
import hashlib
import os
import resource
import sys
import time

import numpy as np

from . import _lib
from .synthetic import BalProblem, project

# docs/bal-performance.md:145-151 (decompressed files)
LADYBUG_SHA256 = {
    "96ca2845519d89d0727953d983427ab38a42c54991cd4d73e46a4221da3c61b4": "problem-49-7776-pre.txt",
    "58525e22314ccd72f5baec575f04cdd0be329dd90a23d82489108aef441af5b4": "problem-73-11032-pre.txt",
    "0921d835d31752af15cec103866535290cb40a9a7aa50102bda47c57ebd387f2": "problem-138-19878-pre.txt",
}

CSV_HEADER = ("run,load_s,build_s,solve_s,status,initial_cost,final_cost,rmse,last_gradient,accepted,rejected,"
              "linearizations,solves,inner_iterations,cost_evaluations,last_rejection,damping,peak_rss_kib,"
              "target_5pct_s,target_2pct_s,prepare_s,factor_s,linear_s,products_s,precondition_s,verification_s")
TERMINATION_NAMES = {0: "NoVariables", 1: "GradientTolerance", 2: "StepTolerance", 3: "CostTolerance"}
REJECTION_NAMES = {-1: "None", 0: "Some(LinearSolve)", 1: "Some(Prediction)", 2: "Some(InvalidEvaluation)",
                   3: "Some(PoorAgreement)"}


class BalError(ValueError):
    """BAL text rejected by Problem::parse; the message is the reference's."""


def _bytes(text):
    return text.encode() if isinstance(text, str) else bytes(text)


def parse(text, name="bal"):
    """Problem::parse on BAL text (str or bytes)."""
    lib = _lib.load()
    data = _bytes(text)
    err = ctypes.create_string_buffer(128)
    sizes = _lib.BalSizes()
    st = lib.fg_bal_scan(data, len(data), ctypes.byref(sizes), err, len(err))
    if st == _lib.INVALID_INPUT:
        raise BalError(err.value.decode())
    _lib.check(st, "fg_bal_scan")
    nc, npts, no = sizes.n_cameras, sizes.n_points, sizes.n_observations
    cams, pts = np.empty((nc, 9)), np.empty((npts, 3))
    oc, op, px = np.empty(no, np.int64), np.empty(no, np.int64), np.empty((no, 2))
    d = ctypes.POINTER(ctypes.c_double)
    i64 = ctypes.POINTER(ctypes.c_int64)
    st = lib.fg_bal_parse(data, len(data), ctypes.byref(sizes), cams.ctypes.data_as(d), pts.ctypes.data_as(d),
                          oc.ctypes.data_as(i64), op.ctypes.data_as(i64), px.ctypes.data_as(d), err, len(err))
    if st == _lib.INVALID_INPUT:
        raise BalError(err.value.decode())
    _lib.check(st, "fg_bal_parse")
    return BalProblem(name=name, cameras=cams, points=pts, obs_camera=oc, obs_point=op, obs_pixel=px)


def read(path):
    with open(path, "rb") as f:
        return parse(f.read(), name=os.path.basename(path))


def format(problem, cameras=None, points=None):  # noqa: A001 (mirrors str.format naming of the writer)
    """BAL text of ``problem`` with optional replacement states (e.g. an optimized solution)."""
    lib = _lib.load()
    cams = np.ascontiguousarray(problem.cameras if cameras is None else cameras, dtype=np.float64)
    pts = np.ascontiguousarray(problem.points if points is None else points, dtype=np.float64)
    oc = np.ascontiguousarray(problem.obs_camera, dtype=np.int64)
    op = np.ascontiguousarray(problem.obs_point, dtype=np.int64)
    px = np.ascontiguousarray(problem.obs_pixel, dtype=np.float64)
    if cams.shape != (problem.n_cameras, 9) or pts.shape != (problem.n_points, 3):
        raise ValueError("state shapes do not match the problem")
    sizes = _lib.BalSizes(problem.n_cameras, problem.n_points, problem.n_observations)
    d = ctypes.POINTER(ctypes.c_double)
    i64 = ctypes.POINTER(ctypes.c_int64)
    args = [ctypes.byref(sizes), cams.ctypes.data_as(d), pts.ctypes.data_as(d), oc.ctypes.data_as(i64),
            op.ctypes.data_as(i64), px.ctypes.data_as(d)]
    need = ctypes.c_size_t()
    st = lib.fg_bal_format(*args, None, 0, ctypes.byref(need))
    if st == _lib.INVALID_OPTIONS:
        raise ValueError("observation index out of range or nonfinite value")
    buf = ctypes.create_string_buffer(need.value + 1)
    _lib.check(lib.fg_bal_format(*args, buf, need.value, ctypes.byref(need)), "fg_bal_format")
    return buf.raw[:need.value].decode()


def write(path, problem, cameras=None, points=None):
    with open(path, "w") as f:
        f.write(format(problem, cameras, points))


def objective(problem, cameras, points):
    """0.5 sum ||pixel(camera, point) - observed||^2 by the published BAL formula."""
    pred = project(np.asarray(cameras, np.float64), np.asarray(points, np.float64), problem.obs_camera,
                   problem.obs_point)
    return 0.5 * float(np.sum((pred - problem.obs_pixel) ** 2))


def identify(path):
    """Name of the Ladybug problem this decompressed file is, or None."""
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return LADYBUG_SHA256.get(h.hexdigest())


def _peak_rss_kib():
    try:
        with open("/proc/self/status") as f:
            for line in f:
                if line.startswith("VmHWM:"):
                    return line.split()[1]
    except OSError:
        pass
    return str(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss)


def run(path, steps=20, inner=1000, repeats=3, mode="scaled", tolerance=1e-6, precision=None, timings=False,
        out=sys.stdout, err=sys.stderr):
    """The reference runner's loop (examples/bal.rs:168-281); returns the per-run rows."""
    import fagra_amd as fa

    precision = fa.FP64 if precision is None else precision
    modes = {"scaled": (True, False), "block": (True, True), "identity": (False, False), "schur": (True, True)}
    if mode not in modes:
        raise ValueError("backend mode must be scaled, block, identity or schur")
    if not np.isfinite(tolerance) or not 0.0 <= tolerance < 1.0:
        raise ValueError("tolerance must be finite and in [0,1)")
    if min(steps, inner, repeats) <= 0:
        raise ValueError("limits and repeats must be positive")
    t0 = time.perf_counter()
    problem = read(path)
    load = time.perf_counter() - t0
    print(f"file={path} cameras={problem.n_cameras} points={problem.n_points} "
          f"observations={problem.n_observations} steps={steps} inner={inner} repeats={repeats} mode={mode} "
          f"rel_tol={tolerance} gradient_tol=1e-8 max_trials=16 gauge=free device=gpu "
          f"precision={'fp32' if precision == fa.FP32 else 'fp64'} ladybug={identify(path)}", file=err, flush=True)
    print(CSV_HEADER, file=out, flush=True)
    rows = []
    for r in range(1, repeats + 1):
        t0 = time.perf_counter()
        ba = fa.BalProblem(problem, precision=precision, huber=0.0)
        if mode == "schur":
            backend = fa.Schur(precision=precision, max_iterations=inner, relative_tolerance=tolerance,
                               collect_timings=timings)
        else:
            scaled, block = modes[mode]
            backend = fa.Lsmr(precision=precision, max_iterations=inner, relative_tolerance=tolerance,
                              diagonal_preconditioning=scaled, block_preconditioning=block,
                              collect_timings=timings)
        build = time.perf_counter() - t0
        initial = objective(problem, problem.cameras, problem.points)
        reached = [None, None]
        start = [0.0]

        def progress(stats):
            for i, frac in enumerate((0.05, 0.02)):
                if reached[i] is None and stats.cost <= initial * frac:
                    reached[i] = time.perf_counter() - start[0]

        lm = fa.LevenbergMarquardt(backend)
        lm.on_accept = progress
        start[0] = time.perf_counter()
        o = lm.solve_batch(ba, fa.OptimizeOptions(steps, 1e-8, 0.0, 0.0))
        elapsed = time.perf_counter() - start[0]
        cams, pts = ba.state()
        cost = objective(problem, cams, pts)
        stats_cost = o.cost if np.isfinite(o.cost) else initial
        if not np.isfinite(cost) or abs(stats_cost - cost) > 1e-8 * max(cost, 1.0):
            raise RuntimeError("independent objective mismatch")
        status = TERMINATION_NAMES.get(o.termination, "?") if o.status == fa.OK else _lib.STATUS_NAMES.get(
            o.status, str(o.status))
        t = backend.timings()
        if mode == "schur":
            tm = (t.preparation, t.factorization, t.linear_solve, t.products, t.preconditioner, t.verification)
        else:
            tm = (t.preparation, t.factorization, t.linear_solve, t.forward + t.transpose, t.preconditioner,
                  t.residual_checks)
        solves, iters = backend.statistics()
        print(f"run={r} last_linear={_diag_text(backend.diagnostics())}", file=err, flush=True)
        tgt = ["NA" if v is None else f"{v:.6f}" for v in reached]
        tms = [f"{v:.6f}" if timings else "NA" for v in tm]
        row = (f"{r},{load:.6f},{build:.6f},{elapsed:.6f},{status},{initial:.12e},{cost:.12e},"
               f"{np.sqrt(2 * cost / problem.n_observations):.8f},{o.gradient_norm:.8e},{o.accepted_steps},"
               f"{o.rejected_steps},{o.linearizations},{solves},{iters},{o.cost_evaluations},"
               f"{REJECTION_NAMES.get(o.last_rejection, '?')},{o.damping:.8e},{_peak_rss_kib()},{tgt[0]},{tgt[1]},"
               + ",".join(tms))
        print(row, file=out, flush=True)
        rows.append(row)
    return rows


def _diag_text(d):
    return "{" + ", ".join(f"{k}: {getattr(d, k)}" for k, _ in d._fields_) + "}"


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    flags = {a for a in argv if a.startswith("--")}
    pos = [a for a in argv if not a.startswith("--")]
    usage = ("usage: python -m fagra_amd.bal problem.txt [steps=20] [inner=1000] [repeats=3] "
             "[scaled|block|identity|schur] [tolerance=1e-6] [--fp32] [--timings]")
    if not pos or len(pos) > 6 or flags - {"--fp32", "--timings"}:
        print(usage, file=sys.stderr)
        return 2
    import fagra_amd as fa

    def number(i, default):
        return int(pos[i]) if len(pos) > i else default

    try:
        run(pos[0], steps=number(1, 20), inner=number(2, 1000), repeats=number(3, 3),
            mode=pos[4] if len(pos) > 4 else "scaled", tolerance=float(pos[5]) if len(pos) > 5 else 1e-6,
            precision=fa.FP32 if "--fp32" in flags else fa.FP64, timings="--timings" in flags)
    except (BalError, ValueError, RuntimeError, OSError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
