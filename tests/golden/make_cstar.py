"""Converged cost of the CPU oracle's f64 LM + Schur-CG on a seeded workload (golden fixture).

Usage: python tests/golden/make_cstar.py rings1500 0 [cg_tol] [max_iter]
Writes tests/golden/cstar_<workload>_<seed>.json with the final cost, the LM statistics and the
wall time of this single-threaded run.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fagra_amd import synthetic  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    workload, seed = sys.argv[1], int(sys.argv[2])
    tol = float(sys.argv[3]) if len(sys.argv) > 3 else 1e-8
    iters = int(sys.argv[4]) if len(sys.argv) > 4 else 40
    p = synthetic.generate(workload, seed=seed)
    t = time.time()
    ba = oracle.BA(p, huber=p.huber)
    out, trace = ba.lm(options=oracle.OptimizeOptions(iters, 1e-8, 0.0, 0.0), tol=tol, trace_capacity=200)
    el = time.time() - t
    res = {"workload": workload, "seed": seed, "cg_relative_tolerance": tol, "max_iterations": iters,
           "n_cameras": p.n_cameras, "n_points": p.n_points, "n_observations": p.n_observations,
           "status": out.status, "initial_cost": out.initial_cost, "final_cost": out.cost,
           "accepted_steps": out.accepted_steps, "rejected_steps": out.rejected_steps,
           "linearizations": out.linearizations, "cg_iterations": out.linear_iterations,
           "gradient_norm": out.gradient_norm, "cost_trace": trace.tolist(), "seconds_single_thread": el}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"cstar_{workload}_{seed}.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps({k: v for k, v in res.items() if k != "cost_trace"}))


if __name__ == "__main__":
    main()
