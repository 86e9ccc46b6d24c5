"""Oracle LM trajectories on the seeded workloads (golden fixtures for the GPU workload tests).

Usage: python tests/golden/make_trajectory.py WORKLOAD SEED CG_TOL STEPS [TAG]
Runs the CPU oracle's f64 LM + Schur-CG (LM defaults, gradient tolerance 1e-8, step and cost
tolerances 0, at most STEPS accepted steps, CG relative tolerance CG_TOL) from the workload's
initial estimates and writes tests/golden/trajectory_<workload>_<seed>_<tag>.json: the LM
statistics, one record per accepted step (cost, attempts so far, damping), and a digest of the
final state -- the 2-norms of the camera and point parameters and 2,048 coordinates of the
flattened state [cameras (9 nc), points (3 np)] at seeded positions -- plus a digest of the
problem (sizes and sums of its arrays) that the tests check before using the fixture.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fagra_amd import synthetic  # noqa: E402
from oracle import oracle  # noqa: E402

SAMPLES = 2048


def problem_digest(p):
    return {"n_cameras": p.n_cameras, "n_points": p.n_points, "n_observations": p.n_observations,
            "cameras_sum": float(np.sum(p.cameras)), "points_sum": float(np.sum(p.points)),
            "pixel_sum": float(np.sum(p.obs_pixel)), "obs_point_sum": int(np.sum(p.obs_point)),
            "obs_camera_sum": int(np.sum(p.obs_camera))}


def sample_index(n):
    return np.sort(np.random.default_rng(12345).choice(n, size=min(SAMPLES, n), replace=False))


def state_digest(cams, pts):
    x = np.concatenate([np.ravel(cams), np.ravel(pts)])
    idx = sample_index(x.size)
    return {"camera_norm": float(np.linalg.norm(cams)), "point_norm": float(np.linalg.norm(pts)),
            "sample_index": idx.tolist(), "sample_values": x[idx].tolist()}


def main():
    workload, seed, tol, steps = sys.argv[1], int(sys.argv[2]), float(sys.argv[3]), int(sys.argv[4])
    tag = sys.argv[5] if len(sys.argv) > 5 else f"cg{tol:g}_{steps}"
    p = synthetic.generate(workload, seed=seed)
    t = time.time()
    ba = oracle.BA(p, huber=p.huber)
    out, trace = ba.lm(options=oracle.OptimizeOptions(steps, 1e-8, 0.0, 0.0), tol=tol, trace_capacity=steps)
    el = time.time() - t
    cams, pts = ba.state()
    res = {"workload": workload, "seed": seed, "cg_relative_tolerance": tol, "max_iterations": steps,
           "problem": problem_digest(p), "status": out.status, "initial_cost": out.initial_cost,
           "final_cost": out.cost, "accepted_steps": out.accepted_steps, "rejected_steps": out.rejected_steps,
           "attempts": out.attempts, "linearizations": out.linearizations, "cg_iterations": out.linear_iterations,
           "gradient_norm": out.gradient_norm, "trace": trace.tolist(), "state": state_digest(cams, pts),
           "oracle_threads": oracle.get_threads(), "seconds": el}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"trajectory_{workload}_{seed}_{tag}.json")
    with open(path, "w") as f:
        json.dump(res, f, separators=(",", ":"))
    print(json.dumps({k: v for k, v in res.items() if k not in ("trace", "state")}))


if __name__ == "__main__":
    main()
