"""Converged cost and state of the CPU oracle's LM on a seeded workload (golden fixture).

Usage: python tests/golden/make_converged.py WORKLOAD SEED TOTAL_STEPS SEGMENT CHECKPOINT_DIR

Runs the oracle's f64 LM + Schur-CG (LM defaults, CG relative tolerance 1e-10, gradient
tolerance 1e-8, step and cost tolerances 0) from the workload's initial estimates in segments of
SEGMENT accepted steps; each segment starts a new LM call from the state the last one accepted
(so its damping restarts at the default 1e-3). After every segment the state is saved in
CHECKPOINT_DIR, and a later run of the same command resumes from there, so a long run survives
interruptions. When TOTAL_STEPS steps are done (or a segment ends by the gradient tolerance),
writes tests/golden/converged_<workload>_<seed>.json: the final cost, the cost after every
segment, the relative cost decrease over the last segment, and the state digest of
make_trajectory.py.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from fagra_amd import synthetic  # noqa: E402
from make_trajectory import problem_digest, state_digest  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    workload, seed, total, segment, ckdir = (sys.argv[1], int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4]),
                                             sys.argv[5])
    os.makedirs(ckdir, exist_ok=True)
    ck = os.path.join(ckdir, f"converged_{workload}_{seed}.npz")
    p = synthetic.generate(workload, seed=seed)
    ba = oracle.BA(p, huber=p.huber)
    done, costs, seconds, status = 0, [], 0.0, None
    if os.path.exists(ck):
        z = np.load(ck)
        ba.set_state(z["cameras"], z["points"])
        done, costs, seconds = int(z["done"]), z["costs"].tolist(), float(z["seconds"])
        print(f"resumed at {done} steps, cost {costs[-1]!r}", flush=True)
    initial = oracle.BA(p, huber=p.huber).cost()[1]
    while done < total:
        t = time.time()
        n = min(segment, total - done)
        out, _ = ba.lm(options=oracle.OptimizeOptions(n, 1e-8, 0.0, 0.0), tol=1e-10)
        seconds += time.time() - t
        done += out.accepted_steps
        costs.append(out.cost)
        status = out.status
        cams, pts = ba.state()
        np.savez(ck, cameras=cams, points=pts, done=done, costs=np.array(costs), seconds=seconds)
        print(f"{done} steps: cost {out.cost!r} status {out.status} ({seconds:.0f}s)", flush=True)
        if out.status != oracle.NO_CONVERGENCE or out.accepted_steps == 0:
            break  # the gradient tolerance (or no further progress) ends the run
    cams, pts = ba.state()
    res = {"workload": workload, "seed": seed, "cg_relative_tolerance": 1e-10, "segment_steps": segment,
           "accepted_steps": done, "last_status": status, "problem": problem_digest(p),
           "initial_cost": initial, "final_cost": costs[-1], "segment_costs": costs,
           "rel_decrease_last_segment": (costs[-2] - costs[-1]) / costs[-1] if len(costs) > 1 else None,
           "state": state_digest(cams, pts), "oracle_threads": oracle.get_threads(), "seconds": seconds}
    path = os.path.join(HERE, f"converged_{workload}_{seed}.json")
    with open(path, "w") as f:
        json.dump(res, f, separators=(",", ":"))
    print(json.dumps({k: v for k, v in res.items() if k not in ("state", "segment_costs")}))


if __name__ == "__main__":
    main()
