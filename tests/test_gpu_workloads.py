"""Final cost and state of complete LM solves on the workload-sized problems (GPU only).

ring16 (16 cameras, 2,000 points, 11,934 observations): both precisions run to convergence
beside the oracle at the reference's default CG tolerance 1e-10. ring256 (256 cameras,
60,000 points, 330,753 observations): the FP64 path follows the oracle's trajectory for six
LM steps run live and for the 40 steps of the committed oracle fixture
(tests/golden/trajectory_ring256_0_cg1e-08_40.json, CG tolerance 1e-8), and the FP32 path runs
to convergence beside the FP64 path and the fixture. Bounds: final cost 1e-9 relative for FP64,
1e-5 for FP32 (the reference's own relative-cost bar for the reduced-precision path). State
error is the relative 2-norm difference of all camera and point parameters (of the fixture's
2,048 sampled coordinates against a fixture); every figure is printed.
"""
import json
import os

import numpy as np
import pytest

import fagra_amd as fa
from fagra_amd import synthetic
from oracle import oracle

pytestmark = pytest.mark.gpu


def state_error(a, b):
    xa = np.concatenate([np.ravel(a[0]), np.ravel(a[1])])
    xb = np.concatenate([np.ravel(b[0]), np.ravel(b[1])])
    return float(np.linalg.norm(xa - xb) / np.linalg.norm(xb))


def gpu_lm(p, precision, max_iter, tol):
    ba = fa.BalProblem(p, precision=precision)
    lm = fa.LevenbergMarquardt(fa.Schur(precision=precision, relative_tolerance=tol))
    out = lm.solve_batch(ba, fa.OptimizeOptions(max_iter, 1e-8, 0.0, 0.0))
    return out, ba.state()


def oracle_lm(p, max_iter):
    ba = oracle.BA(p, huber=p.huber)
    out, _ = ba.lm(options=oracle.OptimizeOptions(max_iter, 1e-8, 0.0, 0.0),
                   lm_options=oracle.LmOptions.default(), tol=1e-10)
    return out, ba.state()


def test_ring16_converged_cost_and_state_match_oracle():
    p = synthetic.generate("ring16", seed=0)
    out_o, x_o = oracle_lm(p, 60)
    assert out_o.status in (oracle.OK, oracle.NO_PROGRESS)
    for precision, tol, cost_bar, state_bar in [(fa.FP64, 1e-10, 1e-9, 1e-6), (fa.FP32, None, 1e-5, 1e-4)]:
        out, x = gpu_lm(p, precision, 60, tol)
        assert out.status in (fa.OK, fa.NO_PROGRESS), (precision, out.status)
        dc = abs(out.cost - out_o.cost) / out_o.cost
        dx = state_error(x, x_o)
        print(f"ring16 precision {precision}: cost {out.cost!r} oracle {out_o.cost!r} rel {dc:.3e} "
              f"state rel {dx:.3e} steps {out.accepted_steps}/{out_o.accepted_steps}")
        assert dc <= cost_bar
        assert dx <= state_bar


def test_ring256_fp64_trajectory_matches_oracle():
    p = synthetic.generate("ring256", seed=0)
    out_o, x_o = oracle_lm(p, 6)
    out, x = gpu_lm(p, fa.FP64, 6, 1e-10)
    assert out.status == out_o.status
    assert out.accepted_steps == out_o.accepted_steps and out.rejected_steps == out_o.rejected_steps
    dc = abs(out.cost - out_o.cost) / out_o.cost
    dx = state_error(x, x_o)
    print(f"ring256 fp64 6 steps: cost {out.cost!r} oracle {out_o.cost!r} rel {dc:.3e} state rel {dx:.3e}")
    assert dc <= 1e-9
    assert dx <= 1e-6


def ring256_fixture():
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "trajectory_ring256_0_cg1e-08_40.json")) as f:
        return json.load(f)


def sampled_error(x, fx):
    v = np.concatenate([np.ravel(x[0]), np.ravel(x[1])])[np.asarray(fx["state"]["sample_index"])]
    ref = np.asarray(fx["state"]["sample_values"])
    return float(np.linalg.norm(v - ref) / np.linalg.norm(ref))


def test_ring256_fp64_follows_oracle_fixture_for_40_steps():
    p = synthetic.generate("ring256", seed=0)
    fx = ring256_fixture()
    assert fx["problem"]["pixel_sum"] == float(np.sum(p.obs_pixel))
    ba = fa.BalProblem(p, precision=fa.FP64)
    lm = fa.LevenbergMarquardt(fa.Schur(precision=fa.FP64, relative_tolerance=fx["cg_relative_tolerance"]))
    trace = []
    lm.on_accept = lambda st: trace.append((st.cost, st.attempts))
    out = lm.solve_batch(ba, fa.OptimizeOptions(fx["max_iterations"], 1e-8, 0.0, 0.0))
    worst = max(abs(c - r[0]) / r[0] for (c, _), r in zip(trace, fx["trace"]))
    dx = sampled_error(ba.state(), fx)
    print(f"ring256 fp64 {len(trace)} steps vs fixture: final {out.cost!r} oracle {fx['final_cost']!r} "
          f"worst step rel {worst:.3e} sampled state rel {dx:.3e}")
    assert len(trace) == len(fx["trace"]) and [a for _, a in trace] == [int(r[1]) for r in fx["trace"]]
    assert worst <= 1e-9
    assert dx <= 1e-6


def test_ring256_fp32_converges_to_fp64_cost():
    p = synthetic.generate("ring256", seed=0)
    out64, x64 = gpu_lm(p, fa.FP64, 200, 1e-10)
    out32, x32 = gpu_lm(p, fa.FP32, 200, None)
    dc = abs(out32.cost - out64.cost) / out64.cost
    dx = state_error(x32, x64)
    fx = ring256_fixture()
    dfx = (out32.cost - fx["final_cost"]) / fx["final_cost"]
    print(f"ring256 converged: fp64 {out64.cost!r} ({out64.status}, {out64.accepted_steps} steps) "
          f"fp32 {out32.cost!r} ({out32.status}, {out32.accepted_steps} steps) rel {dc:.3e} state rel {dx:.3e}; "
          f"fp32 vs the oracle's 40-step cost {dfx:.3e}, sampled state rel {sampled_error(x32, fx):.3e}")
    assert dc <= 1e-5
    assert dx <= 1e-3
    assert dfx <= 1e-5
