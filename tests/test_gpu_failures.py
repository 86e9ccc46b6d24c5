"""Failure semantics of the BA path on the GPU against the oracle (GPU only).

The reference's camera model rejects a projection with zero depth or a nonfinite pixel
(examples/bal/model.rs:147-156) and the sink rejects nonfinite residuals and Jacobians
(src/optimization.rs:556,597): both surface as EvaluationError::InvalidEvaluation from the cost
(src/storage.rs:612-618) and from linearization, and LM aborts on them at the initial state but
turns an invalid trial cost into a rejected trial (src/optimization/levenberg_marquardt.rs:
408-430). Nonfinite pixels never reach the model: the BAL parser rejects them
(examples/bal/model.rs:78-80) and so does the device upload. Each case runs on both
precisions beside the oracle.
"""
import numpy as np
import pytest

import fagra_amd as fa
from fagra_amd import synthetic
from oracle import oracle

pytestmark = pytest.mark.gpu
PRECISIONS = [fa.FP64, fa.FP32]


def small(seed=4):
    return synthetic.generate("ring16", seed=seed, n_cameras=6, n_points=60, huber=1.0)


def zero_depth(p):
    """Camera 0 at the identity rotation with t = (0, 0, 10) and one of its points at z = -10:
    depth P.z = z + 10 is exactly zero."""
    j = int(p.obs_point[p.obs_camera == 0][0])
    p.cameras[0, :6] = [0.0, 0.0, 0.0, 0.0, 0.0, 10.0]
    p.points[j] = [0.3, -0.2, -10.0]
    return p


def gpu_statuses(p, precision, state=None):
    ba = fa.BalProblem(p, precision=precision)
    if state is not None:
        ba.set_state(*state)
    st_cost, _ = ba.cost()
    s = fa.Schur(precision=precision)
    assert s.prepare(ba) == fa.OK
    st_lin = s.linearize()
    if st_lin == fa.OK:  # the evaluation flag of a linearization is read with the gradient norm
        st_lin = s.gradient_norm()[0]
    if state is not None:
        ba.set_state(*state)
    out = fa.LevenbergMarquardt(fa.Schur(precision=precision)).solve_batch(ba, fa.OptimizeOptions(5, 1e-8, 0, 0))
    return st_cost, st_lin, out


def oracle_statuses(p, state=None):
    ba = oracle.BA(p, huber=p.huber)
    if state is not None:
        ba.set_state(*state)
    st_cost, _ = ba.cost()
    st_lin, _ = oracle.ba_schur(ba)
    ba = oracle.BA(p, huber=p.huber)
    if state is not None:
        ba.set_state(*state)
    out, _ = ba.lm(options=oracle.OptimizeOptions(5, 1e-8, 0.0, 0.0))
    return st_cost, st_lin, out


@pytest.mark.parametrize("precision", PRECISIONS)
def test_zero_depth_is_an_invalid_evaluation(precision):
    p = zero_depth(small())
    o_cost, o_lin, o_out = oracle_statuses(p)
    g_cost, g_lin, g_out = gpu_statuses(p, precision)
    print(f"zero depth, precision {precision}: cost {g_cost}/{o_cost}, linearize {g_lin}/{o_lin}, "
          f"LM {g_out.status}/{o_out.status}", flush=True)
    assert o_cost == o_lin == o_out.status == oracle.INVALID_EVALUATION
    assert g_cost == g_lin == g_out.status == fa.INVALID_EVALUATION
    assert g_out.attempts == o_out.attempts == 0 and g_out.cost_evaluations == o_out.cost_evaluations == 1


@pytest.mark.parametrize("precision", PRECISIONS)
def test_nonfinite_state_is_an_invalid_evaluation(precision):
    # a nonfinite state (as a diverged step would leave it), set after a valid upload
    p = small(seed=5)
    pts = p.points.copy()
    pts[int(p.obs_point[0]), 1] = np.nan
    state = (p.cameras, pts)
    o_cost, o_lin, o_out = oracle_statuses(p, state)
    assert o_cost == o_lin == o_out.status == oracle.INVALID_EVALUATION
    # the device evaluator refuses such a state where it is written, with the same status, and
    # keeps the state it had
    ba = fa.BalProblem(p, precision=precision)
    with pytest.raises(fa.FagraError) as err:
        ba.set_state(*state)
    assert err.value.status == fa.INVALID_EVALUATION
    assert ba.cost() == fa.BalProblem(p, precision=precision).cost()
    print(f"nonfinite point, precision {precision}: oracle cost/linearize/LM {o_cost}/{o_lin}/{o_out.status}, "
          f"device set_state {err.value.status}", flush=True)


def test_nonfinite_pixel_is_rejected_before_the_model():
    p = small(seed=6)
    p.obs_pixel[7, 0] = np.nan
    with pytest.raises(fa.FagraError):
        fa.BalProblem(p, precision=fa.FP64)
    text = synthetic.generate("ring16", seed=6, n_cameras=3, n_points=5).to_bal_text().split("\n")
    first_obs = text[1].split()
    first_obs[2] = "NaN"
    text[1] = " ".join(first_obs)
    from fagra_amd import bal
    with pytest.raises(bal.BalError, match="invalid observation index or value"):
        bal.parse("\n".join(text))


def near_camera_plane(seed=21):
    """A point moved to 0.05 behind one of its cameras' image plane; its other observations pull
    it through that plane, where its reprojection in that camera diverges."""
    p = synthetic.generate("ring16", seed=seed, n_cameras=6, n_points=40, huber=1.0)
    R = synthetic.rotation_matrices(p.cameras[:, :3])
    c = int(p.obs_camera[p.obs_point == 0][0])
    P = R[c] @ p.points[0] + p.cameras[c, 3:6]
    P[2] = 0.05
    p.points[0] = np.linalg.solve(R[c], P - p.cameras[c, 3:6])
    return p, c


def test_trials_across_a_camera_plane_are_rejected_as_the_oracle_rejects_them():
    p, c = near_camera_plane()
    ba_o = oracle.BA(p, huber=p.huber)
    out_o, trace_o = ba_o.lm(options=oracle.OptimizeOptions(20, 1e-8, 0.0, 0.0), tol=1e-10, trace_capacity=20)
    ba = fa.BalProblem(p, precision=fa.FP64)
    lm = fa.LevenbergMarquardt(fa.Schur(precision=fa.FP64, relative_tolerance=1e-10))
    trace = []
    lm.on_accept = lambda st: trace.append((st.cost, st.attempts, st.last_rejection))
    out = lm.solve_batch(ba, fa.OptimizeOptions(20, 1e-8, 0.0, 0.0))
    worst = max(abs(t[0] - r[0]) / r[0] for t, r in zip(trace, trace_o))
    print(f"camera plane: gpu {out.accepted_steps} accepted / {out.rejected_steps} rejected, last rejection "
          f"{out.last_rejection}; oracle {out_o.accepted_steps} / {out_o.rejected_steps}, {out_o.last_rejection}; "
          f"worst step cost rel {worst:.2e}", flush=True)
    assert out_o.rejected_steps > 0 and out_o.last_rejection == fa.REJ_POOR_AGREEMENT
    assert (out.status, out.accepted_steps, out.rejected_steps, out.last_rejection) == \
        (out_o.status, out_o.accepted_steps, out_o.rejected_steps, out_o.last_rejection)
    assert [t[1] for t in trace] == [int(r[1]) for r in trace_o]
    # the point's projection in camera c is near-singular along the way, which amplifies the
    # rounding differences of summation order: measured 1.3e-9 at worst over the 20 steps
    assert worst <= 1e-8


@pytest.mark.parametrize("backend", ["schur", "lsmr"])
def test_problem_destroyed_before_its_backend(backend):
    """A backend outlives the problem it was prepared on: the problem's destruction unbinds it,
    later calls report INVALID_OPTIONS instead of reaching freed memory, and the backend
    prepares again on a new problem and solves as a fresh one does."""
    make = (lambda: fa.Schur(precision=fa.FP64)) if backend == "schur" else (lambda: fa.Lsmr(precision=fa.FP64))
    p = small()
    ba = fa.BalProblem(p, precision=fa.FP64)
    s = make()
    assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
    st, first, _ = s.solve_damped(1e-3)
    assert st == fa.OK
    ba.close()
    assert s.linearize() == fa.INVALID_OPTIONS
    assert s.prepare_damping(1e-6) == fa.INVALID_OPTIONS
    assert s.solve_damped(1e-3)[0] == fa.INVALID_OPTIONS
    ba2 = fa.BalProblem(p, precision=fa.FP64)
    assert s.prepare(ba2) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
    st, again, _ = s.solve_damped(1e-3)
    assert st == fa.OK and np.array_equal(again, first)
    ba2.close()
    s.close()  # destroying an unbound backend
