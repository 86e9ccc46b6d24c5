"""What the shipped preconditioner network gains over block-Jacobi, measured (GPU only).

The network (fagra_amd/learned_precond.f32, trained by fagra_amd/train_precond.py on Schur
systems that LM emits on seeded 256- and 400-camera problems, damping 1e-12 .. 1e12) corrects
the block-Jacobi inverse (B_i + lambda I)^-1 (src/schur.rs:658-682) towards the inverse of the
camera's diagonal block of the reduced system, S_cc. Two questions, each on held-out inputs:

* Does it learn that map? Its loss ||G^T A G - I||^2 on held-out systems against G = I
  (block-Jacobi): measured 0.434 -> 0.0034.
* Does it cut PCG iterations to the same tolerance (1e-3, the benchmark's)? On these problems
  block-Jacobi PCG already converges in 4-27 iterations per solve and the exact S_cc blocks
  are no better a preconditioner than B_i, so the network's iterations stay within a few per
  cent of block-Jacobi's: measured 393 against 390 over the held-out grid below (and 917
  against 942 over the benchmark's LM solve to its target). On late long-track states the
  network's blocks converge slower than block-Jacobi's; the fallback (residual not down by 0.2
  within 8 iterations) restarts those solves with block-Jacobi. The test asserts the 5% bound
  over the solves both preconditioners finish, and that cameras outside the network's input
  range fall back to block-Jacobi.

In f32 storage the per-point solves handed between the two halves of the Schur product are
rounded, which perturbs the PCG curvature at first order; at damping 1e-10 and below, along the
similarity gauge of BA, a solve may stop on a nonpositive curvature (NotPositiveDefinite) or
stall until the iteration limit (IterationLimit); LM answers either with a larger damping.
Which of these a given late state meets at those dampings moves with the last bits of the LM
trajectory that produced the state. Such solves are allowed at those dampings only and left out
of the comparison.
"""
import numpy as np
import pytest

import fagra_amd as fa
from fagra_amd import learned, synthetic

pytestmark = pytest.mark.gpu

LAMBDAS = [1e-12, 1e-10, 1e-8, 1e-6, 1e-4, 1e-3, 1e-2, 1e-1]


def iterations(ba, precond, net, tol=1e-3):
    """PCG iterations per damping in LAMBDAS (None: at damping 1e-10 or below, stopped on a
    nonpositive curvature or at the iteration limit)."""
    s = fa.Schur(precision=fa.FP32, relative_tolerance=tol, preconditioner=precond, network=net)
    assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
    out = []
    for lam in LAMBDAS:
        st, _, _ = s.solve_damped(lam)
        d = s.diagnostics()
        if d.termination in (fa.CG_NOT_POSITIVE_DEFINITE, fa.CG_ITERATION_LIMIT) and lam <= 1e-10:
            out.append(None)
            continue
        assert st == fa.OK and d.termination == fa.CG_CONVERGED, (lam, st, d.termination)
        out.append(d.iterations)
    s.close()
    return out


def both_finished(a, b):
    return [(x, y) for x, y in zip(a, b) if x is not None and y is not None]


def at_lm_state(p, steps):
    ba = fa.BalProblem(p, precision=fa.FP32)
    if steps:
        lm = fa.LevenbergMarquardt(fa.Schur(precision=fa.FP32, relative_tolerance=1e-3))
        lm.solve_batch(ba, fa.OptimizeOptions(steps, 0.0, 0.0, 0.0))
    return ba


@pytest.fixture(scope="module")
def net():
    if not learned.available():
        pytest.skip("no trained preconditioner weights")
    return learned.Network.load()


def test_network_fits_the_schur_diagonal_on_held_out_systems(net):
    feats, targs = learned.collect([1000], max_iter=40)
    identity = learned.Network(seed=0)  # zero output layer: G = I, plain block-Jacobi
    base = identity.loss_grad(feats, targs)[0]
    trained = net.loss_grad(feats, targs)[0]
    print(f"held-out loss: block-Jacobi {base:.4f}, network {trained:.5f} over {len(feats)} camera systems")
    assert trained < 0.05 * base


def test_pcg_iterations_against_block_jacobi(net):
    cases = [("ring256", 1000, {}, (0, 20, 60)), ("rings1500", 0, {}, (0, 106)),
             ("zipf4000", 1000, {"n_cameras": 400, "n_points": 60000}, (0, 20))]
    tot_bj = tot_ln = 0
    for name, seed, overrides, states in cases:
        p = synthetic.generate(name, seed=seed, **overrides)
        for steps in states:
            ba = at_lm_state(p, steps)
            bj = iterations(ba, fa.PRECOND_BLOCK_JACOBI, None)
            ln = iterations(ba, fa.PRECOND_LEARNED, net)
            print(f"{name} seed {seed} after {steps} LM steps: block-Jacobi {bj} learned {ln}", flush=True)
            pairs = both_finished(bj, ln)
            tot_bj += sum(x for x, _ in pairs)
            tot_ln += sum(y for _, y in pairs)
            ba.close()
    print(f"PCG iterations to 1e-3 over the grid: block-Jacobi {tot_bj}, learned {tot_ln}")
    assert tot_ln <= 1.05 * tot_bj


def test_cameras_outside_the_input_range_keep_block_jacobi(net):
    p = synthetic.generate("ring256", seed=1001)
    ba = at_lm_state(p, 0)
    gated = learned.Network(params=net.params())
    big = np.full(learned.N_IN, 1e30, dtype=np.float32)
    gated.set_input_range(big, big)  # no camera's features lie in [1e30, 1e30]
    bj = iterations(ba, fa.PRECOND_BLOCK_JACOBI, None)
    ln = iterations(ba, fa.PRECOND_LEARNED, gated)
    print(f"gated out: block-Jacobi {bj} learned {ln}")
    assert all(abs(a - b) <= 1 for a, b in both_finished(bj, ln))
