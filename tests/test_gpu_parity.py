"""HIP path vs the CPU oracle on the same seeded inputs (GPU only).

Tolerances: FP64 GPU path -- products 1e-10 relative, solves and final LM cost 1e-9 relative;
FP32 path (FP32 Jacobians, Schur products and PCG matvecs, FP64 reductions) -- products 2e-5
relative, steps 3e-3 relative (the reference's own f32 Schur bound, src/schur.rs:788), final
LM cost 1e-5 relative. Integer layout is bit-exact.
"""
import numpy as np
import pytest

import fagra_amd as fa
from fagra_amd import synthetic
from oracle import oracle

pytestmark = pytest.mark.gpu

PRECISIONS = [fa.FP64, fa.FP32]
PROD_TOL = {fa.FP64: 1e-10, fa.FP32: 2e-5}


def small(seed=0, nc=8, npts=300, huber=0.0, **kw):
    return synthetic.generate("ring16", seed=seed, n_cameras=nc, n_points=npts, huber=huber, **kw)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def linearized_pair(p, precision, floor=1e-6):
    ba_o = oracle.BA(p, huber=p.huber)
    st, s_o = oracle.ba_schur(ba_o)
    assert st == oracle.OK
    assert s_o.prepare_damping(floor) == oracle.OK
    ba_g = fa.BalProblem(p, precision=precision)
    s_g = fa.Schur(precision=precision)
    assert s_g.prepare(ba_g) == fa.OK
    assert s_g.linearize() == fa.OK
    assert s_g.prepare_damping(floor) == fa.OK
    return ba_o, s_o, ba_g, s_g


def test_layout_is_bit_exact():
    for seed, nc, npts in [(0, 8, 300), (3, 16, 2000), (5, 3, 50)]:
        p = small(seed=seed, nc=nc, npts=npts)
        g = fa.BalProblem(p).layout()
        o = oracle.BA(p).layout()
        for a, b in zip(g, o):
            assert a.dtype == b.dtype and np.array_equal(a, b)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("huber", [0.0, 1.0])
def test_cost_matches_oracle(precision, huber):
    p = small(seed=1, huber=huber)
    st_o, c_o = oracle.BA(p, huber=huber).cost()
    st_g, c_g = fa.BalProblem(p, precision=precision).cost()
    assert st_o == st_g == 0
    tol = 1e-12 if precision == fa.FP64 else 1e-5
    assert abs(c_g - c_o) <= tol * c_o


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("huber", [0.0, 1.0])
def test_linearized_model_matches_oracle(precision, huber):
    p = small(seed=2, huber=huber)
    ba_o, s_o, ba_g, s_g = linearized_pair(p, precision)
    g, d, nr = s_g.model()
    tol = PROD_TOL[precision]
    assert rel(g, s_o.gradient()) < 10 * tol
    assert rel(d, s_o.scales()) < tol
    assert rel(nr, s_o.negative_residual()) < tol
    st, gn = s_g.gradient_norm()
    st_o, gn_o = s_o.gradient_norm()
    assert st == st_o == 0 and abs(gn - gn_o) <= 10 * tol * gn_o
    gram = s_g.camera_gram()
    for c in range(p.n_cameras):
        st, go = s_o.gram_block(9 * c, 9)
        go = np.tril(go) + np.tril(go, -1).T
        assert rel(gram[c], go) < tol


@pytest.mark.parametrize("precision", PRECISIONS)
def test_reduced_apply_matches_oracle(precision):
    p = small(seed=3, huber=1.0)
    ba_o, s_o, ba_g, s_g = linearized_pair(p, precision)
    rng = np.random.default_rng(0)
    for lam in [1e-3, 1.0, 100.0]:
        x = rng.normal(size=9 * p.n_cameras)
        st_o, y_o = s_o.reduced_apply(lam, x)
        st_g, y_g = s_g.reduced_apply(lam, x)
        assert st_o == st_g == 0
        assert rel(y_g, y_o) < 5 * PROD_TOL[precision], lam


@pytest.mark.parametrize("precision", PRECISIONS)
def test_jacobian_products_match_oracle(precision):
    p = small(seed=4, huber=1.0)
    ba_o, s_o, ba_g, s_g = linearized_pair(p, precision)
    rng = np.random.default_rng(1)
    n, rows = 9 * p.n_cameras + 3 * p.n_points, 2 * p.n_observations
    x, y = rng.normal(size=n), rng.normal(size=rows)
    tol = PROD_TOL[precision]
    assert rel(s_g.jacobian_apply(x), s_o.jacobian_apply(x)) < tol
    assert rel(s_g.jacobian_transpose_apply(y), s_o.jacobian_transpose_apply(y)) < tol
    yy = rng.normal(size=rows + n)
    for lam in [0.0, 0.01, 1e8]:
        for scaled in [False, True]:
            assert rel(s_g.damped_apply(lam, scaled, x), s_o.damped_apply(lam, scaled, x)) < tol
            assert rel(s_g.damped_transpose_apply(lam, scaled, yy), s_o.damped_transpose_apply(lam, scaled, yy)) < tol


@pytest.mark.parametrize("precision", PRECISIONS)
def test_solve_damped_matches_oracle(precision):
    p = small(seed=5, huber=1.0)
    ba_o, s_o, ba_g, s_g = linearized_pair(p, precision)
    tol = 1e-10 if precision == fa.FP64 else 1e-6
    s_o.relative_tolerance = tol
    s_g.relative_tolerance = tol
    for lam in [1e-3, 1.0, 100.0, 1e-2]:
        st_o, d_o, pr_o = s_o.solve_damped(lam)
        st_g, d_g, pr_g = s_g.solve_damped(lam)
        assert st_o == st_g == 0, (lam, s_g.diagnostics().termination)
        dg = s_g.diagnostics()
        assert dg.termination == fa.CG_CONVERGED
        bound = 1e-8 if precision == fa.FP64 else 3e-3
        assert rel(d_g, d_o) < bound, lam
        assert abs(pr_g - pr_o) <= bound * abs(pr_o), lam
        # iteration counts agree in order of magnitude (faer's exact CG is not reachable)
        do = s_o.diagnostics()
        assert dg.iterations <= 3 * do.iterations + 3
        assert dg.reduced_relative_residual < 10 * tol * max(1, 10 if precision == fa.FP32 else 1)


def test_iteration_limit_and_invalid_lambda():
    p = small(seed=6)
    ba_o, s_o, ba_g, s_g = linearized_pair(p, fa.FP64)
    s_g.max_iterations = 1
    st, d, pr = s_g.solve_damped(1e-3)
    assert st == fa.LINEAR_SOLVE_FAILED
    diag = s_g.diagnostics()
    assert diag.termination == fa.CG_ITERATION_LIMIT and diag.iterations == 1
    assert np.isnan(diag.estimated_relative_residual)
    assert np.isfinite(diag.reduced_relative_residual) and diag.reduced_relative_residual > s_g.relative_tolerance
    s_g.max_iterations = 1000
    assert s_g.solve_damped(float("nan"))[0] == fa.INVALID_OPTIONS
    assert s_g.solve_damped(-1.0)[0] == fa.INVALID_OPTIONS


@pytest.mark.parametrize("precision", PRECISIONS)
def test_lm_final_cost_matches_oracle(precision):
    p = small(seed=7, nc=8, npts=400, huber=1.0)
    opts = oracle.OptimizeOptions(30, 1e-8, 0.0, 0.0)
    lm_o = oracle.LmOptions.default()
    out_o, _ = oracle.BA(p, huber=1.0).lm(options=opts, lm_options=lm_o, tol=1e-10)
    ba = fa.BalProblem(p, precision=precision)
    lm = fa.LevenbergMarquardt(fa.Schur(precision=precision,
                                        relative_tolerance=1e-10 if precision == fa.FP64 else None))
    o = fa.OptimizeOptions(30, 1e-8, 0.0, 0.0)
    out = lm.solve_batch(ba, o)
    assert out.status in (fa.OK, fa.NO_PROGRESS, fa.NO_CONVERGENCE)
    assert out_o.status in (oracle.OK, oracle.NO_PROGRESS, oracle.NO_CONVERGENCE)
    tol = 1e-9 if precision == fa.FP64 else 1e-5
    assert abs(out.cost - out_o.cost) <= tol * out_o.cost, (out.cost, out_o.cost)
    # LmStatistics invariants (tests/levenberg_marquardt.rs:253-256)
    assert out.attempts == out.accepted_steps + out.rejected_steps or out.status != fa.OK
    st, c = ba.cost()
    assert abs(c - out.cost) <= 1e-12 * c


def test_gauss_newton_first_step_matches_oracle():
    # Undamped BA systems are singular along the 7-dof similarity gauge; the first GN step is
    # well defined (consistent right-hand side), later ones can stall either implementation's
    # CG near convergence. Compare one step exactly and bound the rest.
    p = small(seed=8, nc=6, npts=200)
    ba = fa.BalProblem(p)
    gn = fa.GaussNewton(fa.Schur())
    out = gn.solve_batch(ba, fa.OptimizeOptions(1, 1e-8, 1e-10, 1e-12))
    out_o = oracle.BA(p).gn(options=oracle.OptimizeOptions(1, 1e-8, 1e-10, 1e-12))
    assert out.status == out_o.status == fa.NO_CONVERGENCE
    assert out.accepted_steps == out_o.accepted_steps == 1
    assert abs(out.cost - out_o.cost) <= 1e-9 * out_o.cost
    ba2 = fa.BalProblem(p)
    out = fa.GaussNewton(fa.Schur()).solve_batch(ba2, fa.OptimizeOptions(5, 1e-8, 1e-10, 1e-12))
    assert out.status in (fa.OK, fa.LINEAR_SOLVE_FAILED, fa.NO_CONVERGENCE)
    assert out.cost < out.initial_cost


@pytest.mark.parametrize("precision", PRECISIONS)
def test_graph_batches_match_eager_launches(precision):
    """PCG batches replayed from the captured HIP graph (the default) and launched kernel by
    kernel (collect_timings) produce the same steps, over several solves and damping values
    with eager work between the replays."""
    p = small(seed=9, nc=10, npts=500, huber=1.0)
    runs = {}
    for timed in (False, True):
        ba = fa.BalProblem(p, precision=precision)
        s = fa.Schur(precision=precision, relative_tolerance=1e-8, collect_timings=timed)
        assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
        out = []
        for lam in [1e-3, 1.0, 1e-6, 1e-3, 100.0, 1e-12]:
            st, d, pr = s.solve_damped(lam)
            # lambda = 1e-12 leaves the 7-dimensional gauge of BA nearly singular: in f32 the
            # PCG may stop on a nonpositive curvature there, the same way in both launch modes
            assert st == fa.OK or lam < 1e-9, (timed, lam, s.diagnostics().termination)
            out.append((st, d, pr, s.diagnostics().iterations, s.diagnostics().termination))
            ba.cost()  # eager work between replays
        runs[timed] = out
    for (s0, d0, p0, i0, t0), (s1, d1, p1, i1, t1) in zip(runs[False], runs[True]):
        assert s0 == s1 and t0 == t1 and i0 == i1
        if s0 == fa.OK:
            assert np.array_equal(d0, d1) and p0 == p1


@pytest.mark.skip(reason="the LDS-resident f32 point pass (k_points_lds) left the GPU faulted on its first run "
                         "and its cause is not found; the kernel and its selection hook are out of the library "
                         "until it is")
def test_lds_point_passes_match_global_gathers(monkeypatch):
    """The f32 point passes with every camera held in LDS (k_points_lds: the Schur product's
    elimination and the attempt's recovery/verification sweep) equal the same passes with
    per-observation global gathers (FAGRA_LDS_POINTS=0): the same values go through the same
    arithmetic, so the product and the step agree bit for bit; the verification partials are
    summed over other workgroups and agree to rounding."""
    p = small(seed=21, nc=48, npts=4000, huber=1.0)
    x = np.random.default_rng(5).normal(size=9 * p.n_cameras)
    runs = []
    for flag in (None, "0"):
        if flag is None:
            monkeypatch.delenv("FAGRA_LDS_POINTS", raising=False)
        else:
            monkeypatch.setenv("FAGRA_LDS_POINTS", flag)
        ba = fa.BalProblem(p, precision=fa.FP32)
        s = fa.Schur(precision=fa.FP32)
        assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
        lds = s.lds_workgroups()
        st, y = s.reduced_apply(1e-3, x)
        assert st == fa.OK
        solves = []
        for lam in (1e-3, 1e-1):
            st, d, pr = s.solve_damped(lam)
            assert st == fa.OK
            dg = s.diagnostics()
            solves.append((d, pr, dg.iterations, dg.reduced_relative_residual, dg.original_relative_residual))
        runs.append((lds, y, solves))
    (lds_a, y_a, sa), (lds_b, y_b, sb) = runs
    assert lds_a > 0 and lds_b == 0
    assert np.array_equal(y_a, y_b)
    for (d0, p0, i0, r0, o0), (d1, p1, i1, r1, o1) in zip(sa, sb):
        assert i0 == i1 and np.array_equal(d0, d1)
        assert abs(p0 - p1) <= 1e-12 * abs(p1)
        assert abs(r0 - r1) <= 1e-9 * abs(r1) and abs(o0 - o1) <= 1e-9 * abs(o1)
    # f64 keeps the gathers
    ba = fa.BalProblem(p, precision=fa.FP64)
    s = fa.Schur(precision=fa.FP64)
    assert s.prepare(ba) == fa.OK and s.lds_workgroups() == 0


@pytest.mark.parametrize("precision", PRECISIONS)
def test_camera_major_records_equal_point_major_records(precision):
    """The camera-major linearization pass forms its records from the points again; every
    record and weighted residual must equal the point-major copy bit for bit, so that both
    halves of the Schur products use one Jacobian (also at a second, perturbed state and with
    Huber weights active)."""
    for seed, nc, npts, huber in [(31, 12, 900, 0.0), (32, 40, 5000, 1.0)]:
        p = small(seed=seed, nc=nc, npts=npts, huber=huber)
        ba = fa.BalProblem(p, precision=precision)
        s = fa.Schur(precision=precision)
        assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK
        assert s.record_mismatches() == 0
        cams, pts = ba.state()
        rng = np.random.default_rng(seed)
        ba.set_state(cams + 1e-3 * rng.normal(size=cams.shape), pts + 1e-2 * rng.normal(size=pts.shape))
        assert s.linearize() == fa.OK
        assert s.record_mismatches() == 0


@pytest.mark.parametrize("precision", PRECISIONS)
def test_split_chunks_match_oracle(precision, monkeypatch):
    """Every chunk swept by the four waves of one workgroup (FAGRA_SPLIT_LEN=2, read when the
    problem is uploaded): the reduced product, the Jacobian products, a damped Schur solve and
    an LSMR solve still match the oracle, and the reduced product matches the unsplit mapping
    of the same problem."""
    p = small(seed=33, nc=10, npts=700, huber=1.0)
    monkeypatch.setenv("FAGRA_SPLIT_LEN", "0")
    _, _, _, s_plain = linearized_pair(p, precision)
    monkeypatch.setenv("FAGRA_SPLIT_LEN", "2")
    ba_o, s_o, ba_g, s_g = linearized_pair(p, precision)
    rng = np.random.default_rng(5)
    for lam in [1e-3, 1.0]:
        x = rng.normal(size=9 * p.n_cameras)
        st_o, y_o = s_o.reduced_apply(lam, x)
        st_g, y_g = s_g.reduced_apply(lam, x)
        st_p, y_p = s_plain.reduced_apply(lam, x)
        assert st_o == st_g == st_p == 0
        assert rel(y_g, y_o) < 5 * PROD_TOL[precision], lam
        assert rel(y_g, y_p) < PROD_TOL[precision], lam
    n, rows = 9 * p.n_cameras + 3 * p.n_points, 2 * p.n_observations
    x, y = rng.normal(size=n), rng.normal(size=rows)
    assert rel(s_g.jacobian_apply(x), s_o.jacobian_apply(x)) < PROD_TOL[precision]
    assert rel(s_g.jacobian_transpose_apply(y), s_o.jacobian_transpose_apply(y)) < PROD_TOL[precision]
    tol = 1e-10 if precision == fa.FP64 else 1e-6
    s_o.relative_tolerance = tol
    s_g.relative_tolerance = tol
    bound = 1e-8 if precision == fa.FP64 else 3e-3
    for lam in [1e-3, 1.0]:
        st_o, d_o, pr_o = s_o.solve_damped(lam)
        st_g, d_g, pr_g = s_g.solve_damped(lam)
        assert st_o == st_g == 0, (lam, s_g.diagnostics().termination)
        assert rel(d_g, d_o) < bound, lam
        assert abs(pr_g - pr_o) <= bound * abs(pr_o), lam
    l_g = fa.Lsmr(precision=precision, diagonal_preconditioning=True, block_preconditioning=True,
                  relative_tolerance=1e-12 if precision == fa.FP64 else 1e-8, max_iterations=2000)
    assert l_g.prepare(ba_g) == fa.OK and l_g.linearize() == fa.OK and l_g.prepare_damping(1e-6) == fa.OK
    st, d_l, pr_l = l_g.solve_damped(1.0)
    st_o, d_o, pr_o = s_o.solve_damped(1.0)
    assert st == fa.OK and st_o == 0
    assert rel(d_l, d_o) < (1e-6 if precision == fa.FP64 else 3e-3)
