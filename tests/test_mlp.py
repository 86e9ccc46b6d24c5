"""The preconditioner network: numpy reference pinned by finite differences and closed forms
(CPU), and the GPU's BF16-MFMA forward, gradients and training against it (GPU)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def random_targets(rng, n):
    """Packed lower triangles of SPD matrices with eigenvalues in (0, 1] (like L^-1 S L^-T)."""
    out = np.zeros((n, 45))
    for b in range(n):
        Q, _ = np.linalg.qr(rng.normal(size=(9, 9)))
        A = Q @ np.diag(rng.uniform(0.2, 1.0, 9)) @ Q.T
        out[b] = [A[i, j] for i, j in M.TRIL]
    return out


def test_loss_closed_form_at_identity_correction():
    rng = np.random.default_rng(0)
    p = M.init_params(0).astype(np.float64)
    d = M.unpack(p)
    d["W3"][:] = 0
    d["b3"][:] = 0
    p = M.pack(d)
    x = rng.normal(size=(4, M.N_IN))
    t = random_targets(rng, 4)
    A = M.unpack_sym(t)
    expected = np.mean(np.sum((A - np.eye(9)) ** 2, axis=(1, 2)))
    assert abs(M.loss(p, x, t, exact=True) - expected) < 1e-12


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(1).astype(np.float64)
    x = rng.normal(size=(6, M.N_IN))
    t = random_targets(rng, 6)
    L, g = M.loss_grad(p, x, t, exact=True)
    assert abs(L - M.loss(p, x, t, exact=True)) < 1e-12
    h = 1e-6
    # random directions touching every parameter block, and single coordinates of each block
    dirs = [rng.normal(size=p.size) for _ in range(3)]
    off = 0
    for name, n in M.SIZES:
        e = np.zeros(p.size)
        e[off + rng.integers(0, n)] = 1.0
        if name == "W1":  # a used input column
            e[:] = 0
            e[off + 5 * M.IN_PAD + 7] = 1.0
        if name in ("W3", "b3"):
            e[:] = 0
            e[off + (3 * M.H + 11 if name == "W3" else 7)] = 1.0
        dirs.append(e)
        off += n
    for v in dirs:
        num = (M.loss(p + h * v, x, t, exact=True) - M.loss(p - h * v, x, t, exact=True)) / (2 * h)
        assert abs(num - g @ v) < 1e-6 * max(1.0, abs(num)), (num, g @ v)


# ------------------------------------------------------------------------------------------
# GPU
# ------------------------------------------------------------------------------------------

@pytest.mark.gpu
def test_gpu_forward_matches_reference():
    from fagra_amd import learned
    rng = np.random.default_rng(2)
    p = M.init_params(2)
    net = learned.Network(params=p)
    x = rng.normal(size=(130, M.N_IN)).astype(np.float32)  # not a multiple of 64 rows
    y = net.forward(x)
    ref = M.forward(net.params().astype(np.float64), x.astype(np.float64))["y"]
    assert np.linalg.norm(y - ref) < 1e-2 * np.linalg.norm(ref)
    assert np.abs(y - ref).max() < 2e-2 * np.abs(ref).max()


@pytest.mark.gpu
def test_gpu_gradients_match_reference():
    from fagra_amd import learned
    rng = np.random.default_rng(3)
    p = M.init_params(3)
    net = learned.Network(params=p)
    x = rng.normal(size=(200, M.N_IN)).astype(np.float32)
    t = random_targets(rng, 200).astype(np.float32)
    L, g = net.loss_grad(x, t)
    Lr, gr = M.loss_grad(net.params().astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(L - Lr) < 1e-3 * Lr
    off = 0
    for name, n in M.SIZES:
        a, b = g[off:off + n], gr[off:off + n]
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-7, name
        off += n


@pytest.mark.gpu
def test_gpu_training_reduces_loss():
    from fagra_amd import learned
    rng = np.random.default_rng(4)
    net = learned.Network(seed=4)
    x = rng.normal(size=(2048, M.N_IN)).astype(np.float32)
    # targets depending smoothly on the features: a learnable map
    t = random_targets(np.random.default_rng(5), 2048).astype(np.float32)
    losses = net.train(x, t, epochs=8, batch=256, lr=1e-3, seed=0)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < losses[0]


@pytest.mark.gpu
@pytest.mark.parametrize("force_fallback", [False, True])
def test_learned_preconditioner_solves_the_same_system(force_fallback):
    """With any network attached, the learned PCG must reach the same damped step as
    block-Jacobi (both run to the same tolerance); a forced fallback must restart cleanly."""
    import fagra_amd as fa
    from fagra_amd import learned, synthetic
    p = synthetic.generate("ring16", seed=21, n_cameras=12, n_points=600, huber=1.0)
    net = learned.Network(params=M.init_params(7))   # nonzero output layer: G != I
    steps = {}
    for pc in (fa.PRECOND_BLOCK_JACOBI, fa.PRECOND_LEARNED):
        ba = fa.BalProblem(p, precision=fa.FP32)
        s = fa.Schur(precision=fa.FP32, relative_tolerance=1e-6, preconditioner=pc)
        fa.lib().fg_schur_set_mlp(s.h, net.h)
        if force_fallback:
            s.cfg.fallback_factor = 1e-30
            s.cfg.fallback_iterations = 2
            s._push()
        assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
        st, d, pr = s.solve_damped(1e-2)
        dg = s.diagnostics()
        assert st == fa.OK and dg.termination == fa.CG_CONVERGED
        if pc == fa.PRECOND_LEARNED:
            assert dg.used_fallback == int(force_fallback)
        steps[pc] = (d, pr)
    d0, p0 = steps[fa.PRECOND_BLOCK_JACOBI]
    d1, p1 = steps[fa.PRECOND_LEARNED]
    assert np.linalg.norm(d1 - d0) < 3e-3 * np.linalg.norm(d0)
    assert abs(p1 - p0) < 3e-3 * abs(p0)


@pytest.mark.gpu
def test_learned_preconditioner_fused_and_timed_launches_agree():
    """Without timings the learned inverse is formed in the point factorization's launch; with
    collect_timings it runs as its own kernel, timed as SchurTimings::preconditioner. Both apply
    the same arithmetic to the same operands, so the damped step is the same bit for bit."""
    import fagra_amd as fa
    from fagra_amd import learned, synthetic
    p = synthetic.generate("ring16", seed=23, n_cameras=40, n_points=1500, huber=1.0)
    net = learned.Network(params=M.init_params(11))
    out = {}
    for timed in (False, True):
        ba = fa.BalProblem(p, precision=fa.FP32)
        s = fa.Schur(precision=fa.FP32, relative_tolerance=1e-6, preconditioner=fa.PRECOND_LEARNED,
                     collect_timings=timed)
        fa.lib().fg_schur_set_mlp(s.h, net.h)
        assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
        st, d, pr = s.solve_damped(1e-2)
        assert st == fa.OK and s.diagnostics().termination == fa.CG_CONVERGED
        out[timed] = (d, pr, s.diagnostics().iterations, s.timings().preconditioner)
        s.close()
        ba.close()
    assert np.array_equal(out[False][0], out[True][0]) and out[False][1] == out[True][1]
    assert out[False][2] == out[True][2]
    assert out[True][3] > 0.0 and out[False][3] == 0.0


@pytest.mark.gpu
def test_learned_without_network_is_rejected():
    import fagra_amd as fa
    from fagra_amd import synthetic
    p = synthetic.generate("ring16", seed=22, n_cameras=6, n_points=100)
    ba = fa.BalProblem(p, precision=fa.FP32)
    s = fa.Schur(precision=fa.FP32)  # no network attached (the constructor would attach one)
    s.preconditioner = fa.PRECOND_LEARNED
    assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
    assert s.solve_damped(1e-2)[0] == fa.INVALID_OPTIONS
