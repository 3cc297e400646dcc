"""The oracle against the reference's own known answers, closed forms and
invariants (CPU only)."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

import kat_models as K
from madsdp_amd import instances
from oracle import madsdp_oracle as O

COMBOS = [(d, s, c) for d in (O.NT, O.HKM) for s in (O.EigenStep, O.CholeskyBacktrack)
          for c in (O.Mehrotra, O.AdaptiveMu)]


def test_trivial():
    s = O.Solver(K.trivial(), tol=1e-8, max_iter=60).solve()
    assert s.status == O.Status.SOLVE_SUCCEEDED
    assert abs(s.primal_objective() - 2.0) <= 1e-5


def test_maxcut_lro():
    s = O.Solver(K.maxcut(K.W_LRO), tol=1e-7, max_iter=80).solve()
    assert s.status == O.Status.SOLVE_SUCCEEDED
    assert abs(s.primal_objective() + 18.0) <= 1e-4


def test_trace_one_is_lambda_min():
    s = O.Solver(K.trace_one(), tol=1e-8, max_iter=60).solve()
    assert abs(s.primal_objective() - 2.0) <= 1e-5


def test_mixed():
    s = O.Solver(K.mixed(), tol=1e-7, max_iter=60).solve()
    assert s.status == O.Status.SOLVE_SUCCEEDED
    assert abs(s.primal_objective() - 1.0) <= 1e-5
    assert abs(s.blocks[0].X[0, 0] - 1.0) <= 1e-3 and np.all(s.x < 1e-3)


def test_cycle():
    s = O.Solver(K.maxcut(K.W_CYCLE)).solve()
    assert abs(s.primal_objective() + 4.0) <= 1e-4


@pytest.mark.parametrize("d,st,c", COMBOS, ids=lambda x: x.__name__)
def test_components(d, st, c):
    for model, expected in ((K.maxcut(K.W_LRO), -18.0), (K.mixed(), 1.0)):
        s = O.Solver(model, direction=d(), step=st(), centering=c(), max_iter=100).solve()
        assert s.status == O.Status.SOLVE_SUCCEEDED
        assert abs(s.primal_objective() - expected) <= 1e-4


def test_mc_psd_feasible_start():
    X0, y0 = K.mc_psd_start(K.W_LRO)
    s = O.Solver(K.maxcut(K.W_LRO), direction=O.HKM(), step=O.CholeskyBacktrack(),
                 centering=O.AdaptiveMu(), initial_point=O.InitialPoint(X0, y0)).solve()
    assert s.status == O.Status.SOLVE_SUCCEEDED
    assert abs(s.primal_objective() + 18.0) <= 1e-4
    assert s.inf_du < 1e-12


def test_initial_point_scalars():
    s = O.Solver(K.mixed(), direction=O.HKM(),
                 initial_point=O.InitialPoint([np.ones((1, 1))], [-1.0, -1.0], x=[0.5, 0.5])).solve()
    assert abs(s.primal_objective() - 1.0) <= 1e-5


def test_cholesky_factor_known_answers():
    """test/blocks.jl:9-35."""
    M = np.array([[4.0, 1.0], [3.0, 2.0]])
    L = O.cholesky_factor(M)
    np.testing.assert_allclose(L @ L.T, (M + M.T) / 2)
    assert L[0, 1] == 0
    assert O.cholesky_factor(np.array([[1.0, 2.0], [2.0, 1.0]])) is None
    assert O.cholesky_factor(np.zeros((2, 2))) is None
    eps = np.finfo(float).eps
    F = np.array([[1, 0, 0], [np.finfo(float).tiny, 1, 0], [eps, 1, 1]])
    L = O.cholesky_factor(F @ F.T)
    assert L[1, 0] == 0 and L[2, 0] == eps


def test_identity_scale():
    """test/identity_scale.jl:15-42."""
    for scale in (1, 0.2, 10.0):
        s = O.Solver(K.qr_example(2), initial_point=O.ScaledIdentity(scale))
        s.initialize()
        assert np.all(s.blocks[0].X == scale * np.eye(3)) and np.all(s.x == scale)
    for bad in (0, -1.0, math.inf, math.nan):
        with pytest.raises(ValueError):
            O.ScaledIdentity(bad)
    model = K.qr_example(2)
    model.b = model.b * 1000
    s = O.Solver(model)
    s.initialize()
    np.testing.assert_allclose(s.blocks[0].X, math.sqrt(15000.0) * np.eye(3))


def test_termination_semantics():
    """test/termination.jl:15-40: mu = 0 alone does not certify."""
    one = np.ones((1, 1))
    model = lambda: __import__("madsdp_amd").Model.from_matrices([one], [[one]], [1.0])
    for x, sv, y in ((2.0, 0.0, 1.0), (1.0, 0.0, 0.0), (1.0, 1.0, 0.0)):
        s = O.Solver(model())
        s.initialize()
        s.blocks[0].X[0, 0], s.blocks[0].S[0, 0], s.y[0] = x, sv, y
        s.mu = 0.0
        s.update_residuals()
        s.update_termination()
        assert s.status == O.Status.SEARCH_DIRECTION_BECOMES_TOO_SMALL
    s = O.Solver(model())
    s.initialize()
    s.blocks[0].X[0, 0], s.blocks[0].S[0, 0], s.y[0] = 1.0, 0.0, 1.0
    s.mu = 0.0
    s.update_residuals()
    s.update_termination()
    assert s.status == O.Status.SOLVE_SUCCEEDED
    assert max(s.inf_pr, s.inf_du, s.inf_compl) == 0


@pytest.mark.parametrize("scalars", [0, 2])
@pytest.mark.parametrize("diagonal", [True, False])
def test_qr_newton_directions(scalars, diagonal):
    """test/qr_schur.jl:16-66: Newton equations hold and QR == Cholesky directions."""
    model = K.qr_example(scalars)
    solvers = [O.Solver(model, schur=sc, initial_point=O.ScaledIdentity())
               for sc in (O.DenseCholesky(), O.QRSchur())]
    ops = O._BlockOps(model.A[0], 3)
    Aj = [ops.dense_matrix(j) for j in range(3)]
    for s in solvers:
        s.initialize()
        s.blocks[0].X = np.diag([2.0, 3.0, 4.0]) if diagonal else np.array(
            [[2.0, 0.3, 0.1], [0.3, 3.0, 0.4], [0.1, 0.4, 4.0]])
        s.blocks[0].S = np.diag([4.0, 2.0, 3.0]) if diagonal else np.array(
            [[4.0, -0.2, 0.1], [-0.2, 2.0, 0.3], [0.1, 0.3, 3.0]])
        s.update_residuals()
        assert s.prepare()
        assert s.factorize_schur()
    for corrected in (False, True):
        mu = 0.3 if corrected else 0.0
        for s in solvers:
            if corrected:
                s.correction()
            s.solve_directions(mu, corrected)
            b = s.blocks[0]
            lhs = np.array([np.vdot(A, b.delX) for A in Aj]) + model.C_lin @ s.delx
            np.testing.assert_allclose(lhs, s.Rp, atol=1e-10)
            dual = sum(A * s.dely[j] for j, A in enumerate(Aj))
            np.testing.assert_allclose(b.delS + dual, b.Rd, atol=1e-10)
            center = mu * np.linalg.inv(b.S) - b.X + (b.corr if corrected else 0)
            np.testing.assert_allclose(b.delX + b.W @ b.delS @ b.W, center, atol=1e-10)
        a, c = solvers
        np.testing.assert_allclose(a.dely, c.dely, atol=1e-10)
        np.testing.assert_allclose(a.delx, c.delx, atol=1e-10)
        np.testing.assert_allclose(a.blocks[0].delX, c.blocks[0].delX, atol=1e-10)


def test_qr_solves():
    """test/qr_schur.jl:68-91."""
    for model in (K.qr_example(), K.qr_example(2)):
        s = O.Solver(model, schur=O.QRSchur(), step=O.CholeskyBacktrack()).solve()
        assert s.status == O.Status.SOLVE_SUCCEEDED
        assert max(s.inf_pr, s.inf_du, s.inf_compl) <= s.opt.tol


def test_qr_rejections():
    """test/qr_schur.jl:93-139."""
    with pytest.raises(ValueError):
        O.Solver(K.qr_example(), schur=O.QRSchur(), direction=O.HKM())
    with pytest.raises(ValueError):
        O.Solver(K.qr_example(), schur=O.QRSchur(), algorithm=O.DualCentering())
    from madsdp_amd import Model
    one = np.ones((1, 1))
    m = Model.from_matrices([one], [[np.zeros((1, 1))]], [0.0])
    s = O.Solver(m, schur=O.QRSchur()).solve()
    assert s.status == O.Status.ERROR_IN_STEP_COMPUTATION
    for d, mats in ((1, [one, one]), (2, [np.diag([1.0, 0.0]), np.diag([1.0, 0.0])])):
        m = Model.from_matrices([np.eye(d)], [mats], np.ones(2))
        s = O.Solver(m, schur=O.QRSchur()).solve()
        assert s.status == O.Status.ERROR_IN_STEP_COMPUTATION


def test_dual_centering_kats():
    """test/dual_runtests.jl:191-259."""
    s = O.Solver(K.trivial(), algorithm=O.DualCentering(), tol=1e-7, max_iter=200).solve()
    assert s.status == O.Status.SOLVE_SUCCEEDED and abs(s.dual_objective() - 2.0) <= 1e-4
    s = O.Solver(K.trace_one(), algorithm=O.DualCentering(), tol=1e-7, max_iter=200).solve()
    assert abs(s.dual_objective() - 2.0) <= 1e-4
    s = O.Solver(K.maxcut(K.W_CYCLE), algorithm=O.DualCentering(), tol=1e-6, max_iter=300).solve()
    assert abs(s.dual_objective() + 4.0) <= 1e-3 and s.inf_du < 1e-10
    s = O.Solver(K.maxcut(K.W_LRO), algorithm=O.DualCentering(), tol=1e-6, max_iter=300).solve()
    assert abs(s.dual_objective() + 18.0) <= 1e-3
    assert s.n_newton > 0 and s.n_predictor > 0


def test_dual_centering_errors():
    W2 = np.ones((2, 2)) - np.eye(2)
    with pytest.raises(RuntimeError):
        O.Solver(K.maxcut(W2), algorithm=O.DualCentering(),
                 initial_point=O.InitialPoint([np.eye(2)], [1.0, 1.0])).solve()
    with pytest.raises(RuntimeError):
        O.Solver(K.maxcut(W2), algorithm=O.DualCentering(), initial_point=O.DualShift(1.0, 0.1)).solve()
    with pytest.raises(RuntimeError):
        O.Solver(K.mixed(), algorithm=O.DualCentering())


def test_dualshift_k4_closed_form():
    """K4 max-cut SDP value n^2/4 = 4 with X = (4I - J)/3."""
    W = np.ones((4, 4)) - np.eye(4)
    s = O.Solver(K.maxcut(W), direction=O.HKM(), initial_point=O.DualShift()).solve()
    assert abs(s.primal_objective() + 4.0) <= 1e-4
    np.testing.assert_allclose(s.blocks[0].X, (4 * np.eye(4) - 1) / 3, atol=1e-3)


@pytest.mark.parametrize("n", [5, 8])
def test_complete_graph_closed_form(n):
    """K_n: SDP max-cut value n^2/4."""
    W = np.ones((n, n)) - np.eye(n)
    s = O.Solver(K.maxcut(W)).solve()
    assert abs(-s.primal_objective() - n * n / 4) <= 1e-5 * n * n


def test_bipartite_closed_form():
    """Bipartite graphs: the SDP cuts every edge, value = total weight."""
    rng = np.random.default_rng(0)
    W = np.zeros((10, 10))
    for i in range(5):
        for j in range(5, 10):
            if rng.random() < 0.6:
                W[i, j] = W[j, i] = rng.integers(1, 4)
    s = O.Solver(K.maxcut(W)).solve()
    assert abs(-s.primal_objective() - W.sum() / 2) <= 1e-5 * W.sum()


def test_schur_brute_force():
    """H_jk = tr(A_j W A_k W) by definition, for coo, one-hot and dense blocks."""
    rng = np.random.default_rng(1)
    for model in (instances.dense_sdp(6, 9, 0, 1), K.maxcut(K.W_LRO), K.qr_example()):
        s = O.Solver(model)
        s.initialize()
        n = model.dims[0]
        G = rng.standard_normal((n, n))
        Wm = G @ G.T + np.eye(n)
        s.blocks[0].W = Wm
        H = s.assemble_schur(0.0)
        ops = s.ops[0]
        Aj = [ops.dense_matrix(j) for j in range(model.m)]
        ref = np.array([[np.trace(Aj[j] @ Wm @ Aj[k] @ Wm) for k in range(model.m)] for j in range(model.m)])
        np.testing.assert_allclose(H, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_nt_scaling_invariants():
    """W S W = X and W = S^-1/2 (S^1/2 X S^1/2)^1/2 S^-1/2 (geometric mean)."""
    rng = np.random.default_rng(2)
    s = O.Solver(K.qr_example())
    s.initialize()
    n = 3
    A = rng.standard_normal((n, n)); X = A @ A.T + np.eye(n)
    B = rng.standard_normal((n, n)); S = B @ B.T + np.eye(n)
    s.blocks[0].X, s.blocks[0].S = X, S
    s.update_residuals()
    assert s.prepare()
    W = s.blocks[0].W
    np.testing.assert_allclose(W @ S @ W, X, rtol=1e-12)
    Sh = sla.sqrtm(S).real
    Shi = np.linalg.inv(Sh)
    ref = Shi @ sla.sqrtm(Sh @ X @ Sh).real @ Shi
    np.testing.assert_allclose(W, ref, rtol=1e-10)


def test_lambda_min_brute_force():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 6)); X = A @ A.T + np.eye(6)
    D = rng.standard_normal((6, 6)); D = D + D.T
    L = np.linalg.cholesky(X)
    lam = O.Solver.lambda_min_congruence(L, D)
    # smallest generalized eigenvalue of (D, X)
    assert abs(lam - sla.eigh(D, X, eigvals_only=True)[0]) <= 1e-12 * abs(lam)


def test_regularisation_ladder():
    """A rank-deficient Schur complement succeeds only after the ladder grows
    (src/solver.jl:299-314): reg = 1e-12, 1e-11, ..."""
    from madsdp_amd import Model
    one = np.ones((1, 1))
    m = Model.from_matrices([one], [[one, one]], [1.0, 1.0])
    s = O.Solver(m)
    s.initialize()
    assert s.prepare()
    assert s.factorize_schur()
    assert abs(s.H[0, 0] - s.H[1, 1]) < 1e-30 and s.H[0, 0] - s.H[1, 0] >= 1e-12 * 0.99


def test_workload_instances_feasible():
    """X = I, x = 1 is primal feasible for the dense/block generators."""
    for model in (instances.dense_sdp(8, 20, 5, 2), instances.block_sdp([3, 5, 4], 12, 4, 3)):
        s = O.Solver(model)
        s.initialize()
        for b in s.blocks:
            b.X = np.eye(b.n)
        s.x = np.ones(model.n_scalar)
        s.update_residuals()
        assert np.abs(s.Rp).max() <= 1e-10 * max(1, np.abs(model.b).max())


@pytest.mark.parametrize("model,expected", [(lambda: K.maxcut(K.W_LRO), -18.0), (K.mixed, 1.0),
                                            (K.trivial, 2.0)], ids=["maxcut", "mixed", "trivial"])
def test_madnlp_schur_kats(model, expected):
    """runtests.jl:186-202 runs MadNLPSchur on the same problems."""
    s = O.Solver(model(), schur=O.MadNLPSchur()).solve()
    assert s.status == O.Status.SOLVE_SUCCEEDED
    assert abs(s.primal_objective() - expected) <= 1e-4


def test_madnlp_schur_mc_psd_start_and_dual_centering():
    """runtests.jl:186-202 (dual feasibility kept from the mc_psd start) and
    dual_runtests.jl:88-94 (DualCentering with MadNLPSchur)."""
    X0, y0 = K.mc_psd_start(K.W_LRO)
    s = O.Solver(K.maxcut(K.W_LRO), direction=O.HKM(), step=O.CholeskyBacktrack(),
                 centering=O.AdaptiveMu(), initial_point=O.InitialPoint(X0, y0),
                 schur=O.MadNLPSchur()).solve()
    assert s.status == O.Status.SOLVE_SUCCEEDED and s.inf_du < 1e-12
    s = O.Solver(K.maxcut(K.W_LRO), algorithm=O.DualCentering(), schur=O.MadNLPSchur(),
                 tol=1e-6, max_iter=300).solve()
    assert abs(s.dual_objective() + 18.0) <= 1e-3


def test_madnlp_schur_solve_is_exact():
    """The Bunch-Kaufman solve against a dense solve of the regularised H."""
    s = O.Solver(instances.dense_sdp(8, 20, 3, 6), schur=O.MadNLPSchur())
    s.initialize()
    assert s.prepare() and s.factorize_schur()
    rhs = np.arange(1.0, 21.0)
    np.testing.assert_allclose(s.schur_solve(rhs), np.linalg.solve(s.H, rhs), rtol=1e-10)
