"""QRSchur on the device, after test/qr_schur.jl."""
import numpy as np
import pytest

import kat_models as K
import madsdp_amd as M
from madsdp_amd.model import Model
from oracle import madsdp_oracle as O

pytestmark = pytest.mark.gpu

XS = {True: (np.diag([2.0, 3.0, 4.0]), np.diag([4.0, 2.0, 3.0])),
      False: (np.array([[2.0, 0.3, 0.1], [0.3, 3.0, 0.4], [0.1, 0.4, 4.0]]),
              np.array([[4.0, -0.2, 0.1], [-0.2, 2.0, 0.3], [0.1, 0.3, 3.0]]))}


@pytest.mark.parametrize("scalars", [0, 2])
@pytest.mark.parametrize("diagonal", [True, False])
def test_qr_newton_directions(gpu, scalars, diagonal):
    """Newton equations hold and the QR and Cholesky directions agree to 1e-10
    (qr_schur.jl:16-66); both also agree with the oracle's."""
    model = K.qr_example(scalars)
    X, S = XS[diagonal]
    Aj = [O._BlockOps(model.A[0], 3).dense_matrix(j) for j in range(3)]
    sols = []
    for schur in (M.DenseCholesky(), M.QRSchur()):
        s = M.MadSDPSolver(model, schur=schur, initial_point=M.ScaledIdentity())
        s.initialize()
        s.set_X(0, X)
        s.set_S(0, S)
        s.update_residuals()
        assert s.prepare()
        assert s.factorize_schur()
        sols.append(s)
    ref = O.Solver(model, schur=O.QRSchur(), initial_point=O.ScaledIdentity())
    ref.initialize()
    ref.blocks[0].X, ref.blocks[0].S = X.copy(), S.copy()
    ref.update_residuals()
    assert ref.prepare() and ref.factorize_schur()
    for corrected in (False, True):
        mu = 0.3 if corrected else 0.0
        if corrected:
            ref.correction()
        ref.solve_directions(mu, corrected)
        for s in sols:
            if corrected:
                s.correction()
            s.solve_directions(mu, corrected)
            dX, dS = s.block_array(5, 0), s.block_array(6, 0)
            Rp, Rd = s.Rp, s.block_array(10, 0)
            lhs = np.array([np.vdot(A, dX) for A in Aj]) + model.C_lin @ s.delx
            np.testing.assert_allclose(lhs, Rp, atol=1e-10)
            np.testing.assert_allclose(dS + sum(A * s.dely[j] for j, A in enumerate(Aj)), Rd, atol=1e-10)
            W = s.block_array(13, 0)
            center = mu * np.linalg.inv(S) - X + (s.block_array(16, 0) if corrected else 0)
            np.testing.assert_allclose(dX + W @ dS @ W, center, atol=1e-10)
            np.testing.assert_allclose(s.dely, ref.dely, atol=1e-10)
            np.testing.assert_allclose(dX, ref.blocks[0].delX, atol=1e-10)
        a, b = sols
        np.testing.assert_allclose(a.dely, b.dely, atol=1e-10)
        np.testing.assert_allclose(a.delx, b.delx, atol=1e-10)
        np.testing.assert_allclose(a.delz, b.delz, atol=1e-10)


def pure_lp():
    return Model.from_matrices([], [], [1.0], d_lin=[1.0, 2.0], C_lin=np.array([[1.0, 1.0]]))


@pytest.mark.parametrize("name", ["psd", "mixed", "lp"])
def test_qr_solves(gpu, name):
    """qr_schur.jl:68-91: QRSchur + CholeskyBacktrack solves PSD, mixed and LP problems."""
    model = {"psd": K.qr_example(), "mixed": K.qr_example(2), "lp": pure_lp()}[name]
    s = M.MadSDPSolver(model, schur=M.QRSchur(), initial_point=M.ScaledIdentity(),
                       step=M.CholeskyBacktrack())
    s.solve()
    ref = O.Solver(model, schur=O.QRSchur(), initial_point=O.ScaledIdentity(),
                   step=O.CholeskyBacktrack()).solve()
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert max(s.inf_pr, s.inf_du, s.inf_compl) <= 1e-7
    assert s.iter == ref.iter
    tr = s.trace()
    for k, r in enumerate(ref.trace):
        assert abs(tr[k, 0] - r.pobj) <= 1e-9 * max(1, abs(r.pobj))


def test_qr_unsupported():
    with pytest.raises(ValueError):
        M.MadSDPSolver(K.qr_example(), schur=M.QRSchur(), direction=M.HKM())
    with pytest.raises((ValueError, RuntimeError)):
        M.MadSDPSolver(K.qr_example(), schur=M.QRSchur(), algorithm=M.DualCentering())


def test_qr_zero_column(gpu):
    """An unconstrained multiplier gives a zero column -> ERROR_IN_STEP_COMPUTATION."""
    one = np.ones((1, 1))
    model = Model.from_matrices([one], [[np.zeros((1, 1))]], [0.0])
    s = M.MadSDPSolver(model, schur=M.QRSchur()).solve()
    assert s.status == M.Status.ERROR_IN_STEP_COMPUTATION


@pytest.mark.parametrize("d,mats", [(1, [np.ones((1, 1)), np.ones((1, 1))]),
                                    (2, [np.diag([1.0, 0.0]), np.diag([1.0, 0.0])])])
def test_qr_rank_deficient(gpu, d, mats):
    model = Model.from_matrices([np.eye(d)], [mats], np.ones(2))
    s = M.MadSDPSolver(model, schur=M.QRSchur()).solve()
    assert s.status == M.Status.ERROR_IN_STEP_COMPUTATION


def test_qr_maxcut_parity(gpu):
    """Max-cut G(30, 0.2): per-iteration parity of the QR path with the oracle's."""
    from madsdp_amd import instances
    model = instances.maxcut(30, 0.2, 3)
    s = M.MadSDPSolver(model, schur=M.QRSchur()).solve()
    ref = O.Solver(model, schur=O.QRSchur()).solve()
    assert s.status == M.Status.SOLVE_SUCCEEDED and s.iter == ref.iter
    tr = s.trace()
    for k, r in enumerate(ref.trace):
        assert abs(tr[k, 0] - r.pobj) <= 1e-9 * max(1, abs(r.pobj))


@pytest.mark.parametrize("n,m,ns,seed", [(50, 1100, 10, 7), (90, 1150, 0, 8)],
                         ids=["dense_50_1100_lp10", "dense_90_1150"])
def test_qr_blocked_parity(gpu, n, m, ns, seed):
    """The blocked factorisation (18 panels, the last one partial; B of
    1285 x 1100 and 4095 x 1150) and the per-panel Q'y, Qy and R solves:
    per-iteration parity with the oracle's LAPACK QR to 1e-9."""
    from madsdp_amd import instances
    model = instances.dense_sdp(n, m, ns, seed)
    print(f"QR blocked n={n} m={m} ns={ns}", flush=True)
    s = M.MadSDPSolver(model, schur=M.QRSchur()).solve()
    ref = O.Solver(model, schur=O.QRSchur()).solve()
    assert s.status == M.Status.SOLVE_SUCCEEDED and s.iter == ref.iter, (s.status, s.iter, ref.iter)
    tr = s.trace()
    for k, r in enumerate(ref.trace):
        sc = max(1.0, abs(r.pobj), abs(r.dobj))
        assert abs(tr[k, 0] - r.pobj) <= 1e-9 * sc and abs(tr[k, 1] - r.dobj) <= 1e-9 * sc, k
        assert abs(tr[k, 2] - r.inf_pr) <= 1e-9 * max(1, r.inf_pr), k
        assert abs(tr[k, 3] - r.inf_du) <= 1e-9 * max(1, r.inf_du), k
    np.testing.assert_allclose(s.y, ref.y, rtol=1e-7, atol=1e-9)
