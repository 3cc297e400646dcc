"""Failure semantics on the device, each against the FP64 oracle: a zero
barrier parameter does not certify optimality and an exact solution does
(test/termination.jl:15-40), the iteration limit (MAXIMUM_ITERATIONS_EXCEEDED,
src/solver.jl:349-373), and the Schur regularisation ladder
(src/solver.jl:299-314) on a rank-deficient Schur complement."""
import numpy as np
import pytest

import kat_models as K
import madsdp_amd as M
from madsdp_amd.model import Model
from oracle import madsdp_oracle as O

pytestmark = pytest.mark.gpu


def scalar_model():
    one = np.ones((1, 1))
    return Model.from_matrices([one], [[one]], [1.0])


def _device_state(x, sv, y):
    s = M.MadSDPSolver(scalar_model())
    s.initialize()
    s.set_X(0, np.array([[x]]))
    s.set_S(0, np.array([[sv]]))
    s.set_y([y])
    s.set_mu(0.0)
    s.update_residuals()
    s.update_termination()
    return s


def _oracle_state(x, sv, y):
    r = O.Solver(scalar_model())
    r.initialize()
    r.blocks[0].X[0, 0], r.blocks[0].S[0, 0], r.y[0] = x, sv, y
    r.mu = 0.0
    r.update_residuals()
    r.update_termination()
    return r


@pytest.mark.parametrize("x,sv,y", [(2.0, 0.0, 1.0), (1.0, 0.0, 0.0), (1.0, 1.0, 0.0)])
def test_zero_mu_does_not_certify(gpu, x, sv, y):
    """termination.jl:16-28: mu = 0 with residuals above tol ->
    SEARCH_DIRECTION_BECOMES_TOO_SMALL."""
    s, r = _device_state(x, sv, y), _oracle_state(x, sv, y)
    assert s.status == M.Status.SEARCH_DIRECTION_BECOMES_TOO_SMALL
    assert int(r.status) == int(s.status)
    assert max(s.inf_pr, s.inf_du, s.inf_compl) > 1e-8
    for k in ("inf_pr", "inf_du", "inf_compl"):
        assert getattr(s, k) == pytest.approx(getattr(r, k), rel=1e-14, abs=1e-300), k


def test_exact_solution_certifies(gpu):
    """termination.jl:29-37: X = 1, S = 0, y = 1 is optimal with zero
    residuals, so mu = 0 certifies it."""
    s = _device_state(1.0, 0.0, 1.0)
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert max(s.inf_pr, s.inf_du, s.inf_compl) == 0.0
    assert int(_oracle_state(1.0, 0.0, 1.0).status) == int(s.status)


@pytest.mark.parametrize("max_iter", [0, 3])
def test_iteration_limit(gpu, max_iter):
    """opt.max_iter reached before convergence -> MAXIMUM_ITERATIONS_EXCEEDED
    after exactly max_iter iterations, with the oracle's trace."""
    model = K.maxcut(K.W_LRO)
    s = M.MadSDPSolver(model, max_iter=max_iter).solve()
    r = O.Solver(model, max_iter=max_iter).solve()
    assert s.status == M.Status.MAXIMUM_ITERATIONS_EXCEEDED
    assert int(r.status) == int(s.status) and s.iter == r.iter == max_iter
    tr = s.trace()
    assert tr.shape[0] == len(r.trace)
    for k, rec in enumerate(r.trace):
        assert abs(tr[k, 0] - rec.pobj) <= 1e-9 * max(1, abs(rec.pobj)), k
        assert abs(tr[k, 1] - rec.dobj) <= 1e-9 * max(1, abs(rec.dobj)), k


def test_regularisation_ladder(gpu):
    """Two identical constraints make H singular: the factorisation succeeds
    at the ladder's first shift, 1e-12, as the oracle's does, and the solve
    then follows the oracle per iteration."""
    one = np.ones((1, 1))
    model = Model.from_matrices([one], [[one, one]], [1.0, 1.0])
    s = M.MadSDPSolver(model)
    s.initialize()
    assert s.prepare() and s.factorize_schur()
    r = O.Solver(model)
    r.initialize()
    assert r.prepare() and r.factorize_schur()
    ref_reg = r.H[0, 0] - r.H[1, 0]
    assert s.schur_reg == 1e-12
    assert ref_reg == pytest.approx(1e-12, rel=1e-3)
    s = M.MadSDPSolver(model).solve()
    r = O.Solver(model).solve()
    assert int(s.status) == int(r.status) and s.iter == r.iter
    tr = s.trace()
    for k, rec in enumerate(r.trace):
        assert abs(tr[k, 0] - rec.pobj) <= 1e-9 * max(1, abs(rec.pobj)), k
