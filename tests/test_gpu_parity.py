"""End-to-end parity of the HIP solver with the FP64 oracle.

The bar (FP64 plain path): the same iteration count and status, and every
per-iteration objective and residual within 1e-9 relative of the oracle's.
The known answers of the reference tests are checked as well.
"""
import numpy as np
import pytest

import kat_models as K
from madsdp_amd import instances
import madsdp_amd as M
from oracle import madsdp_oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def to_oracle(component):
    """Same component for the oracle (same name, same parameters)."""
    if component is None:
        return None
    name = type(component).__name__
    cls = getattr(O, name)
    if isinstance(component, (M.NT, M.HKM, M.DenseCholesky, M.QRSchur)):
        return cls()
    if isinstance(component, M.ScaledIdentity):
        return cls(component.scale)
    if isinstance(component, M.InitialPoint):
        return cls(component.X, component.y, component.x)
    return cls(**component.__dict__)


def solve_both(model, **kw):
    okw = {k: (to_oracle(v) if k in ("direction", "step", "centering", "algorithm",
                                     "initial_point", "schur") else v) for k, v in kw.items()}
    print(f"solve {kw}", flush=True)
    ref = O.Solver(model, **okw).solve()
    s = M.MadSDPSolver(model, **kw).solve()
    return s, ref


def assert_parity(s, ref, rtol=RTOL, check_trace=True):
    assert int(s.status) == int(ref.status), (s.status, ref.status)
    assert s.iter == ref.iter, (s.iter, ref.iter)
    if not check_trace:
        return
    tr = s.trace()
    assert tr.shape[0] == len(ref.trace)
    for k, r in enumerate(ref.trace):
        scale = max(1.0, abs(r.pobj), abs(r.dobj))
        assert abs(tr[k, 0] - r.pobj) <= rtol * scale, ("pobj", k, tr[k, 0], r.pobj)
        assert abs(tr[k, 1] - r.dobj) <= rtol * scale, ("dobj", k, tr[k, 1], r.dobj)
        assert abs(tr[k, 2] - r.inf_pr) <= rtol * max(1.0, r.inf_pr), ("inf_pr", k)
        assert abs(tr[k, 3] - r.inf_du) <= rtol * max(1.0, r.inf_du), ("inf_du", k)
        assert abs(tr[k, 4] - r.gap) <= rtol * max(1.0, r.gap), ("gap", k)


COMBOS = [(d, st, c) for d in (M.NT, M.HKM) for st in (M.EigenStep, M.CholeskyBacktrack)
          for c in (M.Mehrotra, M.AdaptiveMu)]


@pytest.mark.parametrize("d,st,c", COMBOS, ids=lambda x: x.__name__)
def test_components_maxcut_lro(gpu, d, st, c):
    """runtests.jl:157-184: every component combination on the LRO max-cut -> -18."""
    s, ref = solve_both(K.maxcut(K.W_LRO), direction=d(), step=st(), centering=c(), max_iter=100)
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert abs(s.primal_objective() + 18.0) <= 1e-4
    assert_parity(s, ref)


@pytest.mark.parametrize("d,st,c", COMBOS, ids=lambda x: x.__name__)
def test_components_mixed(gpu, d, st, c):
    s, ref = solve_both(K.mixed(), direction=d(), step=st(), centering=c(), max_iter=100)
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert abs(s.primal_objective() - 1.0) <= 1e-4
    assert abs(s.X(0)[0, 0] - 1.0) <= 1e-3 and np.all(s.x < 1e-3)
    assert_parity(s, ref)


def test_trivial(gpu):
    s, ref = solve_both(K.trivial(), tol=1e-8, max_iter=60)
    assert s.status == M.Status.SOLVE_SUCCEEDED and abs(s.primal_objective() - 2.0) <= 1e-5
    assert_parity(s, ref)


def test_trace_one(gpu):
    s, ref = solve_both(K.trace_one(), tol=1e-8, max_iter=60)
    assert s.status == M.Status.SOLVE_SUCCEEDED and abs(s.primal_objective() - 2.0) <= 1e-5
    assert_parity(s, ref)


def test_cycle(gpu):
    s, ref = solve_both(K.maxcut(K.W_CYCLE), tol=1e-7, max_iter=80)
    assert abs(s.primal_objective() + 4.0) <= 1e-4
    assert_parity(s, ref)


@pytest.mark.parametrize("schur", ["dense"])
def test_mc_psd_start(gpu, schur):
    """runtests.jl:186-202: HKM + CholeskyBacktrack + AdaptiveMu from the
    dual feasible start keeps inf_du < 1e-12."""
    X0, y0 = K.mc_psd_start(K.W_LRO)
    s, ref = solve_both(K.maxcut(K.W_LRO), direction=M.HKM(), step=M.CholeskyBacktrack(),
                        centering=M.AdaptiveMu(), initial_point=M.InitialPoint(X0, y0))
    assert abs(s.primal_objective() + 18.0) <= 1e-4
    assert s.inf_du < 1e-12
    assert_parity(s, ref)


def test_initial_point_scalars(gpu):
    s, ref = solve_both(K.mixed(), direction=M.HKM(),
                        initial_point=M.InitialPoint([np.ones((1, 1))], [-1.0, -1.0], x=[0.5, 0.5]))
    assert abs(s.primal_objective() - 1.0) <= 1e-5
    assert_parity(s, ref)


def test_dual_shift_k4(gpu):
    """dual_runtests.jl:292-299: PrimalDual from DualShift on K4 -> -4, X = (4I-J)/3."""
    W = np.ones((4, 4)) - np.eye(4)
    s, ref = solve_both(K.maxcut(W), direction=M.HKM(), initial_point=M.DualShift())
    assert abs(s.primal_objective() + 4.0) <= 1e-4
    np.testing.assert_allclose(s.X(0), (4 * np.eye(4) - 1) / 3, atol=1e-3)
    assert_parity(s, ref)


def test_identity_scale_automatic(gpu):
    """identity_scale.jl:32-42: beta = sqrt(15000) for b * 1000."""
    model = K.qr_example(scalars=2)
    model.b = model.b * 1000
    s = M.MadSDPSolver(model)
    s.initialize()
    beta = np.sqrt(15000.0)
    np.testing.assert_allclose(s.X(0), beta * np.eye(3), rtol=1e-14)
    np.testing.assert_allclose(s.x, [beta, beta], rtol=1e-14)


@pytest.mark.parametrize("scalars", [0, 2])
def test_qr_example_cholesky(gpu, scalars):
    s, ref = solve_both(K.qr_example(scalars))
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert_parity(s, ref)


# ---- DualCentering (dual_runtests.jl) ----
def test_dual_trivial(gpu):
    s, ref = solve_both(K.trivial(), algorithm=M.DualCentering(), tol=1e-7, max_iter=200)
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert abs(s.dual_objective() - 2.0) <= 1e-4 and abs(s.primal_objective() - 2.0) <= 1e-4
    assert_parity(s, ref)


def test_dual_trace_one(gpu):
    s, ref = solve_both(K.trace_one(), algorithm=M.DualCentering(), tol=1e-7, max_iter=200)
    assert abs(s.dual_objective() - 2.0) <= 1e-4
    assert_parity(s, ref)


def test_dual_cycle(gpu):
    s, ref = solve_both(K.maxcut(K.W_CYCLE), algorithm=M.DualCentering(), tol=1e-6, max_iter=300)
    assert abs(s.dual_objective() + 4.0) <= 1e-3 and s.inf_du < 1e-10
    assert_parity(s, ref)


def test_dual_lro(gpu):
    s, ref = solve_both(K.maxcut(K.W_LRO), algorithm=M.DualCentering(), tol=1e-6, max_iter=300)
    assert abs(s.dual_objective() + 18.0) <= 1e-3
    assert_parity(s, ref)
    X0, y0 = K.mc_psd_start(K.W_LRO)
    s, ref = solve_both(K.maxcut(K.W_LRO), algorithm=M.DualCentering(A=None),
                        initial_point=M.InitialPoint(X0, y0), tol=1e-6, max_iter=300)
    assert abs(s.dual_objective() + 18.0) <= 1e-3
    assert_parity(s, ref)


# ---- workload-shaped instances ----
def test_maxcut_100(gpu):
    """Max-cut G(100, 0.1) (the CPU-runnable workload): D&C eigensolver path."""
    s, ref = solve_both(instances.maxcut_100())
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert_parity(s, ref)


def test_maxcut_300_pm1(gpu):
    s, ref = solve_both(instances.maxcut(300, 0.02, 11, "pm1"))
    assert_parity(s, ref)


def test_dense_small(gpu):
    """Dense SDP with scalars (the 250/8000 workload's construction, small)."""
    s, ref = solve_both(instances.dense_sdp(20, 60, 10, 2))
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert_parity(s, ref)


def test_dense_small_hkm(gpu):
    s, ref = solve_both(instances.dense_sdp(12, 30, 5, 3), direction=M.HKM())
    assert_parity(s, ref)


def test_blocks_small(gpu):
    """Block-diagonal SOS-style problem (the 96-block workload's construction, small)."""
    model = instances.block_sdp(instances.block_sizes(6, 4, 70, 5), 80, 8, 3)
    s, ref = solve_both(model)
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert_parity(s, ref)


def test_maxcut_1100_pm1(gpu):
    """m = 1100: the Schur solves apply the inverses of two full 512-row
    diagonal blocks and finish the last 76 rows by substitution; n = 1100
    takes the blocked tridiagonalisation and Lanczos step lengths."""
    s, ref = solve_both(instances.maxcut(1100, 0.01, 12, "pm1"))
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert_parity(s, ref)


def test_dense_m1100(gpu):
    """Dense constraints, m = 1100 > 2 x 512, with scalars."""
    s, ref = solve_both(instances.dense_sdp(24, 1100, 12, 13))
    assert s.status == M.Status.SOLVE_SUCCEEDED
    assert_parity(s, ref)
