"""Device Bunch-Kaufman LDL' (MadNLPSchur, src/kkt.jl:7-32) through the C ABI.

Kernel level: the factorisation of symmetric indefinite, positive definite
and 2x2-pivot-forcing matrices against LAPACK's dsytrf (same pivots, same
info) and the solve against numpy; a matrix with a zero column reports its
pivot in info.  Solver level: MadNLPSchur on the reference tests' problems
(test/runtests.jl:186-202, test/dual_runtests.jl:88-94) against the FP64
oracle's MadNLPSchur, per-iteration objective and residuals to 1e-9."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla

import kat_models as K
import madsdp_amd as M
from madsdp_amd import instances
from oracle import madsdp_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(gpu):
    import torch
    from madsdp_amd import lib as L
    lib = L.load()
    return torch, lib, C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _factor(env, A):
    torch, lib, st = env
    n = A.shape[0]
    dA = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    ctl = torch.zeros(int(lib.msdp_k_sytrf_ctl_bytes()) // 8 + 1, dtype=torch.float64, device="cuda")
    work = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    P = lambda t: C.c_void_p(t.data_ptr())
    assert lib.msdp_k_sytrf(st, n, P(dA), n, P(ipiv), P(ctl), P(work), P(info)) == 0
    torch.cuda.synchronize()
    return dA, ipiv, int(info.item())


def _cases():
    rng = np.random.default_rng(11)
    out = []
    A = rng.standard_normal((300, 300))
    out.append(("indefinite_300", (A + A.T) / 2))
    B = rng.standard_normal((1100, 1100))
    out.append(("spd_1100", B @ B.T / 1100 + np.eye(1100)))
    # zero diagonal: every pivot is 2x2 or interchanged
    Z = rng.standard_normal((130, 130))
    Z = (Z + Z.T) / 2
    np.fill_diagonal(Z, 0.0)
    out.append(("zero_diagonal_130", Z))
    out.append(("swap_2x2", np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 2.0], [0.0, 2.0, 1e-3]])))
    return out


@pytest.mark.parametrize("name,A", _cases(), ids=[c[0] for c in _cases()])
def test_sytrf_sytrs_against_lapack(env, name, A):
    torch, lib, st = env
    n = A.shape[0]
    print(f"sytrf {name}", flush=True)
    dA, ipiv, info = _factor(env, A)
    ldu, piv_ref, info_ref = sla.lapack.dsytrf(A, lower=1)
    assert info == info_ref == 0
    np.testing.assert_array_equal(ipiv.cpu().numpy(), piv_ref)
    F = dA.cpu().numpy().reshape((n, n), order="F")
    np.testing.assert_allclose(np.tril(F), np.tril(ldu), rtol=0, atol=1e-10 * max(1, n / 64))
    rng = np.random.default_rng(n)
    b = rng.standard_normal(n)
    db = torch.from_numpy(b.copy()).cuda()
    assert lib.msdp_k_sytrs(st, n, C.c_void_p(dA.data_ptr()), n, C.c_void_p(ipiv.data_ptr()),
                            C.c_void_p(db.data_ptr())) == 0
    torch.cuda.synchronize()
    x = db.cpu().numpy()
    ref = np.linalg.solve(A, b)
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-9 * max(1.0, np.abs(ref).max()))


def test_sytrf_zero_pivot_info(env):
    """A zero column (row) is an exactly zero pivot: info names it (1-based),
    as LAPACK's does, and the factorisation of the rest continues."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((40, 40))
    A = A @ A.T + np.eye(40)
    A[17, :] = 0.0
    A[:, 17] = 0.0
    _, ipiv, info = _factor(env, A)
    _, piv_ref, info_ref = sla.lapack.dsytrf(A, lower=1)
    assert info == info_ref and info > 0
    np.testing.assert_array_equal(ipiv.cpu().numpy(), piv_ref)


def _parity(model, **kw):
    okw = dict(kw)
    for k in ("direction", "step", "centering", "initial_point", "algorithm"):
        if k in okw:
            v = okw[k]
            name = type(v).__name__
            cls = getattr(O, name)
            if isinstance(v, M.InitialPoint):
                okw[k] = cls(v.X, v.y, v.x)
            elif isinstance(v, (M.NT, M.HKM)):
                okw[k] = cls()
            else:
                okw[k] = cls(**v.__dict__)
    ref = O.Solver(model, schur=O.MadNLPSchur(), **okw).solve()
    s = M.MadSDPSolver(model, schur=M.MadNLPSchur(), **kw).solve()
    assert int(s.status) == int(ref.status) and s.iter == ref.iter, (s.status, s.iter, ref.iter)
    tr = s.trace()
    for k, r in enumerate(ref.trace):
        sc = max(1.0, abs(r.pobj), abs(r.dobj))
        assert abs(tr[k, 0] - r.pobj) <= 1e-9 * sc and abs(tr[k, 1] - r.dobj) <= 1e-9 * sc, k
        assert abs(tr[k, 2] - r.inf_pr) <= 1e-9 * max(1, r.inf_pr), k
        assert abs(tr[k, 3] - r.inf_du) <= 1e-9 * max(1, r.inf_du), k
    return s, ref


@pytest.mark.parametrize("name", ["maxcut_lro", "mixed", "dense_40_300"])
def test_madnlp_schur_parity(gpu, name):
    model = {"maxcut_lro": lambda: K.maxcut(K.W_LRO), "mixed": K.mixed,
             "dense_40_300": lambda: instances.dense_sdp(40, 300, 20, 5)}[name]()
    print(f"MadNLPSchur {name}", flush=True)
    s, _ = _parity(model)
    assert s.status == M.Status.SOLVE_SUCCEEDED


def test_madnlp_schur_mc_psd_start(gpu):
    """runtests.jl:186-202: from the mc_psd dual-feasible start the dual stays
    feasible (inf_du < 1e-12)."""
    X0, y0 = K.mc_psd_start(K.W_LRO)
    s, _ = _parity(K.maxcut(K.W_LRO), direction=M.HKM(), step=M.CholeskyBacktrack(),
                   centering=M.AdaptiveMu(), initial_point=M.InitialPoint(X0, y0))
    assert s.status == M.Status.SOLVE_SUCCEEDED and s.inf_du < 1e-12


def test_madnlp_schur_dual_centering(gpu):
    """dual_runtests.jl:88-94: DualCentering with MadNLPSchur -> -18."""
    model = K.maxcut(K.W_LRO)
    ref = O.Solver(model, algorithm=O.DualCentering(), schur=O.MadNLPSchur(), tol=1e-6,
                   max_iter=300).solve()
    s = M.MadSDPSolver(model, algorithm=M.DualCentering(), schur=M.MadNLPSchur(), tol=1e-6,
                       max_iter=300).solve()
    assert s.status == M.Status.SOLVE_SUCCEEDED and s.iter == ref.iter
    assert abs(s.dual_objective() + 18.0) <= 1e-3
    assert abs(s.dual_objective() - ref.dual_objective()) <= 1e-9 * 18
