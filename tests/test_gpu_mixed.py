"""Mixed-precision Schur paths: FP32 MFMA ("mixed") or BF16x3 split-product
("bf16x3") assembly of the dense constraint blocks while mu >= mu_switch,
FP64 Cholesky, and iterative refinement of every Schur solve against the FP64
operator (falling back to an FP64 assembly when the refinement stalls).

The bar: the same termination status and tolerance as the plain FP64 path, in
at most two more iterations, and the same optimum to the tolerance.
"""
import ctypes as C

import numpy as np
import pytest

import madsdp_amd as M
from madsdp_amd import instances

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(gpu):
    import torch
    from madsdp_amd import lib as L
    return torch, L.load(), C.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("Mm,N,K", [(1, 1, 1), (17, 33, 5), (130, 70, 200), (257, 129, 1000)])
def test_gemm32(env, ta, tb, Mm, N, K):
    """FP32 MFMA product widened to FP64 against the FP64 product of the same
    FP32-rounded operands: error within FP32 accumulation roundoff."""
    torch, lib, st = env
    rng = np.random.default_rng(Mm + 3 * N + K)
    A = rng.standard_normal((K, Mm) if ta else (Mm, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    Cm = rng.standard_normal((Mm, N))
    dA = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    dB = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    dC = torch.from_numpy(np.asfortranarray(Cm).ravel(order="F").copy()).cuda()
    print(f"gemm32 ta={ta} tb={tb} M={Mm} N={N} K={K}", flush=True)
    rc = lib.msdp_k_gemm32(st, ta, tb, Mm, N, K, 1.5, C.c_void_p(dA.data_ptr()), A.shape[0],
                           C.c_void_p(dB.data_ptr()), B.shape[0], -0.5, C.c_void_p(dC.data_ptr()), Mm)
    assert rc == 0
    torch.cuda.synchronize()
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = 1.5 * (A64.T if ta else A64) @ (B64.T if tb else B64) - 0.5 * Cm
    got = dC.cpu().numpy().reshape((Mm, N), order="F")
    scale = 1.5 * np.abs(A64.T if ta else A64) @ np.abs(B64.T if tb else B64) + 1.0
    assert np.all(np.abs(got - ref) <= 4 * np.finfo(np.float32).eps * np.sqrt(K) * scale)


LOWP = ["mixed", "bf16x3"]


def _pair(model, precision="mixed", **kw):
    a = M.MadSDPSolver(model, **kw).solve()
    b = M.MadSDPSolver(model, precision=precision, lowp_min_gflop=0.0, **kw).solve()
    return a, b


def _check(a, b, tol=1e-7):
    if b.status != M.Status.SOLVE_SUCCEEDED:
        b.print_trace()
    assert a.status == M.Status.SOLVE_SUCCEEDED
    assert b.status == M.Status.SOLVE_SUCCEEDED
    assert max(b.inf_pr, b.inf_du, b.inf_compl) <= tol
    assert abs(b.iter - a.iter) <= 2
    pa, pb = a.primal_objective(), b.primal_objective()
    assert abs(pa - pb) <= 10 * tol * max(1.0, abs(pa))
    inf = b.info()
    assert inf.lowp_factorizations > 0


@pytest.mark.parametrize("precision", LOWP)
@pytest.mark.parametrize("direction", [M.NT(), M.HKM()])
def test_mixed_dense(gpu, direction, precision):
    a, b = _pair(instances.dense_sdp(40, 300, 20, 5), precision, direction=direction)
    print(f"fp64 {a.iter} it, mixed {b.iter} it, lowp {b.info().lowp_factorizations}, "
          f"refine {b.info().refine_steps}, fallbacks {b.info().fp64_fallbacks}", flush=True)
    _check(a, b)


@pytest.mark.parametrize("precision", LOWP)
def test_mixed_blocks(gpu, precision):
    model = instances.block_sdp(instances.block_sizes(8, 4, 90, 6), 400, 12, 7)
    a, b = _pair(model, precision)
    print(f"fp64 {a.iter} it, mixed {b.iter} it, lowp {b.info().lowp_factorizations}, "
          f"refine {b.info().refine_steps}, fallbacks {b.info().fp64_fallbacks}", flush=True)
    _check(a, b)


@pytest.mark.parametrize("precision", LOWP)
def test_mixed_backtrack_adaptive(gpu, precision):
    model = instances.dense_sdp(30, 150, 0, 9)
    kw = dict(step=M.CholeskyBacktrack(), centering=M.AdaptiveMu())
    a, b = _pair(model, precision, **kw)
    _check(a, b)


def test_mixed_switch_never(gpu):
    """With mu_switch above every mu the mixed path never assembles in FP32 and
    its iterates are the FP64 path's, bit for bit."""
    model = instances.dense_sdp(20, 60, 10, 2)
    a = M.MadSDPSolver(model).solve()
    b = M.MadSDPSolver(model, precision="mixed", lowp_min_gflop=0.0, mu_switch=1e300).solve()
    assert b.info().lowp_factorizations == 0
    assert a.iter == b.iter
    np.testing.assert_array_equal(a.trace(), b.trace())


def test_mixed_maxcut_is_fp64(gpu):
    """Max-cut has only one-hot constraints: nothing is dense, so the mixed
    path is the FP64 path."""
    model = instances.maxcut(60, 0.2, 4)
    a = M.MadSDPSolver(model).solve()
    b = M.MadSDPSolver(model, precision="mixed", lowp_min_gflop=0.0).solve()
    assert b.info().lowp_factorizations == 0
    np.testing.assert_array_equal(a.trace(), b.trace())


def test_precision_option_checked():
    with pytest.raises(ValueError):
        M.MadSDPSolver(instances.maxcut(5, 0.5, 1), precision="fp16")


@pytest.mark.parametrize("precision", LOWP)
def test_mixed_dense_m1100(gpu, precision):
    """m = 1100 (at most n(n+1)/2 + n_s, so H has full rank): the refined
    solves apply the inverses of the factor's full 512-row diagonal blocks."""
    a, b = _pair(instances.dense_sdp(48, 1100, 12, 13), precision)
    print(f"fp64 {a.iter} it, mixed {b.iter} it, lowp {b.info().lowp_factorizations}, "
          f"refine {b.info().refine_steps}, fallbacks {b.info().fp64_fallbacks}", flush=True)
    _check(a, b)


@pytest.mark.parametrize("precision", LOWP)
def test_mixed_error_vs_oracle(gpu, precision):
    """The mixed path's end point against the FP64 oracle's solve of the same
    instance: objective, primal and dual residuals, and X, y, S.  The errors
    are printed (and stated in the design notes); the bounds hold at the 1e-7
    relative gap both solves stop at."""
    from oracle import madsdp_oracle as O
    model = instances.dense_sdp(30, 200, 30, 4)
    ref = O.Solver(model).solve()
    b = M.MadSDPSolver(model, precision=precision, lowp_min_gflop=0.0).solve()
    assert ref.status == O.Status.SOLVE_SUCCEEDED and b.status == M.Status.SOLVE_SUCCEEDED
    assert b.info().lowp_factorizations > 0
    tr = b.trace()
    r = ref.trace[-1]
    e_obj = abs(b.primal_objective() - r.pobj) / max(1.0, abs(r.pobj))
    rel = lambda a, c: np.linalg.norm(a - c) / max(1.0, np.linalg.norm(c))
    e_X = rel(b.X(0), ref.blocks[0].X)
    e_S = rel(b.S(0), ref.blocks[0].S)
    e_y = rel(b.y, ref.y)
    print(f"mixed vs oracle: obj {e_obj:.2e}, inf_pr {tr[-1, 2]:.2e} (oracle {r.inf_pr:.2e}), "
          f"inf_du {tr[-1, 3]:.2e} (oracle {r.inf_du:.2e}), X {e_X:.2e}, y {e_y:.2e}, "
          f"S {e_S:.2e}; iterations {b.iter} vs {ref.iter}", flush=True)
    assert e_obj <= 1e-6
    assert tr[-1, 2] <= 1e-7 and tr[-1, 3] <= 1e-7
    assert max(e_X, e_y, e_S) <= 1e-3
    assert abs(b.iter - ref.iter) <= 2
