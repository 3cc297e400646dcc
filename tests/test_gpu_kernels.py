"""Device building blocks through the C ABI, against numpy/scipy FP64.

Each kernel replaces a LAPACK/BLAS call of the reference (gemm, potrf,
trtri/potri, syevd, lambda_min); the bar is FP64 roundoff relative to the
problem's scale.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(gpu):
    import torch
    from madsdp_amd import lib as L
    lib = L.load()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return torch, lib, stream


def dev(torch, a):
    """Column-major FP64 device copy of a numpy array."""
    return torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).cuda()


def host(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def P(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 33, 5), (64, 64, 16), (130, 70, 200), (257, 3, 129),
                                   (300, 260, 40), (64, 200, 3000), (3, 1, 5000), (600, 520, 64),
                                   (1030, 700, 301), (2200, 2100, 200), (3000, 3000, 3000)])
def test_gemm(env, ta, tb, M, N, K):
    """The two largest shapes have more 128-tiles than CUs: the stream-K
    kernel and its fixup (the max-cut n = 3000 products)."""
    torch, lib, st = env
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    Cm = rng.standard_normal((M, N))
    dA, dB, dC = dev(torch, A), dev(torch, B), dev(torch, Cm)
    print(f"gemm ta={ta} tb={tb} M={M} N={N} K={K}", flush=True)
    rc = lib.msdp_k_gemm(st, ta, tb, M, N, K, 1.5, P(dA), A.shape[0], P(dB), B.shape[0], -0.5, P(dC), M)
    assert rc == 0
    torch.cuda.synchronize()
    ref = 1.5 * (A.T if ta else A) @ (B.T if tb else B) - 0.5 * Cm
    np.testing.assert_allclose(host(dC, (M, N)), ref, rtol=0, atol=1e-12 * max(1, K))


TRI_AL, TRI_AU, TRI_BL, TRI_BU = 1, 2, 4, 8


def _tri_operand(rng, rows, cols, lower):
    """A rows x cols matrix, zero above (lower) or below (upper) its diagonal."""
    X = rng.standard_normal((rows, cols))
    return np.tril(X) if lower else np.triu(X)


@pytest.mark.parametrize("ta,tb,tri,lower,M,N,K", [
    (0, 0, TRI_AL, 0, 3000, 3000, 3000),            # L_X V (stream-K, variable slabs)
    (0, 0, TRI_BL, 0, 3000, 3000, 3000),            # tmp L_X^-1
    (1, 0, TRI_AU | TRI_BL, 1, 3000, 3000, 3000),   # L_S^-T L_S^-1, lower triangle
    (1, 0, TRI_AU | TRI_BL, 0, 3000, 3000, 3000),   # L_S' L_X
    (0, 1, TRI_BU, 0, 3000, 3000, 3000),            # tmp L^-T
    (1, 0, 0, 1, 3000, 3000, 3000),                 # M'M, lower triangle only
    (0, 1, TRI_AL | TRI_BU, 0, 2100, 2100, 2100),
    (0, 0, TRI_AL, 0, 300, 300, 300),               # 64 / 128-tile kernels
    (1, 0, TRI_AU | TRI_BL, 1, 300, 300, 300),
    (0, 1, TRI_BU, 0, 1030, 1030, 1030),
    (0, 0, TRI_BL, 0, 64, 128, 3000),               # split-K over a lower trapezoid
    (1, 1, TRI_AU | TRI_BU, 0, 130, 130, 130),
])
def test_gemm_tri(env, ta, tb, tri, lower, M, N, K):
    """Products with triangular operands run each tile over the K range where
    both operands can be nonzero (and a lower-triangle output over the tiles
    on or below the diagonal only): the result is numpy's full product of the
    same triangular matrices, and a lower product leaves the upper triangle
    as it was."""
    torch, lib, st = env
    rng = np.random.default_rng(M + 3 * K + 11 * tri + 5 * lower)
    # op(A) is M x K, op(B) is K x N; the triangle is that of op(X)
    opA = _tri_operand(rng, M, K, True) if tri & TRI_AL else (
        _tri_operand(rng, M, K, False) if tri & TRI_AU else rng.standard_normal((M, K)))
    opB = _tri_operand(rng, K, N, True) if tri & TRI_BL else (
        _tri_operand(rng, K, N, False) if tri & TRI_BU else rng.standard_normal((K, N)))
    A = opA.T.copy() if ta else opA
    B = opB.T.copy() if tb else opB
    Cm = rng.standard_normal((M, N))
    dA, dB, dC = dev(torch, A), dev(torch, B), dev(torch, Cm)
    print(f"gemm_tri ta={ta} tb={tb} tri={tri} lower={lower} M={M} N={N} K={K}", flush=True)
    rc = lib.msdp_k_gemm_tri(st, ta, tb, M, N, K, 1.5, P(dA), A.shape[0], P(dB), B.shape[0], -0.5,
                             P(dC), M, lower, tri)
    assert rc == 0
    torch.cuda.synchronize()
    ref = 1.5 * opA @ opB - 0.5 * Cm
    got = host(dC, (M, N))
    if lower:
        iu = np.triu_indices(M, 1, N)
        np.testing.assert_array_equal(got[iu], Cm[iu])
        ref = np.tril(ref) + np.triu(Cm, 1)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * max(1, K))


def test_gemm_tri_rejects_bad_flags(env):
    torch, lib, st = env
    d = torch.zeros(4, dtype=torch.float64, device="cuda")
    assert lib.msdp_k_gemm_tri(st, 0, 0, 2, 2, 2, 1.0, P(d), 2, P(d), 2, 0.0, P(d), 2, 0, 16) == -2


@pytest.mark.parametrize("trans,N,K", [(1, 300, 700), (0, 257, 64), (1, 2100, 1100),
                                       (0, 2304, 1030)])
def test_syrk_lower(env, trans, N, K):
    """Lower triangle of A'A (trans) or AA'; the upper triangle is untouched.
    The two large cases take the block-row path (off-diagonal rectangles as
    plain products, diagonal blocks batched)."""
    torch, lib, st = env
    rng = np.random.default_rng(N + K)
    A = rng.standard_normal((K, N) if trans else (N, K))
    Cm = rng.standard_normal((N, N))
    dA, dC = dev(torch, A), dev(torch, Cm)
    print(f"syrk trans={trans} N={N} K={K}", flush=True)
    assert lib.msdp_k_syrk_lower(st, trans, N, K, 0.75, P(dA), A.shape[0], 2.0, P(dC), N) == 0
    torch.cuda.synchronize()
    X = A.T if trans else A
    ref = 0.75 * X @ X.T + 2.0 * Cm
    got = host(dC, (N, N))
    low = np.tril_indices(N)
    np.testing.assert_allclose(got[low], ref[low], rtol=0, atol=1e-12 * K)
    up = np.triu_indices(N, 1)
    assert np.array_equal(got[up], Cm[up])


def _bf16(x):
    """Round-to-nearest-even of float32(x) to bfloat16, as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("N,K", [(1, 1), (129, 31), (300, 1000), (700, 2500), (1300, 96)])
def test_syrk_bf16x3(env, N, K):
    """C = alpha (hi'hi + hi'lo + lo'hi) + beta C on the lower triangle, the
    upper triangle untouched.  Against the same three products formed in
    FP64 from a numpy emulation of the split (hi = bf16(a), lo = bf16(a -
    hi)) the only difference is the FP32 accumulation (a transposed operand
    or a wrong lane would be off by O(1)); against A'A in FP64 the error is
    the split's (about 2^-17 per entry).  N = 1300 spans two 8 x 8
    super-tiles of 128-tiles (the XCD grouping of the tile order)."""
    torch, lib, st = env
    rng = np.random.default_rng(N * 3 + K)
    A = rng.standard_normal((K, N)) * np.exp(rng.uniform(-3, 3, size=(1, N)))
    Cm = rng.standard_normal((N, N))
    dA, dC = dev(torch, A), dev(torch, Cm)
    work = torch.zeros(int(lib.msdp_k_syrk_bf16x3_work_bytes(N, K)) // 8 + 1, dtype=torch.float64,
                       device="cuda")
    print(f"syrk_bf16x3 N={N} K={K}", flush=True)
    assert lib.msdp_k_syrk_bf16x3(st, N, K, 0.5, P(dA), K, 2.0, P(dC), N, P(work)) == 0
    torch.cuda.synchronize()
    got = host(dC, (N, N))
    hi = _bf16(A)
    lo = _bf16(A - hi)
    emul = 0.5 * (hi.T @ hi + hi.T @ lo + lo.T @ hi) + 2.0 * Cm
    scale = 0.5 * np.abs(hi).T @ np.abs(hi) + 2.0 * np.abs(Cm)
    low = np.tril_indices(N)
    # FP32 accumulation of K products (and the FP64 widening of alpha * acc)
    assert np.all(np.abs(got - emul)[low] <= (4 * K + 8) * 2.0 ** -24 * scale[low])
    exact = 0.5 * A.T @ A + 2.0 * Cm
    err = np.abs(got - exact)[low] / scale[low]
    print(f"  max error vs FP64 A'A: {err.max():.2e} (relative to |A|'|A|)", flush=True)
    assert err.max() <= 2.0 ** -14
    up = np.triu_indices(N, 1)
    assert np.array_equal(got[up], Cm[up])


@pytest.mark.parametrize("rows,cols", [(1, 1), (300, 77), (4000, 1100)])
@pytest.mark.parametrize("acc", [0, 1])
def test_panel_apply(env, rows, cols, acc):
    """out (+)= B'(B v) in FP64 (the refinement's operator of a BF16x3 Schur
    complement)."""
    torch, lib, st = env
    rng = np.random.default_rng(rows + cols + acc)
    B = rng.standard_normal((rows, cols))
    v = rng.standard_normal(cols)
    o = rng.standard_normal(cols)
    dB, dv, do = dev(torch, B), dev(torch, v), dev(torch, o)
    work = torch.zeros(int(lib.msdp_k_panel_apply_work(rows, cols)), dtype=torch.float64, device="cuda")
    print(f"panel_apply rows={rows} cols={cols} acc={acc}", flush=True)
    assert lib.msdp_k_panel_apply(st, rows, cols, P(dB), rows, P(dv), P(do), acc, P(work)) == 0
    torch.cuda.synchronize()
    ref = B.T @ (B @ v) + (o if acc else 0.0)
    scale = np.abs(B).T @ (np.abs(B) @ np.abs(v)) + np.abs(o)
    assert np.all(np.abs(do.cpu().numpy() - ref) <= 1e-13 * (rows + cols) * scale)


@pytest.mark.parametrize("n", [1, 5, 64, 65, 100, 128, 200, 300])
def test_cholesky_factor(env, n):
    torch, lib, st = env
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    M = G @ G.T + n * np.eye(n)
    M[np.triu_indices(n, 1)] += 1e-3  # asymmetric input: factor of (M+M')/2
    dM, dL = dev(torch, M), dev(torch, np.zeros((n, n)))
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    scratch = torch.zeros(int(lib.msdp_k_cholesky_factor_scratch(n)), dtype=torch.float64, device="cuda")
    print(f"cholesky n={n}", flush=True)
    assert lib.msdp_k_cholesky_factor(st, n, P(dM), n, P(dL), n, P(info), P(scratch)) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    Lh = host(dL, (n, n))
    ref = np.linalg.cholesky((M + M.T) / 2)
    np.testing.assert_allclose(Lh, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert np.all(np.triu(Lh, 1) == 0)


def test_cholesky_not_pd(env):
    torch, lib, st = env
    for M, pivot in ((np.array([[1.0, 2.0], [2.0, 1.0]]), 2), (np.zeros((2, 2)), 1)):
        dM, dL = dev(torch, M), dev(torch, np.zeros((2, 2)))
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        scratch = torch.zeros(int(lib.msdp_k_cholesky_factor_scratch(2)), dtype=torch.float64, device="cuda")
        assert lib.msdp_k_cholesky_factor(st, 2, P(dM), 2, P(dL), 2, P(info), P(scratch)) == 0
        torch.cuda.synchronize()
        assert int(info.item()) == pivot


@pytest.mark.parametrize("n,bad", [(100, 10), (100, 70), (128, 127), (65, 64)])
def test_cholesky_mid_not_pd(env, n, bad):
    """65 <= n <= 128 (one workgroup): a negative pivot in the first 64
    columns or in the trailing block is reported by its 1-based column."""
    torch, lib, st = env
    M = np.eye(n)
    M[bad, bad] = -1.0
    dM, dL = dev(torch, M), dev(torch, np.zeros((n, n)))
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    scratch = torch.zeros(int(lib.msdp_k_cholesky_factor_scratch(n)), dtype=torch.float64, device="cuda")
    print(f"cholesky not pd n={n} bad={bad}", flush=True)
    assert lib.msdp_k_cholesky_factor(st, n, P(dM), n, P(dL), n, P(info), P(scratch)) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == bad + 1


def test_cholesky_flush(env):
    """test/blocks.jl:24-34: entries below eps^2 max|L_ii| are flushed."""
    torch, lib, st = env
    eps = np.finfo(float).eps
    F = np.array([[1, 0, 0], [np.finfo(float).tiny, 1, 0], [eps, 1, 1]])
    M = F @ F.T
    dM, dL = dev(torch, M), dev(torch, np.zeros((3, 3)))
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    scratch = torch.zeros(int(lib.msdp_k_cholesky_factor_scratch(3)), dtype=torch.float64, device="cuda")
    assert lib.msdp_k_cholesky_factor(st, 3, P(dM), 3, P(dL), 3, P(info), P(scratch)) == 0
    torch.cuda.synchronize()
    Lh = host(dL, (3, 3))
    assert Lh[1, 0] == 0.0
    assert Lh[2, 0] == eps
    np.testing.assert_allclose(Lh @ Lh.T, M, atol=1e-15)


@pytest.mark.parametrize("n", [1, 7, 64, 65, 100, 128, 130, 333])
def test_trtri(env, n):
    torch, lib, st = env
    rng = np.random.default_rng(n + 11)
    Lm = np.tril(rng.standard_normal((n, n))) + 4 * np.eye(n)
    dL, dI = dev(torch, Lm), dev(torch, np.full((n, n), 7.0))
    work = torch.zeros(int(lib.msdp_k_trtri_work(n)) + 1, dtype=torch.float64, device="cuda")
    print(f"trtri n={n}", flush=True)
    assert lib.msdp_k_trtri(st, n, P(dL), n, P(dI), n, P(work)) == 0
    torch.cuda.synchronize()
    ref = sla.solve_triangular(Lm, np.eye(n), lower=True)
    np.testing.assert_allclose(host(dI, (n, n)), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def _spd_clustered(rng, n, clusters):
    """SPD matrix with clustered and repeated eigenvalues (the near-central
    NT scaling point is a multiple of the identity)."""
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sort(rng.choice(clusters, size=n)) * (1 + 1e-13 * rng.standard_normal(n))
    return (Q * lam) @ Q.T


@pytest.mark.parametrize("n", [2, 10, 64, 65, 100, 257, 600, 700, 1100])
@pytest.mark.parametrize("kind", ["random", "clustered", "identity"])
def test_syevd(env, n, kind):
    """n <= 64 Jacobi; <= 640 one-workgroup tridiagonalisation; above, the
    blocked one; then divide and conquer and the back-transform."""
    torch, lib, st = env
    rng = np.random.default_rng(n)
    if kind == "random":
        A = rng.standard_normal((n, n))
        A = (A + A.T) / 2
    elif kind == "clustered":
        A = _spd_clustered(rng, n, np.array([1.0, 1.0 + 1e-9, 2.0, 5.0]))
        A = (A + A.T) / 2
    else:
        A = 3.0 * np.eye(n)
    dA = dev(torch, A)
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    V = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    work = torch.zeros(int(lib.msdp_k_syevd_work(n)), dtype=torch.float64, device="cuda")
    iwork = torch.zeros(int(lib.msdp_k_syevd_iwork(n)), dtype=torch.int32, device="cuda")
    print(f"syevd n={n} kind={kind}", flush=True)
    assert lib.msdp_k_syevd(st, n, P(dA), n, P(w), P(V), n, 1, P(work), P(iwork)) == 0
    torch.cuda.synchronize()
    wh, Vh = w.cpu().numpy(), host(V, (n, n))
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    np.testing.assert_allclose(wh, ref, rtol=0, atol=1e-12 * scale * max(1, n / 64))
    np.testing.assert_allclose(Vh.T @ Vh, np.eye(n), rtol=0, atol=1e-11 * max(1, n / 64))
    np.testing.assert_allclose(A @ Vh, Vh * wh, rtol=0, atol=1e-11 * scale * max(1, n / 64))


def test_syevd_graph_replay(env):
    """n > 640 on a created stream: the blocked tridiagonalisation's column
    sequence is captured as a HIP graph on the first call and replayed on the
    next ones (same workspace).  A second matrix in between, then the first
    again: bit-identical to the first call, every call exact against numpy.
    n = 3000 is the max-cut workload's block."""
    torch, lib, _ = env
    stream = torch.cuda.Stream()
    st = C.c_void_p(stream.cuda_stream)
    for n in (1100, 3000):
        rng = np.random.default_rng(n + 1)
        mats = []
        for _ in range(2):
            A = rng.standard_normal((n, n))
            mats.append((A + A.T) / 2)
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        V = torch.zeros(n * n, dtype=torch.float64, device="cuda")
        work = torch.zeros(int(lib.msdp_k_syevd_work(n)), dtype=torch.float64, device="cuda")
        iwork = torch.zeros(int(lib.msdp_k_syevd_iwork(n)), dtype=torch.int32, device="cuda")
        outs = []
        for k in (0, 1, 0):
            A = mats[k]
            dA = dev(torch, A)
            print(f"syevd graph n={n} call matrix {k}", flush=True)
            with torch.cuda.stream(stream):
                assert lib.msdp_k_syevd(st, n, P(dA), n, P(w), P(V), n, 1, P(work), P(iwork)) == 0
            stream.synchronize()
            wh, Vh = w.cpu().numpy(), host(V, (n, n))
            outs.append((wh, Vh))
            ref = np.linalg.eigvalsh(A)
            scale = max(1.0, np.abs(ref).max())
            np.testing.assert_allclose(wh, ref, rtol=0, atol=1e-12 * scale * n / 64)
            np.testing.assert_allclose(A @ Vh, Vh * wh, rtol=0, atol=1e-11 * scale * n / 64)
        assert np.array_equal(outs[0][0], outs[2][0]) and np.array_equal(outs[0][1], outs[2][1])


def test_syevd_deflation_global_path(env):
    """n = 4500: the last divide-and-conquer merge has K > 4096 entries, so
    the deflation and secular kernels keep their vectors in global memory
    instead of LDS."""
    torch, lib, st = env
    n = 4500
    rng = np.random.default_rng(45)
    A = rng.standard_normal((n, n))
    A = (A + A.T) / 2
    dA = dev(torch, A)
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    V = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    work = torch.zeros(int(lib.msdp_k_syevd_work(n)), dtype=torch.float64, device="cuda")
    iwork = torch.zeros(int(lib.msdp_k_syevd_iwork(n)), dtype=torch.int32, device="cuda")
    print(f"syevd n={n}", flush=True)
    assert lib.msdp_k_syevd(st, n, P(dA), n, P(w), P(V), n, 1, P(work), P(iwork)) == 0
    torch.cuda.synchronize()
    wh, Vh = w.cpu().numpy(), host(V, (n, n))
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    np.testing.assert_allclose(wh, ref, rtol=0, atol=1e-12 * scale * n / 64)
    np.testing.assert_allclose(A @ Vh, Vh * wh, rtol=0, atol=1e-11 * scale * n / 64)


@pytest.mark.parametrize("n,kind", [(3, "random"), (64, "random"), (65, "random"), (200, "random"),
                                    (400, "step"), (500, "random"), (640, "cluster"),
                                    (641, "random"), (1000, "random"), (1000, "cluster"),
                                    (1500, "step")])
def test_eig_min(env, n, kind):
    """Below 400 by tridiagonalisation + Sturm; 400..640 by the one-workgroup
    Lanczos kernel, above by the multi-launch Lanczos."""
    torch, lib, st = env
    rng = np.random.default_rng(n + 3)
    if kind == "random":
        A = rng.standard_normal((n, n))
        A = (A + A.T) / 2
    elif kind == "cluster":  # three nearly equal smallest eigenvalues
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        lam = np.concatenate([[-2.0, -2.0 + 1e-9, -2.0 + 2e-9], rng.uniform(-1, 3, n - 3)])
        A = (Q * lam) @ Q.T
        A = (A + A.T) / 2
    else:  # L^-1 D L^-T of an IPM-like step: X = I + low-rank, D indefinite
        G = rng.standard_normal((n, 20))
        X = np.eye(n) + G @ G.T / n
        D = rng.standard_normal((n, n)) / np.sqrt(n)
        D = (D + D.T) / 2 - 0.3 * np.eye(n)
        Li = np.linalg.inv(np.linalg.cholesky(X))
        A = Li @ D @ Li.T
        A = (A + A.T) / 2
    dA = dev(torch, A)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    work = torch.zeros(int(lib.msdp_k_syevd_work(n)), dtype=torch.float64, device="cuda")
    assert lib.msdp_k_eig_min(st, n, P(dA), n, P(out), P(work)) == 0
    torch.cuda.synchronize()
    ref = np.linalg.eigvalsh(A)[0]
    assert abs(out.item() - ref) <= 1e-12 * max(1, np.abs(A).max() * np.sqrt(n))


@pytest.mark.parametrize("n", [100, 250, 500, 700, 1100])
def test_eig_min2(env, n):
    """The primal/dual pair in shared launches: each result is the single
    call's, bit for bit, and the smallest eigenvalue to 1e-12 (the two
    Lanczos processes of n > 640 converge at different steps)."""
    torch, lib, st = env
    rng = np.random.default_rng(n + 11)
    A0 = rng.standard_normal((n, n))
    A0 = (A0 + A0.T) / 2
    G = rng.standard_normal((n, 10))
    A1 = G @ G.T / n - 0.5 * np.eye(n) + np.diag(rng.uniform(0, 2, n))
    A1 = (A1 + A1.T) / 2
    d0, d1 = dev(torch, A0), dev(torch, A1)
    outs = torch.zeros(4, dtype=torch.float64, device="cuda")
    work = torch.zeros(int(lib.msdp_k_syevd_work(n)), dtype=torch.float64, device="cuda")
    base = outs.data_ptr()
    assert lib.msdp_k_eig_min2(st, n, P(d0), P(d1), n, base, base + 8, P(work)) == 0
    assert lib.msdp_k_eig_min(st, n, P(d0), n, base + 16, P(work)) == 0
    assert lib.msdp_k_eig_min(st, n, P(d1), n, base + 24, P(work)) == 0
    torch.cuda.synchronize()
    o = outs.cpu().numpy()
    assert o[0] == o[2] and o[1] == o[3], o
    for v, A in ((o[0], A0), (o[1], A1)):
        ref = np.linalg.eigvalsh(A)[0]
        assert abs(v - ref) <= 1e-12 * max(1, np.abs(A).max() * np.sqrt(n))


def test_schur_onehot(env):
    """Max-cut Schur complement H = W o W (lower), src/components.jl:102-104."""
    torch, lib, st = env
    rng = np.random.default_rng(5)
    n = 300
    W = rng.standard_normal((n, n))
    W = W + W.T
    dW, dH = dev(torch, W), dev(torch, np.zeros((n, n)))
    assert lib.msdp_k_schur_onehot(st, n, P(dW), n, P(dW), n, P(dH), n) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(np.tril(host(dH, (n, n))), np.tril(W * W))
