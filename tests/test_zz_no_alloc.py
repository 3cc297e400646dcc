"""msdp_k_gemm_noalloc (partials in the library's per-device scratch array)
against msdp_k_gemm (partials from the stream-ordered pool), through the C
ABI: bit-identical results on two calls in a row, the second smaller, for
the stream-K, split-K and plain-tile paths."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(gpu):
    import torch
    from madsdp_amd import lib as L
    lib = L.load()
    return torch, lib, C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(torch, a):
    return torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).cuda()


def _run(env, fn, ta, tb, M, N, K, seed):
    torch, lib, st = env
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    Cm = rng.standard_normal((M, N))
    dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, Cm)
    rc = getattr(lib, fn)(st, ta, tb, M, N, K, 1.25, C.c_void_p(dA.data_ptr()), A.shape[0],
                          C.c_void_p(dB.data_ptr()), B.shape[0], -0.75, C.c_void_p(dC.data_ptr()), M)
    assert rc == 0, lib.msdp_last_error()
    torch.cuda.synchronize()
    return dC.cpu().numpy(), A, B, Cm


# first the largest (stream-K partials fill the scratch on a 256-CU device),
# then smaller ones: split-K, then plain 64-tiles
SHAPES = [(0, 0, 3000, 3000, 3000), (1, 0, 300, 200, 2000), (0, 1, 100, 90, 40)]


def test_noalloc_matches_gemm(env):
    _, lib, _ = env
    assert lib.msdp_k_gemm_noalloc_scratch_bytes() == 128 << 20
    for k, (ta, tb, M, N, K) in enumerate(SHAPES):
        print(f"noalloc gemm ta={ta} tb={tb} M={M} N={N} K={K}", flush=True)
        ref, A, B, Cm = _run(env, "msdp_k_gemm", ta, tb, M, N, K, k)
        got, _, _, _ = _run(env, "msdp_k_gemm_noalloc", ta, tb, M, N, K, k)
        assert np.array_equal(got, ref), (ta, tb, M, N, K)
        exact = 1.25 * (A.T if ta else A) @ (B.T if tb else B) - 0.75 * Cm
        np.testing.assert_allclose(got.reshape((M, N), order="F"), exact, rtol=0,
                                   atol=1e-12 * max(1, K))
