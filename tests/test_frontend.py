"""Front-end ingest/egress (src/moi.jl:53-78): a LowRankOpt-shaped problem with
scipy.sparse / dense / absent constraint matrices mapped to the device upload
format, and the SolverCore statistics and shaped solution the reference
fills, against the FP64 oracle on the reference tests' known answers."""
import numpy as np
import pytest
import scipy.sparse as sp

import kat_models as K
from madsdp_amd import frontend as F
from madsdp_amd.model import Model
from oracle import madsdp_oracle as O


def lro_maxcut(W):
    """min <-L/4, X> s.t. X_kk = 1 with sparse e_k e_k' (test/runtests.jl:72-93)."""
    n = W.shape[0]
    A = [[sp.coo_matrix(([1.0], ([k], [k])), shape=(n, n)) for k in range(n)]]
    return F.LroModel([sp.csr_matrix(-K.laplacian(W) / 4)], A, np.ones(n), dims=[n])


def lro_mixed():
    """1x1 SDP + 2 LP scalars -> 1 (test/runtests.jl:113-129)."""
    one = np.ones((1, 1))
    return F.LroModel([one], [[one, one]], [1.0, 1.0], d_lin=[2.0, 1.0], C_lin=np.eye(2))


def lro_dense(n=6, m=9, seed=3):
    """Dense constraints on one block and None for a second block that no
    constraint touches."""
    rng = np.random.default_rng(seed)
    mats = []
    for _ in range(m):
        A = rng.standard_normal((n, n))
        mats.append((A + A.T) / 2)
    C1 = np.eye(n) * 2 + sum(0.1 * A for A in mats)
    return F.LroModel([C1, np.eye(2)], [mats, [None] * m],
                      [np.trace(A) for A in mats], dims=[n, 2])


def test_ingest_maps_storage_like_datasparsity():
    mc = F.ingest(lro_maxcut(K.W_LRO))
    assert mc.A[0].kind == "coo" and mc.A[0].is_one_hot()
    ref = K.maxcut(K.W_LRO)
    np.testing.assert_array_equal(mc.C[0], ref.C[0])
    for a in ("con", "row", "col", "val"):
        np.testing.assert_array_equal(getattr(mc.A[0], a), getattr(ref.A[0], a))
    d = F.ingest(lro_dense())
    assert d.A[0].kind == "dense" and d.A[0].mats.shape == (9, 6, 6)
    assert d.A[1].kind == "coo" and d.A[1].con.size == 0
    # the same matrices stay COO when datasparsity allows all 36 entries
    assert F.ingest(lro_dense(), datasparsity=36).A[0].kind == "coo"
    mx = F.ingest(lro_mixed())
    assert mx.n_scalar == 2 and np.array_equal(mx.C_lin, np.eye(2))


def test_status_symbols():
    """src/moi.jl:16-27."""
    assert F.solvercore_status(F.Status.SOLVE_SUCCEEDED) == "first_order"
    assert F.solvercore_status(F.Status.ERROR_IN_STEP_COMPUTATION) == "exception"
    assert F.solvercore_status(F.Status.MAXIMUM_ITERATIONS_EXCEEDED) == "max_iter"
    assert F.solvercore_status(F.Status.REGULAR) == "unknown"


def test_ingest_rejects_wrong_dims():
    bad = lro_dense()
    bad.dims = [5, 2]
    with pytest.raises(ValueError):
        F.ingest(bad)


CASES = [("maxcut_lro", lambda: lro_maxcut(K.W_LRO), -18.0), ("mixed", lro_mixed, 1.0),
         ("dense", lro_dense, None)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,known", CASES, ids=[c[0] for c in CASES])
def test_solve_fills_stats(gpu, name, make, known):
    lro = make()
    print(f"frontend {name}", flush=True)
    s = F.Solver(lro)
    st = s.solve()
    ref = O.Solver(F.ingest(lro)).solve()
    assert st.status == "first_order" and int(st.solver_status) == int(ref.status)
    assert st.iter == ref.iter
    scale = max(1.0, abs(ref.primal_objective()))
    assert abs(st.objective - ref.primal_objective()) <= 1e-9 * scale
    assert abs(st.dual_objective - ref.dual_objective()) <= 1e-9 * scale
    np.testing.assert_allclose(st.multipliers, ref.y, rtol=1e-7, atol=1e-9)
    assert st.primal_feas <= 1e-7 and st.dual_feas <= 1e-7
    if known is not None:
        assert abs(st.objective - known) <= 1e-4
    sol = s.solution()
    assert len(sol.X) == len(lro.C) and sol.x.size == s.model.n_scalar
    for X, b in zip(sol.X, ref.blocks):
        np.testing.assert_allclose(X, b.X, rtol=1e-6, atol=1e-8)
