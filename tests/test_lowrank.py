"""Low-rank constraint storage (LowRankOpt's rank-one / factored matrices,
reference README.md:148) on the host: the factors' dense matrices, rank-one
detection, the front end's choice of storage, and the oracle, which takes
A_j = sum_r w_r f_r f_r' by its definition."""
import numpy as np
import scipy.sparse as sp

import kat_models as K
from madsdp_amd import frontend as F
from madsdp_amd import instances
from madsdp_amd.model import BlockConstraints, Model, rank_one_factor
from oracle import madsdp_oracle as O


def test_dense_mats_are_the_weighted_outer_products():
    rng = np.random.default_rng(1)
    f1, f2, f3 = rng.standard_normal((3, 7))
    blk = BlockConstraints.lowrank(7, [2, 5], [[(1.0, f1)], [(-2.0, f2), (0.5, f3)]])
    assert blk.kind == "lowrank" and blk.num_touching == 2 and not blk.is_one_hot()
    assert list(blk.rank_ptr) == [0, 1, 3] and blk.factors.shape == (7, 3)
    M = blk.dense_mats()
    np.testing.assert_array_equal(M[0], np.outer(f1, f1))
    np.testing.assert_allclose(M[1], -2.0 * np.outer(f2, f2) + 0.5 * np.outer(f3, f3), rtol=1e-15)


def test_rank_one_factor():
    rng = np.random.default_rng(2)
    u = rng.standard_normal(9)
    for sign in (1.0, -1.0):
        w, f = rank_one_factor(sign * np.outer(u, u))
        assert w == sign
        np.testing.assert_allclose(w * np.outer(f, f), sign * np.outer(u, u), atol=1e-14)
    v = rng.standard_normal(9)
    assert rank_one_factor(np.outer(u, u) + np.outer(v, v)) is None
    assert rank_one_factor(np.zeros((3, 3))) is None
    # e_2 e_2' is rank one too
    E = np.zeros((4, 4)); E[2, 2] = 3.0
    w, f = rank_one_factor(E)
    np.testing.assert_allclose(w * np.outer(f, f), E)


def test_oracle_lowrank_equals_dense():
    """The oracle solves the factored model exactly as the same matrices
    stored dense (identical iterates: the factors are expanded first)."""
    m = instances.lowrank_sdp(12, 40, 2, 3, rank=2, signed=True)
    blk = m.A[0]
    md = Model(m.C, [BlockConstraints("dense", 12, con_idx=blk.con_idx, mats=blk.dense_mats())],
               m.b, m.d_lin, m.C_lin)
    a, b = O.Solver(m).solve(), O.Solver(md).solve()
    assert a.status == b.status == O.Status.SOLVE_SUCCEEDED and a.iter == b.iter
    assert a.primal_objective() == b.primal_objective()
    np.testing.assert_array_equal(a.y, b.y)


def test_lowrank_instance_is_strictly_feasible():
    """X = I, x = 1 satisfies A(X) + C_lin x = b; S = C - A*(y0) = I."""
    m = instances.lowrank_sdp(10, 30, 4, 7)
    mats = m.A[0].dense_mats()
    Ax = np.array([np.trace(A) for A in mats]) + m.C_lin.sum(axis=1)
    np.testing.assert_allclose(Ax, m.b, rtol=1e-13)


def test_ingest_keeps_factors():
    rng = np.random.default_rng(4)
    n, m = 8, 20
    us = rng.standard_normal((m, n))
    C = np.eye(n)
    # explicit factors
    lro = F.LroModel([C], [[F.LowRankMatrix(u[:, None], [1.0]) for u in us]], np.ones(m))
    mod = F.ingest(lro)
    assert mod.A[0].kind == "lowrank" and mod.A[0].weights.size == m
    # dense rank-one matrices: detected, or kept dense without detection
    lro2 = F.LroModel([C], [[np.outer(u, u) for u in us]], np.ones(m))
    assert F.ingest(lro2).A[0].kind == "lowrank"
    assert F.ingest(lro2, detect_rank_one=False).A[0].kind == "dense"
    np.testing.assert_allclose(F.ingest(lro2).A[0].dense_mats(),
                               np.stack([np.outer(u, u) for u in us]), atol=1e-13)
    # one-hot diagonal constraints stay on the COO / W o W path
    mc = F.ingest(F.LroModel([sp.csr_matrix(-K.laplacian(K.W_LRO) / 4)],
                             [[sp.coo_matrix(([1.0], ([k], [k])), shape=(4, 4)) for k in range(4)]],
                             np.ones(4)))
    assert mc.A[0].kind == "coo" and mc.A[0].is_one_hot()
    # a block mixing rank-one and rank-two matrices is stored dense
    mixed = [np.outer(us[0], us[0]), np.outer(us[1], us[1]) + np.outer(us[2], us[2])]
    assert F.ingest(F.LroModel([C], [mixed], np.ones(2))).A[0].kind == "dense"
