"""Low-rank constraint storage on the device (A(X) = F'XF column dots,
A*(y) = F diag(w y) F', H_jk = sum w_r w_s (f_r' L f_s)(f_s' R f_r), the QR
packing svec(sum w (G'f)(G'f)')) against the FP64 oracle, which takes the
same A_j = sum_r w_r f_r f_r' by its definition: per-iteration objective and
residuals to 1e-9, same iteration count and status."""
import numpy as np
import pytest

import madsdp_amd as M
from madsdp_amd import frontend as F
from madsdp_amd import instances
from madsdp_amd.model import BlockConstraints, Model
from oracle import madsdp_oracle as O

pytestmark = pytest.mark.gpu


def _oracle_kw(kw):
    okw = {}
    for k, v in kw.items():
        if k in ("direction", "step", "centering", "algorithm", "schur"):
            okw[k] = getattr(O, type(v).__name__)()
        else:
            okw[k] = v
    return okw


def _parity(model, **kw):
    s = M.MadSDPSolver(model, **kw).solve()
    r = O.Solver(model, **_oracle_kw(kw)).solve()
    assert int(s.status) == int(r.status) and s.iter == r.iter, (s.status, s.iter, r.iter)
    tr = s.trace()
    for k, rec in enumerate(r.trace):
        sc = max(1.0, abs(rec.pobj), abs(rec.dobj))
        assert abs(tr[k, 0] - rec.pobj) <= 1e-9 * sc and abs(tr[k, 1] - rec.dobj) <= 1e-9 * sc, k
        assert abs(tr[k, 2] - rec.inf_pr) <= 1e-9 * max(1, rec.inf_pr), k
        assert abs(tr[k, 3] - rec.inf_du) <= 1e-9 * max(1, rec.inf_du), k
    return s, r


CASES = {
    "rank1_nt_scalars": (lambda: instances.lowrank_sdp(40, 150, 5, 11), {}),
    "rank2_signed_nt": (lambda: instances.lowrank_sdp(30, 200, 0, 12, rank=2, signed=True), {}),
    "rank1_hkm": (lambda: instances.lowrank_sdp(40, 150, 5, 13), dict(direction=M.HKM())),
    "rank2_hkm_backtrack": (lambda: instances.lowrank_sdp(24, 120, 0, 14, rank=2, signed=True),
                            dict(direction=M.HKM(), step=M.CholeskyBacktrack())),
    "rank1_dualcentering": (lambda: instances.lowrank_sdp(30, 100, 0, 15),
                            dict(algorithm=M.DualCentering())),
    "rank1_qr": (lambda: instances.lowrank_sdp(20, 90, 4, 16), dict(schur=M.QRSchur())),
    "rank1_n200_m800": (lambda: instances.lowrank_sdp(200, 800, 0, 17), {}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_lowrank_parity(gpu, name):
    make, kw = CASES[name]
    print(f"lowrank {name}", flush=True)
    s, _ = _parity(make(), **kw)
    assert s.status == M.Status.SOLVE_SUCCEEDED


def test_lowrank_block_beside_dense_block(gpu):
    """Two blocks sharing constraints: one low-rank, one dense; X = I and
    S = I at y0 are strictly feasible."""
    rng = np.random.default_rng(21)
    n1, n2, m = 18, 10, 60
    fs = rng.standard_normal((m, n1)) / np.sqrt(n1)
    idx = np.arange(0, m, 3)
    mats = instances._sym_normal(rng, idx.size, n2)
    y0 = rng.standard_normal(m)
    C1 = np.eye(n1) + sum(y * np.outer(f, f) for y, f in zip(y0, fs))
    C2 = np.eye(n2) + sum(y0[j] * A for j, A in zip(idx, mats))
    b = np.array([f @ f for f in fs])
    for j, A in zip(idx, mats):
        b[j] += np.trace(A)
    blk1 = BlockConstraints.lowrank(n1, np.arange(m), [[(1.0, f)] for f in fs])
    model = Model([C1, C2], [blk1, BlockConstraints("dense", n2, con_idx=idx, mats=mats)],
                  b, np.zeros(0), np.zeros((m, 0)))
    print("lowrank + dense blocks", flush=True)
    s, _ = _parity(model)
    assert s.status == M.Status.SOLVE_SUCCEEDED


def test_lowrank_matches_dense_storage(gpu):
    """The same matrices stored as factors and dense: same iterations, same
    objective to 1e-9."""
    m = instances.lowrank_sdp(60, 400, 6, 23)
    blk = m.A[0]
    md = Model(m.C, [BlockConstraints("dense", 60, con_idx=blk.con_idx, mats=blk.dense_mats())],
               m.b, m.d_lin, m.C_lin)
    a, b = M.MadSDPSolver(m).solve(), M.MadSDPSolver(md).solve()
    assert a.status == b.status == M.Status.SOLVE_SUCCEEDED and a.iter == b.iter
    sc = max(1.0, abs(b.primal_objective()))
    assert abs(a.primal_objective() - b.primal_objective()) <= 1e-9 * sc
    np.testing.assert_allclose(a.y, b.y, rtol=1e-7, atol=1e-9)


def test_frontend_factored_constraints(gpu):
    """LowRankMatrix entries through ingest, solve and the SolverCore stats."""
    rng = np.random.default_rng(24)
    n, m = 16, 50
    us = rng.standard_normal((m, n)) / 4
    y0 = rng.standard_normal(m)
    C = np.eye(n) + sum(y * np.outer(u, u) for y, u in zip(y0, us))
    b = np.array([u @ u for u in us])
    lro = F.LroModel([C], [[F.LowRankMatrix(u[:, None], [1.0]) for u in us]], b)
    solver = F.Solver(lro)
    st = solver.solve()
    assert solver.model.A[0].kind == "lowrank"
    ref = O.Solver(F.ingest(lro)).solve()
    assert st.status == "first_order" and st.iter == ref.iter
    assert abs(st.objective - ref.primal_objective()) <= 1e-9 * max(1, abs(ref.primal_objective()))
