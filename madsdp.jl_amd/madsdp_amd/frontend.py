"""Front-end ingest and egress: a LowRankOpt-shaped problem in, the SolverCore
statistics and the shaped primal solution out.

The reference's front end hands MadSDP an `LRO.Model(C, A, b, d_lin, C_lin,
dims)` (constructor order of test/runtests.jl:30,55), wraps it in
`LRO.BufferedModelForSchur(model, datasparsity)` (src/solver.jl:98-100), and
after `solve!` fills `SolverCore.GenericExecutionStats` (src/moi.jl:53-72)
and answers `MOI.get(::Solver, ::LRO.Solution)` with an
`LRO.ShapedSolution(x, [X_i])` (src/moi.jl:76-78).  This module is that
plumbing for the device solver:

* `LroModel` holds the same fields: per-block cost matrices C_i (dense or
  scipy.sparse), the nblocks x m table A[i][j] of constraint matrices
  (scipy.sparse, dense or None for zero), b, d_lin and C_lin (m x n_s).
* `ingest(lro, datasparsity, detect_rank_one)` maps it once to the device
  upload format (`Model`): a block whose constraint matrices are all given as
  factors (`LowRankMatrix`) -- or, with `detect_rank_one`, are all rank one,
  LowRankOpt's detection (README.md:148) -- keeps them as factors; otherwise a
  block whose matrices have more than `datasparsity` nonzeros is stored dense
  (its touching matrices contiguous, the BufferedModelForSchur dense path),
  else as COO entries, and a block of diagonal one-hot constraints is
  recognised for the W o W Schur kernel.
* `Solver(lro).solve(**kws)` creates the device solver with the component
  keywords `SolverCore.solve!` forwards (src/moi.jl:53-60), solves, and
  returns `Stats` with the fields src/moi.jl:62-70 sets: status (the
  SolverCore symbol of src/moi.jl:16-27), objective, dual_objective,
  multipliers (y), primal / dual feasibility, iteration count and time;
  `solution()` returns `ShapedSolution(x, [X_i])`.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Any, List, Optional, Sequence

import numpy as np

from .model import BlockConstraints, Model, rank_one_factor
from .solver import MadSDPSolver, Status

# MadNLP status -> SolverCore status symbol (src/moi.jl:16-27)
SOLVERCORE_STATUS = {
    Status.SOLVE_SUCCEEDED: "first_order",
    Status.SOLVED_TO_ACCEPTABLE_LEVEL: "acceptable",
    Status.SEARCH_DIRECTION_BECOMES_TOO_SMALL: "small_step",
    Status.MAXIMUM_ITERATIONS_EXCEEDED: "max_iter",
    Status.MAXIMUM_WALLTIME_EXCEEDED: "max_time",
    Status.INFEASIBLE_PROBLEM_DETECTED: "infeasible",
    Status.ERROR_IN_STEP_COMPUTATION: "exception",
    Status.USER_REQUESTED_STOP: "user",
    Status.INITIAL: "unknown",
    Status.REGULAR: "unknown",
}


def solvercore_status(status) -> str:
    """`_solvercore_status` (src/moi.jl:29)."""
    try:
        return SOLVERCORE_STATUS.get(Status(int(status)), "unknown")
    except ValueError:
        return "unknown"


@dataclass
class LowRankMatrix:
    """A constraint matrix given by its factors, sum_r weights[r] f_r f_r'
    with f_r = factors[:, r] (LowRankOpt's factored matrix types)."""
    factors: Any
    weights: Any

    def toarray(self) -> np.ndarray:
        F = np.asarray(self.factors, dtype=np.float64)
        F = F.reshape(F.shape[0], -1)
        return (F * np.asarray(self.weights, dtype=np.float64).reshape(1, -1)) @ F.T

    def terms(self):
        F = np.asarray(self.factors, dtype=np.float64)
        F = F.reshape(F.shape[0], -1)
        w = np.asarray(self.weights, dtype=np.float64).reshape(-1)
        return [(w[r], F[:, r]) for r in range(w.size)]


def _dense(M) -> np.ndarray:
    if M is None:
        return None
    if hasattr(M, "toarray"):
        return np.asarray(M.toarray(), dtype=np.float64)
    return np.asarray(M, dtype=np.float64)


def _nnz(M) -> int:
    if isinstance(M, LowRankMatrix):
        return int(np.count_nonzero(M.toarray()))
    if hasattr(M, "nnz"):
        M = M.tocoo()
        return int(np.count_nonzero(M.data))
    return int(np.count_nonzero(M))


@dataclass
class LroModel:
    """The fields of `LRO.Model(C, A, b, d_lin, C_lin, dims)`."""
    C: Sequence[Any]
    A: Sequence[Sequence[Any]]       # A[i][j]: block i, constraint j (None = zero)
    b: Any
    d_lin: Any = None
    C_lin: Any = None
    dims: List[int] = field(default_factory=list)


def ingest(lro: LroModel, datasparsity: int = 8, detect_rank_one: bool = True) -> Model:
    """LroModel -> device upload format, with the datasparsity switch of
    `BufferedModelForSchur` (src/options.jl:12-14) and the rank-one
    detection of LowRankOpt deciding each block's storage."""
    b = np.asarray(lro.b, dtype=np.float64).reshape(-1)
    m = b.size
    d_lin = np.zeros(0) if lro.d_lin is None else np.asarray(lro.d_lin, dtype=np.float64).reshape(-1)
    C_lin = (np.zeros((m, d_lin.size)) if lro.C_lin is None
             else np.asarray(_dense(lro.C_lin), dtype=np.float64).reshape(m, d_lin.size))
    Cs, blocks = [], []
    for i, Ci in enumerate(lro.C):
        C = _dense(Ci)
        n = C.shape[0]
        if lro.dims and lro.dims[i] != n:
            raise ValueError(f"block {i}: C is {n} x {n} but dims says {lro.dims[i]}")
        touching = []
        for j in range(m):
            Aij = lro.A[i][j] if j < len(lro.A[i]) else None
            if Aij is None or _nnz(Aij) == 0:
                continue
            touching.append((j, Aij))
        maxnnz = max((_nnz(A) for _, A in touching), default=0)
        explicit = bool(touching) and all(isinstance(A, LowRankMatrix) for _, A in touching)
        terms = None
        if explicit:
            terms = [A.terms() for _, A in touching]
        elif detect_rank_one and touching and maxnnz > datasparsity:
            facs = [A.terms() if isinstance(A, LowRankMatrix) and len(A.terms()) == 1
                    else rank_one_factor(_dense(A)) for _, A in touching]
            facs = [f[0] if isinstance(f, list) else f for f in facs]
            if all(f is not None for f in facs):
                terms = [[f] for f in facs]
        if terms is not None:
            blocks.append(BlockConstraints.lowrank(n, [j for j, _ in touching], terms))
        elif maxnnz > datasparsity:
            mats = np.stack([_dense(A) for _, A in touching]) if touching else np.zeros((0, n, n))
            if not np.allclose(mats, np.swapaxes(mats, 1, 2)):
                raise ValueError(f"block {i}: constraint matrices must be symmetric")
            blocks.append(BlockConstraints("dense", n,
                                           con_idx=np.array([j for j, _ in touching], dtype=np.int64),
                                           mats=np.ascontiguousarray(mats)))
        else:
            con, row, col, val = [], [], [], []
            for j, A in touching:
                if hasattr(A, "tocoo"):
                    c = A.tocoo()
                    r, k, v = c.row, c.col, c.data
                    keep = v != 0
                    r, k, v = r[keep], k[keep], v[keep]
                else:
                    Ad = _dense(A)
                    r, k = np.nonzero(Ad)
                    v = Ad[r, k]
                con.extend([j] * len(v)); row.extend(r); col.extend(k); val.extend(v)
            blocks.append(BlockConstraints("coo", n, con=np.array(con, dtype=np.int64),
                                           row=np.array(row, dtype=np.int64),
                                           col=np.array(col, dtype=np.int64),
                                           val=np.array(val, dtype=np.float64)))
        Cs.append(C)
    return Model(Cs, blocks, b, d_lin, C_lin)


@dataclass
class ShapedSolution:
    """`LRO.ShapedSolution(x, [X_i])` (src/moi.jl:76-78)."""
    x: np.ndarray
    X: List[np.ndarray]


@dataclass
class Stats:
    """The fields of `SolverCore.GenericExecutionStats` that src/moi.jl:62-70 sets."""
    status: str
    objective: float
    dual_objective: float
    multipliers: np.ndarray
    primal_feas: float
    dual_feas: float
    iter: int
    elapsed_time: float
    solver_status: Status


class Solver:
    """`MadSDP.Solver(model)` (src/moi.jl:42-51) over the device solver."""

    def __init__(self, lro: LroModel, datasparsity: int = 8):
        self.lro = lro
        self.model = ingest(lro, datasparsity)
        self.inner: Optional[MadSDPSolver] = None
        self.stats: Optional[Stats] = None

    def solve(self, device: int = 0, **kws) -> Stats:
        """`SolverCore.solve!` (src/moi.jl:53-72): component keywords go to the
        solver, as LRO.Optimizer's raw attributes do."""
        t0 = time.perf_counter()
        self.inner = MadSDPSolver(self.model, device=device, **kws)
        self.inner.solve()
        s = self.inner
        self.stats = Stats(status=solvercore_status(s.status), objective=s.primal_objective(),
                           dual_objective=s.dual_objective(), multipliers=np.array(s.y, copy=True),
                           primal_feas=s.inf_pr, dual_feas=s.inf_du, iter=int(s.iter),
                           elapsed_time=time.perf_counter() - t0, solver_status=s.status)
        return self.stats

    def solution(self) -> ShapedSolution:
        """`MOI.get(::Solver, ::LRO.Solution)` (src/moi.jl:76-78)."""
        s = self.inner
        return ShapedSolution(np.array(s.x, copy=True),
                              [np.array(s.X(i), copy=True) for i in range(self.model.nblocks)])
