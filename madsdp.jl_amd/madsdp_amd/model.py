"""Problem data of a mixed SDP/LP, laid out the way the device library takes it.

The problem is the one `LowRankOpt.Model(C, A, b, d_lin, C_lin, dims)` encodes
for MadSDP (reference `test/runtests.jl:16-31,36-56`):

    min  sum_i <C_i, X_i> + d_lin' x
    s.t. sum_i <A_ij, X_i> + (C_lin x)_j = b_j      j = 1..m
         X_i PSD (n_i x n_i),  x >= 0 (n_s scalars)

Each PSD block keeps its constraint matrices in one of three storages, chosen
per block like the `datasparsity` threshold of `LRO.BufferedModelForSchur`
(reference `src/options.jl:12-14`, `src/solver.jl:98-100`) and LowRankOpt's
rank-one detection (reference `README.md:148`):

* ``coo``: every nonzero entry of every A_ij, both triangles listed, as
  parallel arrays (con, row, col, val).  One-hot diagonal constraints
  (max-cut's A_j = e_j e_j') are the common case.
* ``dense``: the constraints that touch the block, `con_idx` (m_i,), and
  their dense matrices `mats` (m_i, n_i, n_i), symmetric.
* ``lowrank``: A_j = sum_r w_r f_r f_r' -- `con_idx` (m_i,), `rank_ptr`
  (m_i + 1,) the column range of each constraint in `factors` (n_i, R), and
  `weights` (R,).

All arrays are float64 / int64 numpy arrays; matrices are symmetric so row-
and column-major orders coincide.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np


@dataclass
class BlockConstraints:
    kind: str                      # "coo", "dense" or "lowrank"
    n: int
    con: Optional[np.ndarray] = None   # coo: constraint index per entry
    row: Optional[np.ndarray] = None
    col: Optional[np.ndarray] = None
    val: Optional[np.ndarray] = None
    con_idx: Optional[np.ndarray] = None  # dense / lowrank: constraints touching the block
    mats: Optional[np.ndarray] = None     # dense: (m_i, n, n)
    rank_ptr: Optional[np.ndarray] = None  # lowrank: (m_i + 1,) column ranges
    factors: Optional[np.ndarray] = None   # lowrank: (n, R)
    weights: Optional[np.ndarray] = None   # lowrank: (R,)

    @property
    def num_touching(self) -> int:
        if self.kind in ("dense", "lowrank"):
            return int(self.con_idx.size)
        return int(np.unique(self.con).size) if self.con.size else 0

    def dense_mats(self) -> np.ndarray:
        """(m_i, n, n) matrices of a dense or low-rank block (sum_r w_r f_r f_r')."""
        if self.kind == "dense":
            return self.mats
        if self.kind != "lowrank":
            raise ValueError("dense_mats of a coo block")
        out = np.zeros((self.con_idx.size, self.n, self.n))
        for j in range(self.con_idx.size):
            for r in range(int(self.rank_ptr[j]), int(self.rank_ptr[j + 1])):
                f = self.factors[:, r]
                out[j] += self.weights[r] * np.outer(f, f)
        return out

    @staticmethod
    def lowrank(n: int, con_idx, terms) -> "BlockConstraints":
        """`terms[j]` = list of (weight, vector) of constraint `con_idx[j]`."""
        ptr, cols, w = [0], [], []
        for t in terms:
            for wt, f in t:
                cols.append(np.asarray(f, dtype=np.float64).reshape(n))
                w.append(float(wt))
            ptr.append(len(cols))
        F = np.stack(cols, axis=1) if cols else np.zeros((n, 0))
        return BlockConstraints("lowrank", n, con_idx=np.asarray(con_idx, dtype=np.int64),
                                rank_ptr=np.array(ptr, dtype=np.int64),
                                factors=np.asfortranarray(F), weights=np.array(w))

    def is_one_hot(self) -> bool:
        """True when each constraint touching the block is c * e_r e_r'."""
        if self.kind != "coo" or self.con.size == 0:
            return False
        if not np.all(self.row == self.col):
            return False
        return np.unique(self.con).size == self.con.size


@dataclass
class Model:
    C: List[np.ndarray]
    A: List[BlockConstraints]
    b: np.ndarray
    d_lin: np.ndarray
    C_lin: np.ndarray              # (m, n_s) dense
    dims: List[int] = field(default_factory=list)

    def __post_init__(self):
        self.b = np.ascontiguousarray(self.b, dtype=np.float64)
        self.d_lin = np.ascontiguousarray(self.d_lin, dtype=np.float64).reshape(-1)
        m = self.b.size
        self.C_lin = np.ascontiguousarray(self.C_lin, dtype=np.float64).reshape(m, -1)
        self.C = [np.ascontiguousarray(c, dtype=np.float64) for c in self.C]
        if not self.dims:
            self.dims = [c.shape[0] for c in self.C]
        if self.C_lin.shape[1] != self.d_lin.size:
            raise ValueError("C_lin must have one column per scalar variable")
        if len(self.A) != len(self.C):
            raise ValueError("one BlockConstraints per PSD block")

    @property
    def m(self) -> int:
        return int(self.b.size)

    @property
    def n_scalar(self) -> int:
        return int(self.d_lin.size)

    @property
    def nblocks(self) -> int:
        return len(self.C)

    # -- construction helpers -------------------------------------------------
    @staticmethod
    def from_matrices(C: Sequence[np.ndarray], A: Sequence[Sequence[Optional[np.ndarray]]],
                      b, d_lin=None, C_lin=None, datasparsity: int = 8,
                      detect_rank_one: bool = False) -> "Model":
        """Builds a model from `A[i][j]` (block i, constraint j; None = zero),
        as `LRO.Model(C, A, b, d_lin, C_lin, dims)` does.  A block whose
        constraints have more than `datasparsity` nonzeros is stored dense,
        or, with `detect_rank_one`, as factors when every one of them is
        rank one (LowRankOpt's `detect_rank_one`)."""
        b = np.asarray(b, dtype=np.float64).reshape(-1)
        m = b.size
        d_lin = np.zeros(0) if d_lin is None else np.asarray(d_lin, dtype=np.float64)
        C_lin = np.zeros((m, d_lin.size)) if C_lin is None else np.asarray(C_lin, dtype=np.float64)
        blocks = []
        for i, Ci in enumerate(C):
            n = np.asarray(Ci).shape[0]
            mats = [(j, np.asarray(A[i][j], dtype=np.float64)) for j in range(m)
                    if A[i][j] is not None and np.any(np.asarray(A[i][j]) != 0)]
            maxnnz = max((int(np.count_nonzero(M)) for _, M in mats), default=0)
            factors = ([rank_one_factor(M) for _, M in mats]
                       if detect_rank_one and maxnnz > datasparsity else [])
            if factors and all(f is not None for f in factors):
                blocks.append(BlockConstraints.lowrank(n, [j for j, _ in mats],
                                                       [[f] for f in factors]))
            elif maxnnz > datasparsity:
                blocks.append(BlockConstraints(
                    "dense", n,
                    con_idx=np.array([j for j, _ in mats], dtype=np.int64),
                    mats=np.ascontiguousarray(np.stack([M for _, M in mats]) if mats
                                              else np.zeros((0, n, n)))))
            else:
                con, row, col, val = [], [], [], []
                for j, M in mats:
                    r, c = np.nonzero(M)
                    con.extend([j] * r.size); row.extend(r); col.extend(c); val.extend(M[r, c])
                blocks.append(BlockConstraints(
                    "coo", n, con=np.array(con, dtype=np.int64), row=np.array(row, dtype=np.int64),
                    col=np.array(col, dtype=np.int64), val=np.array(val, dtype=np.float64)))
        return Model(list(C), blocks, b, d_lin, C_lin)

    def footprint_bytes(self) -> int:
        """Device bytes the problem data and a PrimalDual+NT iterate need."""
        s = 8 * (self.m * self.m + 8 * self.m + 4 * self.n_scalar + self.m * self.n_scalar)
        for i, n in enumerate(self.dims):
            s += 8 * 24 * n * n
            blk = self.A[i]
            if blk.kind == "dense":
                s += 8 * blk.mats.size + 8 * blk.con_idx.size
                s += 8 * blk.con_idx.size * n * (n + 1) // 2  # svec(G'AG) panel
            elif blk.kind == "lowrank":
                R = blk.weights.size
                s += 8 * (2 * n * R + 2 * R * R + R) + 8 * blk.con_idx.size
            else:
                s += 32 * blk.con.size
        return s


def rank_one_factor(M: np.ndarray, rtol: float = 1e-13):
    """(w, f) with M = w f f' (w = +-1) when the symmetric M is rank one to
    `rtol` of its largest entry, else None."""
    M = np.asarray(M, dtype=np.float64)
    d = np.diag(M)
    p = int(np.argmax(np.abs(d)))
    if d[p] == 0.0:
        return None
    w = 1.0 if d[p] > 0 else -1.0
    f = M[:, p] / np.sqrt(abs(d[p]))
    scale = np.abs(M).max()
    if np.abs(w * np.outer(f, f) - M).max() > rtol * scale:
        return None
    return (w, f)
