"""Seeded synthetic SDP instances of the five workloads.

The graphs and data are drawn with numpy's PCG64 `default_rng(seed)`; they are
this project's own instances (the reference benchmark's `sprand` stream,
`benchmark/maxcut.jl:19-26`, is Julia-specific), bit-identical on every
machine: the derived data (C, b, d) are accumulated in a fixed order with
correctly rounded elementwise operations, never by a BLAS call whose
summation order depends on the host CPU.

* ``maxcut(n, p, seed, weights)`` -- min <-L/4, X>  s.t. diag(X) = 1, X PSD,
  with L the Laplacian of an Erdos-Renyi graph (strict upper triangle drawn,
  mirrored).  Constraints A_j = e_j e_j' (one-hot, stored coo).
* ``dense_sdp(n, m, n_s, seed)`` -- one n x n block with dense symmetric
  N(0,1) constraint matrices, n_s nonnegative scalars with U[0,1] rows,
  b_j = tr(A_j) + sum(a_j), C = sum_j y0_j A_j + I, d = C_lin' y0 + 1, so
  X = I, x = 1 is primal feasible and y0 gives S = I, z = 1.
* ``block_sdp(sizes, m, n_s, seed)`` -- block-diagonal SOS-style problem:
  each constraint touches 1-3 random blocks densely, data built as above.
"""
from __future__ import annotations

import math

import numpy as np

from .model import BlockConstraints, Model


def maxcut_laplacian(n: int, p: float, seed: int, weights: str = "unit") -> np.ndarray:
    rng = np.random.default_rng(seed)
    mask = np.triu(rng.random((n, n)) < p, 1)
    if weights == "unit":
        w = np.ones((n, n))
    elif weights == "pm1":
        w = rng.choice(np.array([-1.0, 1.0]), size=(n, n))
    else:
        raise ValueError(weights)
    Wm = np.where(mask, w, 0.0)
    Wm = Wm + Wm.T
    return np.diag(Wm.sum(axis=1)) - Wm


def maxcut_from_laplacian(L: np.ndarray) -> Model:
    n = L.shape[0]
    j = np.arange(n, dtype=np.int64)
    blk = BlockConstraints("coo", n, con=j.copy(), row=j.copy(), col=j.copy(), val=np.ones(n))
    return Model([-L / 4.0], [blk], np.ones(n), np.zeros(0), np.zeros((n, 0)))


def maxcut(n: int, p: float, seed: int, weights: str = "unit") -> Model:
    return maxcut_from_laplacian(maxcut_laplacian(n, p, seed, weights))


def _sym_normal(rng, k, n):
    """k symmetric n x n matrices: the upper triangle (diagonal included) of
    iid N(0,1) draws, mirrored."""
    A = rng.standard_normal((k, n, n))
    low = np.tril(np.ones((n, n), dtype=bool), -1)
    np.copyto(A, np.swapaxes(A, 1, 2), where=low)
    return A


def _combination(coef, mats):
    """sum_j coef_j mats_j accumulated in index order, one correctly rounded
    multiply and add per entry and term: the same bits on every machine (a
    BLAS product's summation order depends on the host CPU)."""
    out = np.zeros(mats.shape[1:])
    for c, M in zip(coef, mats):
        out += c * M
    return out


def _rhs(mats, Clin):
    """b_j = tr(A_j) + sum(a_j), each exactly rounded (math.fsum)."""
    return np.array([math.fsum(np.concatenate((np.diagonal(A), a))) for A, a in zip(mats, Clin)])


def dense_sdp(n: int, m: int, n_s: int, seed: int) -> Model:
    rng = np.random.default_rng(seed)
    mats = _sym_normal(rng, m, n)
    Clin = rng.random((m, n_s))
    b = _rhs(mats, Clin)
    y0 = rng.standard_normal(m)
    C = _combination(y0, mats) + np.eye(n)
    d = _combination(y0, Clin) + 1.0
    blk = BlockConstraints("dense", n, con_idx=np.arange(m, dtype=np.int64), mats=mats)
    return Model([C], [blk], b, d, Clin)


def block_sdp(sizes, m: int, n_s: int, seed: int) -> Model:
    rng = np.random.default_rng(seed)
    sizes = [int(s) for s in sizes]
    nb = len(sizes)
    touch = [[] for _ in range(nb)]
    for j in range(m):
        k = int(rng.integers(1, 4))
        for i in rng.choice(nb, size=min(k, nb), replace=False):
            touch[int(i)].append(j)
    blocks, Cs = [], []
    b = np.zeros(m)
    y0 = rng.standard_normal(m)
    parts = [[] for _ in range(m)]  # the terms of each b_j, summed exactly below
    for i, n in enumerate(sizes):
        idx = np.array(sorted(touch[i]), dtype=np.int64)
        mats = _sym_normal(rng, idx.size, n)
        for j, A in zip(idx, mats):
            parts[j].append(np.diagonal(A))
        Cs.append(_combination(y0[idx], mats) + np.eye(n))
        blocks.append(BlockConstraints("dense", n, con_idx=idx, mats=mats))
    Clin = rng.random((m, n_s))
    for j in range(m):
        b[j] = math.fsum(np.concatenate(parts[j] + [Clin[j]]))
    d = _combination(y0, Clin) + 1.0
    return Model(Cs, blocks, b, d, Clin)


def lowrank_sdp(n: int, m: int, n_s: int, seed: int, rank: int = 1, signed: bool = False) -> Model:
    """Constraints A_j = sum_{r<rank} w_jr f_jr f_jr' (f ~ N(0, 1/n), w = 1 or
    +-1), stored as factors; X = I, x = 1 and S = I, z = 1 are strictly
    feasible (b_j = sum_r w_jr |f_jr|^2 + sum(C_lin[j]), C = I + sum_j y0_j A_j)."""
    rng = np.random.default_rng(seed)
    F = rng.standard_normal((m, rank, n)) / math.sqrt(n)
    w = np.where(rng.random((m, rank)) < 0.5, -1.0, 1.0) if signed else np.ones((m, rank))
    Clin = rng.random((m, n_s))
    y0 = rng.standard_normal(m)
    C = np.eye(n)
    b = np.zeros(m)
    for j in range(m):
        for r in range(rank):
            C += (y0[j] * w[j, r]) * np.outer(F[j, r], F[j, r])
        b[j] = math.fsum(np.concatenate([w[j, r] * F[j, r] * F[j, r] for r in range(rank)]
                                        + [Clin[j]]))
    d = _combination(y0, Clin) + 1.0
    blk = BlockConstraints.lowrank(n, np.arange(m), [[(w[j, r], F[j, r]) for r in range(rank)]
                                                     for j in range(m)])
    return Model([C], [blk], b, d, Clin)


def block_sizes(count: int, lo: int, hi: int, seed: int):
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.integers(lo, hi + 1, size=count)]


# -- the five workloads -------------------------------------------------------
def maxcut_100(seed: int = 0) -> Model:
    """Max-cut, G(100, 0.1), unit weights: one 100x100 block, m = 100."""
    return maxcut(100, 0.1, seed, "unit")


def maxcut_3000(seed: int = 1) -> Model:
    """Max-cut, G(3000, 0.01), weights +-1: one 3000x3000 block, m = 3000."""
    return maxcut(3000, 0.01, seed, "pm1")


def dense_250_m8000(seed: int = 2) -> Model:
    """Dense 250x250 block + 400 scalars, m = 8000."""
    return dense_sdp(250, 8000, 400, seed)


def blocks_96_m6000(seed: int = 3) -> Model:
    """96 blocks of 8..120, 200 scalars, m = 6000."""
    return block_sdp(block_sizes(96, 8, 120, seed + 1000), 6000, 200, seed)


def serving_stream(count: int = 4096, seed: int = 4, rate: float = 50.0):
    """(arrival_time, kind, params) for a Poisson stream; half max-cut with
    n ~ U{20..600}, p = 0.1, half dense SDPs with n ~ U{16..200},
    m ~ U{50..3000} and n_s = n scalars."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=count)
    t = np.cumsum(gaps)
    out = []
    for k in range(count):
        s = int(rng.integers(0, 2**31 - 1))
        if rng.random() < 0.5:
            out.append((float(t[k]), "maxcut", dict(n=int(rng.integers(20, 601)), p=0.1, seed=s)))
        else:
            n = int(rng.integers(16, 201))
            out.append((float(t[k]), "dense", dict(n=n, m=int(rng.integers(50, 3001)), n_s=n, seed=s)))
    return out


def build_stream_instance(kind: str, params: dict) -> Model:
    if kind == "maxcut":
        return maxcut(params["n"], params["p"], params["seed"], "unit")
    return dense_sdp(params["n"], params["m"], params["n_s"], params["seed"])
