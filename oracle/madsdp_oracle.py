"""FP64 CPU restatement of MadSDP's interior-point methods -- the ORACLE.

TEST INFRASTRUCTURE ONLY.  Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import this module, and only as the
checker / the timed CPU baseline.  The product (the HIP library and its host
driver) never imports, calls or falls back to it.

It restates, line by line, the algorithm of the reference (`src/*.jl`) with
numpy/scipy LAPACK: potrf/potri/trsm/gemm/syevd and a dense Householder QR
in place of SuiteSparse SPQR.  The LowRankOpt callbacks that the reference
calls but does not vendor are restated from their contracts:

    A(X)_j    = sum_i <A_ij, X_i>                       (LRO.add_jprod!)
    A*(y)_i   = sum_j y_j A_ij                          (LRO.unsafe_jtprod)
    H_jk      = sum_i <A_ij, L_i A_ik R_i> + sum_l C_lin[j,l] x_l z_l^-1 C_lin[k,l]
                                                        (LRO.schur_complement!)
    lambda_min(L^-1 D L^-T) exactly (eigvalsh)          (LRO.lanczos_eigmin!)

How each part is pinned (what would catch a wrong sign, index or term):
* end-to-end known answers of the reference tests (`test/runtests.jl:60-212`,
  `test/dual_runtests.jl:33-141`, `test/qr_schur.jl:68-139`);
* the Newton-equation invariants of `test/qr_schur.jl:16-66` on its 3x3
  example (both Schur solvers), which pin rhs!/directions!/E/correction;
* closed forms: K_n max-cut = n^2/4 with X = (nI-J)/(n-1), bipartite graphs
  cut every edge, trace-1 SDPs give lambda_min(C);
* unit known answers: Cholesky flush (`test/blocks.jl:24-34`), the automatic
  identity scale sqrt(15000) (`test/identity_scale.jl:32-42`), termination
  semantics (`test/termination.jl`);
* brute force: H_jk = tr(A_j W A_k W) by definition, W S W = X, lambda_min
  by dense eig, Schur regularisation ladder;
* independent mathematics (tests/test_oracle_pins.py): the NT correction from
  the Lyapunov equation in the W^1/2 scaling (scipy solve_sylvester), the
  commuting closed form -dX dS / S of both corrections, whole Mehrotra
  iterations through the unreduced KKT matrix with step lengths by bisection
  on eigvalsh(X + a dX), EigenStep against the brute-force largest step,
  DualCentering's alpha by brentq on a log-det form of xi and its predictor
  step with H_jk = tr(A_j S^-1 A_k S^-1); planted mistakes (correction halved
  or zeroed, alpha = tau, DC alpha x 0.7, xi x 2) fail them.
Parity unpinned: LRO.lanczos_eigmin!'s exact output (the restatement uses the
exact lambda_min, which Lanczos approximates).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Optional

import numpy as np
import scipy.linalg as sla

EPS = float(np.finfo(np.float64).eps)
SQRT2 = math.sqrt(2.0)


class Status(IntEnum):
    """MadNLP.Status codes used by MadSDP (`src/moi.jl:16-27`)."""
    SOLVE_SUCCEEDED = 1
    SOLVED_TO_ACCEPTABLE_LEVEL = 2
    SEARCH_DIRECTION_BECOMES_TOO_SMALL = 3
    DIVERGING_ITERATES = 4
    INFEASIBLE_PROBLEM_DETECTED = 5
    MAXIMUM_ITERATIONS_EXCEEDED = 6
    MAXIMUM_WALLTIME_EXCEEDED = 7
    INITIAL = 11
    REGULAR = 12
    ERROR_IN_STEP_COMPUTATION = -3
    USER_REQUESTED_STOP = -5


# --------------------------------------------------------------------------
# Components (`src/components.jl`)
# --------------------------------------------------------------------------
class NT:                       # components.jl:96
    pass


class HKM:                      # components.jl:110
    pass


@dataclass
class EigenStep:                # components.jl:132-135
    tau: float = 0.95


@dataclass
class CholeskyBacktrack:        # components.jl:146-150
    shrink: float = 0.8
    boundary: float = 0.95


@dataclass
class Mehrotra:                 # components.jl:169-172
    sigma_exponent: float = 3.0


@dataclass
class AdaptiveMu:               # components.jl:188-192
    initial_ratio: float = 0.25
    ratio: float = 0.5


class ScaledIdentity:           # components.jl:208-217
    def __init__(self, scale=None):
        if scale is not None:
            scale = float(scale)
            if not math.isfinite(scale) or scale <= 0:
                raise ValueError("The identity scale must be finite and positive.")
        self.scale = scale


@dataclass
class DualShift:                # components.jl:228-232
    initial_shift: float = 1.0
    max_shift: float = 1e10


class InitialPoint:             # components.jl:243-251
    def __init__(self, X, y, x=None):
        self.X = [np.array(Xi, dtype=np.float64) for Xi in X]
        self.y = np.array(y, dtype=np.float64)
        self.x = None if x is None else np.array(x, dtype=np.float64)


class DenseCholesky:            # components.jl:273
    pass


class MadNLPSchur:              # components.jl:275-283 (LDL' Bunch-Kaufman, kkt.jl:7-32)
    pass


class QRSchur:                  # components.jl:300
    pass


@dataclass
class PrimalDual:               # components.jl:30-36
    direction: object = field(default_factory=NT)
    step: object = field(default_factory=EigenStep)
    centering: object = field(default_factory=Mehrotra)


@dataclass
class DualCentering:            # components.jl:57-65
    beta: float = 0.2
    delta: float = 1.0
    A: Optional[float] = 2.0
    centering_max_iter: int = 50


@dataclass
class Options:                  # src/options.jl:3-15
    tol: float = 1e-7
    max_iter: int = 100
    max_wall_time: float = 3600.0
    mu_min: float = 1e-14
    regularize_schur: float = 1e-12
    max_schur_reg: int = 8


# --------------------------------------------------------------------------
# LowRankOpt callbacks, restated from their contracts
# --------------------------------------------------------------------------
class _BlockOps:
    """A(X), A*(y) and the Schur contribution of one PSD block."""

    def __init__(self, blk, m: int):
        if blk.kind == "lowrank":
            # A_j = sum_r w_r f_r f_r' taken by its definition, as dense matrices
            # (LowRankOpt's factored matrices are ordinary A_ij to the algorithm)
            blk = type(blk)("dense", blk.n, con_idx=blk.con_idx, mats=blk.dense_mats())
        self.blk = blk
        self.n = blk.n
        self.m = m
        if blk.kind == "dense":
            self.idx = np.asarray(blk.con_idx, dtype=np.int64)
            self.flat = blk.mats.reshape(self.idx.size, self.n * self.n)
            self.empty = self.idx.size == 0
            self.one_hot = False
        else:
            self.empty = blk.con.size == 0
            self.one_hot = blk.is_one_hot()
            if self.one_hot:
                self.oh_con = blk.con
                self.oh_row = blk.row
                self.oh_val = blk.val
            else:
                import scipy.sparse as sp
                self.sflat = sp.csr_matrix(
                    (blk.val, (blk.con, blk.row * self.n + blk.col)), shape=(m, self.n * self.n))
                self.cons = np.unique(blk.con)

    def jprod(self, X, out):
        """out += A_i(X)   (LRO.add_jprod!, reference src/step.jl:19)."""
        if self.empty:
            return out
        if self.blk.kind == "dense":
            out[self.idx] += self.flat @ X.reshape(-1)
        elif self.one_hot:
            np.add.at(out, self.oh_con, self.oh_val * X[self.oh_row, self.oh_row])
        else:
            out += self.sflat @ X.reshape(-1)
        return out

    def jtprod(self, y):
        """A*_i(y) = sum_j y_j A_ij; None means the FillArrays.Zeros of a block
        no constraint touches (LRO.unsafe_jtprod, reference src/step.jl:30)."""
        if self.empty:
            return None
        n = self.n
        if self.blk.kind == "dense":
            return np.tensordot(y[self.idx], self.blk.mats, axes=1)
        if self.one_hot:
            out = np.zeros((n, n))
            np.add.at(out, (self.oh_row, self.oh_row), self.oh_val * y[self.oh_con])
            return out
        return (self.sflat.T @ y).reshape(n, n)

    def schur(self, L, R, H):
        """H_jk += <A_j, L A_k R> for the constraints touching the block."""
        if self.empty:
            return H
        if self.blk.kind == "dense":
            P = np.matmul(np.matmul(L, self.blk.mats), R)
            Hb = self.flat @ P.reshape(self.idx.size, -1).T
            H[np.ix_(self.idx, self.idx)] += Hb
        elif self.one_hot:
            r = self.oh_row
            v = self.oh_val
            Hb = (v[:, None] * v[None, :]) * L[np.ix_(r, r)] * R[np.ix_(r, r)].T
            H[np.ix_(self.oh_con, self.oh_con)] += Hb
        else:
            n = self.n
            blk = self.blk
            for k in self.cons:
                sel = blk.con == k
                Ak = np.zeros((n, n))
                np.add.at(Ak, (blk.row[sel], blk.col[sel]), blk.val[sel])
                Pk = L @ Ak @ R
                H[:, k] += self.sflat @ Pk.reshape(-1)
        return H

    def dense_matrix(self, j):
        """A_ij as a dense matrix (zero if the constraint does not touch it)."""
        n = self.n
        out = np.zeros((n, n))
        if self.empty:
            return out
        if self.blk.kind == "dense":
            hit = np.nonzero(self.idx == j)[0]
            if hit.size:
                out[:] = self.blk.mats[hit[0]]
            return out
        blk = self.blk
        sel = blk.con == j
        np.add.at(out, (blk.row[sel], blk.col[sel]), blk.val[sel])
        return out


# --------------------------------------------------------------------------
# Block kernels (`src/blocks.jl`)
# --------------------------------------------------------------------------
def symmetrize(A):
    """A = (A + A')/2 (blocks.jl:77-84); the diagonal is left untouched."""
    return (A + A.T) / 2


def cholesky_factor(M):
    """Lower Cholesky factor of (M+M')/2 with the flush of entries below
    eps^2 * max|L_ii| (blocks.jl:94-113).  None if not positive definite."""
    S = symmetrize(M)
    if not np.all(np.isfinite(S)):
        return None
    L, info = sla.lapack.dpotrf(S, lower=1, clean=1, overwrite_a=0)
    if info != 0:
        return None
    L = np.tril(L)
    tol = EPS ** 2 * (np.max(np.abs(np.diag(L))) if L.size else 0.0)
    low = np.tril(np.ones_like(L, dtype=bool), -1)
    L[low & (np.abs(L) < tol)] = 0.0
    return L


def inverse_from_cholesky(L):
    """S^-1 from its Cholesky factor with potri, lower mirrored (blocks.jl:116-123)."""
    Si, info = sla.lapack.dpotri(L, lower=1)
    Si = np.tril(Si)
    return Si + np.tril(Si, -1).T


def is_add_mul_posdef(X, alpha, dX):
    """Whether X + alpha dX has a Cholesky factorization (blocks.jl:133-137)."""
    B = X + alpha * dX
    if not np.all(np.isfinite(B)):
        return False
    _, info = sla.lapack.dpotrf(np.tril(B) + np.tril(B, -1).T, lower=1)
    return info == 0


def _trsm_left_lower(L, B):
    return sla.solve_triangular(L, B, lower=True, check_finite=False)


def _trsm_right_lower(B, L):
    """B L^-1."""
    return sla.solve_triangular(L, B.T, lower=True, trans="T", check_finite=False).T


# --------------------------------------------------------------------------
# Solver state
# --------------------------------------------------------------------------
class _Block:
    def __init__(self, n):
        z = lambda: np.zeros((n, n))
        self.n = n
        self.X, self.S, self.Rd, self.Ls, self.Si = z(), z(), z(), z(), z()
        self.delX, self.delS, self.Lx, self.corr = z(), z(), z(), z()
        # NT workspace (directions.jl:16-43)
        self.W, self.G, self.Gi, self.D, self.RNT = z(), z(), z(), np.zeros(n), z()
        self.ntLx = z()
        # DualCentering workspace (dual_centering.jl:48-55)
        self.Shat, self.Dc, self.eigvals = z(), z(), np.zeros(n)


@dataclass
class IterRecord:
    iter: int
    pobj: float
    dobj: float
    inf_pr: float
    inf_du: float
    gap: float
    mu: float
    alpha: float
    beta: float
    sigma: float


class Solver:
    """`MadSDPSolver` (reference src/solver.jl:7-160)."""

    def __init__(self, model, direction=None, step=None, centering=None, algorithm=None,
                 initial_point=None, schur=None, **options):
        direction = NT() if direction is None else direction
        step = EigenStep() if step is None else step
        centering = Mehrotra() if centering is None else centering
        if algorithm is None:
            algorithm = PrimalDual(direction, step, centering)
        if initial_point is None:
            initial_point = DualShift() if isinstance(algorithm, DualCentering) else ScaledIdentity()
        schur = DenseCholesky() if schur is None else schur
        self.opt = Options(**options)
        self.model = model
        self.algorithm = algorithm
        self.initial_point = initial_point
        self.schur = schur
        nb, ns = model.nblocks, model.n_scalar
        if nb == 0 and ns == 0:
            raise RuntimeError("MadSDP: model has no variables.")
        if isinstance(algorithm, DualCentering):          # dual_centering.jl:68-76
            if ns > 0:
                raise RuntimeError("MadSDP: `DualCentering` does not support scalar (LP) blocks yet.")
            if nb == 0:
                raise RuntimeError("MadSDP: `DualCentering` needs at least one PSD block.")
        if isinstance(schur, QRSchur):                     # qr_schur.jl:12-22
            if not (isinstance(algorithm, PrimalDual) and isinstance(algorithm.direction, NT)):
                raise ValueError("QRSchur requires PrimalDual with the NT direction.")
        m = model.m
        self.m, self.nblocks, self.n_scalar = m, nb, ns
        self.ops = [_BlockOps(model.A[i], m) for i in range(nb)]
        self.blocks = [_Block(n) for n in model.dims]
        self.y, self.dely, self.Rp = np.zeros(m), np.zeros(m), np.zeros(m)
        self.x, self.z = np.zeros(ns), np.zeros(ns)
        self.delx, self.delz, self.Rd_lin = np.zeros(ns), np.zeros(ns), np.zeros(ns)
        self.zinv, self.corr_lin = np.zeros(ns), np.zeros(ns)
        self.rd_nonzero = np.ones(nb, dtype=bool)
        self.cnorm = np.array([np.max(np.abs(C)) if C.size else 0.0 for C in model.C])
        self.mu = self.sigma = self.alpha = self.beta = 0.0
        self.inf_pr = self.inf_du = self.inf_compl = self.obj_val = 0.0
        self.iter = 0
        self.status = Status.INITIAL
        self.start_time = time.time()
        self.trace: List[IterRecord] = []
        # DualCentering state (dual_centering.jl:12-41)
        self.dc_d, self.dc_g, self.dc_grad, self.dc_dy, self.dc_yhat = (np.zeros(m) for _ in range(5))
        self.dc_nu, self.dc_A, self.dc_t, self.dc_lambda = 0, 0.0, 0.0, 0.0
        self.n_predictor = self.n_newton = 0
        # Schur workspace
        self.H = np.zeros((m, m))
        self.cho = None
        self.qr = None

    # -- objectives (solver.jl:164-190) ---------------------------------------
    def primal_objective(self):
        s = 0.0
        for C, b in zip(self.model.C, self.blocks):
            s += float(np.vdot(C, b.X))
        if self.n_scalar > 0:
            s += float(self.model.d_lin @ self.x)
        return s

    def dual_objective(self):
        return float(self.model.b @ self.y)

    def barrier_dim(self):
        return sum(b.n for b in self.blocks) + self.n_scalar

    def total_complementarity(self):
        s = sum(float(np.vdot(b.X, b.S)) for b in self.blocks)
        if self.n_scalar > 0:
            s += float(self.x @ self.z)
        return s

    # -- slacks and initial points (solver.jl:192-293) ------------------------
    def _slack(self, i, y):
        out = self.model.C[i].copy()
        AT = self.ops[i].jtprod(y)
        if AT is not None:
            out -= AT
        return symmetrize(out)

    def _set_dual_slacks(self):
        for i, b in enumerate(self.blocks):
            b.S = self._slack(i, self.y)
        if self.n_scalar > 0:
            self.z = self.model.d_lin - self.model.C_lin.T @ self.y

    def _initial_point(self):
        p = self.initial_point
        model = self.model
        if isinstance(p, ScaledIdentity):                  # solver.jl:214-238
            bscale = np.max(np.abs(model.b)) if model.m else 0.0
            cscale = np.max(self.cnorm) if self.cnorm.size else 0.0
            if self.n_scalar > 0:
                cscale = max(cscale, np.max(np.abs(model.d_lin)))
            if p.scale is None:
                beta = max(10.0, math.sqrt(max(bscale, 1.0) * max(cscale, 1.0))
                           * math.sqrt(float(self.barrier_dim())))
            else:
                beta = p.scale
            for b in self.blocks:
                b.X = beta * np.eye(b.n)
                b.S = beta * np.eye(b.n)
            self.y[:] = 0.0
            self.x[:] = beta
            self.z[:] = beta
        elif isinstance(p, InitialPoint):                  # solver.jl:240-252
            self.y = p.y.copy()
            for i, b in enumerate(self.blocks):
                b.X = p.X[i].copy()
            self.x = np.ones(self.n_scalar) if p.x is None else p.x.copy()
            self._set_dual_slacks()
        elif isinstance(p, DualShift):                     # solver.jl:260-284
            for b in self.blocks:
                b.X = np.eye(b.n)
            self.x[:] = 1.0
            rho = 0.0
            while rho <= p.max_shift:
                self.y = np.full(self.m, -rho)
                self._set_dual_slacks()
                if self._dual_interior():
                    if rho != 0.0:
                        self.y = np.full(self.m, -2 * rho)
                        self._set_dual_slacks()
                    return
                rho = p.initial_shift if rho == 0.0 else 2 * rho
            raise RuntimeError("MadSDP: could not find a dual interior point `y = -rho 1` "
                               f"with rho <= {p.max_shift}; use `InitialPoint` to give one.")
        else:
            raise TypeError(p)

    def _dual_interior(self):
        for b in self.blocks:
            L = cholesky_factor(b.S)
            if L is None:
                return False
            b.Ls = L
        return bool(np.all(self.z > 0))

    def initialize(self):                                  # solver.jl:286-293
        self._initial_point()
        self.iter = 0
        self.status = Status.REGULAR
        self.start_time = time.time()
        if isinstance(self.algorithm, PrimalDual):
            self.update_residuals()                        # primal_dual.jl:21-25
            c = self.algorithm.centering
            nu = max(self.barrier_dim(), 1)
            if isinstance(c, Mehrotra):
                self.mu = self.total_complementarity() / nu
            else:
                self.mu = c.initial_ratio * self.total_complementarity() / nu
        else:
            self._dc_initialize()

    # -- Schur complement (solver.jl:295-314, schur_solvers.jl) --------------
    def _scaling(self, b):
        alg = self.algorithm
        if isinstance(alg, DualCentering):
            return b.Si, b.Si
        if isinstance(alg.direction, NT):
            return b.W, b.W
        return b.Si, b.X

    def assemble_schur(self, reg):
        H = np.zeros((self.m, self.m))
        for ops, b in zip(self.ops, self.blocks):
            L, R = self._scaling(b)
            ops.schur(L, R, H)
        if self.n_scalar > 0:
            # scalar scalings x and 1 ./ z (solver.jl:30-31, 116-118)
            Cl = self.model.C_lin
            H += (Cl * (self.x * self.zinv)) @ Cl.T
        H = np.tril(H) + np.tril(H, -1).T    # only the lower triangle is used
        if reg > 0:
            H[np.diag_indices_from(H)] += reg
        return H

    def factorize_schur(self):
        if isinstance(self.schur, QRSchur):
            return self._qr_factorize()
        opt = self.opt
        reg = opt.regularize_schur
        if isinstance(self.schur, MadNLPSchur):
            # MadNLP's LAPACK solver in BUNCHKAUFMAN mode (kkt.jl:15-25): sytrf on
            # the lower triangle; an exactly zero pivot (info > 0) is the
            # SingularException schur_solvers.jl:27-37 turns into `false`
            for _ in range(opt.max_schur_reg):
                H = self.assemble_schur(reg)
                ldu, piv, info = sla.lapack.dsytrf(H, lower=1)
                if info == 0 and np.all(np.isfinite(ldu)):
                    self.H = H
                    self.cho = (ldu, piv)
                    return True
                reg = 1e-12 if reg == 0 else 10 * reg
            return False
        for _ in range(opt.max_schur_reg):
            H = self.assemble_schur(reg)
            Lh, info = sla.lapack.dpotrf(H, lower=1, clean=1)
            if info == 0 and np.all(np.isfinite(Lh)):
                self.H = H
                self.cho = Lh
                return True
            reg = 1e-12 if reg == 0 else 10 * reg
        return False

    def schur_solve(self, rhs):
        if isinstance(self.schur, MadNLPSchur):           # sytrs (kkt.jl:27-32)
            ldu, piv = self.cho
            x, info = sla.lapack.dsytrs(ldu, piv, rhs, lower=1)
            return x
        L = self.cho
        t = sla.solve_triangular(L, rhs, lower=True, check_finite=False)
        return sla.solve_triangular(L, t, lower=True, trans="T", check_finite=False)

    # -- PrimalDual components (step.jl, directions.jl, primal_dual.jl) -------
    def update_residuals(self):                            # step.jl:147-182
        model = self.model
        Rp = np.zeros(self.m)
        for ops, b in zip(self.ops, self.blocks):
            ops.jprod(b.X, Rp)
        if self.n_scalar > 0:
            Rp += model.C_lin @ self.x
        self.Rp = model.b - Rp
        for i, (ops, b) in enumerate(zip(self.ops, self.blocks)):
            Rd = model.C[i] - b.S
            AT = ops.jtprod(self.y)
            if AT is not None:
                Rd = Rd - AT
            b.Rd = Rd
            thr = 100 * EPS * max(1.0, self.cnorm[i])
            self.rd_nonzero[i] = (np.max(np.abs(Rd)) if Rd.size else 0.0) > thr
        if self.n_scalar > 0:
            self.Rd_lin = model.d_lin - self.z - model.C_lin.T @ self.y

    def _prepare_direction(self, b):
        if isinstance(self.algorithm.direction, HKM):
            return True
        Lx = cholesky_factor(b.X)                          # directions.jl:45-75
        if Lx is None:
            return False
        b.ntLx = Lx
        M = b.Ls.T @ Lx
        N = symmetrize(M.T @ M)
        lam, V = np.linalg.eigh(N)
        if not np.all(lam > 0):
            return False
        b.D = np.sqrt(lam)
        sc = np.sqrt(b.D)
        G = (Lx @ V) / sc[None, :]
        Gi = V.T * sc[:, None]
        Gi = _trsm_right_lower(Gi, Lx)
        b.G, b.Gi = G, Gi
        b.W = symmetrize(G @ G.T)
        return True

    def apply_E(self, b, M):
        if M is None:
            return np.zeros((b.n, b.n))
        if isinstance(self.algorithm.direction, NT):       # directions.jl:80-84
            return symmetrize((b.W @ M) @ b.W)
        return symmetrize((b.X @ M) @ b.Si)                # directions.jl:116-120

    def _block_correction(self, b):
        if isinstance(self.algorithm.direction, NT):       # directions.jl:90-100
            tmp = b.Gi @ b.delX
            RNT = tmp @ b.delS
            Mq = RNT @ b.G
            RNT = -(Mq + Mq.T) / (b.D[:, None] + b.D[None, :])
            b.RNT = RNT
            return symmetrize((b.G @ RNT) @ b.G.T)
        return symmetrize(-((b.delX @ b.delS) @ b.Si))    # directions.jl:123-127

    def prepare(self):                                     # primal_dual.jl:29-37
        step = self.algorithm.step
        for b in self.blocks:
            Ls = cholesky_factor(b.S)
            if Ls is None:
                return False
            b.Ls = Ls
            b.Si = inverse_from_cholesky(Ls)
            if not self._prepare_direction(b):
                return False
            if isinstance(step, EigenStep):                # step.jl:177
                Lx = cholesky_factor(b.X)
                if Lx is None:
                    return False
                b.Lx = Lx
        self.zinv = 1.0 / self.z
        return True

    def rhs(self, mu, corrected):                          # step.jl:198-224
        h = self.Rp.copy()
        for i, (ops, b) in enumerate(zip(self.ops, self.blocks)):
            buf = b.X - mu * b.Si
            if self.rd_nonzero[i]:
                buf = buf + self.apply_E(b, b.Rd)
            if corrected:
                buf = buf - b.corr
            ops.jprod(buf, h)
        if self.n_scalar > 0:
            x, zi = self.x, self.zinv
            v = x - mu * zi + x * zi * self.Rd_lin
            if corrected:
                v = v - self.corr_lin
            h += self.model.C_lin @ v
        return h

    def directions(self, dely, mu, corrected):             # step.jl:229-259
        for i, (ops, b) in enumerate(zip(self.ops, self.blocks)):
            AT = ops.jtprod(dely)
            dS = b.Rd.copy() if AT is None else b.Rd - AT
            b.delS = symmetrize(dS)
            dX = mu * b.Si - b.X + self.apply_E(b, AT)
            if self.rd_nonzero[i]:
                dX = dX - self.apply_E(b, b.Rd)
            if corrected:
                dX = dX + b.corr
            b.delX = symmetrize(dX)
        if self.n_scalar > 0:
            self.delz = self.Rd_lin - self.model.C_lin.T @ dely
            x, zi = self.x, self.zinv
            self.delx = mu * zi - x - x * zi * self.delz
            if corrected:
                self.delx = self.delx + self.corr_lin

    def solve_directions(self, mu, corrected):             # primal_dual.jl:44-54
        if isinstance(self.schur, QRSchur):
            return self._qr_solve_directions(mu, corrected)
        h = self.rhs(mu, corrected)
        self.dely = self.schur_solve(h)
        self.directions(self.dely, mu, corrected)

    def correction(self):                                  # step.jl:263-271
        for b in self.blocks:
            b.corr = self._block_correction(b)
        self.corr_lin = -self.delx * self.delz * self.zinv

    # -- step lengths (step.jl:139-228) --------------------------------------
    @staticmethod
    def max_scalar_step(x, dx):
        neg = dx < 0
        if not np.any(neg):
            return math.inf
        return float(np.min(-x[neg] / dx[neg]))

    @staticmethod
    def lambda_min_congruence(L, Delta):
        """lambda_min(L^-1 Delta L^-T), Delta read from its upper triangle."""
        D = np.triu(Delta) + np.triu(Delta, 1).T
        T = _trsm_left_lower(L, D)
        T = _trsm_left_lower(L, T.T).T
        T = symmetrize(T)
        return float(np.linalg.eigvalsh(T)[0])

    @classmethod
    def eigen_step(cls, L, Delta, tau):                    # step.jl:180-184
        lam = cls.lambda_min_congruence(L, Delta)
        return 1.0 if lam > -math.sqrt(EPS) else min(1.0, -tau / lam)

    def _backtrack(self, s, mats, dmats, x, dx):           # step.jl:200-216
        a = 1.0
        for X, dX in zip(mats, dmats):
            while not is_add_mul_posdef(X, a, dX):
                a *= s.shrink
                if a < EPS:
                    return 0.0
        if a != 1.0:
            a *= s.boundary
        if x.size > 0:
            a = min(a, s.boundary * self.max_scalar_step(x, dx))
        return a

    def step_lengths(self):
        s = self.algorithm.step
        if isinstance(s, EigenStep):                       # step.jl:186-198
            a = bt = 1.0
            for b in self.blocks:
                a = min(a, self.eigen_step(b.Lx, b.delX, s.tau))
                bt = min(bt, self.eigen_step(b.Ls, b.delS, s.tau))
            if self.n_scalar > 0:
                a = min(a, s.tau * self.max_scalar_step(self.x, self.delx))
                bt = min(bt, s.tau * self.max_scalar_step(self.z, self.delz))
            return a, bt
        a = self._backtrack(s, [b.X for b in self.blocks], [b.delX for b in self.blocks],
                            self.x, self.delx)
        bt = self._backtrack(s, [b.S for b in self.blocks], [b.delS for b in self.blocks],
                             self.z, self.delz)
        return a, bt

    def affine_complementarity(self, a, bt):               # step.jl:231-243
        s = 0.0
        for b in self.blocks:
            s += float(np.vdot(b.X, b.S))
            s += bt * float(np.vdot(b.X, b.delS))
            s += a * float(np.vdot(b.delX, b.S))
            s += a * bt * float(np.vdot(b.delX, b.delS))
        if self.n_scalar > 0:
            s += float(np.sum((self.x + a * self.delx) * (self.z + bt * self.delz)))
        return s / max(self.barrier_dim(), 1)

    def take_step(self, a, bt):                            # primal_dual.jl:56-72
        self.alpha, self.beta = a, bt
        for b in self.blocks:
            b.X = symmetrize(b.X + a * b.delX)
            b.S = symmetrize(b.S + bt * b.delS)
        self.y = self.y + bt * self.dely
        if self.n_scalar > 0:
            self.x = self.x + a * self.delx
            self.z = self.z + bt * self.delz
        self.iter += 1
        self.update_residuals()

    def _start_iteration(self):                            # primal_dual.jl:74-80
        if not self.prepare() or not self.factorize_schur():
            self.status = Status.ERROR_IN_STEP_COMPUTATION
            return False
        return True

    def iterate(self):
        alg = self.algorithm
        if isinstance(alg, DualCentering):
            return self._dc_iterate()
        c = alg.centering
        if isinstance(c, Mehrotra):                        # primal_dual.jl:85-101
            if not self._start_iteration():
                return
            mu = self.total_complementarity() / max(self.barrier_dim(), 1)
            self.mu = mu
            self.solve_directions(0.0, False)
            a_aff, b_aff = self.step_lengths()
            mu_aff = self.affine_complementarity(a_aff, b_aff)
            sigma = min(max((mu_aff / max(mu, EPS)) ** c.sigma_exponent, 1e-8), 1.0)
            self.sigma = sigma
            self.correction()
            self.solve_directions(sigma * mu, True)
            a, bt = self.step_lengths()
            self.take_step(a, bt)
        else:                                              # primal_dual.jl:104-121
            if not self._start_iteration():
                return
            self.solve_directions(self.mu, False)
            a, bt = self.step_lengths()
            self.take_step(a, bt)
            mu = c.ratio * self.total_complementarity() / max(self.barrier_dim(), 1)
            if min(a, bt) < 0.5:
                mu *= 1.5
            if a + bt > 1.6:
                mu *= 0.75
            if a + bt > 1.9:
                mu /= 1 + 0.1 * self.iter
            self.mu = mu

    # -- termination (primal_dual.jl:123-135, solver.jl:319-347) -------------
    def _update_infeasibilities(self):
        bn = np.max(np.abs(self.model.b)) if self.m else 0.0
        self.inf_pr = (np.max(np.abs(self.Rp)) if self.m else 0.0) / max(1.0, bn)
        rd = 0.0
        for b in self.blocks:
            if b.Rd.size:
                rd = max(rd, float(np.max(np.abs(b.Rd))))
        cn = float(np.max(self.cnorm)) if self.cnorm.size else 0.0
        if self.n_scalar > 0:
            rd = max(rd, float(np.max(np.abs(self.Rd_lin))))
            cn = max(cn, float(np.max(np.abs(self.model.d_lin))))
        self.inf_du = rd / max(1.0, cn)

    def _update_status(self, converged):
        if converged:
            self.status = Status.SOLVE_SUCCEEDED
        elif self.iter >= self.opt.max_iter:
            self.status = Status.MAXIMUM_ITERATIONS_EXCEEDED
        elif time.time() - self.start_time >= self.opt.max_wall_time:
            self.status = Status.MAXIMUM_WALLTIME_EXCEEDED

    def update_termination(self):
        pobj = self.primal_objective()
        dobj = self.dual_objective()
        self.obj_val = pobj
        if isinstance(self.algorithm, DualCentering):      # dual_centering.jl:293-303
            self.update_residuals()
            self._update_infeasibilities()
            self.inf_compl = self.dc_nu / max(self.dc_t, EPS) / max(1.0, abs(pobj) + abs(dobj))
            self._update_status(self.inf_compl <= self.opt.tol)
        else:
            self._update_infeasibilities()
            self.inf_compl = abs(pobj - dobj) / max(1.0, abs(pobj) + abs(dobj))
            conv = max(self.inf_pr, self.inf_du, self.inf_compl) <= self.opt.tol
            self._update_status(conv)
            if self.status == Status.REGULAR and self.mu < self.opt.mu_min:
                self.status = Status.SEARCH_DIRECTION_BECOMES_TOO_SMALL
        self.trace.append(IterRecord(self.iter, pobj, dobj, self.inf_pr, self.inf_du,
                                     self.inf_compl, self.mu, self.alpha, self.beta, self.sigma))

    def solve(self):                                       # solver.jl:385-396
        self.initialize()
        self.update_termination()
        while self.status == Status.REGULAR:
            self.iterate()
            if self.status != Status.REGULAR:
                break
            self.update_termination()
        return self

    # -- QRSchur (qr_schur.jl) ------------------------------------------------
    def _qr_factorize(self):                               # qr_schur.jl:24-88
        rows = sum(b.n * (b.n + 1) // 2 for b in self.blocks) + self.n_scalar
        B = np.zeros((rows, self.m))
        off = 0
        for ops, b in zip(self.ops, self.blocks):
            n = b.n
            iu = np.triu_indices(n)          # column-major svec order: col, row<=col
            order = np.lexsort((iu[0], iu[1]))
            r_idx, c_idx = iu[0][order], iu[1][order]
            w = np.where(r_idx == c_idx, 1.0, SQRT2)
            for j in range(self.m):
                Aj = ops.dense_matrix(j)
                if not np.any(Aj):
                    continue
                P = b.G.T @ (Aj @ b.G)
                B[off:off + r_idx.size, j] = w * P[r_idx, c_idx]
            off += n * (n + 1) // 2
        if self.n_scalar > 0:
            wgt = np.sqrt(self.x / self.z)
            B[off:off + self.n_scalar, :] = (self.model.C_lin * wgt[None, :]).T
        scale = np.ones(self.m)
        for j in range(self.m):
            mag = np.linalg.norm(B[:, j])
            if mag == 0 or not np.isfinite(mag):
                return False
            scale[j] = 1.0 / mag
            B[:, j] *= scale[j]
        if rows < self.m:
            return False
        Q, R = np.linalg.qr(B, mode="reduced")
        dR = np.abs(np.diag(R))
        # A dense Householder QR leaves roundoff where SPQR's exact zero pivot
        # would be; a pivot below this tolerance counts as zero.
        if np.any(dR <= 64 * EPS * max(1.0, float(np.max(dR)) if dR.size else 1.0)):
            return False
        self.qr = (Q, R, scale, rows)
        return True

    def _qr_solve_directions(self, mu, corrected):         # qr_schur.jl:90-160
        Q, R, scale, rows = self.qr
        base = np.zeros(rows)
        off = 0
        svecs = []
        for b in self.blocks:
            n = b.n
            iu = np.triu_indices(n)
            order = np.lexsort((iu[0], iu[1]))
            r_idx, c_idx = iu[0][order], iu[1][order]
            svecs.append((r_idx, c_idx))
            P = -(b.G.T @ (b.Rd @ b.G))
            P[np.diag_indices(n)] += mu / b.D - b.D
            if corrected:
                P = P + b.RNT
            w = np.where(r_idx == c_idx, 1.0, SQRT2)
            base[off:off + r_idx.size] = w * P[r_idx, c_idx]
            off += n * (n + 1) // 2
        if self.n_scalar > 0:
            wgt = np.sqrt(self.x / self.z)
            prod = np.sqrt(self.x * self.z)
            v = mu / prod - prod - wgt * self.Rd_lin
            if corrected:
                v = v + self.corr_lin / wgt
            base[off:off + self.n_scalar] = v
        # B' dxhat = Rp and dxhat = base + B dy, with B = Q R diag(1/scale)
        q1 = Q.T @ base
        c = sla.solve_triangular(R, scale * self.Rp, trans="T", check_finite=False)
        t = c - q1
        self.dely = scale * sla.solve_triangular(R, t, check_finite=False)
        direction = base + Q @ t
        off = 0
        for i, (ops, b) in enumerate(zip(self.ops, self.blocks)):
            n = b.n
            r_idx, c_idx = svecs[i]
            v = direction[off:off + r_idx.size]
            v = np.where(r_idx == c_idx, v, v / SQRT2)
            Mm = np.zeros((n, n))
            Mm[r_idx, c_idx] = v
            Mm[c_idx, r_idx] = v
            b.delX = symmetrize((b.G @ Mm) @ b.G.T)
            AT = ops.jtprod(self.dely)
            b.delS = symmetrize(b.Rd - (0 if AT is None else AT))
            off += n * (n + 1) // 2
        if self.n_scalar > 0:
            self.delz = self.Rd_lin - self.model.C_lin.T @ self.dely
            self.delx = np.sqrt(self.x / self.z) * direction[off:off + self.n_scalar]

    # -- DualCentering (dual_centering.jl) -----------------------------------
    def _refresh_slacks(self):                             # :83-86
        self._set_dual_slacks()
        for b in self.blocks:
            Ls = cholesky_factor(b.S)
            if Ls is None:
                return False
            b.Ls = Ls
            b.Si = inverse_from_cholesky(Ls)
        return True

    def _compute_gradient(self):                           # :89-98
        g = np.zeros(self.m)
        for ops, b in zip(self.ops, self.blocks):
            ops.jprod(b.Si, g)
        self.dc_grad = g
        self.dc_g = g - self.dc_t * self.model.b

    def _newton_direction(self):                           # :102-110
        self._compute_gradient()
        if not self.factorize_schur():
            return False
        self.dc_d = self.schur_solve(self.dc_g)
        self.dc_lambda = math.sqrt(max(float(self.dc_g @ self.dc_d), 0.0))
        return True

    def _dc_initialize(self):                              # :112-143
        a = self.algorithm
        self.dc_nu = sum(b.n for b in self.blocks)
        if a.A is None:
            tau = 2 * a.beta / (1 - a.beta) ** 2
            self.dc_A = a.delta + (-tau - math.log1p(-tau))
        else:
            self.dc_A = float(a.A)
        self.n_predictor = self.n_newton = 0
        if not self._refresh_slacks():
            raise RuntimeError("MadSDP: the initial `y` does not give a dual slack `S > 0`.")
        self.dc_t = 0.0
        if not self._newton_direction():
            raise RuntimeError("MadSDP: failed to factorize the dual Hessian at y0.")
        bvec = self.model.b
        self.dc_dy = self.schur_solve(bvec)
        bBb = max(float(bvec @ self.dc_dy), EPS)
        bBg = float(bvec @ self.dc_d)
        self.dc_t = max(bBg / bBb, a.beta / math.sqrt(bBb))
        for _ in range(a.centering_max_iter):
            if not self._newton_direction():
                raise RuntimeError("MadSDP: failed to factorize the dual Hessian during the initial centering.")
            if self.dc_lambda <= a.beta / 2:
                break
            self.y = self.y - self.dc_d / (1 + self.dc_lambda)
            self.iter += 1
            self.n_newton += 1
            if not self._refresh_slacks():
                raise RuntimeError("MadSDP: damped Newton left the interior during the initial centering.")
        self.mu = 1.0 / self.dc_t

    def _xi_value(self, a):                                # :146-159
        s = 0.0
        b2 = a / (1 - a)
        for b in self.blocks:
            u = 1 - a * b.eigvals
            v = 1 + b2 * b.eigvals
            if np.any(u <= 0) or np.any(v <= 0):
                return math.inf
            s -= float(np.sum(np.log(u)))
            s -= float(np.sum(np.log(v)))
        return s

    def _xi_alpha_max(self):                               # :162-176
        amax = 1.0
        for b in self.blocks:
            for mu in b.eigvals:
                if mu > 0:
                    amax = min(amax, 1.0 / mu)
                if mu < 0:
                    amax = min(amax, 1.0 / (1 - mu))
        return amax

    def _solve_alpha(self, A):                             # :180-201
        amax = self._xi_alpha_max()
        hi = amax * 0.99999
        xhi = self._xi_value(hi)
        if not math.isfinite(xhi) or xhi < A:
            return hi
        lo = 0.0
        for _ in range(80):
            mid = (lo + hi) / 2
            if self._xi_value(mid) < A:
                lo = mid
            else:
                hi = mid
            if hi - lo <= EPS * max(1.0, hi):
                break
        return (lo + hi) / 2

    def _proximity_eigvals(self, b):                       # :205-215
        L = cholesky_factor(b.Shat)
        if L is None:
            return False
        T = _trsm_left_lower(L, b.Dc)
        T = _trsm_left_lower(L, T.T).T
        b.eigvals = np.linalg.eigvalsh(symmetrize(T))
        return True

    def _shat_norm2(self):                                 # :219-229
        s = 0.0
        for b in self.blocks:
            T = _trsm_left_lower(b.Ls, b.Shat)
            T = _trsm_left_lower(b.Ls, T.T).T
            s += float(np.sum(T * T))
        return s

    def _predictor_step(self):                             # :232-264
        self.dc_yhat = self.y + self.dc_d
        self.dc_dy = self.dc_t * self.schur_solve(self.model.b)
        for i, (ops, b) in enumerate(zip(self.ops, self.blocks)):
            b.Shat = self._slack(i, self.dc_yhat)
            AT = ops.jtprod(self.dc_dy)
            b.Dc = symmetrize(np.zeros((b.n, b.n)) if AT is None else AT)
            if not self._proximity_eigvals(b):
                return False
        shat_n2 = self._shat_norm2()
        a = self._solve_alpha(self.dc_A)
        self.alpha = self.beta = a
        for b in self.blocks:
            buf = (1 - a) * b.Shat + a * b.Dc
            b.X = symmetrize((b.Si @ buf) @ b.Si * (1.0 / self.dc_t))
        self.y = self.dc_yhat + a * self.dc_dy
        self.dc_t = self.dc_nu * self.dc_t / ((1 - a) * shat_n2)
        self.n_predictor += 1
        return True

    def _dc_iterate(self):                                 # :266-289
        a = self.algorithm
        if not self._newton_direction():
            self.status = Status.ERROR_IN_STEP_COMPUTATION
            return
        if self.dc_lambda > a.beta:
            self.y = self.y - self.dc_d / (1 + self.dc_lambda)
            self.alpha = self.beta = 1.0 / (1 + self.dc_lambda)
            self.n_newton += 1
        elif not self._predictor_step():
            self.status = Status.ERROR_IN_STEP_COMPUTATION
            return
        self.iter += 1
        if not self._refresh_slacks():
            self.status = Status.ERROR_IN_STEP_COMPUTATION
            return
        self.mu = 1.0 / self.dc_t


def madsdp(model, **kw):
    """`madsdp(model; kwargs...)` (reference src/solver.jl:404)."""
    return Solver(model, **kw).solve()
