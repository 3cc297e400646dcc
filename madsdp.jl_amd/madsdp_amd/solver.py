"""Host mirror of MadSDP's solver API over the gfx950 library.

Same names, arguments, defaults and error behaviour as the reference:
`MadSDPSolver(model; direction, step, centering, algorithm, initial_point,
schur, options...)` (src/solver.jl:87-160), `solve!`, `initialize!`,
`iterate!`, `update_termination!` (src/solver.jl:286-396), the component
types of src/components.jl and the fields of MadSDPOptions
(src/options.jl:3-15).  All arithmetic runs on the device; this module only
moves data and calls the C ABI.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Optional

import numpy as np

from . import lib as L
from .model import Model


class Status(IntEnum):
    """MadNLP.Status codes (src/moi.jl:16-27)."""
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


# ----------------------------------------------------------- components ----
class NT:
    """Nesterov-Todd direction (src/components.jl:96)."""
    code = 0


class HKM:
    """HKM direction (src/components.jl:110)."""
    code = 1


@dataclass
class EigenStep:
    """Fraction tau of the largest step keeping X, S PD (components.jl:132-135)."""
    tau: float = 0.95


@dataclass
class CholeskyBacktrack:
    """Cholesky backtracking line search (components.jl:146-150)."""
    shrink: float = 0.8
    boundary: float = 0.95


@dataclass
class Mehrotra:
    sigma_exponent: float = 3.0   # components.jl:169-172


@dataclass
class AdaptiveMu:
    initial_ratio: float = 0.25   # components.jl:188-192
    ratio: float = 0.5


class ScaledIdentity:
    """X = S = beta I (components.jl:208-217); explicit scale must be finite and positive."""

    def __init__(self, scale=None):
        if scale is not None:
            scale = float(scale)
            if not math.isfinite(scale) or scale <= 0:
                raise ValueError("The identity scale must be finite and positive.")
        self.scale = scale


@dataclass
class DualShift:
    initial_shift: float = 1.0    # components.jl:228-232
    max_shift: float = 1e10


class InitialPoint:
    """InitialPoint(X, y; x) (components.jl:243-251)."""

    def __init__(self, X, y, x=None):
        self.X = [np.array(Xi, dtype=np.float64) for Xi in X]
        self.y = np.array(y, dtype=np.float64)
        self.x = None if x is None else np.array(x, dtype=np.float64)


class DenseCholesky:
    """Dense Cholesky of the Schur complement (components.jl:273)."""


class MadNLPSchur:
    """`MadNLPSchur()` (components.jl:275-283): the Schur complement factored by
    a symmetric indefinite LDL' with Bunch-Kaufman pivoting (MadNLP's LAPACK
    solver in BUNCHKAUFMAN mode, kkt.jl:15-25), here on the device."""


class QRSchur:
    """QR of the NT-scaled constraint matrix (components.jl:285-300)."""


@dataclass
class PrimalDual:
    direction: object = field(default_factory=NT)
    step: object = field(default_factory=EigenStep)
    centering: object = field(default_factory=Mehrotra)


@dataclass
class DualCentering:
    beta: float = 0.2
    delta: float = 1.0
    A: Optional[float] = 2.0
    centering_max_iter: int = 50


OPTION_FIELDS = ("tol", "max_iter", "max_wall_time", "mu_min", "regularize_schur", "max_schur_reg",
                 "mu_switch", "refine_max_iter", "lowp_min_gflop", "shared_device")
# Schur assembly precision of the dense constraint blocks: "fp64" (plain),
# "mixed" (FP32 MFMA assembly while mu >= mu_switch, FP64 Cholesky and
# iterative refinement against the FP64 operator; FP64 below mu_switch, and
# throughout when one dense assembly costs less than lowp_min_gflop GFLOP) or
# "bf16x3" (the same with H = B'B formed from the FP64 scaled panel B split
# into bf16 hi + lo by three bf16 MFMA products, refined through B).
PRECISIONS = {"fp64": 0, "mixed": 1, "bf16x3": 2}

X_, S_, Y_, XLIN, ZLIN, DELX, DELS, DELY, DELXLIN, DELZLIN, RD, RP, RDLIN, W_, G_, SI, CORR, LS, LX, H_, D_ = range(21)


class MadSDPSolver:
    """`MadSDPSolver(model; kwargs...)` on one HIP device."""

    def __init__(self, model: Model, direction=None, step=None, centering=None, algorithm=None,
                 initial_point=None, schur=None, device: int = 0, **options):
        for k in options:
            if k not in OPTION_FIELDS and k not in ("print_level", "output_file", "datasparsity",
                                                    "precision"):
                raise TypeError(f"unknown option {k}")
        prec = options.get("precision", "fp64")
        if prec not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        direction = NT() if direction is None else direction
        step = EigenStep() if step is None else step
        centering = Mehrotra() if centering is None else centering
        if algorithm is None:
            algorithm = PrimalDual(direction, step, centering)
        if initial_point is None:
            initial_point = DualShift() if isinstance(algorithm, DualCentering) else ScaledIdentity()
        schur = DenseCholesky() if schur is None else schur
        if model.nblocks == 0 and model.n_scalar == 0:
            raise RuntimeError("MadSDP: model has no variables.")
        if isinstance(algorithm, DualCentering):
            if model.n_scalar > 0:
                raise RuntimeError("MadSDP: `DualCentering` does not support scalar (LP) blocks yet.")
        if isinstance(schur, QRSchur) and not (isinstance(algorithm, PrimalDual)
                                               and isinstance(algorithm.direction, NT)):
            raise ValueError("QRSchur requires PrimalDual with the NT direction.")
        self.model, self.algorithm, self.initial_point, self.schur = model, algorithm, initial_point, schur
        self.lib = L.load()
        h = C.c_void_p()
        dims = np.array(model.dims, dtype=np.int64)
        L.check(self.lib.msdp_solver_create(C.byref(h), device, model.nblocks, L.iptr(dims),
                                            model.m, model.n_scalar), "msdp_solver_create")
        self._h = h
        self._upload(model)
        opt = L.Options()
        self.lib.msdp_default_options(C.byref(opt))
        for k in OPTION_FIELDS:
            if k in options:
                setattr(opt, k, options[k])
        opt.precision = PRECISIONS[prec]
        if isinstance(algorithm, DualCentering):
            opt.algorithm = 1
            opt.dc_beta, opt.dc_delta = algorithm.beta, algorithm.delta
            opt.dc_A = -1.0 if algorithm.A is None else algorithm.A
            opt.dc_centering_max_iter = algorithm.centering_max_iter
        else:
            opt.algorithm = 0
            opt.direction = 0 if isinstance(algorithm.direction, NT) else 1
            if isinstance(algorithm.step, EigenStep):
                opt.step, opt.tau = 0, algorithm.step.tau
            else:
                opt.step, opt.shrink, opt.boundary = 1, algorithm.step.shrink, algorithm.step.boundary
            if isinstance(algorithm.centering, Mehrotra):
                opt.centering, opt.sigma_exponent = 0, algorithm.centering.sigma_exponent
            else:
                opt.centering = 1
                opt.initial_ratio, opt.ratio = algorithm.centering.initial_ratio, algorithm.centering.ratio
        p = initial_point
        if isinstance(p, ScaledIdentity):
            opt.initial_point, opt.identity_scale = 0, (0.0 if p.scale is None else p.scale)
        elif isinstance(p, InitialPoint):
            opt.initial_point = 1
            Xcat = np.concatenate([np.asfortranarray(Xi).ravel(order="F") for Xi in p.X]) \
                if p.X else np.zeros(0)
            xv = None if p.x is None else np.ascontiguousarray(p.x, dtype=np.float64)
            L.check(self.lib.msdp_solver_set_initial_point(
                self._h, L.dptr(np.ascontiguousarray(Xcat)), L.dptr(np.ascontiguousarray(p.y)),
                None if xv is None else L.dptr(xv)), "set_initial_point")
        else:
            opt.initial_point = 2
            opt.initial_shift, opt.max_shift = p.initial_shift, p.max_shift
        opt.schur = 1 if isinstance(schur, QRSchur) else 2 if isinstance(schur, MadNLPSchur) else 0
        L.check(self.lib.msdp_solver_configure(self._h, C.byref(opt)), "configure")
        self.opt = opt

    def _upload(self, model: Model):
        lib = self.lib
        for i, Ci in enumerate(model.C):
            Cf = np.asfortranarray(Ci, dtype=np.float64)
            L.check(lib.msdp_solver_set_cost_block(self._h, i, Cf.ctypes.data_as(L.c_double_p)))
            blk = model.A[i]
            if blk.kind == "dense":
                # constraint matrices are symmetric: row- and column-major agree
                mats = np.ascontiguousarray(blk.mats, dtype=np.float64)
                idx = np.ascontiguousarray(blk.con_idx, dtype=np.int64)
                L.check(lib.msdp_solver_set_constraints_dense(self._h, i, idx.size, L.iptr(idx), L.dptr(mats)))
            elif blk.kind == "lowrank":
                idx = np.ascontiguousarray(blk.con_idx, dtype=np.int64)
                ptr = np.ascontiguousarray(blk.rank_ptr, dtype=np.int64)
                F = np.asfortranarray(blk.factors, dtype=np.float64)
                w = np.ascontiguousarray(blk.weights, dtype=np.float64)
                L.check(lib.msdp_solver_set_constraints_lowrank(
                    self._h, i, idx.size, L.iptr(idx), L.iptr(ptr),
                    F.ctypes.data_as(L.c_double_p), L.dptr(w)))
            else:
                con = np.ascontiguousarray(blk.con, dtype=np.int64)
                row = np.ascontiguousarray(blk.row, dtype=np.int64)
                col = np.ascontiguousarray(blk.col, dtype=np.int64)
                val = np.ascontiguousarray(blk.val, dtype=np.float64)
                L.check(lib.msdp_solver_set_constraints_coo(self._h, i, con.size, L.iptr(con), L.iptr(row),
                                                            L.iptr(col), L.dptr(val)))
        b = np.ascontiguousarray(model.b, dtype=np.float64)
        L.check(lib.msdp_solver_set_rhs(self._h, L.dptr(b)))
        if model.n_scalar > 0:
            d = np.ascontiguousarray(model.d_lin)
            Cl = np.asfortranarray(model.C_lin)
            L.check(lib.msdp_solver_set_scalar_data(self._h, L.dptr(d), Cl.ctypes.data_as(L.c_double_p)))

    def close(self):
        if getattr(self, "_h", None):
            self.lib.msdp_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- driver -------------------------------------------------------------
    def initialize(self):
        L.check(self.lib.msdp_solver_initialize(self._h), "initialize!")

    def iterate(self):
        L.check(self.lib.msdp_solver_iterate(self._h), "iterate!")

    def update_termination(self):
        L.check(self.lib.msdp_solver_update_termination(self._h), "update_termination!")

    def update_residuals(self):
        L.check(self.lib.msdp_solver_update_residuals(self._h), "update_residuals!")

    def set_mu(self, mu: float):
        """`solver.mu = mu` (the field test/termination.jl:22 assigns)."""
        L.check(self.lib.msdp_solver_set_mu(self._h, float(mu)), "mu")

    def solve(self, verbose: bool = False, host_loop: bool = False):
        """`solve!(solver)` (src/solver.jl:385-396).  The loop runs inside
        the library unless `host_loop`, which runs the same loop here."""
        if not host_loop:
            L.check(self.lib.msdp_solver_solve(self._h), "solve!")
        else:
            self.initialize()
            self.update_termination()
            while self.status == Status.REGULAR:
                self.iterate()
                if self.status != Status.REGULAR:
                    break
                self.update_termination()
        if verbose:
            self.print_trace()
        return self

    # -- components ---------------------------------------------------------
    def prepare(self) -> bool:
        ok = C.c_int32()
        L.check(self.lib.msdp_solver_prepare(self._h, C.byref(ok)))
        return bool(ok.value)

    def factorize_schur(self) -> bool:
        ok = C.c_int32()
        L.check(self.lib.msdp_solver_factorize_schur(self._h, C.byref(ok)))
        return bool(ok.value)

    def solve_directions(self, mu: float, corrected: bool):
        L.check(self.lib.msdp_solver_solve_directions(self._h, float(mu), int(corrected)))

    def correction(self):
        L.check(self.lib.msdp_solver_correction(self._h))

    def step_lengths(self):
        a, b = C.c_double(), C.c_double()
        L.check(self.lib.msdp_solver_step_lengths(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def affine_complementarity(self, a, b):
        out = C.c_double()
        L.check(self.lib.msdp_solver_affine_complementarity(self._h, a, b, C.byref(out)))
        return out.value

    def take_step(self, a, b):
        L.check(self.lib.msdp_solver_take_step(self._h, a, b))

    # -- state --------------------------------------------------------------
    def footprint(self) -> int:
        """Device bytes the solver holds once initialized (estimate)."""
        v = C.c_int64()
        L.check(self.lib.msdp_solver_footprint(self._h, C.byref(v)))
        return v.value

    def info(self) -> L.Info:
        inf = L.Info()
        L.check(self.lib.msdp_solver_get_info(self._h, C.byref(inf)))
        return inf

    @property
    def status(self) -> Status:
        return Status(self.info().status)

    def __getattr__(self, name):
        if name in ("mu", "sigma", "alpha", "beta", "inf_pr", "inf_du", "inf_compl", "obj_val",
                    "iter", "n_newton", "n_predictor", "schur_reg"):
            return getattr(self.info(), name)
        raise AttributeError(name)

    def primal_objective(self) -> float:
        return self.info().pobj

    def dual_objective(self) -> float:
        return self.info().dobj

    def _get(self, which, block, shape):
        out = np.zeros(int(np.prod(shape)) if shape else 0, dtype=np.float64)
        if out.size:
            L.check(self.lib.msdp_solver_get(self._h, which, block, L.dptr(out)))
        return out.reshape(shape, order="F") if len(shape) == 2 else out

    def _set(self, which, block, arr):
        a = np.asfortranarray(arr, dtype=np.float64).ravel(order="F")
        L.check(self.lib.msdp_solver_set(self._h, which, block, L.dptr(np.ascontiguousarray(a))))

    def block_array(self, which: int, i: int):
        n = self.model.dims[i]
        return self._get(which, i, (n, n))

    def X(self, i=0):
        return self.block_array(X_, i)

    def S(self, i=0):
        return self.block_array(S_, i)

    def set_X(self, i, X):
        self._set(X_, i, X)

    def set_S(self, i, S):
        self._set(S_, i, S)

    def set_y(self, y):
        self._set(Y_, -1, np.asarray(y, dtype=np.float64))

    @property
    def y(self):
        return self._get(Y_, -1, (self.model.m,))

    @property
    def x(self):
        return self._get(XLIN, -1, (self.model.n_scalar,))

    @property
    def z(self):
        return self._get(ZLIN, -1, (self.model.n_scalar,))

    @property
    def dely(self):
        return self._get(DELY, -1, (self.model.m,))

    @property
    def delx(self):
        return self._get(DELXLIN, -1, (self.model.n_scalar,))

    @property
    def delz(self):
        return self._get(DELZLIN, -1, (self.model.n_scalar,))

    @property
    def Rp(self):
        return self._get(RP, -1, (self.model.m,))

    def trace(self) -> np.ndarray:
        """Per-termination-check records: pobj, dobj, inf_pr, inf_du, gap,
        mu, alpha, beta, sigma."""
        n = C.c_int32()
        L.check(self.lib.msdp_solver_trace(self._h, 0, None, C.byref(n)))
        out = np.zeros(9 * max(n.value, 1))
        L.check(self.lib.msdp_solver_trace(self._h, n.value, L.dptr(out), C.byref(n)))
        return out[: 9 * n.value].reshape(n.value, 9)

    def print_trace(self):
        """The iteration table of print_iter (src/solver.jl:353-383)."""
        for k, r in enumerate(self.trace()):
            if k % 10 == 0:
                print(f"{'iter':>4} {'p_obj':>14} {'d_obj':>14} {'inf_pr':>9} {'inf_du':>9} "
                      f"{'gap':>9} {'mu':>9} {'a/b':>9}")
            print(f"{k:4d} {r[0]:14.6e} {r[1]:14.6e} {r[2]:9.2e} {r[3]:9.2e} {r[4]:9.2e} "
                  f"{r[5]:9.2e} {r[6]:4.2f}/{r[7]:4.2f}")


def madsdp(model: Model, **kw) -> MadSDPSolver:
    """`madsdp(model; kwargs...)` (src/solver.jl:404)."""
    return MadSDPSolver(model, **kw).solve()
