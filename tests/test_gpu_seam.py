"""The reference's Schur-solver seam driven the way MadSDP's own driver drives
it, through the C ABI (the calls madsdp.jl_amd/julia/MadSDPHIP.jl makes).

The host owns the iterate, as the stock Julia driver does, and runs one
Mehrotra predictor-corrector iteration (src/primal_dual.jl:85-101) with the
device as the Schur component:

    factorize_schur!(ws, solver)            src/solver.jl:295-314 / qr_schur.jl:24
        set X, S, y, x, z -> update_residuals -> prepare -> factorize_schur
    mu = <X,S> + x'z over nu                (host)
    solve_directions!(ws, solver, 0, false) src/primal_dual.jl:44-54 / qr_schur.jl:90
        solve_directions -> get dy, dX, dS, dx, dz
    step_lengths (HIPEigenStep)             src/step.jl:186-198, on the device
    mu_aff, sigma                           (host, src/step.jl:231-243)
    solve_directions!(ws, solver, sigma mu, true)
        correction -> solve_directions -> get directions
    step_lengths; _take_step!               (host, src/primal_dual.jl:56-72)

Each iteration's iterate is compared with the FP64 oracle's iterate after the
same iteration (1e-9 relative), and the host's termination test must stop at
the oracle's iteration.
"""
import ctypes as C

import numpy as np
import pytest

import kat_models as K
import madsdp_amd as M
from madsdp_amd import instances
from madsdp_amd import lib as L
from oracle import madsdp_oracle as O

pytestmark = pytest.mark.gpu

X_, S_, Y_, XLIN, ZLIN, DELX, DELS, DELY, DELXLIN, DELZLIN = range(10)


class HostDriver:
    """The stock driver's host side for one solve, with the device solver
    handle as its Schur workspace."""

    def __init__(self, model, direction):
        self.model = model
        self.dev = M.MadSDPSolver(model, direction=direction())
        self.lib, self.h = self.dev.lib, self.dev._h
        ref = O.Solver(model, direction=getattr(O, type(direction()).__name__)())
        ref.initialize()                       # the host's initial point (ScaledIdentity)
        self.X = [b.X.copy() for b in ref.blocks]
        self.S = [b.S.copy() for b in ref.blocks]
        self.y, self.x, self.z = ref.y.copy(), ref.x.copy(), ref.z.copy()
        self.nu = sum(model.dims) + model.n_scalar
        self.dX = [np.zeros_like(X) for X in self.X]
        self.dS = [np.zeros_like(X) for X in self.X]
        self.dx, self.dz = np.zeros(model.n_scalar), np.zeros(model.n_scalar)
        ops = [O._BlockOps(model.A[i], model.m) for i in range(model.nblocks)]
        self.Aj = [[o.dense_matrix(j) for j in range(model.m)] for o in ops]

    def _set(self, which, block, a):
        a = np.asfortranarray(a, dtype=np.float64)
        L.check(self.lib.msdp_solver_set(self.h, which, block, L.dptr(a.ravel(order="F"))), "set")

    def _get(self, which, block, shape):
        out = np.zeros(int(np.prod(shape)))
        L.check(self.lib.msdp_solver_get(self.h, which, block, L.dptr(out)), "get")
        return out.reshape(shape, order="F")

    def factorize_schur(self):
        for i in range(len(self.X)):
            self._set(X_, i, self.X[i])
            self._set(S_, i, self.S[i])
        self._set(Y_, -1, self.y)
        if self.model.n_scalar:
            self._set(XLIN, -1, self.x)
            self._set(ZLIN, -1, self.z)
        L.check(self.lib.msdp_solver_update_residuals(self.h), "update_residuals")
        ok = C.c_int32()
        L.check(self.lib.msdp_solver_prepare(self.h, C.byref(ok)), "prepare")
        if not ok.value:
            return False
        L.check(self.lib.msdp_solver_factorize_schur(self.h, C.byref(ok)), "factorize_schur")
        return bool(ok.value)

    def solve_directions(self, mu, corrected):
        if corrected:
            L.check(self.lib.msdp_solver_correction(self.h), "correction")
        L.check(self.lib.msdp_solver_solve_directions(self.h, mu, int(corrected)), "solve_directions")
        self.dy = self._get(DELY, -1, (self.model.m,))
        self.dX = [self._get(DELX, i, X.shape) for i, X in enumerate(self.X)]
        self.dS = [self._get(DELS, i, X.shape) for i, X in enumerate(self.X)]
        ns = self.model.n_scalar
        self.dx = self._get(DELXLIN, -1, (ns,)) if ns else np.zeros(0)
        self.dz = self._get(DELZLIN, -1, (ns,)) if ns else np.zeros(0)

    def step_lengths(self):
        a, b = C.c_double(), C.c_double()
        L.check(self.lib.msdp_solver_step_lengths(self.h, C.byref(a), C.byref(b)), "step_lengths")
        return a.value, b.value

    def complementarity(self, a=0.0, b=0.0):
        s = 0.0
        for X, S, dX, dS in zip(self.X, self.S, self.dX, self.dS):
            s += np.vdot(X + a * dX, S + b * dS)
        if self.model.n_scalar:
            s += (self.x + a * self.dx) @ (self.z + b * self.dz)
        return s / max(self.nu, 1)

    def iterate(self):
        assert self.factorize_schur()
        mu = self.complementarity()
        self.solve_directions(0.0, False)
        aa, ba = self.step_lengths()
        mu_aff = self.complementarity(aa, ba)
        sigma = min(max((mu_aff / max(mu, O.EPS)) ** 3, 1e-8), 1.0)
        self.solve_directions(sigma * mu, True)
        a, b = self.step_lengths()
        for i in range(len(self.X)):
            X = self.X[i] + a * self.dX[i]
            S = self.S[i] + b * self.dS[i]
            self.X[i], self.S[i] = (X + X.T) / 2, (S + S.T) / 2
        self.y = self.y + b * self.dy
        if self.model.n_scalar:
            self.x, self.z = self.x + a * self.dx, self.z + b * self.dz

    def converged(self, tol=1e-7):
        m = self.model
        Rp = m.b - sum(np.array([np.vdot(A, X) for A in Aj]) for Aj, X in zip(self.Aj, self.X))
        if m.n_scalar:
            Rp = Rp - m.C_lin @ self.x
        Rd = max(np.abs(C - S - sum(v * A for v, A in zip(self.y, Aj))).max()
                 for C, S, Aj in zip(m.C, self.S, self.Aj))
        cn = max(np.abs(C).max() for C in m.C)
        if m.n_scalar:
            Rd = max(Rd, np.abs(m.d_lin - self.z - m.C_lin.T @ self.y).max())
            cn = max(cn, np.abs(m.d_lin).max())
        pobj = sum(np.vdot(C, X) for C, X in zip(m.C, self.X)) + (m.d_lin @ self.x if m.n_scalar else 0)
        dobj = m.b @ self.y
        inf_pr = np.abs(Rp).max() / max(1.0, np.abs(m.b).max())
        inf_du = Rd / max(1.0, cn)
        gap = abs(pobj - dobj) / max(1.0, abs(pobj) + abs(dobj))
        return max(inf_pr, inf_du, gap) <= tol


CASES = [("qr_example_2", lambda: K.qr_example(2)), ("maxcut_lro", lambda: K.maxcut(K.W_LRO)),
         ("mixed", K.mixed), ("dense_12_30_4", lambda: instances.dense_sdp(12, 30, 4, 3)),
         ("blocks", lambda: instances.block_sdp([5, 9, 3], 20, 3, 4))]


@pytest.mark.parametrize("direction", [M.NT, M.HKM], ids=["NT", "HKM"])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_stock_driver_through_schur_seam(gpu, name, make, direction):
    model = make()
    print(f"seam {name} {direction.__name__}", flush=True)
    drv = HostDriver(model, direction)
    ref = O.Solver(model, direction=getattr(O, direction.__name__)())
    ref.initialize()
    ref.update_termination()
    it = 0
    while ref.status == O.Status.REGULAR:
        ref.iterate()
        assert ref.status == O.Status.REGULAR or ref.status == O.Status.SOLVE_SUCCEEDED
        drv.iterate()
        it += 1
        scale = max(1.0, max(np.abs(X).max() for X in ref_blocks_X(ref)))
        for X, Xr in zip(drv.X, ref_blocks_X(ref)):
            assert np.abs(X - Xr).max() <= 1e-9 * scale, (it, np.abs(X - Xr).max())
        assert np.abs(drv.y - ref.y).max() <= 1e-9 * max(1.0, np.abs(ref.y).max()), it
        ref.update_termination()
        assert drv.converged() == (ref.status == O.Status.SOLVE_SUCCEEDED), it
    assert ref.status == O.Status.SOLVE_SUCCEEDED and it == ref.iter
    drv.dev.close()


def ref_blocks_X(ref):
    return [b.X for b in ref.blocks]
