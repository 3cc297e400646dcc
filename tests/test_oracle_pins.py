"""Independent mathematical pins of the oracle's Mehrotra correction, step-length
rules and DualCentering predictor (CPU only).

Each check recomputes a quantity of one IPM step by a route that shares no
code or formula with the oracle:

* the NT correction (reference src/directions.jl:86-100) from the Lyapunov
  equation V R + R V = -(dX^ dS^ + dS^ dX^) in the symmetric scaling
  W^1/2 (scipy `solve_sylvester`), W from the matrix geometric mean;
* the corrections in the commuting (diagonal) case: -dx dz / z;
* a whole Mehrotra predictor-corrector iteration (src/primal_dual.jl:85-101)
  with every Newton system solved as the full unreduced KKT matrix
  (np.linalg.solve over vec(dX), vec(dS), dy, dx, dz) instead of the Schur
  complement, step lengths by bisection on eigvalsh(X + a dX);
* EigenStep (src/step.jl:139-198) against the brute-force largest feasible
  step: a = min(1, tau * a_max), with a_max found by bisection;
* DualCentering's alpha (src/dual_centering.jl:146-201) by brentq on
  xi(a) = 2 logdet S^ - logdet(S^ - a D) - logdet(S^ + a/(1-a) D), and the
  whole predictor step (:232-264) with H_jk = tr(A_j S^-1 A_k S^-1) by
  definition and the local norm as a trace.

The mutation tests plant each plausible mistake in the oracle (in memory, on
the CPU) and check that a pin fails.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.optimize as sopt

import kat_models as K
from madsdp_amd import instances
from madsdp_amd.model import Model
from oracle import madsdp_oracle as O

EPS = np.finfo(float).eps


def _spd(rng, n, shift=1.0):
    A = rng.standard_normal((n, n))
    return A @ A.T / n + shift * np.eye(n)


def _sym(rng, n, scale=1.0):
    A = rng.standard_normal((n, n)) * scale
    return (A + A.T) / 2


def _psd_sqrt(M):
    w, V = np.linalg.eigh((M + M.T) / 2)
    return (V * np.sqrt(w)) @ V.T


def _geometric_mean_W(X, S):
    """The NT scaling W = S^-1/2 (S^1/2 X S^1/2)^1/2 S^-1/2 (W S W = X)."""
    Sh = _psd_sqrt(S)
    Shi = np.linalg.inv(Sh)
    return Shi @ _psd_sqrt(Sh @ X @ Sh) @ Shi


def nt_correction_lyapunov(X, S, dX, dS):
    """Mehrotra second-order term of the NT direction in the W^1/2 scaling:
    V R + R V = -(dX^ dS^ + dS^ dX^), corr = W^1/2 R W^1/2."""
    W = _geometric_mean_W(X, S)
    Wh = _psd_sqrt(W)
    Whi = np.linalg.inv(Wh)
    V = Wh @ S @ Wh
    dXh = Whi @ dX @ Whi
    dSh = Wh @ dS @ Wh
    C = -(dXh @ dSh + dSh @ dXh)
    R = sla.solve_sylvester(V, V, C)
    return Wh @ R @ Wh


# --------------------------------------------------------------------------
# corrections
# --------------------------------------------------------------------------
def _nt_block_state(seed, n=6):
    rng = np.random.default_rng(seed)
    model = instances.dense_sdp(n, 5, 0, seed)
    s = O.Solver(model)
    s.initialize()
    b = s.blocks[0]
    b.X, b.S = _spd(rng, n), _spd(rng, n, 0.5)
    s.update_residuals()
    assert s.prepare()
    b.delX, b.delS = _sym(rng, n), _sym(rng, n)
    return s, b


def nt_correction_error(seed=0):
    s, b = _nt_block_state(seed)
    s.correction()
    ref = nt_correction_lyapunov(b.X, b.S, b.delX, b.delS)
    return float(np.abs(b.corr - ref).max() / np.abs(ref).max())


def diagonal_correction_error(direction, seed=1):
    """Commuting case: X, S, dX, dS diagonal -> corr_ii = -dX_ii dS_ii / S_ii;
    scalars: corr_lin = -dx dz / z (src/step.jl:128-136)."""
    rng = np.random.default_rng(seed)
    n = 5
    model = instances.dense_sdp(n, 4, 3, seed)
    s = O.Solver(model, direction=direction)
    s.initialize()
    b = s.blocks[0]
    x, sv = rng.uniform(0.5, 2, n), rng.uniform(0.5, 2, n)
    b.X, b.S = np.diag(x), np.diag(sv)
    s.x, s.z = rng.uniform(0.5, 2, 3), rng.uniform(0.5, 2, 3)
    s.update_residuals()
    assert s.prepare()
    dx, ds = rng.standard_normal(n), rng.standard_normal(n)
    b.delX, b.delS = np.diag(dx), np.diag(ds)
    s.delx, s.delz = rng.standard_normal(3), rng.standard_normal(3)
    s.correction()
    ref = np.diag(-dx * ds / sv)
    ref_lin = -s.delx * s.delz / s.z
    return max(float(np.abs(b.corr - ref).max() / np.abs(ref).max()),
               float(np.abs(s.corr_lin - ref_lin).max() / np.abs(ref_lin).max()))


def test_nt_correction_matches_lyapunov():
    for seed in range(3):
        assert nt_correction_error(seed) <= 1e-9


@pytest.mark.parametrize("direction", [O.NT, O.HKM], ids=["NT", "HKM"])
def test_correction_diagonal_closed_form(direction):
    assert diagonal_correction_error(direction()) <= 1e-12


# --------------------------------------------------------------------------
# step lengths
# --------------------------------------------------------------------------
def max_psd_step(X, dX, hi=1e12):
    """sup{a >= 0 : X + a dX is positive definite}, by bisection on the
    smallest eigenvalue (inf if dX keeps X positive definite up to `hi`)."""
    f = lambda a: np.linalg.eigvalsh(X + a * dX)[0] > 0
    if f(hi):
        return math.inf
    lo = 0.0
    while hi - lo > 1e-13 * max(1.0, lo):
        mid = (lo + hi) / 2
        if f(mid):
            lo = mid
        else:
            hi = mid
    return (lo + hi) / 2


def max_scalar_step(x, dx):
    neg = dx < 0
    return math.inf if not np.any(neg) else float(np.min(x[neg] / -dx[neg]))


def eigen_step_error(seed=0):
    rng = np.random.default_rng(seed)
    tau = 0.95
    worst = 0.0
    n = 7
    X = _spd(rng, n)
    L = np.linalg.cholesky(X)
    cases = [_sym(rng, n, 3.0),                    # a_max well below 1
             _sym(rng, n, 0.4),                    # a_max near 1/tau
             _spd(rng, n),                         # a_max = inf -> 1
             -1e-10 * np.eye(n),                   # lambda in (-sqrt(eps), 0): a_max huge -> 1
             -X / 0.9]                             # a_max = 0.9 exactly
    for dX in cases:
        got = O.Solver.eigen_step(L, dX, tau)
        ref = min(1.0, tau * max_psd_step(X, dX))
        worst = max(worst, abs(got - ref) / ref)
    return worst


def test_eigen_step_brute_force():
    assert eigen_step_error() <= 1e-9


def step_lengths_error(seed=3):
    rng = np.random.default_rng(seed)
    model = instances.block_sdp([4, 6], 8, 3, seed)
    s = O.Solver(model)
    s.initialize()
    for b in s.blocks:
        b.X, b.S = _spd(rng, b.n), _spd(rng, b.n)
    s.x, s.z = rng.uniform(0.5, 2, 3), rng.uniform(0.5, 2, 3)
    s.update_residuals()
    assert s.prepare()
    for b in s.blocks:
        b.delX, b.delS = _sym(rng, b.n, 2.0), _sym(rng, b.n, 2.0)
    s.delx, s.delz = rng.standard_normal(3), rng.standard_normal(3)
    a, bt = s.step_lengths()
    tau = s.algorithm.step.tau
    ra = min([1.0] + [tau * max_psd_step(b.X, b.delX) for b in s.blocks] +
             [tau * max_scalar_step(s.x, s.delx)])
    rb = min([1.0] + [tau * max_psd_step(b.S, b.delS) for b in s.blocks] +
             [tau * max_scalar_step(s.z, s.delz)])
    return max(abs(a - ra) / ra, abs(bt - rb) / rb)


def test_step_lengths_brute_force():
    assert step_lengths_error() <= 1e-9


# --------------------------------------------------------------------------
# a whole Mehrotra iteration through the unreduced KKT system
# --------------------------------------------------------------------------
def _vec(M):
    return M.reshape(-1, order="F")


def _unvec(v, n):
    return v.reshape(n, n, order="F")


def kkt_directions(model, X, S, y, x, z, mu, corr, corr_lin):
    """Solves the NT-linearised Newton system of one PSD block and the
    scalars as one dense linear system:
        <A_j, dX> + (C_lin dx)_j = Rp_j,       dS + sum_j dy_j A_j = Rd,
        dz + C_lin' dy = Rd_lin,               dX + W dS W = mu S^-1 - X + corr,
        dx + (x/z) dz = mu/z - x + corr_lin."""
    ops = O._BlockOps(model.A[0], model.m)
    n, m, ns = X.shape[0], model.m, model.n_scalar
    Aj = [ops.dense_matrix(j) for j in range(m)]
    C, Cl = model.C[0], model.C_lin
    Rp = model.b - np.array([np.vdot(A, X) for A in Aj]) - Cl @ x
    Rd = C - S - sum(yj * A for yj, A in zip(y, Aj))
    Rdl = model.d_lin - z - Cl.T @ y
    W = _geometric_mean_W(X, S)
    N2 = n * n
    nv = 2 * N2 + m + 2 * ns
    iX, iS, iy, ix, iz = 0, N2, 2 * N2, 2 * N2 + m, 2 * N2 + m + ns
    K_ = np.zeros((nv, nv))
    r = np.zeros(nv)
    row = 0
    for j in range(m):
        K_[row, iX:iX + N2] = _vec(Aj[j])
        K_[row, ix:ix + ns] = Cl[j]
        r[row] = Rp[j]
        row += 1
    K_[row:row + N2, iS:iS + N2] = np.eye(N2)
    for j in range(m):
        K_[row:row + N2, iy + j] = _vec(Aj[j])
    r[row:row + N2] = _vec(Rd)
    row += N2
    K_[row:row + ns, iz:iz + ns] = np.eye(ns)
    K_[row:row + ns, iy:iy + m] = Cl.T
    r[row:row + ns] = Rdl
    row += ns
    K_[row:row + N2, iX:iX + N2] = np.eye(N2)
    K_[row:row + N2, iS:iS + N2] = np.kron(W, W)
    r[row:row + N2] = _vec(mu * np.linalg.inv(S) - X + corr)
    row += N2
    K_[row:row + ns, ix:ix + ns] = np.eye(ns)
    K_[row:row + ns, iz:iz + ns] = np.diag(x / z)
    r[row:row + ns] = mu / z - x + corr_lin
    sol = np.linalg.solve(K_, r)
    dX = _unvec(sol[iX:iX + N2], n)
    dS = _unvec(sol[iS:iS + N2], n)
    return (dX + dX.T) / 2, (dS + dS.T) / 2, sol[iy:iy + m], sol[ix:ix + ns], sol[iz:iz + ns]


def mehrotra_iteration_error(seed=5, iters=3):
    """Runs the oracle's Mehrotra iterations on a small mixed SDP/LP and, from
    the same iterate, one independent predictor-corrector step each."""
    model = instances.dense_sdp(4, 6, 2, seed)
    s = O.Solver(model)
    s.initialize()
    s.update_termination()
    tau = s.algorithm.step.tau
    nu = model.dims[0] + model.n_scalar
    worst = 0.0
    for _ in range(iters):
        b = s.blocks[0]
        X, S, y, x, z = b.X.copy(), b.S.copy(), s.y.copy(), s.x.copy(), s.z.copy()
        s.iterate()
        mu = (np.vdot(X, S) + x @ z) / nu
        zero = np.zeros_like(X)
        dX, dS, dy, dx, dz = kkt_directions(model, X, S, y, x, z, 0.0, zero, np.zeros_like(x))

        def steps(dX, dS, dx, dz):
            a = min(1.0, tau * max_psd_step(X, dX), tau * max_scalar_step(x, dx))
            bt = min(1.0, tau * max_psd_step(S, dS), tau * max_scalar_step(z, dz))
            return a, bt

        a, bt = steps(dX, dS, dx, dz)
        mu_aff = (np.vdot(X + a * dX, S + bt * dS) + (x + a * dx) @ (z + bt * dz)) / nu
        sigma = min(max((mu_aff / mu) ** 3, 1e-8), 1.0)
        corr = nt_correction_lyapunov(X, S, dX, dS)
        corr_lin = -dx * dz / z
        dX, dS, dy, dx, dz = kkt_directions(model, X, S, y, x, z, sigma * mu, corr, corr_lin)
        a, bt = steps(dX, dS, dx, dz)
        Xn, Sn, yn = X + a * dX, S + bt * dS, y + bt * dy
        xn, zn = x + a * dx, z + bt * dz
        for got, ref in ((s.blocks[0].X, Xn), (s.blocks[0].S, Sn), (s.y, yn), (s.x, xn), (s.z, zn)):
            worst = max(worst, float(np.abs(got - ref).max() / max(1.0, np.abs(ref).max())))
        worst = max(worst, abs(s.sigma - sigma) / sigma)
        s.update_termination()
    return worst


def test_mehrotra_iteration_kkt():
    assert mehrotra_iteration_error() <= 1e-8


# --------------------------------------------------------------------------
# DualCentering
# --------------------------------------------------------------------------
def xi_logdet(Shat, D, a):
    """xi(a) = -log det(I - a S^-1/2 D S^-1/2) - log det(I + a/(1-a) S^-1/2 D S^-1/2)."""
    s0 = np.linalg.slogdet(Shat)[1]
    s1, l1 = np.linalg.slogdet(Shat - a * D)
    s2, l2 = np.linalg.slogdet(Shat + a / (1 - a) * D)
    if s1 <= 0 or s2 <= 0:
        return math.inf
    return 2 * s0 - l1 - l2


def alpha_brentq(pairs, A):
    """Root of sum_blocks xi = A on (0, 0.99999 a_max) with a_max from the
    generalized eigenvalues of (D, S^); the largest interior step when xi stays
    below A there (src/dual_centering.jl:180-191)."""
    amax = 1.0
    for Shat, D in pairs:
        mu = sla.eigh(D, Shat, eigvals_only=True)
        for v in mu:
            if v > 0:
                amax = min(amax, 1 / v)
            elif v < 0:
                amax = min(amax, 1 / (1 - v))
    hi = 0.99999 * amax
    f = lambda a: sum(xi_logdet(Sh, D, a) for Sh, D in pairs) - A
    fh = f(hi)
    if not math.isfinite(fh) or fh < 0:
        return hi
    return sopt.brentq(f, 0.0, hi, xtol=1e-15, rtol=4 * EPS)


def dc_alpha_error(seed=0):
    rng = np.random.default_rng(seed)
    worst = 0.0
    s = O.Solver(K.maxcut(K.W_LRO), algorithm=O.DualCentering())
    for scale, A in ((0.3, 2.0), (1.5, 2.0), (0.05, 2.0), (0.8, 0.5)):
        pairs = []
        for b, n in zip(s.blocks, (4,)):
            b.Shat = _spd(rng, n)
            b.Dc = _sym(rng, n, scale)
            assert s._proximity_eigvals(b)
            pairs.append((b.Shat, b.Dc))
        got = s._solve_alpha(A)
        ref = alpha_brentq(pairs, A)
        worst = max(worst, abs(got - ref) / ref)
    return worst


def test_dc_alpha_brentq():
    assert dc_alpha_error() <= 1e-9


def dc_predictor_error(model=None, steps=4):
    """From the iterates of a DualCentering run on the 4-node max-cut KAT,
    recompute each Newton decrement and predictor step independently."""
    model = K.maxcut(K.W_LRO) if model is None else model
    s = O.Solver(model, algorithm=O.DualCentering())
    s.initialize()
    ops = O._BlockOps(model.A[0], model.m)
    Aj = [ops.dense_matrix(j) for j in range(model.m)]
    C = model.C[0]
    nu = model.dims[0]
    worst = 0.0
    done = 0
    while done < steps:
        y, t = s.y.copy(), s.dc_t
        Sy = C - sum(v * A for v, A in zip(y, Aj))
        Si = np.linalg.inv(Sy)
        H = np.array([[np.trace(Ai @ Si @ Ak @ Si) for Ak in Aj] for Ai in Aj])
        g = np.array([np.vdot(A, Si) for A in Aj]) - t * model.b
        d = np.linalg.solve(H, g)
        lam = math.sqrt(g @ d)
        s.iterate()
        assert s.status == O.Status.REGULAR
        worst = max(worst, abs(s.dc_lambda - lam) / lam)
        if lam > s.algorithm.beta:
            ref_y = y - d / (1 + lam)
            worst = max(worst, float(np.abs(s.y - ref_y).max() / max(1, np.abs(ref_y).max())))
            continue
        yhat = y + d
        dy = t * np.linalg.solve(H, model.b)
        Shat = C - sum(v * A for v, A in zip(yhat, Aj))
        D = sum(v * A for v, A in zip(dy, Aj))
        a = alpha_brentq([(Shat, D)], s.dc_A)
        nrm2 = float(np.trace(Si @ Shat @ Si @ Shat))
        t_new = nu * t / ((1 - a) * nrm2)
        X = Si @ ((1 - a) * Shat + a * D) @ Si / t
        y_new = yhat + a * dy
        for got, ref in ((s.alpha, a), (s.dc_t, t_new)):
            worst = max(worst, abs(got - ref) / abs(ref))
        for got, ref in ((s.blocks[0].X, X), (s.y, y_new)):
            worst = max(worst, float(np.abs(got - ref).max() / max(1.0, np.abs(ref).max())))
        done += 1
    return worst


def test_dc_predictor_independent():
    assert dc_predictor_error() <= 1e-8


# --------------------------------------------------------------------------
# mutations: each plausible slip in the oracle must fail a pin
# --------------------------------------------------------------------------
def _scaled_correction(factor):
    orig = O.Solver._block_correction

    def mutated(self, b):
        return factor * orig(self, b)
    return mutated


def _scaled_alpha(factor):
    orig = O.Solver._solve_alpha

    def mutated(self, A):
        return factor * orig(self, A)
    return mutated


def _scaled_xi(factor):
    orig = O.Solver._xi_value

    def mutated(self, a):
        return factor * orig(self, a)
    return mutated


def _tau_always(cls, L, Delta, tau):
    lam = cls.lambda_min_congruence(L, Delta)
    return tau if lam > -math.sqrt(EPS) else min(1.0, -tau / lam)


MUTATIONS = {
    "correction x0.5": ("_block_correction", _scaled_correction(0.5),
                        [nt_correction_error, mehrotra_iteration_error]),
    "correction x0": ("_block_correction", _scaled_correction(0.0),
                      [nt_correction_error, mehrotra_iteration_error,
                       lambda: diagonal_correction_error(O.HKM())]),
    "eigen step alpha = tau": ("eigen_step", classmethod(_tau_always), [eigen_step_error]),
    "dc alpha x0.7": ("_solve_alpha", _scaled_alpha(0.7), [dc_predictor_error]),
    "dc xi x2": ("_xi_value", _scaled_xi(2.0), [dc_alpha_error, dc_predictor_error]),
}


@pytest.mark.parametrize("name", list(MUTATIONS))
def test_mutation_is_caught(name, monkeypatch):
    attr, repl, pins = MUTATIONS[name]
    monkeypatch.setattr(O.Solver, attr, repl)
    for pin in pins:
        try:
            err = pin()
        except AssertionError:
            continue            # the pin itself refused the mutated state
        assert err > 1e-6, (name, pin, err)
