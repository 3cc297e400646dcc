"""The small problems of the reference's own tests, as models.

test/runtests.jl:58-212, test/dual_runtests.jl:191-299,
test/identity_scale.jl:8-14, test/qr_schur.jl:8-14.
"""
import numpy as np

from madsdp_amd.model import Model

W_LRO = np.array([[0, 5, 7, 6], [5, 0, 0, 1], [7, 0, 0, 1], [6, 1, 1, 0]], dtype=float)
W_CYCLE = np.array([[0, 1, 0, 1], [1, 0, 1, 0], [0, 1, 0, 1], [1, 0, 1, 0]], dtype=float)


def laplacian(W):
    return np.diag(W.sum(axis=1)) - W


def maxcut(W):
    """min <-L/4, X> s.t. X_kk = 1 (runtests.jl:72-93)."""
    n = W.shape[0]
    As = [[None] * n]
    for k in range(n):
        A = np.zeros((n, n))
        A[k, k] = 1.0
        As[0][k] = A
    return Model.from_matrices([-laplacian(W) / 4], As, np.ones(n))


def trivial():
    """min <diag(2,1), X> s.t. X_11 = 1 -> 2 (runtests.jl:60-69)."""
    return Model.from_matrices([np.diag([2.0, 1.0])], [[np.diag([1.0, 0.0])]], [1.0])


def trace_one():
    """min <diag(3,2), X> s.t. tr X = 1 -> 2 (runtests.jl:96-107)."""
    return Model.from_matrices([np.diag([3.0, 2.0])], [[np.eye(2)]], [1.0])


def mixed():
    """1x1 SDP + 2 LP scalars -> 1 (runtests.jl:113-129)."""
    one = np.ones((1, 1))
    return Model.from_matrices([one], [[one, one]], [1.0, 1.0], d_lin=[2.0, 1.0], C_lin=np.eye(2))


def qr_example(scalars=0):
    """3x3 example of test/qr_schur.jl:8-14 and test/identity_scale.jl:8-14."""
    C = np.array([[3, 0.1, 0], [0.1, 2, 0.2], [0, 0.2, 1]])
    As = [[None] * 3]
    for k in range(3):
        A = np.zeros((3, 3))
        A[k, k] = 1.0
        As[0][k] = A
    Lm = np.arange(1, 3 * scalars + 1, dtype=float).reshape(3, scalars, order="F")
    return Model.from_matrices([C], As, np.ones(3), d_lin=np.ones(scalars), C_lin=Lm)


def mc_psd_start(W):
    """Dual feasible start of mc_psd (runtests.jl:186-187)."""
    L = laplacian(W)
    return [np.eye(W.shape[0])], -(1 + np.abs(L / 4).sum(axis=0))
