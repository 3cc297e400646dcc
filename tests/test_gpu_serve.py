"""Serving loop: continuous batching of independent solves on one device.

Each result must be the same alone and in the mix: every kernel of a solve is
deterministic and runs on that solve's own stream, so the served iterates are
compared bit for bit with a solve of the same instance on its own.
"""
import numpy as np
import pytest

import madsdp_amd as M
from madsdp_amd import instances
from oracle import madsdp_oracle as O

pytestmark = pytest.mark.gpu


def mix():
    return [
        (instances.maxcut(40, 0.2, 1), {}),
        (instances.dense_sdp(20, 60, 10, 2), {}),
        (instances.maxcut(150, 0.05, 3, "pm1"), {}),
        (instances.dense_sdp(30, 200, 30, 4), dict(precision="mixed", lowp_min_gflop=0.0)),
        (instances.maxcut(90, 0.1, 5), dict(direction=M.HKM())),
        (instances.block_sdp(instances.block_sizes(5, 4, 40, 6), 70, 6, 7), {}),
        (instances.dense_sdp(16, 50, 16, 8), dict(step=M.CholeskyBacktrack(), centering=M.AdaptiveMu())),
        (instances.maxcut(300, 0.03, 9), {}),
        (instances.maxcut(400, 0.02, 12), dict(shared_device=1)),
        (instances.dense_sdp(120, 1500, 120, 13), dict(shared_device=1)),
        (instances.dense_sdp(60, 400, 0, 10), dict(precision="mixed", lowp_min_gflop=0.0)),
        (instances.dense_sdp(40, 300, 20, 14), dict(precision="bf16x3", lowp_min_gflop=0.0,
                                                    shared_device=1)),
        (instances.maxcut(25, 0.3, 11), dict(algorithm=M.DualCentering())),
    ]


def snapshot(s, model):
    return dict(status=s.status, iter=s.iter, trace=s.trace().copy(), y=s.y.copy(),
                X=[s.X(i).copy() for i in range(model.nblocks)],
                x=s.x.copy() if model.n_scalar else None)


def test_same_alone_and_in_mix(gpu):
    cases = mix()
    alone = []
    for model, kw in cases:
        s = M.MadSDPSolver(model, **kw).solve()
        alone.append(snapshot(s, model))
        s.close()
    solvers = [M.MadSDPSolver(model, **kw) for model, kw in cases]
    print(f"serving {len(solvers)} requests on 4 slots", flush=True)
    res, summary = M.serve(solvers, [0.0] * len(solvers), slots=4)
    assert summary["peak_active"] >= 2
    for k, ((model, _), s) in enumerate(zip(cases, solvers)):
        got, ref = snapshot(s, model), alone[k]
        assert got["status"] == ref["status"] == M.Status.SOLVE_SUCCEEDED, k
        assert got["iter"] == ref["iter"], k
        np.testing.assert_array_equal(got["trace"], ref["trace"])
        np.testing.assert_array_equal(got["y"], ref["y"])
        for a, b in zip(got["X"], ref["X"]):
            np.testing.assert_array_equal(a, b)
        if ref["x"] is not None:
            np.testing.assert_array_equal(got["x"], ref["x"])


def test_served_matches_oracle(gpu):
    """Served solves against the FP64 oracle per iteration, among them one
    whose products are single launches on the shared device (n = 300: the
    n x n products would otherwise be split over K)."""
    models = [instances.maxcut(60, 0.15, 21), instances.dense_sdp(24, 80, 12, 22),
              instances.maxcut(300, 0.03, 23)]
    kws = [{}, {}, dict(shared_device=1)]
    solvers = [M.MadSDPSolver(m, **kw) for m, kw in zip(models, kws)]
    M.serve(solvers, [0.0] * len(models), slots=3)
    for m, s in zip(models, solvers):
        ref = O.Solver(m).solve()
        assert s.status == M.Status.SOLVE_SUCCEEDED and s.iter == ref.iter
        tr = s.trace()
        for k, r in enumerate(ref.trace):
            assert abs(tr[k, 0] - r.pobj) <= 1e-9 * max(1, abs(r.pobj))


def test_admission_by_footprint(gpu):
    """With a budget that holds two of these requests, never more than two run."""
    models = [instances.dense_sdp(40, 300, 0, 30 + k) for k in range(6)]
    solvers = [M.MadSDPSolver(m) for m in models]
    fp = max(s.footprint() for s in solvers)
    res, summary = M.serve(solvers, [0.0] * 6, slots=6, mem_budget=2 * fp + fp // 2)
    assert summary["peak_active"] <= 2
    assert summary["peak_bytes"] <= 2 * fp + fp // 2
    assert all(s.status == M.Status.SOLVE_SUCCEEDED for s in solvers)


def test_arrivals_respected(gpu):
    models = [instances.maxcut(30 + 10 * k, 0.2, 40 + k) for k in range(4)]
    solvers = [M.MadSDPSolver(m) for m in models]
    arrivals = [0.0, 0.15, 0.3, 0.45]
    res, _ = M.serve(solvers, arrivals, slots=2)
    for r, a in zip(res, arrivals):
        assert r.arrival == a and r.admitted >= a and r.done >= r.admitted


def test_failed_solve_does_not_stop_the_loop(gpu):
    """A numerical failure (status ERROR_IN_STEP_COMPUTATION) is a result like
    any other; the other requests complete."""
    one = np.ones((1, 1))
    bad = M.Model.from_matrices([one], [[np.zeros((1, 1))]], [0.0])
    solvers = [M.MadSDPSolver(bad, schur=M.QRSchur()), M.MadSDPSolver(instances.maxcut(20, 0.3, 50))]
    M.serve(solvers, [0.0, 0.0], slots=2)
    assert solvers[0].status == M.Status.ERROR_IN_STEP_COMPUTATION
    assert solvers[1].status == M.Status.SOLVE_SUCCEEDED


def test_released_solver_rejects_device_calls(gpu):
    s = M.MadSDPSolver(instances.maxcut(20, 0.3, 60))
    M.serve([s], [0.0], slots=1)
    assert s.X(0).shape == (20, 20)
    with pytest.raises(RuntimeError):
        s.iterate()


def test_footprint_bounds_allocations(gpu):
    """The admission estimate is an upper bound of what a solver allocates:
    single dense block with scalars, many small blocks (auxiliary-stream
    scratch, recorder buffer), max-cut, the mixed path's FP32 copies and the
    QR path's dense constraint matrix, T factors and workspace."""
    cases = [(instances.dense_sdp(40, 300, 20, 5), {}),
             (instances.block_sdp(instances.block_sizes(12, 4, 60, 3), 200, 6, 3), {}),
             (instances.maxcut(300, 0.05, 2), {}),
             (instances.dense_sdp(30, 200, 30, 4), dict(precision="mixed", lowp_min_gflop=0.0)),
             (instances.dense_sdp(40, 300, 20, 5), dict(schur=M.QRSchur()))]
    for model, kw in cases:
        s = M.MadSDPSolver(model, **kw)
        est = s.footprint()
        s.solve()
        used = s.info().device_bytes
        print(f"footprint {est / 2**20:.1f} MiB, allocated {used / 2**20:.1f} MiB", flush=True)
        assert 0 < used <= est


def test_first_fit_admission(gpu):
    """A request that does not fit next to the running one is passed over by
    smaller ones that do (first fit), and still runs once memory frees."""
    big = [instances.dense_sdp(60, 700, 0, 60 + k) for k in range(2)]
    small = [instances.maxcut(20, 0.3, 70 + k) for k in range(2)]
    solvers = [M.MadSDPSolver(m) for m in big + small]
    fb = min(s.footprint() for s in solvers[:2])
    fs = max(s.footprint() for s in solvers[2:])
    budget = max(s.footprint() for s in solvers[:2]) + 2 * fs + fs // 2
    assert budget < 2 * fb, "the two large requests must not fit together"
    res, summary = M.serve(solvers, [0.0] * 4, slots=4, mem_budget=budget)
    assert all(s.status == M.Status.SOLVE_SUCCEEDED for s in solvers)
    big2 = res[1]
    assert big2.admitted >= res[0].done - 1e-3          # it waited for memory
    assert all(r.admitted < big2.admitted for r in res[2:])  # the small ones went ahead
