"""The device path on the benchmark workloads at their own sizes, against the
FP64 oracle's committed traces (tests/golden/config_*.npz, written by
tests/golden/make_config_traces.py).

* max-cut G(3000, 0.01), plain FP64: same status and iteration count, every
  per-iteration objective and residual within 1e-9 relative.
* dense 250-block + 400 scalars, m = 8000, and 96 blocks of 8..120 + 200
  scalars, m = 6000: the same for plain FP64; the mixed paths (FP32 MFMA or
  BF16x3 split-product assembly, FP64 refinement) reach the same status in at
  most 2 extra iterations, their final objective, residuals and y are stated
  against the oracle's.
* the first 24 requests of the serving stream: the plain FP64 path matches
  the oracle per request (status, iterations, trace); served concurrently
  (continuous batching) each result equals the same instance solved alone
  bit for bit; the mixed batched path stays within 2 extra iterations.

Each instance is regenerated from its seed; its SHA-256 digest must equal the
one stored with the trace.
"""
import os
import sys

import numpy as np
import pytest

import madsdp_amd as M
from madsdp_amd import instances, serve

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_config_traces import digest, stream_models  # noqa: E402

pytestmark = pytest.mark.gpu
RTOL = 1e-9


def _load(name):
    return np.load(os.path.join(HERE, "golden", f"config_{name}.npz"))


def assert_trace(tr, ref, rtol=RTOL):
    assert tr.shape[0] == ref.shape[0], (tr.shape, ref.shape)
    for k in range(ref.shape[0]):
        scale = max(1.0, abs(ref[k, 0]), abs(ref[k, 1]))
        assert abs(tr[k, 0] - ref[k, 0]) <= rtol * scale, ("pobj", k, tr[k, 0], ref[k, 0])
        assert abs(tr[k, 1] - ref[k, 1]) <= rtol * scale, ("dobj", k, tr[k, 1], ref[k, 1])
        for col, name in ((2, "inf_pr"), (3, "inf_du"), (4, "gap")):
            assert abs(tr[k, col] - ref[k, col]) <= rtol * max(1.0, ref[k, col]), (name, k)


def check_plain(model, z):
    assert digest(model) == str(z["digest"]), "regenerated instance differs from the traced one"
    s = M.MadSDPSolver(model).solve()
    print(f"plain: status {int(s.status)}, {s.iter} it (oracle {int(z['status'])}, {int(z['iter'])})",
          flush=True)
    assert int(s.status) == int(z["status"])
    assert s.iter == int(z["iter"])
    assert_trace(s.trace(), z["trace"])
    np.testing.assert_allclose(s.y, z["y"], rtol=1e-6, atol=1e-8 * max(1.0, np.abs(z["y"]).max()))
    return s


def check_mixed(model, z, precision="mixed", **kw):
    s = M.MadSDPSolver(model, precision=precision, **kw).solve()
    info = s.info()
    ref = z["trace"]
    tr = s.trace()
    e_obj = abs(tr[-1, 0] - ref[-1, 0]) / max(1.0, abs(ref[-1, 0]))
    e_y = np.abs(s.y - z["y"]).max() / max(1.0, np.abs(z["y"]).max())
    print(f"{precision}: status {int(s.status)}, {s.iter} it (oracle {int(z['iter'])}), "
          f"low-precision assemblies {info.lowp_factorizations}, refinement steps {info.refine_steps}, "
          f"FP64 fallbacks {info.fp64_fallbacks}; vs oracle: objective {e_obj:.2e}, "
          f"inf_pr {tr[-1, 2]:.2e} (oracle {ref[-1, 2]:.2e}), inf_du {tr[-1, 3]:.2e} "
          f"(oracle {ref[-1, 3]:.2e}), gap {tr[-1, 4]:.2e}, y {e_y:.2e}", flush=True)
    assert int(s.status) == int(z["status"]) == int(M.Status.SOLVE_SUCCEEDED)
    assert s.iter <= int(z["iter"]) + 2
    assert max(tr[-1, 2], tr[-1, 3], tr[-1, 4]) <= 1e-7
    assert e_obj <= 1e-6
    assert info.lowp_factorizations > 0
    return s


def test_maxcut_3000_plain(gpu):
    """G(3000, 0.01) with +-1 weights, seed 1."""
    print("maxcut_3000 plain", flush=True)
    check_plain(instances.maxcut_3000(), _load("maxcut_3000"))


@pytest.fixture(scope="module")
def dense_250():
    return instances.dense_250_m8000()


def test_dense_250_plain(gpu, dense_250):
    print("dense_250 plain", flush=True)
    check_plain(dense_250, _load("dense_250"))


def test_dense_250_mixed(gpu, dense_250):
    print("dense_250 mixed", flush=True)
    check_mixed(dense_250, _load("dense_250"))


def test_dense_250_bf16x3(gpu, dense_250):
    print("dense_250 bf16x3", flush=True)
    check_mixed(dense_250, _load("dense_250"), precision="bf16x3")


@pytest.fixture(scope="module")
def blocks_96():
    return instances.blocks_96_m6000()


def test_blocks_96_plain(gpu, blocks_96):
    print("blocks_96 plain", flush=True)
    check_plain(blocks_96, _load("blocks_96"))


def test_blocks_96_mixed(gpu, blocks_96):
    print("blocks_96 mixed (FP32 assembly forced)", flush=True)
    check_mixed(blocks_96, _load("blocks_96"), lowp_min_gflop=0.0)


def test_blocks_96_bf16x3(gpu, blocks_96):
    """Constraints touching a subset of the blocks: the panel operator gathers
    and scatters by the block's constraint list."""
    print("blocks_96 bf16x3 (forced)", flush=True)
    check_mixed(blocks_96, _load("blocks_96"), precision="bf16x3", lowp_min_gflop=0.0)


def test_stream_prefix(gpu):
    z = _load("stream_prefix")
    count = int(z["count"])
    items = stream_models(count)
    models = [m for _, _, m in items]
    for k, m in enumerate(models):
        assert digest(m) == str(z[f"{k}_digest"]), k
    print(f"stream prefix: {count} requests, plain FP64 one at a time", flush=True)
    plain = [M.MadSDPSolver(m) for m in models]
    serve(plain, [0.0] * count, slots=1)
    for k, s in enumerate(plain):
        assert int(s.status) == int(z[f"{k}_status"]), (k, items[k][:2])
        assert s.iter == int(z[f"{k}_iter"]), (k, items[k][:2], s.iter, int(z[f"{k}_iter"]))
        assert_trace(s.trace(), z[f"{k}_trace"])
    print("served concurrently (8 slots), FP64", flush=True)
    conc = [M.MadSDPSolver(m) for m in models]
    serve(conc, [0.002 * k for k in range(count)], slots=8)
    for k in range(count):
        a, b = plain[k], conc[k]
        assert a.iter == b.iter and int(a.status) == int(b.status), k
        assert np.array_equal(a.trace(), b.trace()), k
        assert np.array_equal(a.y, b.y), k
    print("served concurrently (8 slots), mixed", flush=True)
    mixed = [M.MadSDPSolver(m, precision="mixed") for m in models]
    serve(mixed, [0.002 * k for k in range(count)], slots=8)
    for k in range(count):
        assert int(mixed[k].status) == int(z[f"{k}_status"]), k
        assert mixed[k].iter <= int(z[f"{k}_iter"]) + 2, k
