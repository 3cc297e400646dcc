"""Writes the FP64 oracle's traces of the benchmark workloads at their own sizes.

    python tests/golden/make_config_traces.py [name ...]

The instances are regenerated from their seeds (madsdp_amd.instances) on any
machine, so only the oracle's outputs are stored, keyed by the seed and a
SHA-256 digest of the regenerated problem data:

* ``maxcut_3000``  -- max-cut G(3000, 0.01), +-1 weights, seed 1
* ``dense_250``    -- one 250 x 250 dense block + 400 scalars, m = 8000, seed 2
* ``blocks_96``    -- 96 blocks of 8..120 + 200 scalars, m = 6000, seed 3
* ``stream_prefix``-- the first STREAM_COUNT instances of the seed-4 serving stream

Each record holds the per-iteration trace (pobj, dobj, inf_pr, inf_du, gap,
mu, alpha, beta, sigma), the status, the iteration count and the final y.
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "madsdp.jl_amd"), HERE]

from madsdp_amd import instances  # noqa: E402

STREAM_COUNT = 24


def digest(model) -> str:
    """SHA-256 of a model's problem data (C, A, b, d_lin, C_lin)."""
    h = hashlib.sha256()
    for C, A in zip(model.C, model.A):
        h.update(np.ascontiguousarray(C).tobytes())
        if A.kind == "dense":
            h.update(np.ascontiguousarray(A.con_idx).tobytes())
            h.update(np.ascontiguousarray(A.mats).tobytes())
        else:
            for a in (A.con, A.row, A.col, A.val):
                h.update(np.ascontiguousarray(a).tobytes())
    for a in (model.b, model.d_lin, model.C_lin):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def config_models():
    return {
        "maxcut_3000": (lambda: instances.maxcut_3000(), 1),
        "dense_250": (lambda: instances.dense_250_m8000(), 2),
        "blocks_96": (lambda: instances.blocks_96_m6000(), 3),
    }


def stream_models(count=STREAM_COUNT):
    stream = instances.serving_stream(count=count, seed=4)
    return [(kind, params, instances.build_stream_instance(kind, params))
            for _, kind, params in stream]


def solve_record(model):
    from oracle import madsdp_oracle as O
    from make_golden import trace_array
    t = time.perf_counter()
    s = O.Solver(model).solve()
    return dict(trace=trace_array(s), status=np.array(int(s.status)), iter=np.array(s.iter),
                y=s.y.copy(), digest=np.array(digest(model)),
                oracle_seconds=np.array(time.perf_counter() - t))


def main(names):
    cfg = config_models()
    for name in names:
        if name == "stream_prefix":
            recs = {}
            for k, (kind, params, model) in enumerate(stream_models()):
                r = solve_record(model)
                for key, v in r.items():
                    recs[f"{k}_{key}"] = v
                print(f"stream {k} {kind} {params}: status {int(r['status'])}, {int(r['iter'])} it, "
                      f"{float(r['oracle_seconds']):.1f} s", flush=True)
            recs["count"] = np.array(STREAM_COUNT)
            np.savez_compressed(os.path.join(HERE, "config_stream_prefix.npz"), **recs)
            continue
        make, seed = cfg[name]
        model = make()
        r = solve_record(model)
        r["seed"] = np.array(seed)
        np.savez_compressed(os.path.join(HERE, f"config_{name}.npz"), **r)
        print(f"{name}: status {int(r['status'])}, {int(r['iter'])} iterations, "
              f"{float(r['oracle_seconds']):.1f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["blocks_96", "maxcut_3000", "stream_prefix", "dense_250"])
