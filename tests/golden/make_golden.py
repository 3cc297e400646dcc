"""Writes the golden fixtures: small problems (inputs) with the FP64 oracle's
per-iteration trace, final status, iteration count and solution (outputs).

    python tests/golden/make_golden.py

The cases are the known-answer problems of the reference's tests
(test/runtests.jl: trivial, trace-one, LRO max-cut -> -18, the 4-cycle -> -4,
the mixed PSD + scalar problem -> 1; test/qr_schur.jl's 3 x 3 example with two
scalars), the seeded max-cut G(100, 0.1) and a small seeded block-diagonal
problem.  The fixtures pin the oracle against regressions (tests/test_golden.py)
and give the device path a committed expected trace (tests/test_gpu_golden.py).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "madsdp.jl_amd"), os.path.dirname(HERE)]

import kat_models as K  # noqa: E402
from madsdp_amd import instances  # noqa: E402
from oracle import madsdp_oracle as O  # noqa: E402

FIELDS = ("pobj", "dobj", "inf_pr", "inf_du", "gap", "mu", "alpha", "beta", "sigma")


def cases():
    return {
        "trivial": (K.trivial(), dict(tol=1e-8, max_iter=60)),
        "trace_one": (K.trace_one(), dict(tol=1e-8, max_iter=60)),
        "maxcut_lro": (K.maxcut(K.W_LRO), dict(max_iter=100)),
        "cycle": (K.maxcut(K.W_CYCLE), dict(tol=1e-7, max_iter=80)),
        "mixed": (K.mixed(), dict(max_iter=100)),
        "qr_example_2": (K.qr_example(2), {}),
        "maxcut_100": (instances.maxcut_100(), {}),
        "blocks_small": (instances.block_sdp(instances.block_sizes(4, 3, 12, 5), 24, 4, 3), {}),
    }


def pack(model, prefix="m_"):
    out = {prefix + "b": model.b, prefix + "d_lin": model.d_lin, prefix + "C_lin": model.C_lin,
           prefix + "nblocks": np.array(model.nblocks)}
    for i, (C, A) in enumerate(zip(model.C, model.A)):
        p = f"{prefix}{i}_"
        out[p + "C"] = C
        out[p + "kind"] = np.array(A.kind)
        if A.kind == "dense":
            out[p + "con_idx"], out[p + "mats"] = A.con_idx, A.mats
        else:
            out[p + "con"], out[p + "row"], out[p + "col"], out[p + "val"] = A.con, A.row, A.col, A.val
    return out


def unpack(z, prefix="m_"):
    from madsdp_amd.model import BlockConstraints, Model
    Cs, As = [], []
    for i in range(int(z[prefix + "nblocks"])):
        p = f"{prefix}{i}_"
        C = z[p + "C"]
        n = C.shape[0]
        if str(z[p + "kind"]) == "dense":
            A = BlockConstraints("dense", n, con_idx=z[p + "con_idx"], mats=z[p + "mats"])
        else:
            A = BlockConstraints("coo", n, con=z[p + "con"], row=z[p + "row"], col=z[p + "col"],
                                 val=z[p + "val"])
        Cs.append(C)
        As.append(A)
    return Model(Cs, As, z[prefix + "b"], z[prefix + "d_lin"], z[prefix + "C_lin"])


def trace_array(solver):
    return np.array([[getattr(r, f) for f in FIELDS] for r in solver.trace])


def main():
    for name, (model, opts) in cases().items():
        s = O.Solver(model, **opts).solve()
        rec = pack(model)
        rec.update(trace=trace_array(s), status=np.array(int(s.status)), iter=np.array(s.iter),
                   y=s.y, X0=s.blocks[0].X if s.blocks else np.zeros((0, 0)),
                   opts_tol=np.array(opts.get("tol", 1e-7)),
                   opts_max_iter=np.array(opts.get("max_iter", 100)))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **rec)
        print(f"{name}: status {int(s.status)}, {s.iter} iterations")


if __name__ == "__main__":
    main()
