"""The oracle against its committed golden traces (tests/golden/*.npz, written
by tests/golden/make_golden.py): a change to the oracle that moves any
iterate shows up here before it can move a parity test's reference."""
import glob
import os
import sys

import numpy as np
import pytest

from oracle import madsdp_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import FIELDS, trace_array, unpack  # noqa: E402

FILES = sorted(f for f in glob.glob(os.path.join(HERE, "golden", "*.npz"))
               if not os.path.basename(f).startswith("config_"))


def load(path):
    z = np.load(path)
    opts = dict(tol=float(z["opts_tol"]), max_iter=int(z["opts_max_iter"]))
    return z, unpack(z), opts


def test_fixtures_present():
    assert len(FILES) >= 8


@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_golden(path):
    z, model, opts = load(path)
    s = O.Solver(model, **opts).solve()
    assert int(s.status) == int(z["status"])
    assert s.iter == int(z["iter"])
    np.testing.assert_allclose(trace_array(s), z["trace"], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(s.y, z["y"], rtol=1e-8, atol=1e-10)
    assert FIELDS[0] == "pobj"


CONFIG = sorted(glob.glob(os.path.join(HERE, "golden", "config_*.npz")))


@pytest.mark.parametrize("path", [p for p in CONFIG if "dense_250" not in p],
                         ids=lambda p: os.path.basename(p)[7:-4])
def test_config_traces_match_regenerated_instances(path):
    """The workload traces were made from the instances the generators still
    produce (same digest), and each ends converged."""
    from make_config_traces import config_models, digest, stream_models
    z = np.load(path)
    name = os.path.basename(path)[7:-4]
    if name == "stream_prefix":
        items = stream_models(int(z["count"]))
        keys = [f"{k}_" for k in range(len(items))]
        models = [m for _, _, m in items]
    else:
        keys = [""]
        models = [config_models()[name][0]()]
    for key, model in zip(keys, models):
        assert digest(model) == str(z[key + "digest"])
        tr = z[key + "trace"]
        assert int(z[key + "status"]) == int(O.Status.SOLVE_SUCCEEDED)
        assert tr.shape[0] == int(z[key + "iter"]) + 1
        assert max(tr[-1, 2], tr[-1, 3], tr[-1, 4]) <= 1e-7
