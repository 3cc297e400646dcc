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

FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


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
