"""The device path against the committed golden traces (no oracle run needed):
same status and iteration count, per-iteration primal objective to 1e-9
relative, final y to 1e-6 relative."""
import glob
import os
import sys

import numpy as np
import pytest

import madsdp_amd as M

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import unpack  # noqa: E402

pytestmark = pytest.mark.gpu
FILES = sorted(f for f in glob.glob(os.path.join(HERE, "golden", "*.npz"))
               if not os.path.basename(f).startswith("config_"))


@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_device_matches_golden(gpu, path):
    z = np.load(path)
    model = unpack(z)
    print(f"golden {os.path.basename(path)}", flush=True)
    s = M.MadSDPSolver(model, tol=float(z["opts_tol"]), max_iter=int(z["opts_max_iter"])).solve()
    assert int(s.status) == int(z["status"])
    assert s.iter == int(z["iter"])
    tr, ref = s.trace(), z["trace"]
    assert tr.shape == ref.shape
    for k in range(ref.shape[0]):
        assert abs(tr[k, 0] - ref[k, 0]) <= 1e-9 * max(1.0, abs(ref[k, 0])), k
    np.testing.assert_allclose(s.y, z["y"], rtol=1e-6, atol=1e-8)
