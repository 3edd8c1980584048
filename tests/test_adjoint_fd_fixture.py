"""The sixteen-voxel finite-difference fixture of the adjoint gradient check
(tests/golden/adjoint_fd16.npz, written by tests/golden/make_adjoint_fd.py),
on the CPU:

* its h = 1e-2 and h = 2e-2 central differences agree to 5e-4 on every
  voxel, so the h = 2e-2 values the GPU test compares with are not limited
  by truncation or by the rounding of the FP32 coefficient tables;
* one voxel's h = 2e-2 difference is recomputed from two FP64 oracle runs
  and equals the stored one, and the stepped permittivities are the FP32
  values of eps +- h.
"""
import os

import numpy as np

from tests import test_gpu_adjoint as T
from alpineem_amd.solver import Plan


def _fixture(golden_dir):
    return np.load(os.path.join(golden_dir, "adjoint_fd16.npz"))


def test_step_sizes_agree(golden_dir):
    fx = _fixture(golden_dir)
    g = fx["g_fd"]
    assert g.shape == (3, 16) and list(fx["h"]) == [1e-3, 1e-2, 2e-2]
    rel = np.abs(g[2] / g[1] - 1)
    print("h=2e-2 against h=1e-2:", np.array2string(rel, precision=2))
    assert rel.max() <= 5e-4
    # h = 1e-3 is limited by the FP32 tables' rounding, at the 1e-3 level
    assert np.abs(g[0] / g[1] - 1).max() <= 5e-3
    cells = [tuple(int(x) for x in c) for c in fx["cells"]]
    assert len(set(cells)) == 16


def test_one_voxel_recomputed(golden_dir):
    fx = _fixture(golden_dir)
    lo, hi = tuple(int(x) for x in fx["box_lo"]), tuple(int(x) for x in fx["box_hi"])
    eps = T.design_eps(5, lo, hi)
    np.testing.assert_array_equal(eps, fx["eps"])
    q = 3
    c = tuple(int(x) for x in fx["cells"][q])
    h = float(fx["h"][2])
    J = []
    for s, sgn in enumerate((1, -1)):
        e = eps.copy()
        e[c] = np.float32(e[c] + sgn * h)
        assert float(e[c]) == float(fx["eps_pm"][2, q, s])
        plan = Plan.from_text(T.design_deck(e, lo, hi).text())
        J.append(abs(T.oracle_objective(plan, True, T.THETA, T.PHI, T.FREQ)) ** 2)
    np.testing.assert_allclose(J, fx["J_pm"][2, q], rtol=1e-12)
