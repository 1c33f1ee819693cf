"""The synthetic workload generators the benchmark and the GPU tests draw from: shapes, and the
tree ones consistent with the octree's own geometry."""
import numpy as np

import orc
from gpu_util import amrhydro  # noqa: F401  (puts the package on the path)
import workloads


def test_sedov_tree_shapes():
    U, parent, child = workloads.sedov_tree(maxlevel=1)
    assert U.shape == (9, 6, 12, 12, 12) and list(parent) == [-1] + [0] * 8 and list(child) == [-1] + list(range(8))
    assert np.all(U[:, 0] == 1.0)


def test_binary_tree_state_matches_the_tree_geometry():
    parent, child, lv, loc = workloads.binary_tree(2)
    S = workloads.binary_tree_state(lv, loc)
    T = orc.RotTree(orc.rot_cfg(omega=S["omega"]), parent, child)
    for b in range(len(parent)):
        x0, h = T.x0(b)
        assert np.array_equal(x0, S["x0"][b]) and h == S["h"][b]
    assert S["U"].shape == (len(parent), 13, 14, 14, 14) and S["phi"].shape == (len(parent), 10, 10, 10)
    assert np.all(S["U"][:, 0] > 0)
    src = workloads.binary_source(lv, loc)
    np.testing.assert_allclose(src[:, 1:-1, 1:-1, 1:-1], 4 * np.pi * S["U"][:, 0, 3:-3, 3:-3, 3:-3], rtol=1e-14)


def test_multilevel_and_flat_batches():
    assert workloads.flat_batch(4, n=8, g=2).shape == (4, 6, 12, 12, 12)
    lev, org = workloads.multilevel_layout(counts=(8, 60))
    assert len(lev) == 68 and org.shape == (68, 3)
