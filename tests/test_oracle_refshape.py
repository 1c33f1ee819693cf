"""The reference-shaped CPU restatement (oracle: orc_refshape_*), which bench.py times as a second
CPU baseline: AoS nodes of 56-byte states and per-node OpenMP passes, as grid/grid.cpp and
grid/grid_node.cpp arrange the work.  It must compute exactly what the oracle's batch path
computes: the same RK3 step bit for bit, and the same CFL maximum."""
import numpy as np
import pytest

import orc
from gpu_util import amrhydro  # noqa: F401  (puts the package on the path)
import workloads


@pytest.mark.parametrize("n", [8, 16])
def test_refshape_step_equals_the_batch_oracle(n):
    cfg = orc.euler_cfg(n=n)
    U = workloads.flat_batch(6, n=n, seed=21)
    h = np.full(U.shape[0], 1.0 / n)
    orc.outflow_ghosts(cfg, U)
    m = orc.batch_max_eig(cfg, U)
    R = orc.RefShaped(cfg, U, h)
    try:
        assert R.max_eig_over_h() == (m / h).max()
        dt = 0.4 * (h / m).min()
        ref = U.copy()
        orc.rk_steps(cfg, ref, h, dt, 3, 1)
        R.rk_step(dt, 3)
        np.testing.assert_array_equal(R.download(), ref)
    finally:
        R.close()
