"""Diagnostics as device reductions (GridNode::state_sum / state_max / state_min,
grid/grid_node.cpp:1489-1589) through the C ABI, against numpy over the same cells: flat batches
(every cell uncovered, positions from the block origins), the rotating 13-field batch, and an
octree (covered octants excluded, positions from the tree); and the conservation ledger of the
octree measured with them on the device."""
import numpy as np
import pytest

import cases
import orc
import tree_cases as tc
from gpu_util import amrhydro, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def numpy_diag(U, g, h, x0, covered=None):
    """(4, nf) moments, max, min over uncovered interior cells."""
    nb, nf, G = U.shape[0], U.shape[1], U.shape[2]
    n = G - 2 * g
    I = U[:, :, g:-g, g:-g, g:-g]
    sums = np.zeros((4, nf))
    mx = np.full(nf, -1e99)
    mn = np.full(nf, 1e99)
    c = np.arange(n) + 0.5
    for b in range(nb):
        m = np.ones((n, n, n), bool)
        if covered is not None:
            for ci in range(8):
                if covered[b] >> ci & 1:
                    cx, cy, cz = ci & 1, ci >> 1 & 1, ci >> 2 & 1
                    h2 = n // 2
                    m[cz * h2:(cz + 1) * h2, cy * h2:(cy + 1) * h2, cx * h2:(cx + 1) * h2] = False
        z, y, x = np.meshgrid(x0[b, 2] + c * h[b], x0[b, 1] + c * h[b], x0[b, 0] + c * h[b], indexing="ij")
        ds = I[b][:, m] * h[b] ** 3
        sums[0] += ds.sum(axis=1)
        sums[1] += (ds * x[m]).sum(axis=1)
        sums[2] += (ds * y[m]).sum(axis=1)
        sums[3] += (ds * z[m]).sum(axis=1)
        if m.any():
            mx = np.maximum(mx, I[b][:, m].max(axis=1))
            mn = np.minimum(mn, I[b][:, m].min(axis=1))
    return sums, mx, mn


def check(B, U, g, h, x0, covered=None):
    s, mx, mn = numpy_diag(U, g, h, x0, covered)
    got = B.state_sum()
    np.testing.assert_allclose(got, s, rtol=1e-12, atol=1e-14 * np.abs(s).max())
    np.testing.assert_array_equal(B.state_max(), mx)
    np.testing.assert_array_equal(B.state_min(), mn)


def test_flat_batch_diagnostics():
    n, nb = 16, 24
    U = cases.c2_fast(nb, n=n, seed=41)
    rng = np.random.default_rng(42)
    h = rng.choice([1 / 16, 1 / 32, 1 / 64], nb)
    x0 = rng.integers(-8, 8, (nb, 3)) * h[:, None] * n
    B = amrhydro.Batch(amrhydro.euler_config(n=n), nb, h)
    B.upload(U)
    with pytest.raises(amrhydro.AmrhError):
        B.state_sum()  # moments need positions
    B.set_block_origin(x0)
    check(B, U, 2, h, x0)


def test_rotating_batch_diagnostics():
    import rot_cases as rc
    Bn = rc.batch(counts=(8, 24), seed=43)
    U, h, x0 = Bn["U"], Bn["h"], Bn["x0"]
    cfg = amrhydro.rotating_config(omega=Bn["omega"])
    B = amrhydro.Batch(cfg, U.shape[0], h)
    B.upload(U)
    B.set_block_origin(x0)
    check(B, U, 3, h, x0)


def test_octree_diagnostics_and_ledger():
    cfg_o = orc.euler_cfg(n=8, g=2, grid_dim=0.5, z_reflect=0)
    T, parent, child = tc.build(cfg_o, tc.partial_spec())
    tc.blob_ic(T)
    T.enforce_boundaries()
    T.inject_from_children()
    U = tc.padded_state(T)
    nn = len(parent)
    h = np.array([T.info(b)[3] for b in range(nn)])
    x0 = np.zeros((nn, 3))
    for b in range(nn):
        X, Y, Z, hb = tc.cell_centres(T, b)
        x0[b] = [X[0, 0, 2] - 0.5 * hb, Y[0, 2, 0] - 0.5 * hb, Z[2, 0, 0] - 0.5 * hb]
    covered = np.zeros(nn, np.int64)
    for b in range(1, nn):
        covered[parent[b]] |= 1 << child[b]
    cfg = amrhydro.euler_config(n=8, grid_dim=0.5)
    B = amrhydro.Batch(cfg, nn, 1.0)
    B.tree_set(parent, child)
    B.upload(U)
    check(B, U, 2, h, x0, covered)
    # mass and energy: uncovered state sum + the root's flow-off ledger stay constant over steps
    total0 = B.state_sum()[0] + B.tree_flow_off()
    dt = 2e-4
    for _ in range(3):
        B.tree_step(dt, 3)
    total = B.state_sum()[0] + B.tree_flow_off()
    rel = np.abs(total - total0) / np.abs(total0)
    say("octree conservation after 3 RK3 steps: rel change rho %.1e, E %.1e" % (rel[0], rel[4]))
    assert rel[0] < 1e-13 and rel[4] < 1e-13
