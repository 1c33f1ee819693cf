"""GPU octree path against the oracle's GridNode restatement.

Config 1: Sedov-type blast in the unit cube (ideal gas, rho = 1, p = 1e-5, energy 1 in a sphere
of 2 finest cells), outflow on all faces, 2 refinement levels of 8^3 blocks (73 nodes, fully
refined), 10 RK substeps (3 RK3 steps + 1), compared after every substep at <= 1e-12 relative L1
per field over the leaf cells; mass and energy ledger conserved to roundoff.  A partially refined
tree with coarse/fine faces inside the domain checks the coarse->fine ghosts, (nested) reflux and
restriction the same way."""
import numpy as np
import pytest

import orc
import tree_cases as tc
from gpu_util import amrhydro, say

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_l1_leaves(a, b, leaves, g=2):
    A = a[leaves][:, :, g:-g, g:-g, g:-g]
    B = b[leaves][:, :, g:-g, g:-g, g:-g]
    return np.array([np.abs(A[:, f] - B[:, f]).sum() / np.abs(B[:, f]).sum() for f in range(A.shape[1])])


def gpu_tree(cfg_o, parent, child, U):
    cfg = amrhydro.euler_config(n=cfg_o.n, g=cfg_o.g, grid_dim=cfg_o.grid_dim, z_reflect=cfg_o.z_reflect,
                                recon=amrhydro.AMRH_RECON_PPM if cfg_o.recon == 1 else amrhydro.AMRH_RECON_MINMOD)
    B = amrhydro.Batch(cfg, len(parent), 1.0)
    B.tree_set(parent, child)
    B.upload(U)
    return B


def run_and_compare(cfg, spec, ic, nsub, label):
    T, parent, child = tc.build(cfg, spec)
    ic(T)
    T.enforce_boundaries()
    T.inject_from_children()
    B = gpu_tree(cfg, parent, child, tc.padded_state(T))
    B.tree_fill_ghosts()
    leaves = tc.leaf_mask(parent)
    h = np.array([T.info(i)[3] for i in range(T.nnodes)])
    s0 = tc.state_sum(tc.padded_state(T), h, parent, child, cfg.n, cfg.g) + T.flow_off()
    betas = (1.0, 0.25, 2.0 / 3.0)
    dt = None
    done = 0
    while done < nsub:
        # HydroDriver::step_to_time's dt (euler/hydro_driver.cpp:107-113) on the oracle tree
        T.enforce_boundaries()
        m = T.max_dt()
        mg = B.tree_max_dt()
        assert abs(mg - m) <= 1e-14 * m, (mg, m)
        dt = 1e-2 * m if dt is None else min(1.25 * dt, m)
        T.store()
        B.store()
        for beta in betas:
            if done == nsub:
                break
            T.substep(dt, beta)
            B.tree_substep(dt, beta)
            done += 1
            out = B.download()
            ref = tc.padded_state(T)
            e = rel_l1_leaves(out, ref, leaves, cfg.g)
            say("%s substep %d beta=%.3f rel-L1 %s" % (label, done, beta, np.array2string(e, precision=2)))
            assert np.all(e <= TOL), e
        B.tree_fill_ghosts()
    fo_g = B.tree_flow_off()
    fo_o = T.flow_off()
    np.testing.assert_allclose(fo_g, fo_o, rtol=1e-12, atol=1e-15 * np.abs(fo_o).max())
    out = B.download()
    s1 = tc.state_sum(out, h, parent, child, cfg.n, cfg.g) + fo_g
    for f in (0, 4):
        say("%s field %d: sum+flow_off %.17g -> %.17g" % (label, f, s0[f], s1[f]))
        assert abs(s1[f] - s0[f]) <= 1e-13 * abs(s0[f])
    assert B.error_flags() == 0


def test_config1_sedov_tree_matches_oracle():
    cfg = orc.euler_cfg(n=8, g=2, grid_dim=0.5, z_reflect=0)
    run_and_compare(cfg, tc.full_spec(2), lambda T: tc.sedov_ic(T, 2), 10, "sedov")


def test_partial_tree_reflux_matches_oracle():
    cfg = orc.euler_cfg(n=8, g=2, grid_dim=0.5, z_reflect=0)
    run_and_compare(cfg, tc.partial_spec(), tc.blob_ic, 6, "partial")


def test_partial_tree_z_reflect_matches_oracle():
    cfg = orc.euler_cfg(n=8, g=2, grid_dim=0.5, z_reflect=1)
    run_and_compare(cfg, tc.partial_spec(), tc.blob_ic, 3, "zreflect")


def test_root_only_tree_matches_oracle():
    """A tree of one node: every face is physical, nothing to reflux or restrict."""
    cfg = orc.euler_cfg(n=8, g=2, grid_dim=0.5, z_reflect=0)
    run_and_compare(cfg, [], tc.blob_ic, 6, "root")


@pytest.mark.parametrize("z_reflect", [0, 1])
def test_partial_tree_ppm_matches_oracle(z_reflect):
    """PPM (g = 3) on the octree: the block kernel level by level with its reflux exports"""
    cfg = orc.euler_cfg(n=8, g=3, recon=1, grid_dim=0.5, z_reflect=z_reflect)
    run_and_compare(cfg, tc.partial_spec(), tc.blob_ic, 6, "ppm zr=%d" % z_reflect)


def test_config1_sedov_tree_ppm_matches_oracle():
    cfg = orc.euler_cfg(n=8, g=3, recon=1, grid_dim=0.5, z_reflect=0)
    run_and_compare(cfg, tc.full_spec(2), lambda T: tc.sedov_ic(T, 2), 6, "sedov ppm")


def test_n16_tree_matches_oracle():
    """16^3 blocks on the octree: one refined level, coarse/fine faces at the root's children."""
    cfg = orc.euler_cfg(n=16, g=2, grid_dim=0.5, z_reflect=0)
    run_and_compare(cfg, [(0, c) for c in range(8)] + [(1, c) for c in range(8)], tc.blob_ic, 3, "n16")
