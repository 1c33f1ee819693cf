"""Grid::restore (grid/grid.cpp:187-199, grid/grid_node.cpp:486-493) and BinaryDriver::step's dt
rejection (REDO, binary/binary_driver.cpp:414-423) through the C ABI; captured step graphs that
follow a new tree; GridNode::set_origin (grid/grid_node.cpp:20-26) in the rotating state's positions
(Grid::xf, grid/grid.cpp:512-524) and through checkpoints."""
import numpy as np
import pytest

import orc
import tree_cases as tc
from gpu_util import amrhydro, say
from test_gpu_rot_tree import TOL, device_batch
from test_gpu_rotating import rel_l1_rot
from test_oracle_rot_tree import binary_tree_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAXDTINC = 1.25  # binary/binary_defs.h
BETAS = (1.0, 0.25, 2.0 / 3.0)


def euler_tree(spec):
    cfg = orc.euler_cfg(n=8, g=2, grid_dim=0.5)
    T, parent, child = tc.build(cfg, spec)
    tc.sedov_ic(T, 2)
    for b in range(T.nnodes):  # a uniform drift in x, so mass and energy cross the domain boundary
        U = T.U(b)
        U[4] += 0.5 * 0.04 * U[0]
        U[1] = 0.2 * U[0]
    T.enforce_boundaries()
    T.inject_from_children()
    d = amrhydro.euler_config(n=8, g=2, grid_dim=0.5)
    B = amrhydro.Batch(d, len(parent), 1.0)
    B.tree_set(parent, child)
    B.upload(tc.padded_state(T))
    return T, B, parent, child


def test_euler_tree_restore_returns_the_stored_state_and_ledger():
    """store, three tree substeps (ledger updated at the root), restore: U and FO are back bit for
    bit, and the restored state steps on exactly as the stored one did"""
    T, B, parent, child = euler_tree(tc.partial_spec())
    B.tree_fill_ghosts()
    dt = 0.5 * B.tree_max_dt()
    U_s, fo_s = B.download(), B.tree_flow_off()
    B.store()
    for beta in BETAS:
        B.tree_substep(dt, beta)
    U_1, fo_1 = B.download(), B.tree_flow_off()
    assert np.any(fo_1 != fo_s) and np.any(U_1 != U_s)
    B.restore()
    np.testing.assert_array_equal(B.download(), U_s)
    np.testing.assert_array_equal(B.tree_flow_off(), fo_s)
    # the step again from the restored state: the same bits as the first time
    for beta in BETAS:
        B.tree_substep(dt, beta)
    np.testing.assert_array_equal(B.download(), U_1)
    np.testing.assert_array_equal(B.tree_flow_off(), fo_1)
    B.close()


def test_restore_needs_a_store():
    T, B, parent, child = euler_tree(tc.partial_spec())
    with pytest.raises(amrhydro.AmrhError, match="amrh_store"):
        B.restore()
    B.close()


def test_rotating_redo_step_matches_the_oracle():
    """BinaryDriver::step with a first dt 1.6x above the CFL limit: the three substeps run, the new
    CFL dt is below dt/MAXDTINC, restore, dt /= MAXDTINC, again, until the step is accepted; the
    accepted step equals the oracle's step at that dt from the stored state"""
    T, cfg = binary_tree_case("partial", seed=31)
    B = device_batch(T, cfg)
    T.fill_ghosts()
    B.tree_fill_ghosts()
    dt = 1.6 * B.tree_max_dt()
    B.store()
    redo = 0
    while True:
        for beta in BETAS:
            B.tree_substep(dt, beta)
            B.tree_inject()
        B.tree_fill_ghosts()
        test_dt = B.tree_max_dt()
        if test_dt < dt / MAXDTINC:
            say("  REDO %d: dt %.4e too big (CFL now %.4e)" % (redo, dt, test_dt))
            dt = dt / MAXDTINC
            B.restore()
            redo += 1
            assert redo < 10
            continue
        break
    assert redo >= 1
    B.error_flags()  # the rejected attempts may have flagged states; the accepted one is checked below
    T.store()
    for beta in BETAS:
        assert T.substep(dt, beta) == 0
        assert T.inject() == 0
    T.fill_ghosts()
    e = rel_l1_rot(B.download(), T.state(), cfg.g)
    say("rotating REDO: %d rejections, accepted dt %.4e, rel-L1 max %.1e" % (redo, dt, e.max()))
    assert np.all(e <= TOL), e
    fo, fo_ref = B.tree_flow_off(), T.flow_off()
    np.testing.assert_allclose(fo, fo_ref, rtol=1e-11, atol=1e-14 * np.abs(fo_ref).max())
    B.close()


def test_graph_replays_follow_a_new_tree():
    """run_steps as graphs on one tree, then amrh_tree_set with another topology of the same node
    count: the replays must use the new tree's tables (not the freed ones) and equal the direct path"""
    spec_a = tc.full_spec(1)                                          # root + 8 children
    spec_b = [(0, c) for c in range(4)] + [(1, c) for c in range(4)]  # root + 4, child 0 + 4
    assert len(spec_a) == len(spec_b)
    T_a, B, parent_a, child_a = euler_tree(spec_a)
    T_b, D, parent_b, child_b = euler_tree(spec_b)
    U_b = D.download()
    dt = torch.full((1,), 1e-4, dtype=torch.float64, device="cuda")
    B.run_steps(2, dt.data_ptr(), rk=3, use_graph=True)
    B.tree_set(parent_b, child_b)
    B.upload(U_b)
    dt_b = torch.full((1,), 1e-4, dtype=torch.float64, device="cuda")
    B.run_steps(2, dt_b.data_ptr(), rk=3, use_graph=True)
    dt_d = torch.full((1,), 1e-4, dtype=torch.float64, device="cuda")
    D.run_steps(2, dt_d.data_ptr(), rk=3, use_graph=False)
    np.testing.assert_array_equal(B.download(), D.download())
    np.testing.assert_array_equal(B.tree_flow_off(), D.tree_flow_off())
    assert dt_b.item() == dt_d.item()
    B.close()
    D.close()


ORIGIN = (0.13, -0.07, 0.05)


@pytest.mark.parametrize("z_reflect", [0, 1])
def test_rotating_tree_at_an_origin_matches_the_oracle(z_reflect):
    """The rotating state's positions, sources and outflow ghosts with a non-zero origin: one RK3
    step on the GPU against the oracle with the same origin"""
    T, cfg = binary_tree_case("partial", seed=32, z_reflect=z_reflect)
    T.set_origin(ORIGIN)
    B = device_batch(T, cfg)
    B.tree_set_origin(ORIGIN)
    np.testing.assert_array_equal(B.tree_origin(), ORIGIN)
    T.fill_ghosts()
    B.tree_fill_ghosts()
    dt = 0.5 * T.max_dt()
    np.testing.assert_allclose(B.tree_max_dt(), 2 * dt, rtol=1e-14)
    T.store()
    B.store()
    for beta in BETAS:
        T.substep(dt, beta)
        T.inject()
        B.tree_substep(dt, beta)
        B.tree_inject()
    e = rel_l1_rot(B.download(), T.state(), cfg.g)
    say("rotating tree at origin %s zr=%d: rel-L1 max %.1e" % (ORIGIN, z_reflect, e.max()))
    assert np.all(e <= TOL), e
    B.close()


def test_origin_survives_a_checkpoint(tmp_path):
    T, cfg = binary_tree_case("partial", seed=33)
    B = device_batch(T, cfg)
    B.tree_set_origin(ORIGIN)
    path = str(tmp_path / "o.chk")
    B.tree_checkpoint_write(path)
    d = amrhydro.rotating_config(eos=cfg.eos, recon=cfg.recon, omega=cfg.Omega, omega0=cfg.Omega0)
    B2, t, alpha = amrhydro.Batch.tree_checkpoint_read(path, d)
    np.testing.assert_array_equal(B2.tree_origin(), ORIGIN)
    x, y, z, _, _ = B.tree_output_mesh()
    x2, y2, z2, _, _ = B2.tree_output_mesh()
    np.testing.assert_array_equal(np.sort(x), np.sort(x2))
    np.testing.assert_array_equal(np.sort(y), np.sort(y2))
    np.testing.assert_array_equal(np.sort(z), np.sort(z2))
    B.close()
    B2.close()
