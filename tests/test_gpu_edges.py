"""Edge cases through the C ABI on the GPU: a batch of one block, a one-node Poisson tree, and the
tree checks that run once a context exists."""
import numpy as np
import pytest

import orc
from gpu_util import amrhydro, say
from test_oracle_poisson import centres, sphere_source

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import workloads  # noqa: E402


@pytest.mark.parametrize("n", [8, 16])
def test_single_block_substeps(n):
    U = workloads.flat_batch(1, n=n, seed=3)
    cfg = orc.euler_cfg(n=n)
    orc.outflow_ghosts(cfg, U)
    h = np.full(1, 1.0 / n)
    dt = 0.3 * orc.block_dt(cfg, orc.batch_max_eig(cfg, U).max(), h[0])
    ref, ref0 = U.copy(), U.copy()
    outs = {}
    for prec in (amrhydro.AMRH_FP64, amrhydro.AMRH_MIXED):
        B = amrhydro.Batch(amrhydro.euler_config(n=n, precision=prec), 1, h)
        B.upload(U)
        B.store()
        for beta in (1.0, 0.25, 2.0 / 3.0):
            B.fill_outflow_ghosts()
            B.substep(dt, beta)
        outs[prec] = B.download()
        B.close()
    for beta in (1.0, 0.25, 2.0 / 3.0):
        orc.outflow_ghosts(cfg, ref)
        orc.batch_substep(cfg, ref, ref0, h, dt, beta)
    g = 2
    I = (slice(None), slice(0, 5), slice(g, -g), slice(g, -g), slice(g, -g))
    np.testing.assert_array_equal(outs[amrhydro.AMRH_FP64][I], ref[I])
    e = np.abs(outs[amrhydro.AMRH_MIXED][I] - ref[I]).sum() / np.abs(ref[I]).sum()
    say("single block n=%d: mixed rel-L1 %.2e" % (n, e))
    assert e < 1e-6


@pytest.mark.parametrize("z_reflect", [0, 1])
def test_one_node_poisson(z_reflect):
    T = orc.Poisson(8, [-1], [0], z_reflect=z_reflect)
    D = amrhydro.Poisson(8, [-1], [0], z_reflect=z_reflect)
    C = centres(T, z_reflect=z_reflect)
    sphere_source(T, C, r0=0.5)
    D.upload("S", T.fields("S"))
    it_ref, _ = T.solve(1e-9)
    it, _ = D.solve(1e-9)
    phi_ref, phi = T.fields("phi"), D.download("phi")
    err = np.max(np.abs(phi - phi_ref)) / np.max(np.abs(phi_ref))
    say("one-node poisson zr=%d: %d / %d V-cycles, phi max rel diff %.2e" % (z_reflect, it, it_ref, err))
    assert it == it_ref and err < 1e-12


@pytest.mark.parametrize("parent,child", [
    ([-1, 1], [0, 0]),          # parent after child
    ([-1, 0, 0], [0, 2, 2]),    # one slot twice
    ([-1, 0], [0, 9]),          # child index out of range
])
def test_tree_set_rejects_bad_arrays(parent, child):
    B = amrhydro.Batch(amrhydro.euler_config(n=8), len(parent), 1.0)
    with pytest.raises(amrhydro.AmrhError, match="bad parent"):
        B.tree_set(np.array(parent, np.int32), np.array(child, np.int32))
    B.close()


def test_rotating_entry_points_reject_misuse(tmp_path):
    """The rotating-state and self-gravity entry points check their preconditions before any
    device work: forces and pot_to_grid belong to the rotating state, the substep needs the
    potential, phi from pot needs a rotating state, re-packing needs one block shape and one physics
    on both sides, and
    a checkpoint shorter than its root prefix is refused."""
    E = amrhydro.Batch(amrhydro.euler_config(n=8), 9, 1.0)
    E.tree_set(np.array([-1] + [0] * 8, np.int32), np.array([-1] + list(range(8)), np.int32))
    with pytest.raises(amrhydro.AmrhError, match="ROTATING"):
        E.set_forces(np.zeros((3, 9, 10, 10, 10)))
    with pytest.raises(amrhydro.AmrhError, match="ROTATING"):
        E.pot_to_grid()
    P = amrhydro.Poisson(8, [-1] + [0] * 8, [-1] + list(range(8)))
    E.upload(np.ones((9, 6, 12, 12, 12)))
    with pytest.raises(amrhydro.AmrhError, match="rotating"):
        P.phi_from_state(E)
    R = amrhydro.Batch(amrhydro.rotating_config(), 9, 1.0)
    R.tree_set(np.array([-1] + [0] * 8, np.int32), np.array([-1] + list(range(8)), np.int32))
    U = np.zeros((9, 13, 14, 14, 14))
    U[:, 0] = 1.0
    U[:, 7] = 1.0
    U[:, 8] = 1.0
    U[:, 9] = 1.0
    U[:, 10] = 1.0
    R.upload(U)
    with pytest.raises(amrhydro.AmrhError, match="potential"):
        R.tree_substep(1e-4, 1.0)
    with pytest.raises(amrhydro.AmrhError, match="shapes differ|different physics"):
        R.tree_repack_from(E, np.arange(9, dtype=np.int32))
    path = str(tmp_path / "short.chk")
    R.tree_checkpoint_write(path, root_prefix=b"\x00" * 16)
    with pytest.raises(amrhydro.AmrhError):
        amrhydro.Batch.tree_checkpoint_read(path, amrhydro.rotating_config(), prefix_bytes=1 << 30)
    for X in (E, R, P):
        X.close()
