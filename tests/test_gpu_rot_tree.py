"""The rotating-frame state on the octree on the GPU (amrh_tree_* with AMRH_PHYS_ROTATING) against
the oracle's restatement of BinaryDriver::sub_step (binary/binary_driver.cpp:346-382) through the
C ABI: ghost fill (sibling copy, coarse->fine interpolation, outflow and z mirror) bit for bit;
substeps with reflux, sources, floor and the potential energy, then restriction floored at the
parent cell, to <= 1e-12 relative L1 per field (pow ulps as on flat batches); the root's flow-off
ledger; the Poisson face forces handed over by device pointer; whole RK3 steps as direct launches
and as captured graphs; the mixed form's error against the FP64 oracle."""
import numpy as np
import pytest

import orc
from gpu_util import amrhydro, say
from test_gpu_rotating import helm_eos, rel_l1_rot
from test_oracle_rot_tree import TREES, binary_tree_case, levels

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-12
NF = 13


def device_batch(T, cfg, precision=amrhydro.AMRH_FP64):
    parent, child = T.parent, T.child
    d = amrhydro.rotating_config(eos=cfg.eos, recon=cfg.recon, omega=cfg.Omega, omega0=cfg.Omega0,
                                 z_reflect=cfg.z_reflect, precision=precision)
    B = amrhydro.Batch(d, T.nn, 1.0)
    B.tree_set(parent, child)
    B.upload(T.state())
    B.set_potential(np.stack([T.phi(b).copy() for b in range(T.nn)]))
    if cfg.eos == 1:
        B.attach_eos(helm_eos())
    return B


def slab_mask(g, G):
    m = np.zeros((G, G, G), bool)
    inner = slice(g, G - g)
    m[:, inner, inner] = True
    m[inner, :, inner] = True
    m[inner, inner, :] = True
    return m


@pytest.mark.parametrize("name", list(TREES))
@pytest.mark.parametrize("recon", [1, 0])
@pytest.mark.parametrize("z_reflect", [0, 1])
def test_rot_tree_ghosts_match_oracle(name, recon, z_reflect):
    T, cfg = binary_tree_case(name, recon=recon, z_reflect=z_reflect, seed=21)
    say("rot tree ghosts %s recon=%d zr=%d: %d nodes" % (name, recon, z_reflect, T.nn))
    B = device_batch(T, cfg)
    T.fill_ghosts()
    B.tree_fill_ghosts()
    g, G = cfg.g, cfg.n + 2 * cfg.g
    m = slab_mask(g, G)
    np.testing.assert_array_equal(B.download()[:, :, m], T.state()[:, :, m])
    assert B.error_flags() == 0
    B.close()


@pytest.mark.parametrize("name,recon,z_reflect,eos", [("partial", 1, 0, 0), ("mid", 1, 1, 0), ("partial", 0, 0, 0),
                                                      ("mid", 0, 1, 0), ("partial", 1, 0, 1)])
def test_rot_tree_rk3_matches_oracle(name, recon, z_reflect, eos):
    if eos == 1:
        helm_eos()  # loads the oracle's table too
    T, cfg = binary_tree_case(name, recon=recon, z_reflect=z_reflect, seed=22, eos=eos)
    B = device_batch(T, cfg)
    g = cfg.g
    say("rot tree rk3 %s recon=%d zr=%d eos=%d: %d nodes" % (name, recon, z_reflect, eos, T.nn))
    for step in range(2):
        T.fill_ghosts()
        dt = 0.5 * T.max_dt()
        B.tree_fill_ghosts()
        dt_dev = B.tree_max_dt()
        np.testing.assert_allclose(dt_dev, 2 * dt, rtol=1e-14)
        T.store()
        B.store()
        for beta in (1.0, 0.25, 2.0 / 3.0):
            assert T.substep(dt, beta) == 0
            assert T.inject() == 0
            B.tree_substep(dt, beta)
            B.tree_inject()
            out, ref = B.download(), T.state()
            e = rel_l1_rot(out, ref, g)
            I = (slice(None), slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
            same = np.mean(out[I] == ref[I])
            say("  step %d beta=%.4f rel-L1 max %.1e (field %d)  bit-identical %.6f" % (step, beta, e.max(),
                                                                                      int(np.argmax(e)), same))
            assert np.all(e <= TOL), e
            fo, fo_ref = B.tree_flow_off(), T.flow_off()
            np.testing.assert_allclose(fo, fo_ref, rtol=1e-11, atol=1e-14 * np.abs(fo_ref).max())
    assert B.error_flags() == 0
    B.close()


def test_rot_tree_forces_and_phi_from_the_device_poisson_solver():
    """BinaryDriver's coupling: S = 4 pi G rho from the tree state, the V-cycles, then the hydro
    substep reading the solver's face forces and phi by device pointer, pot_to_grid, restriction"""
    T, cfg = binary_tree_case("one", seed=23)
    B = device_batch(T, cfg)
    P = amrhydro.Poisson(8, T.parent, T.child, grid_dim=1.0, bw=cfg.g)
    P.source_from(B, 1.0)
    it, err = P.solve(1e-8)
    phi = P.download("phi")
    f = np.stack([P.download(k) for k in ("fx", "fy", "fz")])
    say("rot tree + poisson: %d V-cycles, error %.1e" % (it, err))
    B.set_potential_device(P.phi_device())
    B.set_forces_device(P.forces_device())
    T.set_phi(phi)
    T.set_forces(f[0], f[1], f[2])
    T.use_forces(True)
    dt = 1e-4
    T.store()
    B.store()
    T.substep(dt, 1.0)
    T.pot_to_grid()
    T.inject()
    B.tree_substep(dt, 1.0)
    B.pot_to_grid()
    B.tree_inject()
    e = rel_l1_rot(B.download(), T.state(), cfg.g)
    assert np.all(e <= TOL), e
    # the forces differ from phi's differences at coarse-fine faces: the two paths part there
    B.set_forces_device(None)
    B.close()


@pytest.mark.parametrize("use_graph", [0, 1])
def test_rot_tree_whole_steps_conserve(use_graph):
    """amrh_run_steps on the octree (fill, device dt, store, 3 substeps each followed by the
    restriction), direct and as graphs: with Omega = phi = 0 the leaf sums plus the ledger keep rho,
    l_z, s_z and E to roundoff"""
    T, cfg = binary_tree_case("partial", quiet=True, seed=24)
    B = device_batch(T, cfg)
    B.tree_fill_ghosts()
    B.tree_inject()
    s0 = B.state_sum()[0] + B.tree_flow_off()
    B.tree_fill_ghosts()
    # the first step's dt, as HydroDriver's first step (euler/hydro_driver.cpp:108-111); later
    # steps grow it by at most 1.25 up to the CFL limit
    dt = torch.full((1,), 0.01 * B.tree_max_dt(), dtype=torch.float64, device="cuda")
    B.run_steps(3, dt.data_ptr(), rk=3, use_graph=bool(use_graph))
    assert dt.item() > 0
    s1 = B.state_sum()[0] + B.tree_flow_off()
    scale = np.abs(B.state_sum()[0])
    for f in (0, 4, 5, 7):
        assert abs(s1[f] - s0[f]) <= 1e-13 * max(scale[f], scale[0]), (f, s0[f], s1[f])
    say("rot tree run_steps graph=%d: dt %.3e, drift rho %.1e E %.1e" % (use_graph, dt.item(), abs(s1[0] - s0[0]),
                                                                          abs(s1[7] - s0[7])))
    B.close()


def test_rot_tree_graph_steps_equal_direct_steps():
    outs = []
    for use_graph in (0, 1):
        T, cfg = binary_tree_case("partial", seed=25)
        B = device_batch(T, cfg)
        B.tree_fill_ghosts()
        dt = torch.full((1,), 0.01 * B.tree_max_dt(), dtype=torch.float64, device="cuda")
        B.run_steps(2, dt.data_ptr(), rk=3, use_graph=bool(use_graph))
        assert dt.item() > 0
        outs.append((B.download(), B.tree_flow_off()))
        B.close()
    g = 3
    I = (slice(None), slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    np.testing.assert_array_equal(outs[0][0][I], outs[1][0][I])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("eos,recon", [(0, 1), (0, 0), (1, 1)])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("z_reflect", [0, 1])
def test_rot_tree_deferred_steps_equal_plain_steps(eos, recon, precision, z_reflect):
    """The octree's deferred-dt step (the first substep of every level storing D and reducing its
    face speeds to the CFL maximum, then dt, the store and the first stage of every node) gives the
    plain step's bits (ghost fill, CFL scan, store, substeps): the interiors, the flow-off ledger
    and dt, direct and as graphs"""
    outs = []
    g = 3 if recon == 1 else 2
    I = (slice(None), slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    for sched, use_graph in ((amrhydro.AMRH_SCHEDULE_PLAIN, 0), (amrhydro.AMRH_SCHEDULE_DEFERRED, 0),
                             (amrhydro.AMRH_SCHEDULE_DEFERRED, 1)):
        T, cfg = binary_tree_case("partial", recon=recon, z_reflect=z_reflect, seed=27, eos=eos)
        B = device_batch(T, cfg, precision)
        B.set_schedule(sched)
        assert B.schedule() == sched
        B.tree_fill_ghosts()
        dt = torch.full((1,), 0.01 * B.tree_max_dt(), dtype=torch.float64, device="cuda")
        B.run_steps(3, dt.data_ptr(), rk=3, use_graph=bool(use_graph))
        B.sync()
        outs.append((B.download()[I], B.tree_flow_off(), dt.item()))
        assert B.error_flags() == 0
        if sched == amrhydro.AMRH_SCHEDULE_DEFERRED and not use_graph:
            st = B.step_timed(dt.data_ptr(), rk=3)
            assert [k for k, _ in st] == [amrhydro.AMRH_LAUNCH_FIRST, amrhydro.AMRH_LAUNCH_DT,
                                          amrhydro.AMRH_LAUNCH_FIRST_STAGE, amrhydro.AMRH_LAUNCH_SUBSTEP,
                                          amrhydro.AMRH_LAUNCH_SUBSTEP]
        B.close()
    say("rot tree deferred eos=%d recon=%d prec=%d zr=%d: dt %.6e" % (eos, recon, precision, z_reflect, outs[0][2]))
    for u, fo, d in outs[1:]:
        np.testing.assert_array_equal(u, outs[0][0])
        np.testing.assert_array_equal(fo, outs[0][1])
        assert d == outs[0][2]


def test_rot_tree_mixed_stays_near_fp64_oracle():
    T, cfg = binary_tree_case("partial", seed=26)
    B = device_batch(T, cfg, precision=amrhydro.AMRH_MIXED)
    g = cfg.g
    for step in range(5):
        T.fill_ghosts()
        dt = T.max_dt()
        T.store()
        B.store()
        for beta in (1.0, 0.25, 2.0 / 3.0):
            T.substep(dt, beta)
            T.inject()
            B.tree_substep(dt, beta)
            B.tree_inject()
    e = rel_l1_rot(B.download(), T.state(), g)
    say("rot tree mixed after 5 RK3 steps: rel-L1 rho %.2e, max over fields %.2e" % (e[0], e.max()))
    assert e[0] <= 1e-6 and e.max() <= 1e-5
    B.close()


def test_rot_tree_repack_prolongs_like_the_oracle():
    """A regrid of the rotating state: two leaves refined (a level-1 and a level-2 node); surviving
    nodes carry over bit for bit, the new nodes take GridNode::inject_from_parent with the rotating
    floor at their own cells (grid/grid_node.cpp:193-233), as the oracle does"""
    from test_oracle_poisson import tree
    from test_oracle_rot_tree import partial_spec
    T1, cfg = binary_tree_case("partial", seed=27)
    T1.fill_ghosts()
    U1 = T1.state()
    spec2 = partial_spec() + [(p, c) for p in (3, 10) for c in range(8)]
    p2, c2 = tree(spec2)
    T2 = orc.RotTree(cfg, p2, c2)
    n1 = T1.nn
    for b in range(n1):
        T2.U(b)[:] = U1[b]
    for b in range(n1, T2.nn):
        assert T2.prolong(b) == 0
    src = device_batch(T1, cfg)
    d = amrhydro.rotating_config(eos=cfg.eos, recon=cfg.recon, omega=cfg.Omega, omega0=cfg.Omega0)
    dst = amrhydro.Batch(d, T2.nn, 1.0)
    dst.tree_set(p2, c2)
    dst.tree_repack_from(src, np.array(list(range(n1)) + [-1] * (T2.nn - n1), np.int32))
    out, ref = dst.download(), T2.state()
    g = cfg.g
    I = (slice(None), slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    np.testing.assert_array_equal(out[:n1][I], U1[I])
    e = rel_l1_rot(out[n1:], ref[n1:], g)
    same = np.mean(out[n1:][I] == ref[n1:][I])
    say("rot repack: %d carried, %d prolongated, rel-L1 max %.1e, bit-identical %.6f" % (n1, T2.nn - n1, e.max(), same))
    assert np.all(e <= TOL), e
    assert dst.error_flags() == 0
    src.close()
    dst.close()


def test_poisson_phi_from_the_rotating_state():
    """Binary::pot_from_grid (binary/binary.cpp:857-872): phi = pot / rho + R^2 Omega0^2 / 2"""
    T, cfg = binary_tree_case("one", seed=28)
    B = device_batch(T, cfg)
    P = amrhydro.Poisson(8, T.parent, T.child, grid_dim=1.0, bw=cfg.g)
    P.phi_from_state(B)
    phi = P.download("phi")
    g = cfg.g
    for b in range(T.nn):
        x0, h = T.x0(b)
        ax = [x0[d] + (np.arange(8) + 0.5) * h for d in range(3)]
        Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        U = T.U(b)[:, g:-g, g:-g, g:-g]
        want = U[6] / U[0] + 0.5 * (X * X + Y * Y) * cfg.Omega0 * cfg.Omega0
        np.testing.assert_allclose(phi[b, 1:-1, 1:-1, 1:-1], want, rtol=1e-15, atol=1e-15 * np.abs(want).max())
    B.close()
    P.close()


def test_rot_tree_checkpoint_with_the_binary_header(tmp_path):
    """A rotating-state checkpoint: Binary::write's root parameters (binary/binary.cpp:70-93) before
    GridNode::write's depth-first records (grid/grid_node.cpp:29-43), parsed independently; read
    back, the run continues bit for bit"""
    import struct
    T, cfg = binary_tree_case("partial", seed=29)
    B = device_batch(T, cfg)
    B.tree_fill_ghosts()
    dt = 0.5 * B.tree_max_dt()
    B.store()
    B.tree_substep(dt, 1.0)
    B.tree_inject()
    header = struct.pack("<?15d2i", False, *[0.125 * (q + 1) for q in range(12)], cfg.Omega, 0.3, cfg.Omega0, 1, 2)
    assert len(header) == 129
    path = str(tmp_path / "X.chk")
    B.tree_checkpoint_write(path, time=0.5, root_prefix=header)
    U, fo = B.download(), B.tree_flow_off()
    buf = open(path, "rb").read()
    assert buf[:129] == header
    G, nf = 14, 13
    rec = 3 * 8 + 3 * (4 + 8) + nf * 8 + G ** 3 * nf * 8 + 8 + 8 + 4
    assert len(buf) == 129 + T.nn * rec
    root = np.frombuffer(buf, "<f8", count=G ** 3 * nf, offset=129 + 3 * 8 + 36 + nf * 8).reshape(G, G, G, nf)
    g = cfg.g
    np.testing.assert_array_equal(np.moveaxis(root, [0, 1, 2, 3], [3, 2, 1, 0])[:, g:-g, g:-g, g:-g],
                                  U[0][:, g:-g, g:-g, g:-g])
    np.testing.assert_array_equal(np.frombuffer(buf, "<f8", count=nf, offset=129 + 60), fo)
    d = amrhydro.rotating_config(eos=cfg.eos, recon=cfg.recon, omega=cfg.Omega, omega0=cfg.Omega0)
    B2, t, alpha, pre = amrhydro.Batch.tree_checkpoint_read(path, d, prefix_bytes=129)
    assert pre == header and t == 0.5
    # the file's node order is depth first: map it back to the context's order
    order = []

    def dfs(b):
        order.append(b)
        for q in range(8):
            for x in range(T.nn):
                if T.parent[x] == b and T.child[x] == q:
                    dfs(x)
    dfs(0)
    U2 = B2.download()
    I = (slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    for new, old in enumerate(order):
        np.testing.assert_array_equal(U2[new][I], U[old][I])
    np.testing.assert_array_equal(B2.tree_flow_off(), fo)
    B2.set_potential(np.stack([T.phi(b).copy() for b in order]))
    for X in (B, B2):
        X.tree_fill_ghosts()
        X.store()
        X.tree_substep(dt, 1.0)
        X.tree_inject()
    Ua, Ub = B.download(), B2.download()
    for new, old in enumerate(order):
        np.testing.assert_array_equal(Ub[new][I], Ua[old][I])
    B.close()
    B2.close()


@pytest.mark.parametrize("z_reflect", [0, 1])
def test_rot_tree_output_mesh(tmp_path, z_reflect):
    """GridNode::output's mesh of a rotating-state tree with PPM blocks (g = 3): corner points from
    each node's origin, the uncovered cells' zones depth first in k, j, i order, and the 13 fields
    named as State::field_name (binary/state_ztwd.cpp:58-89) in the VTK file"""
    from test_gpu_output import read_vtk
    T, cfg = binary_tree_case("partial", seed=30, z_reflect=z_reflect)
    B = device_batch(T, cfg)
    x, y, z, nl, v = B.tree_output_mesh()
    g, G = cfg.g, 14
    kids = {b: [None] * 8 for b in range(T.nn)}
    for b in range(1, T.nn):
        kids[int(T.parent[b])][int(T.child[b])] = b
    xs, vals, zones = [], [], 0

    def rec(b):
        nonlocal zones
        x0, h = T.x0(b)
        f = [x0[d] + (np.arange(G + 1) - g) * h for d in range(3)]
        Z, Y, X = np.meshgrid(f[2], f[1], f[0], indexing="ij")
        xs.append(np.stack([X.ravel(), Y.ravel(), Z.ravel()]))
        U = T.U(b)
        for k in range(g, G - g):
            for j in range(g, G - g):
                for i in range(g, G - g):
                    c = (2 * i) // G | ((2 * j) // G) << 1 | ((2 * k) // G) << 2
                    if kids[b][c] is None:
                        vals.append(U[:, k, j, i].copy())
        for c in range(8):
            if kids[b][c] is not None:
                rec(kids[b][c])

    rec(0)
    P = np.concatenate(xs, axis=1)
    np.testing.assert_array_equal(x, P[0])
    np.testing.assert_array_equal(y, P[1])
    np.testing.assert_array_equal(z, P[2])
    np.testing.assert_array_equal(v, np.array(vals).T)
    assert nl.shape == (len(vals), 8)
    path = str(tmp_path / "rot.vtk")
    B.tree_output_vtk(path, time=0.5)
    _, pts, cells, _, fields = read_vtk(path)
    assert list(fields) == ["d_tot", "d_acc", "d_don", "sr", "lz", "sz", "pot", "et", "tau", "T", "s1", "s2", "s3"]
    np.testing.assert_array_equal(fields["lz"], np.array(vals)[:, 4])
    vol = np.prod(pts[cells[:, 1:]][:, 6] - pts[cells[:, 1:]][:, 0], axis=1).sum()
    assert abs(vol - 8.0) < 1e-12  # [-1, 1]^2 x [-1, 1] or [0, 2] under z reflection
    B.close()


@pytest.mark.parametrize("rk", [1, 2])
def test_rot_tree_steps_cycle_through_three_state_buffers(rk):
    """The rotating octree keeps three state buffers (its substep never writes its input): RK1 and
    RK2 steps end in a different buffer than the one they started from, so successive steps start
    from each buffer in turn and capture one graph per starting buffer.  Five steps as graphs equal
    five direct steps, and a restore afterwards returns the last step's stored state bit for bit."""
    g = 3
    I = (slice(None), slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    outs = []
    for use_graph in (0, 1):
        T, cfg = binary_tree_case("partial", seed=30)
        B = device_batch(T, cfg)
        B.tree_fill_ghosts()
        dt = torch.full((1,), 0.01 * B.tree_max_dt(), dtype=torch.float64, device="cuda")
        B.run_steps(4, dt.data_ptr(), rk=rk, use_graph=bool(use_graph))
        B.sync()
        before = B.download()
        fo_before = B.tree_flow_off()
        B.run_steps(1, dt.data_ptr(), rk=rk, use_graph=bool(use_graph))
        B.sync()
        outs.append((B.download(), B.tree_flow_off(), dt.item()))
        B.restore()
        np.testing.assert_array_equal(B.download()[I], before[I])
        np.testing.assert_array_equal(B.tree_flow_off(), fo_before)
        assert B.error_flags() == 0
        B.close()
    say("rot tree rk%d steps through three buffers: dt %.6e" % (rk, outs[0][2]))
    np.testing.assert_array_equal(outs[0][0][I], outs[1][0][I])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2]


def test_rot_tree_substeps_on_a_large_octree_match_oracle():
    """The binary's 2,473-node octree (levels of up to ~1,700 nodes, so a level's workgroups run in
    several rounds): the substeps read the ghosts across same-level faces from the neighbours'
    interiors, which must still hold the substep's input when a neighbour has already finished
    (the update writes another buffer).  One RK3 step, FP64, against the oracle, then whole steps of
    the deferred schedule, direct and as graphs, against the plain one."""
    import workloads
    parent, child, lv, loc = workloads.binary_tree(5)
    S = workloads.binary_tree_state(lv, loc)
    say("large rotating octree: %d nodes, level sizes %s" % (len(parent), np.bincount(lv).tolist()))
    cfg = orc.rot_cfg(omega=S["omega"])
    T = orc.RotTree(cfg, parent, child)
    T.set_state(S["U"])
    T.set_phi(S["phi"])
    T.fill_ghosts()
    dt = T.max_dt()
    d = amrhydro.rotating_config(omega=S["omega"])
    B = amrhydro.Batch(d, len(parent), 1.0)
    B.tree_set(parent, child)
    B.upload(S["U"])
    B.set_potential(S["phi"])
    B.tree_fill_ghosts()
    T.store()
    B.store()
    g = cfg.g
    for beta in (1.0, 0.25, 2.0 / 3.0):
        assert T.substep(dt, beta) == 0
        assert T.inject() == 0
        B.tree_substep(dt, beta)
        B.tree_inject()
        e = rel_l1_rot(B.download(), T.state(), g)
        say("  beta=%.4f rel-L1 max %.1e (field %d)" % (beta, e.max(), int(np.argmax(e))))
        assert np.all(e <= TOL), e
    assert B.error_flags() == 0
    B.close()
    I = (slice(None), slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    outs = []
    for prec in (amrhydro.AMRH_FP64, amrhydro.AMRH_MIXED):
        for sched, use_graph in ((amrhydro.AMRH_SCHEDULE_PLAIN, 0), (amrhydro.AMRH_SCHEDULE_DEFERRED, 0),
                                 (amrhydro.AMRH_SCHEDULE_DEFERRED, 1)):
            B = amrhydro.Batch(amrhydro.rotating_config(omega=S["omega"], precision=prec), len(parent), 1.0)
            B.tree_set(parent, child)
            B.upload(S["U"])
            B.set_potential(S["phi"])
            B.set_schedule(sched)
            B.tree_fill_ghosts()
            B.tree_inject()
            dtd = torch.full((1,), 0.01 * B.tree_max_dt(), dtype=torch.float64, device="cuda")
            B.run_steps(3, dtd.data_ptr(), rk=3, use_graph=bool(use_graph))
            B.sync()
            outs.append((prec, B.download()[I], dtd.item()))
            assert B.error_flags() == 0
            B.close()
    for k in (1, 2, 4, 5):
        base = outs[0 if k < 3 else 3]
        assert outs[k][0] == base[0]
        np.testing.assert_array_equal(outs[k][1], base[1])
        assert outs[k][2] == base[2]
