"""The oracle's rotating-frame state on the octree (BinaryDriver::sub_step, binary/binary_driver.cpp:
346-382: enforce_boundaries, fluxes with adjust_*_flux, sum_*_difs, Binary::add_difs; then the
potential energy and inject_from_children), pinned by what the geometry and the conservation laws
fix:
  * a root-only tree is the flat batch of one block bit for bit (outflow and z-mirror ghosts,
    fluxes, sources, floor);
  * coarse->fine ghosts (grid/grid_amr_bound.cpp:27-78) continue a linear profile along the face
    normal exactly and take the parent cell's value along the face; sibling ghosts reproduce any
    profile: this fixes the parity, sign and side of the interpolation and the sibling mapping;
  * the coarse face fluxes next to a refined region are the means of the four fine face fluxes that
    tile them, found by position (grid/grid_node.cpp:1076-1274), in every direction and plane;
  * with Omega = phi = 0 the leaf sums plus the root's boundary-flux ledger of rho, l_z, E (and s_z
    without z reflection) are constant across RK3 steps: reflux, restriction and the ledger
    (grid/grid.cpp:350-404, 419) keep the scheme conservative;
  * restriction is the mean of 8 children, floored with T solved at the parent cell's position
    (grid/grid_node.cpp:1035-1074, binary/state_ztwd.cpp:482-554);
  * the Poisson face forces path (binary/binary.cpp:225-227) equals the phi path when the forces are
    phi's differences."""
import numpy as np
import pytest

import orc
import rot_cases as rc
import workloads
from test_oracle_poisson import full, tree

NF = 13
GAMMA = 5.0 / 3.0


def mid_spec():
    """root, all 8 level-1 nodes, node 1 (the low corner octant) with its lower-z half refined only:
    coarse-fine faces at node 1's middle z plane and at its children's side faces."""
    return full(1) + [(1, c) for c in range(4)]


def partial_spec():
    """root, all 8 level-1 nodes, node 1 refined, and node 1's child 7 (next to the domain centre)
    refined again: nested reflux."""
    return full(1) + [(1, c) for c in range(8)] + [(9 + 7, c) for c in range(8)]


# "one": a properly nested tree (every refined node's 26 neighbours exist), as the Poisson solver needs
TREES = {"mid": tree(mid_spec()), "partial": tree(partial_spec()), "one": tree(full(1) + [(3, c) for c in range(8)])}


def levels(parent, child):
    lv, loc = [0] * len(parent), [(0, 0, 0)] * len(parent)
    for b in range(1, len(parent)):
        p, c = parent[b], child[b]
        lv[b] = lv[p] + 1
        loc[b] = tuple(2 * loc[p][d] + ((c >> d) & 1) for d in range(3))
    return lv, loc


def centres(x0, h, G, g):
    ax = [x0[d] + (np.arange(G) - g + 0.5) * h for d in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return X, Y, Z


def binary_tree_case(name, recon=1, omega=None, z_reflect=0, seed=7, amp=0.3, quiet=False, eos=0):
    """the binary model on tree `name` with random rotating-frame velocities; quiet: Omega = 0 and
    phi = pot = 0 (no sources in rho, l_z, s_z, E); eos 1: the Helmholtz EOS (its table loaded)"""
    parent, child = TREES[name]
    lv, loc = levels(parent, child)
    g = 3 if recon == 1 else 2
    S = workloads.binary_tree_state(lv, loc, n=8, g=g, eos="helm" if eos == 1 else "ideal")
    om = 0.0 if quiet else (S["omega"] if omega is None else omega)
    cfg = orc.rot_cfg(recon=recon, omega=om, z_reflect=z_reflect, eos=eos)
    T = orc.RotTree(cfg, parent, child)
    U, phi = S["U"], S["phi"]
    G = 8 + 2 * g
    rng = np.random.default_rng(seed)
    for b in range(len(parent)):
        x0, h = T.x0(b)
        X, Y, _ = centres(x0, h, G, g)
        cs = np.sqrt(GAMMA * (U[b, 7] / 1.5) / U[b, 0])
        v = [amp * cs * rng.standard_normal(cs.shape) for _ in range(3)]
        rc.set_velocity(U[b:b + 1], X, Y, om, *v)
    if quiet:
        U[:, 6] = 0.0
        phi[:] = 0.0
    T.set_state(U)
    T.set_phi(phi)
    return T, cfg


@pytest.mark.parametrize("recon", [1, 0])
@pytest.mark.parametrize("z_reflect", [0, 1])
def test_root_only_tree_is_the_flat_batch(recon, z_reflect):
    g = 3 if recon == 1 else 2
    S = workloads.binary_tree_state([0], [(0, 0, 0)], n=8, g=g)
    cfg = orc.rot_cfg(recon=recon, omega=S["omega"], z_reflect=z_reflect)
    T = orc.RotTree(cfg, [-1], [0])
    x0, h = T.x0(0)
    assert h == 0.25 and list(x0) == [-1.0, -1.0, 0.0 if z_reflect else -1.0]
    U = S["U"].copy()
    T.set_state(U)
    T.set_phi(S["phi"])
    flat = U.copy()
    X0, H = x0[None].copy(), np.array([h])
    orc.rot_outflow_ghosts(cfg, flat, X0, H)
    base = flat.copy()
    dt = 2e-4
    T.store()
    for beta in (1.0, 0.25):
        T.substep(dt, beta)
        orc.rot_outflow_ghosts(cfg, flat, X0, H)
        orc.rot_substep(cfg, flat, base, S["phi"], X0, H, dt, beta)
        I = (slice(None),) + (slice(g, -g),) * 3
        np.testing.assert_array_equal(T.U(0)[I], flat[0][I])


def _ghost_case(name, coef):
    """fill every node's interior with the linear profiles coef (NF, 4), poison the ghosts, fill
    them; returns the tree and, per node and non-physical face, the ghost slab and its source"""
    parent, child = TREES[name]
    cfg = orc.rot_cfg(recon=1)
    T = orc.RotTree(cfg, parent, child)
    g, G = 3, 14
    lv, loc = levels(parent, child)
    where = {(lv[b],) + loc[b]: b for b in range(len(parent))}
    f = lambda X, Y, Z: np.stack([c[0] + c[1] * X + c[2] * Y + c[3] * Z for c in coef])
    for b in range(T.nn):
        x0, h = T.x0(b)
        T.U(b)[:] = f(*centres(x0, h, G, g))
        T.U(b)[:, :g], T.U(b)[:, -g:] = -99.0, -99.0
        T.U(b)[:, :, :g], T.U(b)[:, :, -g:] = -99.0, -99.0
        T.U(b)[:, :, :, :g], T.U(b)[:, :, :, -g:] = -99.0, -99.0
    T.fill_ghosts()

    def kind(b, face):
        d, up = face >> 1, face & 1
        nb = list(loc[b])
        nb[d] += 1 if up else -1
        if nb[d] < 0 or nb[d] >= 2 ** lv[b]:
            return "physical"
        return "sibling" if (lv[b],) + tuple(nb) in where else "coarse"

    out = []
    for b in range(T.nn):
        for face in range(6):
            k = kind(b, face)
            if k == "physical":
                continue
            d, up = face >> 1, face & 1
            sl = [slice(g, G - g)] * 3
            sl[2 - d] = slice(G - g, G) if up else slice(0, g)
            out.append((b, face, k, kind(parent[b], face) if b else "physical", tuple(sl)))
    return T, f, out


@pytest.mark.parametrize("name", list(TREES))
def test_coarse_fine_ghosts_continue_normal_linear_profiles(name):
    """a profile linear along the face normal is continued exactly into every non-physical ghost,
    through nested coarse-fine faces (the parent's own ghosts feed its children's)"""
    rng = np.random.default_rng(2)
    g, G = 3, 14
    counts = {"sibling": 0, "coarse": 0}
    for d in range(3):
        coef = np.zeros((NF, 4))
        coef[:, 0] = rng.uniform(-1, 1, NF)
        coef[:, 1 + d] = rng.uniform(-1, 1, NF)
        T, f, faces = _ghost_case(name, coef)
        for b, face, k, _, sl in faces:
            if face >> 1 != d:
                continue
            x0, h = T.x0(b)
            want = f(*centres(x0, h, G, g))[(slice(None),) + sl]
            got = T.U(b)[(slice(None),) + sl]
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-13, err_msg="node %d face %d" % (b, face))
            counts[k] += 1
    assert counts["sibling"] > 0 and counts["coarse"] > 0


@pytest.mark.parametrize("name", list(TREES))
def test_coarse_fine_ghosts_are_piecewise_constant_along_the_face(name):
    """along the face a coarse->fine ghost takes the parent cell's value (the slope is only taken
    along the normal); sibling ghosts reproduce any profile"""
    rng = np.random.default_rng(3)
    g, G = 3, 14
    T, f, faces = _ghost_case(name, rng.uniform(-1, 1, (NF, 4)))
    counts = {"sibling": 0, "coarse": 0}
    for b, face, k, pk, sl in faces:
        d = face >> 1
        x0, h = T.x0(b)
        X, Y, Z = centres(x0, h, G, g)
        if k == "sibling":
            want = f(X, Y, Z)[(slice(None),) + sl]
        elif pk == "coarse":
            continue  # the parent's own ghosts there are coarsened once more
        else:
            px0, ph = T.x0(T.parent[b])
            P = [X, Y, Z]
            for t in range(3):
                if t != d:
                    P[t] = px0[t] + (np.floor((P[t] - px0[t]) / ph) + 0.5) * ph
            want = f(*P)[(slice(None),) + sl]
        got = T.U(b)[(slice(None),) + sl]
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13, err_msg="node %d face %d" % (b, face))
        counts[k] += 1
    assert counts["sibling"] > 0 and counts["coarse"] > 0


@pytest.mark.parametrize("name", list(TREES))
def test_coarse_face_fluxes_are_the_fine_means(name):
    T, cfg = binary_tree_case(name)
    parent, child = TREES[name]
    lv, loc = levels(parent, child)
    n, g, G = 8, 3, 14
    T.store()
    T.substep(1e-4, 1.0)
    adjusted = 0
    for d in range(3):
        tang = [t for t in range(3) if t != d]
        # fine boundary faces by global integer position at their own level: (level, plane, a, b)
        fine = {}
        for b in range(T.nn):
            F = T.F(b, d)
            for p in (g, G - g):
                plane = loc[b][d] * n + p - g
                for ib in range(n):
                    for ia in range(n):
                        idx = [0, 0, 0]
                        idx[d], idx[tang[0]], idx[tang[1]] = p, ia + g, ib + g
                        key = (lv[b], plane, loc[b][tang[0]] * n + ia, loc[b][tang[1]] * n + ib)
                        fine.setdefault(key, []).append(F[:, idx[2], idx[1], idx[0]].copy())
        for b in range(T.nn):
            F = T.F(b, d)
            for p in (g, G // 2, G - g):
                plane = 2 * (loc[b][d] * n + p - g)
                for ib in range(n):
                    for ia in range(n):
                        A = 2 * (loc[b][tang[0]] * n + ia)
                        B = 2 * (loc[b][tang[1]] * n + ib)
                        quad = [fine.get((lv[b] + 1, plane, A + da, B + db)) for db in (0, 1) for da in (0, 1)]
                        if any(q is None for q in quad) or any(len(q) != 1 for q in quad):
                            continue  # not next to a refined region, or fine on both sides
                        v = 0.0
                        for q in quad:
                            v = v + q[0]
                        want = v * 0.25
                        idx = [0, 0, 0]
                        idx[d], idx[tang[0]], idx[tang[1]] = p, ia + g, ib + g
                        np.testing.assert_array_equal(F[:, idx[2], idx[1], idx[0]], want)
                        adjusted += 1
    assert adjusted > 0


@pytest.mark.parametrize("name", list(TREES))
@pytest.mark.parametrize("z_reflect", [0, 1])
def test_conservation_with_reflux_and_restriction(name, z_reflect):
    T, cfg = binary_tree_case(name, quiet=True, z_reflect=z_reflect)
    T.fill_ghosts()
    T.inject()
    s0 = T.state_sum() + T.flow_off()
    for _ in range(2):
        dt = T.max_dt()
        T.store()
        for beta in (1.0, 0.25, 2.0 / 3.0):
            assert T.substep(dt, beta) == 0
            T.inject()
    s1 = T.state_sum() + T.flow_off()
    fields = [0, 4, 7] + ([] if z_reflect else [5])
    scale = np.abs(T.state_sum())
    for f in fields:
        assert abs(s1[f] - s0[f]) <= 1e-13 * max(scale[f], scale[0]), (f, s0[f], s1[f])
    # something moved: the ledger is not empty
    assert np.abs(T.flow_off()[[0, 7]]).min() > 0


def test_restriction_is_the_child_mean_floored_at_the_parent_cell():
    T, cfg = binary_tree_case("partial")
    n, g, G = 8, 3, 14
    T.U(1)[:, g:-g, g:-g, g:-g] = -1.0
    assert T.inject() == 0
    kids = {c: T.U(9 + c).copy() for c in range(8)}  # child 7 restricted from its own children first
    par = T.U(1)
    x0, h = T.x0(1)
    X, Y, Z = centres(x0, h, G, g)
    for c in range(8):
        cx, cy, cz = c & 1, (c >> 1) & 1, (c >> 2) & 1
        K = kids[c][:, g:-g, g:-g, g:-g]
        s = 0.0
        for dk, dj, di in ((0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)):
            s = s + K[:, dk::2, dj::2, di::2]
        mean = s * 0.125
        sl = (slice(g + 4 * cz, g + 4 * cz + 4), slice(g + 4 * cy, g + 4 * cy + 4), slice(g + 4 * cx, g + 4 * cx + 4))
        blk = par[(slice(None),) + sl]
        for f in (0, 3, 4, 5, 6, 7):  # untouched by the floor above the density floor
            np.testing.assert_array_equal(blk[f], mean[f])
        # T from the internal energy at the parent cell's position (ideal gas, closed form)
        Xp, Yp = X[sl], Y[sl]
        R = np.hypot(Xp, Yp)
        st = blk[4] / R - blk[0] * R * cfg.Omega
        ek = 0.5 * (blk[3] ** 2 + st ** 2 + blk[5] ** 2) / blk[0]
        ei = blk[7] - ek
        ei = np.where(ei < 0.001 * blk[7], blk[8], ei)
        abar = blk[0] / (blk[10] / 4.0 + blk[11] / 14.12 + blk[12] / 17.52)
        T_want = (ei / blk[0]) * (GAMMA - 1.0) * abar / (1.0 + abar / 2.0)
        np.testing.assert_allclose(blk[9], T_want, rtol=1e-12)
        np.testing.assert_allclose(blk[10:13].sum(axis=0), blk[0], rtol=1e-14)


def test_poisson_forces_path_equals_the_phi_path():
    T1, cfg = binary_tree_case("partial")
    T2, _ = binary_tree_case("partial")
    P = 10
    for b in range(T2.nn):
        _, h = T2.x0(b)
        ph = T2.phi(b)
        Fz = np.zeros((3, P, P, P))
        Fz[0, :, :, 1:] = -(ph[:, :, 1:] - ph[:, :, :-1]) / h
        Fz[1, :, 1:, :] = -(ph[:, 1:, :] - ph[:, :-1, :]) / h
        Fz[2, 1:, :, :] = -(ph[1:, :, :] - ph[:-1, :, :]) / h
        T2.force(b)[:] = Fz
    T2.use_forces(True)
    for T in (T1, T2):
        T.store()
        T.substep(1e-4, 1.0)
    for b in range(T1.nn):
        np.testing.assert_array_equal(T1.U(b)[:, 3:-3, 3:-3, 3:-3], T2.U(b)[:, 3:-3, 3:-3, 3:-3])


def test_max_dt_is_the_min_over_nodes():
    T, cfg = binary_tree_case("partial")
    dt = T.max_dt()  # fills the ghosts first
    dts = []
    for b in range(T.nn):
        x0, h = T.x0(b)
        m = orc.rot_max_eig(cfg, T.U(b)[None].copy(), x0[None].copy(), np.array([h]))[0]
        dts.append(cfg.cfl * h / m)
    assert abs(dt - min(dts)) <= 4e-16 * min(dts)


def test_prolongation_continues_linear_profiles_and_restricts_back():
    """GridNode::inject_from_parent (grid/grid_node.cpp:193-233) of the rotating state: a new child
    of node 16 (the partial tree's level-2 node, siblings on its low faces, coarser nodes on its
    high ones) takes its parent's linear profile at its own cell centres (rho, s_R, l_z, s_z, pot,
    E, which the floor keeps), T solved at the child's cell, and the 8 children restrict back to
    the parent (grid/grid_node.cpp:1035-1074)"""
    parent, child = TREES["partial"]
    T = orc.RotTree(orc.rot_cfg(recon=1, omega=0.4), parent, child)
    g, G = 3, 14
    coef = {0: (1.0, 0.125, -0.0625, 0.25), 3: (0.5, 0.25, 0.125, -0.125), 4: (0.75, -0.125, 0.25, 0.0625),
            5: (-0.25, 0.0625, 0.125, 0.25), 6: (-0.5, 0.125, 0.0, -0.25), 7: (8.0, 0.5, -0.25, 0.125)}
    lin = lambda c, X, Y, Z: c[0] + c[1] * X + c[2] * Y + c[3] * Z
    for b in range(T.nn):
        x0, h = T.x0(b)
        X, Y, Z = centres(x0, h, G, g)
        U = T.U(b)
        for f, c in coef.items():
            U[f] = lin(c, X, Y, Z)
        U[1], U[2] = 0.5 * U[0], 0.5 * U[0]
        U[8] = 1.0
        U[9] = 1.0
        U[10], U[11], U[12] = U[0], 0.0 * U[0], 0.0 * U[0]
    T.fill_ghosts()
    node16 = T.U(16).copy()
    for b in range(17, 25):
        T.U(b)[:] = -7.0
        assert T.prolong(b) == 0
    checked = 0
    for b in range(17, 25):
        x0, h = T.x0(b)
        X, Y, Z = centres(x0, h, G, g)
        U = T.U(b)
        q = T.child[b]
        # child cells whose parent stencil stays off node 16's high (coarse-fine) faces
        keep = np.zeros((G, G, G), bool)
        for k in range(g, G - g):
            for j in range(g, G - g):
                for i in range(g, G - g):
                    p = [(g + c) // 2 + ((q >> d) & 1) * (G // 2 - g) for d, c in enumerate((i - (i - g) % 2, j - (j - g) % 2, k - (k - g) % 2))]
                    keep[k, j, i] = all(pp < G - g - 1 for pp in p)
        for f, c in coef.items():
            want = lin(c, X, Y, Z)
            np.testing.assert_allclose(U[f][keep], want[keep], rtol=0, atol=1e-14, err_msg="field %d" % f)
        checked += keep.sum()
        # T at the child's own cell (ideal gas, closed form), every interior cell
        I = (slice(g, -g),) * 3
        R = np.hypot(X, Y)[I]
        st = U[4][I] / R - U[0][I] * R * 0.4
        ei = U[7][I] - 0.5 * (U[3][I] ** 2 + st ** 2 + U[5][I] ** 2) / U[0][I]
        ei = np.where(ei < 0.001 * U[7][I], U[8][I], ei)  # the dual-energy switch: tau
        abar = U[0][I] / (U[10][I] / 4.0 + U[11][I] / 14.12 + U[12][I] / 17.52)
        np.testing.assert_allclose(U[9][I], (ei / U[0][I]) * (GAMMA - 1.0) * abar / (1.0 + abar / 2.0), rtol=1e-12)
    assert checked > 1000
    T.inject()
    I = (slice(None),) + (slice(g, -g),) * 3
    for f in coef:
        np.testing.assert_allclose(T.U(16)[I][f], node16[I][f], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("z_reflect", [0, 1])
def test_origin_enters_the_positions_as_grid_xf(z_reflect):
    """GridNode::set_origin (grid/grid_node.cpp:20-26) with Grid::xf = (offset + i) h - GRID_DIM -
    bw h0 - origin (grid/grid.cpp:512-524): x and y move by -origin; z moves by -origin_z only under
    Z_REFLECT (without it zf has no origin term)"""
    parent, child = TREES["partial"]
    cfg = orc.rot_cfg(z_reflect=z_reflect)
    T0 = orc.RotTree(cfg, parent, child)
    T1 = orc.RotTree(cfg, parent, child)
    o = np.array([0.13, -0.07, 0.05])
    T1.set_origin(o)
    for b in range(len(parent)):
        (a, h0), (c, h1) = T0.x0(b), T1.x0(b)
        assert h0 == h1
        assert c[0] == a[0] - o[0] and c[1] == a[1] - o[1]
        assert c[2] == (a[2] - o[2] if z_reflect else a[2])
