"""Pins the oracle's block and tree operations (grid/grid.cpp, grid/grid_node.cpp,
grid/grid_amr_bound.cpp), which cannot be built from the reference here (they include <silo.h>),
by what the mathematics fixes: exact cancellation, closed forms, symmetry, conservation."""
import numpy as np
import pytest

import cases
import orc

G2 = 2


def blocks(n, nb, seed):
    U = cases.c2_fast(nb, n=n, seed=seed)
    orc.outflow_ghosts(orc.euler_cfg(n=n), U)
    return U


def test_uniform_state_has_zero_divergence():
    cfg = orc.euler_cfg(n=8)
    cell = cases.prim_to_con(np.array(0.7), np.array(0.3), np.array(-0.2), np.array(0.1), np.array(0.4))
    U = np.empty((2, 6, 12, 12, 12))
    U[:] = cell[None, :, None, None, None]
    D = orc.batch_substep(cfg, U, U.copy(), np.full(2, 0.1), 1e-3, 1.0, want_D=True)
    assert np.all(D[:, :, 2:-2, 2:-2, 2:-2] == 0.0)


def test_linear_profile_faces_are_exact():
    """A primitive that is linear along the pencil has minmod-theta slope equal to the gradient, so
    the face values are the midpoints (dyadic data: exact in binary)."""
    cfg = orc.euler_cfg(n=8)
    G = 12
    rho = 1.0 + np.arange(G) / 8.0
    q = cases.prim_to_con(rho, np.zeros(G), np.zeros(G), np.zeros(G), np.full(G, 0.5))
    ql = np.zeros((G, 6))
    qr = np.zeros((G, 6))
    orc.lib().orc_pencil_faces(orc.C.byref(cfg), G, orc.dptr(q), orc.dptr(ql), orc.dptr(qr))
    for i in range(2, 11):
        want = 1.0 + (i - 0.5) / 8.0
        assert ql[i, 0] == want and qr[i, 0] == want


def test_beta_one_is_forward_euler():
    """beta = 1: U' = U + dt*D before the floor, which leaves rho, s and E untouched here."""
    cfg = orc.euler_cfg(n=8)
    U = blocks(8, 3, 4)
    U0 = U.copy()
    dt = 1e-4
    D = orc.batch_substep(cfg, U, U0, np.full(3, 0.125), dt, 1.0, want_D=True)
    I = (slice(None), slice(0, 5), slice(2, -2), slice(2, -2), slice(2, -2))
    np.testing.assert_array_equal(U[I], (U0[I] + D[I] * dt) * 1.0)


def test_rk_combination_closed_form():
    cfg = orc.euler_cfg(n=8)
    U = blocks(8, 2, 8)
    U0 = blocks(8, 2, 9)
    Ua = U.copy()
    D = orc.batch_substep(cfg, Ua, U0, np.full(2, 0.125), 1e-4, 0.25, want_D=True)
    I = (slice(None), slice(0, 5), slice(2, -2), slice(2, -2), slice(2, -2))
    np.testing.assert_array_equal(Ua[I], (U[I] + D[I] * 1e-4) * 0.25 + U0[I] * (1.0 - 0.25))


def permute_xy(U):
    V = np.swapaxes(U, -1, -2).copy()  # (f, k, j, i) -> (f, k, i, j): x <-> y
    V[:, [1, 2]] = V[:, [2, 1]]
    return V


def test_axis_permutation_symmetry():
    cfg = orc.euler_cfg(n=8)
    U = blocks(8, 3, 21)
    Dx = orc.batch_substep(cfg, U.copy(), U.copy(), np.full(3, 0.125), 1e-4, 1.0, want_D=True)
    V = permute_xy(U)
    Dy = orc.batch_substep(cfg, V.copy(), V.copy(), np.full(3, 0.125), 1e-4, 1.0, want_D=True)
    a = permute_xy(Dx)[:, :, 2:-2, 2:-2, 2:-2]
    b = Dy[:, :, 2:-2, 2:-2, 2:-2]
    # only the order of the x/y contributions to D (and of sx^2 + sy^2) differs: roundoff
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_mirror_symmetry():
    cfg = orc.euler_cfg(n=8)
    U = blocks(8, 3, 22)
    D = orc.batch_substep(cfg, U.copy(), U.copy(), np.full(3, 0.125), 1e-4, 1.0, want_D=True)
    V = U[..., ::-1].copy()
    V[:, 1] = -V[:, 1]
    Dm = orc.batch_substep(cfg, V.copy(), V.copy(), np.full(3, 0.125), 1e-4, 1.0, want_D=True)
    Dm = Dm[..., ::-1].copy()
    Dm[:, 1] = -Dm[:, 1]
    a, b = D[:, :, 2:-2, 2:-2, 2:-2], Dm[:, :, 2:-2, 2:-2, 2:-2]
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_periodic_block_divergence_sums_to_zero():
    cfg = orc.euler_cfg(n=8)
    U = cases.c2_fast(2, n=8, seed=3)
    g = 2
    I = U[:, :, g:-g, g:-g, g:-g]
    U[:, :, :g, g:-g, g:-g] = I[:, :, -g:]
    U[:, :, -g:, g:-g, g:-g] = I[:, :, :g]
    U[:, :, g:-g, :g, g:-g] = I[:, :, :, -g:]
    U[:, :, g:-g, -g:, g:-g] = I[:, :, :, :g]
    U[:, :, g:-g, g:-g, :g] = I[:, :, :, :, -g:]
    U[:, :, g:-g, g:-g, -g:] = I[:, :, :, :, :g]
    D = orc.batch_substep(cfg, U.copy(), U.copy(), np.full(2, 0.125), 1e-4, 1.0, want_D=True)
    Di = D[:, :, g:-g, g:-g, g:-g]
    for f in range(5):
        s = Di[:, f].sum(axis=(1, 2, 3))
        scale = np.abs(Di[:, f]).sum(axis=(1, 2, 3))
        assert np.all(np.abs(s) <= 1e-13 * scale)


# ---- tree ------------------------------------------------------------------------------------

def fill_tree(T, fn):
    """Set every cell (ghosts included) of every node from fn(x, y, z) -> (6,) conserved."""
    n, g = T.cfg.n, T.cfg.g
    G = n + 2 * g
    for node in range(T.nnodes):
        lev, loc, off, h = T.info(node)
        c = (off[0] + np.arange(G)) * h - T.cfg.grid_dim - g * (2 * T.cfg.grid_dim / n) + 0.5 * h
        cy = (off[1] + np.arange(G)) * h - T.cfg.grid_dim - g * (2 * T.cfg.grid_dim / n) + 0.5 * h
        cz = (off[2] + np.arange(G)) * h - T.cfg.grid_dim - g * (2 * T.cfg.grid_dim / n) + 0.5 * h
        Z, Y, X = np.meshgrid(cz, cy, c, indexing="ij")
        T.U(node)[:] = np.moveaxis(fn(X, Y, Z), -1, 0)


def partial_tree(cfg):
    """Root refined once; child 0 refined again: coarse/fine interfaces inside the domain."""
    T = orc.Tree(cfg)
    for c in range(8):
        T.add_child(0, c)
    for c in range(8):
        T.add_child(1, c)
    return T


def test_tree_offsets_and_widths():
    cfg = orc.euler_cfg(n=8, grid_dim=0.5)
    T = partial_tree(cfg)
    lev, loc, off, h = T.info(9 + 7)  # child 7 of node 1 (= child 0 of the root)
    assert lev == 2 and loc == (1, 1, 1) and h == 0.5 * 0.5 * (1.0 / 8)
    # grid/grid_node.cpp:276: offset = 2*parent_offset + bw + c*(GNX-2bw)
    # parent (child 0 of the root) has offset 2*0 + 2 + 0 = 2
    assert off == (2 * 2 + 2 + 8,) * 3


@pytest.mark.parametrize("g,recon", [(2, 0), (3, 1)])
def test_amr_ghosts_reproduce_normal_linear_profile(g, recon):
    """Coarse->fine ghost = u -/+ minmod(normal slope)/4 (grid/grid_amr_bound.cpp:27-78): exact for a
    profile linear along the face normal; piecewise constant (the parent cell's value) along the
    face.  Node 9+1 (child 1 of the root's child 0) has a coarse neighbour across its x-upper face.
    Ghost width 2 (minmod) and 3 (PPM)."""
    cfg = orc.euler_cfg(n=8, g=g, recon=recon, grid_dim=0.5)
    T = partial_tree(cfg)
    node = 9 + 1

    def lin(X, Y, Z):
        rho = 2.0 + 0.25 * X + 0 * Y
        return cases.prim_to_con(rho, 0 * X, 0 * X, 0 * X, 1.0 + 0 * X)

    fill_tree(T, lin)
    want = T.U(node).copy()
    T.U(node)[:, :, :, -g:] = -99.0
    T.enforce_boundaries()
    I = (0, slice(g, -g), slice(g, -g), slice(-g, None))
    np.testing.assert_allclose(T.U(node)[I], want[I], rtol=0, atol=1e-14)
    if g != 2:
        return

    def tang(X, Y, Z):
        rho = 2.0 + 0.25 * Y + 0 * X
        return cases.prim_to_con(rho, 0 * X, 0 * X, 0 * X, 1.0 + 0 * X)

    fill_tree(T, tang)
    T.enforce_boundaries()
    par = T.U(1)  # the parent: fine j -> parent j1 = (off_j - 2*off_parent_j + j)//2
    lev, loc, off, h = T.info(node)
    plev, ploc, poff, ph = T.info(1)
    for j in range(2, 10):
        j1 = (off[1] - 2 * poff[1] + j) // 2
        k1 = (off[2] - 2 * poff[2] + 5) // 2
        assert T.U(node)[0, 5, j, 10] == par[0, k1, j1, 10]


def test_restriction_is_child_mean():
    cfg = orc.euler_cfg(n=8, grid_dim=0.5)
    T = partial_tree(cfg)
    rng = np.random.default_rng(0)
    for i in range(T.nnodes):
        U = T.U(i)
        U[0] = np.exp(rng.normal(0, 0.3, U[0].shape))
        U[1:4] = rng.normal(0, 0.1, U[1:4].shape)
        U[4] = 2.0 + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]
        U[5] = (U[4] - 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]) ** 0.6
    kids = [T.U(9 + c).copy() for c in range(8)]
    T.inject_from_children()
    P = T.U(1)
    for c in range(8):
        cx, cy, cz = c & 1, (c >> 1) & 1, (c >> 2) & 1
        K = kids[c][:, 2:-2, 2:-2, 2:-2]
        mean = K.reshape(6, 4, 2, 4, 2, 4, 2).mean(axis=(2, 4, 6))
        blk = P[:, 2 + 4 * cz:6 + 4 * cz, 2 + 4 * cy:6 + 4 * cy, 2 + 4 * cx:6 + 4 * cx]
        for f in range(5):  # summation order differs from numpy's mean: roundoff of the field scale
            np.testing.assert_allclose(blk[f], mean[f], rtol=0, atol=2e-15 * np.abs(K[f]).max())


@pytest.mark.parametrize("g,recon", [(2, 0), (3, 1)])
def test_tree_conserves_mass_and_energy_with_reflux(g, recon):
    """Sum over leaf cells of U*h^3 plus the root's boundary-flux ledger is constant to roundoff
    across RK3 steps of a tree with coarse/fine interfaces: reflux (grid/grid_node.cpp:1076-1274)
    and restriction keep the scheme conservative, with minmod and with PPM."""
    cfg = orc.euler_cfg(n=8, g=g, recon=recon, grid_dim=0.5)
    T = partial_tree(cfg)

    def blob(X, Y, Z):
        r2 = X * X + Y * Y + Z * Z
        rho = 1.0 + 3.0 * np.exp(-r2 / 0.02)
        return cases.prim_to_con(rho, 0.3 + 0 * X, -0.1 + 0 * X, 0.2 * np.sin(3 * X), 1.0 + 0.5 * np.exp(-r2 / 0.05))

    fill_tree(T, blob)
    T.enforce_boundaries()
    T.inject_from_children()
    s0 = T.state_sum() + T.flow_off()
    for _ in range(3):
        T.enforce_boundaries()
        dt = T.max_dt()
        T.store()
        for beta in (1.0, 0.25, 2.0 / 3.0):
            T.substep(dt, beta)
    s1 = T.state_sum() + T.flow_off()
    for f in (0, 1, 2, 3, 4):
        scale = abs(s0[f]) + abs(T.state_sum()[f]) + 1e-300
        assert abs(s1[f] - s0[f]) <= 1e-13 * max(scale, abs(s0[0])), (f, s0[f], s1[f])


def test_tree_uniform_state_is_stationary():
    cfg = orc.euler_cfg(n=8, grid_dim=0.5)
    T = partial_tree(cfg)
    cell = cases.prim_to_con(np.array(0.7), np.array(0.0), np.array(0.0), np.array(0.0), np.array(0.4))
    orc.lib().orc_floor(orc.C.byref(cfg), orc.dptr(cell))
    fill_tree(T, lambda X, Y, Z: np.broadcast_to(cell, X.shape + (6,)).copy())
    before = [T.U(i).copy() for i in range(T.nnodes)]
    T.enforce_boundaries()
    T.store()
    T.substep(1e-3, 1.0)
    # leaves (nodes 2..16): exact.  Covered parent cells are the 8-cell average, which rounds.
    for i in range(2, T.nnodes):
        np.testing.assert_array_equal(T.U(i)[:, 2:-2, 2:-2, 2:-2], before[i][:, 2:-2, 2:-2, 2:-2])


def test_tree_max_dt_is_min_over_nodes():
    cfg = orc.euler_cfg(n=8, grid_dim=0.5)
    T = partial_tree(cfg)
    fill_tree(T, lambda X, Y, Z: cases.prim_to_con(1.0 + 0.5 * np.sin(4 * X), 0.2 + 0 * X, 0 * X, 0 * X, 1.0 + 0 * X))
    T.enforce_boundaries()
    dts = []
    for i in range(T.nnodes):
        lev, loc, off, h = T.info(i)
        m = orc.batch_max_eig(cfg, T.U(i)[None].copy())[0]
        dts.append(orc.block_dt(cfg, m, h))
    assert abs(T.max_dt() - min(dts)) <= 4e-16 * min(dts)
