"""Octree cases shared by the oracle and GPU tree tests."""
import numpy as np

import cases
import orc

GAMMA = 5.0 / 3.0


def build(cfg, spec):
    """spec: list of (parent, child) in creation order (node 0 = root).  Returns the oracle tree
    and the parent/child arrays of the C ABI."""
    T = orc.Tree(cfg)
    parent, child = [-1], [-1]
    for p, c in spec:
        T.add_child(p, c)
        parent.append(p)
        child.append(c)
    return T, np.array(parent, np.int32), np.array(child, np.int32)


def full_spec(maxlevel):
    """Fully refined tree, breadth first: 1 + 8 + 64 + ... nodes."""
    spec, frontier, nxt = [], [0], 1
    for _ in range(maxlevel):
        new = []
        for p in frontier:
            for c in range(8):
                spec.append((p, c))
                new.append(nxt)
                nxt += 1
        frontier = new
    return spec


def partial_spec():
    """Root refined once, its child 0 refined again (coarse/fine faces inside the domain), and one
    grandchild refined a third time (nested reflux)."""
    spec = [(0, c) for c in range(8)] + [(1, c) for c in range(8)]
    spec += [(9 + 7, c) for c in range(8)]
    return spec


def cell_centres(T, node):
    n, g = T.cfg.n, T.cfg.g
    G = n + 2 * g
    lev, loc, off, h = T.info(node)
    h0 = 2 * T.cfg.grid_dim / n
    base = -T.cfg.grid_dim - g * h0
    x = (off[0] + np.arange(G)) * h + base + 0.5 * h
    y = (off[1] + np.arange(G)) * h + base + 0.5 * h
    z = (off[2] + np.arange(G)) * h + base + 0.5 * h
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z, h


def sedov_ic(T, maxlevel):
    """Sedov-type blast in the unit cube: ambient rho = 1, p = 1e-5; energy 1.0 deposited uniformly in
    the finest-level cells whose centres lie within 2 finest cells of the origin."""
    n = T.cfg.n
    hf = 2 * T.cfg.grid_dim / n / 2 ** maxlevel
    r = 2 * hf
    # count finest cells inside r over the whole domain
    m = int(round(2 * T.cfg.grid_dim / hf))
    c = -T.cfg.grid_dim + (np.arange(m) + 0.5) * hf
    Zc, Yc, Xc = np.meshgrid(c, c, c, indexing="ij")
    ndep = int((Xc ** 2 + Yc ** 2 + Zc ** 2 <= r * r).sum())
    edep = 1.0 / (ndep * hf ** 3)
    for node in range(T.nnodes):
        X, Y, Z, h = cell_centres(T, node)
        E = np.full(X.shape, 1e-5 / (GAMMA - 1.0))
        E = np.where(X * X + Y * Y + Z * Z <= r * r, E + edep, E)
        U = T.U(node)
        U[0] = 1.0
        U[1:4] = 0.0
        U[4] = E
        U[5] = E ** (1.0 / GAMMA)
    return ndep


def blob_ic(T):
    for node in range(T.nnodes):
        X, Y, Z, h = cell_centres(T, node)
        r2 = X * X + Y * Y + Z * Z
        rho = 1.0 + 3.0 * np.exp(-r2 / 0.02)
        U = cases.prim_to_con(rho, 0.3 + 0 * X, -0.1 + 0 * X, 0.2 * np.sin(3 * X), 1.0 + 0.5 * np.exp(-r2 / 0.05))
        T.U(node)[:] = np.moveaxis(U, -1, 0)


def padded_state(T):
    return np.stack([T.U(i).copy() for i in range(T.nnodes)])


def leaf_mask(parent):
    """True for nodes without children."""
    nn = len(parent)
    has = np.zeros(nn, bool)
    for b in range(1, nn):
        has[parent[b]] = True
    return ~has


def state_sum(U, h, parent, child, n, g):
    """Sum of U*h^3 over uncovered cells (grid/grid_node.cpp:1489-1529, a0)."""
    nn = U.shape[0]
    covered = np.zeros((nn, 8), bool)
    for b in range(1, nn):
        covered[parent[b], child[b]] = True
    s = np.zeros(U.shape[1])
    half = n // 2
    for b in range(nn):
        I = U[b][:, g:-g, g:-g, g:-g]
        for c in range(8):
            if covered[b, c]:
                continue
            cx, cy, cz = c & 1, (c >> 1) & 1, (c >> 2) & 1
            blk = I[:, cz * half:(cz + 1) * half, cy * half:(cy + 1) * half, cx * half:(cx + 1) * half]
            s += blk.sum(axis=(1, 2, 3)) * h[b] ** 3
    return s
