"""Regrid re-packing, the refinement density test and checkpoints through the C ABI:
  * a tree refined by one more level: surviving nodes carry over bit for bit, new nodes match the
    oracle's inject_from_parent (grid/grid_node.cpp:193-233) bit for bit in rho, momenta and E,
    and to an ulp (rtol 1e-15) in tau, which the floor resets through pow;
  * the density flags of GridNode::set_refine_flags (:647-667) match numpy;
  * a checkpoint written by the library parses with the reference's layout (GridNode::write /
    Grid::write, grid/grid_node.cpp:29-43, grid/grid.cpp:104-124), reads back to the same state,
    and the run continues identically from it."""
import numpy as np
import pytest

import orc
import tree_cases as tc
from gpu_util import amrhydro, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

n, g = 8, 2
G = n + 2 * g


def oracle_tree(spec, g=g):
    cfg = orc.euler_cfg(n=n, g=g, recon=1 if g == 3 else 0, grid_dim=0.5, z_reflect=0)
    return tc.build(cfg, spec)


def gpu_tree(parent, child, U=None, g=g):
    recon = amrhydro.AMRH_RECON_PPM if g == 3 else amrhydro.AMRH_RECON_MINMOD
    B = amrhydro.Batch(amrhydro.euler_config(n=n, g=g, grid_dim=0.5, recon=recon), len(parent), 1.0)
    B.tree_set(parent, child)
    if U is not None:
        B.upload(U)
    return B


def slab_mask(g=g):
    G = n + 2 * g
    m = np.zeros((G, G, G), bool)
    s = slice(g, G - g)
    m[:, s, s] = True
    m[s, :, s] = True
    m[s, s, :] = True
    return m


@pytest.mark.parametrize("g", [2, 3])
def test_repack_after_refinement_matches_oracle(g):
    spec1 = tc.partial_spec()
    T1, p1, c1 = oracle_tree(spec1, g)
    tc.blob_ic(T1)
    T1.enforce_boundaries()
    T1.inject_from_children()
    T1.enforce_boundaries()
    U1 = tc.padded_state(T1)
    # refine two leaves: node 3 (level 1) and node 20 (level 2)
    leaves = [3, 20]
    spec2 = spec1 + [(p, c) for p in leaves for c in range(8)]
    T2, p2, c2 = oracle_tree(spec2, g)
    n1 = len(p1)
    for b in range(n1):
        T2.U(b)[:] = U1[b]
    new = list(range(n1, len(p2)))
    for b in new:  # parents are old nodes with filled ghosts
        T2.prolong(b)
    ref = tc.padded_state(T2)
    src = gpu_tree(p1, c1, U1, g)
    dst = gpu_tree(p2, c2, g=g)
    idx = np.array(list(range(n1)) + [-1] * len(new), np.int32)
    dst.tree_repack_from(src, idx)
    out = dst.download()
    I = (slice(None), slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    np.testing.assert_array_equal(out[:n1][I], U1[I])
    # rho, momenta, E bit for bit; tau is reset by the floor through pow (device vs libm: an ulp)
    J = (slice(None), slice(0, 5), slice(g, -g), slice(g, -g), slice(g, -g))
    np.testing.assert_array_equal(out[n1:][J], ref[n1:][J])
    T_ = (slice(None), 5, slice(g, -g), slice(g, -g), slice(g, -g))
    np.testing.assert_allclose(out[n1:][T_], ref[n1:][T_], rtol=1e-15, atol=0)
    say("repack g=%d: %d old nodes carried, %d new nodes prolongated, match" % (g, n1, len(new)))
    # the new tree steps on: ghosts were filled by the repack (same as the oracle's fill)
    T2.enforce_boundaries()
    m = slab_mask(g)
    np.testing.assert_allclose(out[:, :, m], tc.padded_state(T2)[:, :, m], rtol=1e-15, atol=0)


@pytest.mark.parametrize("g", [2, 3])
def test_density_flags_match_numpy(g):
    spec = tc.partial_spec()
    T, p, c = oracle_tree(spec, g)
    tc.blob_ic(T)
    U = tc.padded_state(T)
    B = gpu_tree(p, c, U, g)
    thr = 1.5
    got = B.tree_density_flags(thr)
    rho = U[:, 0, g:-g, g:-g, g:-g]
    want = np.zeros(len(p), np.uint8)
    h2 = n // 2
    for ci in range(8):
        cx, cy, cz = ci & 1, ci >> 1 & 1, ci >> 2 & 1
        octv = rho[:, cz * h2:(cz + 1) * h2, cy * h2:(cy + 1) * h2, cx * h2:(cx + 1) * h2]
        want |= ((octv > thr).any(axis=(1, 2, 3)).astype(np.uint8) << ci)
    np.testing.assert_array_equal(got, want)


def parse_checkpoint(path, nf=6):
    """The reference's GridNode::write layout, read independently (depth first)."""
    buf = open(path, "rb").read()
    pos = [0]

    def take(fmt_dtype, count):
        a = np.frombuffer(buf, dtype=fmt_dtype, count=count, offset=pos[0])
        pos[0] += a.nbytes
        return a

    nodes = []

    def node(parent, ci):
        rec = {"parent": parent, "child": ci}
        rec["h"], rec["h0"], rec["time"] = take("<f8", 3)
        offs, orig = [], []
        for _ in range(3):
            offs.append(int(take("<i4", 1)[0]))
            orig.append(float(take("<f8", 1)[0]))
        rec["off"], rec["origin"] = offs, orig
        rec["FO"] = take("<f8", nf).copy()
        rec["U"] = take("<f8", G * G * G * nf).reshape(G, G, G, nf).copy()  # [i][j][k][l]
        rec["c"] = take("u1", 8).copy()
        rec["alpha"] = float(take("<f8", 1)[0])
        rec["age"] = int(take("<i4", 1)[0])
        me = len(nodes)
        nodes.append(rec)
        for q in range(8):
            if rec["c"][q]:
                node(me, q)

    node(-1, -1)
    assert pos[0] == len(buf)
    return nodes


def test_checkpoint_roundtrip_and_layout(tmp_path):
    spec = tc.partial_spec()
    T, p, c = oracle_tree(spec)
    tc.blob_ic(T)
    T.enforce_boundaries()
    T.inject_from_children()
    B = gpu_tree(p, c, tc.padded_state(T))
    for _ in range(2):
        B.tree_step(1e-4, 3)
    path = str(tmp_path / "X.chk")
    B.tree_checkpoint_write(path, time=0.75, refine_alpha=1.0)
    U = B.download()
    fo = B.tree_flow_off()
    nodes = parse_checkpoint(path)
    assert len(nodes) == len(p)
    # depth-first order of the file vs the breadth-first node order of the context
    m = slab_mask()
    info = [T.info(b) for b in range(len(p))]
    order = []

    def dfs(b):
        order.append(b)
        for q in range(8):
            kids = [x for x in range(len(p)) if p[x] == b and c[x] == q]
            for x in kids:
                dfs(x)

    dfs(0)
    for rec, b in zip(nodes, order):
        assert rec["h"] == info[b][3] and rec["h0"] == info[0][3] and rec["time"] == 0.75
        assert tuple(rec["off"]) == tuple(info[b][2]) and rec["origin"] == [0.0, 0.0, 0.0]
        Ufile = np.moveaxis(rec["U"], [0, 1, 2, 3], [3, 2, 1, 0])  # -> [l][k][j][i]
        np.testing.assert_array_equal(Ufile[:, m], U[b][:, m])
        if b == 0:
            np.testing.assert_array_equal(rec["FO"], fo)
        else:
            assert not rec["FO"].any()
    cfg = amrhydro.euler_config(n=n, grid_dim=0.5)
    B2, t, alpha = amrhydro.Batch.tree_checkpoint_read(path, cfg)
    assert t == 0.75 and alpha == 1.0 and B2.nb == len(p)
    B2.tree_fill_ghosts()
    U2 = B2.download()
    # the read-back tree is in depth-first order: compare node by node
    for k, b in enumerate(order):
        np.testing.assert_array_equal(U2[k][:, m], U[b][:, m])
    np.testing.assert_array_equal(B2.tree_flow_off(), fo)
    # both continue identically
    for X in (B, B2):
        X.tree_step(1e-4, 3)
    U, U2 = B.download(), B2.download()
    for k, b in enumerate(order):
        np.testing.assert_array_equal(U2[k][:, m], U[b][:, m])
    np.testing.assert_array_equal(B2.tree_flow_off(), B.tree_flow_off())
