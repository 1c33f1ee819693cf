"""Silo-equivalent output through the C ABI: the mesh and zone fields GridNode::output hands to Silo
(grid/grid_node.cpp:1315-1356, :1422-1487), against a numpy restatement of that recursion on the
oracle's tree, and the legacy VTK file that carries them, parsed back independently."""
import numpy as np
import pytest

import orc
import tree_cases as tc
from gpu_util import amrhydro, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

n, g = 8, 2
GNX = n + 2 * g
G1 = GNX + 1


def reference_output(T, nn, grid_dim, z_reflect):
    """GridNode::output (grid/grid_node.cpp:1315-1356) in numpy, depth first from the root."""
    info = [T.info(b) for b in range(nn)]
    kids = {b: [None] * 8 for b in range(nn)}
    for b in range(1, nn):
        lv, loc, off, h = info[b]
        p = [q for q in range(nn) if info[q][0] == lv - 1 and all(info[q][1][d] == loc[d] // 2 for d in range(3))][0]
        c = sum(((loc[d] & 1) << d) for d in range(3))
        kids[p][c] = b
    h0 = info[0][3]
    xs, ys, zs, nl, vals = [], [], [], [], []
    pi = [0]

    def rec(b):
        lv, loc, off, h = info[b]
        ii = np.arange(G1)
        xf = (off[0] + ii) * h - grid_dim - g * h0 - 0.0
        yf = (off[1] + ii) * h - grid_dim - g * h0 - 0.0
        zf = ((off[2] + ii) * h - g * h0 - 0.0) if z_reflect else ((off[2] + ii) * h - grid_dim - g * h0)
        Z, Y, X = np.meshgrid(zf, yf, xf, indexing="ij")
        xs.append(X.ravel()), ys.append(Y.ravel()), zs.append(Z.ravel())
        U = T.U(b)
        for k in range(g, GNX - g):
            for j in range(g, GNX - g):
                for i in range(g, GNX - g):
                    c = (2 * i) // GNX | ((2 * j) // GNX) << 1 | ((2 * k) // GNX) << 2
                    if kids[b][c] is not None:
                        continue
                    at = lambda a, bb, cc: a + G1 * (bb + G1 * cc) + pi[0]
                    nl.append([at(i, j, k), at(i + 1, j, k), at(i + 1, j + 1, k), at(i, j + 1, k),
                               at(i, j, k + 1), at(i + 1, j, k + 1), at(i + 1, j + 1, k + 1), at(i, j + 1, k + 1)])
                    vals.append(U[:, k, j, i].copy())
        pi[0] += G1 ** 3
        for c in range(8):
            if kids[b][c] is not None:
                rec(kids[b][c])

    rec(0)
    return (np.concatenate(xs), np.concatenate(ys), np.concatenate(zs), np.array(nl, np.int32),
            np.array(vals).T.copy())


def read_vtk(path):
    """Legacy binary VTK unstructured grid, big-endian (independent of the library's writer)."""
    buf = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        e = buf.index(b"\n", pos)
        s = buf[pos:e].decode()
        pos = e + 1
        return s

    def arr(dtype, count):
        nonlocal pos
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=pos)
        pos += a.nbytes
        return a

    assert line().startswith("# vtk DataFile")
    title = line()
    assert line() == "BINARY" and line() == "DATASET UNSTRUCTURED_GRID"
    npts = int(line().split()[1])
    pts = arr(">f8", 3 * npts).reshape(npts, 3)
    while True:
        s = line()
        if s:
            break
    nz = int(s.split()[1])
    cells = arr(">i4", 9 * nz).reshape(nz, 9)
    while not line():
        pass
    types = arr(">i4", nz)
    out = {}
    while pos < len(buf):
        s = line()
        if not s or s.startswith("CELL_DATA"):
            continue
        name = s.split()[1]
        assert line() == "LOOKUP_TABLE default"
        out[name] = arr(">f8", nz)
    return title, pts, cells, types, out


@pytest.mark.parametrize("z_reflect", [0, 1])
def test_output_mesh_matches_grid_node_output(tmp_path, z_reflect):
    cfg = orc.euler_cfg(n=n, g=g, grid_dim=0.5, z_reflect=z_reflect)
    T, p, c = tc.build(cfg, tc.partial_spec())
    tc.blob_ic(T)
    T.enforce_boundaries()
    T.inject_from_children()
    B = amrhydro.Batch(amrhydro.euler_config(n=n, grid_dim=0.5, z_reflect=z_reflect), len(p), 1.0)
    B.tree_set(p, c)
    B.upload(tc.padded_state(T))
    x, y, z, nl, v = B.tree_output_mesh()
    rx, ry, rz, rnl, rv = reference_output(T, len(p), 0.5, z_reflect)
    np.testing.assert_array_equal(x, rx)
    np.testing.assert_array_equal(y, ry)
    np.testing.assert_array_equal(z, rz)
    np.testing.assert_array_equal(nl, rnl)
    np.testing.assert_array_equal(v, rv)
    path = str(tmp_path / "out.vtk")
    B.tree_output_vtk(path, time=0.125)
    title, pts, cells, types, fields = read_vtk(path)
    assert "0.125" in title
    np.testing.assert_array_equal(pts, np.stack([rx, ry, rz], axis=1))
    assert (cells[:, 0] == 8).all() and (types == 12).all()
    np.testing.assert_array_equal(cells[:, 1:], rnl)
    for l, name in enumerate(("d", "sx", "sy", "sz", "et", "tau")):
        np.testing.assert_array_equal(fields[name], rv[l])
    # the zones tile the domain exactly once: total zone volume = (2 grid_dim)^3
    P = pts[cells[:, 1:]]
    vol = np.prod(P[:, 6] - P[:, 0], axis=1).sum()
    assert abs(vol - 1.0) < 1e-12
    say("output: %d points, %d zones, %d nodes" % (len(pts), len(cells), len(p)))
