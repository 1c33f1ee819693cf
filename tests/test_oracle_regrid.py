"""The oracle's GridNode::inject_from_parent (grid/grid_node.cpp:193-233) pinned by what the
mathematics fixes: a uniform parent gives uniform children; a parent profile linear along one axis
(dyadic values, so every operation is exact) gives children on the same line, which fixes the sign
of every slope term of every child; restricting the children gives the parent back."""
import numpy as np
import pytest

import cases
import orc
import tree_cases as tc

G, g, n = 12, 2, 8


def two_level(axis_profile=None, const=None):
    cfg = orc.euler_cfg(n=n, g=g, grid_dim=0.5)
    T = orc.Tree(cfg)
    for c in range(8):
        T.add_child(0, c)
    U = T.U(0)
    X, Y, Z, h = tc.cell_centres(T, 0)
    if const is not None:
        cell = cases.prim_to_con(np.array(1.3), np.array(0.2), np.array(-0.1), np.array(0.05), np.array(0.7))
        U[:] = cell[:, None, None, None]
    else:
        d, slope = axis_profile
        coord = (X, Y, Z)[d]
        # a linear profile on a dyadic grid: rho = 1 + slope*x is exact in binary
        U[:] = 0.0
        U[0] = 1.0 + slope * coord
        U[4] = 2.0 + slope * coord
        U[5] = (U[4] ** 0.6)  # tau is not linear, but floor leaves it alone only if E - ek <= 0.1 E
    return T, h


def test_uniform_parent_gives_uniform_children():
    T, _ = two_level(const=True)
    for c in range(1, 9):
        T.prolong(c)
        I = T.U(c)[:, g:-g, g:-g, g:-g]
        for f in range(5):
            np.testing.assert_array_equal(I[f], T.U(0)[f, g, g, g])


@pytest.mark.parametrize("d", [0, 1, 2])
def test_linear_parent_gives_children_on_the_line(d):
    slope = 0.25
    T, h0 = two_level(axis_profile=(d, slope))
    for c in range(1, 9):
        T.prolong(c)
        X, Y, Z, h = tc.cell_centres(T, c)
        coord = (X, Y, Z)[d][g:-g, g:-g, g:-g]
        got = T.U(c)[0, g:-g, g:-g, g:-g]
        np.testing.assert_array_equal(got, 1.0 + slope * coord)


def test_restriction_of_prolongation_is_the_parent():
    T, _ = two_level(axis_profile=(1, 0.125))
    T.U(0)[1] = 0.3 * T.U(0)[0]  # add momentum, still dyadic-exact? keep comparison to roundoff
    before = T.U(0)[:, g:-g, g:-g, g:-g].copy()
    for c in range(1, 9):
        T.prolong(c)
    T.inject_from_children()
    after = T.U(0)[:, g:-g, g:-g, g:-g]
    np.testing.assert_allclose(after[:5], before[:5], rtol=1e-15, atol=1e-15)
