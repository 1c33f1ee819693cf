"""The multigrid Poisson solver on the GPU (amrh_poisson_*) against the oracle's restatement of
poisson/poisson.cpp, poisson/poisson_amr_bound.cpp, poisson/poisson_phys_bound.cpp and
binary/binary.cpp:938-1019, through the C ABI:
  * ghost fill (faces, edges, corners; coarse->fine, reflection, tables), restriction and the
    forces with their coarse-fine averaging are bit-identical;
  * the monopole boundary tables agree to 1e-13 (the per-octant sums are added in another order);
  * single V-cycles and whole solves agree to 1e-12 (residual sums in another order);
  * the quadratic fixed point is reached on the device as in the oracle;
  * S = 4 pi G rho from a hydro batch, and phi handed to a rotating-state batch by device pointer."""
import numpy as np
import pytest

import orc
from gpu_util import amrhydro, say
from test_oracle_poisson import centre_refined, centres, full, refined_trees, sphere_source, tree

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

n = 8
P = n + 2
GHOST = np.ones((P, P, P), bool)
GHOST[1:-1, 1:-1, 1:-1] = False

TREES = {"centre": centre_refined(), **refined_trees()}


def pair(name, z_reflect):
    parent, child = TREES[name]
    T = orc.Poisson(n, parent, child, z_reflect=z_reflect)
    D = amrhydro.Poisson(n, parent, child, z_reflect=z_reflect)
    return T, D


def oracle_tables(T):
    t = np.zeros((T.nn, 6, P, P))
    for b in range(T.nn):
        for f in range(6):
            x = T.table(b, f)
            if x is not None:
                t[b, f] = x
    return t


def random_state(T, D, seed, fields=("phi", "S")):
    rng = np.random.default_rng(seed)
    for name in fields:
        a = rng.standard_normal((T.nn, P, P, P))
        T.set_fields(name, a)
        D.upload(name, a)


@pytest.mark.parametrize("z_reflect", [0, 1])
def test_boundary_tables_match(z_reflect):
    parent, child = tree(full(2))
    T = orc.Poisson(n, parent, child, z_reflect=z_reflect)
    D = amrhydro.Poisson(n, parent, child, z_reflect=z_reflect)
    C = centres(T, z_reflect=z_reflect)
    sphere_source(T, C, centre=(0.1, -0.05, 0.02))
    D.upload("S", T.fields("S"))
    T.prepare()
    D.prepare()
    want = oracle_tables(T)
    got = D.tables()
    m = want != 0
    assert m.sum() > 1000
    np.testing.assert_allclose(got[m], want[m], rtol=1e-13, atol=0)
    say("tables: %d entries, max rel %.2e" % (m.sum(), np.max(np.abs(got[m] / want[m] - 1))))


@pytest.mark.parametrize("name", list(TREES))
@pytest.mark.parametrize("z_reflect", [0, 1])
def test_ghosts_restriction_forces_bitwise(name, z_reflect):
    T, D = pair(name, z_reflect)
    random_state(T, D, 11)
    T.prepare()
    rng = np.random.default_rng(5)
    tabs = oracle_tables(T)
    tabs[tabs != 0] = 0.0
    for b in range(T.nn):  # random boundary data in both
        for f in range(6):
            t = T.table(b, f)
            if t is not None:
                t[:] = rng.standard_normal((P, P))
                tabs[b, f] = t
    D.tables(tabs)
    T.phi_from_children()
    D.phi_from_children()
    np.testing.assert_array_equal(D.download("phi"), T.fields("phi"))
    T.enforce_phi()
    D.phi_boundaries()
    np.testing.assert_array_equal(D.download("phi"), T.fields("phi"))
    T.forces()
    e_ref = T.error()
    e = D.forces()
    for f in ("fx", "fy", "fz"):
        got, want = D.download(f), T.fields(f)
        m = np.zeros((P, P, P), bool)
        m[1:, 1:, 1:] = True  # compute_forces writes faces 1..P-1 on each axis
        np.testing.assert_array_equal(got[:, m], want[:, m])
    np.testing.assert_allclose(e, e_ref, rtol=1e-12)


@pytest.mark.parametrize("name", ["centre", "pair", "nested"])
@pytest.mark.parametrize("z_reflect", [0, 1])
def test_vcycles_match(name, z_reflect):
    T, D = pair(name, z_reflect)
    C = centres(T, z_reflect=z_reflect)
    sphere_source(T, C, r0=0.45, centre=(-0.2, 0.15, 0.3 if z_reflect else 0.05))
    D.upload("S", T.fields("S"))
    T.prepare()
    D.tables(oracle_tables(T))  # identical boundary data on both sides
    T.forces()
    D.forces()
    for cyc in range(4):
        e_ref = T.vcycle()
        e = D.vcycle()
        phi_ref, phi = T.fields("phi"), D.download("phi")
        scale = np.max(np.abs(phi_ref))
        err = np.max(np.abs(phi - phi_ref)) / scale
        say("%s zr=%d cycle %d: error %.3e / %.3e, phi max rel diff %.2e" % (name, z_reflect, cyc, e, e_ref, err))
        assert err < 1e-12
        np.testing.assert_allclose(e, e_ref, rtol=1e-10)


@pytest.mark.parametrize("z_reflect", [0, 1])
def test_solve_matches_oracle(z_reflect):
    parent, child = TREES["nested"]
    T = orc.Poisson(n, parent, child, z_reflect=z_reflect)
    D = amrhydro.Poisson(n, parent, child, z_reflect=z_reflect)
    C = centres(T, z_reflect=z_reflect)
    sphere_source(T, C, r0=0.4, centre=(0.05, -0.1, 0.1))
    D.upload("S", T.fields("S"))
    tol = 1e-8
    it_ref, e_ref = T.solve(tol)
    it, e = D.solve(tol)
    phi_ref, phi = T.fields("phi"), D.download("phi")
    err = np.max(np.abs(phi - phi_ref)) / np.max(np.abs(phi_ref))
    say("solve zr=%d: %d / %d V-cycles, error %.2e / %.2e, phi max rel diff %.2e" % (z_reflect, it, it_ref, e, e_ref,
                                                                                    err))
    assert it == it_ref
    assert err < 1e-11


def test_quadratic_fixed_point_on_device():
    parent, child = TREES["nested"]
    T = orc.Poisson(n, parent, child, z_reflect=1)
    D = amrhydro.Poisson(n, parent, child, z_reflect=1)
    C = centres(T, z_reflect=1)
    q = np.stack([X * X + Y * Y + Z * Z + 0.5 * X * Y + 0.3 * X - 0.2 * Y for X, Y, Z, _ in C])
    D.upload("S", np.full((T.nn, P, P, P), 6.0))
    tabs = np.zeros((T.nn, 6, P, P))
    for b in range(T.nn):
        for f in range(6):
            d, g = f >> 1, (P - 1 if f & 1 else 0)
            sl = [slice(None)] * 3
            sl[2 - d] = g
            tabs[b, f] = q[b][tuple(sl)].T
    D.tables(tabs)
    it, e = D.solve(1e-10, prepare=False)
    I = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    worst = np.max(np.abs(D.download("phi")[I] - q[I]))
    say("fixed point: %d V-cycles, error %.2e, max |phi - q| %.2e" % (it, e, worst))
    assert worst < 1e-9 and it <= 60


def test_source_from_hydro_state_and_phi_to_rotating_batch():
    # S = 4 pi G rho from a tree batch of the ideal gas
    import tree_cases as tc
    cfg_o = orc.euler_cfg(n=n, g=2, grid_dim=1.0, z_reflect=0)
    # properly nested, as the reference's regrid keeps every tree (enforce_proper_nesting,
    # grid/grid_node.cpp:687-900): a coarse-fine ghost never needs a missing coarse neighbour
    To, p, c = tc.build(cfg_o, full(1) + [(3, q) for q in range(8)])
    tc.blob_ic(To)
    U = tc.padded_state(To)
    B = amrhydro.Batch(amrhydro.euler_config(n=n, grid_dim=1.0), len(p), 1.0)
    B.tree_set(p, c)
    B.upload(U)
    D = amrhydro.Poisson(n, p, c)
    G = 0.75
    D.source_from(B, G)
    S = D.download("S")
    np.testing.assert_array_equal(S[:, 1:-1, 1:-1, 1:-1], (4.0 * np.pi * G) * U[:, 0, 2:-2, 2:-2, 2:-2])
    it, e = D.solve(1e-8)
    assert e <= 1e-8
    # the solved phi feeds a rotating-state batch of the same nodes by device pointer
    rcfg = amrhydro.rotating_config(eos=amrhydro.AMRH_EOS_IDEAL, omega=0.0)
    g = rcfg.g
    Gr = n + 2 * g
    R1 = amrhydro.Batch(rcfg, len(p), 1.0)
    R2 = amrhydro.Batch(rcfg, len(p), 1.0)
    rng = np.random.default_rng(3)
    Ur = np.zeros((len(p), 13, Gr, Gr, Gr))
    Ur[:, 0] = 1.0 + 0.1 * rng.random((len(p), Gr, Gr, Gr))
    Ur[:, 1] = 0.5 * Ur[:, 0]
    Ur[:, 2] = 0.5 * Ur[:, 0]
    Ur[:, 7] = 2.0
    Ur[:, 8] = 2.0 ** 0.6
    Ur[:, 9] = 1.0
    Ur[:, 10] = Ur[:, 0]
    hs = np.array([To.info(b)[3] for b in range(len(p))])
    lv = [To.info(b)[0] for b in range(len(p))]
    loc = [To.info(b)[1] for b in range(len(p))]
    x0 = np.array([[-1.0 + loc[b][d] * n * hs[b] for d in range(3)] for b in range(len(p))])
    phi = D.download("phi")
    for R in (R1, R2):
        R.set_h(hs)
        R.set_block_origin(x0)
        R.upload(Ur)
    R1.set_potential(phi)
    R2.set_potential_device(D.phi_device())
    for R in (R1, R2):
        R.store()
        R.substep(1e-4, 1.0)
    np.testing.assert_array_equal(R1.download(), R2.download())
    say("poisson -> rotating: %d nodes, %d V-cycles, levels %d" % (len(p), it, max(lv)))
