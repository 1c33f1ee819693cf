"""Pins the oracle's per-pencil chain (to_prim, reconstruction, to_con, eigenvalues, State flux,
KT combine), floor and outflow fix against the reference's own Reconstruct / State classes,
compiled from reconstruct.cpp and euler/state.cpp by oracle/ref/Makefile.  Bit-exact.

The reference build only exists in the container that has the reference sources; elsewhere
these tests are skipped and the committed golden vectors (test_oracle_golden.py) pin the same
functions."""
import numpy as np
import pytest

import cases
import orc

# The reference-built libraries are loaded by the tests that use them, never at import: collecting
# this module (e.g. for a GPU-only run) must not map them into the process.
pytestmark = pytest.mark.skipif(not orc.ref_lib_built("euler"),
                                reason="oracle/_ref not built (no reference sources)")


@pytest.fixture(scope="module")
def REF():
    return orc.ref_lib("euler")


@pytest.fixture(scope="module")
def REFPPM():
    L = orc.ref_lib("ppm")
    if L is None:
        pytest.skip("PPM reference build missing")
    return L


def ref_pencil(L, d, q):
    G = q.shape[0]
    F = np.full((G, 6), np.nan)
    e = np.full(2 * G, np.nan)
    L.ref_pencil_flux(d, orc.dptr(np.ascontiguousarray(q)), orc.dptr(F), orc.dptr(e))
    return F, e.reshape(G, 2)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_minmod_pencils_bitexact(d, REF):
    rng = np.random.default_rng(100 + d)
    cfg = orc.euler_cfg(n=8, g=2)
    P = cases.random_pencils(rng, 400, 12)
    for q in P:
        Fr, er = ref_pencil(REF, d, q)
        Fo, eo = orc.pencil_flux(cfg, d, q)
        np.testing.assert_array_equal(Fo[2:11], Fr[2:11])
        np.testing.assert_array_equal(eo[2:11], er[2:11])


def test_mach_pencils_hit_dual_energy_branch():
    """The 'mach' family really exercises pg's tau branch (euler/state.cpp:158)."""
    rng = np.random.default_rng(7)
    P = cases.random_pencils(rng, 50, 12, kind="mach")
    hit = 0
    for q in P:
        for u in q:
            ek = 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / u[0]
            hit += (u[4] - ek) < 0.001 * u[4]
    assert hit > 50


def test_ppm_pencils_bitexact(REFPPM):
    rng = np.random.default_rng(5)
    cfg = orc.euler_cfg(n=8, g=2, recon=1)
    P = cases.random_pencils(rng, 300, 12)
    for q in P:
        for d in range(3):
            Fr, er = ref_pencil(REFPPM, d, q)
            Fo, eo = orc.pencil_flux(cfg, d, q)
            # faces 3..GNX-3 are the ones PPM defines (the caller uses BW=3 with PPM)
            np.testing.assert_array_equal(Fo[3:10], Fr[3:10])
            np.testing.assert_array_equal(eo[3:10], er[3:10])


def test_floor_bitexact(REF):
    rng = np.random.default_rng(11)
    cfg = orc.euler_cfg()
    P = cases.random_pencils(rng, 200, 12).reshape(-1, 6)
    # cells on both sides of the 0.1*E threshold and with a negative density
    extra = P[:40].copy()
    extra[:, 4] = 0.5 * (extra[:, 1] ** 2 + extra[:, 2] ** 2 + extra[:, 3] ** 2) / extra[:, 0] * 1.1111
    extra[:10, 0] = -1e-3
    for u in np.concatenate([P, extra]):
        a = u.copy()
        b = u.copy()
        REF.ref_floor(orc.dptr(a))
        orc.lib().orc_floor(orc.C.byref(cfg), orc.dptr(b))
        np.testing.assert_array_equal(a, b)


def test_outflow_bitexact(REF):
    rng = np.random.default_rng(12)
    cfg = orc.euler_cfg()
    P = cases.random_pencils(rng, 50, 12).reshape(-1, 6)
    for u in P:
        for face in range(6):
            a = u.copy()
            b = u.copy()
            REF.ref_outflow(face, orc.dptr(a))
            orc.lib().orc_outflow(orc.C.byref(cfg), face, orc.dptr(b))
            np.testing.assert_array_equal(a, b)
