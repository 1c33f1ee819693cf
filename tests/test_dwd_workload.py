"""The double-white-dwarf workload and its table (workloads.dwd_batch, helmholtz_table_dwd) pinned by
what the mathematics fixes, on the CPU:
  * the cold ideal electron gas of cold_electrons equals a direct quadrature of its defining
    integrals, and tends to its non-relativistic and ultra-relativistic limits;
  * every column of the table is the derivative of the f column its name says (finite differences
    on the table's own nodes), so the columns' order and signs are right;
  * the oracle's helmholtz_eos on the table reproduces the closed-form pressure and energy of the
    free energy it tabulates (ions analytic) to interpolation error, radiation included;
  * the Newton solve at the workload's 1e-10 inverts e(T) to that tolerance;
  * each white dwarf has its target mass and satisfies dP/dr = -G m rho / r^2;
  * the batch's fields are consistent (species and donor/accretor sums, table range)."""
import numpy as np
import pytest

import orc
from gpu_util import amrhydro  # noqa: F401  (puts the package on the path)
import workloads as w

ABAR = w.DWD_ABAR
ZBAR = ABAR / 2.0


@pytest.fixture(scope="module")
def rows():
    r = w.helmholtz_table_dwd()
    orc.helm_table(r)
    orc.helm_set_units()
    return r


def test_cold_electrons_equal_their_integrals():
    from scipy.integrate import quad
    d = np.array([1e-9, 1e-2, 30.0, 1e3, 9.7e4, 1.0e5, 3e6, 1e8, 1e12, 1e15])
    P, u, dP = w.cold_electrons(d)
    x = np.cbrt(d / w.FERMI_B)
    fq = np.array([8 * quad(lambda t: t ** 4 / np.sqrt(1 + t * t), 0, xi, epsabs=0, epsrel=1e-13, limit=200)[0]
                   for xi in x])
    gq = np.array([24 * quad(lambda t: t ** 4 / (1 + np.sqrt(1 + t * t)), 0, xi, epsabs=0, epsrel=1e-13, limit=200)[0]
                   for xi in x])
    np.testing.assert_allclose(P, w.FERMI_A * fq, rtol=1e-11)
    np.testing.assert_allclose(u, w.FERMI_A * gq, rtol=1e-11)
    # dP/dd against a centred difference
    h = 1e-5
    P2, _, _ = w.cold_electrons(d * (1 + h))
    P1, _, _ = w.cold_electrons(d * (1 - h))
    np.testing.assert_allclose((P2 - P1) / (2 * h * d), dP, rtol=1e-8)
    # limits: non-relativistic P = A (8/5 x^5 - 4/7 x^7), u = A (12/5 x^5 - 3/7 x^7);
    # ultra-relativistic P = A (2 x^4 - 2 x^2 + ...), u = A (6 x^4 - 8 x^3 + ...)
    x0 = x[0]
    assert abs(P[0] / (w.FERMI_A * (1.6 * x0 ** 5 - 4.0 / 7.0 * x0 ** 7)) - 1) < 1e-14
    assert abs(u[0] / (w.FERMI_A * (2.4 * x0 ** 5 - 3.0 / 7.0 * x0 ** 7)) - 1) < 1e-14
    x1 = x[-1]
    assert abs(P[-1] / (w.FERMI_A * (2 * x1 ** 4 - 2 * x1 ** 2)) - 1) < 1e-5
    assert abs(u[-1] / (w.FERMI_A * (6 * x1 ** 4 - 8 * x1 ** 3)) - 1) < 1e-5


def test_table_columns_are_the_named_derivatives(rows):
    """f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt: each column against centred differences of the
    column it derives from, on interior nodes (second-order differences on a log grid)"""
    I, J = w.HELM_IMAX, w.HELM_JMAX
    R = rows.reshape(J, I, 9)
    d = 10.0 ** (-12.0 + np.arange(I) * (27.0 / (I - 1)))
    T = 10.0 ** (3.0 + np.arange(J) * (10.0 / (J - 1)))
    jj = np.arange(40, J - 40, 37)
    ii = np.arange(40, I - 40, 41)

    def dd(col):  # d/dd on nodes ii (non-uniform centred difference)
        a = R[np.ix_(jj, ii + 1)][..., col]
        b = R[np.ix_(jj, ii - 1)][..., col]
        c = R[np.ix_(jj, ii)][..., col]
        h1, h2 = d[ii] - d[ii - 1], d[ii + 1] - d[ii]
        return (a * h1 ** 2 - b * h2 ** 2 - c * (h1 ** 2 - h2 ** 2)) / (h1 * h2 * (h1 + h2))

    def dT(col):
        a = R[np.ix_(jj + 1, ii)][..., col]
        b = R[np.ix_(jj - 1, ii)][..., col]
        c = R[np.ix_(jj, ii)][..., col]
        h1, h2 = (T[jj] - T[jj - 1])[:, None], (T[jj + 1] - T[jj])[:, None]
        return (a * h1 ** 2 - b * h2 ** 2 - c * (h1 ** 2 - h2 ** 2)) / (h1 * h2 * (h1 + h2))

    at = lambda col: R[np.ix_(jj, ii)][..., col]  # noqa: E731
    # (column, derivative of column, axis)
    for col, base, ax in ((1, 0, "d"), (2, 0, "t"), (3, 1, "d"), (4, 2, "t"), (5, 1, "t"), (6, 3, "t"),
                          (7, 4, "d"), (8, 7, "d")):
        num = dd(base) if ax == "d" else dT(base)
        ref = at(col)
        scale = np.abs(ref) + 1e-30
        err = np.abs(num - ref) / scale
        assert np.median(err) < 2e-2, (col, np.median(err))  # O(h^2) of a 6% log grid; a wrong column is O(1)


def test_eos_reproduces_the_tabulated_free_energy(rows):
    rng = np.random.default_rng(5)
    n = 3000
    st = np.zeros((n, 10))
    st[:, 4], st[:, 5] = ABAR, ZBAR
    st[:, 6] = 10.0 ** rng.uniform(-3.0, 8.5, n)
    st[:, 7] = 10.0 ** rng.uniform(5.0, 9.5, n)
    q = st.copy()
    assert not orc.helm_eos(q).any()
    p, e = w.dwd_eos(st[:, 6], st[:, 7], ABAR, ZBAR)
    assert np.max(np.abs(q[:, 0] / p - 1)) < 1e-6
    assert np.max(np.abs(q[:, 1] / e - 1)) < 1e-6
    assert np.all(q[:, 3] > 0) and np.all(q[:, 2] > 0)
    # radiation: at rho = 1e-2, T = 3e9 it dominates -- p -> a T^4 / 3 within the gas terms
    s2 = np.zeros((1, 10))
    s2[0, 4:8] = ABAR, ZBAR, 1e-2, 3e9
    assert not orc.helm_eos(s2).any()
    prad = w.HELM_ASOL * 3e9 ** 4 / 3.0
    assert 1.0 < s2[0, 0] / prad < 1.01


def test_newton_to_1e10_inverts_e(rows):
    rng = np.random.default_rng(6)
    n = 2000
    st = np.zeros((n, 10))
    st[:, 4], st[:, 5] = ABAR, ZBAR
    st[:, 6] = 10.0 ** rng.uniform(0.0, 7.5, n)
    st[:, 7] = 10.0 ** rng.uniform(6.0, 9.0, n)
    q = st.copy()
    assert not orc.helm_eos(q).any()
    e0 = q[:, 1].copy()
    cfg = orc.rot_cfg(eos=1, abar=w.DWD_ROT_ABAR, newton_tol=1e-10)
    # the rotating floor's Newton through the oracle: a state with E = rho e at the starting T * 1.2
    for k in range(0, n, 200):
        u = np.zeros(13)
        u[0] = st[k, 6]
        u[1] = u[2] = 0.5 * u[0]
        u[7] = u[8] = st[k, 6] * e0[k]
        u[9] = 1.2 * st[k, 7]
        u[11] = u[0]
        X = np.array([0.3, 0.2, 0.1])
        out, _, err = orc.rot_state(cfg, 5, u.copy(), X)  # op 5: the floor (Newton for T)
        assert err == 0
        r = np.zeros((1, 10))
        r[0, 4:8] = ABAR, ZBAR, st[k, 6], out[9]
        orc.helm_eos(r)
        assert abs(r[0, 1] / e0[k] - 1) < 3e-10, (k, r[0, 1] / e0[k] - 1)


@pytest.mark.parametrize("mass", [0.6, 0.9])
def test_white_dwarf_is_in_hydrostatic_equilibrium(mass):
    r, rho, m, phi = w.white_dwarf(mass)
    assert abs(m[-1] / w.HELM_MSOL - mass) < 1e-6
    P, _, _ = w.cold_electrons(0.5 * rho)
    dPdr = np.gradient(P, r)
    rhs = -w.HELM_G * m * rho / r ** 2
    inner = slice(10, len(r) * 9 // 10)
    assert np.max(np.abs(dPdr[inner] - rhs[inner]) / np.abs(rhs[inner]).max()) < 2e-3
    # dphi/dr = G m / r^2 and phi continuous into -G M / r
    dphi = np.gradient(phi, r)
    assert np.max(np.abs(dphi[inner] - w.HELM_G * m[inner] / r[inner] ** 2)) < 2e-3 * np.abs(dphi).max()
    assert abs(phi[-1] / (-w.HELM_G * m[-1] / r[-1]) - 1) < 1e-12


def test_dwd_batch_fields_are_consistent():
    B = w.dwd_batch(counts=(8, 40))
    U = B["U"]
    rho = U[:, 0]
    np.testing.assert_allclose(U[:, 1] + U[:, 2], rho, rtol=1e-14)
    np.testing.assert_allclose(U[:, 10] + U[:, 11] + U[:, 12], rho, rtol=1e-14)
    np.testing.assert_array_equal(U[:, 7], U[:, 8])
    assert set(np.unique(U[:, 9])) <= {w.DWD_T, w.DWD_T_BG}
    L, M, t, _ = B["units"]
    rho_cgs = rho * M / L ** 3
    assert rho_cgs.min() >= w.DWD_RHO_BG * (1 - 1e-12) and rho_cgs.max() < 2e7
    assert B["abar"][1] == ABAR
