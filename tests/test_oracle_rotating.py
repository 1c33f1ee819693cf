"""The rotating-frame state restatement (oracle/rotating.cpp) pinned by what the physics fixes.
The reference's binary/state_ztwd.cpp cannot be built here on its own (its Binary:: statics live in
binary/binary.cpp, which needs Silo), so these checks stand in for a compiled comparison:
  * to_con inverts to_prim;
  * the cylindrical-component fluxes equal the Cartesian momentum flux projected on e_R and
    the inertial angular momentum flux (signs and position components of every term);
  * a y-mirrored block gives the y-mirrored update bit for bit (Omega = 0: s_R even, l_z odd);
  * the floor renormalises species and donor/accretor densities to rho and solves the EOS for T;
  * outflow removes the inflow velocity and keeps the internal energy;
  * a uniform gravitational field adds rho g dt to s_z."""
import numpy as np
import pytest

import orc
import rot_cases as rc

NF = 13


def rand_state(rng, cfg, X, omega):
    u = np.zeros(NF)
    rho = rng.uniform(0.5, 2.0)
    v = rng.normal(0, 0.3, 3)
    u[0] = rho
    w = rng.uniform(0.1, 0.9)
    u[1], u[2] = rho * w, rho * (1 - w)
    R = np.hypot(X[0], X[1])
    vt = (X[0] * v[1] - X[1] * v[0]) / R
    u[3] = rho * (X[0] * v[0] + X[1] * v[1]) / R
    u[4] = rho * R * (vt + omega * R)
    u[5] = rho * v[2]
    u[6] = rho * rng.uniform(-1, 0)
    eps = rng.uniform(0.5, 2.0)
    u[7] = eps + 0.5 * rho * (v @ v)
    u[8] = eps
    u[9] = eps / rho * 4.0 / 3.0 * 0.6  # any positive T
    spec = rng.uniform(0.1, 1.0, 3)
    u[10:13] = rho * spec / spec.sum()
    return u, v


@pytest.mark.parametrize("omega", [0.0, 0.7])
def test_to_con_inverts_to_prim(omega):
    cfg = orc.rot_cfg(omega=omega)
    rng = np.random.default_rng(1)
    for _ in range(200):
        X = rng.uniform(-1, 1, 3)
        u, _ = rand_state(rng, cfg, X, omega)
        p, _, _ = orc.rot_state(cfg, 0, u.copy(), X)
        back, _, _ = orc.rot_state(cfg, 1, p.copy(), X)
        np.testing.assert_allclose(back, u, rtol=1e-13, atol=1e-15)
        # primitive momenta are the rotating-frame velocity components
        assert abs(p[5] - u[5] / u[0]) <= 1e-15 * abs(u[5] / u[0]) + 1e-300


@pytest.mark.parametrize("omega,omega0", [(0.0, 0.0), (0.45, 0.45), (0.45, 0.3), (0.0, -0.2)])
def test_fluxes_are_the_cartesian_fluxes_in_cylindrical_components(omega, omega0):
    """Omega0 != Omega: the energy flux carries the pressure work of the frame velocity difference
    (Omega - Omega0) x r (State::dvx/dvy, binary/state_ztwd.cpp:256-262, used at :360-388)"""
    cfg = orc.rot_cfg(omega=omega, omega0=omega0)
    rng = np.random.default_rng(2)
    for _ in range(100):
        X = rng.uniform(-1, 1, 3)
        u, v = rand_state(rng, cfg, X, omega)
        R = np.hypot(X[0], X[1])
        rho = u[0]
        abar = u[0] / (u[10] / 4.0 + u[11] / 14.12 + u[12] / 17.52)
        p = rho * u[9] * (1 + abar / 2) / abar  # the ideal EOS
        e_r = np.array([X[0], X[1]]) / R
        for d in range(3):
            _, F, _ = orc.rot_state(cfg, 2 + d, u.copy(), X)
            vd = v[d]
            # Cartesian rotating-frame momentum flux through a face normal to d
            mom = rho * v * vd
            mom[d] += p
            # s_R = rho v.e_R: its flux is the e_R projection
            assert np.isclose(F[3], mom[:2] @ e_r, rtol=1e-12, atol=1e-14)
            # l_z = x (rho v_y + rho Omega x) - y (rho v_x - rho Omega y): flux x F_y - y F_x with
            # the rotation term advected
            lz_flux = X[0] * mom[1] - X[1] * mom[0] + rho * omega * R * R * vd
            assert np.isclose(F[4], lz_flux, rtol=1e-12, atol=1e-14)
            assert np.isclose(F[5], mom[2], rtol=1e-12, atol=1e-14)
            assert np.isclose(F[0], rho * vd, rtol=1e-13)
            # energy: E v_d + p (v_d + dv_d), dv = (Omega - Omega0) z x r
            dv = np.array([-X[1], X[0], 0.0]) * (omega - omega0)
            assert np.isclose(F[7], u[7] * vd + p * (vd + dv[d]), rtol=1e-12, atol=1e-14)
            cs = np.sqrt(5.0 / 3.0 * p / rho)
            assert np.isclose(F[13], abs(vd) + cs, rtol=1e-13)


def mirror_y(A, lz_sign=-1.0, sr_sign=1.0):
    """Reflect padded block arrays in y (axis -2); l_z changes sign at Omega = 0."""
    B = A[..., ::-1, :].copy()
    if A.ndim == 5:
        B[:, 4] *= lz_sign
        B[:, 3] *= sr_sign
    return B


def test_y_mirror_symmetry_is_bitwise():
    cfg = orc.rot_cfg(omega=0.0, omega0=0.0)
    B = rc.batch(counts=(8,), seed=4)
    U, phi, x0, h = B["U"], B["phi"], B["x0"], B["h"]
    U[:, 4] -= U[:, 0] * 0  # omega = 0 in this test: l_z is the plain angular momentum already
    # mirrored geometry: y -> -y, so the block origin's y goes to -(y0 + n h)
    x0m = x0.copy()
    x0m[:, 1] = -(x0[:, 1] + 8 * h)
    Um = mirror_y(U)
    phim = phi[..., ::-1, :].copy()
    dt = 1e-4
    a, am = U.copy(), Um.copy()
    orc.rot_outflow_ghosts(cfg, a, x0, h)
    orc.rot_outflow_ghosts(cfg, am, x0m, h)
    np.testing.assert_array_equal(mirror_y(a), am)
    a0, am0 = a.copy(), am.copy()
    orc.rot_substep(cfg, a, a0, phi, x0, h, dt, 1.0)
    orc.rot_substep(cfg, am, am0, phim, x0m, h, dt, 1.0)
    I = (slice(None), slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    np.testing.assert_array_equal(mirror_y(a)[I], am[I])


def test_floor_renormalises_and_solves_T():
    cfg = orc.rot_cfg(omega=0.3)
    rng = np.random.default_rng(5)
    X = np.array([0.3, -0.2, 0.1])
    u, _ = rand_state(rng, cfg, X, 0.3)
    u[10:13] *= np.array([1.2, 0.7, 1.1])   # species no longer sum to rho
    u[1] *= 1.3                             # nor rho_acc + rho_don
    out, _, err = orc.rot_state(cfg, 5, u.copy(), X)
    assert err == 0
    assert np.isclose(out[10:13].sum(), out[0], rtol=1e-15)
    assert np.isclose(out[1] + out[2], out[0], rtol=1e-15)
    # T: the ideal gas e = T (1 + zbar)/(abar (gamma - 1)) per unit mass equals ei/rho
    abar = out[0] / (out[10] / 4.0 + out[11] / 14.12 + out[12] / 17.52)
    R = np.hypot(X[0], X[1])
    st = out[4] / R - out[0] * R * 0.3
    ei = out[7] - 0.5 * (out[3] ** 2 + st ** 2 + out[5] ** 2) / out[0]
    T = ei / out[0] * abar * (2.0 / 3.0) / (1 + abar / 2)
    assert np.isclose(out[9], T, rtol=1e-9)
    assert out[8] == ei  # tau takes the internal energy when it is > 0.1 E


@pytest.mark.parametrize("face", range(6))
def test_outflow_removes_inflow_velocity(face):
    cfg = orc.rot_cfg(omega=0.0, omega0=0.0)
    rng = np.random.default_rng(10 + face)
    X = np.array([0.35, 0.25, 0.15])
    u, v = rand_state(rng, cfg, X, 0.0)
    d, up = face >> 1, face & 1
    v[d] = (-0.2 if up else 0.2)  # inflow through the node face
    R = np.hypot(X[0], X[1])
    rho = u[0]
    vt = (X[0] * v[1] - X[1] * v[0]) / R
    eps = u[7] - 0.5 * (u[3] ** 2 + (u[4] / R) ** 2 + u[5] ** 2) / rho
    u[3] = rho * (X[0] * v[0] + X[1] * v[1]) / R
    u[4] = rho * R * vt
    u[5] = rho * v[2]
    u[7] = eps + 0.5 * rho * (v @ v)
    out, _, _ = orc.rot_state(cfg, 6, u.copy(), X, np.array([face] + [0.0] * 13))
    sr, st, sz = out[3], out[4] / R, out[5]
    vx = (X[0] * sr - X[1] * st) / R / rho
    vy = (X[1] * sr + X[0] * st) / R / rho
    got = np.array([vx, vy, sz / rho])
    want = v.copy()
    want[d] = 0.0
    np.testing.assert_allclose(got, want, atol=1e-14)
    e_in = out[7] - 0.5 * (sr ** 2 + st ** 2 + sz ** 2) / rho
    assert np.isclose(e_in, eps, rtol=1e-13)


def test_uniform_gravity_adds_rho_g_dt_to_sz():
    cfg = orc.rot_cfg(omega=0.0, omega0=0.0)
    B = rc.batch(counts=(8,), seed=6)
    U, x0, h = B["U"], B["x0"], B["h"]
    orc.rot_outflow_ghosts(cfg, U, x0, h)
    P = 10
    gz = 0.75
    k = np.arange(P) - 1 + 0.5
    zc = x0[:, 2][:, None] + k[None, :] * h[:, None]
    phi_lin = np.broadcast_to((-gz * zc)[:, :, None, None], (U.shape[0], P, P, P)).copy()
    phi0 = np.zeros_like(phi_lin)
    dt = 1e-5
    a, b = U.copy(), U.copy()
    # the pot field enters the flux: set it consistently for each potential first
    orc.rot_substep(cfg, a, U.copy(), phi_lin, x0, h, 0.0, 1.0)
    orc.rot_substep(cfg, b, U.copy(), phi0, x0, h, 0.0, 1.0)
    orc.rot_outflow_ghosts(cfg, a, x0, h)
    orc.rot_outflow_ghosts(cfg, b, x0, h)
    a1, b1 = a.copy(), b.copy()
    orc.rot_substep(cfg, a1, a.copy(), phi_lin, x0, h, dt, 1.0)
    orc.rot_substep(cfg, b1, b.copy(), phi0, x0, h, dt, 1.0)
    I = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    dsz = (a1[:, 5] - a[:, 5])[I] - (b1[:, 5] - b[:, 5])[I]
    np.testing.assert_allclose(dsz, a[:, 0][I] * gz * dt, rtol=1e-6)
