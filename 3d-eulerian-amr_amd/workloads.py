"""Seeded synthetic inputs for the benchmark configurations (ideal gas, gamma = 5/3).

All arrays are padded SoA float64: (nblocks, 6, G, G, G), G = n + 2g, x fastest; conserved fields
rho, sx, sy, sz, E, tau with tau = eps^(1/gamma) (the value State::floor sets, euler/state.cpp:86-92).
"""
import numpy as np

GAMMA = 5.0 / 3.0
SEED = 20261203


def flat_batch(nb, n=16, g=2, seed=SEED, jump_frac=0.05):
    """Batched flat stencil: lognormal density (mu=-3, sigma=1.5); entropy K ~ U[0.5, 2] and
    p = K rho^gamma; velocity components ~ N(0, 0.3 c_s); a fraction `jump_frac` of the blocks carry
    a planar 10x density and pressure jump on a plane of random orientation through the block.
    Full size: nb = 4096, n = 16."""
    rng = np.random.default_rng(seed)
    G = n + 2 * g
    shp = (nb, G, G, G)
    rho = np.exp(rng.normal(-3.0, 1.5, shp))
    K = rng.uniform(0.5, 2.0, shp)
    p = K * rho ** GAMMA
    jb = np.nonzero(rng.random(nb) < jump_frac)[0]
    ii = (np.arange(G) - g + 0.5) / n
    Z, Y, X = np.meshgrid(ii, ii, ii, indexing="ij")
    for b in jb:
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        off = rng.uniform(0.3, 0.7)
        side = (X * nrm[0] + Y * nrm[1] + Z * nrm[2]) > off * np.abs(nrm).sum()
        rho[b] = np.where(side, rho[b] * 10.0, rho[b])
        p[b] = np.where(side, p[b] * 10.0, p[b])
    cs = np.sqrt(GAMMA * p / rho)
    U = np.empty((nb, 6, G, G, G))
    eps = p / (GAMMA - 1.0)
    U[:, 0] = rho
    ek = np.zeros(shp)
    for d in range(3):
        v = rng.normal(0.0, 1.0, shp) * 0.3 * cs
        U[:, 1 + d] = rho * v
        ek += 0.5 * rho * v * v
    U[:, 4] = eps + ek
    U[:, 5] = eps ** (1.0 / GAMMA)
    return U


# Uneven multilevel octree leaves: 8 at level 0 up to 30k at the finest, 59,518 in all.  The
# endpoints and the ~60k total are not one geometric series, so the counts are fixed here.
LEVEL_COUNTS = (8, 60, 450, 2000, 7000, 20000, 30000)


def multilevel_layout(counts=LEVEL_COUNTS, seed=SEED):
    """Leaf positions per level.  Level l tiles the cube [-1, 1]^3 with 2^(l+1) blocks per edge; the
    leaves of a level are the blocks nearest the centre (ties broken by a seeded jitter), so finer
    levels sit on the density core.  Returns (level, origin[3]) arrays."""
    rng = np.random.default_rng(seed)
    lev, org = [], []
    for l, cnt in enumerate(counts):
        m = 2 ** (l + 1)
        if cnt > m ** 3:
            raise ValueError("level %d has only %d block positions" % (l, m ** 3))
        edge = 2.0 / m
        idx = np.arange(m)
        c = -1.0 + (idx + 0.5) * edge
        Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
        r = np.sqrt(X * X + Y * Y + Z * Z).ravel() + rng.uniform(0, 1e-6, m ** 3)
        pick = np.argsort(r, kind="stable")[:cnt]
        o = np.stack([X.ravel()[pick], Y.ravel()[pick], Z.ravel()[pick]], axis=1) - 0.5 * edge
        lev.append(np.full(cnt, l))
        org.append(o)
    return np.concatenate(lev), np.concatenate(org)


def multilevel(counts=LEVEL_COUNTS, n=8, g=2, sigma=0.2, omega=0.5, seed=SEED):
    """Uneven multilevel batch: leaves of 7 levels, n^3 interior + g ghosts, cell width
    h_l = 2/(n 2^(l+1)).  Smooth Gaussian density core rho = 1e-6 + exp(-r^2/(2 sigma^2)), solid-body
    rotation v = omega z_hat x r, pressure p = 0.1 rho^gamma + 1e-6 (subsonic everywhere).  Ghost
    cells hold the same analytic field.  Returns (U, h, level)."""
    lev, org = multilevel_layout(counts, seed)
    nb = len(lev)
    G = n + 2 * g
    U = np.empty((nb, 6, G, G, G))
    h = 2.0 / (n * 2.0 ** (lev + 1))
    off = (np.arange(G) - g + 0.5)
    for s in range(0, nb, 4096):
        e = min(nb, s + 4096)
        hh = h[s:e, None]
        x = org[s:e, 0:1] + off[None, :] * hh
        y = org[s:e, 1:2] + off[None, :] * hh
        z = org[s:e, 2:3] + off[None, :] * hh
        X = x[:, None, None, :]
        Y = y[:, None, :, None]
        Z = z[:, :, None, None]
        r2 = X * X + Y * Y + Z * Z
        rho = 1e-6 + np.exp(-r2 / (2 * sigma * sigma))
        vx = -omega * Y + 0 * X
        vy = omega * X + 0 * Y
        vz = 0 * r2
        p = 0.1 * rho ** GAMMA + 1e-6
        eps = p / (GAMMA - 1.0)
        U[s:e, 0] = rho
        U[s:e, 1] = rho * vx
        U[s:e, 2] = rho * vy
        U[s:e, 3] = rho * vz
        U[s:e, 4] = eps + 0.5 * rho * (vx * vx + vy * vy)
        U[s:e, 5] = eps ** (1.0 / GAMMA)
    return U, h, lev


# ---- Helmholtz EOS table ---------------------------------------------------------------------
HELM_IMAX, HELM_JMAX = 1081, 401  # helmholtz/helmeos.cpp:7-8
HELM_AVO, HELM_KERG = 6.0221417930e23, 1.380650424e-16  # helmholtz/const.h
HELM_H, HELM_ME = 6.6260689633e-27, 9.1093821545e-28


def helmholtz_table():
    """Synthetic electron free-energy table at the reference's dimensions and ranges
    (log10(ye*rho) in [-12, 15] over 1081 nodes, log10 T in [3, 13] over 401 nodes,
    helmholtz/helmeos.cpp:7-8, :78-86), built from the free energy of a non-degenerate ideal
    electron gas per unit ye*rho:
        f(d, T) = A T (ln d - 1.5 ln T + C),  A = N_A k,  C = ln(N_A/2 (h^2/(2 pi m_e k))^1.5) - 1,
    with its derivatives in closed form.  Rows are (f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt),
    temperature outer, density inner: the helmholtz.table layout.  The EOS built on it is an ideal
    gas of ions and electrons, p = N_A k T rho (1/abar + ye), to the biquintic interpolation error."""
    I, J = HELM_IMAX, HELM_JMAX
    d = 10.0 ** (-12.0 + np.arange(I) * (27.0 / (I - 1)))
    T = 10.0 ** (3.0 + np.arange(J) * (10.0 / (J - 1)))
    A = HELM_AVO * HELM_KERG
    C = np.log(HELM_AVO / 2.0 * (HELM_H ** 2 / (2.0 * np.pi * HELM_ME * HELM_KERG)) ** 1.5) - 1.0
    dd, TT = np.meshgrid(d, T)  # shape (J, I): temperature outer
    L = np.log(dd) - 1.5 * np.log(TT) + C
    rows = np.empty((J, I, 9))
    rows[..., 0] = A * TT * L               # f
    rows[..., 1] = A * TT / dd              # f_d
    rows[..., 2] = A * (L - 1.5)            # f_t
    rows[..., 3] = -A * TT / dd ** 2        # f_dd
    rows[..., 4] = -1.5 * A / TT            # f_tt
    rows[..., 5] = A / dd                   # f_dt
    rows[..., 6] = -A / dd ** 2             # f_ddt
    rows[..., 7] = 0.0                      # f_dtt
    rows[..., 8] = 0.0                      # f_ddtt
    return np.ascontiguousarray(rows.reshape(J * I, 9))


# cgs constants of helmholtz/const.h used by the double-white-dwarf table and model
HELM_CLIGHT, HELM_ASOL = 2.99792458e10, 4.0 * 5.67051e-5 / 2.99792458e10
HELM_G, HELM_MSOL = 6.6742867e-8, 1.9892e33
# the cold ideal Fermi gas of electrons: n_e = N_A d (d = ye rho), x = p_F/(m_e c) = (d/FERMI_B)^(1/3),
# P = FERMI_A * 8 int_0^x t^4/sqrt(1+t^2) dt (Chandrasekhar)
FERMI_B = 8.0 * np.pi * HELM_ME ** 3 * HELM_CLIGHT ** 3 / (3.0 * HELM_H ** 3 * HELM_AVO)
FERMI_A = np.pi * HELM_ME ** 4 * HELM_CLIGHT ** 5 / (3.0 * HELM_H ** 3)


def cold_electrons(d):
    """The T = 0 ideal (non-interacting) electron gas at electron density N_A d (any degeneracy and
    relativity), x = (d/B)^(1/3): pressure P = A f(x) and kinetic energy density u = A g(x) with
        f(x) = 8 int_0^x t^4 (1+t^2)^-1/2 dt = x (2x^2 - 3) sqrt(1+x^2) + 3 asinh(x),
        g(x) = 24 int_0^x t^2 (sqrt(1+t^2) - 1) dt = 8 x^3 (sqrt(1+x^2) - 1) - f(x),
    and dP/dd = 8 A x^5 / (3 d sqrt(1+x^2)).  Below x = 0.1 the closed forms cancel, so the
    integrands' binomial series are integrated term by term there (truncation below 1e-20); above,
    sqrt(1+x^2) - 1 is formed as x^2/(sqrt(1+x^2) + 1).  Returns (P, u, dP/dd) in erg/cm^3 and
    erg/cm^3 per g/cm^3."""
    from scipy.special import binom
    d = np.asarray(d, dtype=np.float64)
    x = np.cbrt(d / FERMI_B)
    sq = np.sqrt(1.0 + x * x)
    f = x * (2.0 * x * x - 3.0) * sq + 3.0 * np.arcsinh(x)
    g = 8.0 * x ** 3 * (x * x / (sq + 1.0)) - f
    small = x < 0.1
    if np.any(small):
        k = np.arange(12)
        cf = 8.0 * binom(-0.5, k) / (2 * k + 5)        # t^4 (1+t^2)^-1/2
        cg = 24.0 * binom(0.5, k + 1) / (2 * k + 5)    # t^2 (sqrt(1+t^2) - 1)
        xs = x[small][..., None]
        f = np.where(small, 0.0, f)
        g = np.where(small, 0.0, g)
        f[small] = (cf * xs ** (2 * k + 5)).sum(axis=-1)
        g[small] = (cg * xs ** (2 * k + 5)).sum(axis=-1)
    P, u = FERMI_A * f, FERMI_A * g
    dPdd = 8.0 * FERMI_A * x ** 5 / (3.0 * d * sq)
    return P, u, dPdd


def helmholtz_table_dwd():
    """The electron + radiation free-energy table of the double-white-dwarf workload, at the
    reference's dimensions and ranges (helmholtz/helmeos.cpp:7-8, :78-86), per unit ye*rho = d as
    helmholtz_eos reads it (p = d^2 f_d, e = ye (f - T f_t), :380-387):
      * the cold ideal Fermi gas, f = u(d)/d with u the kinetic energy density of cold_electrons:
        f_d = P/d^2, f_dd = P'(d)/d^2 - 2P/d^3, no T dependence -- the degenerate part;
      * the non-degenerate ideal electron gas of helmholtz_table (A T (ln d - 1.5 ln T + C));
      * radiation, f = -a T^4/(3d): p = a T^4/3, e = a T^4/rho, s = 4 a T^3/(3 rho) per gram.
    Every column is the closed-form derivative of that sum, so the table is thermodynamically
    consistent.  Ions stay analytic in the EOS."""
    I, J = HELM_IMAX, HELM_JMAX
    rows = helmholtz_table().reshape(J, I, 9)
    d = 10.0 ** (-12.0 + np.arange(I) * (27.0 / (I - 1)))
    T = 10.0 ** (3.0 + np.arange(J) * (10.0 / (J - 1)))
    P, u, dPdd = cold_electrons(d)
    rows[..., 0] += (u / d)[None, :]
    rows[..., 1] += (P / d ** 2)[None, :]
    rows[..., 3] += (dPdd / d ** 2 - 2.0 * P / d ** 3)[None, :]
    a = HELM_ASOL
    dd, TT = np.meshgrid(d, T)
    rows[..., 0] += -a * TT ** 4 / (3.0 * dd)              # f
    rows[..., 1] += a * TT ** 4 / (3.0 * dd ** 2)          # f_d
    rows[..., 2] += -4.0 * a * TT ** 3 / (3.0 * dd)        # f_t
    rows[..., 3] += -2.0 * a * TT ** 4 / (3.0 * dd ** 3)   # f_dd
    rows[..., 4] += -4.0 * a * TT ** 2 / dd                # f_tt
    rows[..., 5] += 4.0 * a * TT ** 3 / (3.0 * dd ** 2)    # f_dt
    rows[..., 6] += -8.0 * a * TT ** 3 / (3.0 * dd ** 3)   # f_ddt
    rows[..., 7] += 4.0 * a * TT ** 2 / dd ** 2            # f_dtt
    rows[..., 8] += -8.0 * a * TT ** 2 / dd ** 3           # f_ddtt
    return np.ascontiguousarray(rows.reshape(J * I, 9))


def dwd_eos(rho, T, abar, zbar):
    """Closed-form p (dyn/cm^2) and e (erg/g) of the free energy helmholtz_table_dwd tabulates, with
    the EOS's analytic ions (helmholtz/helmeos.cpp:186-198): the values its interpolation approaches."""
    rho, T = np.asarray(rho, np.float64), np.asarray(T, np.float64)
    ye = zbar / abar
    d = ye * rho
    A = HELM_AVO * HELM_KERG
    P, u, _ = cold_electrons(d)
    p = P + A * T * d + HELM_ASOL * T ** 4 / 3.0 + A * T * rho / abar
    e = u / rho + 1.5 * A * T * ye + HELM_ASOL * T ** 4 / rho + 1.5 * A * T / abar
    return p, e


def write_helmholtz_table(path, rows):
    """The rows as a helmholtz.table text file (9 values per line, 17 significant digits, so
    %le parsing restores them exactly)."""
    np.savetxt(path, rows, fmt="%.17e")


# ---- rotating binary (13-field state of binary/state_ztwd.cpp) -------------------------------
BINARY_COUNTS = (8, 64, 400, 2000, 6000, 11528)  # 20,000 leaves of 8^3 over 6 levels
ROT_ABAR = (4.0, 14.12, 17.52)                   # binary/state_ztwd.cpp:28
# code time unit (s) of the Helmholtz runs: with cm = g = 1 it puts the stars' centres at ~1e7 K
HELM_UNITS = (1.0, 1.0, 4.9e-9, 1.0)


def lane_emden(n=1.5, dxi=1e-3):
    """theta(xi) of the Lane-Emden equation of index n by RK4; returns (xi, theta, xi1, omega)
    with xi1 the first zero and omega = -xi1^2 theta'(xi1)."""
    def rhs(x, y):
        th, dth = y
        return np.array([dth, -max(th, 0.0) ** n - 2.0 * dth / x])
    x, y = dxi, np.array([1.0 - dxi * dxi / 6.0, -dxi / 3.0])
    xs, ts = [0.0, x], [1.0, y[0]]
    while y[0] > 0.0:
        k1 = rhs(x, y)
        k2 = rhs(x + dxi / 2, y + dxi / 2 * k1)
        k3 = rhs(x + dxi / 2, y + dxi / 2 * k2)
        k4 = rhs(x + dxi, y + dxi * k3)
        yn = y + dxi / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        if yn[0] <= 0.0:
            f = y[0] / (y[0] - yn[0])
            xi1 = x + f * dxi
            return np.array(xs), np.array(ts), xi1, -xi1 * xi1 * (y[1] + f * (yn[1] - y[1]))
        x, y = x + dxi, yn
        xs.append(x)
        ts.append(y[0])
    raise RuntimeError("no surface")


def _binary_model(q=0.7, r_donor=0.2):
    """The two n = 1.5 polytropes of binary_batch: Lane-Emden profile, radii, central densities,
    masses, centres, the polytropic K, the Keplerian Omega and the background."""
    xi, th, xi1, om = lane_emden(1.5)
    a = 2.5 * r_donor
    r_acc = r_donor * q ** (1.0 / 3.0)
    rc_d, rc_a = 1.0, 1.0 / (q * q)
    m_d = 4.0 * np.pi * rc_d * r_donor ** 3 * om / xi1 ** 3
    m_a = 4.0 * np.pi * rc_a * r_acc ** 3 * om / xi1 ** 3
    K = 4.0 * np.pi * r_donor ** 2 * rc_d ** (1.0 - 1.0 / 1.5) / (2.5 * xi1 ** 2)
    omega = np.sqrt((m_a + m_d) / a ** 3)
    xa, xd = -a * m_d / (m_a + m_d), a * m_a / (m_a + m_d)
    rho_bg, p_bg = 1e-4 * rc_d, 1e-3 * K * rc_d ** (5.0 / 3.0)
    return dict(xi=xi, th=th, xi1=xi1, r_acc=r_acc, r_donor=r_donor, rc_a=rc_a, rc_d=rc_d, m_a=m_a, m_d=m_d,
                K=K, omega=omega, xa=xa, xd=xd, rho_bg=rho_bg, p_bg=p_bg)


def _binary_fields(M, X, Y, Z, eos):
    """The 13 conserved fields of the binary model at cell centres (X, Y, Z), at rest in the frame."""
    def star(r, R, rc):
        x = np.clip(r / R, 0.0, 1.0) * M["xi1"]
        t = np.interp(x, M["xi"], M["th"])
        return np.where(r < R, rc * np.maximum(t, 0.0) ** 1.5, 0.0)

    xa, xd, omega = M["xa"], M["xd"], M["omega"]
    ra = np.sqrt((X - xa) ** 2 + Y ** 2 + Z ** 2)
    rd = np.sqrt((X - xd) ** 2 + Y ** 2 + Z ** 2)
    rho_a, rho_d = star(ra, M["r_acc"], M["rc_a"]), star(rd, M["r_donor"], M["rc_d"])
    rho = rho_a + rho_d + M["rho_bg"]
    p = M["K"] * (rho_a + rho_d) ** (5.0 / 3.0) + M["p_bg"]
    w_a = rho_a / rho
    R2 = X * X + Y * Y
    abar = 1.0 / ((1.0 - w_a) / ROT_ABAR[0] + w_a / ROT_ABAR[1])
    zbar = abar / 2.0
    eps = 1.5 * p
    if eos == "ideal":
        T = p * abar / (rho * (1.0 + zbar))
    else:
        cm, gr, sec, _ = HELM_UNITS
        p_cgs = p * (gr / (cm * sec * sec))
        rho_cgs = rho * (gr / cm ** 3)
        T = p_cgs / (HELM_AVO * HELM_KERG * rho_cgs * (1.0 / abar + zbar / abar))
    u = np.empty((13,) + np.broadcast(X, Y, Z).shape)
    u[0] = rho
    u[1] = rho * w_a
    u[2] = rho * (1.0 - w_a)
    u[3] = 0.0
    u[4] = rho * R2 * omega
    u[5] = 0.0
    u[6] = rho * (_binary_phi(M, X, Y, Z) - 0.5 * R2 * omega * omega)
    u[7] = eps
    u[8] = eps
    u[9] = T
    u[10] = rho * (1.0 - w_a)
    u[11] = rho * w_a
    u[12] = 0.0
    return u


def _binary_phi(M, X, Y, Z):
    """the two stars' Plummer-softened point-mass potential (softening R/2)"""
    ra = np.sqrt((X - M["xa"]) ** 2 + Y ** 2 + Z ** 2)
    rd = np.sqrt((X - M["xd"]) ** 2 + Y ** 2 + Z ** 2)
    return -M["m_a"] / np.sqrt(ra ** 2 + (0.5 * M["r_acc"]) ** 2) - M["m_d"] / np.sqrt(rd ** 2 + (0.5 * M["r_donor"]) ** 2)


def binary_batch(counts=BINARY_COUNTS, n=8, g=3, eos="ideal", q=0.7, r_donor=0.2, seed=SEED):
    """Two n = 1.5 polytropes (Lane-Emden, one K) with mass ratio q = M_donor/M_accretor, separation
    2.5 R_donor, centre of mass at the origin, at rest in the frame rotating at the Keplerian
    Omega (G = 1); a hot tenuous background.  Accretor C/O (species 1), donor and background He
    (species 0).  Leaves as in multilevel_layout (nearest the origin finest).  The potential is the
    two stars' Plummer-softened point-mass potential (softening R/2), given on (n+2)^3 per block.
    T from p with the ideal EOS (p = rho T (1 + zbar)/abar) or the synthetic Helmholtz table in
    HELM_UNITS.  Returns a dict with U (nb, 13, G, G, G), phi, x0 (nb, 3), h (nb), level, omega."""
    lev, org = multilevel_layout(counts, seed)
    nb = len(lev)
    G = n + 2 * g
    M = _binary_model(q, r_donor)
    h = 2.0 / (n * 2.0 ** (lev + 1))
    U = np.empty((nb, 13, G, G, G))
    P = n + 2
    phi = np.empty((nb, P, P, P))
    for s in range(0, nb, 2048):
        e = min(nb, s + 2048)
        hh = h[s:e, None]

        def grid(off):
            x = org[s:e, 0:1] + off[None, :] * hh
            y = org[s:e, 1:2] + off[None, :] * hh
            z = org[s:e, 2:3] + off[None, :] * hh
            return x[:, None, None, :], y[:, None, :, None], z[:, :, None, None]

        X, Y, Z = grid(np.arange(G) - g + 0.5)
        U[s:e] = np.moveaxis(_binary_fields(M, X, Y, Z, eos), 0, 1)
        phi[s:e] = _binary_phi(M, *grid(np.arange(P) - 1 + 0.5))
    x0 = np.ascontiguousarray(org)
    return dict(U=U, phi=phi, x0=x0, h=h, level=lev, omega=M["omega"])


# ---- the double white dwarf of the Helmholtz-EOS workload ------------------------------------------
DWD_MASSES = (0.6, 0.9)  # solar masses: donor, accretor
DWD_ABAR = 96.0 / 7.0    # C/O 50/50 by mass: 1/abar = 0.5/12 + 0.5/16; zbar = abar/2 = 6.857
DWD_ROT_ABAR = (4.0, DWD_ABAR, 17.52)  # the C/O species slot carries abar 13.714
DWD_T = 1.0e7            # K, isothermal stars
DWD_RHO_BG, DWD_T_BG = 1.0, 1.0e8  # g/cm^3, K: a hot tenuous envelope


def white_dwarf(mass, ye=0.5, n=4000):
    """A cold white dwarf of `mass` (solar masses) in hydrostatic equilibrium, dP/dr = -G m rho/r^2,
    dm/dr = 4 pi r^2 rho, on the T = 0 ideal electron gas of cold_electrons (electron pressure only:
    the ions and the finite temperature add < 1% at 1e7 K).  The central density is found by
    bisection.  Returns (r, rho, m, phi) on n radii from the centre to the surface (cgs), phi the
    star's own potential, -G M/r outside."""
    from scipy.integrate import solve_ivp

    def rhs(r, y):
        m, x = y
        x = max(x, 0.0)
        rho = FERMI_B * x ** 3 / ye
        dPdx = 8.0 * FERMI_A * x ** 4 / np.sqrt(1.0 + x * x)
        return [4.0 * np.pi * r * r * rho, -HELM_G * m * rho / (r * r * dPdx) if x > 0.0 else 0.0]

    def integrate(rho_c, dense=False):
        xc = np.cbrt(ye * rho_c / FERMI_B)
        r0 = 1.0e3
        m0 = 4.0 / 3.0 * np.pi * r0 ** 3 * rho_c
        surf = lambda r, y: y[1] - 1e-4 * xc  # noqa: E731
        surf.terminal = True
        sol = solve_ivp(rhs, (r0, 1e11), [m0, xc], events=surf, rtol=1e-10, atol=[1e-12 * HELM_MSOL, 1e-14],
                        dense_output=dense, max_step=1e7)
        return sol

    lo, hi = 1e4, 1e10
    for _ in range(48):  # log-bisection: 6 decades to 1e-13 relative
        mid = np.sqrt(lo * hi)
        M = integrate(mid).y[0, -1] / HELM_MSOL
        lo, hi = (mid, hi) if M < mass else (lo, mid)
    sol = integrate(np.sqrt(lo * hi), dense=True)
    R = sol.t[-1]
    r = np.linspace(sol.t[0], R, n)
    m, x = sol.sol(r)
    rho = FERMI_B * np.maximum(x, 0.0) ** 3 / ye
    # phi(r) = -G M/R - int_r^R G m/r'^2 dr'
    gm = HELM_G * m / (r * r)
    tail = np.concatenate([np.cumsum((0.5 * (gm[1:] + gm[:-1]) * np.diff(r))[::-1])[::-1], [0.0]])
    phi = -HELM_G * m[-1] / R - tail
    return r, rho, m, phi


_DWD_MODEL = None


def dwd_model():
    """The 0.6 + 0.9 solar-mass C/O pair: radii, centres (centre of mass at the origin), separation
    3 donor radii (the donor near its Roche lobe, 0.344 a at q = 2/3), the Keplerian Omega, and the
    code units: length L = R_donor / 0.2 (so the donor spans 0.2 of the root domain's half edge, as
    in the polytrope workload), mass 1 solar mass, time (L^3 / (G Msun))^1/2; T in K.  units =
    (cm, g, s, K) as amrh_eos_set_units takes them."""
    global _DWD_MODEL
    if _DWD_MODEL is not None:
        return _DWD_MODEL
    stars = [white_dwarf(mm) for mm in DWD_MASSES]
    (rd, rhod, md, phid), (ra, rhoa, ma, phia) = stars
    Rd, Ra = rd[-1], ra[-1]
    Md, Ma = md[-1], ma[-1]
    a = 3.0 * Rd
    xd, xa = a * Ma / (Ma + Md), -a * Md / (Ma + Md)
    L = Rd / 0.2
    t = np.sqrt(L ** 3 / (HELM_G * HELM_MSOL))
    omega = np.sqrt(HELM_G * (Ma + Md) / a ** 3) * t  # code units
    _DWD_MODEL = dict(stars=stars, R=(Rd, Ra), M=(Md, Ma), centres=(xd, xa), a=a, L=L, t=t, omega=omega,
                      units=(L, HELM_MSOL, t, 1.0))
    return _DWD_MODEL


def _dwd_phi(M, X, Y, Z):
    """The two stars' own spherical potentials at (X, Y, Z) (code units), in code units."""
    L, tu = M["L"], M["t"]
    phi = 0.0
    for (r_, rho_, m_, phi_), xc in zip(M["stars"], M["centres"]):
        rr = np.sqrt((X * L - xc) ** 2 + (Y * L) ** 2 + (Z * L) ** 2)
        phi = phi + np.where(rr < r_[-1], np.interp(rr, r_, phi_), -HELM_G * m_[-1] / np.maximum(rr, 1.0))
    return phi / (L / tu) ** 2


def _dwd_fields(M, X, Y, Z):
    """The 13 conserved fields (code units) of the double white dwarf at cell centres (code units),
    at rest in the frame rotating at Omega; T = DWD_T in the stars, DWD_T_BG in the envelope."""
    L, tu = M["L"], M["t"]
    rho_u, e_u = HELM_MSOL / L ** 3, (L / tu) ** 2
    rho = np.full(np.broadcast(X, Y, Z).shape, DWD_RHO_BG)
    rho_acc = np.zeros_like(rho)
    for (r_, rho_, m_, phi_), xc, acc in zip(M["stars"], M["centres"], (False, True)):
        rr = np.sqrt((X * L - xc) ** 2 + (Y * L) ** 2 + (Z * L) ** 2)
        rs = np.where(rr < r_[-1], np.interp(rr, r_, rho_), 0.0)
        rho = rho + rs
        if acc:
            rho_acc = rho_acc + rs
    phi_c = _dwd_phi(M, X, Y, Z)
    star = rho > 10.0 * DWD_RHO_BG
    T = np.where(star, DWD_T, DWD_T_BG)
    e = np.empty_like(rho)
    e[star] = dwd_eos(rho[star], DWD_T, DWD_ABAR, DWD_ABAR / 2.0)[1]
    rest = ~star
    e[rest] = dwd_eos(rho[rest], DWD_T_BG, DWD_ABAR, DWD_ABAR / 2.0)[1]
    omega = M["omega"]
    R2 = X * X + Y * Y
    rho_c = rho / rho_u
    u = np.empty((13,) + rho.shape)
    u[0] = rho_c
    u[1] = rho_acc / rho_u + 0.5 * DWD_RHO_BG / rho_u
    u[2] = rho_c - u[1]
    u[3] = 0.0
    u[4] = rho_c * R2 * omega
    u[5] = 0.0
    u[6] = rho_c * (phi_c - 0.5 * R2 * omega * omega)
    u[7] = rho_c * e / e_u
    u[8] = u[7]
    u[9] = T
    u[10] = 0.0
    u[11] = rho_c
    u[12] = 0.0
    return u


def dwd_batch(counts=BINARY_COUNTS, n=8, g=3, seed=SEED):
    """The Helmholtz-EOS workload: a 0.6 + 0.9 solar-mass C/O double white dwarf (abar 13.714,
    zbar 6.857), each star a cold white dwarf in hydrostatic equilibrium (white_dwarf) at 1e7 K,
    in a 1 g/cm^3, 1e8 K envelope, at rest in the frame rotating at the Keplerian Omega, on
    BINARY_COUNTS leaves of 8^3 (multilevel_layout, nearest the origin finest) in the code units of
    dwd_model.  The given potential is the two stars' own spherical potentials.  Energies follow
    from the closed-form free energy that helmholtz_table_dwd tabulates.  Returns binary_batch's
    dict plus units (for amrh_eos_set_units) and abar (the species' mass numbers)."""
    lev, org = multilevel_layout(counts, seed)
    nb = len(lev)
    G = n + 2 * g
    M = dwd_model()
    h = 2.0 / (n * 2.0 ** (lev + 1))
    U = np.empty((nb, 13, G, G, G))
    P = n + 2
    phi = np.empty((nb, P, P, P))
    for s in range(0, nb, 1024):
        e = min(nb, s + 1024)
        hh = h[s:e, None]

        def grid(off):
            x = org[s:e, 0:1] + off[None, :] * hh
            y = org[s:e, 1:2] + off[None, :] * hh
            z = org[s:e, 2:3] + off[None, :] * hh
            return x[:, None, None, :], y[:, None, :, None], z[:, :, None, None]

        U[s:e] = np.moveaxis(_dwd_fields(M, *grid(np.arange(G) - g + 0.5)), 0, 1)
        phi[s:e] = _dwd_phi(M, *grid(np.arange(P) - 1 + 0.5))
    return dict(U=U, phi=phi, x0=np.ascontiguousarray(org), h=h, level=lev, omega=M["omega"], units=M["units"],
                abar=DWD_ROT_ABAR)


def binary_tree_state(level, loc, n=8, g=3, eos="ideal"):
    """binary_batch's model on the nodes of an octree over [-1, 1]^3 (binary_tree): U (nodes, 13,
    G, G, G) at the padded cell centres, the point-mass phi on (n+2)^3, the low interior corner x0
    and the cell width h of every node, and Omega."""
    M = _binary_model()
    G, P = n + 2 * g, n + 2
    nn = len(level)
    level = np.asarray(level)
    loc = np.asarray(loc, dtype=np.float64).reshape(nn, 3)
    h = 2.0 / n / 2.0 ** level
    x0 = -1.0 + loc * n * h[:, None]
    U = np.empty((nn, 13, G, G, G))
    phi = np.empty((nn, P, P, P))
    for s in range(0, nn, 2048):
        e = min(nn, s + 2048)
        hh = h[s:e, None]

        def grid(off):
            x = x0[s:e, 0:1] + off[None, :] * hh
            y = x0[s:e, 1:2] + off[None, :] * hh
            z = x0[s:e, 2:3] + off[None, :] * hh
            return x[:, None, None, :], y[:, None, :, None], z[:, :, None, None]

        U[s:e] = np.moveaxis(_binary_fields(M, *grid(np.arange(G) - g + 0.5), eos), 0, 1)
        phi[s:e] = _binary_phi(M, *grid(np.arange(P) - 1 + 0.5))
    return dict(U=U, phi=phi, x0=np.ascontiguousarray(x0), h=h, omega=M["omega"])


def sedov_tree(maxlevel=2, n=8, g=2, grid_dim=0.5):
    """Fully refined octree (1 + 8 + 64 nodes for maxlevel 2, breadth first: every parent before
    its children, child c with x, y, z bits c & 1, c >> 1 & 1, c >> 2 & 1) over the cube
    [-grid_dim, grid_dim]^3: ambient rho = 1, p = 1e-5 at rest; energy 1.0 deposited uniformly in
    the finest-level cells whose centres lie within 2 finest cells of the origin.  The node at
    level l with integer location L has padded cell p centred at -grid_dim + (L n + p - g + 1/2) h_l
    (grid/grid.cpp:500-526).  Every padded cell holds the analytic field; the library's ghost fill
    and restriction make them consistent.  Returns (U (nodes, 6, G, G, G), parent, child)."""
    parent, child, level, loc = [-1], [-1], [0], [(0, 0, 0)]
    frontier = [0]
    for _ in range(maxlevel):
        new = []
        for p in frontier:
            for c in range(8):
                parent.append(p)
                child.append(c)
                level.append(level[p] + 1)
                loc.append(tuple(2 * loc[p][d] + ((c >> d) & 1) for d in range(3)))
                new.append(len(parent) - 1)
        frontier = new
    G = n + 2 * g
    h0 = 2.0 * grid_dim / n
    hf = h0 / 2 ** maxlevel
    r = 2.0 * hf
    m = int(round(2.0 * grid_dim / hf))
    c1 = -grid_dim + (np.arange(m) + 0.5) * hf
    Zc, Yc, Xc = np.meshgrid(c1, c1, c1, indexing="ij")
    ndep = int((Xc ** 2 + Yc ** 2 + Zc ** 2 <= r * r).sum())
    edep = 1.0 / (ndep * hf ** 3)
    U = np.empty((len(parent), 6, G, G, G))
    for b, (l, L) in enumerate(zip(level, loc)):
        h = h0 / 2 ** l
        x, y, z = (-grid_dim + (L[d] * n + np.arange(G) - g + 0.5) * h for d in range(3))
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        E = np.full(X.shape, 1e-5 / (GAMMA - 1.0))
        E = np.where(X * X + Y * Y + Z * Z <= r * r, E + edep, E)
        U[b, 0] = 1.0
        U[b, 1:4] = 0.0
        U[b, 4] = E
        U[b, 5] = E ** (1.0 / GAMMA)
    return U, np.array(parent, np.int32), np.array(child, np.int32)


# ---- self-gravity: the binary's density on a properly nested octree ---------------------------
def _binary_stars(q=0.7, r_donor=0.2):
    """centres, radii and central densities of the two n = 1.5 polytropes of binary_batch"""
    xi, th, xi1, om = lane_emden(1.5)
    a = 2.5 * r_donor
    r_acc = r_donor * q ** (1.0 / 3.0)
    rc_d, rc_a = 1.0, 1.0 / (q * q)
    m_d = 4.0 * np.pi * rc_d * r_donor ** 3 * om / xi1 ** 3
    m_a = 4.0 * np.pi * rc_a * r_acc ** 3 * om / xi1 ** 3
    xa, xd = -a * m_d / (m_a + m_d), a * m_a / (m_a + m_d)
    return (xi, th, xi1), ((xa, r_acc, rc_a, m_a), (xd, r_donor, rc_d, m_d))


def binary_tree(maxlevel=5, n=8, margin=1.25):
    """A properly nested octree over [-1, 1]^3 refined to maxlevel where a node's box comes within
    margin radii of either star of binary_batch, closed so that every refined node's 26 same-level
    neighbours inside the domain exist (the nesting the reference's regrid enforces,
    grid/grid_node.cpp:687-900).  Breadth first, children in index order.  Returns
    (parent, child, level, loc)."""
    _, stars = _binary_stars()
    refine = [set() for _ in range(maxlevel)]
    for l in range(maxlevel):
        w = 2.0 / 2 ** l
        m = 2 ** l
        lo = -1.0 + np.arange(m) * w
        for (xc, R, _, _) in stars:
            r = margin * R
            # box-sphere distance per axis
            dx = np.maximum(np.maximum(lo - xc, xc - (lo + w)), 0.0)
            dy = np.maximum(np.maximum(lo, -(lo + w)), 0.0)
            d2 = dx[None, None, :] ** 2 + dy[None, :, None] ** 2 + dy[:, None, None] ** 2
            for k, j, i in zip(*np.nonzero(d2 <= r * r)):
                refine[l].add((int(i), int(j), int(k)))
    changed = True
    while changed:
        changed = False
        for l in range(1, maxlevel):
            m = 2 ** l
            for loc in list(refine[l]):
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            nb = (loc[0] + dx, loc[1] + dy, loc[2] + dz)
                            if min(nb) < 0 or max(nb) >= m:
                                continue
                            pa = (nb[0] // 2, nb[1] // 2, nb[2] // 2)
                            if pa not in refine[l - 1]:
                                refine[l - 1].add(pa)
                                changed = True
    parent, child, level, loc = [-1], [0], [0], [(0, 0, 0)]
    frontier = [0]
    for l in range(maxlevel):
        new = []
        for p in frontier:
            if loc[p] not in refine[l]:
                continue
            for c in range(8):
                parent.append(p)
                child.append(c)
                level.append(l + 1)
                loc.append(tuple(2 * loc[p][d] + ((c >> d) & 1) for d in range(3)))
                new.append(len(parent) - 1)
        frontier = new
    return np.array(parent, np.int32), np.array(child, np.int32), np.array(level), loc


def binary_source(level, loc, n=8, G=1.0):
    """S = 4 pi G rho of binary_batch's two stars and background at the (n+2)^3 cell centres of
    every node (one ghost layer): (nodes, n+2, n+2, n+2)."""
    (xi, th, xi1), stars = _binary_stars()
    P = n + 2
    S = np.empty((len(level), P, P, P))
    for b, (l, L) in enumerate(zip(level, loc)):
        h = 2.0 / n / 2 ** l
        x, y, z = (-1.0 + (L[d] * n + np.arange(P) - 0.5) * h for d in range(3))
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        rho = np.full(X.shape, 1e-4)
        for (xc, R, rc, _) in stars:
            r = np.sqrt((X - xc) ** 2 + Y ** 2 + Z ** 2)
            t = np.interp(np.clip(r / R, 0.0, 1.0) * xi1, xi, th)
            rho += np.where(r < R, rc * np.maximum(t, 0.0) ** 1.5, 0.0)
        S[b] = 4.0 * np.pi * G * rho
    return S
