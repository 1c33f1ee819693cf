"""Seeded state generators for tests (ideal gas, gamma = 5/3).  Conserved order: rho, sx, sy, sz,
E, tau with tau = (E - ek)^(1/gamma) (euler/state.cpp:86-92)."""
import numpy as np

GAMMA = 5.0 / 3.0


def prim_to_con(rho, vx, vy, vz, p, tau=None):
    eps = p / (GAMMA - 1.0)
    E = eps + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    if tau is None:
        tau = np.maximum(eps, 1e-20) ** (1.0 / GAMMA)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, E, tau], axis=-1)


def random_pencils(rng, npencil, length, kind="mixed"):
    """Pencils of conserved states covering the branches of the chain:
    smooth / jumpy / flat (zero slopes) / high-Mach (dual-energy branch of pg) / opposite-sign
    velocities."""
    out = np.empty((npencil, length, 6))
    for p in range(npencil):
        k = kind if kind != "mixed" else ["smooth", "jump", "flat", "mach", "noise"][p % 5]
        x = np.linspace(0.0, 1.0, length)
        if k == "smooth":
            rho = 1.0 + 0.5 * np.sin(2 * np.pi * (x + rng.random()))
            v = [rng.normal(0, 0.5) * np.cos(2 * np.pi * (x + rng.random())) for _ in range(3)]
            pr = 0.6 + 0.3 * np.cos(2 * np.pi * (x + rng.random()))
        elif k == "jump":
            s = rng.integers(2, length - 2)
            rho = np.where(np.arange(length) < s, 1.0, 0.125) * np.exp(rng.normal(0, 0.1))
            v = [np.where(np.arange(length) < s, rng.normal(0, 1), rng.normal(0, 1)) for _ in range(3)]
            pr = np.where(np.arange(length) < s, 1.0, 0.1)
        elif k == "flat":
            rho = np.full(length, np.exp(rng.normal()))
            rho[rng.integers(0, length)] *= 1.0 + rng.random()
            v = [np.full(length, rng.normal()) for _ in range(3)]
            pr = np.full(length, np.exp(rng.normal()))
        elif k == "mach":
            # kinetic energy >> internal: E - ek < 0.001 E selects the tau branch of pg
            rho = np.exp(rng.normal(0, 0.3, length))
            v = [rng.normal(30, 5, length), rng.normal(0, 3, length), rng.normal(0, 3, length)]
            pr = np.exp(rng.normal(-3, 0.5, length))
        else:
            rho = np.exp(rng.normal(-3, 1.5, length))
            cs = np.sqrt(GAMMA * 0.01)
            v = [rng.normal(0, 0.3 * cs, length) for _ in range(3)]
            pr = rho ** GAMMA * rng.uniform(0.5, 2.0, length)
        out[p] = prim_to_con(rho, v[0], v[1], v[2], pr)
        if k == "mach":
            # tau carries its own (entropy) history, not exactly (E-ek)^(1/gamma)
            out[p, :, 5] *= rng.uniform(0.9, 1.1, length)
    return out


def c2_blocks(nb, n=16, g=2, seed=20261203, jump_frac=0.05):
    """Batched flat stencil (4096 x 16^3 + 2 ghosts in the full size): lognormal density
    (mu=-3, sigma=1.5), entropy K ~ U[0.5,2], p = K rho^gamma, v ~ N(0, 0.3 c_s); 5% of blocks carry
    a planar 10x density/pressure jump at a random orientation.  Padded SoA (nb, 6, G, G, G);
    ghosts are filled with the same field (the caller may refresh them with outflow ghosts)."""
    rng = np.random.default_rng(seed)
    G = n + 2 * g
    U = np.empty((nb, 6, G, G, G))
    ii = (np.arange(G) - g + 0.5) / n
    Z, Y, X = np.meshgrid(ii, ii, ii, indexing="ij")
    for b in range(nb):
        rho = np.exp(rng.normal(-3.0, 1.5, (G, G, G)))
        K = rng.uniform(0.5, 2.0, (G, G, G))
        p = K * rho ** GAMMA
        if rng.random() < jump_frac:
            nrm = rng.normal(size=3)
            nrm /= np.linalg.norm(nrm)
            off = rng.uniform(0.3, 0.7)
            side = (X * nrm[0] + Y * nrm[1] + Z * nrm[2]) > off * (np.abs(nrm).sum())
            rho = np.where(side, rho * 10.0, rho)
            p = np.where(side, p * 10.0, p)
        cs = np.sqrt(GAMMA * p / rho)
        v = [rng.normal(0.0, 1.0, (G, G, G)) * 0.3 * cs for _ in range(3)]
        U[b] = np.moveaxis(prim_to_con(rho, v[0], v[1], v[2], p), -1, 0)
    return U


def c2_fast(nb, n=16, g=2, seed=20261203, jump_frac=0.05):
    """Vectorised version of c2_blocks for large nb (same distribution; different draw order)."""
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
        side = (X * nrm[0] + Y * nrm[1] + Z * nrm[2]) > off * (np.abs(nrm).sum())
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
