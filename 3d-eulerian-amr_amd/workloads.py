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


def flat_batch_sample_rate():
    """Cells per block of the flat batch, for throughput arithmetic."""
    return 16 ** 3
