"""Seeded state generators for tests (ideal gas, gamma = 5/3).  Conserved order: rho, sx, sy, sz,
E, tau with tau = (E - ek)^(1/gamma) (euler/state.cpp:86-92)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "3d-eulerian-amr_amd"))
import workloads  # noqa: E402

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


def c2_fast(nb, n=16, g=2, seed=20261203, jump_frac=0.05):
    """The benchmark's flat-batch generator (3d-eulerian-amr_amd/workloads.py)."""
    return workloads.flat_batch(nb, n=n, g=g, seed=seed, jump_frac=jump_frac)
