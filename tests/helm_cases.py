"""Inputs of the Helmholtz EOS tests: the synthetic table and seeded states (eos_t records)."""
import numpy as np

import orc  # noqa: F401  (puts nothing on the path; kept for symmetry with the other helpers)
from gpu_util import amrhydro  # noqa: F401  (makes the package importable)
import workloads

A_NK = workloads.HELM_AVO * workloads.HELM_KERG

_rows = None


def table_rows():
    global _rows
    if _rows is None:
        _rows = workloads.helmholtz_table()
    return _rows


def states(n, seed, log_rho=(-4.0, 8.0), log_T=(4.0, 9.5)):
    """n states (n, 10) in eos_t order: rho and T log-uniform, abar in [4, 17.5] with zbar = abar/2
    (ye = 1/2, He/C/O-like), the other fields zero."""
    rng = np.random.default_rng(seed)
    st = np.zeros((n, 10))
    st[:, 4] = rng.uniform(4.0, 17.5, n)
    st[:, 5] = st[:, 4] / 2.0
    st[:, 6] = 10.0 ** rng.uniform(*log_rho, n)
    st[:, 7] = 10.0 ** rng.uniform(*log_T, n)
    return st


def ideal_pressure(st):
    """Ions plus non-degenerate electrons: p = N_A k T rho (1/abar + ye)."""
    return A_NK * st[:, 7] * st[:, 6] * (1.0 / st[:, 4] + st[:, 5] / st[:, 4])
