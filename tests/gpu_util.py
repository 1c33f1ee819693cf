"""Helpers shared by the GPU tests: import of the binding, comparison metrics."""
import os
import sys

import numpy as np

import orc

PKG = os.path.join(orc.ROOT, "3d-eulerian-amr_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import amrhydro  # noqa: E402

FIELDS = ["rho", "sx", "sy", "sz", "E", "tau"]


def interior(U, g):
    return U[..., g:-g, g:-g, g:-g]


def rel_l1(a, b, g):
    """Per-field relative L1 over interior cells: sum|a-b| / sum|b| (equal cell volumes)."""
    A = interior(a, g)
    B = interior(b, g)
    out = []
    for f in range(A.shape[1]):
        num = np.abs(A[:, f] - B[:, f]).sum()
        den = np.abs(B[:, f]).sum()
        out.append(num / den if den > 0 else num)
    return np.array(out)


def say(msg):
    print(msg, flush=True)
