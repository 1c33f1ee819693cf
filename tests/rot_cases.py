"""Inputs of the rotating-frame state tests: small binary batches with seeded velocity
perturbations, and the conversions between Cartesian rotating-frame velocities and the state's
cylindrical momenta (binary/state_ztwd.cpp:241-267)."""
import numpy as np

from gpu_util import amrhydro  # noqa: F401  (puts the package on the path)
import workloads

G3 = 3  # PPM ghost width


def cell_centres(x0, h, n, g):
    off = np.arange(n + 2 * g) - g + 0.5
    x = x0[:, 0:1] + off[None, :] * h[:, None]
    y = x0[:, 1:2] + off[None, :] * h[:, None]
    z = x0[:, 2:3] + off[None, :] * h[:, None]
    return x[:, None, None, :], y[:, None, :, None], z[:, :, None, None]


def set_velocity(U, X, Y, omega, vx, vy, vz):
    """Momenta of the rotating-frame velocity (vx, vy, vz) at cell centres; E keeps its internal
    energy E - ek."""
    rho = U[:, 0]
    R = np.sqrt(X * X + Y * Y)
    st_old = U[:, 4] / R - rho * R * omega
    ek_old = 0.5 * (U[:, 3] ** 2 + st_old ** 2 + U[:, 5] ** 2) / rho
    eps = U[:, 7] - ek_old
    vt = (X * vy - Y * vx) / R
    U[:, 3] = rho * (X * vx + Y * vy) / R
    U[:, 4] = rho * R * (vt + omega * R)
    U[:, 5] = rho * vz
    U[:, 7] = eps + 0.5 * rho * (vx * vx + vy * vy + vz * vz)


def batch(counts=(8, 40), eos="ideal", seed=3, amp=0.3, g=G3):
    """A binary batch with random rotating-frame velocities of amplitude amp * c_s."""
    B = workloads.binary_batch(counts=counts, n=8, g=g, eos=eos)
    U = B["U"]
    rng = np.random.default_rng(seed)
    X, Y, _ = cell_centres(B["x0"], B["h"], 8, g)
    X = np.broadcast_to(X, U.shape[:1] + U.shape[2:])
    Y = np.broadcast_to(Y, U.shape[:1] + U.shape[2:])
    cs = np.sqrt(5.0 / 3.0 * (U[:, 7] / 1.5) / U[:, 0])
    v = [amp * cs * rng.standard_normal(U[:, 0].shape) for _ in range(3)]
    set_velocity(U, X, Y, B["omega"], *v)
    return B
