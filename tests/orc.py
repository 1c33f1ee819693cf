"""ctypes access to the FP64 CPU oracle (oracle/liboracle.so) and, in this container only, to the
reference-built harness (oracle/_ref/*.so).  Test infrastructure: the product never imports this."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class OrcCfg(C.Structure):
    _fields_ = [("n", C.c_int), ("g", C.c_int), ("nf", C.c_int), ("recon", C.c_int),
                ("z_reflect", C.c_int), ("gamma", C.c_double), ("rho_floor", C.c_double),
                ("ei_floor", C.c_double), ("theta", C.c_double), ("cfl", C.c_double),
                ("grid_dim", C.c_double)]


def euler_cfg(n=8, g=2, recon=0, z_reflect=0, grid_dim=1.0):
    """Ideal-gas constants of euler/state.cpp:9-11 and euler/euler_defs.h:20,46."""
    return OrcCfg(n=n, g=g, nf=6, recon=recon, z_reflect=z_reflect, gamma=5.0 / 3.0,
                  rho_floor=1.0e-20, ei_floor=1.0e-20, theta=1.3, cfl=0.4, grid_dim=grid_dim)


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        cfgp = C.POINTER(OrcCfg)
        L.orc_pencil_flux.argtypes = [cfgp, C.c_int, C.c_int, DP, DP, DP]
        L.orc_pencil_faces.argtypes = [cfgp, C.c_int, DP, DP, DP]
        L.orc_floor.argtypes = [cfgp, DP]
        L.orc_floor.restype = None
        L.orc_outflow.argtypes = [cfgp, C.c_int, DP]
        L.orc_outflow.restype = None
        L.orc_batch_substep.argtypes = [cfgp, C.c_int, DP, DP, DP, C.c_double, C.c_double, DP]
        L.orc_batch_max_eig.argtypes = [cfgp, C.c_int, DP, DP]
        L.orc_block_dt.argtypes = [cfgp, C.c_double, C.c_double]
        L.orc_block_dt.restype = C.c_double
        L.orc_batch_outflow_ghosts.argtypes = [cfgp, C.c_int, DP]
        L.orc_batch_rk_steps.argtypes = [cfgp, C.c_int, DP, DP, DP, C.c_double, C.c_int, C.c_int]
        L.orc_refshape_create.argtypes = [cfgp, C.c_int, DP, DP]
        L.orc_refshape_create.restype = C.c_void_p
        L.orc_refshape_destroy.argtypes = [C.c_void_p]
        L.orc_refshape_destroy.restype = None
        L.orc_refshape_download.argtypes = [C.c_void_p, DP]
        L.orc_refshape_max_eig.argtypes = [C.c_void_p]
        L.orc_refshape_max_eig.restype = C.c_double
        L.orc_refshape_rk_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_tree_create.argtypes = [cfgp]
        L.orc_tree_create.restype = C.c_void_p
        L.orc_tree_destroy.argtypes = [C.c_void_p]
        L.orc_tree_destroy.restype = None
        L.orc_tree_add_child.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_tree_nnodes.argtypes = [C.c_void_p]
        L.orc_tree_node_info.argtypes = [C.c_void_p, C.c_int, IP, IP, IP, DP]
        L.orc_tree_node_U.argtypes = [C.c_void_p, C.c_int]
        L.orc_tree_node_U.restype = DP
        L.orc_tree_node_F.argtypes = [C.c_void_p, C.c_int]
        L.orc_tree_node_F.restype = DP
        for fn in ("orc_tree_enforce_boundaries", "orc_tree_inject_from_children", "orc_tree_store"):
            getattr(L, fn).argtypes = [C.c_void_p]
        L.orc_tree_max_dt.argtypes = [C.c_void_p]
        L.orc_tree_max_dt.restype = C.c_double
        L.orc_tree_substep.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_tree_state_sum.argtypes = [C.c_void_p, DP]
        L.orc_tree_prolong.argtypes = [C.c_void_p, C.c_int]
        L.orc_tree_flow_off.argtypes = [C.c_void_p, DP]
        L.orc_helm_table.argtypes = [DP]
        L.orc_helm_table.restype = None
        L.orc_helm_set_units.argtypes = [C.c_double] * 4
        L.orc_helm_set_units.restype = None
        L.orc_helm_eos.argtypes = [C.c_int, DP, IP]
        L.orc_helm_compute_T.argtypes = [C.c_int, DP, IP]
        _lib = L
    return _lib


def ref_lib_built(kind="euler"):
    return os.path.exists(os.path.join(REF_DIR, "libref_%s.so" % kind))


def ref_lib(kind="euler"):
    """The reference-built harness; only in this container (needs /root/reference to build)."""
    path = os.path.join(REF_DIR, "libref_%s.so" % kind)
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ref_pencil_flux.argtypes = [C.c_int, DP, DP, DP]
    L.ref_pencil_flux.restype = None
    L.ref_pencil_faces.argtypes = [DP, DP, DP]
    L.ref_pencil_faces.restype = None
    L.ref_floor.argtypes = [DP]
    L.ref_floor.restype = None
    L.ref_outflow.argtypes = [C.c_int, DP]
    L.ref_outflow.restype = None
    L.ref_pg.argtypes = [DP]
    L.ref_pg.restype = C.c_double
    return L


# ---- convenience wrappers -------------------------------------------------------------------

def pencil_flux(cfg, d, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = q.shape[0]
    F = np.full((n, cfg.nf), np.nan)
    eig = np.full(2 * n, np.nan)
    rc = lib().orc_pencil_flux(C.byref(cfg), d, n, dptr(q), dptr(F), dptr(eig))
    assert rc == 0
    return F, eig.reshape(n, 2)


def batch_substep(cfg, U, U0, h, dt, beta, want_D=False):
    nb = U.shape[0]
    h = np.ascontiguousarray(h, dtype=np.float64)
    D = np.zeros_like(U) if want_D else None
    rc = lib().orc_batch_substep(C.byref(cfg), nb, dptr(U), dptr(U0), dptr(h), dt, beta,
                                 dptr(D) if want_D else None)
    assert rc == 0
    return D


def batch_max_eig(cfg, U):
    e = np.zeros(U.shape[0])
    assert lib().orc_batch_max_eig(C.byref(cfg), U.shape[0], dptr(U), dptr(e)) == 0
    return e


def block_dt(cfg, m, h):
    return lib().orc_block_dt(C.byref(cfg), m, h)


def outflow_ghosts(cfg, U):
    assert lib().orc_batch_outflow_ghosts(C.byref(cfg), U.shape[0], dptr(U)) == 0


class RefShaped:
    """The reference-shaped restatement (orc_refshape_*): AoS nodes, one OpenMP parallel-for per
    pass inside each node."""

    def __init__(self, cfg, U, h):
        self.shape = U.shape
        h = np.ascontiguousarray(h, dtype=np.float64)
        self.r = lib().orc_refshape_create(C.byref(cfg), U.shape[0], dptr(np.ascontiguousarray(U)), dptr(h))
        assert self.r

    def rk_step(self, dt, rk=3):
        assert lib().orc_refshape_rk_step(self.r, dt, rk) == 0

    def max_eig_over_h(self):
        return lib().orc_refshape_max_eig(self.r)

    def download(self):
        U = np.empty(self.shape)
        assert lib().orc_refshape_download(self.r, dptr(U)) == 0
        return U

    def close(self):
        if self.r:
            lib().orc_refshape_destroy(self.r)
            self.r = None


def rk_steps(cfg, U, h, dt, rk, nsteps):
    U0 = np.empty_like(U)
    h = np.ascontiguousarray(h, dtype=np.float64)
    assert lib().orc_batch_rk_steps(C.byref(cfg), U.shape[0], dptr(U), dptr(U0), dptr(h), dt, rk,
                                    nsteps) == 0


class Tree:
    """Octree of the oracle (GridNode recursion restated)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        self.t = self.L.orc_tree_create(C.byref(cfg))
        assert self.t
        self.G = cfg.n + 2 * cfg.g

    def __del__(self):
        if getattr(self, "t", None):
            self.L.orc_tree_destroy(self.t)
            self.t = None

    def add_child(self, parent, c):
        r = self.L.orc_tree_add_child(self.t, parent, c)
        assert r >= 0
        return r

    @property
    def nnodes(self):
        return self.L.orc_tree_nnodes(self.t)

    def info(self, i):
        lev = C.c_int()
        loc = (C.c_int * 3)()
        off = (C.c_int * 3)()
        h = C.c_double()
        assert self.L.orc_tree_node_info(self.t, i, C.byref(lev), loc, off, C.byref(h)) == 0
        return lev.value, tuple(loc), tuple(off), h.value

    def U(self, i):
        p = self.L.orc_tree_node_U(self.t, i)
        G = self.G
        return np.ctypeslib.as_array(p, shape=(self.cfg.nf, G, G, G))

    def F(self, i):
        p = self.L.orc_tree_node_F(self.t, i)
        G = self.G
        return np.ctypeslib.as_array(p, shape=(self.cfg.nf, G, G, G))

    def enforce_boundaries(self):
        self.L.orc_tree_enforce_boundaries(self.t)

    def inject_from_children(self):
        self.L.orc_tree_inject_from_children(self.t)

    def store(self):
        self.L.orc_tree_store(self.t)

    def max_dt(self):
        return self.L.orc_tree_max_dt(self.t)

    def prolong(self, node):
        assert self.L.orc_tree_prolong(self.t, node) == 0

    def substep(self, dt, beta):
        self.L.orc_tree_substep(self.t, dt, beta)

    def state_sum(self):
        s = np.zeros(self.cfg.nf)
        self.L.orc_tree_state_sum(self.t, dptr(s))
        return s

    def flow_off(self):
        s = np.zeros(self.cfg.nf)
        self.L.orc_tree_flow_off(self.t, dptr(s))
        return s


# ---- Helmholtz EOS ---------------------------------------------------------------------------

def helm_table(rows):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    lib().orc_helm_table(dptr(rows))


def helm_set_units(cm=1.0, g=1.0, s=1.0, K=1.0):
    lib().orc_helm_set_units(cm, g, s, K)


def helm_eos(st, solve_T=False):
    """States (n, 10) in eos_t order, updated in place; returns the per-state error flags."""
    assert st.dtype == np.float64 and st.flags["C_CONTIGUOUS"] and st.shape[1] == 10
    flags = np.zeros(st.shape[0], dtype=np.int32)
    fn = lib().orc_helm_compute_T if solve_T else lib().orc_helm_eos
    assert fn(st.shape[0], dptr(st), flags.ctypes.data_as(IP)) == 0
    return flags


def ref_helm_lib(table_dir):
    """The reference's helmeos.cpp (oracle/_ref/libref_helm.so), which reads ./helmholtz.table when
    it is loaded: table_dir must hold that file.  None where the harness is not built."""
    path = os.path.join(REF_DIR, "libref_helm.so")
    if not os.path.exists(path):
        return None
    if not os.path.exists(os.path.join(table_dir, "helmholtz.table")):
        raise RuntimeError("helmholtz.table missing in %s" % table_dir)
    cwd = os.getcwd()
    os.chdir(table_dir)
    try:
        L = C.CDLL(path)
    finally:
        os.chdir(cwd)
    L.ref_helm_set_units.argtypes = [C.c_double] * 4
    L.ref_helm_set_units.restype = None
    L.ref_helm_eos.argtypes = [C.c_int, DP]
    L.ref_helm_eos.restype = None
    L.ref_helm_compute_T.argtypes = [C.c_int, DP]
    L.ref_helm_compute_T.restype = None
    return L


# ---- rotating-frame state --------------------------------------------------------------------

class OrcRotCfg(C.Structure):
    _fields_ = [("n", C.c_int), ("g", C.c_int), ("recon", C.c_int), ("z_reflect", C.c_int),
                ("eos", C.c_int), ("theta", C.c_double), ("rho_floor", C.c_double),
                ("Omega", C.c_double), ("Omega0", C.c_double), ("fR", C.c_double),
                ("ftheta", C.c_double), ("gamma", C.c_double), ("cfl", C.c_double),
                ("abar", C.c_double * 3), ("newton_tol", C.c_double)]


def rot_cfg(eos=0, recon=1, omega=0.0, omega0=None, f_r=0.0, f_theta=0.0, z_reflect=0, abar=(4.0, 14.12, 17.52),
            newton_tol=0.0):
    """The constants of binary/state_ztwd.cpp:28-31 and binary/binary_defs.h:29-30."""
    c = OrcRotCfg(n=8, g=3 if recon == 1 else 2, recon=recon, z_reflect=z_reflect, eos=eos, theta=1.3,
                  rho_floor=1.0e-10, Omega=omega, Omega0=omega if omega0 is None else omega0, fR=f_r,
                  ftheta=f_theta, gamma=5.0 / 3.0, cfl=0.3, newton_tol=newton_tol)
    c.abar[0], c.abar[1], c.abar[2] = abar
    return c


def _rot_lib():
    L = lib()
    if not getattr(L, "_rot_ready", False):
        cp = C.POINTER(OrcRotCfg)
        L.orc_rot_batch_substep.argtypes = [cp, C.c_int, DP, DP, DP, DP, DP, C.c_double, C.c_double, IP]
        L.orc_rot_batch_max_eig.argtypes = [cp, C.c_int, DP, DP, DP, DP]
        L.orc_rot_batch_outflow_ghosts.argtypes = [cp, C.c_int, DP, DP, DP]
        L.orc_rot_state.argtypes = [cp, C.c_int, DP, DP, DP]
        L._rot_ready = True
    return L


def rot_substep(cfg, U, U0, phi, x0, h, dt, beta):
    nb = U.shape[0]
    h = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (nb,)))
    flags = np.zeros(nb, dtype=np.int32)
    assert _rot_lib().orc_rot_batch_substep(C.byref(cfg), nb, dptr(U), dptr(U0), dptr(phi), dptr(x0), dptr(h),
                                            dt, beta, flags.ctypes.data_as(IP)) == 0
    return flags


def rot_max_eig(cfg, U, x0, h):
    nb = U.shape[0]
    h = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (nb,)))
    eig = np.empty(nb)
    _rot_lib().orc_rot_batch_max_eig(C.byref(cfg), nb, dptr(U), dptr(x0), dptr(h), dptr(eig))
    return eig


def rot_outflow_ghosts(cfg, U, x0, h):
    nb = U.shape[0]
    h = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (nb,)))
    _rot_lib().orc_rot_batch_outflow_ghosts(C.byref(cfg), nb, dptr(U), dptr(x0), dptr(h))


def rot_state(cfg, op, u, X, F=None):
    """op 0 to_prim, 1 to_con, 2-4 flux in x/y/z (returns F with the eigenvalue in F[13]), 5 floor,
    6 outflow on node face F[0].  u is updated in place."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    F = np.zeros(14) if F is None else np.ascontiguousarray(F, dtype=np.float64)
    err = _rot_lib().orc_rot_state(C.byref(cfg), op, dptr(u), dptr(X), dptr(F))
    return u, F, err


class RotTree:
    """The oracle's rotating-frame state on an octree given as (parent, child index) arrays.
    U(node) is (13, G, G, G) [field, k, j, i]; phi(node) and the Poisson face forces
    force(node) (3, n+2, n+2, n+2) are views into the oracle's arrays."""

    def __init__(self, cfg, parent, child, grid_dim=1.0):
        L = _rot_lib()
        if not getattr(L, "_rot_tree_ready", False):
            vp = C.c_void_p
            L.orc_rot_tree_create.argtypes = [C.POINTER(OrcRotCfg), C.c_double, C.c_int, IP, IP]
            L.orc_rot_tree_create.restype = vp
            L.orc_rot_tree_destroy.argtypes = [vp]
            L.orc_rot_tree_destroy.restype = None
            for f in ("U", "phi", "force", "F"):
                getattr(L, "orc_rot_tree_" + f).argtypes = [vp, C.c_int]
                getattr(L, "orc_rot_tree_" + f).restype = DP
            L.orc_rot_tree_F.argtypes = [vp, C.c_int, C.c_int]
            L.orc_rot_tree_node_x0.argtypes = [vp, C.c_int, DP, DP]
            L.orc_rot_tree_use_forces.argtypes = [vp, C.c_int]
            L.orc_rot_tree_use_forces.restype = None
            L.orc_rot_tree_set_origin.argtypes = [vp, DP]
            L.orc_rot_tree_set_origin.restype = None
            for f in ("fill_ghosts", "store", "pot_to_grid", "inject"):
                getattr(L, "orc_rot_tree_" + f).argtypes = [vp]
            L.orc_rot_tree_prolong.argtypes = [vp, C.c_int]
            L.orc_rot_tree_substep.argtypes = [vp, C.c_double, C.c_double]
            L.orc_rot_tree_max_dt.argtypes = [vp]
            L.orc_rot_tree_max_dt.restype = C.c_double
            L.orc_rot_tree_flow_off.argtypes = [vp, DP]
            L.orc_rot_tree_state_sum.argtypes = [vp, DP]
            L._rot_tree_ready = True
        self.L = L
        self.cfg = cfg
        self.n, self.g = cfg.n, cfg.g
        self.G, self.P = cfg.n + 2 * cfg.g, cfg.n + 2
        self.parent = np.ascontiguousarray(parent, np.int32)
        self.child = np.ascontiguousarray(child, np.int32)
        self.nn = len(self.parent)
        self.t = L.orc_rot_tree_create(C.byref(cfg), grid_dim, self.nn, self.parent.ctypes.data_as(IP),
                                       self.child.ctypes.data_as(IP))
        if not self.t:
            raise ValueError("bad rotating tree")

    def __del__(self):
        if getattr(self, "t", None):
            self.L.orc_rot_tree_destroy(self.t)
            self.t = None

    def U(self, node):
        G = self.G
        return np.ctypeslib.as_array(self.L.orc_rot_tree_U(self.t, node), shape=(13, G, G, G))

    def F(self, node, d):
        """the face fluxes of direction d of the last substep; face p lies below cell p"""
        G = self.G
        return np.ctypeslib.as_array(self.L.orc_rot_tree_F(self.t, node, d), shape=(13, G, G, G))

    def phi(self, node):
        P = self.P
        return np.ctypeslib.as_array(self.L.orc_rot_tree_phi(self.t, node), shape=(P, P, P))

    def force(self, node):
        P = self.P
        return np.ctypeslib.as_array(self.L.orc_rot_tree_force(self.t, node), shape=(3, P, P, P))

    def state(self):
        return np.stack([self.U(b).copy() for b in range(self.nn)])

    def set_state(self, U):
        for b in range(self.nn):
            self.U(b)[:] = U[b]

    def set_phi(self, phi):
        for b in range(self.nn):
            self.phi(b)[:] = phi[b]

    def set_forces(self, fx, fy, fz):
        for b in range(self.nn):
            self.force(b)[0], self.force(b)[1], self.force(b)[2] = fx[b], fy[b], fz[b]

    def x0(self, node):
        x0, h = np.zeros(3), C.c_double(0.0)
        self.L.orc_rot_tree_node_x0(self.t, node, dptr(x0), C.byref(h))
        return x0, h.value

    def use_forces(self, on):
        self.L.orc_rot_tree_use_forces(self.t, int(on))

    def set_origin(self, origin):
        o = np.ascontiguousarray(origin, dtype=np.float64).reshape(3)
        self.L.orc_rot_tree_set_origin(self.t, dptr(o))

    def fill_ghosts(self):
        self.L.orc_rot_tree_fill_ghosts(self.t)

    def store(self):
        self.L.orc_rot_tree_store(self.t)

    def substep(self, dt, beta):
        return self.L.orc_rot_tree_substep(self.t, dt, beta)

    def pot_to_grid(self):
        self.L.orc_rot_tree_pot_to_grid(self.t)

    def inject(self):
        return self.L.orc_rot_tree_inject(self.t)

    def prolong(self, node):
        return self.L.orc_rot_tree_prolong(self.t, node)

    def max_dt(self):
        return self.L.orc_rot_tree_max_dt(self.t)

    def flow_off(self):
        fo = np.zeros(13)
        self.L.orc_rot_tree_flow_off(self.t, dptr(fo))
        return fo

    def state_sum(self):
        s = np.zeros(13)
        self.L.orc_rot_tree_state_sum(self.t, dptr(s))
        return s


# ---- self-gravity: the multigrid Poisson solver (poisson/) --------------------------------------
_poisson_ready = False


def _poisson_lib():
    global _poisson_ready
    L = lib()
    if not _poisson_ready:
        vp = C.c_void_p
        L.orc_poisson_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, IP, IP]
        L.orc_poisson_create.restype = vp
        L.orc_poisson_destroy.argtypes = [vp]
        L.orc_poisson_destroy.restype = None
        L.orc_poisson_field.argtypes = [vp, C.c_int, C.c_int]
        L.orc_poisson_field.restype = DP
        L.orc_poisson_prepare.argtypes = [vp]
        L.orc_poisson_boundary_table.argtypes = [vp, C.c_int, C.c_int]
        L.orc_poisson_boundary_table.restype = DP
        L.orc_poisson_boundary_phi.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.orc_poisson_boundary_phi.restype = C.c_double
        L.orc_poisson_cell_centre.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.orc_poisson_cell_centre.restype = C.c_double
        for f in ("phi_from_children", "enforce_phi", "forces"):
            getattr(L, "orc_poisson_" + f).argtypes = [vp]
        L.orc_poisson_enforce_phi_level.argtypes = [vp, C.c_int]
        L.orc_poisson_error.argtypes = [vp]
        L.orc_poisson_error.restype = C.c_double
        L.orc_poisson_vcycle.argtypes = [vp]
        L.orc_poisson_vcycle.restype = C.c_double
        L.orc_poisson_solve.argtypes = [vp, C.c_double, C.c_int, C.c_int, IP, DP]
        L.orc_poisson_store_phi.argtypes = [vp, C.c_int]
        _poisson_ready = True
    return L


POISSON_FIELDS = {"phi": 0, "S": 1, "fx": 2, "fy": 3, "fz": 4, "dphi": 5, "dphi1": 6, "phi0": 7}


class Poisson:
    """The oracle's octree Poisson solver; arrays are (n+2)^3 per node, indexed [k, j, i]."""

    def __init__(self, n, parent, child, grid_dim=1.0, z_reflect=0, bw=2):
        self.L = _poisson_lib()
        self.n, self.P = n, n + 2
        self.parent = np.ascontiguousarray(parent, np.int32)
        self.child = np.ascontiguousarray(child, np.int32)
        self.nn = len(self.parent)
        self.p = self.L.orc_poisson_create(n, bw, grid_dim, z_reflect, self.nn,
                                           self.parent.ctypes.data_as(IP), self.child.ctypes.data_as(IP))
        if not self.p:
            raise ValueError("bad Poisson tree")

    def __del__(self):
        if getattr(self, "p", None):
            self.L.orc_poisson_destroy(self.p)
            self.p = None

    def field(self, node, name):
        P = self.P
        ptr = self.L.orc_poisson_field(self.p, node, POISSON_FIELDS[name])
        return np.ctypeslib.as_array(ptr, shape=(P, P, P))

    def fields(self, name):
        return np.stack([self.field(b, name).copy() for b in range(self.nn)])

    def set_fields(self, name, a):
        for b in range(self.nn):
            self.field(b, name)[:] = a[b]

    def table(self, node, face):
        ptr = self.L.orc_poisson_boundary_table(self.p, node, face)
        if not ptr:
            return None
        return np.ctypeslib.as_array(ptr, shape=(self.P, self.P))

    def centre(self, node, axis, i):
        return self.L.orc_poisson_cell_centre(self.p, node, axis, i)

    def prepare(self):
        self.L.orc_poisson_prepare(self.p)

    def boundary_phi(self, x, y, z):
        return self.L.orc_poisson_boundary_phi(self.p, x, y, z)

    def phi_from_children(self):
        self.L.orc_poisson_phi_from_children(self.p)

    def enforce_phi(self, level=None):
        if level is None:
            self.L.orc_poisson_enforce_phi(self.p)
        else:
            assert self.L.orc_poisson_enforce_phi_level(self.p, level) == 0

    def forces(self):
        self.L.orc_poisson_forces(self.p)

    def error(self):
        return self.L.orc_poisson_error(self.p)

    def vcycle(self):
        return self.L.orc_poisson_vcycle(self.p)

    def solve(self, toler, max_iter=10000, prepare=True):
        it = C.c_int(0)
        err = C.c_double(0.0)
        rc = self.L.orc_poisson_solve(self.p, toler, max_iter, int(prepare), C.byref(it), C.byref(err))
        if rc:
            raise RuntimeError("Poisson solve did not converge in %d vcycles" % max_iter)
        return it.value, err.value

    def store_phi(self, mode):
        self.L.orc_poisson_store_phi(self.p, mode)
