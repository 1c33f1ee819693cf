"""The C-ABI library loads (no GPU call is made) and exports every entry point include/*.h
declares; the Python binding covers the same surface."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import orc

ROOT = orc.ROOT
PKG = os.path.join(ROOT, "3d-eulerian-amr_amd")
HEADER = os.path.join(ROOT, "include", "amrhydro.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(amrh_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "amrh_substep" in names and "amrh_create" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    so = os.path.join(PKG, "libamrhydro.so")
    if not os.path.exists(so):
        pytest.fail("libamrhydro.so not built (run __graft_entry__.build())")
    L = C.CDLL(so)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import sys
    sys.path.insert(0, PKG)
    import amrhydro
    assert set(declared_functions()) == set(amrhydro.SIGNATURES)


@pytest.mark.parametrize("n,g,recon,msg", [
    (8, 2, 1, "PPM reconstruction needs g = 3"),
    (16, 3, 1, "PPM reconstruction runs on n = 8"),
    (8, 3, 0, "minmod reconstruction needs g = 2"),
    (8, 2, 7, "recon 7 unknown"),
    (12, 2, 0, "n must be 8 or 16"),
])
def test_create_rejects_unsupported_schemes(n, g, recon, msg):
    """Configuration checks run before any device call, so they are exercised without a GPU."""
    import sys
    sys.path.insert(0, PKG)
    import amrhydro
    cfg = amrhydro.euler_config(n=n, g=g, recon=recon)
    ctx = C.c_void_p()
    rc = amrhydro.lib().amrh_create(C.byref(cfg), 4, C.byref(ctx))
    assert rc == 1 and not ctx
    assert msg in amrhydro.lib().amrh_last_error().decode()


def _amrhydro():
    import sys
    sys.path.insert(0, PKG)
    import amrhydro
    return amrhydro


@pytest.mark.parametrize("nblocks", [0, -3])
def test_create_rejects_empty_batches(nblocks):
    amrhydro = _amrhydro()
    cfg = amrhydro.euler_config(n=8)
    ctx = C.c_void_p()
    assert amrhydro.lib().amrh_create(C.byref(cfg), nblocks, C.byref(ctx)) == 1 and not ctx
    assert "nblocks must be positive" in amrhydro.lib().amrh_last_error().decode()


@pytest.mark.parametrize("parent,child,msg", [
    ([0, 0], [0, 0], "node 0 must be the root"),
    ([-1, 1], [0, 0], "bad parent"),           # a node's parent after it
    ([-1, 0, 0], [0, 3, 3], "bad parent"),     # two children in one slot
    ([-1, 0], [0, 8], "bad parent"),           # child index out of range
])
def test_poisson_rejects_bad_trees(parent, child, msg):
    """The tree is checked before anything is allocated on the device."""
    amrhydro = _amrhydro()
    p = np.array(parent, np.int32)
    c = np.array(child, np.int32)
    h = C.c_void_p()
    IP = C.POINTER(C.c_int)
    rc = amrhydro.lib().amrh_poisson_create(8, 2, 1.0, 0, len(p), p.ctypes.data_as(IP), c.ctypes.data_as(IP),
                                            None, C.byref(h))
    assert rc == 1 and not h
    assert msg in amrhydro.lib().amrh_last_error().decode()


@pytest.mark.parametrize("n,bw,msg", [(10, 2, "n = 8, 12 or 16"), (8, 0, "ghost width")])
def test_poisson_rejects_bad_shapes(n, bw, msg):
    amrhydro = _amrhydro()
    p = np.array([-1], np.int32)
    h = C.c_void_p()
    IP = C.POINTER(C.c_int)
    rc = amrhydro.lib().amrh_poisson_create(n, bw, 1.0, 0, 1, p.ctypes.data_as(IP), p.ctypes.data_as(IP), None,
                                            C.byref(h))
    assert rc == 1 and not h
    assert msg in amrhydro.lib().amrh_last_error().decode()
