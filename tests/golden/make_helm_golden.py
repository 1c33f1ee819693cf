"""Writes golden Helmholtz EOS vectors from the reference's own helmholtz/helmeos.cpp.

Run in a container that has the reference sources (make -C oracle ref first):
    python tests/golden/make_helm_golden.py
For each synthetic table -- the non-degenerate electron table (workloads.helmholtz_table) and the
cold-Fermi-gas + radiation table of the double white dwarf (workloads.helmholtz_table_dwd) -- the
reference-built harness (oracle/_ref/libref_helm.so, which reads ./helmholtz.table when it is
loaded; one copy of the library per table) evaluates helmholtz_eos and helmholtz_compute_T on
seeded states in the table's code units.  Inputs and outputs (eos_t records of 10 doubles) are
saved to tests/golden/ref_helm.npz.
"""
import ctypes as C
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import orc  # noqa: E402
from gpu_util import amrhydro  # noqa: E402,F401  (puts the package on the path)
import workloads  # noqa: E402

N = 300


def states(rng, kind):
    st = np.zeros((N, 10))
    if kind == "ideal":
        st[:, 4] = rng.uniform(4.0, 17.5, N)
        st[:, 6] = 10.0 ** rng.uniform(-4.0, 8.0, N)
        st[:, 7] = 10.0 ** rng.uniform(4.0, 9.5, N)
    else:
        st[:, 4] = workloads.DWD_ABAR
        st[:, 6] = 10.0 ** rng.uniform(-1.0, 7.5, N)
        st[:, 7] = 10.0 ** rng.uniform(5.5, 9.5, N)
    st[:, 5] = st[:, 4] / 2.0
    return st


def main():
    src = os.path.join(orc.REF_DIR, "libref_helm.so")
    if not os.path.exists(src):
        raise SystemExit("oracle/_ref/libref_helm.so missing: run `make -C oracle ref`")
    out = {}
    cwd = os.getcwd()
    for kind, rows, units in (("ideal", workloads.helmholtz_table(), workloads.HELM_UNITS),
                              ("dwd", workloads.helmholtz_table_dwd(), workloads.dwd_model()["units"])):
        d = tempfile.mkdtemp()
        try:
            workloads.write_helmholtz_table(os.path.join(d, "helmholtz.table"), rows)
            lib = os.path.join(d, "libref_helm_%s.so" % kind)
            shutil.copy(src, lib)
            os.chdir(d)
            L = C.CDLL(lib)
        finally:
            os.chdir(cwd)
        L.ref_helm_set_units.argtypes = [C.c_double] * 4
        L.ref_helm_eos.argtypes = [C.c_int, orc.DP]
        L.ref_helm_compute_T.argtypes = [C.c_int, orc.DP]
        L.ref_helm_set_units(*units)
        rng = np.random.default_rng(7 if kind == "ideal" else 8)
        st = states(rng, kind)
        # code units: rho given in cgs above, converted as eos_to_code does
        cm, g, s, K = units
        st[:, 6] /= g / cm ** 3
        st[:, 7] /= K
        a = np.ascontiguousarray(st.copy())
        L.ref_helm_eos(N, orc.dptr(a))
        # compute_T from each state's own e, starting 30% off
        tin = a.copy()
        tin[:, 7] *= rng.uniform(0.7, 1.3, N)
        b = np.ascontiguousarray(tin.copy())
        L.ref_helm_compute_T(N, orc.dptr(b))
        out[kind + "_units"] = np.array(units)
        out[kind + "_eos_in"] = st
        out[kind + "_eos_out"] = a
        out[kind + "_T_in"] = tin
        out[kind + "_T_out"] = b
        shutil.rmtree(d)
    path = os.path.join(HERE, "ref_helm.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
