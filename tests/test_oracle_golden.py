"""The oracle against the committed golden vectors written by the reference's own Reconstruct /
State classes (tests/golden/make_golden.py).  Runs anywhere; bit-exact."""
import os

import numpy as np
import pytest

import orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_pencils.npz")


@pytest.fixture(scope="module")
def gold():
    if not os.path.exists(GOLD):
        pytest.fail("golden fixture missing: run tests/golden/make_golden.py")
    return np.load(GOLD)


@pytest.mark.parametrize("kind,recon,faces", [("euler", 0, slice(2, 11)), ("ppm", 1, slice(3, 10))])
def test_pencils_match_reference_vectors(gold, kind, recon, faces):
    cfg = orc.euler_cfg(n=8, g=2, recon=recon)
    Q, F, E = gold[kind + "_q"], gold[kind + "_F"], gold[kind + "_eig"]
    for p in range(Q.shape[0]):
        for d in range(3):
            Fo, eo = orc.pencil_flux(cfg, d, Q[p])
            np.testing.assert_array_equal(Fo[faces], F[p, d][faces])
            np.testing.assert_array_equal(eo[faces], E[p, d][faces])


def test_floor_and_outflow_match_reference_vectors(gold):
    cfg = orc.euler_cfg()
    for u, want in zip(gold["cells"], gold["floor"]):
        a = u.copy()
        orc.lib().orc_floor(orc.C.byref(cfg), orc.dptr(a))
        np.testing.assert_array_equal(a, want)
    for c in range(gold["outflow"].shape[0]):
        for face in range(6):
            a = gold["cells"][c].copy()
            orc.lib().orc_outflow(orc.C.byref(cfg), face, orc.dptr(a))
            np.testing.assert_array_equal(a, gold["outflow"][c, face])


HELM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_helm.npz")


@pytest.mark.parametrize("kind", ["ideal", "dwd"])
def test_helmholtz_matches_reference_vectors(kind):
    """helmholtz_eos and helmholtz_compute_T of the reference's helmeos.cpp
    (tests/golden/make_helm_golden.py) on the two synthetic tables, in code units: the oracle gives
    the same bits in all ten eos_t fields"""
    from gpu_util import amrhydro  # noqa: F401  (puts the package on the path)
    import workloads
    if not os.path.exists(HELM):
        pytest.fail("golden fixture missing: run tests/golden/make_helm_golden.py")
    gold = np.load(HELM)
    rows = workloads.helmholtz_table() if kind == "ideal" else workloads.helmholtz_table_dwd()
    orc.helm_table(rows)
    orc.helm_set_units(*gold[kind + "_units"])
    try:
        a = np.ascontiguousarray(gold[kind + "_eos_in"].copy())
        assert not orc.helm_eos(a).any()
        np.testing.assert_array_equal(a, gold[kind + "_eos_out"])
        b = np.ascontiguousarray(gold[kind + "_T_in"].copy())
        assert not orc.helm_eos(b, solve_T=True).any()
        np.testing.assert_array_equal(b, gold[kind + "_T_out"])
    finally:
        orc.helm_set_units()
