"""Config 3 (uneven multilevel leaves, per-level cell width) and the whole-step machinery: the
autotuned persistent launch and hipGraph-captured RK3 steps must give the same bits as direct,
one-workgroup-per-block launches; FP64 matches the oracle at <= 1e-12; mixed stays <= 1e-6 on
density at full size."""
import numpy as np
import pytest

import orc
from gpu_util import amrhydro, rel_l1, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import workloads  # noqa: E402  (on sys.path through gpu_util)


def dt_buffer(value):
    return torch.full((1,), value, dtype=torch.float64, device="cuda")


def batch(U, h, precision, n=8):
    cfg = amrhydro.euler_config(n=n, precision=precision, boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW)
    B = amrhydro.Batch(cfg, U.shape[0], h)
    B.upload(U)
    B.fill_outflow_ghosts()
    return B


def small_multilevel():
    U, h, lev = workloads.multilevel(counts=(8, 60, 120, 60))
    orc.outflow_ghosts(orc.euler_cfg(n=8), U)
    return U, h


def test_multilevel_fp64_matches_oracle():
    U, h = small_multilevel()
    say("multilevel subset: %d blocks, h in %s" % (U.shape[0], np.unique(h)))
    cfg = orc.euler_cfg(n=8)
    ref = U.copy()
    B = batch(U, h, amrhydro.AMRH_FP64)
    dt = None
    for _ in range(4):
        m = orc.batch_max_eig(cfg, ref)
        cfl = min(orc.block_dt(cfg, m[b], h[b]) for b in range(len(h)))
        dt = 1e-2 * cfl if dt is None else min(1.25 * dt, cfl)
        orc.rk_steps(cfg, ref, h, dt, 3, 1)
        B.rk_step(dt, 3)
    out = B.download()
    e = rel_l1(out, ref, 2)
    say("multilevel fp64 rel-L1 %s" % np.array2string(e, precision=2))
    assert np.all(e <= 1e-12), e


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_and_persistent_launch_match_direct(precision):
    U, h = small_multilevel()
    outs = []
    for mode in ("direct", "graph", "persistent+graph"):
        B = batch(U, h, precision)
        if mode == "persistent+graph":
            gs, gm = B.autotune()
            say("autotuned grids: substep %d, cfl scan %d (0 = one workgroup per block)" % (gs, gm))
            B.set_launch(64, 64)  # a persistent grid smaller than the batch, whatever was chosen
        dt = dt_buffer(1e-5)
        B.run_steps(5, dt.data_ptr(), rk=3, use_graph=(mode != "direct"))
        B.sync()
        outs.append((B.download(), float(dt.cpu()[0])))
        B.close()
    for o, d in outs[1:]:
        np.testing.assert_array_equal(o, outs[0][0])
        assert d == outs[0][1]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [8, 16])
def test_register_budget_variants_give_the_same_bits(precision, n):
    """Every register-budget variant of the flat sweep (2, 3, 4 waves per SIMD) and the autotuner's
    pick give the bits of the default launch; amrh_launch_info reports the pick"""
    if n == 8:
        U, h = small_multilevel()
    else:
        U = workloads.flat_batch(40, n=16, seed=4)
        h = np.full(40, 1.0 / 16)
    outs = []
    for v in range(4):
        B = amrhydro.Batch(amrhydro.euler_config(n=n, precision=precision,
                                                 boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW), U.shape[0], h)
        B.upload(U)
        B.fill_outflow_ghosts()
        B.set_variant(v)
        dt = dt_buffer(1e-5)
        B.run_steps(3, dt.data_ptr(), rk=3, use_graph=True)
        B.sync()
        outs.append(B.download())
        if v == 0:
            B2 = amrhydro.Batch(amrhydro.euler_config(n=n, precision=precision,
                                                      boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW), U.shape[0], h)
            B2.upload(U)
            B2.fill_outflow_ghosts()
            B2.autotune()
            gs, gm, var, ms0, ms1 = B2.launch_info()
            say("n=%d prec=%d: autotuned grid %d, variant %d, substep %.3f -> %.3f ms" % (n, precision, gs, var, ms0, ms1))
            assert 0 <= var < 4 and ms1 <= ms0 and ms1 > 0.0
            dt2 = dt_buffer(1e-5)
            B2.run_steps(3, dt2.data_ptr(), rk=3, use_graph=True)
            B2.sync()
            np.testing.assert_array_equal(B2.download(), outs[0])
            B2.close()
        B.close()
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


@pytest.mark.parametrize("precision", [0, 1])
def test_tree_graph_matches_direct(precision):
    import tree_cases as tc
    cfg_o = orc.euler_cfg(n=8, g=2, grid_dim=0.5, z_reflect=0)
    T, parent, child = tc.build(cfg_o, tc.partial_spec())
    tc.blob_ic(T)
    T.enforce_boundaries()
    T.inject_from_children()
    U = tc.padded_state(T)
    outs = []
    for use_graph in (False, True):
        cfg = amrhydro.euler_config(n=8, grid_dim=0.5, precision=precision)
        B = amrhydro.Batch(cfg, len(parent), 1.0)
        B.tree_set(parent, child)
        B.upload(U)
        dt = dt_buffer(1e-4)
        B.run_steps(4, dt.data_ptr(), rk=3, use_graph=use_graph)
        B.sync()
        outs.append((B.download(), B.tree_flow_off()))
        B.close()
    np.testing.assert_array_equal(outs[1][0], outs[0][0])
    np.testing.assert_array_equal(outs[1][1], outs[0][1])


def test_full_multilevel_mixed_vs_fp64():
    """The full uneven batch (59,518 leaves of 8^3 over 7 levels): 20 graph-captured RK3 steps per
    form with the same dt sequence; density rel-L1 of the mixed form against FP64 <= 1e-6."""
    U, h, lev = workloads.multilevel()
    say("full multilevel: %d leaves, per level %s" % (U.shape[0], np.bincount(lev)))
    B64 = batch(U, h, amrhydro.AMRH_FP64)
    B32 = batch(U, h, amrhydro.AMRH_MIXED)
    del U
    for B in (B64, B32):
        B.autotune()
    dt = None
    for _ in range(20):
        dt = B64.cfl_dt(dt)  # the FP64 form's dt drives both (euler/hydro_driver.cpp:107-113)
        for B in (B64, B32):
            B.rk_step(dt, 3)
    res = []
    for B in (B64, B32):
        assert B.error_flags() == 0
        res.append(B.download())
        B.close()
    e = rel_l1(res[1], res[0], 2)
    say("full multilevel mixed vs fp64 after 20 RK3 steps: rel-L1 %s" % np.array2string(e, precision=2))
    assert e[0] <= 1e-6, e
