"""GPU parity of the PPM scheme (reconstruct.cpp:22-58, g = 3) on 8^3 blocks: the block kernel
through the C ABI against the FP64 CPU oracle on the same seeded inputs.  FP64 form: <= 1e-12
relative L1 per conserved field (bit-identical on rho, s, E); mixed form: density <= 1e-6."""
import numpy as np
import pytest

import cases
import orc
from gpu_util import amrhydro, rel_l1, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")  # device dt buffers; imported before the library touches the GPU

N, G = 8, 3
FP64_TOL = 1e-12


def ocfg(z_reflect=0):
    return orc.euler_cfg(n=N, g=G, recon=1, z_reflect=z_reflect)


def make_blocks(nb, seed, z_reflect=0):
    U = cases.c2_fast(nb, n=N, g=G, seed=seed)
    orc.outflow_ghosts(ocfg(z_reflect), U)
    return U


def gpu_batch(nb, h, precision=amrhydro.AMRH_FP64, z_reflect=0, boundary=amrhydro.AMRH_BOUNDARY_EXTERNAL):
    cfg = amrhydro.euler_config(n=N, g=G, precision=precision, z_reflect=z_reflect, boundary=boundary,
                                recon=amrhydro.AMRH_RECON_PPM)
    return amrhydro.Batch(cfg, nb, h)


def test_ppm_substeps_fp64_match_oracle():
    nb = 40
    say("ppm fp64 substeps nb=%d" % nb)
    U = make_blocks(nb, seed=5)
    h = np.full(nb, 1.0 / N)
    cfg = ocfg()
    dt = 0.2 * orc.block_dt(cfg, orc.batch_max_eig(cfg, U).max(), h[0])
    B = gpu_batch(nb, h)
    B.upload(U)
    B.store()
    ref, ref0 = U.copy(), U.copy()
    for beta in (1.0, 0.25, 2.0 / 3.0):
        orc.outflow_ghosts(cfg, ref)
        orc.batch_substep(cfg, ref, ref0, h, dt, beta)
        B.fill_outflow_ghosts()
        B.substep(dt, beta)
        out = B.download()
        e = rel_l1(out, ref, G)
        I = (slice(None), slice(0, 5), slice(G, -G), slice(G, -G), slice(G, -G))
        same = np.mean(out[I] == ref[I])
        say("  beta=%.4f rel-L1 %s  rho/s/E bit-identical %.6f" % (beta, np.array2string(e, precision=2), same))
        assert np.all(e <= FP64_TOL), e
        assert same == 1.0
        assert B.error_flags() == 0


@pytest.mark.parametrize("precision", [0, 1])
def test_ppm_block_max_eig_matches_oracle(precision):
    nb = 12
    U = make_blocks(nb, seed=6)
    ref = orc.batch_max_eig(ocfg(), U)
    B = gpu_batch(nb, 1.0 / N, precision=precision)
    B.upload(U)
    got = B.block_max_eig()
    np.testing.assert_allclose(got, ref, rtol=1e-14 if precision == 0 else 1e-5, atol=0)


def oracle_dt(cfg, U, h, last):
    dt = orc.block_dt(cfg, orc.batch_max_eig(cfg, U).max(), h)
    return 1e-2 * dt if last is None else min(1.25 * last, dt)


@pytest.mark.parametrize("z_reflect", [0, 1])
def test_ppm_rk3_trajectory_fp64(z_reflect):
    """10 RK3 steps with the outflow (z-mirror) ghost refresh done by the library after each
    substep (AMRH_BOUNDARY_OUTFLOW) and the reference's dt control."""
    nb, steps = 16, 10
    U = make_blocks(nb, seed=7, z_reflect=z_reflect)
    h = np.full(nb, 1.0 / N)
    cfg = ocfg(z_reflect)
    ref = U.copy()
    B = gpu_batch(nb, h, z_reflect=z_reflect, boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW)
    B.upload(U)
    dt = None
    for _ in range(steps):
        dt = oracle_dt(cfg, ref, h[0], dt)
        orc.rk_steps(cfg, ref, h, dt, 3, 1)
        B.rk_step(dt, 3)
    orc.outflow_ghosts(cfg, ref)
    out = B.download()
    e = rel_l1(out, ref, G)
    say("ppm RK3 x%d z_reflect=%d fp64 rel-L1 %s" % (steps, z_reflect, np.array2string(e, precision=2)))
    assert np.all(e <= FP64_TOL), e


def test_ppm_mixed_error_after_rk3_steps():
    nb, steps = 16, 30
    U = make_blocks(nb, seed=8)
    h = np.full(nb, 1.0 / N)
    cfg = ocfg()
    ref = U.copy()
    B = gpu_batch(nb, h, precision=amrhydro.AMRH_MIXED, boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW)
    B.upload(U)
    dt = None
    for _ in range(steps):
        dt = oracle_dt(cfg, ref, h[0], dt)
        orc.rk_steps(cfg, ref, h, dt, 3, 1)
        B.rk_step(dt, 3)
    out = B.download()
    e = rel_l1(out, ref, G)
    say("ppm mixed %d RK3 steps rel-L1 %s" % (steps, np.array2string(e, precision=2)))
    assert e[0] <= 1e-6, e


@pytest.mark.parametrize("precision", [0, 1])
def test_ppm_graph_and_persistent_match_direct(precision):
    nb = 24
    U = make_blocks(nb, seed=9)
    h = np.full(nb, 1.0 / N)
    outs = []
    for mode in ("direct", "graph", "persistent+graph"):
        B = gpu_batch(nb, h, precision=precision, boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW)
        B.upload(U)
        if mode == "persistent+graph":
            B.set_launch(5, 7)
        dt = amrhydro_dt(1e-4)
        B.run_steps(4, dt.data_ptr(), rk=3, use_graph=(mode != "direct"))
        B.sync()
        outs.append(B.download())
        B.close()
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def amrhydro_dt(v):
    return torch.full((1,), v, dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("precision", [0, 1])
def test_ppm_uniform_state_is_stationary(precision):
    nb = 2
    Gp = N + 2 * G
    cell = cases.prim_to_con(np.array(0.7), np.array(0.3), np.array(-0.2), np.array(0.1), np.array(0.4))
    orc.lib().orc_floor(orc.C.byref(ocfg()), orc.dptr(cell))
    U = np.empty((nb, 6, Gp, Gp, Gp))
    U[:] = cell[None, :, None, None, None]
    B = gpu_batch(nb, 0.125, precision=precision)
    B.upload(U)
    B.store()
    B.substep(1e-3, 1.0)
    out = B.download()
    I = (slice(None), slice(0, 5), slice(G, -G), slice(G, -G), slice(G, -G))
    np.testing.assert_array_equal(out[I], U[I])
