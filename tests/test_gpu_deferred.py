"""Self-ghosting outflow blocks and the deferred-dt RK step of flat ideal-gas batches.

An AMRH_BOUNDARY_OUTFLOW block of the minmod ideal gas forms every ghost from the boundary cell it
copies (grid/grid_phys_bound.cpp:27-54, euler/state.cpp:45-84) where the kernel reads it, so the
ghost slabs in memory play no part.  An RK2/RK3 step of such a batch (amrh_run_steps) runs with its
dt deferred: the first substep reduces its own KT face speeds to the CFL maximum and stores the
divergence, and the second (minmod) or an elementwise pass (PPM) forms the first stage from U^n
and D once dt is known.  Both must give the bits of the plain sequence (CFL scan, store,
substeps), and the FP64 form the oracle's."""
import numpy as np
import pytest

import cases
import orc
from gpu_util import amrhydro, rel_l1, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BETAS = {2: (1.0, 0.5), 3: (1.0, 0.25, 2.0 / 3.0)}


def dt_buffer(value):
    return torch.full((1,), value, dtype=torch.float64, device="cuda")


def blocks(n, nb, seed, z_reflect=0):
    U = cases.c2_fast(nb, n=n, g=2, seed=seed)
    orc.outflow_ghosts(orc.euler_cfg(n=n, z_reflect=z_reflect), U)
    return U


def outflow_batch(U, h, precision, n, z_reflect=0):
    cfg = amrhydro.euler_config(n=n, precision=precision, z_reflect=z_reflect,
                                boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW)
    B = amrhydro.Batch(cfg, U.shape[0], h)
    B.upload(U)
    return B


def spoil_ghosts(U, n, g=2):
    """U with every ghost-slab cell replaced by a value no ghost fill would give"""
    V = U.copy()
    G = n + 2 * g
    inner = np.zeros(G, bool)
    inner[g:G - g] = True
    ghost = ~(inner[:, None, None] & inner[None, :, None] & inner[None, None, :])
    V[:, :, ghost] = 3.0
    return V


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("z_reflect", [0, 1])
@pytest.mark.parametrize("n", [8, 16])
def test_outflow_blocks_form_their_own_ghosts(precision, z_reflect, n):
    """Stored ghost slabs are never read: with them spoilt, the substeps and the CFL scan of an
    outflow batch equal those of a batch whose ghosts are filled before every substep."""
    nb = 6
    U = blocks(n, nb, seed=50 + n, z_reflect=z_reflect)
    h = np.full(nb, 1.0 / n)
    cfg_o = orc.euler_cfg(n=n, z_reflect=z_reflect)
    dt = 0.1 * orc.block_dt(cfg_o, orc.batch_max_eig(cfg_o, U).max(), h[0])
    A = outflow_batch(spoil_ghosts(U, n), h, precision, n, z_reflect)
    ext = amrhydro.Batch(amrhydro.euler_config(n=n, precision=precision, z_reflect=z_reflect), nb, h)
    ext.upload(U)
    np.testing.assert_array_equal(A.block_max_eig(), ext.block_max_eig())
    for B in (A, ext):
        B.rk_step(dt, 3)
    ext.fill_outflow_ghosts()
    a, b = A.download(), ext.download()  # download forms the outflow batch's ghosts
    np.testing.assert_array_equal(a, b)
    assert A.error_flags() == 0
    if precision == 0:
        ref = U.copy()
        orc.rk_steps(cfg_o, ref, h, dt, 3, 1)
        e = rel_l1(a, ref, 2)
        say("self-ghosting n=%d zr=%d fp64 vs oracle %s" % (n, z_reflect, np.array2string(e, precision=2)))
        assert np.all(e <= 1e-12), e


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("z_reflect", [0, 1])
@pytest.mark.parametrize("n,rk", [(8, 3), (16, 3), (16, 2)])
def test_deferred_step_equals_scan_store_substeps(precision, z_reflect, n, rk):
    nb = 10 if n == 16 else 40
    U = blocks(n, nb, seed=60 + n + rk, z_reflect=z_reflect)
    h = np.full(nb, 1.0 / n) * np.where(np.arange(nb) % 3 == 0, 0.5, 1.0)  # two cell widths
    steps = 3
    # the deferred step, direct launches and as graphs
    outs = []
    for use_graph in (False, True):
        A = outflow_batch(U, h, precision, n, z_reflect)
        dA = dt_buffer(1e-4)  # the growth cap binds on the first step
        A.run_steps(steps, dA.data_ptr(), rk=rk, use_graph=use_graph)
        A.sync()
        # the deferred steps leave the CFL word cleared: a scan afterwards sees only the new state
        outs.append((A.download(), float(dA.cpu()[0]), A.error_flags(), A.max_speed()))
        A.close()
    # the plain sequence: scan -> dt, store, substeps
    B = outflow_batch(U, h, precision, n, z_reflect)
    dB = dt_buffer(1e-4)
    for _ in range(steps):
        B.dt_device(1.0, 1.25, dB.data_ptr())
        B.store()
        for beta in BETAS[rk]:
            B.substep_dev(dB.data_ptr(), beta)
    B.sync()
    ref = B.download()
    say("deferred n=%d rk=%d prec=%d zr=%d: dt %.6e / %.6e" % (n, rk, precision, z_reflect, outs[0][1], float(dB.cpu()[0])))
    m = B.max_speed()
    for out, d, err, ma in outs:
        assert err == 0
        assert d == float(dB.cpu()[0])
        np.testing.assert_array_equal(out, ref)
        assert ma == m


def test_deferred_step_fp64_matches_oracle():
    """FP64 deferred RK3 steps against the oracle, whose CFL dt (grid/grid.cpp:422-483 with the
    growth cap of euler/hydro_driver.cpp:107-113) is computed from its own state.  The first step
    is 1e-2 of the CFL dt, as the reference's first step (MAXINITDT): started at the full CFL dt a
    12-block sample drives E negative in the reference's algorithm and the oracle's alike."""
    n, nb = 16, 12
    U = blocks(n, nb, seed=71)
    h = np.full(nb, 1.0 / n)
    cfg = orc.euler_cfg(n=n)
    A = outflow_batch(U, h, 0, n)
    last = 0.8e-2 * orc.block_dt(cfg, orc.batch_max_eig(cfg, U).max(), h[0])
    d = dt_buffer(last)  # the first step's dt is 1.25 x this
    ref = U.copy()
    for s in range(5):
        A.run_steps(1, d.data_ptr(), rk=3, use_graph=False)
        A.sync()
        dt_gpu = float(d.cpu()[0])
        # HydroDriver::step ends with enforce_boundaries (euler/hydro_driver.cpp:190), so the
        # next max_dt sees the ghosts of the new state
        orc.outflow_ghosts(cfg, ref)
        cfl = orc.block_dt(cfg, orc.batch_max_eig(cfg, ref).max(), h[0])
        dt_ref = min(1.25 * last, cfl)
        assert abs(dt_gpu - dt_ref) <= 1e-15 * dt_ref, (s, dt_gpu, dt_ref)
        orc.rk_steps(cfg, ref, h, dt_gpu, 3, 1)
        last = dt_gpu
    out = A.download()
    assert A.error_flags() == 0
    e = rel_l1(out, ref, 2)
    same = np.mean(out[:, :5, 2:-2, 2:-2, 2:-2] == ref[:, :5, 2:-2, 2:-2, 2:-2])
    say("deferred fp64 5 RK3 steps vs oracle %s, rho..E bit-identical %.6f" % (np.array2string(e, precision=2), same))
    assert np.all(e <= 1e-12), e
    assert same == 1.0


def test_step_timed_stages():
    """amrh_step_timed: the deferred step's stages for minmod and PPM outflow batches (the PPM one
    with its first-stage pass), the plain ones for RK1"""
    n, nb = 16, 8
    U = blocks(n, nb, seed=81)
    A = outflow_batch(U, 1.0 / n, 1, n)
    d = dt_buffer(1e-4)
    st = A.step_timed(d.data_ptr(), rk=3)
    say("minmod rk3 stages %s" % st)
    assert [k for k, _ in st] == [amrhydro.AMRH_LAUNCH_FIRST, amrhydro.AMRH_LAUNCH_DT,
                                  amrhydro.AMRH_LAUNCH_SECOND, amrhydro.AMRH_LAUNCH_SUBSTEP]
    assert all(ms > 0.0 for _, ms in st)
    st = A.step_timed(d.data_ptr(), rk=1)
    assert [k for k, _ in st] == [amrhydro.AMRH_LAUNCH_SCAN, amrhydro.AMRH_LAUNCH_SUBSTEP]
    import workloads
    V = workloads.flat_batch(8, n=8, g=3, seed=2)
    cfg = amrhydro.euler_config(n=8, g=3, recon=amrhydro.AMRH_RECON_PPM, boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW)
    P = amrhydro.Batch(cfg, 8, 1.0 / 8)
    P.upload(V)
    P.fill_outflow_ghosts()
    st = P.step_timed(d.data_ptr(), rk=3)
    assert [k for k, _ in st] == [amrhydro.AMRH_LAUNCH_FIRST, amrhydro.AMRH_LAUNCH_DT,
                                  amrhydro.AMRH_LAUNCH_FIRST_STAGE] + [amrhydro.AMRH_LAUNCH_SUBSTEP] * 2


@pytest.mark.parametrize("precision", [0, 1])
def test_autotuned_deferred_kernels_give_the_same_bits(precision):
    n, nb = 16, 600
    U = blocks(n, nb, seed=91)
    h = np.full(nb, 1.0 / n)
    outs = []
    for tune in (False, True):
        A = outflow_batch(U, h, precision, n)
        if tune:
            A.autotune()
            for k in (0, 1):
                g, v, ms0, ms1 = A.deferred_launch_info(k)
                say("deferred kernel %d: grid %d variant %d, %.3f -> %.3f ms" % (k, g, v, ms0, ms1))
                assert 0 <= v < 4 and 0.0 < ms1 <= ms0
        d = dt_buffer(1e-4)
        A.run_steps(2, d.data_ptr(), rk=3, use_graph=True)
        A.sync()
        outs.append((A.download(), float(d.cpu()[0])))
        A.close()
    np.testing.assert_array_equal(outs[1][0], outs[0][0])
    assert outs[1][1] == outs[0][1]


@pytest.mark.parametrize("n", [8, 16])
def test_schedules_give_the_same_bits(n):
    """amrh_set_schedule: deferred and plain RK3 steps of the same batch are bit-identical; the
    autotuner's pick is one of the two; a batch without outflow boundaries reports the plain
    schedule"""
    nb = 300 if n == 8 else 64
    U = blocks(n, nb, seed=93 + n)
    h = np.full(nb, 1.0 / n)
    outs = []
    for sched in (amrhydro.AMRH_SCHEDULE_PLAIN, amrhydro.AMRH_SCHEDULE_DEFERRED, None):
        A = outflow_batch(U, h, 1, n)
        if sched is None:
            A.autotune()
            say("n=%d: autotuned schedule %d" % (n, A.schedule()))
            assert A.schedule() in (amrhydro.AMRH_SCHEDULE_PLAIN, amrhydro.AMRH_SCHEDULE_DEFERRED)
        else:
            A.set_schedule(sched)
            assert A.schedule() == sched
        d = dt_buffer(1e-4)
        A.run_steps(3, d.data_ptr(), rk=3, use_graph=True)
        A.sync()
        outs.append((A.download(), float(d.cpu()[0])))
        A.close()
    for o, dt in outs[1:]:
        np.testing.assert_array_equal(o, outs[0][0])
        assert dt == outs[0][1]
    P = amrhydro.Batch(amrhydro.euler_config(n=n), 4, 1.0 / n)
    assert P.schedule() == amrhydro.AMRH_SCHEDULE_PLAIN


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("z_reflect", [0, 1])
def test_ppm_deferred_step_equals_plain(precision, z_reflect):
    """The PPM block kernel's deferred step (its first substep's face speeds as the CFL maximum, D
    stored, the first stage formed by an elementwise pass) gives the bits of the plain schedule,
    direct, as graphs and with the autotuner's launch shapes, dt included"""
    import workloads
    nb = 96
    V = workloads.flat_batch(nb, n=8, g=3, seed=7 + z_reflect)
    h = np.full(nb, 1.0 / 8) * np.where(np.arange(nb) % 4 == 0, 0.5, 1.0)
    cfg = amrhydro.euler_config(n=8, g=3, recon=amrhydro.AMRH_RECON_PPM, precision=precision,
                                z_reflect=z_reflect, boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW)
    outs = []
    for sched, use_graph in ((amrhydro.AMRH_SCHEDULE_PLAIN, False), (amrhydro.AMRH_SCHEDULE_DEFERRED, False),
                             (amrhydro.AMRH_SCHEDULE_DEFERRED, True), (None, True)):
        P = amrhydro.Batch(cfg, nb, h)
        P.upload(V)
        if sched is None:
            P.autotune()
            g, _, ms0, ms1 = P.deferred_launch_info(0)
            say("ppm deferred substep: grid %d, %.3f -> %.3f ms; schedule %d" % (g, ms0, ms1, P.schedule()))
            assert 0.0 < ms1 <= ms0
        else:
            P.set_schedule(sched)
            assert P.schedule() == sched
        d = dt_buffer(1e-4)
        P.run_steps(3, d.data_ptr(), rk=3, use_graph=use_graph)
        P.sync()
        outs.append((P.download(), float(d.cpu()[0])))
        assert P.error_flags() == 0
        P.close()
    say("ppm prec=%d zr=%d: dt after 3 steps %.6e" % (precision, z_reflect, outs[0][1]))
    for o, dt in outs[1:]:
        np.testing.assert_array_equal(o, outs[0][0])
        assert dt == outs[0][1]
