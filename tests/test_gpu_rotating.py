"""GPU parity of the rotating-frame binary-star state (binary/state_ztwd.cpp with the substep of
BinaryDriver::sub_step and a given potential) through the C ABI, against the FP64 oracle on the
same seeded binary batches.  FP64: <= 1e-12 relative L1 per field (the geometric source's
pow(l_z/R, 2) may differ from glibc by an ulp); ideal-gas and Helmholtz EOS; PPM and minmod."""
import numpy as np
import pytest

import helm_cases as hc
import orc
import rot_cases as rc
from gpu_util import amrhydro, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import workloads  # noqa: E402

TOL = 1e-12


def rel_l1_rot(a, b, g):
    A = a[:, :, g:-g, g:-g, g:-g]
    B = b[:, :, g:-g, g:-g, g:-g]
    out = []
    for f in range(A.shape[1]):
        den = np.abs(B[:, f]).sum()
        out.append(np.abs(A[:, f] - B[:, f]).sum() / den if den > 0 else np.abs(A[:, f] - B[:, f]).sum())
    return np.array(out)


_eos = {}


def helm_eos():
    """The device table of the polytrope workload in its units; the oracle's (global) table and
    units are set to the same on every call, since other tests load their own."""
    rows = hc.table_rows()
    orc.helm_table(rows)
    orc.helm_set_units(*workloads.HELM_UNITS)
    if "E" not in _eos:
        E = amrhydro.HelmholtzEos(rows)
        E.set_units(*workloads.HELM_UNITS)
        _eos["E"] = E
    return _eos["E"]


def setup(eos, recon, seed, counts=(8, 40), f_r=0.0, boundary=amrhydro.AMRH_BOUNDARY_EXTERNAL, precision=0):
    g = 3 if recon == 1 else 2
    B = rc.batch(counts=counts, eos="helm" if eos == 1 else "ideal", seed=seed, g=g)
    om = B["omega"]
    cfg_o = orc.rot_cfg(eos=eos, recon=recon, omega=om, omega0=om, f_r=f_r, f_theta=0.3)
    cfg = amrhydro.rotating_config(eos=eos, recon=recon, omega=om, omega0=om, f_r=f_r, f_theta=0.3,
                                   boundary=boundary, precision=precision)
    G = amrhydro.Batch(cfg, B["U"].shape[0], B["h"])
    G.set_block_origin(B["x0"])
    G.set_potential(B["phi"])
    if eos == 1:
        G.attach_eos(helm_eos())
    return B, cfg_o, G


@pytest.mark.parametrize("eos,recon", [(0, 1), (1, 1), (0, 0)])
def test_rotating_substeps_match_oracle(eos, recon):
    B, cfg, G = setup(eos, recon, seed=11)
    g = cfg.g
    U, phi, x0, h = B["U"], B["phi"], B["x0"], B["h"]
    say("rotating eos=%d recon=%d: %d blocks" % (eos, recon, U.shape[0]))
    m = orc.rot_max_eig(cfg, U, x0, h)
    dt = 0.1 * cfg.cfl * (h / m).min()
    ref, ref0 = U.copy(), U.copy()
    G.upload(U)
    G.store()
    for beta in (1.0, 0.25, 2.0 / 3.0):
        orc.rot_outflow_ghosts(cfg, ref, x0, h)
        flags = orc.rot_substep(cfg, ref, ref0, phi, x0, h, dt, beta)
        assert not flags.any()
        G.fill_outflow_ghosts()
        G.substep(dt, beta)
        out = G.download()
        e = rel_l1_rot(out, ref, g)
        I = (slice(None), slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
        same = np.mean(out[I] == ref[I])
        say("  beta=%.4f rel-L1 %s  bit-identical %.6f" % (beta, np.array2string(e, precision=1), same))
        assert np.all(e <= TOL), e
        assert G.error_flags() == 0
    G.close()


@pytest.mark.parametrize("eos", [0, 1])
def test_full_size_rotating_batch_rk3_step(eos):
    """C4 and C5 at their full size (workloads.binary_batch: 20,000 leaves of 8^3, ideal gas or
    Helmholtz EOS, PPM, outflow): one RK3 step of the whole batch on the GPU; every 40th block
    (blocks are independent) against the oracle, the FP64 form <= 1e-12 relative L1 per field, the
    mixed form density <= 1e-6."""
    Bn = workloads.binary_batch(eos="helm" if eos == 1 else "ideal")
    U, phi, x0, h = Bn["U"], Bn["phi"], Bn["x0"], Bn["h"]
    om = Bn["omega"]
    cfg_o = orc.rot_cfg(eos=eos, recon=1, omega=om)
    E = helm_eos() if eos == 1 else None  # the table, for the oracle and the GPU
    sel = np.arange(0, U.shape[0], 40)
    ref = np.ascontiguousarray(U[sel])
    ps, xs, hs = np.ascontiguousarray(phi[sel]), np.ascontiguousarray(x0[sel]), np.ascontiguousarray(h[sel])
    orc.rot_outflow_ghosts(cfg_o, ref, xs, hs)
    m = orc.rot_max_eig(cfg_o, ref, xs, hs)
    dt = 0.1 * cfg_o.cfl * (hs / m).min()
    ref0 = ref.copy()
    for beta in (1.0, 0.25, 2.0 / 3.0):
        orc.rot_outflow_ghosts(cfg_o, ref, xs, hs)
        assert not orc.rot_substep(cfg_o, ref, ref0, ps, xs, hs, dt, beta).any()
    say("full batch, eos=%d: %d blocks, %d checked against the oracle" % (eos, U.shape[0], len(sel)))
    outs = []
    for prec in (amrhydro.AMRH_FP64, amrhydro.AMRH_MIXED):
        cfg = amrhydro.rotating_config(eos=eos, omega=om, precision=prec)
        G = amrhydro.Batch(cfg, U.shape[0], h)
        G.set_block_origin(x0)
        G.set_potential(phi)
        if E is not None:
            G.attach_eos(E)
        G.upload(U)
        G.store()
        for beta in (1.0, 0.25, 2.0 / 3.0):
            G.fill_outflow_ghosts()
            G.substep(dt, beta)
        assert G.error_flags() == 0
        outs.append(G.download()[sel])
        G.close()
    e64 = rel_l1_rot(outs[0], ref, 3)
    e32 = rel_l1_rot(outs[1], ref, 3)
    say("  fp64 vs oracle %s" % np.array2string(e64, precision=1))
    say("  mixed vs oracle %s" % np.array2string(e32, precision=1))
    assert np.all(e64 <= TOL), e64
    assert e32[rot_field("rho")] <= 1e-6, e32


def test_rotating_ghosts_match_oracle():
    B, cfg, G = setup(0, 1, seed=12)
    U = B["U"].copy()
    ref = U.copy()
    orc.rot_outflow_ghosts(cfg, ref, B["x0"], B["h"])
    G.upload(U)
    G.fill_outflow_ghosts()
    out = G.download()
    g, Gp = 3, 14
    mask = np.zeros((Gp, Gp, Gp), bool)
    inner = slice(g, Gp - g)
    mask[:, inner, inner] = True
    mask[inner, :, inner] = True
    mask[inner, inner, :] = True
    np.testing.assert_array_equal(out[:, :, mask], ref[:, :, mask])


@pytest.mark.parametrize("eos", [0, 1])
def test_rotating_max_eig_matches_oracle(eos):
    B, cfg, G = setup(eos, 1, seed=13)
    ref = orc.rot_max_eig(cfg, B["U"], B["x0"], B["h"])
    G.upload(B["U"])
    got = G.block_max_eig()
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)


def test_rotating_com_force_terms():
    """A nonzero centre-of-mass correction force (State::fR, ftheta) goes through atan2/cos/sin,
    which may differ from glibc by an ulp."""
    B, cfg, G = setup(0, 1, seed=14, f_r=0.05)
    U, phi, x0, h = B["U"], B["phi"], B["x0"], B["h"]
    dt = 1e-4
    ref = U.copy()
    orc.rot_outflow_ghosts(cfg, ref, x0, h)
    orc.rot_substep(cfg, ref, ref.copy(), phi, x0, h, dt, 1.0)
    G.upload(U)
    G.fill_outflow_ghosts()
    G.store()
    G.substep(dt, 1.0)
    e = rel_l1_rot(G.download(), ref, 3)
    assert np.all(e <= TOL), e


@pytest.mark.parametrize("eos", [0, 1])
def test_rotating_rk3_graph_matches_direct_and_oracle(eos):
    """Four RK3 steps with the device dt control, outflow ghosts written after every substep:
    direct launches, captured graphs and a persistent grid agree bit for bit; the FP64 path stays
    on the oracle's trajectory (dt read back from the device)."""
    outs = []
    for mode in ("direct", "graph", "persistent+graph"):
        B, cfg, G = setup(eos, 1, seed=15, boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW)
        G.upload(B["U"])
        G.fill_outflow_ghosts()
        if mode == "persistent+graph":
            G.set_launch(7, 9)
        dt = torch.full((1,), 1e-5, dtype=torch.float64, device="cuda")
        G.run_steps(4, dt.data_ptr(), rk=3, use_graph=(mode != "direct"))
        G.sync()
        outs.append((G.download(), float(dt.cpu()[0])))
        assert G.error_flags() == 0
        G.close()
    for o, d in outs[1:]:
        np.testing.assert_array_equal(o, outs[0][0])
        assert d == outs[0][1]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("eos,recon", [(0, 1), (1, 1), (0, 0)])
def test_rotating_deferred_step_equals_plain(eos, recon, precision):
    """The rotating state's deferred-dt step (the first substep's face speeds as the CFL maximum,
    its D stored, the first stage formed by an elementwise pass; with the Helmholtz EOS the later
    substeps split the same way) gives the bits of the plain schedule (CFL scan, store, substeps),
    direct, as graphs and with the autotuner's launch shapes and split choice, dt included"""
    outs = []
    for sched, use_graph in ((amrhydro.AMRH_SCHEDULE_PLAIN, False), (amrhydro.AMRH_SCHEDULE_DEFERRED, False),
                             (amrhydro.AMRH_SCHEDULE_DEFERRED, True), (None, True)):
        B, cfg, G = setup(eos, recon, seed=17, boundary=amrhydro.AMRH_BOUNDARY_OUTFLOW, precision=precision)
        G.upload(B["U"])
        if sched is None:
            G.autotune()
            g2, split, ms_sub, ms_split = G.deferred_launch_info(2)
            say("rotating eos=%d: later substeps split %d (%.3f vs %.3f ms)" % (eos, split, ms_split, ms_sub))
            assert split in (0, 1) and ms_split > 0.0
        else:
            G.set_schedule(sched)
            assert G.schedule() == sched
        dt = torch.full((1,), 1e-5, dtype=torch.float64, device="cuda")
        G.run_steps(3, dt.data_ptr(), rk=3, use_graph=use_graph)
        G.sync()
        outs.append((G.download(), float(dt.cpu()[0])))
        assert G.error_flags() == 0
        if sched == amrhydro.AMRH_SCHEDULE_DEFERRED and not use_graph:
            st = G.step_timed(dt.data_ptr(), rk=3)
            assert [k for k, _ in st] == [amrhydro.AMRH_LAUNCH_FIRST, amrhydro.AMRH_LAUNCH_DT,
                                          amrhydro.AMRH_LAUNCH_FIRST_STAGE, amrhydro.AMRH_LAUNCH_SUBSTEP,
                                          amrhydro.AMRH_LAUNCH_SUBSTEP]
        G.close()
    say("rotating eos=%d recon=%d prec=%d: dt after 3 steps %.6e" % (eos, recon, precision, outs[0][1]))
    for o, d in outs[1:]:
        np.testing.assert_array_equal(o, outs[0][0])
        assert d == outs[0][1]


@pytest.mark.parametrize("eos,steps", [(0, 100), (1, 20)])
def test_rotating_mixed_rk3_accuracy(eos, steps):
    """The mixed form (FP32 reconstruction, face states, EOS inputs and fluxes; FP64 flux
    differences, sources, update and floor) against the FP64 oracle over RK3 steps at a fixed dt:
    density within 1e-6 relative L1."""
    B, cfg, G = setup(eos, 1, seed=21, precision=amrhydro.AMRH_MIXED)
    U, phi, x0, h = B["U"], B["phi"], B["x0"], B["h"]
    m = orc.rot_max_eig(cfg, U, x0, h)
    dt = 0.2 * cfg.cfl * (h / m).min()
    ref = U.copy()
    G.upload(U)
    for _ in range(steps):
        ref0 = ref.copy()
        for beta in (1.0, 0.25, 2.0 / 3.0):
            orc.rot_outflow_ghosts(cfg, ref, x0, h)
            assert not orc.rot_substep(cfg, ref, ref0, phi, x0, h, dt, beta).any()
        G.rk_step(dt, 3)
    out = G.download()
    assert G.error_flags() == 0
    e = rel_l1_rot(out, ref, 3)
    say("mixed eos=%d, %d RK3 steps: rel-L1 per field %s" % (eos, steps, np.array2string(e, precision=2)))
    assert e[rot_field("rho")] <= 1e-6
    for f in ("l_z", "E", "rho_acc", "rho_don"):
        assert e[rot_field(f)] <= 1e-5, f


def rot_field(name):
    return amrhydro.ROT_FIELDS.index(name)


@pytest.mark.parametrize("eos", [0, 1])
def test_rotating_mixed_max_eig(eos):
    B, cfg, G = setup(eos, 1, seed=13, precision=amrhydro.AMRH_MIXED)
    ref = orc.rot_max_eig(cfg, B["U"], B["x0"], B["h"])
    G.upload(B["U"])
    got = G.block_max_eig()
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=0)
