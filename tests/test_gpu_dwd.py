"""The Helmholtz-EOS workload as specified: a 0.6 + 0.9 solar-mass C/O double white dwarf (abar
13.714, zbar 6.857; workloads.dwd_batch) on the cold-Fermi-gas + radiation table
(workloads.helmholtz_table_dwd), the floor's Newton solve to 1e-10, through the C ABI:
  * the full 20,000-leaf batch, one RK3 step: every 40th block (blocks are independent) against
    the oracle -- FP64 <= 1e-12 relative L1 per field, the mixed form (FP32 face EOS) density
    <= 1e-6;
  * the mixed form after 20 RK3 steps of a 48-block batch against the FP64 oracle: density
    <= 1e-6 relative L1."""
import numpy as np
import pytest

import orc
from gpu_util import amrhydro, say
import workloads
from test_gpu_rotating import rel_l1_rot, rot_field

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-12
NEWTON = 1e-10
BETAS = (1.0, 0.25, 2.0 / 3.0)
_cache = {}


def dwd_table(units):
    """The device table and the oracle's (global) table and units for the double white dwarf; the
    oracle's are set again on every call, since other tests load their own."""
    if "rows" not in _cache:
        _cache["rows"] = workloads.helmholtz_table_dwd()
    rows = _cache["rows"]
    orc.helm_table(rows)
    orc.helm_set_units(*units)
    if "E" not in _cache:
        E = amrhydro.HelmholtzEos(rows)
        E.set_units(*units)
        _cache["E"] = E
    return _cache["E"]


def configs(Bn, precision):
    om = Bn["omega"]
    cfg = amrhydro.rotating_config(eos=amrhydro.AMRH_EOS_HELMHOLTZ, omega=om, precision=precision, abar=Bn["abar"],
                                   newton_tol=NEWTON)
    cfg_o = orc.rot_cfg(eos=1, recon=1, omega=om, abar=Bn["abar"], newton_tol=NEWTON)
    return cfg, cfg_o


def device_batch(Bn, cfg, E):
    G = amrhydro.Batch(cfg, Bn["U"].shape[0], Bn["h"])
    G.set_block_origin(Bn["x0"])
    G.set_potential(Bn["phi"])
    G.attach_eos(E)
    G.upload(Bn["U"])
    return G


@pytest.mark.timeout(900)
def test_full_size_dwd_rk3_step():
    say("building the 20,000-leaf double-white-dwarf batch")
    Bn = workloads.dwd_batch()
    U, phi, x0, h = Bn["U"], Bn["phi"], Bn["x0"], Bn["h"]
    E = dwd_table(Bn["units"])
    _, cfg_o = configs(Bn, amrhydro.AMRH_FP64)
    sel = np.arange(0, U.shape[0], 40)
    ref = np.ascontiguousarray(U[sel])
    ps, xs, hs = np.ascontiguousarray(phi[sel]), np.ascontiguousarray(x0[sel]), np.ascontiguousarray(h[sel])
    orc.rot_outflow_ghosts(cfg_o, ref, xs, hs)
    m = orc.rot_max_eig(cfg_o, ref, xs, hs)
    dt = 0.1 * cfg_o.cfl * (hs / m).min()
    ref0 = ref.copy()
    for beta in BETAS:
        orc.rot_outflow_ghosts(cfg_o, ref, xs, hs)
        assert not orc.rot_substep(cfg_o, ref, ref0, ps, xs, hs, dt, beta).any()
    say("DWD full batch: %d blocks, %d checked against the oracle, dt %.3e" % (U.shape[0], len(sel), dt))
    outs = []
    for prec in (amrhydro.AMRH_FP64, amrhydro.AMRH_MIXED):
        cfg, _ = configs(Bn, prec)
        G = device_batch(Bn, cfg, E)
        G.store()
        for beta in BETAS:
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


@pytest.mark.timeout(600)
def test_dwd_mixed_20_steps_near_fp64_oracle():
    Bn = workloads.dwd_batch(counts=(8, 40))
    U, phi, x0, h = Bn["U"], Bn["phi"], Bn["x0"], Bn["h"]
    E = dwd_table(Bn["units"])
    cfg, cfg_o = configs(Bn, amrhydro.AMRH_MIXED)
    G = device_batch(Bn, cfg, E)
    ref = U.copy()
    orc.rot_outflow_ghosts(cfg_o, ref, x0, h)
    dt = 0.5 * cfg_o.cfl * (h / orc.rot_max_eig(cfg_o, ref, x0, h)).min()
    steps = 20
    for s in range(steps):
        ref0 = ref.copy()
        for beta in BETAS:
            orc.rot_outflow_ghosts(cfg_o, ref, x0, h)
            assert not orc.rot_substep(cfg_o, ref, ref0, phi, x0, h, dt, beta).any()
        G.rk_step(dt, 3)
    out = G.download()
    assert G.error_flags() == 0
    e = rel_l1_rot(out, ref, 3)
    say("DWD mixed, %d RK3 steps (dt %.3e): rel-L1 per field %s" % (steps, dt, np.array2string(e, precision=2)))
    assert e[rot_field("rho")] <= 1e-6, e
    for f in ("l_z", "E", "rho_acc", "rho_don"):
        assert e[rot_field(f)] <= 1e-5, f
    G.close()
