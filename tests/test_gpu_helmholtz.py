"""GPU parity of the Helmholtz EOS (helmholtz/helmeos.cpp:164-486) through the C ABI: the device
EOS and Newton solve for T against the oracle (itself bit-identical to the reference's compiled
helmeos.cpp) on the synthetic table.  p, e, cs, cv, and the Newton T are bit-identical; the
entropy goes through log and may differ by an ulp."""
import numpy as np
import pytest

import helm_cases as hc
import orc
from gpu_util import amrhydro, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BITWISE = (0, 1, 2, 3, 4, 5, 6, 7, 9)  # every eos_t field but s


@pytest.fixture(scope="module")
def eos():
    rows = hc.table_rows()
    orc.helm_table(rows)
    orc.helm_set_units()
    E = amrhydro.HelmholtzEos(rows)
    yield E
    E.close()


def device_run(E, st, solve_T=False):
    t = torch.from_numpy(st.copy()).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    (E.compute_T if solve_T else E.eval)(st.shape[0], t.data_ptr(), stream)
    torch.cuda.synchronize()
    return t.cpu().numpy()


def compare(got, ref):
    for c in BITWISE:
        np.testing.assert_array_equal(got[:, c], ref[:, c], err_msg=amrhydro.EOS_FIELDS[c])
    np.testing.assert_allclose(got[:, 8], ref[:, 8], rtol=1e-15, atol=0)


def test_eos_matches_oracle(eos):
    st = hc.states(100000, seed=11)
    say("helmholtz_eos on %d states" % len(st))
    ref = st.copy()
    assert not orc.helm_eos(ref).any()
    got = device_run(eos, st)
    assert eos.error_flags() == 0
    compare(got, ref)


def test_eos_in_code_units_matches_oracle(eos):
    units = (3.1e8, 2.2e33, 27.0, 1.0)
    st = hc.states(20000, seed=12)
    st[:, 6] /= units[1] / units[0] ** 3
    try:
        orc.helm_set_units(*units)
        eos.set_units(*units)
        ref = st.copy()
        assert not orc.helm_eos(ref).any()
        got = device_run(eos, st)
    finally:
        orc.helm_set_units()
        eos.set_units(1.0, 1.0, 1.0, 1.0)
    compare(got, ref)


def test_compute_T_matches_oracle(eos):
    st = hc.states(50000, seed=13)
    orc.helm_eos(st)
    rng = np.random.default_rng(14)
    st[:, 7] *= rng.uniform(0.3, 3.0, len(st))
    ref = st.copy()
    assert not orc.helm_eos(ref, solve_T=True).any()
    got = device_run(eos, st, solve_T=True)
    assert eos.error_flags() == 0
    compare(got, ref)


def test_out_of_table_states_raise_flags(eos):
    st = hc.states(4, seed=15)
    st[1, 6] = 1e30
    got = device_run(eos, st)
    assert eos.error_flags() & 8
    assert np.isnan(got[1, 0]) and np.isfinite(got[0, 0])


def test_table_file_loads_like_rows(eos, tmp_path):
    import workloads
    path = str(tmp_path / "helmholtz.table")
    workloads.write_helmholtz_table(path, hc.table_rows())
    E2 = amrhydro.HelmholtzEos(path=path)
    st = hc.states(5000, seed=16)
    np.testing.assert_array_equal(device_run(E2, st), device_run(eos, st))
    E2.close()


def f32_pair(E, st):
    """the FP32 face EOS's p and cs (amrh_eos_eval_f32) beside the oracle's FP64 ones"""
    ref = st.copy()
    assert not orc.helm_eos(ref).any()
    t = torch.from_numpy(st.copy()).cuda()
    E.eval_f32(st.shape[0], t.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    assert E.error_flags() == 0
    return got, ref


def test_fp32_face_eos_near_fp64(eos):
    """the mixed form's face EOS (monomial coefficients in FP32, FP64 assembly) against the FP64
    EOS on the non-degenerate electron table: p within 1e-6, cs within 4e-6 relative"""
    st = hc.states(50000, seed=15)
    got, ref = f32_pair(eos, st)
    ep = np.abs(got[:, 0] / ref[:, 0] - 1)
    ec = np.abs(got[:, 2] / ref[:, 2] - 1)
    say("FP32 face EOS, first table: max rel p %.2e cs %.2e" % (ep.max(), ec.max()))
    assert ep.max() < 1e-6 and ec.max() < 4e-6
    np.testing.assert_array_equal(got[:, 4:8], st[:, 4:8])  # inputs untouched


def test_fp32_face_eos_near_fp64_on_the_dwd_table():
    """the same on the double white dwarf's degenerate + radiation table, C/O 50/50, over the
    densities and temperatures of the two stars and their surroundings"""
    import workloads
    rows = workloads.helmholtz_table_dwd()
    orc.helm_table(rows)
    orc.helm_set_units()
    E = amrhydro.HelmholtzEos(rows)
    try:
        rng = np.random.default_rng(16)
        n = 50000
        st = np.zeros((n, 10))
        st[:, 4] = workloads.DWD_ABAR
        st[:, 5] = workloads.DWD_ABAR / 2.0
        st[:, 6] = 10.0 ** rng.uniform(-1.0, 7.5, n)
        st[:, 7] = 10.0 ** rng.uniform(5.5, 9.5, n)
        got, ref = f32_pair(E, st)
        ep = np.abs(got[:, 0] / ref[:, 0] - 1)
        ec = np.abs(got[:, 2] / ref[:, 2] - 1)
        say("FP32 face EOS, DWD table: max rel p %.2e cs %.2e" % (ep.max(), ec.max()))
        assert ep.max() < 1e-6 and ec.max() < 4e-6
    finally:
        E.close()
        orc.helm_table(hc.table_rows())  # the module's table again for the oracle
