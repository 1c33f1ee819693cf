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
