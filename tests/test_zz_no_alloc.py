"""amrh_state_sum_dev, the allocation-free form of amrh_state_sum, through the C ABI: the same bits
as amrh_state_sum on two calls in a row (the second on a smaller batch, so stale partials of the
first would show), on a flat batch and on an octree with covered cells; and its size limit."""
import numpy as np
import pytest

from gpu_util import amrhydro, say  # noqa: I001 (puts the package on the path)
import workloads
from test_gpu_rot_tree import device_batch
from test_oracle_rot_tree import binary_tree_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def flat(nb, seed):
    U = workloads.flat_batch(nb, n=8, seed=seed)
    B = amrhydro.Batch(amrhydro.euler_config(n=8), nb, np.full(nb, 1.0 / 8))
    B.upload(U)
    B.set_block_origin(np.random.default_rng(seed).uniform(-1.0, 1.0, (nb, 3)))
    return B


def dev_sums(B, nf):
    out = torch.full((4 * nf,), float("nan"), dtype=torch.float64, device="cuda")
    B.state_sum_dev(out.data_ptr())
    B.sync()
    return out.cpu().numpy().reshape(4, nf)


def test_state_sum_dev_equals_state_sum_twice_flat():
    for nb, seed in ((700, 1), (37, 2)):
        say("no-alloc state_sum: flat batch of %d blocks" % nb)
        B = flat(nb, seed)
        want = B.state_sum()
        np.testing.assert_array_equal(dev_sums(B, 6), want)
        B.close()


def test_state_sum_dev_equals_state_sum_twice_on_octrees():
    for name in ("mid", "partial"):
        T, cfg = binary_tree_case(name, seed=40)
        say("no-alloc state_sum: rotating octree %s, %d nodes" % (name, T.nn))
        B = device_batch(T, cfg)
        want = B.state_sum()
        np.testing.assert_array_equal(dev_sums(B, 13), want)
        B.close()


def test_state_sum_dev_rejects_batches_beyond_its_scratch():
    nb = amrhydro.STATE_SUM_DEV_MAX_PARTIALS // (6 * 6) + 1
    B = amrhydro.Batch(amrhydro.euler_config(n=8), nb, np.full(nb, 1.0 / 8))
    out = torch.zeros(24, dtype=torch.float64, device="cuda")
    with pytest.raises(amrhydro.AmrhError, match="no-allocation scratch"):
        B.state_sum_dev(out.data_ptr())
    B.close()
