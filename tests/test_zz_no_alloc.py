"""pmc_gpu_size_dist_noalloc, the size distribution that allocates nothing,
against pmc_gpu_size_dist on the same particles.

Both run the same kernels, the first with its partial histograms in the
library's fixed scratch array, the second in memory it allocates per call,
so their results are bit-identical.  Two calls in a row, the second on a
smaller box with fewer bins, show that what the first call left in the
scratch does not reach the second's result."""
import numpy as np
import pytest
import torch

from helpers import gpu_boxes, small
from partmc_amd import MODE_PARALLEL, PREC_FP32_RECHECK, populations as P

pytestmark = pytest.mark.gpu


def noalloc(g, box, n_bin, d_min, d_max, quantity):
    out = torch.full((n_bin,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.size_dist_noalloc(box, n_bin, d_min, d_max, out.data_ptr(), quantity)
    g.sync()
    return out.cpu().numpy()


@pytest.mark.parametrize("quantity", ["num", "mass"])
def test_noalloc_equals_size_dist_on_two_calls(quantity):
    big = small(P.config_c2(), 30000)
    little = small(P.config_c2(), 3000, seed=2)
    g = gpu_boxes([big, little], 5)
    g.set_mode(MODE_PARALLEL, PREC_FP32_RECHECK, 1e-5)
    for _ in range(3):
        g.mc_coag(P.KERNEL_BROWN, 60.0)
    print(f"noalloc {quantity}: boxes of {g.n_part(0)} and {g.n_part(1)} particles", flush=True)
    calls = [(0, 1500, 1e-9, 1e-4), (1, 60, 2e-8, 5e-7)]
    for box, n_bin, d_min, d_max in calls:
        got = noalloc(g, box, n_bin, d_min, d_max, quantity)
        ref = g.size_dist(box, n_bin, d_min, d_max, quantity)
        assert np.count_nonzero(ref) > 10
        np.testing.assert_array_equal(got, ref)
    g.close()


def test_noalloc_rejects_bins_beyond_its_scratch_and_bad_args():
    pop = small(P.config_c1prime(), 1000)
    g = gpu_boxes([pop], 1)
    out = torch.zeros(4096, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="2048"):
        g.size_dist_noalloc(0, 2049, 1e-9, 1e-4, out.data_ptr())
    with pytest.raises(RuntimeError):
        g.size_dist_noalloc(0, 10, 1e-4, 1e-9, out.data_ptr())
    with pytest.raises(RuntimeError):
        g.size_dist_noalloc(0, 10, 1e-9, 1e-4, 0)
    # the largest size it takes still equals the allocating entry point
    got = noalloc(g, 0, 2048, 1e-12, 1.0, "num")
    np.testing.assert_array_equal(got, g.size_dist(0, 2048, 1e-12, 1.0))
    g.close()
