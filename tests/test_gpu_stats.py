"""In-situ size distributions against bin_grid_histogram_1d of the same
particles (src/bin_grid.F90:225-248, src/extract_aero_size.F90:132-138).

The device sums each bin after a stable sort by bin, in a tree rather than
in particle order, so values agree to rounding (relative 1e-12)."""
import numpy as np
import pytest

import oracle
from helpers import gpu_boxes, radius, small
from partmc_amd import MODE_PARALLEL, PREC_FP32_RECHECK, populations as P

pytestmark = pytest.mark.gpu


def expected(g, box, pop, n_bin, d_min, d_max, quantity):
    d = g.download(box)
    e, _, wd = oracle.bin_grid_make_log(n_bin, d_min, d_max)
    r = radius(d["vol"])
    w = oracle.Weights([[pop.weight_type]], [[pop.weight_magnitude]], [[pop.weight_exponent]])
    nc = np.array([w.num_conc_at_radius(int(k), float(x)) for k, x in zip(d["klass"], r)])
    if quantity == "mass":
        nc = nc * (d["vol"] * np.asarray(pop.density)[None, :]).sum(axis=1)
    elif quantity == "vol":
        nc = nc * d["vol"].sum(axis=1)
    return oracle.histogram_1d(e, wd, 2.0 * r, nc)


@pytest.mark.parametrize("quantity", ["num", "mass", "vol"])
def test_size_dist_matches_histogram_after_coagulation(quantity):
    pop = small(P.config_c2(), 20000)
    g = gpu_boxes([pop], 3)
    g.set_mode(MODE_PARALLEL, PREC_FP32_RECHECK, 1e-5)
    for _ in range(5):
        g.mc_coag(P.KERNEL_BROWN, 60.0)
    # a grid narrower than the population, so both ends drop particles
    n_bin, d_min, d_max = 70, 2e-8, 5e-7
    got = g.size_dist(0, n_bin, d_min, d_max, quantity)
    ref = expected(g, 0, pop, n_bin, d_min, d_max, quantity)
    assert np.count_nonzero(ref) > n_bin // 2
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)


def test_size_dist_empty_box_and_bad_args():
    pop = small(P.config_c1prime(), 1000)
    g = gpu_boxes([pop, pop], 1)
    g.upload(1, np.zeros((0, pop.vol.shape[1])), np.zeros(0, dtype=np.int64),
             np.zeros(0), np.zeros(0), capacity=64)
    assert not g.size_dist(1, 10, 1e-9, 1e-5).any()
    tot = g.size_dist(0, 200, 1e-12, 1.0).sum() * (np.log(1.0) - np.log(1e-12)) / 200
    ref = pop.n_part * pop.weight_magnitude  # flat weight: magnitude is m^-3 per particle
    np.testing.assert_allclose(tot, ref, rtol=1e-12)
    with pytest.raises(Exception):
        g.size_dist(0, 10, 1e-5, 1e-9)
    with pytest.raises(Exception):
        g.size_dist(0, 0, 1e-9, 1e-5)
