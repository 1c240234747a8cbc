"""Exact (sequential-order replay) mode vs the oracle: bit-identical state.

Both sides draw every call site from the same Philox stream, so the GPU's
exact sweep must reproduce the oracle's mc_coag step by step: the same
tot_n_samp / tot_n_coag, the same particle ids, groups, bins and list
order, and species volumes equal (relative 1e-12; in practice bitwise,
since merges only add stored doubles).
"""
import numpy as np
import pytest

from helpers import compare_state, gpu_boxes, oracle_box, small
from partmc_amd import populations as P

pytestmark = pytest.mark.gpu


def run_pair(pop, kernel, dt, n_steps, seed):
    ob = oracle_box(pop, seed)
    g = gpu_boxes([pop], seed)
    for step in range(n_steps):
        os_, oc = ob.mc_coag(kernel, dt, step)
        gs, gc = g.mc_coag(kernel, dt)
        assert (os_, oc) == (int(gs[0]), int(gc[0])), f"step {step}: counters differ"
    return ob, g


def test_grid_and_bins_match_oracle():
    pop = small(P.config_c2(), 20000)
    ob = oracle_box(pop, 7)
    g = gpu_boxes([pop], 7)
    oe, oc, ow = ob.grid()
    ge, gc, gw = g.grid(0)
    np.testing.assert_allclose(ge, oe, rtol=1e-15)
    np.testing.assert_allclose(gc, oc, rtol=1e-15)
    ocount, oent = ob.lists()
    gcount, gent = g.lists(0)
    np.testing.assert_array_equal(ocount, gcount)
    np.testing.assert_array_equal(oent, gent)


def test_kbounds_match_oracle_brownian():
    pop = small(P.config_c2(), 5000)
    ob, g = run_pair(pop, P.KERNEL_BROWN, 60.0, 1, 11)
    okmin, okmax = ob.kmax()
    gkmin, gkmax = g.kbounds(0)
    np.testing.assert_allclose(gkmin, okmin, rtol=1e-12)
    np.testing.assert_allclose(gkmax, okmax, rtol=1e-12)


def test_exact_additive_c1prime():
    pop = P.config_c1prime(n_part=4000)
    ob, g = run_pair(pop, P.KERNEL_ADDITIVE, 1.0, 120, 42)
    compare_state(ob, g)


def test_exact_brownian_power_weighted():
    pop = small(P.config_c2(), 6000)
    ob, g = run_pair(pop, P.KERNEL_BROWN, 60.0, 20, 5)
    compare_state(ob, g)


def test_exact_sedi():
    pop = small(P.config_c5(n_box=1, n_part=20000)[0], 4000)
    ob, g = run_pair(pop, P.KERNEL_SEDI, 1.0, 30, 9)
    compare_state(ob, g)


def test_exact_brownian_20_species():
    pop = small(P.config_c3(n_part=20000), 3000)
    pop.weight_magnitude *= 1e4  # dense enough that events happen
    ob, g = run_pair(pop, P.KERNEL_BROWN, 60.0, 10, 3)
    compare_state(ob, g)


@pytest.mark.parametrize("ibox", [0, 3])
def test_exact_brownian_fractional_power_weight(ibox):
    pop = small(P.config_c4(n_box=4, n_min=20000, n_max=40000)[ibox], 4000)
    ob, g = run_pair(pop, P.KERNEL_BROWN, 30.0, 15, 21 + ibox)
    compare_state(ob, g)
