"""The Fortran host binding (partmc_amd/fortran/pmc_coag_gpu.F90) drives the
GPU step the way run_part_timestep drives mc_coag; in exact mode its run is
bit-identical to the oracle's."""
import os
import subprocess

import numpy as np
import pytest

from helpers import oracle_box
from partmc_amd import populations as P
from partmc_amd.build import DRIVER

pytestmark = pytest.mark.gpu


def test_fortran_driver_matches_oracle(tmp_path):
    assert os.path.exists(DRIVER), "build first: python -m partmc_amd.build"
    pop = P.config_c1prime(n_part=3000)
    seed, n_steps, dt = 31, 60, 1.0
    pop.density.astype(np.float64).tofile(tmp_path / "dens.bin")
    np.ascontiguousarray(pop.vol).tofile(tmp_path / "vol.bin")  # vol(n_spec, n_part)
    pop.ids().tofile(tmp_path / "ids.bin")
    n_bin, rmin, rmax = pop.grid
    args = [DRIVER, str(tmp_path), str(pop.n_spec), str(pop.n_part), str(n_steps),
            str(P.KERNEL_ADDITIVE), repr(dt), str(seed), repr(pop.temp), repr(pop.pressure),
            repr(pop.beta), repr(pop.weight_magnitude), str(n_bin), repr(rmin), repr(rmax), "0"]
    subprocess.run(args, check=True, timeout=120)
    counters = np.loadtxt(tmp_path / "counters.txt", dtype=np.int64).reshape(-1, 3)
    ob = oracle_box(pop, seed)
    for step in range(n_steps):
        s, c = ob.mc_coag(P.KERNEL_ADDITIVE, dt, step)
        assert (counters[step, 1], counters[step, 2]) == (s, c), step
    n = int(open(tmp_path / "out_n.txt").read())
    o = ob.particles()
    assert n == len(o["id"])
    np.testing.assert_array_equal(np.fromfile(tmp_path / "out_ids.bin", dtype=np.int64), o["id"])
    vol = np.fromfile(tmp_path / "out_vol.bin", dtype=np.float64).reshape(n, pop.n_spec)
    np.testing.assert_allclose(vol, o["vol"], rtol=1e-12, atol=0)


def test_fortran_driver_with_rebalance_matches_oracle(tmp_path):
    """Each step followed by aero_state_rebalance, as run_part_timestep
    does: the population is halved early and doubled once coagulation has
    thinned it, bit-identical to the oracle."""
    pop = P.config_c1prime(n_part=3000)
    seed, n_steps, dt, ideal = 17, 80, 8.0, 1200.0
    pop.density.astype(np.float64).tofile(tmp_path / "dens.bin")
    np.ascontiguousarray(pop.vol).tofile(tmp_path / "vol.bin")
    pop.ids().tofile(tmp_path / "ids.bin")
    n_bin, rmin, rmax = pop.grid
    args = [DRIVER, str(tmp_path), str(pop.n_spec), str(pop.n_part), str(n_steps),
            str(P.KERNEL_ADDITIVE), repr(dt), str(seed), repr(pop.temp), repr(pop.pressure),
            repr(pop.beta), repr(pop.weight_magnitude), str(n_bin), repr(rmin), repr(rmax), "0",
            repr(ideal)]
    subprocess.run(args, check=True, timeout=120)
    counters = np.loadtxt(tmp_path / "counters.txt", dtype=np.int64).reshape(-1, 3)
    ob = oracle_box(pop, seed)
    for step in range(n_steps):
        s, c = ob.mc_coag(P.KERNEL_ADDITIVE, dt, step)
        assert (counters[step, 1], counters[step, 2]) == (s, c), step
        ob.rebalance(ideal, step=step + 1)
    n = int(open(tmp_path / "out_n.txt").read())
    o = ob.particles()
    assert n == len(o["id"])
    np.testing.assert_array_equal(np.fromfile(tmp_path / "out_ids.bin", dtype=np.int64), o["id"])
    vol = np.fromfile(tmp_path / "out_vol.bin", dtype=np.float64).reshape(n, pop.n_spec)
    np.testing.assert_allclose(vol, o["vol"], rtol=1e-12, atol=0)
    wm = float(open(tmp_path / "out_wmag.txt").read())
    assert wm == ob.weight_magnitudes()[0, 0]
    acts = ob.info()[1]
    assert (acts == 3).sum() > 0  # halved at the first step
    assert wm < pop.weight_magnitude  # then doubled more often than halved
