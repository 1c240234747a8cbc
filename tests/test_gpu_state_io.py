"""Restart from the on-disk state: the GPU box and the oracle box loaded
from the same file (volumes = mass / density, as aero_state_input_netcdf
reads them) continue identically."""
import numpy as np
import pytest

import oracle
from helpers import compare_state, gpu_boxes, small
from partmc_amd import CoagGPU, populations as P
from partmc_amd.state_io import read_aero_state

pytestmark = pytest.mark.gpu


def test_exact_restart_from_file_matches_oracle(tmp_path):
    pop = small(P.config_c2(), 5000)
    g = gpu_boxes([pop], 4)
    for _ in range(4):
        g.mc_coag(P.KERNEL_BROWN, 60.0)
    path = tmp_path / "state_0001.nc"
    g.save_state(0, path, species_names=["SO4", "NO3", "OC", "H2O"])
    d = g.download(0)
    s = read_aero_state(path)
    np.testing.assert_array_equal(s["id"], d["id"])
    np.testing.assert_allclose(s["mass"] / pop.density[None, :], d["vol"], rtol=2.3e-16)
    assert len(s["removed"][0]) == len(g.removed(0)[0])
    # restart both sides from the file, on the same grid and streams
    e, _, _ = g.grid(0)
    h = CoagGPU(1, pop.density, max_bin=256)
    h.set_rng(4, 4)
    h.set_env(0, pop.temp, pop.pressure, pop.beta)
    h.load_state(0, path, capacity=2 * len(d["id"]) + 64)
    h.set_next_id(0, g.next_id(0))
    h.set_bin_grid(0, len(e) - 1, e[0], e[-1])
    ob = oracle.Box(pop.density, oracle.Weights(*s["weights"]), pop.temp, pop.pressure,
                    pop.beta, seed=4, box_index=0)
    ob.add_particles(s["mass"] / pop.density[None, :], s["id"], s["group"], s["klass"],
                     s["n_primary"], s["n_comp"], s["comp_source"], s["comp_time"])
    ob.next_id = g.next_id(0)
    ob.sort_with_grid(len(e) - 1, e[0], e[-1])
    for step in range(4, 8):
        os_, oc = ob.mc_coag(P.KERNEL_BROWN, 60.0, step)
        hs, hc = h.mc_coag(P.KERNEL_BROWN, 60.0)
        assert (os_, oc) == (int(hs[0]), int(hc[0]))
    compare_state(ob, h)
