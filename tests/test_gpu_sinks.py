"""Sources and sinks of scenario_update_aero_state on the device against the
oracle (src/scenario.F90:259-341): dilution out, particle loss, weight
scaling and appended particles.  Removal is sequential by definition
(swap-last from the last particle down), so states are compared bit for
bit, and in exact mode the steps that follow as well."""
import numpy as np
import pytest

import oracle
from helpers import compare_state, gpu_boxes, oracle_box, small
from partmc_amd import CoagError, MODE_PARALLEL, PREC_FP32_RECHECK, populations as P

pytestmark = pytest.mark.gpu


def new_particles(n, n_spec, seed):
    rng = np.random.default_rng(seed)
    return 10 ** rng.uniform(-23.5, -21.5, size=(n, 1)) * rng.dirichlet(np.ones(n_spec), size=n)


def steps(ob, g, n, first):
    for s in range(first, first + n):
        gs, gc = g.mc_coag(P.KERNEL_BROWN, 60.0)
        os_, oc = ob.mc_coag(P.KERNEL_BROWN, 60.0, s)
        assert (os_, oc) == (int(gs[0]), int(gc[0])), f"step {s}: counters differ"
    return first + n


def test_exact_sinks_and_sources_match_oracle():
    pop = small(P.config_c2(), 6000)
    ob = oracle_box(pop, 13)
    g = gpu_boxes([pop], 13, capacity_factor=3)
    s = steps(ob, g, 3, 0)
    # dilution out, then constant-rate and dry-deposition loss
    ob.dilute(0.2, step=s)
    g.dilute([0.2])
    compare_state(ob, g)
    ob.particle_loss(oracle.LOSS_CONSTANT, 300.0, 1000.0, step=s)
    g.particle_loss(oracle.LOSS_CONSTANT, 300.0)
    compare_state(ob, g)
    n0 = ob.n_part
    ob.particle_loss(oracle.LOSS_DRYDEP, 3000.0, 0.5, step=s)
    g.particle_loss(oracle.LOSS_DRYDEP, 3000.0, [0.5])
    compare_state(ob, g)
    assert ob.n_part < n0
    # the computational volume follows T and p; emitted particles join
    ob.scale_weights(0.97)
    g.scale_weights(0, 0.97)
    vol = new_particles(300, pop.n_spec, 4)
    nid = ob.next_id
    ob.add_particles(vol, np.arange(nid, nid + 300), np.ones(300), np.ones(300))
    ob.next_id = nid + 300
    g.append(0, vol, np.ones(300), np.ones(300))
    compare_state(ob, g)
    np.testing.assert_array_equal(g.weights(0)[1], ob.weight_magnitudes())
    s = steps(ob, g, 3, s)
    compare_state(ob, g)


def test_parallel_mode_sinks_match_oracle_on_the_same_state():
    pop = small(P.config_c2(), 30000)
    g = gpu_boxes([pop], 6, capacity_factor=3)
    g.set_mode(MODE_PARALLEL, PREC_FP32_RECHECK, 1e-5)
    for _ in range(3):
        g.mc_coag(P.KERNEL_BROWN, 60.0)
    d = g.download(0)
    e, _, _ = g.grid(0)
    _, mag, _ = g.weights(0)
    ob = oracle.Box(pop.density, oracle.Weights([[pop.weight_type]], mag, [[pop.weight_exponent]]),
                    pop.temp, pop.pressure, pop.beta, seed=6, box_index=0)
    ob.add_particles(d["vol"], d["id"], d["group"], d["klass"], d["n_primary"])
    ob.next_id = g.next_id(0)
    ob.sort_with_grid(len(e) - 1, e[0], e[-1])
    g.clear_removed(0)
    ob.dilute(0.35, step=3)
    g.dilute([0.35])
    ob.particle_loss(oracle.LOSS_VOLUME, 3e5, 1000.0, step=3)
    g.particle_loss(oracle.LOSS_VOLUME, 3e5)
    o, d2 = ob.particles(), g.download(0)
    np.testing.assert_array_equal(o["id"], d2["id"])
    np.testing.assert_array_equal(o["vol"], d2["vol"])
    for a, b in zip(ob.info(), g.removed(0)):
        np.testing.assert_array_equal(a, b)
    assert len(ob.info()[0]) > 0.3 * len(d["id"])
    # appended particles outside the grid: the next step remakes it
    g.append(0, np.full((1, pop.n_spec), 1e-13), [1], [1])
    _, c = g.mc_coag(P.KERNEL_BROWN, 60.0)
    e2, _, _ = g.grid(0)
    r = (3 * g.download(0)["vol"].sum(axis=1) / (4 * np.pi)) ** (1 / 3)
    assert e2[0] < r.min() and r.max() < e2[-1]


def test_sink_argument_errors():
    pop = small(P.config_c1prime(), 500)
    g = gpu_boxes([pop], 2)
    with pytest.raises(CoagError):
        g.particle_loss(oracle.LOSS_CHAMBER, 1.0)
    with pytest.raises(CoagError):
        g.particle_loss(oracle.LOSS_DRYDEP, 1.0)  # no height
    with pytest.raises(CoagError):
        g.dilute([1.5])
    with pytest.raises(CoagError):
        g.append(0, np.ones((1, pop.n_spec)) * 1e-21, [2], [1])  # no group 2
    g.particle_loss(oracle.LOSS_NONE, 1.0)
    assert g.n_part(0) == pop.n_part
