"""Population control on the device against the oracle's restatement of
aero_state_rebalance / double / halve (src/aero_state.F90:1722-1852).

Halving is sequential by definition (swap-last removals from the last
particle down), so the device result is compared bit for bit: particle
order, ids, removal log, weights and, in exact mode, the bin lists and the
steps that follow."""
import numpy as np
import pytest

import oracle
from helpers import compare_state, gpu_boxes, oracle_box, small
from partmc_amd import MODE_PARALLEL, PREC_FP32_RECHECK, CoagError, populations as P

pytestmark = pytest.mark.gpu


def steps(ob, g, n, first, kernel=P.KERNEL_BROWN, dt=60.0, n_box=1):
    for s in range(first, first + n):
        gs, gc = g.mc_coag(kernel, dt)
        for b in range(n_box):
            os_, oc = ob[b].mc_coag(kernel, dt, s)
            assert (os_, oc) == (int(gs[b]), int(gc[b])), f"box {b} step {s}: counters differ"
    return first + n


def same_weights(ob, g, box=0):
    np.testing.assert_array_equal(g.weights(box)[1], ob.weight_magnitudes())


def test_exact_halve_then_double_match_oracle():
    pop = small(P.config_c2(), 6000)
    ob = oracle_box(pop, 21)
    g = gpu_boxes([pop], 21, capacity_factor=12)
    s = steps([ob], g, 4, 0)
    n = ob.n_part
    # halving: ideal n / 8 -> at least two passes
    ob.rebalance(n / 8, step=s)
    g.set_n_part_ideal(0, n / 8)
    g.rebalance()
    compare_state(ob, g)
    same_weights(ob, g)
    assert ob.n_part <= n / 4
    s = steps([ob], g, 4, s)
    compare_state(ob, g)
    # doubling: ideal 8 n' -> the sort is remade at the next step
    n = ob.n_part
    ob.rebalance(8 * n, step=s)
    g.set_n_part_ideal(0, 8 * n)
    g.rebalance()
    compare_state(ob, g, lists=False)
    same_weights(ob, g)
    assert ob.n_part >= 4 * n
    s = steps([ob], g, 3, s)
    compare_state(ob, g)


def test_exact_rebalance_per_box_plans():
    """Two boxes: one doubles while the other halves; each matches its
    own oracle box."""
    pops = [small(P.config_c2(), 3000, seed=1), small(P.config_c2(), 5000, seed=2)]
    obs = [oracle_box(p, 8, box_index=i) for i, p in enumerate(pops)]
    g = gpu_boxes(pops, 8, capacity_factor=12)
    s = steps(obs, g, 3, 0, n_box=2)
    ideal = [4.0 * obs[0].n_part, obs[1].n_part / 5.0]
    for b in range(2):
        obs[b].rebalance(ideal[b], step=s)
        g.set_n_part_ideal(b, ideal[b])
    g.rebalance()
    for b in range(2):
        compare_state(obs[b], g, box=b, lists=(b == 1))
        same_weights(obs[b], g, b)
    s = steps(obs, g, 3, s, n_box=2)
    for b in range(2):
        compare_state(obs[b], g, box=b)


def test_parallel_mode_rebalance_matches_oracle_on_the_same_state():
    pop = small(P.config_c2(), 30000)  # several tiles of particles
    g = gpu_boxes([pop], 5, capacity_factor=12)
    g.set_mode(MODE_PARALLEL, PREC_FP32_RECHECK, 1e-5)
    for _ in range(4):
        g.mc_coag(P.KERNEL_BROWN, 60.0)
    d = g.download(0)
    e, _, _ = g.grid(0)
    _, mag, _ = g.weights(0)
    ob = oracle.Box(pop.density, oracle.Weights([[pop.weight_type]], mag, [[pop.weight_exponent]]),
                    pop.temp, pop.pressure, pop.beta, seed=5, box_index=0)
    ob.add_particles(d["vol"], d["id"], d["group"], d["klass"], d["n_primary"])
    ob.next_id = g.next_id(0)
    ob.sort_with_grid(len(e) - 1, e[0], e[-1])
    ob.clear_info()
    g.clear_removed(0)
    n = ob.n_part
    ob.rebalance(n / 6, step=4)
    g.set_n_part_ideal(0, n / 6)
    g.rebalance()
    o, d2 = ob.particles(), g.download(0)
    np.testing.assert_array_equal(o["id"], d2["id"])
    np.testing.assert_array_equal(o["vol"], d2["vol"])
    np.testing.assert_array_equal(o["n_primary"], d2["n_primary"])
    for a, b in zip(ob.info(), g.removed(0)):
        np.testing.assert_array_equal(a, b)
    same_weights(ob, g)
    # the parallel step carries on from the new population
    tot = 0
    for _ in range(3):
        _, c = g.mc_coag(P.KERNEL_BROWN, 60.0)
        tot += int(c[0])
    d3 = g.download(0)
    assert tot > 0 and len(np.unique(d3["id"])) == len(d3["id"])
    assert d3["id"].max() < g.next_id(0)


def test_doubling_beyond_capacity_fails_before_any_change():
    pop = small(P.config_c1prime(), 1000)
    g = gpu_boxes([pop], 3, capacity_factor=2)
    g.set_n_part_ideal(0, 100 * pop.n_part)
    with pytest.raises(CoagError):
        g.rebalance()
    assert g.n_part(0) == pop.n_part
    np.testing.assert_array_equal(g.weights(0)[1], [[pop.weight_magnitude]])
    with pytest.raises(CoagError):
        gpu_boxes([pop], 3).rebalance()  # no n_part_ideal


def test_exact_two_classes_halve_one_double_the_other():
    """Weight classes are rebalanced separately: with 90% of the particles
    in class 1 and an ideal of 0.3 n per class, class 1 halves and class 2
    doubles in the same call; the state and the steps after it match the
    oracle."""
    from partmc_amd import CoagGPU
    rng = np.random.default_rng(12)
    n = 6000
    vol = 10 ** rng.uniform(-24, -21, size=(n, 1)) * rng.dirichlet([1, 1], size=n)
    klass = np.where(rng.uniform(size=n) < 0.9, 1, 2).astype(np.int32)
    dens = np.array([1800.0, 1000.0])
    wt, wm, we = [[1, 1]], [[3e8, 1e8]], [[0.0, 0.0]]
    ob = oracle.Box(dens, oracle.Weights(wt, wm, we), seed=12)
    ob.add_particles(vol, np.arange(1, n + 1), np.ones(n), klass)
    ob.next_id = n + 1
    ob.sort_with_grid(60, 1e-10, 1e-5)
    g = CoagGPU(1, dens, max_bin=256)
    g.set_rng(12, 0)
    g.set_weights(0, wt, wm, we)
    g.upload(0, vol, np.arange(1, n + 1), np.ones(n), klass, capacity=4 * n)
    g.set_next_id(0, n + 1)
    g.set_bin_grid(0, 60, 1e-10, 1e-5)
    for s in range(2):
        os_, oc = ob.mc_coag(P.KERNEL_BROWN, 10.0, s)
        gs, gc = g.mc_coag(P.KERNEL_BROWN, 10.0)
        assert (os_, oc) == (int(gs[0]), int(gc[0]))
    ob.rebalance(0.6 * ob.n_part, step=2)
    g.set_n_part_ideal(0, 0.6 * g.n_part(0))
    g.rebalance()
    compare_state(ob, g, n_class=2, lists=False)
    same_weights(ob, g)
    m = ob.weight_magnitudes()
    assert m[0, 0] > 3e8 and m[0, 1] < 1e8  # class 1 halved, class 2 doubled
    for s in range(2, 4):
        os_, oc = ob.mc_coag(P.KERNEL_BROWN, 10.0, s)
        gs, gc = g.mc_coag(P.KERNEL_BROWN, 10.0)
        assert (os_, oc) == (int(gs[0]), int(gc[0]))
    compare_state(ob, g, n_class=2)
