"""Shared set-up: the same synthetic box in the oracle and on the GPU."""
import numpy as np

import oracle
from partmc_amd import populations as P


def oracle_box(pop, seed, box_index=0):
    w = oracle.Weights([[pop.weight_type]], [[pop.weight_magnitude]], [[pop.weight_exponent]])
    b = oracle.Box(pop.density, w, pop.temp, pop.pressure, pop.beta, seed=seed,
                   box_index=box_index)
    n = pop.n_part
    b.add_particles(pop.vol, pop.ids(), np.ones(n), np.ones(n))
    b.next_id = n + 1
    if pop.grid is not None:
        b.sort_with_grid(*pop.grid)
    return b


def gpu_boxes(pops, seed, capacity_factor=2, max_bin=256):
    from partmc_amd import CoagGPU
    g = CoagGPU(len(pops), pops[0].density, max_bin=max_bin)
    g.set_rng(seed, 0)
    for i, pop in enumerate(pops):
        g.load_population(i, pop, capacity=capacity_factor * pop.n_part + 64)
    return g


def small(pop, n, seed=1):
    """A random subset of n particles of a population (same environment)."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(pop.n_part, size=min(n, pop.n_part), replace=False)
    out = P.Population(**{**pop.__dict__})
    out.vol = pop.vol[idx].copy()
    # keep the physical number concentration: each particle stands for more
    out.weight_magnitude = pop.weight_magnitude * pop.n_part / out.vol.shape[0]
    out.source = None if pop.source is None else pop.source[idx].copy()
    return out


def radius(vol):
    v = np.asarray(vol).sum(axis=1)
    return (3.0 * v / 4.0 / np.pi) ** (1.0 / 3.0)


def compare_state(ob, g, n_class=1, box=0, lists=True):
    """The GPU box equals the oracle's: particles, lists (when the sort is
    valid), next id, removal log."""
    o = ob.particles()
    d = g.download(box)
    assert len(o["id"]) == len(d["id"])
    np.testing.assert_array_equal(o["id"], d["id"])
    np.testing.assert_array_equal(o["group"], d["group"])
    np.testing.assert_array_equal(o["klass"], d["klass"])
    np.testing.assert_array_equal(o["n_primary"], d["n_primary"])
    np.testing.assert_array_equal(o["n_comp"], d["n_comp"])
    np.testing.assert_array_equal(o["comp_source"], d["comp_source"])
    np.testing.assert_allclose(d["vol"], o["vol"], rtol=1e-12, atol=0)
    if lists:
        oc, oe = ob.lists()
        gc, ge = g.lists(box, n_class)
        np.testing.assert_array_equal(oc, gc)
        np.testing.assert_array_equal(oe, ge)
    assert g.next_id(box) == ob.next_id
    oi = ob.info()
    gi = g.removed(box)
    for a, b in zip(oi, gi):
        np.testing.assert_array_equal(a, b)
