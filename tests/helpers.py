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
