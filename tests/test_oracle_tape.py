"""The oracle's tape: its uniforms in call order, recorded or replayed
(CPU only).  A box that replays its own recording reaches the same state;
a trace from another generator gives another chain, the same one each time;
the reference's rounded-normal Poisson sampler (src/rand.F90:298-302) on a
tape follows Box-Muller as numpy computes it."""
import math

import numpy as np

from helpers import oracle_box, small
from partmc_amd import populations as P


def state(ob):
    d = ob.particles()
    return d["id"].copy(), d["vol"].copy()


def test_replay_of_own_recording_is_identical():
    pop = small(P.config_c2(), 3000)
    pop.weight_magnitude *= 30.0
    a = oracle_box(pop, 8)
    a.record(2_000_000)
    ca = [a.mc_coag(P.KERNEL_BROWN, 60.0, s) for s in range(4)]
    n, u, ran_out = a.trace(2_000_000)
    assert not ran_out and n == len(u) > 1000
    b = oracle_box(pop, 77)  # another seed: every draw comes from the tape
    b.replay(u)
    cb = [b.mc_coag(P.KERNEL_BROWN, 60.0, s) for s in range(4)]
    assert ca == cb and b.trace()[0] == n
    for x, y in zip(state(a), state(b)):
        np.testing.assert_array_equal(x, y)


def test_independent_trace_gives_a_reproducible_other_chain():
    pop = P.config_c1prime(n_part=3000)
    u = np.random.default_rng(3).random(500_000)
    runs = []
    for _ in range(2):
        b = oracle_box(pop, 8)
        b.replay(u, ref_samplers=True)
        for s in range(30):
            b.mc_coag(P.KERNEL_ADDITIVE, 1.0, s)
        runs.append(state(b))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    c = oracle_box(pop, 8)
    for s in range(30):
        c.mc_coag(P.KERNEL_ADDITIVE, 1.0, s)
    assert not np.array_equal(state(c)[0], runs[0][0])


def test_reference_poisson_sampler_on_a_tape():
    """A box of two equal particles has one non-empty bin pair, whose count
    is the step's first draw: with a mean above 10 the reference's sampler
    gives nint(mean + sqrt(mean) sqrt(-2 ln u1) cos(2 pi u2))."""
    r = 5.5e-5
    pop = P.config_c1prime(n_part=2)
    pop.vol[:] = 4 * math.pi / 3 * r ** 3
    pop.weight_magnitude = 1e10
    u = np.random.default_rng(4).random(100_000)
    d = oracle_box(pop, 1)
    d.mc_coag(P.KERNEL_ADDITIVE, 0.0, 0)  # the sort and its bounds, no draws
    e, _, _ = d.grid()
    d.replay(u, ref_samplers=True)
    n_samp, _ = d.mc_coag(P.KERNEL_ADDITIVE, 1.0, 1)
    b = int(np.searchsorted(e, r, side="right")) - 1
    _, kmax = d.kmax()
    mean = kmax[b, b] * pop.weight_magnitude  # flat: f = nc; n_possible = 1; dt = 1
    assert mean > 10.0
    z = math.sqrt(-2.0 * math.log(u[0])) * math.cos(2.0 * math.pi * u[1])
    assert n_samp == max(int(math.floor(mean + math.sqrt(mean) * z + 0.5)), 0)
