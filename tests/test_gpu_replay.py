"""Record and replay of a box's random draws (exact mode) against the oracle.

A trace is the sequence of uniforms the sequential chain draws, in call
order: what pmc_random returns in a reference run without GSL
(src/rand.F90:138-166), where every other draw is made from it.

- The GPU records its draws, and the oracle records its own: the two traces
  are equal value for value, so both consume draws in the same order.
- The GPU replays the oracle's recorded trace and reaches the oracle's
  state; it replays a trace from an independent generator (numpy), with
  the reference's rounded-normal samplers, and reaches the state the oracle
  reaches on the same trace, which differs from the state of the call-site
  streams.
"""
import numpy as np
import pytest

from helpers import compare_state, gpu_boxes, oracle_box, small
from partmc_amd import MODE_PARALLEL, PREC_FP64, CoagError
from partmc_amd import populations as P

pytestmark = pytest.mark.gpu

CASES = {
    "additive": (lambda: P.config_c1prime(n_part=3000), P.KERNEL_ADDITIVE, 1.0, 40),
    "brownian-power": (lambda: small(P.config_c2(), 3000), P.KERNEL_BROWN, 60.0, 4),
    "sedi": (lambda: small(P.config_c5(n_box=1, n_part=20000)[0], 5000), P.KERNEL_SEDI, 1.0, 20),
}


def boost(pop, name):
    if name == "sedi":
        pop.weight_magnitude *= 100.0
    if name == "brownian-power":
        pop.weight_magnitude *= 30.0
    return pop


@pytest.mark.parametrize("name", sorted(CASES))
def test_recorded_traces_equal(name):
    make, kernel, dt, steps = CASES[name]
    pop = boost(make(), name)
    ob = oracle_box(pop, 41)
    g = gpu_boxes([pop], 41)
    cap = 4_000_000
    ob.record(cap)
    g.record(0, cap)
    for s in range(steps):
        os_, oc = ob.mc_coag(kernel, dt, s)
        gs, gc = g.mc_coag(kernel, dt)
        assert (os_, oc) == (int(gs[0]), int(gc[0])), s
    n_o, u_o, ro = ob.trace(cap)
    n_g, u_g, rg = g.trace(0, cap)
    print(f"{name}: {n_g} uniforms recorded", flush=True)
    assert not ro and not rg and n_o == n_g > 100
    np.testing.assert_array_equal(u_g, u_o)
    compare_state(ob, g)


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_replays_oracle_trace(name):
    make, kernel, dt, steps = CASES[name]
    pop = boost(make(), name)
    ob = oracle_box(pop, 43)
    cap = 4_000_000
    ob.record(cap)
    for s in range(steps):
        ob.mc_coag(kernel, dt, s)
    n, u, _ = ob.trace(cap)
    # a fresh oracle box replays its own recording to the same state
    oc = oracle_box(pop, 99)
    oc.replay(u)
    g = gpu_boxes([pop], 99)  # another seed: every draw must come from the trace
    g.replay(0, u)
    for s in range(steps):
        a = oc.mc_coag(kernel, dt, s)
        b = g.mc_coag(kernel, dt)
        assert a == (int(b[0][0]), int(b[1][0])), s
    assert g.trace(0)[0] == n == oc.trace()[0]
    compare_state(ob, g)
    compare_state(oc, g)


@pytest.mark.parametrize("name", sorted(CASES))
def test_independent_trace_with_reference_samplers(name):
    make, kernel, dt, steps = CASES[name]
    pop = boost(make(), name)
    u = np.random.default_rng(2026).random(3_000_000)
    ob = oracle_box(pop, 5)
    ob.replay(u, ref_samplers=True)
    g = gpu_boxes([pop], 5)
    g.replay(0, u, ref_samplers=True)
    tot = 0
    for s in range(steps):
        a = ob.mc_coag(kernel, dt, s)
        b = g.mc_coag(kernel, dt)
        assert a == (int(b[0][0]), int(b[1][0])), s
        tot += a[1]
    assert tot > 0
    compare_state(ob, g)
    # the call-site streams give another chain
    h = gpu_boxes([pop], 5)
    for _ in range(steps):
        h.mc_coag(kernel, dt)
    assert not np.array_equal(h.download(0)["id"], g.download(0)["id"])


def test_replay_errors():
    pop = P.config_c1prime(n_part=2000)
    g = gpu_boxes([pop], 3)
    g.replay(0, np.full(10, 0.5))
    with pytest.raises(CoagError, match="ran out"):
        for _ in range(50):
            g.mc_coag(P.KERNEL_ADDITIVE, 1.0)
    with pytest.raises(CoagError):
        g.replay(0, np.array([0.3, 1.5]))
    h = gpu_boxes([pop], 3)
    h.record(0, 1000)
    h.set_mode(MODE_PARALLEL, PREC_FP64, 1e-5)
    with pytest.raises(CoagError, match="exact mode"):
        h.mc_coag(P.KERNEL_ADDITIVE, 1.0)
    h.trace_off(0)
    h.mc_coag(P.KERNEL_ADDITIVE, 1.0)
