"""Device numerics vs the oracle: kernel values and the RNG streams."""
import numpy as np
import pytest

import oracle
from partmc_amd import CoagGPU, populations as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev4():
    return CoagGPU(1, P.C2_DENSITY)


def random_pairs(rng, n, n_spec, lo=-27, hi=-14):
    v1 = 10 ** rng.uniform(lo, hi, size=(n, 1)) * rng.dirichlet(np.ones(n_spec), size=n)
    v2 = 10 ** rng.uniform(lo, hi, size=(n, 1)) * rng.dirichlet(np.ones(n_spec), size=n)
    return v1, v2


@pytest.mark.parametrize("kind", [P.KERNEL_ADDITIVE, P.KERNEL_BROWN, P.KERNEL_CONSTANT])
def test_kernel_fp64_matches_oracle(dev4, kind):
    rng = np.random.default_rng(kind)
    v1, v2 = random_pairs(rng, 2000, 4)
    k64, _ = dev4.eval_kernel(kind, v1, v2, 283.0, 9.5e4, 1000.0)
    ref = np.array([oracle.kernel(kind, P.C2_DENSITY, a, b, 283.0, 9.5e4, 1000.0)
                    for a, b in zip(v1, v2)])
    np.testing.assert_allclose(k64, ref, rtol=1e-13)


def test_sedi_fp64_matches_oracle():
    g = CoagGPU(1, [1000.0])
    rng = np.random.default_rng(3)
    r1 = 10 ** rng.uniform(-6, -3.3, size=3000)
    r2 = 10 ** rng.uniform(-6, -3.3, size=3000)
    v1 = (4 * np.pi * r1 ** 3 / 3)[:, None]
    v2 = (4 * np.pi * r2 ** 3 / 3)[:, None]
    k64, k32 = g.eval_kernel(P.KERNEL_SEDI, v1, v2)
    ref = np.array([oracle.kernel(P.KERNEL_SEDI, [1000.0], a, b) for a, b in zip(v1, v2)])
    np.testing.assert_allclose(k64, ref, rtol=1e-12)
    # FP32 pair combine: relative error bounded away from equal sizes
    sep = np.abs(np.log(r1 / r2)) > 0.05
    rel = np.abs(k32[sep] - ref[sep]) / ref[sep]
    assert rel.max() < 1e-4


def test_brownian_fp32_error(dev4):
    rng = np.random.default_rng(11)
    v1, v2 = random_pairs(rng, 4000, 4, lo=-27, hi=-15)
    k64, k32 = dev4.eval_kernel(P.KERNEL_BROWN, v1, v2)
    rel = np.abs(k32.astype(np.float64) - k64) / k64
    assert rel.max() < 2e-6, rel.max()


def test_uniform_stream_matches_oracle(dev4):
    u = dev4.rng_draw(123456789, 3, 7, 11, 13, 2, 0, n=257)
    s = oracle.Stream(123456789, 3, 7, 11, 13, 2)
    ref = np.array([s.uniform() for _ in range(257)])
    np.testing.assert_array_equal(u, ref)


@pytest.mark.parametrize("mean", [0.0, 0.3, 4.0, 10.0, 10.5, 37.0, 1e4])
def test_poisson_matches_oracle(dev4, mean):
    d = dev4.rng_draw(99, 1, 2, 3, 0, 1, 1, p1=mean, n=2000)
    ref = np.array([oracle.Stream(99, 1, 2, 3, i, 1).poisson(mean) for i in range(2000)])
    np.testing.assert_array_equal(d, ref)


@pytest.mark.parametrize("n,p", [(10, 0.3), (30, 0.5), (1000, 0.02), (1000, 0.7)])
def test_binomial_matches_oracle(dev4, n, p):
    d = dev4.rng_draw(5, 0, 1, 2, 0, 6, 2, p1=n, p2=p, n=2000)
    ref = np.array([oracle.Stream(5, 0, 1, 2, i, 6).binomial(n, p) for i in range(2000)])
    np.testing.assert_array_equal(d, ref)
