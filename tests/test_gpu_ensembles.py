"""Free-running ensembles: the parallel mode against the sequential chain.

The exact mode reproduces the oracle's sequential chain bit for bit
(tests/test_gpu_exact.py), so its ensemble stands for the reference's.  Over
100 independent runs of each mode (their own seeds), the mean particle
count must agree within its 99% confidence interval, and the mean number
and volume size distributions within theirs: the per-bin differences of the
means, in units of their standard errors, must pass a joint chi-square
test at the 99% level.  The file makes 15 such comparisons; at 95% each, a
correct parallel mode would fail one of them in about half of the runs
whenever a change reshuffles its chain, at 99% in about one run in seven.

Cases: a C5-type sedimentation box (flat weights, cloud and drizzle drops)
at its own concentration and at 30 times it, where one drizzle drop
collects about 170 cloud drops per step and grows through several bins
within a step; and a C2-type Brownian box (trimodal, power weights of
exponent -3, per-particle path) at 1, 10 and 30 times its concentration.
"""
import numpy as np
import pytest
from scipy import stats

from helpers import gpu_boxes, radius, small
from partmc_amd import MODE_EXACT, MODE_PARALLEL, PREC_FP32_RECHECK, PREC_FP64
from partmc_amd import populations as P

pytestmark = pytest.mark.gpu

N_REP = 100


def distributions(vol, pop, edges):
    """Number and volume concentration per bin of radius (the weight
    num_conc = magnitude * r**exponent, src/aero_weight.F90:127-146)."""
    r = radius(vol)
    nc = pop.weight_magnitude * r ** pop.weight_exponent
    v = vol.sum(axis=1)
    b = np.searchsorted(edges, r, side="right") - 1
    ok = (b >= 0) & (b < len(edges) - 1)
    hn = np.bincount(b[ok], weights=nc[ok], minlength=len(edges) - 1)
    hv = np.bincount(b[ok], weights=(nc * v)[ok], minlength=len(edges) - 1)
    return hn, hv


def ensemble(base, mode, precision, seed, kernel, dt, steps, edges):
    g = gpu_boxes([base] * N_REP, seed)
    g.set_mode(mode, precision, 1e-5)
    for _ in range(steps):
        g.mc_coag(kernel, dt)
    count = np.array([g.n_part(k) for k in range(N_REP)], dtype=float)
    h = [distributions(g.download(k)["vol"], base, edges) for k in range(N_REP)]
    extra = {}
    if mode == MODE_PARALLEL:
        extra = g.arrival_stats()
    g.close()
    return count, np.array([x[0] for x in h]), np.array([x[1] for x in h]), extra


def z_scores(a, b):
    d = a.mean(axis=0) - b.mean(axis=0)
    se = np.sqrt(a.var(axis=0, ddof=1) / len(a) + b.var(axis=0, ddof=1) / len(b))
    return d, se


def assert_agree(name, par, seq):
    (cp, np_, vp, _), (cs, ns, vs, _) = par, seq
    d, se = z_scores(cp, cs)
    print(f"{name}: particles {cp.mean():.1f} vs {cs.mean():.1f} (diff {d:.1f}, se {se:.1f})",
          flush=True)
    assert abs(d) <= stats.norm.ppf(0.995) * se, (name, "count", d, se)
    for label, a, b in (("number", np_, ns), ("volume", vp, vs)):
        d, se = z_scores(a, b)
        use = se > 0
        z = d[use] / se[use]
        chi2 = float(np.sum(z ** 2))
        lim = stats.chi2.ppf(0.99, use.sum())
        print(f"  {label}: {use.sum()} bins, chi2 {chi2:.1f} (99% limit {lim:.1f}), "
              f"max |z| {np.abs(z).max():.2f}", flush=True)
        assert chi2 <= lim, (name, label, chi2, lim, z)


@pytest.mark.parametrize("factor", [1.0, 30.0])
def test_sedimentation_ensembles_agree(factor):
    base = P.config_c5(n_box=1, n_part=20000)[0]
    base.weight_magnitude *= factor
    edges = np.logspace(np.log10(0.5e-6), np.log10(2.5e-3), 61)
    args = (P.KERNEL_SEDI, 1.0, 30, edges)
    par = ensemble(base, MODE_PARALLEL, PREC_FP32_RECHECK, 4321, *args)
    seq = ensemble(base, MODE_EXACT, PREC_FP64, 1234, *args)
    assert par[3]["n_arrivals"].sum() > 0
    assert_agree(f"sedimentation x{factor:g}", par, seq)


@pytest.mark.parametrize("factor", [1.0, 10.0, 30.0])
def test_power_weighted_brownian_ensembles_agree(factor):
    base = small(P.config_c2(), 6000)
    base.weight_magnitude *= factor
    edges = np.logspace(np.log10(0.5e-9), np.log10(0.5e-4), 51)
    args = (P.KERNEL_BROWN, 60.0, 10, edges)
    par = ensemble(base, MODE_PARALLEL, PREC_FP32_RECHECK, 4321, *args)
    seq = ensemble(base, MODE_EXACT, PREC_FP64, 1234, *args)
    assert_agree(f"power-weighted Brownian x{factor:g}", par, seq)
