"""More pins of the CPU oracle (CPU only), each from something the
mathematics or independent data fixes, chosen so that a plausible slip in
the restatement fails:

- the Brownian kernel's delta (Jacobson eq. 15.34) against a quadrature of
  its geometric definition, the particle diffusivity against the published
  table of Seinfeld & Pandis (2006, Table 9.5), the air mean free path
  against the ISO 15900 standard value, and the whole transition-regime
  kernel against Dahneke's (1983) independent interpolation;
- the weighted flow, f_max and the per-particle path under power weights,
  through the analytic Golovin solution, which does not depend on the
  weighting, at the reference test's tolerances (test/additive: 0.05 in
  number, 0.2 in mass, relative 2-norm at every output time);
- per-set sampling against the bidisperse ODE of the reference's
  test/bidisperse (tolerance 0.3);
- est_k_minmax's sample points and over-scale for the additive kernel in
  closed form, and its bound for the Brownian kernel;
- rand_group's distribution (chi-square) and component resampling's
  invariants and subset frequencies.
"""
import math
import multiprocessing as mp
import os

import numpy as np
import pytest
from scipy import stats
from scipy.integrate import solve_ivp

import oracle
from partmc_amd import populations as P


# ------------------------------------------------------------ Brownian
def mean_hemisphere_distance(R, lam, n=200):
    """Mean distance from the centre of a sphere of radius R reached by a
    point leaving its surface outwards (directions uniform over the outward
    hemisphere) and travelling lam, by Gauss-Legendre quadrature over the
    polar angle and the trapezoid rule in azimuth of |R n + lam u|."""
    x, wx = np.polynomial.legendre.leggauss(n)
    theta = (x + 1.0) * (math.pi / 4.0)          # [0, pi/2]
    wt = wx * (math.pi / 4.0)
    phi = np.linspace(0.0, 2.0 * math.pi, 64, endpoint=False)
    th, ph = np.meshgrid(theta, phi, indexing="ij")
    u = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=-1)
    pos = np.array([0.0, 0.0, R]) + lam * u
    d = np.linalg.norm(pos, axis=-1)
    dA = np.sin(th) * wt[:, None] * (2.0 * math.pi / len(phi))
    return float((d * dA).sum() / dA.sum())


@pytest.mark.parametrize("diam", [2e-9, 1e-8, 5e-8, 2e-7, 1e-6])
def test_brownian_delta_is_the_mean_distance_reached_from_the_surface(diam):
    """Jacobson eq. 15.34 (doc/coagulation_kernels/brownian.tex:59): delta
    is the mean distance from the centre of a sphere of radius 2 r reached
    by particles leaving its surface and travelling their mean free path,
    less 2 r."""
    t = oracle.brown_terms(math.pi / 6 * diam ** 3, 1500.0, 288.0, 1e5)
    R, lam = 2.0 * t["radius"], t["mfp"]
    delta = mean_hemisphere_distance(R, lam) - R
    assert math.sqrt(t["delta_sq"]) == pytest.approx(delta, rel=1e-7)


def test_air_mean_free_path_standard_value_and_pressure_scaling():
    """ISO 15900: 67.3 nm at 296.15 K and 101.325 kPa; inversely
    proportional to pressure at fixed temperature."""
    v = math.pi / 6 * 1e-21
    lam = oracle.brown_terms(v, 1000.0, 296.15, 101325.0)["air_mfp"]
    assert lam == pytest.approx(67.3e-9, rel=0.03)
    lam2 = oracle.brown_terms(v, 1000.0, 296.15, 2 * 101325.0)["air_mfp"]
    assert lam2 == pytest.approx(lam / 2, rel=1e-12)


@pytest.mark.parametrize("diam,diff_cm2", [(1e-9, 5.14e-2), (1e-8, 5.24e-4), (1e-7, 6.82e-6),
                                           (1e-6, 2.74e-7), (1e-5, 2.38e-8)])
def test_particle_diffusivity_against_published_table(diam, diff_cm2):
    """Seinfeld & Pandis (2006), Table 9.5: particle diffusion coefficients
    in air at 298 K and 1 atm.  Their slip correction and mean free path
    differ from Jacobson's by up to ~4% in the transition regime."""
    t = oracle.brown_terms(math.pi / 6 * diam ** 3, 1000.0, 298.15, 101325.0)
    assert t["diffusivity"] == pytest.approx(diff_cm2 * 1e-4, rel=0.05)


def dahneke(a, b):
    """Dahneke's (1983) transition-regime kernel from the continuum kernel
    and the diffusion Knudsen number."""
    rs = a["radius"] + b["radius"]
    ds = a["diffusivity"] + b["diffusivity"]
    c = math.sqrt(a["speed_sq"] + b["speed_sq"])
    kn = 2.0 * ds / (c * rs)
    return 4.0 * math.pi * rs * ds * (1.0 + kn) / (1.0 + 2.0 * kn + 2.0 * kn * kn)


def test_brownian_transition_regime_against_dahneke():
    """Fuchs' interpolation (the reference's) and Dahneke's agree within
    4.5% from the free-molecular to the continuum regime; the pairs span
    diffusion Knudsen numbers 1e-2 to 1e2.  A factor of 2 either way in
    delta^2 moves the worst pair to 6-10%."""
    T, p, rho = 288.0, 1e5, 1000.0
    ds = np.logspace(-9, -5, 17)
    worst = 0.0
    for d1 in ds:
        for d2 in ds:
            v1, v2 = math.pi / 6 * d1 ** 3, math.pi / 6 * d2 ** 3
            k = oracle.kernel(P.KERNEL_BROWN, [rho], [v1], [v2], T, p)
            a = oracle.brown_terms(v1, rho, T, p)
            b = oracle.brown_terms(v2, rho, T, p)
            worst = max(worst, abs(k / dahneke(a, b) - 1.0))
    assert worst < 0.045, worst


# ------------------------------------------------------- est_k_minmax
def test_kbounds_additive_closed_form():
    """3 x 3 samples per bin pair that include the bin edges, k_max scaled
    by 2 (src/coag_kernel.F90:296-336): for the additive kernel
    k_min = beta (v1_lo + v2_lo) and k_max = 2 beta (v1_hi + v2_hi)."""
    e, _, _ = oracle.bin_grid_make_log(40, 1e-8, 1e-4)
    kmin, kmax = oracle.est_k_minmax(P.KERNEL_ADDITIVE, [1000.0], e, beta=700.0)
    v = 4 * math.pi / 3 * e ** 3
    lo = v[:-1][:, None] + v[:-1][None, :]
    hi = v[1:][:, None] + v[1:][None, :]
    np.testing.assert_allclose(kmin, 700.0 * lo, rtol=1e-13)
    np.testing.assert_allclose(kmax, 2 * 700.0 * hi, rtol=1e-13)


def test_kbounds_brownian_bound_inside_each_bin_pair():
    """k_max bounds the Brownian kernel of any pair of particles inside the
    two bins with densities between the species' extremes, and k_min is
    below it."""
    dens = [1000.0, 1800.0]
    e, _, _ = oracle.bin_grid_make_log(24, 1e-9, 1e-5)
    kmin, kmax = oracle.est_k_minmax(P.KERNEL_BROWN, dens, e, 270.0, 8e4)
    rng = np.random.default_rng(11)
    for b1 in range(0, 24, 3):
        for b2 in range(b1, 24, 2):
            for _ in range(12):
                r1 = math.exp(rng.uniform(math.log(e[b1]), math.log(e[b1 + 1])))
                r2 = math.exp(rng.uniform(math.log(e[b2]), math.log(e[b2 + 1])))
                f1, f2 = rng.uniform(0, 1, 2)
                v1 = 4 * math.pi / 3 * r1 ** 3 * np.array([f1, 1 - f1])
                v2 = 4 * math.pi / 3 * r2 ** 3 * np.array([f2, 1 - f2])
                k = oracle.kernel(P.KERNEL_BROWN, dens, v1, v2, 270.0, 8e4)
                assert k <= kmax[b1, b2] and kmin[b1, b2] <= kmax[b1, b2] / 2


# ------------------------------------------------ weighting functions
def test_num_conc_factor_closed_form_and_fmax_bound():
    """coag_num_conc_factor (src/coag_kernel.F90:346-377) with one power
    weight of exponent -3: nc(r) = m r^-3 falls with r, so the merged
    particle has the least and f = nc_i nc_j / nc_ij =
    m (4 pi / 3) (v_i + v_j) / (v_i v_j).  max_coag_num_conc_factor bounds f
    for radii inside the bins."""
    m = 3.0e-9
    w = oracle.Weights([[P.WEIGHT_POWER]], [[m]], [[-3.0]])
    rng = np.random.default_rng(2)
    for _ in range(50):
        ri, rj = 10 ** rng.uniform(-8, -5, 2)
        vi, vj = 4 * math.pi / 3 * ri ** 3, 4 * math.pi / 3 * rj ** 3
        f = w.coag_num_conc_factor(ri, rj, 1, 1, 1)
        assert f == pytest.approx(m * 4 * math.pi / 3 * (vi + vj) / (vi * vj), rel=1e-12)
    e, _, _ = oracle.bin_grid_make_log(20, 1e-8, 1e-5)
    for wt in (w, oracle.Weights([[P.WEIGHT_POWER], [P.WEIGHT_POWER]], [[1e-6], [4e-3]],
                                 [[-1.5], [-2.5]])):
        for ib in range(0, 20, 3):
            for jb in range(ib, 20, 4):
                fmax = wt.max_factor(e, ib + 1, jb + 1, 1, 1, 1)
                for _ in range(10):
                    ri = math.exp(rng.uniform(math.log(e[ib]), math.log(e[ib + 1])))
                    rj = math.exp(rng.uniform(math.log(e[jb]), math.log(e[jb + 1])))
                    assert wt.coag_num_conc_factor(ri, rj, 1, 1, 1) <= fmax * (1 + 1e-12)


# ------------------------------------------------ Golovin, power weights
N0, D_MEAN, BETA = 1e9, 2e-5, 1000.0
V_MEAN = math.pi / 6 * D_MEAN ** 3
EDGES, CENTERS, WIDTHS = oracle.bin_grid_make_log(160, 0.5e-8, 0.5e-3)  # diameters 1e-8..1e-3


def golovin_run(args):
    """One run of test/additive's case (exponential initial condition at a
    20 um mean-volume diameter, N0 = 1e9 m^-3, beta = 1000 s^-1, 600 s) with
    power weighting of exponent alpha and doubling/halving to n_c
    particles every step.  The computational particles sample
    n(v) / nc(v): for nc = m r^alpha that is a Gamma(1 - alpha/3, v_mean)
    law, and m makes the expected number concentration N0.  Returns the
    number and mass distributions per ln D on 160 bins at every 60 s, and
    the bin pairs that took the per-particle path."""
    alpha, n_c, seed, dt = args
    rng = np.random.default_rng(seed)
    k = 1.0 - alpha / 3.0
    v = rng.gamma(k, V_MEAN, size=n_c)
    c_r = (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)
    mag = N0 * math.gamma(k) / (n_c * c_r ** alpha * V_MEAN ** (alpha / 3.0))
    w = oracle.Weights([[P.WEIGHT_POWER]], [[mag]], [[alpha]])
    b = oracle.Box([1000.0], w, 288.0, 1e5, BETA, seed=seed)
    b.add_particles(v[:, None], np.arange(1, n_c + 1), np.ones(n_c), np.ones(n_c))
    b.next_id = n_c + 1
    out = []

    def hist():
        d = b.particles()
        vv = d["vol"].sum(axis=1)
        r = (3.0 * vv / (4.0 * math.pi)) ** (1.0 / 3.0)
        nc = b.weight_magnitudes()[0, 0] * r ** alpha
        bi = np.searchsorted(EDGES, r, side="right") - 1
        ok = (bi >= 0) & (bi < 160)
        hn, hm = np.zeros(160), np.zeros(160)
        np.add.at(hn, bi[ok], nc[ok] / WIDTHS[bi[ok]])
        np.add.at(hm, bi[ok], nc[ok] * vv[ok] * 1000.0 / WIDTHS[bi[ok]])
        return hn, hm

    out.append(hist())
    per_out = int(round(60.0 / dt))
    for s in range(int(round(600.0 / dt))):
        b.mc_coag(P.KERNEL_ADDITIVE, dt, s)
        b.rebalance(n_c, s)
        if (s + 1) % per_out == 0:
            out.append(hist())
    return np.array(out), b.n_accel


def golovin_ensemble(alpha, n_c, n_rep, dt):
    args = [(alpha, n_c, 5000 + i, dt) for i in range(n_rep)]
    procs = min(8, os.cpu_count() or 1)
    with mp.get_context("fork").Pool(procs) as pool:
        runs = pool.map(golovin_run, args)
    hist = np.mean([r[0] for r in runs], axis=0)  # [time][num/mass][bin]
    return hist, sum(r[1] for r in runs)


@pytest.mark.slow
@pytest.mark.parametrize("alpha,n_c,n_rep,dt,min_accel", [
    (-3.0, 10000, 192, 1.0, 1000),    # the reference test's step
    (-3.0, 10000, 192, 5.0, 20000),   # many bin pairs on the per-particle path
    (-1.5, 20000, 64, 1.0, 0),        # fractional exponent: general pow
])
def test_golovin_under_power_weights(alpha, n_c, n_rep, dt, min_accel):
    """The analytic solution does not depend on the weighting, so the mean
    distributions under power weights must match it at test/additive's
    tolerances at every output time."""
    hist, n_accel = golovin_ensemble(alpha, n_c, n_rep, dt)
    assert n_accel >= min_accel
    for j in range(11):
        nr, vr = np.zeros(160), np.zeros(160)
        oracle.lib().orc_soln_additive_exp(160, oracle._ptr(CENTERS), 60.0 * j, N0,
                                           D_MEAN / 2, BETA, oracle._ptr(nr), oracle._ptr(vr))
        mr = vr * 1000.0
        err_n = np.linalg.norm(hist[j, 0] - nr) / np.linalg.norm(nr)
        err_m = np.linalg.norm(hist[j, 1] - mr) / np.linalg.norm(mr)
        assert err_n < 0.05 and err_m < 0.2, (60 * j, err_n, err_m)


# ---------------------------------------------------------- bidisperse
def test_bidisperse_matches_the_ode():
    """test/bidisperse: 10^4 water drops of 10 um radius and one of 100 um
    in 1e-5 m^3, sedimentation kernel, 600 s at dt = 1 s.  Small drops do
    not collect each other (equal fall speeds), so the small-drop number
    follows dn/dt = -K(v_small, v_big(n)) n / V with
    v_big = v_big0 + (n0 - n) v_small (test/bidisperse/test_bidisperse_ode.F90).
    The reference's tolerance is 0.3 (relative 2-norm per column); the
    oracle meets 0.05."""
    r_s, r_b, n0 = 1e-5, 1e-4, 10000
    comp_vol = (n0 + 1) / (1e9 * (n0 + 1) / n0)
    v_s, v_b0 = 4 * math.pi / 3 * r_s ** 3, 4 * math.pi / 3 * r_b ** 3

    def rhs(_t, y):
        vb = v_b0 + (n0 - y[0]) * v_s
        return [-oracle.kernel(P.KERNEL_SEDI, [1000.0], [v_s], [vb]) / comp_vol * y[0]]

    ts = np.arange(0.0, 601.0, 10.0)
    n_ode = solve_ivp(rhs, (0.0, 600.0), [float(n0)], t_eval=ts, rtol=1e-10, atol=1e-8).y[0]
    m_ode = (v_b0 + (n0 - n_ode) * v_s) * 1000.0 / comp_vol
    w = oracle.Weights([[P.WEIGHT_NONE]], [[1.0 / comp_vol]], [[0.0]])
    b = oracle.Box([1000.0], w, 288.0, 1e5, 1000.0, seed=17)
    vol = np.array([v_s] * n0 + [v_b0])[:, None]
    b.add_particles(vol, np.arange(1, n0 + 2), np.ones(n0 + 1), np.ones(n0 + 1))
    b.next_id = n0 + 2
    n_small, m_big = [n0], [v_b0 * 1000.0]
    for s in range(600):
        b.mc_coag(P.KERNEL_SEDI, 1.0, s)
        if (s + 1) % 10 == 0:
            v = b.particles()["vol"][:, 0]
            assert np.sum(v > 1.5 * v_s) == 1  # one large drop
            n_small.append(int(np.sum(v < 1.5 * v_s)))
            m_big.append(v.max() * 1000.0)
    n_mc = np.array(n_small) / comp_vol
    m_mc = np.array(m_big) / comp_vol
    for got, ref in ((n_mc, n_ode / comp_vol), (m_mc, m_ode)):
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 0.3 and err < 0.05, err


# ---------------------------------------------------------- rand_group
def test_rand_group_distribution_chi_square():
    """aero_weight_array_rand_group picks group k with probability
    proportional to 1 / nc_k(r) (src/aero_weight_array.F90:374-393) by
    sample_cts_pdf's accept-reject (src/rand.F90:550-576)."""
    wt = oracle.Weights([[P.WEIGHT_POWER], [P.WEIGHT_POWER], [P.WEIGHT_NONE]],
                        [[200.0], [1.79e-8], [1e9]], [[-1.0], [-2.5], [0.0]])
    b = oracle.Box([1000.0], wt)
    n = 30000
    for r in (1e-7, 2e-7, 4e-7):
        inv = np.array([1.0 / (200.0 * r ** -1.0), 1.0 / (1.79e-8 * r ** -2.5), 1.0 / 1e9])
        p = inv / inv.sum()
        assert p.min() > 0.02
        counts = np.bincount([b.rand_group(77, i, 1, r) for i in range(n)], minlength=4)[1:]
        chi2 = ((counts - n * p) ** 2 / (n * p)).sum()
        assert chi2 < stats.chi2.ppf(0.999, 2), (r, counts, n * p)
    two = oracle.Box([1000.0], oracle.Weights([[P.WEIGHT_NONE], [P.WEIGHT_NONE]],
                                              [[1.0], [3.0]], [[0.0], [0.0]]))
    counts = np.bincount([two.rand_group(78, i, 1, 1e-7) for i in range(n)], minlength=3)[1:]
    p = np.array([0.75, 0.25])
    assert ((counts - n * p) ** 2 / (n * p)).sum() < stats.chi2.ppf(0.999, 1)


# ------------------------------------------------- component resampling
def test_component_concatenation_without_resampling():
    """At most 20 components in all: concatenated in order, no draws."""
    b = oracle.Box([1000.0], oracle.Weights([[P.WEIGHT_NONE]], [[1.0]], [[0.0]]))
    out = b.coagulate_components(1, 0, [1, 2, 3], 5, list(range(101, 118)), 30)
    np.testing.assert_array_equal(out, [1, 2, 3] + list(range(101, 118)))


def test_component_resampling_invariants_and_frequencies():
    """Above 20 components (src/aero_particle.F90:973-993): prob_round
    gives the first particle round(20 np1 / (np1 + np2)) slots, each
    particle's kept components are a uniformly random subset in their
    original order (sample_without_replacement, src/rand.F90:515-542), and
    the merged particle has exactly 20."""
    b = oracle.Box([1000.0], oracle.Weights([[P.WEIGHT_NONE]], [[1.0]], [[0.0]]))
    src1 = list(range(1, 7))           # 6 components, 6 primary parts: 20 * 6 / 40 = 3 slots
    src2 = list(range(101, 118))       # 17 components, 34 primary parts: the other 17 slots
    n = 40000
    subsets = {}
    for i in range(n):
        out = b.coagulate_components(9, i, src1, 6, src2, 34)
        assert len(out) == 20
        first, second = out[:3], out[3:]
        np.testing.assert_array_equal(second, src2)       # all 17 fit
        assert all(1 <= x <= 6 for x in first) and list(first) == sorted(set(first))
        subsets[tuple(first)] = subsets.get(tuple(first), 0) + 1
    assert len(subsets) == math.comb(6, 3)
    counts = np.array(list(subsets.values()))
    exp = n / math.comb(6, 3)
    assert ((counts - exp) ** 2 / exp).sum() < stats.chi2.ppf(0.999, 19)
    # prob_round of a fractional share: 20 * 14 / 32 = 8.75 slots for the
    # first particle on average, never fewer than 8 or more than 9
    k1 = []
    for i in range(8000):
        out = b.coagulate_components(10, i, list(range(1, 13)), 14, list(range(101, 113)), 18)
        assert len(out) == 20
        k = int(np.sum(out < 100))
        assert k in (8, 9)
        assert list(out[:k]) == sorted(set(out[:k])) and list(out[k:]) == sorted(set(out[k:]))
        k1.append(k)
    p9 = np.mean(np.array(k1) == 9)
    assert abs(p9 - 0.75) < 4 * math.sqrt(0.75 * 0.25 / 8000)
