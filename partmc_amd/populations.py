"""Synthetic particle populations for the coagulation workloads.

The sampling laws follow the reference's initial-distribution code:
exponential modes draw ``v = -v_mean * ln(U)`` (``src/aero_mode.F90:609-612``);
log-normal modes under power weighting shift the log10-median by
``-alpha * sigma10**2 * ln 10`` and scale the computational-particle count by
``exp((x'^2 - x^2) / (2 sigma10^2))`` (``src/aero_mode.F90:496-544,569-578``).
Radii are geometric radii in m; weights are ``w(r) = magnitude * r**alpha``
(``src/aero_weight.F90:127-146``).  Species volume fractions are drawn from a
Dirichlet distribution (the workload definition; the reference has no such
mode).  Everything is deterministic in ``seed``.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

WEIGHT_NONE, WEIGHT_POWER = 1, 2
KERNEL_SEDI, KERNEL_ADDITIVE, KERNEL_CONSTANT, KERNEL_BROWN, KERNEL_ZERO = 1, 2, 3, 4, 5


@dataclasses.dataclass
class Population:
    """One box: particles plus the box's environment and weighting."""

    vol: np.ndarray              # (n_part, n_spec) float64, m^3
    density: np.ndarray          # (n_spec,) kg/m^3
    weight_type: int
    weight_magnitude: float
    weight_exponent: float
    temp: float = 288.0
    pressure: float = 1e5
    beta: float = 1000.0         # additive kernel coefficient (s^-1)
    source: Optional[np.ndarray] = None   # (n_part,) int32 mode index + 1
    grid: Optional[tuple] = None  # (n_bin, r_min, r_max) fixed sort grid

    @property
    def n_part(self) -> int:
        return int(self.vol.shape[0])

    @property
    def n_spec(self) -> int:
        return int(self.vol.shape[1])

    def ids(self, start: int = 1) -> np.ndarray:
        return np.arange(start, start + self.n_part, dtype=np.int64)


def sphere_rad2vol(r):
    return 4.0 * math.pi * np.asarray(r) ** 3 / 3.0


def exp_mode(rng: np.random.Generator, n_part: int, num_conc: float,
             diam_at_mean_vol: float, n_spec: int = 1,
             fractions: Optional[np.ndarray] = None) -> tuple:
    """Exponential mode with flat weighting (src/aero_mode.F90:609-612)."""
    v_mean = sphere_rad2vol(diam_at_mean_vol / 2.0)
    u = 1.0 - rng.random(n_part)           # in (0, 1]
    vtot = -v_mean * np.log(u)
    if fractions is None:
        fractions = np.full((n_part, n_spec), 1.0 / n_spec)
    vol = vtot[:, None] * fractions
    magnitude = num_conc / n_part
    return vol, magnitude


def lognormal_modes(rng: np.random.Generator, n_part: int,
                    modes: Sequence[tuple], alpha: float,
                    alphas_dirichlet: Sequence[float]) -> tuple:
    """Sample ``n_part`` computational particles from log-normal modes.

    ``modes`` holds (num_conc m^-3, geometric-mean diameter m, sigma_g).
    Weighting is ``magnitude * r**alpha`` (alpha = 0 is flat).  Returns
    (vol, source, magnitude).
    """
    xs, s10s, numbers = [], [], []
    for (nc, dg, sg) in modes:
        x = math.log10(dg / 2.0)
        s10 = math.log10(sg)
        xp = x - alpha * s10 * s10 * math.log(10.0)
        # aero_mode_number with magnitude 1
        num = nc * math.exp((xp * xp - x * x) / (2.0 * s10 * s10))
        xs.append(xp)
        s10s.append(s10)
        numbers.append(num)
    total = sum(numbers)
    magnitude = total / n_part
    counts = [int(round(n / magnitude)) for n in numbers]
    counts[int(np.argmax(counts))] += n_part - sum(counts)
    vols, srcs = [], []
    n_spec = len(alphas_dirichlet)
    for m, (xp, s10, c) in enumerate(zip(xs, s10s, counts)):
        if c <= 0:
            continue
        r = 10.0 ** rng.normal(xp, s10, size=c)
        vtot = sphere_rad2vol(r)
        frac = rng.dirichlet(alphas_dirichlet, size=c) if n_spec > 1 else np.ones((c, 1))
        vols.append(vtot[:, None] * frac)
        srcs.append(np.full(c, m + 1, dtype=np.int32))
    vol = np.concatenate(vols, axis=0)
    src = np.concatenate(srcs)
    perm = rng.permutation(vol.shape[0])
    return vol[perm], src[perm], magnitude


C2_DENSITY = np.array([1800.0, 1400.0, 1800.0, 1000.0])


def config_c1(seed: int = 20261203, n_part: int = 10_000,
              diam_at_mean_vol: float = 2e-6) -> Population:
    """Additive (Golovin) kernel, exponential mode, flat weighting."""
    rng = np.random.default_rng(seed)
    vol, mag = exp_mode(rng, n_part, 1e9, diam_at_mean_vol)
    return Population(vol=vol, density=np.array([1000.0]),
                      weight_type=WEIGHT_NONE, weight_magnitude=mag,
                      weight_exponent=0.0, temp=288.0, pressure=1e5,
                      beta=1000.0, source=np.ones(n_part, dtype=np.int32),
                      grid=(100, 0.5e-8, 0.5e-3))


def config_c1prime(seed: int = 20261203, n_part: int = 10_000) -> Population:
    """The reference additive test's 20 um mean-volume diameter."""
    return config_c1(seed=seed, n_part=n_part, diam_at_mean_vol=2e-5)


def config_c2(seed: int = 20261204, n_part: int = 100_000) -> Population:
    """Brownian, 4 species, trimodal log-normal, power weighting -3."""
    rng = np.random.default_rng(seed)
    modes = [(1e10, 30e-9, 1.6), (1e9, 150e-9, 1.8), (1e6, 2e-6, 2.0)]
    vol, src, mag = lognormal_modes(rng, n_part, modes, -3.0, [2, 2, 1, 1])
    return Population(vol=vol, density=C2_DENSITY.copy(),
                      weight_type=WEIGHT_POWER, weight_magnitude=mag,
                      weight_exponent=-3.0, temp=288.0, pressure=1e5,
                      source=src, grid=(150, 0.5e-9, 0.5e-4))


def config_c3(seed: int = 20261205, n_part: int = 1_000_000,
              n_spec: int = 20) -> Population:
    """Brownian precision case: exponential mode, flat, 20 species."""
    rng = np.random.default_rng(seed)
    density = rng.uniform(1000.0, 2000.0, size=n_spec)
    frac = rng.dirichlet(np.ones(n_spec), size=n_part)
    vol, mag = exp_mode(rng, n_part, 1e9, 2e-6, n_spec=n_spec, fractions=frac)
    return Population(vol=vol, density=density, weight_type=WEIGHT_NONE,
                      weight_magnitude=mag, weight_exponent=0.0, temp=288.0,
                      pressure=1e5, source=np.ones(n_part, dtype=np.int32),
                      grid=(100, 0.5e-8, 0.5e-3))


def config_c4(seed: int = 20261206, n_box: int = 256,
              n_min: int = 10_000, n_max: int = 1_000_000,
              scale: float = 1.0) -> List[Population]:
    """Batched Brownian boxes of uneven sizes and environments."""
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(n_box):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max))) * scale))
        n = max(n, 16)
        temp = rng.uniform(260.0, 310.0)
        pres = rng.uniform(7e4, 1.05e5)
        alpha = rng.uniform(-3.0, 0.0)
        modes = []
        for _m in range(2):
            dg = math.exp(rng.uniform(math.log(20e-9), math.log(500e-9)))
            sg = rng.uniform(1.4, 2.2)
            nc = math.exp(rng.uniform(math.log(1e8), math.log(1e11)))
            modes.append((nc, dg, sg))
        sub = np.random.default_rng(rng.integers(1 << 62))
        vol, src, mag = lognormal_modes(sub, n, modes, alpha, [2, 2, 1, 1])
        boxes.append(Population(vol=vol, density=C2_DENSITY.copy(),
                                weight_type=WEIGHT_POWER, weight_magnitude=mag,
                                weight_exponent=alpha, temp=temp, pressure=pres,
                                source=src, grid=None))
    return boxes


def config_c5(seed: int = 20261207, n_box: int = 64,
              n_part: int = 2_000_000) -> List[Population]:
    """Batched sedimentation boxes: cloud + drizzle water drops, flat."""
    rng = np.random.default_rng(seed)
    boxes = []
    modes = [(1e8, 15e-6, 1.3), (1e4, 200e-6, 1.5)]
    for _ in range(n_box):
        sub = np.random.default_rng(rng.integers(1 << 62))
        vol, src, mag = lognormal_modes(sub, n_part, modes, 0.0, [1])
        boxes.append(Population(vol=vol, density=np.array([1000.0]),
                                weight_type=WEIGHT_NONE, weight_magnitude=mag,
                                weight_exponent=0.0, temp=283.0, pressure=9e4,
                                source=src, grid=(200, 0.5e-6, 2.5e-3)))
    return boxes
