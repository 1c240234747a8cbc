#!/usr/bin/env python3
"""Benchmark of the gfx950 coagulation step (PartMC mc_coag).

``python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]``

A step is one mc_coag timestep over every box of the workload.  Inputs are
uploaded to HBM before the timed region; the timed region is exactly K
steps bracketed by device synchronisation.  The JSON line reports test
pairs per second (the reference's tot_n_samp, src/coagulation.F90:234,627)
as ``value`` and particle-steps per second beside it, a roofline entry for
the dominant kernel measured with HIP events on the library's stream, and
the CPU oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import partmc_amd  # noqa: E402
from partmc_amd import CoagGPU, populations as P  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

WORKLOADS = {
    # name: (builder, kernel, dt, description)
    "c2": (lambda s: [P.config_c2(n_part=int(1e5 * s))], P.KERNEL_BROWN, 60.0,
           "Brownian, 1 box of 1e5 particles, 4 species, trimodal log-normal, "
           "power weighting -3, T=288 K, p=1e5 Pa, dt=60 s"),
    "c4": (lambda s: P.config_c4(scale=s), P.KERNEL_BROWN, 30.0,
           "Brownian, 256 boxes of 1e4-1e6 particles (log-uniform), 4 species, "
           "bimodal log-normal, power weighting alpha in [-3,0], dt=30 s"),
    "c5": (lambda s: P.config_c5(n_part=int(2e6 * s)), P.KERNEL_SEDI, 1.0,
           "sedimentation, 64 boxes of 2e6 water drops, cloud+drizzle modes, flat "
           "weighting, 200 bins over 1 um-5 mm, T=283 K, p=9e4 Pa, dt=1 s"),
}


def per_test_bytes(kernel, n_spec):
    """Algorithmic HBM bytes of one test pair evaluated from raw particle data:
    two particles' species volumes plus two perm entries."""
    return 2 * (8 * n_spec + 4)


def cpu_baseline(pops, kernel, dt, budget_s=15.0):
    """The CPU oracle (sequential FP64 restatement of mc_coag), one core,
    on the first box of the workload for as many steps as fit the budget."""
    import oracle  # the checker, used here only as the timed CPU baseline
    pop = pops[0]
    w = oracle.Weights([[pop.weight_type]], [[pop.weight_magnitude]], [[pop.weight_exponent]])
    b = oracle.Box(pop.density, w, pop.temp, pop.pressure, pop.beta, seed=1)
    n = pop.n_part
    b.add_particles(pop.vol, pop.ids(), np.ones(n), np.ones(n))
    b.next_id = n + 1
    if pop.grid is not None:
        b.sort_with_grid(*pop.grid)
    b.mc_coag(kernel, dt, 0)  # first call sorts and builds the k bounds
    samp = 0
    psteps = 0
    steps = 0
    t_run = 0.0
    while t_run < budget_s and steps < 100000:
        psteps += b.n_part
        t0 = time.perf_counter()
        s, _ = b.mc_coag(kernel, dt, steps + 1)
        t_run += time.perf_counter() - t0
        samp += s
        steps += 1
    return {"value": samp / t_run, "unit": "test-pairs/s", "cores": 1, "kind": "port",
            "sample": f"box 0 of the workload ({n} particles), {steps} steps, "
                      f"{t_run:.1f} s of mc_coag",
            "particle_steps_per_s": psteps / t_run}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="exact", choices=["exact", "parallel"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box of the workload runs on cuda:0; "
              "multi-GPU runs are not part of this build (see DESIGN.md for sharding)",
              file=sys.stderr)
        sys.exit(2)

    build, kernel, dt, desc = WORKLOADS[args.workload]
    t_gen = time.perf_counter()
    pops = build(args.scale)
    t_gen = time.perf_counter() - t_gen
    g = CoagGPU(len(pops), pops[0].density, max_bin=256)
    g.reserve([2 * p.n_part + 1024 for p in pops])
    g.set_rng(20261202, 0)
    for i, pop in enumerate(pops):
        g.load_population(i, pop)
    g.set_mode(partmc_amd.MODE_EXACT if args.mode == "exact" else partmc_amd.MODE_PARALLEL,
               partmc_amd.PREC_FP64 if args.precision == "fp64" else partmc_amd.PREC_FP32_RECHECK,
               1e-5)
    for _ in range(args.warmup):
        g.mc_coag_async(kernel, dt)
    g.sync()
    g.reset_counters()
    g.set_timing(True)
    g.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.mc_coag_async(kernel, dt)
    g.sync()
    elapsed = time.perf_counter() - t0
    c = g.counters()
    psteps = int(g.part_steps().sum())
    n_samp = int(c["n_samp"].sum())
    n_coag = int(c["n_coag"].sum())
    timings = {}
    for name in ("k_prepare", "k_exact_sweep", "k_par_sort", "k_par_tests", "k_par_pairs"):
        ms, n = g.kernel_time(name)
        if n:
            timings[name] = (ms, n)
    dom = max(timings, key=lambda k: timings[k][0]) if timings else None
    roofline = None
    if dom:
        ms, n = timings[dom]
        avg_s = ms / n / 1e3
        n_spec = pops[0].n_spec
        if dom in ("k_exact_sweep", "k_par_tests"):
            traffic = n_samp * per_test_bytes(kernel, n_spec) / n
        else:
            traffic = psteps / n * 12.0
        achieved = traffic / avg_s / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                    "avg_launch_ms": ms / n, "algorithmic_bytes_per_launch": traffic}
    base = None if args.no_cpu_baseline else cpu_baseline(pops, kernel, dt, args.cpu_budget)
    total_particles = sum(p.n_part for p in pops)
    line = {
        "metric": "coag test-pairs/s",
        "value": n_samp / elapsed,
        "unit": "test-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64-recheck",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}", "boxes": len(pops),
                   "particles": total_particles, "mode": args.mode,
                   "precision": args.precision, "scale": args.scale},
        "particle_steps_per_s": psteps / elapsed,
        "sim_s_per_wall_s": dt * args.steps / elapsed,
        "test_pairs": n_samp,
        "coag_events": n_coag,
        "kmax_warnings": int(c["n_warn"].sum()),
        "fp64_rechecks": int(c["n_recheck"].sum()),
        "decision_flips": int(c["n_flip"].sum()),
        "max_rel_kernel_err": float(c["max_rel_err"].max()),
        "kernel_ms": {k: v[0] / v[1] for k, v in timings.items()},
        "roofline": roofline,
        "cpu_baseline": base,
        "setup_s": {"generate": t_gen},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
