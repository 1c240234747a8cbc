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
import hashlib
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


# Algorithmic HBM bytes per unit of work of each timed kernel (the minimum
# a kernel must move): keys, perm entries, per-particle terms, species
# volumes and the per-pair tables it reads or writes.
def kernel_bytes(name, n_spec, precision, part_steps_per_launch, tests_per_launch,
                 pairs_per_launch, events_per_launch):
    if name == "k_par_hist":
        return 4.0 * part_steps_per_launch                   # read key
    if name == "k_par_scatter":
        return 8.0 * part_steps_per_launch                   # read key, write perm
    if name == "k_par_pairs":
        return 20.0 * pairs_per_launch                       # 2 counts, k_max, n_samp
    if name in ("k_par_tests", "k_exact_sweep", "k_par_tail"):
        per = 8.0 + (32.0 if precision == "fp32" else 2 * 8.0 * n_spec)
        return per * tests_per_launch                        # 2 perm + 2 particles
    if name == "k_par_apply":
        return 3.0 * (8.0 * n_spec + 28.0) * events_per_launch
    if name == "k_prepare":
        return 20.0 * part_steps_per_launch
    return 0.0


# What limits each kernel: the bandwidth roofline bounds the streaming
# passes; the target events and the tail are chains of dependent loads and
# FP64 special functions, one event or chain per lane.
LIMITER = {
    "k_par_acc_tests": "latency: each target event is one lane's sample loop of dependent "
                       "loads and FP64 special functions",
    "k_par_tail": "latency: one workgroup per box runs the remaining claim rounds, one link "
                  "of a chain per round",
    "k_exact_sweep": "latency: one wavefront per box follows the sequential chain",
    "k_par_tests": "latency of random gathers of two particles per test",
    "k_par_apply": "hbm: random gathers of the merged particles' records",
}

TIMED = ("k_prepare", "k_exact_sweep", "k_par_hist", "k_par_scatter", "k_par_pairs",
         "k_par_tests", "k_par_acc_tests", "k_par_apply", "k_par_tail", "k_par_compact")


def state_digest(g, n_box):
    """SHA-1 over every box's particle ids and species volumes, in storage
    order: equal digests mean equal final states."""
    h = hashlib.sha1()
    for b in range(n_box):
        ids, vol = g.download_ids_vols(b)
        h.update(ids.tobytes())
        h.update(vol.tobytes())
    return h.hexdigest()


def run_gpu(pops, kernel, dt, mode, precision, steps, warmup, seed=20261202, rebalance=False,
            validation=False, digest=False):
    g = CoagGPU(len(pops), pops[0].density, max_bin=256)
    g.reserve([p.n_part + p.n_part // 4 + 1024 for p in pops])
    g.set_rng(seed, 0)
    for i, pop in enumerate(pops):
        g.load_population(i, pop)
    g.set_mode(partmc_amd.MODE_EXACT if mode == "exact" else partmc_amd.MODE_PARALLEL,
               partmc_amd.PREC_FP64 if precision == "fp64" else partmc_amd.PREC_FP32_RECHECK,
               1e-5)
    g.set_validation(validation)
    for _ in range(warmup):
        g.mc_coag_async(kernel, dt)
    g.sync()
    g.reset_counters()
    g.set_timing(not validation)
    g.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.mc_coag_async(kernel, dt)
    g.sync()
    elapsed = time.perf_counter() - t0
    c = g.counters()
    st = g.path_stats()
    ar = g.arrival_stats()
    out = {
        "claim_rounds": st["n_rounds"],
        "tail_tests": st["n_tail"],
        "tail_rounds": st["n_tail_rounds"],
        "accel_pairs": int(st["n_accel_pairs"].sum()),
        "acc_samples": st["acc_samples"],
        "n_dropped": int(st["n_dropped"].sum()),
        "arrivals": int(ar["n_arrivals"].sum()),
        "arrival_overflow": int(ar["n_overflow"].sum()),
        "elapsed": elapsed,
        "n_samp": int(c["n_samp"].sum()),
        "n_coag": int(c["n_coag"].sum()),
        "n_warn": int(c["n_warn"].sum()),
        "n_recheck": int(c["n_recheck"].sum()),
        "n_flip": int(c["n_flip"].sum()),
        "max_rel_err": float(c["max_rel_err"].max()),
        "max_p_err": float(g.error_stats()["max_p_err"].max()),
        "part_steps": int(g.part_steps().sum()),
        "timings": {},
    }
    for name in TIMED:
        ms, n = g.kernel_time(name)
        if n:
            out["timings"][name] = (ms, n)
    g.set_timing(False)
    if digest:
        out["digest"] = state_digest(g, len(pops))
    if rebalance:
        out["rebalance"] = time_rebalance(g, len(pops))
    g.close()
    return out


def time_rebalance(g, n_box):
    """aero_state_rebalance on the evolved boxes, after the timed steps:
    one halving pass per box (ideal count a third of the current one), then
    one doubling pass back (ideal 2.5 times), then a call with nothing to
    do.  Wall clock of the whole call, counts and host reads included."""
    n0 = [g.n_part(b) for b in range(n_box)]
    for b in range(n_box):
        g.set_n_part_ideal(b, n0[b] / 3.0)
    t = time.perf_counter()
    g.rebalance()
    t_halve = time.perf_counter() - t
    n1 = [g.n_part(b) for b in range(n_box)]
    for b in range(n_box):
        g.set_n_part_ideal(b, n1[b] * 2.5)
    t = time.perf_counter()
    g.rebalance()
    t_double = time.perf_counter() - t
    n2 = [g.n_part(b) for b in range(n_box)]
    t = time.perf_counter()
    g.rebalance()
    t_noop = time.perf_counter() - t
    return {"particles_before": int(sum(n0)), "after_halving": int(sum(n1)),
            "after_doubling": int(sum(n2)), "halve_ms": t_halve * 1e3,
            "double_ms": t_double * 1e3, "no_op_ms": t_noop * 1e3,
            "halve_particles_per_s": sum(n0) / t_halve,
            "double_particles_per_s": sum(n1) / t_double}


def oracle_run(args):
    """The CPU oracle (the sequential FP64 restatement of mc_coag, built
    -O3) on one box for about budget_s seconds of mc_coag calls: test pairs,
    particle-steps, steps and seconds.  Runs in a process that never loads
    the HIP library."""
    pop, kernel, dt, budget_s, seed = args
    import oracle  # the checker, used here only as the timed CPU baseline
    lib_fast = os.path.join(ROOT, "oracle", "libcoag_oracle_fast.so")
    if os.path.exists(lib_fast):
        oracle.LIB_PATH = lib_fast
    w = oracle.Weights([[pop.weight_type]], [[pop.weight_magnitude]], [[pop.weight_exponent]])
    b = oracle.Box(pop.density, w, pop.temp, pop.pressure, pop.beta, seed=seed)
    n = pop.n_part
    b.add_particles(pop.vol, pop.ids(), np.ones(n), np.ones(n))
    b.next_id = n + 1
    if pop.grid is not None:
        b.sort_with_grid(*pop.grid)
    b.mc_coag(kernel, dt, 0)  # first call sorts and builds the k bounds
    samp = psteps = steps = 0
    t_run = 0.0
    while t_run < budget_s and steps < 100000:
        psteps += b.n_part
        t0 = time.perf_counter()
        s, _ = b.mc_coag(kernel, dt, steps + 1)
        t_run += time.perf_counter() - t0
        samp += s
        steps += 1
    return samp, psteps, steps, t_run, n, os.path.basename(oracle.LIB_PATH)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pops, kernel, dt, budget_s=15.0, procs=16):
    """The CPU oracle timed on the host, before any GPU call: one process on
    box 0, then `procs` processes side by side on boxes 0..procs-1 (boxes
    cycle when the workload has fewer), each with its own seed, as the
    reference runs independent repeats one per core.  Rates are sums over
    the processes of their own test pairs (particle-steps) per second."""
    import multiprocessing as mp
    one = oracle_run((pops[0], kernel, dt, budget_s, 1))
    n_cpu = os.cpu_count() or 1
    procs = max(1, min(procs, n_cpu))
    ctx = mp.get_context("spawn")
    jobs = [(pops[i % len(pops)], kernel, dt, budget_s, 1 + i) for i in range(procs)]
    with ctx.Pool(procs) as pool:
        many = pool.map(oracle_run, jobs)
    rate = sum(r[0] / r[3] for r in many)
    prate = sum(r[1] / r[3] for r in many)
    return {"value": one[0] / one[3], "unit": "test-pairs/s", "cores": 1, "kind": "port",
            "sample": f"box 0 of the workload ({one[4]} particles), {one[2]} steps, "
                      f"{one[3]:.1f} s of mc_coag in one process ({one[5]})",
            "particle_steps_per_s": one[1] / one[3],
            "cpu_model": cpu_model(), "host_cpus": n_cpu,
            "per_core_processes": {
                "processes": procs, "value": rate, "unit": "test-pairs/s",
                "particle_steps_per_s": prate,
                "sample": f"boxes 0..{procs - 1} of the workload, one process each "
                          f"({sum(r[4] for r in many)} particles), about {budget_s:.0f} s "
                          "of mc_coag each, side by side"}}


def copy_probe(nbytes, launches=5):
    """A plain device-to-device copy of nbytes, for the scale of FETCH_SIZE and
    WRITE_SIZE: each launch reads nbytes and writes nbytes."""
    import torch
    src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(launches):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": launches}))


def probe_rates():
    """Measured FP32 FMA, FP64 FMA and FP32 transcendental rates of this GPU
    (partmc_amd/libpmc_probe.so): the ceilings of kernel evaluation."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "partmc_amd", "libpmc_probe.so"))
    rates = (ctypes.c_double * 3)()
    if lib.pmc_probe_rates(0, rates) != 0:
        raise RuntimeError("pmc_probe_rates failed")
    out = {"fp32_fma_per_s": rates[0], "fp32_tflops": 2 * rates[0] / 1e12,
           "fp64_fma_per_s": rates[1], "fp64_tflops": 2 * rates[1] / 1e12,
           "fp32_transcendental_per_s": rates[2]}
    print(json.dumps({"probe": out}))
    return out


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, copy-scaled), or None"""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        ks = json.load(f)["kernels"]
    k = ks.get(kernel)
    if k is not None:
        return k["bytes"]
    # a templated kernel appears once per instantiation ("void k<false>"):
    # the launches of all of them, averaged
    parts = [v for name, v in ks.items()
             if name.split("<")[0].replace("void ", "").strip() == kernel and "<" in name]
    n = sum(v["dispatches"] for v in parts)
    return sum(v["bytes"] * v["dispatches"] for v in parts) / n if n else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c4", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="parallel", choices=["exact", "parallel"])
    ap.add_argument("--precision", default=None, choices=["fp64", "fp32"],
                    help="fp32 (with the FP64 recheck) by default; the exact mode is FP64 only")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--no-rebalance", action="store_true")
    ap.add_argument("--no-validation", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="processes of the per-core CPU baseline (at most the host's CPUs)")
    ap.add_argument("--probe", action="store_true",
                    help="only measure the FP32/FP64 FMA and FP32 transcendental rates")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many bytes (the scale of "
                         "the HBM counters)")
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "fp64" if args.mode == "exact" else "fp32"
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.probe:
        probe_rates()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box of the workload runs on cuda:0; "
              "multi-GPU runs are not part of this build (see DESIGN.md for sharding)",
              file=sys.stderr)
        sys.exit(2)

    build, kernel, dt, desc = WORKLOADS[args.workload]
    t_gen = time.perf_counter()
    pops = build(args.scale)
    t_gen = time.perf_counter() - t_gen
    n_spec = pops[0].n_spec
    # the CPU baseline first, in processes that never touch the GPU
    base = None if args.no_cpu_baseline else cpu_baseline(pops, kernel, dt, args.cpu_budget,
                                                         args.cpu_procs)
    compare64 = (args.mode == "parallel" and args.precision == "fp32"
                 and not args.no_fp64_compare)
    r = run_gpu(pops, kernel, dt, args.mode, args.precision, args.steps, args.warmup,
                seed=args.seed, rebalance=not args.no_rebalance, digest=compare64)
    elapsed = r["elapsed"]
    n_samp, n_coag, psteps = r["n_samp"], r["n_coag"], r["part_steps"]
    timings = r["timings"]
    # the plain-FP64 comparison pass on the same populations and streams
    r64 = None
    if compare64:
        r64 = run_gpu(pops, kernel, dt, args.mode, "fp64", args.steps, args.warmup,
                      seed=args.seed, digest=True)
    # the validation pass: FP32 with every test also evaluated in FP64
    rv = None
    if args.precision == "fp32" and not args.no_validation:
        rv = run_gpu(pops, kernel, dt, args.mode, "fp32", args.steps, args.warmup,
                     seed=args.seed, validation=True)
    roofline = None
    if timings:
        dom = max(timings, key=lambda k: timings[k][0])
        ms, n = timings[dom]
        pairs = len(pops) * 256 * 256
        traffic = kernel_bytes(dom, n_spec, args.precision, psteps / args.steps,
                               n_samp / max(n, 1), pairs, n_coag / max(n, 1))
        if dom in ("k_par_tests", "k_par_apply"):
            traffic = kernel_bytes(dom, n_spec, args.precision, psteps / args.steps,
                                   n_samp / n, pairs, n_coag / n)
        if dom == "k_par_acc_tests":
            # per source sample: its list entry and its species volumes
            traffic = (4.0 + 8.0 * n_spec) * r["acc_samples"] / n
        if dom == "k_par_tail":
            # the tests the tail finished, per launch (one launch per step)
            traffic = kernel_bytes(dom, n_spec, args.precision, psteps / args.steps,
                                   r["tail_tests"] / n, pairs, n_coag / n)
        avg_s = ms / n / 1e3
        achieved = traffic / avg_s / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS,
                    "traffic": measured_traffic(args.workload, dom) if args.mode == "parallel"
                    and args.scale == 1.0 else None, "kernel": dom,
                    "limiter": LIMITER.get(dom, "hbm"),
                    "avg_launch_ms": ms / n, "launches": n,
                    "algorithmic_bytes_per_launch": traffic}
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
        "kmax_warnings": r["n_warn"],
        "claim_rounds_per_step": r["claim_rounds"] / args.steps,
        "tail_tests_per_step": r["tail_tests"] / args.steps,
        "tail_rounds_per_step": r["tail_rounds"] / args.steps,
        "accel_pairs_per_step": r["accel_pairs"] / args.steps,
        "acc_samples_per_step": r["acc_samples"] / args.steps,
        "fp64_rechecks": r["n_recheck"],
        "dropped_draws": r["n_dropped"],
        "arrivals_per_step": r["arrivals"] / args.steps,
        "arrival_overflow_per_step": r["arrival_overflow"] / args.steps,
        "precision_check": None if rv is None else {
            "decision_flips": rv["n_flip"],
            "max_rel_kernel_err": rv["max_rel_err"],
            "max_abs_p_err": rv["max_p_err"],
            "recheck_band": 1e-5,
            "tests_checked": rv["n_samp"],
            "fp64_rechecks": rv["n_recheck"],
            "note": "validation pass: the same steps with every FP32 test also evaluated in "
                    "FP64; flips count tests whose FP32+recheck decision differs from FP64's"},
        "kernel_ms": {k: v[0] / v[1] for k, v in timings.items()},
        "kernel_total_ms": {k: v[0] for k, v in timings.items()},
        "fp64_compare": None if r64 is None else {
            "test_pairs_per_s": r64["n_samp"] / r64["elapsed"],
            "ms_per_step": r64["elapsed"] / args.steps * 1e3,
            "k_par_tests_ms": r64["timings"].get("k_par_tests", (0.0, 1))[0] / max(
                r64["timings"].get("k_par_tests", (0.0, 1))[1], 1),
            "same_final_state": r64["digest"] == r["digest"],
            "state_digest": r["digest"],
            "ratio_step": (n_samp / elapsed) / max(r64["n_samp"] / r64["elapsed"], 1e-30),
            "ratio_test_kernel": (r64["timings"].get("k_par_tests", (0.0, 1))[0] /
                                  max(timings.get("k_par_tests", (1e-30, 1))[0], 1e-30)),
        },
        "roofline": roofline,
        "cpu_baseline": base,
        "rebalance": r.get("rebalance"),
        "seed": args.seed,
        "setup_s": {"generate": t_gen},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
