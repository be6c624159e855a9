"""Benchmark: ARGOT's point-source radiation step on one MI355X.

Workload (the size the metric is quoted on): 256^3 lognormal H+He mesh,
4096 sources in 8 Gaussian clumps (sigma = 8 cells), 1e5 K blackbody spectra,
256 frequency bins, theta = 0.7, one rank owning the whole box.  A step is one
step_radiation_tree call through the C-ABI: tree build on the host, then the
fused GPU kernel (tree walk, segment, optical-depth walk, photon-conserving
rates, per-cell sums).  Inputs are resident in HBM before timing starts.

Prints one JSON line: segments/s (one segment per ray on one rank), the
fused kernel's roofline against HBM, and the CPU oracle timed on a bounded
sample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(nray, ncross, ncell):
    """Per launch: walk 24 B (segment endpoints) + 12 B (3 tau) + 12 B per
    cell crossing (n_HI, n_HeI, n_HeII); rate 12 B (tau) + 8 B (target, node)
    + 16 B (target dens + 3 fractions) per ray; 24 B (6 rates) per cell."""
    return nray * (24 + 12) + 12 * ncross + nray * (12 + 8 + 16) + 24 * ncell


def cpu_baseline(w, seconds_target=15.0):
    """Oracle (plain-C restatement, OpenMP) on contiguous target ranges spread
    over the box, sized to about seconds_target of host time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_problem
    P = oracle_problem(w)
    n = w.n
    ncell = n ** 3
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    strip = n * n // 4          # a quarter x-slab per range
    nrange = 64
    starts = [int(ncell * (k + 0.5) / nrange) // strip * strip for k in range(nrange)]
    tot_ray = 0
    tot_t = 0.0
    done = 0
    for s0 in starts:
        t0 = time.perf_counter()
        _, _, st = P.step(s0, min(s0 + strip, ncell), nthreads=threads)
        tot_t += time.perf_counter() - t0
        tot_ray += st["nray"]
        done += 1
        if tot_t > seconds_target:
            break
    return {
        "value": tot_ray / tot_t,
        "unit": "segments/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{done} ranges of {strip} target cells ({done * strip} of {ncell}), "
                  f"{tot_ray} segments, {tot_t:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="clumped256")
    ap.add_argument("--kahan", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from argot_amd import ic
    from argot_amd import radiation as R

    t0 = time.perf_counter()
    w = ic.WORKLOADS[args.workload]()
    step = R.PointSourceStep(w, tau_mode=1 if args.kahan else 0)
    setup_s = time.perf_counter() - t0
    print(f"# setup {setup_s:.1f} s: {args.workload} n={w.n} nsrc={w.nsrc}", file=sys.stderr,
          flush=True)

    for i in range(args.warmup):
        st = step.step()
        print(f"# warmup {i}: {st['nray']} segments, kernel {st['ms_kernel']:.1f} ms, "
              f"total {st['ms_total']:.1f} ms", file=sys.stderr, flush=True)
    ker = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = step.step()
        ker.append(st["ms_kernel"])
    wall = time.perf_counter() - t0
    ms_step = 1000.0 * wall / args.steps
    nray, ncross, ncell = st["nray"], st["ncross"], st["ntarget"]
    print(f"# timed: {args.steps} steps, {ms_step:.1f} ms/step, kernel {np.mean(ker):.1f} ms, "
          f"{ncross / max(nray, 1):.1f} crossings/segment", file=sys.stderr, flush=True)

    kernel_ms = float(np.mean(ker))
    bytes_launch = algorithmic_bytes(nray, ncross, ncell)
    achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9
    # memory traffic of this kernel on this workload, from the committed
    # rocprofv3 FETCH_SIZE/WRITE_SIZE passes (profiles/traffic.json)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath) and not args.kahan:
        with open(tpath) as f:
            rec = json.load(f).get("fused_step_kernel", {})
        if rec.get("workload") == args.workload:
            traffic = rec.get("traffic_bytes")
    from argot_amd._lib import check, lib
    copy_gbs = lib().argot_hip_copy_probe(1 << 30, 10)
    check("argot_hip_copy_probe")
    print(f"# copy probe: {copy_gbs:.0f} GB/s (1 GiB copy, read + write)", file=sys.stderr,
          flush=True)
    out = {
        "metric": "ray segments/s (point-source step, exact rates)",
        "value": nray / (ms_step * 1e-3),
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 walk, f32/f64 rates",
        "data": "synthetic (seeded lognormal field, clumped sources)",
        "config": {"workload": f"{args.workload}: {w.n}^3 mesh, {w.nsrc} sources, "
                               "256 nu bins, theta 0.7, one rank",
                   "segments_per_step": int(nray), "crossings_per_step": int(ncross),
                   "tau_sum": "kahan" if args.kahan else "plain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "fused_step_kernel", "kernel_ms": kernel_ms,
                     "bytes_per_launch": int(bytes_launch),
                     "measured_copy_gbs": copy_gbs},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
