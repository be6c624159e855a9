"""Benchmark: ARGOT's point-source radiation step on one MI355X.

Workload (the size the metric is quoted on): 256^3 lognormal H+He mesh,
4096 sources in 8 Gaussian clumps (sigma = 8 cells), 1e5 K blackbody spectra,
256 frequency bins, theta = 0.7, one rank owning the whole box.  A step is one
step_radiation_tree call through the C-ABI: tree build on the host, then the
device work (field packing, rate zeroing, the fused persistent kernel: tree
walk, segment, optical-depth walk, photon-conserving rates, per-cell sums)
replayed as one HIP graph.  Inputs are resident in HBM before timing starts.
The autotuner picks the tile shape and grid size during the untimed warmup.

Prints one JSON line: segments/s (one segment per ray on one rank), the
fused kernel's roofline against HBM, the CPU oracle timed on a bounded
sample of the same workload on this host's cores, and two side measurements:
the MLP rate surrogate (training, validated accuracy, rays it may serve at
1e-3, step speed with it) and the ART diffuse transfer on 128^3 x 768
directions.
"""
import argparse
import hashlib
import json
import os
import platform
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


# Sources that define the fused kernel's code: a committed profile figure is
# used only when it was measured on these exact sources.
KERNEL_SOURCES = ("kernels.hip", "argot_device.h", "argot_internal.h", "kernels.h", "mlp.h",
                  "surrogate.h", "Makefile")


def kernel_source_hash():
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "argot_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def profile_record(name, workload):
    """(record, None) for a committed rocprofv3 figure of kernel `name` on
    `workload` (profiles/traffic.json), or (None, reason) when there is none
    or it was measured on other kernel sources."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, "no profile"
    with open(path) as f:
        rec = json.load(f).get(name)
    if not rec or rec.get("workload") != workload:
        return None, f"no profile of {name} on {workload}"
    if rec.get("source_hash") != kernel_source_hash():
        return None, "stale: the kernel sources changed after the profile"
    return rec, None


def oracle_problem_for(w, fast=False):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_problem
    return oracle_problem(w, fast=fast)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def time_oracle(P, n, seconds_target, threads, nrange=64):
    """The oracle's step over contiguous target ranges (a quarter x-slab
    each) spread over the box, until about seconds_target of host time."""
    ncell = n ** 3
    strip = n * n // 4
    starts = [int(ncell * (k + 0.5) / nrange) // strip * strip for k in range(nrange)]
    tot_ray, tot_t, done = 0, 0.0, 0
    for s0 in starts:
        t0 = time.perf_counter()
        _, _, st = P.step(s0, min(s0 + strip, ncell), nthreads=threads)
        tot_t += time.perf_counter() - t0
        tot_ray += st["nray"]
        done += 1
        if tot_t > seconds_target:
            break
    return tot_ray, tot_t, done, strip


def cpu_baseline(w, seconds_target=12.0):
    """The oracle (plain-C restatement, OpenMP on all host cores) on a
    bounded sample of the workload, built twice: with IEEE flags (the
    validated checker; `value`) and with the reference's own Makefile flags
    (-O3 -ffast-math -funroll-loops, src/Makefile:69; `reference_flags`)."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = None
    for fast in (False, True):
        P = oracle_problem_for(w, fast=fast)
        nray, t, done, strip = time_oracle(P, w.n, seconds_target, threads)
        rec = {"value": nray / t, "unit": "segments/s", "cores": threads, "kind": "port",
               "sample": f"{done} ranges of {strip} target cells ({done * strip} of {w.n ** 3}), "
                         f"{nray} segments, {t:.1f} s",
               "build": "-O3 -ffast-math -funroll-loops" if fast else "-O2 IEEE (no fast-math)"}
        if not fast:
            out = rec
            out["cpu"] = cpu_model()
        else:
            out["reference_flags"] = rec
    return out


def fp64_error(w, rates, rec, ntau=20000, size=2048, others=None):
    """Relative error against the FP64 oracle (exact chords, expm1): the
    step's six per-cell rates on two ranges of target cells (the middle of
    the box and the cell of the first source), beside the reference
    arithmetic's own (the oracle's FP32 path); and the FP32 optical depths
    of ntau captured rays against the exact-chord walk of the same segment."""
    P = oracle_problem_for(w)
    ncell = w.n ** 3
    i, j, k = np.minimum((np.asarray(w.src_pos[0], np.float64) * w.n).astype(np.int64), w.n - 1)
    starts = [ncell // 2, int((i * w.n + j) * w.n + k) - size // 2]
    others = {k: v for k, v in (others or {}).items() if v is not None}
    got, r32s, r64s = [], [], []
    oth = {k: [] for k in others}
    for s0 in starts:
        s0 = int(min(max(s0, 0), ncell - size))
        r32, r64, _ = P.step(s0, s0 + size, f64=True)
        got.append(rates[s0:s0 + size])
        for k in others:
            oth[k].append(others[k][s0:s0 + size])
        r32s.append(r32[s0:s0 + size])
        r64s.append(r64[s0:s0 + size])
    got, r32, r64 = np.concatenate(got), np.concatenate(r32s), np.concatenate(r64s)
    fin = np.isfinite(r64).all(1) & np.isfinite(got).all(1) & np.isfinite(r32).all(1)
    names = ("GammaHI", "HeatHI", "GammaHeI", "GammaHeII", "HeatHeI", "HeatHeII")

    def err(a, b):
        a = np.asarray(a, np.float64)
        return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)

    eg, er = err(got[fin], r64[fin]), err(r32[fin], r64[fin])
    out = {"cells": int(fin.sum()),
           "rates": {n: {"max": float(eg[:, c].max()), "median": float(np.median(eg[:, c])),
                         "reference_max": float(er[:, c].max()),
                         "reference_median": float(np.median(er[:, c]))}
                     for c, n in enumerate(names)}}
    for k, v in oth.items():
        v = np.concatenate(v)
        ok = fin & np.isfinite(v).all(1)
        e = err(v[ok], r64[ok])
        out[k] = {"cells": int(ok.sum()),
                  "rates": {n: {"max": float(e[:, c].max()), "median": float(np.median(e[:, c]))}
                            for c, n in enumerate(names)}}
    if rec is not None and len(rec):
        idx = np.linspace(0, len(rec) - 1, min(ntau, len(rec))).astype(np.int64)
        n = w.n
        lo = np.float64(np.float32(0.0))
        h = np.float64(np.float32(np.float32(1.0) / np.float32(n)))
        e3 = np.zeros((len(idx), 3))
        for q, r in enumerate(idx):
            t = int(rec["target"][r])
            ijk = np.array([t // (n * n), (t // n) % n, t % n], np.float32).astype(np.float64)
            end = (lo + (ijk + 0.5) * h).astype(np.float32)
            t64, _ = P.walk_exact(rec["start"][r], end)
            d = np.abs(rec["tau"][r].astype(np.float64) - t64)
            e3[q] = np.where(d == 0.0, 0.0, d / np.maximum(np.abs(t64), 1e-30))
        out["tau"] = {"rays": int(len(idx)), "max": float(e3.max()),
                      "median": float(np.median(e3))}
    return out


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def surrogate_leg(step, plain_seg_s, nray, plain_kernel_ms, exact_rates):
    """The MLP rate surrogate: trained on the C4 grid (2^22 uniform samples
    over [-6, 4]^3, 128 bins), then on the step's own grid around the
    optical depths of 2^20 rays the step emits (captured; cubic baseline,
    validated at 5e-4 on 2^21 points drawn the same way), both by the
    bf16-MFMA training kernels on the GPU; then the step with rate mode 1
    (the network where validated, the exact integral elsewhere), timed like
    the headline, with its per-cell error against the exact path.  Returns
    (report, rate-mode-1 rates)."""
    from argot_amd import surrogate as S
    out = {}
    nu, cs, spec = S.spectral_grid(128, S.C4_LOG_NU_MAX)
    rep = S.train(S.Config(nu, cs, spec, nsample=1 << 22, batch=16384, nstep=20000,
                           trust_n=32, trust_tol=1e-3))
    out["c4_128bin"] = {k: rep[k] for k in ("ms_train", "err_median", "err_p99", "err_max",
                                            "trusted_frac", "loss_last")}
    log(f"surrogate C4: {out['c4_128bin']}")
    rec, _ = step.capture(1 << 20)
    seeds = S.seeds_from_rays(rec["tau"], step.run, step.cs)
    del rec
    nu, cs, spec = S.step_grid(step.freq, step.cs,
                               np.ctypeslib.as_array(step.src[0].photon_rate))
    tol = 5e-4
    rep = S.train(S.Config(nu, cs, spec, nsample=1 << 22, batch=16384, nstep=150000,
                           trust_n=64, trust_tol=tol, log_tau_lo=-3.5, log_tau_hi=3.0,
                           seeds=seeds, jitter=0.1, poly_degree=3, nvalidate=1 << 21,
                           lr_max=3e-3, seed=8))
    out["step_grid"] = {k: rep[k] for k in ("ms_train", "err_median", "err_p99", "err_max",
                                            "trusted_frac", "loss_last")}
    out["step_grid"].update({"samples": "2^22 around 2^20 captured rays, jitter 0.1 dex",
                             "baseline": "cubic", "trust_tol": tol})
    log(f"surrogate step grid: {out['step_grid']}")
    step.rate_mode = 1
    try:
        ms, kms, st = timed_steps(step, 3)
        rates = step.rates()
    finally:
        step.rate_mode = 0
    ex = np.asarray(exact_rates, np.float64)
    fin = np.isfinite(ex) & (ex > 0)
    rel = np.abs(rates.astype(np.float64)[fin] - ex[fin]) / ex[fin]
    out["step_rate_mode1"] = {"segments_per_s": nray / (ms * 1e-3), "ms_per_step": ms,
                              "kernel_ms": kms, "exact_kernel_ms": plain_kernel_ms,
                              "rays_served": int(st["nsurrogate"]),
                              "served_frac": st["nsurrogate"] / max(st["nray"], 1),
                              "vs_exact": (nray / (ms * 1e-3)) / plain_seg_s,
                              "cell_error_vs_exact_max": float(rel.max()),
                              "tolerance": tol}
    log(f"surrogate step: {out['step_rate_mode1']}")
    # Amdahl ceiling of any rate surrogate: the fused kernel with the
    # frequency loop taken out (a timing-only build, profiled on the same
    # workload) bounds how fast a free, perfect surrogate could make it
    prof, note = profile_record("fused_step_kernel_no_rate_loop", "clumped256")
    if prof:
        out["amdahl"] = {"no_rate_loop_kernel_ms": prof["kernel_ms"],
                         "plain_kernel_ms": plain_kernel_ms,
                         "max_speedup": plain_kernel_ms / prof["kernel_ms"],
                         "summary": prof.get("summary")}
    else:
        out["amdahl"] = {"note": note}
    return out, rates


def walk_leg(step, nsample=8 << 20):
    """Walk only: calc_optical_depth_cuda (the reference's walk entry point:
    56-B ray_segment records, one lane per segment) on the first nsample
    segments the step emits, captured from the fused kernel.  Their optical
    depths must equal the fused kernel's bit for bit."""
    import ctypes as C
    from argot_amd import abi
    from argot_amd._lib import check, lib
    rec, _ = step.capture(nsample)
    n = len(rec)
    side = step.w.n
    run = step.run
    t = rec["target"]
    idx = np.stack([t // (side * side), (t // side) % side, t % side], axis=1)
    lo = np.array([run.xmin, run.ymin, run.zmin], np.float32).astype(np.float64)
    h = np.array([run.delta_x, run.delta_y, run.delta_z], np.float32).astype(np.float64)
    # target centre as the kernels form it: lo + ((float)i + 0.5) h in
    # double, rounded to float
    centre = (lo + (idx.astype(np.float32).astype(np.float64) + 0.5) * h).astype(np.float32)
    L = lib()
    segp = C.c_void_p()
    L.allocate_pinned_segment(C.byref(segp), n)
    check("allocate_pinned_segment")
    try:
        seg = abi.view_at(abi.RaySegment, segp.value, n)
        seg["ray_indx"] = np.arange(n, dtype=np.uint64)
        for k, a in enumerate("xyz"):
            seg[f"{a}pos_start"] = rec["start"][:, k]
            seg[f"{a}pos_end"] = centre[:, k]
        run.nseg = n
        ms = []
        for _ in range(3):
            L.calc_optical_depth_cuda(segp, C.byref(step.cmem), C.byref(step.cpar), C.byref(run))
            check("calc_optical_depth_cuda")
            ms.append(step.stats()["ms_walk"])
        tau = np.stack([seg["optical_depth_HI"], seg["optical_depth_HeI"],
                        seg["optical_depth_HeII"]], axis=1)
        same = bool(np.array_equal(tau.view(np.uint32), rec["tau"].view(np.uint32)))
    finally:
        L.deallocate_pinned_segment(segp)
    ncross = int(rec["ncross"].astype(np.int64).sum())
    kms = float(np.mean(ms[1:]))
    # per segment: the 56-B record read, 12 B of optical depths written;
    # 12 B gathered per crossing
    bytes_ = n * (56 + 12) + 12 * ncross
    out = {"entry_point": "calc_optical_depth_cuda", "segments": n, "crossings": ncross,
           "ms_kernel": kms, "segments_per_s": n / (kms * 1e-3),
           "crossings_per_s": ncross / (kms * 1e-3),
           "roofline": {"bound": "hbm", "achieved": bytes_ / (kms * 1e-3) / 1e9,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": bytes_ / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "tau_equal_to_fused_kernel": same}
    log(f"walk only: {out}")
    return out, rec


def rate_call_leg(step, rec):
    """calc_photoion_rate_cuda (the reference's per-call rate entry point) on
    the rays the walk leg captured, grouped by target, each with its tree
    node's or source's spectrum row: wall time of the call (host regrouping
    and spectrum deduplication, transfer, kernel) and the rate kernel's own
    time.  Afterwards the step runs again, so the mesh holds its rates."""
    import ctypes as C
    from argot_amd import abi
    from argot_amd import radiation as R
    from argot_amd._lib import check, lib
    L = lib()
    w, run = step.w, step.run
    n = len(rec)
    pos = np.ascontiguousarray(w.src_pos, np.float32)
    src_rows = np.array([np.ctypeslib.as_array(step.src[i].photon_rate) for i in range(w.nsrc)])
    box = (C.c_float * 6)(run.xmin, run.ymin, run.zmin, run.xmax, run.ymax, run.zmax)
    h = L.argot_hip_tree_build(pos.ctypes.data, np.ascontiguousarray(src_rows).ctypes.data,
                               w.nsrc, src_rows.shape[1], box, 0.7)
    try:
        nn = L.argot_hip_tree_nnode(h)
        node_rows = np.array([np.ctypeslib.as_array(L.argot_hip_tree_spec(h, u),
                                                    shape=(src_rows.shape[1],))
                              for u in range(nn)])
    finally:
        L.argot_hip_tree_free(h)
    table = np.concatenate([node_rows, src_rows])
    index = np.where(rec["kind"] == 0, rec["id"], nn + rec["id"])
    side = w.n
    t = rec["target"]
    ijk = np.stack([t // (side * side), (t // side) % side, t % side], axis=1)
    arr = R.light_rays(rec["src"], (table, index), ijk, rec["tau"])
    iop = C.c_void_p()
    L.allocate_pinned_light_ray_IO(C.byref(iop), n)
    check("allocate_pinned_light_ray_IO")
    walls, kms = [], []
    try:
        for _ in range(3):
            L.zero_out_photoion_rate_cuda(C.byref(step.cmem), C.byref(step.cpar), C.byref(run))
            check("zero_out_photoion_rate_cuda")
            run.nray = n
            t0 = time.perf_counter()
            L.calc_photoion_rate_cuda(C.byref(step.cmem), arr, iop, C.byref(step.cpar),
                                      C.byref(run))
            check("calc_photoion_rate_cuda")
            walls.append(1000.0 * (time.perf_counter() - t0))
            st = step.stats()
            kms.append(st["ms_rate"])
    finally:
        L.deallocate_pinned_light_ray_IO(iop)
    del arr
    step.step()
    ntarget = len(np.unique(t))
    kms_ = float(np.mean(kms[1:]))
    # per ray: optical depths 12 B, target and spectrum index 8 B, target
    # density and fractions 16 B; per target cell 24 B of rates
    bytes_ = 36 * n + 24 * ntarget
    out = {"entry_point": "calc_photoion_rate_cuda", "kernel": "target_rates_kernel",
           "rays": n, "targets": ntarget, "distinct_spectra": int(st["nspec_call"]),
           "ms_call": float(np.mean(walls[1:])), "ms_kernel": kms_,
           "rays_per_s_kernel": n / (kms_ * 1e-3),
           "rays_per_s_call": n / (np.mean(walls[1:]) * 1e-3),
           "roofline": {"bound": "valu (FP32 frequency loop, 256 bins per ray)",
                        "achieved": bytes_ / (kms_ * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": bytes_ / (kms_ * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    log(f"rate call: {out}")
    return out


def chem_leg(step, dtime=1.0e-3):
    """Chemistry of the whole mesh over dtime (code units) after the pass, as
    the reference's loop runs it: smoothed rates, then step_chemistry; run
    twice, the second timed.  The cells that hold a source have infinite
    rates in this workload (sources at cell centres); they are set to zero
    first, since such a cell's chemistry would stop on its NaN state."""
    from argot_amd import abi, ic
    from argot_amd import radiation as R
    step.mesh_bytes()
    v = abi.mesh_view(step.mesh)
    pc = v["prev_chem"]
    for f in R.RATE_FIELDS:
        x = pc[f]
        x[~np.isfinite(x)] = 0.0
        pc[f] = x
    v["prev_chem"] = pc
    step.send_mesh()
    step.smooth()
    step.chemistry(np.float32(dtime))
    t0 = time.perf_counter()
    dc, du = step.chemistry(np.float32(dtime))
    ms = 1000.0 * (time.perf_counter() - t0)
    cs = step.chem_stats()
    ncell = int(cs["ncell"])
    out = {"workload": f"{step.w.name} mesh after the point-source pass",
           "dtime_code": dtime, "dtime_yr": dtime * step.run.tunit / ic.YEAR_S,
           "ms_kernel": cs["ms_kernel"], "ms_call": ms,
           "cells_per_s": ncell / (cs["ms_kernel"] * 1e-3),
           "substeps": int(cs["nsub"]), "bisections": int(cs["niter"]),
           "substeps_per_s": cs["nsub"] / (cs["ms_kernel"] * 1e-3),
           "heatcool_evals_per_s": cs["niter"] / (cs["ms_kernel"] * 1e-3),
           "cells_stopped": int(cs["nstopped"]),
           "max_diff_chem": float(dc), "max_diff_uene": float(du)}
    log(f"chemistry: {out}")
    return out


def art_leg():
    """ART diffuse transfer, 128^3 lognormal mesh, 768 HEALPix directions."""
    from argot_amd import art as A
    from argot_amd import ic
    w = ic.art128()
    st = A.DiffuseStep(w, nside=8)
    st.step()
    t0 = time.perf_counter()
    s = st.step()
    ms = 1000.0 * (time.perf_counter() - t0)
    st.close()
    # per crossing: absorption + source (3 bands) 24 B read, band intensities
    # 12 B and path length 4 B read-modify-write (32 B)
    bytes_ = s["ncross"] * (24 + 32)
    return {"workload": "art128: 128^3, 768 directions, 3 bands", "ncross": int(s["ncross"]),
            "ms_per_pass": ms, "ms_trace": s["ms_trace"], "ms_rates": s["ms_gh"],
            "crossings_per_s": s["ncross"] / (s["ms_trace"] * 1e-3),
            "trace_roofline": {"bound": "hbm", "achieved": bytes_ / (s["ms_trace"] * 1e-3) / 1e9,
                               "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": bytes_ / (s["ms_trace"] * 1e-3) / 1e9 / HBM_PEAK_GBS}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="clumped256")
    ap.add_argument("--kahan", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tile", default=None, help="fixed target tile, e.g. 4x4x4 (skips tuning)")
    ap.add_argument("--blocks-per-cu", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip every side leg (the same as --legs none)")
    ap.add_argument("--legs", default="all",
                    help="side legs to run, comma-separated: kahan, surrogate, walk, rate_call, "
                         "chemistry, art, lognormal128 (default: all)")
    ap.add_argument("--art-only", action="store_true",
                    help="run only the ART leg and print its JSON (profiling)")
    ap.add_argument("--chem-only", action="store_true",
                    help="run one point-source pass and the chemistry leg, print its JSON")
    ap.add_argument("--walk-only", action="store_true",
                    help="run the walk-only leg (calc_optical_depth_cuda), print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library drives one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    if args.art_only:
        print(json.dumps(art_leg()), flush=True)
        return

    from argot_amd import ic
    from argot_amd import radiation as R
    from argot_amd._lib import check, lib

    t0 = time.perf_counter()
    w = ic.WORKLOADS[args.workload]()
    step = R.PointSourceStep(w, tau_mode=1 if args.kahan else 0)
    log(f"setup {time.perf_counter() - t0:.1f} s: {args.workload} n={w.n} nsrc={w.nsrc}")
    if args.chem_only:
        step.step()
        print(json.dumps(chem_leg(step)), flush=True)
        return
    if args.walk_only:
        print(json.dumps(walk_leg(step)[0]), flush=True)
        return

    for i in range(args.warmup):
        st = step.step()
        log(f"warmup {i}: {st['nray']} segments, kernel {st['ms_kernel']:.1f} ms, "
            f"total {st['ms_total']:.1f} ms")
    tuned = None
    if args.tile or args.blocks_per_cu:
        if args.tile:
            lib().argot_hip_set_tile(*[int(v) for v in args.tile.split("x")])
            check("argot_hip_set_tile")
        if args.blocks_per_cu:
            lib().argot_hip_set_blocks_per_cu(args.blocks_per_cu)
        tuned = {"tile": args.tile, "blocks_per_cu": args.blocks_per_cu, "candidates": 0}
        step.step()
    elif not args.no_tune:
        rows, best = step.autotune()
        tuned = {"tile": list(rows[best]["tile"]), "blocks_per_cu": rows[best]["blocks_per_cu"],
                 "candidates": len(rows),
                 "ms_kernel_range": [min(r["ms_kernel"] for r in rows),
                                     max(r["ms_kernel"] for r in rows)]}
        log(f"autotune: {tuned}")
        step.step()   # graph capture for the selected configuration
    ker = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = step.step()
        ker.append(st["ms_kernel"])
    wall = time.perf_counter() - t0
    ms_step = 1000.0 * wall / args.steps
    nray, ncross, ncell = st["nray"], st["ncross"], st["ntarget"]
    util = nray / max(64 * st["nemit_iter"], 1)
    log(f"timed: {args.steps} steps, {ms_step:.1f} ms/step, kernel {np.mean(ker):.1f} ms, "
        f"{ncross / max(nray, 1):.1f} crossings/segment, {st['nvisit']} tree visits, "
        f"lane utilisation {util:.3f}")

    kernel_ms = float(np.mean(ker))
    bytes_launch = algorithmic_bytes(nray, ncross, ncell)
    achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9
    # memory traffic of this kernel on this workload, from the committed
    # rocprofv3 FETCH_SIZE/WRITE_SIZE passes, when they were measured on
    # the kernel sources of this tree
    traffic, traffic_note = None, None
    if args.kahan:
        traffic_note = "profiled in plain-sum mode only"
    else:
        prof, traffic_note = profile_record("fused_step_kernel", args.workload)
        if prof:
            traffic = prof.get("traffic_bytes")
    copy_gbs = lib().argot_hip_copy_probe(1 << 30, 10)
    check("argot_hip_copy_probe")
    log(f"copy probe: {copy_gbs:.0f} GB/s (1 GiB copy, read + write)")
    seg_s = nray / (ms_step * 1e-3)
    rates_plain = step.rates() if not args.no_cpu_baseline else None
    out = {
        "metric": "ray segments/s + % HBM peak @256^3,4096 src; surrogate vs exact rate @1e-3",
        "metric_detail": "segments/s of the point-source step with exact rates; % HBM peak in roofline.frac; surrogate vs exact in surrogate",
        "value": seg_s,
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
                   "tau_sum": "kahan" if args.kahan else "plain", "hip_graph": True,
                   "tree_visits": int(st["nvisit"]), "emitting_visits": int(st["nemit_iter"]),
                   "lane_utilisation": util,
                   "tuned": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_note": traffic_note or "rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                                     "profiles/traffic.json",
                     "kernel": "fused_step_kernel", "kernel_ms": kernel_ms,
                     "bytes_per_launch": int(bytes_launch),
                     "measured_copy_gbs": copy_gbs},
    }
    rec_sample = None
    all_legs = ("kahan", "surrogate", "walk", "rate_call", "chemistry", "art", "lognormal128")
    legs = set() if args.no_extras or args.legs == "none" else (
        set(all_legs) if args.legs == "all" else set(args.legs.split(",")))
    if legs - set(all_legs):
        raise SystemExit(f"unknown legs {sorted(legs - set(all_legs))}")
    if "kahan" in legs and not args.kahan:
        out["kahan"] = step_leg(step, "kahan sums on the same workload", tau_mode=1)
    rates_surr = None
    if "surrogate" in legs:
        out["surrogate"], rates_surr = surrogate_leg(
            step, seg_s, nray, kernel_ms,
            rates_plain if rates_plain is not None else step.rates())
    if "walk" in legs or "rate_call" in legs:
        walk, rec = walk_leg(step)
        if "walk" in legs:
            out["walk_only"] = walk
        rec_sample = rec[np.linspace(0, len(rec) - 1, min(len(rec), 20000)).astype(np.int64)]
        if "rate_call" in legs:
            out["rate_call"] = rate_call_leg(step, rec)
        del rec
    if "chemistry" in legs:
        out["chemistry"] = chem_leg(step)
    if "art" in legs:
        out["art"] = art_leg()
        log(f"art: {out['art']}")
    if "lognormal128" in legs:
        out["lognormal128"] = lognormal128_leg()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w)
        log(f"cpu baseline: {out['cpu_baseline']}")
        out["fp64_error"] = fp64_error(w, rates_plain, rec_sample,
                                       others={"surrogate": rates_surr})
        log(f"fp64 error: {out['fp64_error']}")
    print(json.dumps(out), flush=True)


def timed_steps(step, nsteps):
    """nsteps timed calls after one untimed one (graph capture for the
    current settings): (ms per step, mean kernel ms, last stats)."""
    step.step()
    ker = []
    t0 = time.perf_counter()
    for _ in range(nsteps):
        st = step.step()
        ker.append(st["ms_kernel"])
    return 1000.0 * (time.perf_counter() - t0) / nsteps, float(np.mean(ker)), st


def roofline_of(st, kernel_ms):
    b = algorithmic_bytes(st["nray"], st["ncross"], st["ntarget"])
    a = b / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": a / HBM_PEAK_GBS, "bytes_per_launch": int(b)}


def step_leg(step, what, tau_mode, nsteps=3):
    """The headline step with other settings (e.g. Kahan sums), timed like it."""
    old = step.tau_mode
    step.tau_mode = tau_mode
    try:
        ms, kms, st = timed_steps(step, nsteps)
    finally:
        step.tau_mode = old
    out = {"what": what, "tau_sum": "kahan" if tau_mode else "plain", "steps": nsteps,
           "segments_per_s": st["nray"] / (ms * 1e-3), "ms_per_step": ms, "kernel_ms": kms,
           "segments": int(st["nray"]), "crossings": int(st["ncross"]),
           "roofline": roofline_of(st, kms)}
    log(f"{what}: {out}")
    return out


def lognormal128_leg(nsteps=3):
    """The 128^3 lognormal mesh with 1024 power-law sources and compensated
    optical-depth sums: the step, timed like the headline."""
    from argot_amd import ic
    from argot_amd import radiation as R
    w = ic.lognormal128()
    step = R.PointSourceStep(w, tau_mode=1)
    ms, kms, st = timed_steps(step, nsteps)
    out = {"workload": f"lognormal128: {w.n}^3 mesh, {w.nsrc} sources, kahan sums",
           "steps": nsteps, "segments_per_s": st["nray"] / (ms * 1e-3), "ms_per_step": ms,
           "kernel_ms": kms, "segments": int(st["nray"]), "crossings": int(st["ncross"]),
           "roofline": roofline_of(st, kms)}
    log(f"lognormal128: {out}")
    return out


if __name__ == "__main__":
    main()
