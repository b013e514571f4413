#!/usr/bin/env python3
"""Stage-2 gamma-estimation benchmark on one MI355X.

Metric: PS-candidate x interferogram x gamma-iteration per second of the
stage-2 gamma loop (first fill .. final stop decision, then the write_pm1
outputs of the last iteration in the caller's layouts), with the patch's
inputs already resident in HBM.  A step is one whole gamma loop over the
patch (to convergence or gamma_max_iterations) with its outputs.

Default workload: 500,000 candidates in 15x15 km, 60 interferograms, bperp
U[-250, 250] m, 40 % coherent pixels, StaMPS defaults (50 m grid => 299x299
cells, 2,209 CLAP windows per interferogram, 13 trials), 3 iterations max.
The FP32 form is timed; the plain FP64 GPU form and the CPU restatement of the
reference (the oracle in its reference-precision mode, all host cores) are
timed beside it.

    python bench.py [--steps K] [--warmup W] [--config 2] [--no-cpu] [--no-fp64]
"""
from __future__ import annotations

import argparse
import json
import tempfile
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# rocprofv3 --pmc on a plain 128 MiB device copy reads FETCH_SIZE = 65,548 KiB
# and WRITE_SIZE = 131,072 KiB on gfx950: HBM bytes = 2.0 x FETCH + 1.0 x WRITE
FETCH_SCALE = 2.0
WRITE_SCALE = 1.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def clap_windows(n):
    return max(0, math.ceil(n / 6) - 3)


def kernel_work(prep, name):
    """Algorithmic work per launch: (flops, HBM bytes).

    CLAP: 143,360 flop per window and interferogram (forward + inverse 32x32
    FFT2 at 2 * 5 N^2 log2 N^2 = 102,400, 7-tap separable smoothing 28,672,
    pointwise gain 12,288); bytes = grid read + the partial planes written
    (each window row writes its 24 output rows: ~4x the grid, complex64); the
    combine reads those planes and writes the filtered grid.  Trial scan: one
    complex multiply-add (8 flop) per pixel, interferogram and trial; bytes =
    ph_nm (8 B) + the unit trial step per pixel and interferogram (8 B from the
    step table, or 4 B of bperp where the 8-trial scan of a large patch
    computes the step), the filtered grid read once.  Refine: two passes over
    ph_nm + bperp (12 B) and the filtered grid.  Fill: ph_nm + bperp read,
    grid written.  With row-invariant bperp the per-element bperp / step bytes
    drop out (one row, cache resident)."""
    P, I, T = prep.n_ps, prep.n_ifg, len(prep.trial_values)
    C = prep.n_i * prep.n_j
    rowinv = bool(prep.row_invariant_bperp)
    bpb = 0.0 if rowinv else 4.0                                   # bperp bytes per element
    stepb = 0.0 if rowinv else (4.0 if (T - 1) // 2 < 8 and P >= 65536 else 8.0)
    W = clap_windows(prep.n_i) * clap_windows(prep.n_j)
    part = 8.0 * clap_windows(prep.n_i) * 24 * prep.n_j * I
    if name == "clap_strip":
        return 143_360.0 * W * I, 8.0 * C * I + part
    if name == "scan":
        return 8.0 * P * I * T, (8.0 + stepb) * P * I + 8.0 * C * I
    if name == "refine":
        return 2 * 20.0 * P * I, 2 * (8.0 + bpb) * P * I + 8.0 * C * I
    if name == "fill_grid":
        return 16.0 * P * I, (8.0 + bpb) * P * I + 8.0 * C * I
    if name == "clap_combine":
        return 0.0, part + 8.0 * C * I
    if name == "weights":
        return 0.0, 24.0 * P
    if name == "post":   # coh, coh_prev read, coh_prev written, re-decided rows
        return 0.0, 24.0 * P
    if name == "outputs":
        # once per loop: ph_nm + bperp read, ph_weight + ph_patch + ph_res
        # written (8 + 8 + 4 B) in row order, the filtered grid read, the grid
        # transposed (read + write), five f64 vectors and the trial index
        return 0.0, (8.0 + bpb + 20.0) * P * I + 24.0 * C * I + 92.0 * P
    return 0.0, 0.0


def aggregate_roofline(prep, kt, peak_tflops):
    """Time-weighted roofline over the loop's kernels: sum of per-kernel ideal
    times max(F / peak_VALU, B / peak_HBM) over the sum of measured times."""
    ideal = meas = 0.0
    for name, (ms, n) in kt.items():
        if not n:
            continue
        f, b = kernel_work(prep, name)
        ideal += max(f / (peak_tflops * 1e12), b / (PEAK_HBM_GBS * 1e9)) * n
        meas += ms * 1e-3
    return ideal / meas if meas > 0 else None


KERNEL_SYMBOLS = {"clap_strip": "k_clap_strip", "scan": "k_scan", "refine": "k_refine",
                  "fill_grid": "k_fill_grid", "clap_combine": "k_clap_combine",
                  "rescore": "k_rescore", "weights": "k_weights", "outputs": "k_final_outputs",
                  "post": "k_post"}


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "stage2_fp32_500k_60ifg_pmc_traffic.md")
STATS_SUMMARY = os.path.join(ROOT, "profiles", "stage2_fp32_500k_60ifg_kernel_stats.md")


def summary_avg_us(path, kernel):
    """Average launch duration (us) of `kernel` in a committed rocprofv3
    --kernel-trace --stats summary table (| kernel | calls | total ms | avg us | ...)."""
    sym = KERNEL_SYMBOLS.get(kernel, kernel)
    try:
        with open(path) as fh:
            for line in fh:
                cols = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
                if len(cols) >= 4 and sym in cols[0]:
                    try:
                        return float(cols[3])
                    except ValueError:
                        continue
    except OSError:
        pass
    return None


def cpu_threads():
    """Host threads for the CPU baseline: every core this process may run on
    (its affinity set), within the job's CPU share when the host sets one
    through OMP_NUM_THREADS."""
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def summary_traffic(path, kernel):
    """Calibrated HBM bytes per launch of `kernel` from a committed rocprofv3
    --pmc summary table (| symbol | FETCH_SIZE KiB | WRITE_SIZE KiB | ...)."""
    sym = KERNEL_SYMBOLS.get(kernel, kernel)
    try:
        with open(path) as fh:
            for line in fh:
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cols) >= 3 and sym in cols[0]:
                    try:
                        return FETCH_SCALE * float(cols[1]) * 1024.0 + WRITE_SCALE * float(cols[2]) * 1024.0
                    except ValueError:
                        continue
    except OSError:
        pass
    return None


def copy_calibration():
    """A plain 128 MiB device-to-device copy, three times: the reference
    kernel for scaling FETCH_SIZE / WRITE_SIZE to HBM bytes."""
    import torch
    n = 128 * 1024 * 1024 // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": a.numel() * 4, "copies": 3}), flush=True)


def pmc_traffic(pmc_dir, kernel):
    """HBM bytes per dispatch of `kernel` from rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE counter CSVs found under pmc_dir (calibrated)."""
    import csv
    import glob
    sym = KERNEL_SYMBOLS.get(kernel, kernel)
    tot = {"FETCH_SIZE": [], "WRITE_SIZE": []}
    for f in glob.glob(os.path.join(pmc_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if sym in r["Kernel_Name"] and r["Counter_Name"] in tot:
                    tot[r["Counter_Name"]].append(float(r["Counter_Value"]))
    if not tot["FETCH_SIZE"] or not tot["WRITE_SIZE"]:
        return None
    fetch = sum(tot["FETCH_SIZE"]) / len(tot["FETCH_SIZE"]) * 1024.0
    write = sum(tot["WRITE_SIZE"]) / len(tot["WRITE_SIZE"]) * 1024.0
    return FETCH_SCALE * fetch + WRITE_SCALE * write


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5],
                    help="3 = 16 uneven patches in one batched call")
    ap.add_argument("--n-ps", type=int, default=0, help="override the candidate count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--nr-rows", type=int, default=300_000)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--pmc-dir", default="", help="rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE outputs "
                    "of this workload, for the roofline traffic field")
    ap.add_argument("--autotune-cache", default=os.path.join(tempfile.gettempdir(), "pystamps_hip_tiles.txt"),
                    help="FP32 tile cache (measured on a miss); 'none' keeps the built-in tile model")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the plain-copy kernel that calibrates FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs stage 2 on one GPU; --gpus must be 1 "
                                   "(multi-GPU patch sharding is a design note, not built)"}))
        sys.exit(2)

    import numpy as np
    from pystamps_amd import synth
    from pystamps_amd.prepare import prepare_stage2_inputs
    from pystamps_amd.stage2 import Stage2Session

    if args.config == 3:
        return bench_batch(args)
    cfg = synth.CONFIGS[args.config]
    if args.n_ps:
        cfg = synth.scaled(cfg, args.n_ps)
    t = time.perf_counter()
    raw = synth.generate(cfg)
    log(f"[bench] generated {cfg.name}: {cfg.n_ps} x {cfg.n_ifg} in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    prep = prepare_stage2_inputs(raw, nr_rows=args.nr_rows)
    prep_s = time.perf_counter() - t
    del raw
    P, I = prep.n_ps, prep.n_ifg
    log(f"[bench] prep (incl. GPU Nr, {args.nr_rows} rows) {prep_s:.2f}s grid {prep.n_i}x{prep.n_j} "
        f"trials {len(prep.trial_values)}")

    cache = None if args.autotune_cache == "none" else args.autotune_cache

    def time_form(form, steps, warmup):
        t0 = time.perf_counter()
        s = Stage2Session(prep, form=form, autotune_cache=cache if form == "fp32" else None)
        setup = time.perf_counter() - t0
        tiles = s.tiles()
        log(f"[bench] {form} tiles {tiles}")
        for _ in range(warmup):
            s.run()
        # timed loops: the default launch path (one iteration captured as a
        # graph, one host wait per loop)
        iters = 0
        dev_s = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            i_loop, secs = s.run()
            iters += i_loop
            dev_s += secs
        wall = time.perf_counter() - t0
        # per-kernel breakdown from separate loops whose kernels are launched
        # one by one between HIP event pairs (the same kernels, every
        # iteration queued before one host wait, so the device runs them back
        # to back as in the timed loops)
        s.set_timing(True)
        n_timing = max(1, min(steps, 3))
        for _ in range(n_timing):
            s.run()
        kt = s.kernel_times()
        s.set_timing(False)
        t0 = time.perf_counter()
        out = s.download()
        download = time.perf_counter() - t0
        s.close()
        log(f"[bench] {form}: {steps} loops, {iters} iterations, wall {wall * 1e3:.2f} ms, "
            f"device {dev_s * 1e3:.2f} ms, setup {setup:.2f}s")
        for name, (ms, n) in kt.items():
            if n:
                log(f"[bench]   {name:13s} {ms / n * 1e3:9.1f} us/launch x {n}")
        # device time per loop of the timed (graph-launched) loops over the sum
        # of the per-kernel event times per timing loop (the control step,
        # the starting state and the control-block write are not timed per
        # kernel; they are ~20 us of a loop)
        ksum = sum(ms for ms, n in kt.values()) * 1e-3 / n_timing
        return {"wall": wall, "iters": iters, "dev": dev_s, "kt": kt, "out": out, "setup": setup,
                "download": download, "tiles": tiles,
                "graph_over_kernel_sum": (dev_s / steps) / ksum if ksum > 0 else None}

    r32 = time_form("fp32", args.steps, args.warmup)
    e2e = prep_s + r32["setup"] + r32["wall"] / args.steps + r32["download"]
    value = P * I * r32["iters"] / r32["wall"]
    ms_per_step = r32["wall"] / args.steps * 1e3

    # roofline of the dominant kernel (largest summed device time)
    kt = r32["kt"]
    dom = max((k for k in kt if kt[k][1] > 0), key=lambda k: kt[k][0])
    avg_s = kt[dom][0] / kt[dom][1] * 1e-3
    flops, nbytes = kernel_work(prep, dom)
    if flops / max(nbytes, 1.0) > PEAK_FP32_TFLOPS * 1e3 / PEAK_HBM_GBS:
        roof = {"bound": "valu", "engine": "fp32 vector ALU (no matrix op on this path)",
                "achieved": flops / avg_s / 1e12, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": nbytes / avg_s / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    if roof["unit"] == "TFLOP/s":
        # the library is built without packed FP32 instructions: one FP32
        # FMA per lane per cycle, half the published packed peak
        roof["frac_of_unpacked_fp32_peak"] = roof["achieved"] / (PEAK_FP32_TFLOPS / 2)
    if args.pmc_dir:
        roof["traffic"] = pmc_traffic(args.pmc_dir, dom)
        roof["traffic_source"] = "rocprofv3 --pmc run given by --pmc-dir"
    elif args.config == 2 and not args.n_ps:
        roof["traffic"] = summary_traffic(TRAFFIC_SUMMARY, dom)
        roof["traffic_source"] = os.path.relpath(TRAFFIC_SUMMARY, ROOT) if roof["traffic"] else None
    else:
        roof["traffic"] = None
    roof["kernel"] = dom
    roof["avg_launch_us"] = avg_s * 1e6
    roof["avg_source"] = ("HIP events on the session stream, per launch, over loops launched "
                          "kernel by kernel after the timed loops, all iterations queued before "
                          "one host wait")
    if args.config == 2 and not args.n_ps:
        pavg = summary_avg_us(STATS_SUMMARY, dom)
        if pavg:
            roof["profile_avg_launch_us"] = pavg
            roof["profile_source"] = os.path.relpath(STATS_SUMMARY, ROOT)
            roof["frac_from_profile"] = (flops / (pavg * 1e-6) / 1e12 / PEAK_FP32_TFLOPS
                                         if roof["unit"] == "TFLOP/s"
                                         else nbytes / (pavg * 1e-6) / 1e9 / PEAK_HBM_GBS)
    roof["loop_time_weighted_frac"] = aggregate_roofline(prep, kt, PEAK_FP32_TFLOPS)
    roof["timed_loop_device_over_kernel_event_sum"] = r32["graph_over_kernel_sum"]

    result = {
        "metric": "stage-2 PS·ifg·iter/s, FP32+batched vs FP64 GPU at matched K/gamma error",
        "value": value,
        "unit": "PS·ifg·iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, random-init patch of the named workload)",
        "config": {"workload": cfg.name, "n_ps": P, "n_ifg": I, "grid": [prep.n_i, prep.n_j],
                   "clap_windows_per_ifg": clap_windows(prep.n_i) * clap_windows(prep.n_j),
                   "n_trial": len(prep.trial_values), "iterations_per_step": r32["iters"] // args.steps,
                   "form": "fp32 bulk + fp64 near-tie re-score and P-square re-decision",
                   "tiles": r32["tiles"]},
        "roofline": roof,
        "rescored_rows_per_loop": r32["out"]["n_rescored"],
        "rechecked_rows_per_loop": r32["out"]["n_rechecked"],
        "end_to_end_seconds": {"total": e2e, "prep_incl_gpu_nr": prep_s, "upload": r32["setup"],
                               "loop": r32["wall"] / args.steps, "download": r32["download"]},
        "kernel_us_per_launch": {k: round(v[0] / v[1] * 1e3, 1) for k, v in r32["kt"].items() if v[1]},
    }

    if not args.no_fp64:
        r64 = time_form("fp64", args.fp64_steps, 1)
        v64 = P * I * r64["iters"] / r64["wall"]
        o32, o64 = r32["out"], r64["out"]
        agree = float(np.mean(o32["trial_ix"] == o64["trial_ix"]))
        dcoh = np.abs(o32["coh_ps"] - o64["coh_ps"])
        ok = np.isfinite(o32["K_ps"]) & np.isfinite(o64["K_ps"])
        dk = np.abs(o32["K_ps"][ok] - o64["K_ps"][ok])
        dc = np.abs(np.angle(np.exp(1j * (o32["C_ps"] - o64["C_ps"]))))
        dpp = np.abs(np.angle(o32["ph_patch"] * np.conj(o64["ph_patch"])))
        result["fp64_plain"] = {"value": v64, "ms_per_step": r64["wall"] / args.fp64_steps * 1e3,
                                "fp32_speedup": value / v64,
                                "trial_ix_agreement": agree,
                                "coh_abs_err_max": float(dcoh.max()),
                                "coh_abs_err_p999": float(np.quantile(dcoh, 0.999)),
                                "coh_frac_above_1e-5": float(np.mean(dcoh > 1e-5)),
                                "K_abs_err_p999": float(np.quantile(dk, 0.999)) if dk.size else None,
                                "C_abs_err_p999": float(np.quantile(dc, 0.999)),
                                "ph_patch_phase_err_p999": float(np.quantile(dpp, 0.999)),
                                "rescored_rows_per_loop": r32["out"]["n_rescored"],
                                "i_loop_equal": bool(o32["i_loop"] == o64["i_loop"])}

    if not args.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads()
        # one warm-up, then the median of args.cpu_runs timed loops
        O.run(prep, mode=O.REF32, threads=threads, diagnostics=False)
        runs = []
        for _ in range(args.cpu_runs):
            t0 = time.perf_counter()
            ro = O.run(prep, mode=O.REF32, threads=threads, diagnostics=False)
            runs.append(time.perf_counter() - t0)
        cpu_s = float(np.median(runs))
        cpu_v = P * I * ro["i_loop"] / cpu_s
        result["cpu_baseline"] = {
            "value": cpu_v, "unit": "PS·ifg·iter/s", "cores": threads, "kind": "port",
            "cores_visible": len(os.sched_getaffinity(0)), "host_cpus": os.cpu_count(),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
            "sample": f"one whole gamma loop ({ro['i_loop']} iterations) of the same {P} x {I} "
                      f"workload, median of {args.cpu_runs} after a warm-up; oracle in the "
                      f"reference's precision map, OpenMP over rows/ifgs",
            "seconds": cpu_s,
            "cpu": _cpu_model(),
        }
        result["gpu_vs_cpu"] = value / cpu_v
        log(f"[bench] cpu ref32: {cpu_s:.2f}s (median of {args.cpu_runs}) for {ro['i_loop']} "
            f"iterations on {threads} threads")
    print(json.dumps(result), flush=True)


def bench_batch(args):
    """16 uneven patches: all sessions' loops run concurrently (one call), timed
    against the same sessions run one after another."""
    import numpy as np
    from pystamps_amd import synth
    from pystamps_amd.prepare import prepare_stage2_inputs
    from pystamps_amd.stage2 import Stage2Session, run_sessions

    cfgs = synth.c3_patches()
    if args.n_ps:
        f = args.n_ps / sum(c.n_ps for c in cfgs)
        cfgs = [synth.scaled(c, max(2000, int(c.n_ps * f))) for c in cfgs]
    preps = []
    t = time.perf_counter()
    for c in cfgs:
        raw = synth.generate(c)
        preps.append(prepare_stage2_inputs(raw, nr_rows=args.nr_rows))
        del raw
    log(f"[bench] generated + prepared {len(preps)} patches, {sum(p.n_ps for p in preps)} PS "
        f"in {time.perf_counter() - t:.1f}s")
    ss = [Stage2Session(p, form="fp32") for p in preps]
    work = np.array([p.n_ps * p.n_ifg for p in preps], np.float64)
    for _ in range(args.warmup):
        run_sessions(ss)
    t0 = time.perf_counter()
    tot = 0.0
    for _ in range(args.steps):
        loops, _ = run_sessions(ss)
        tot += float(np.dot(work, loops))
    wall = time.perf_counter() - t0
    value = tot / wall
    t0 = time.perf_counter()
    tot_seq = 0.0
    for _ in range(args.steps):
        for k, s in enumerate(ss):
            il, _ = s.run()
            tot_seq += work[k] * il
    wall_seq = time.perf_counter() - t0
    for s in ss:
        s.close()
    result = {
        "metric": "stage-2 PS·ifg·iter/s, FP32+batched vs FP64 GPU at matched K/gamma error",
        "value": value, "unit": "PS·ifg·iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator, 16 uneven patches)",
        "config": {"workload": "batched uneven patches", "n_patches": len(preps),
                   "n_ps": [p.n_ps for p in preps], "n_ifg": preps[0].n_ifg,
                   "form": "fp32 bulk + fp64 near-tie re-score, patches concurrent"},
        "sequential": {"value": tot_seq / wall_seq, "ms_per_step": wall_seq / args.steps * 1e3},
        "batch_speedup_vs_sequential": value / (tot_seq / wall_seq),
    }
    print(json.dumps(result), flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
