"""Throughput of the HIP tile correlator on one MI355X.

Workload: the scene-sized FP32 case -- 8192 x 4096 complex64 speckle, the
secondary resampled by a smooth field (range ramp 0-5 px, azimuth quadratic
up to 2 px) plus 15 dB noise; a 120 x 60 grid of pairs (6844 valid under the
plan geometry), chip 64, window 128, refine factor 2, refine margin 8, zoom 8.
A step is one adjust over all pairs with both rasters already resident in
device memory: gather + detect, chip statistics, FP64 SATs, the contraction,
normalization, peak/SNR/nudge, upsampling, refined correlation, zoom and
sub-pixel offsets, which stay in device memory.  The rate with the offsets
copied to the host and returned as the reference's tuple of (dLine, dSample)
is reported beside it (config.host_inclusive_pairs_per_sec).

Prints one JSON line.  The roofline figure is for the dominant kernel, timed
with HIP events on the plan's stream during the timed steps; the CPU baseline
is the FP64 oracle on a bounded sample of the same pairs.  The "serving"
object reports the continuous-batching server on a Poisson stream of 256 jobs
of mixed geometries (ampcor_amd/serving.py): p50/p99 job latency and pairs/s
at a fixed arrival rate, with every job submitted at once (the saturated
throughput), and at a fixed rate of 75 % of that throughput, for each form.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tile pairs/sec @64²/128² zoom8, mixed vs FP32 at ≤0.01px; p99 job latency"
FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 (matrix) peak
MIXED_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 / FP16 matrix peak
FP8_PEAK_TFLOPS = 5000.0    # MI355X dense FP8 matrix peak
HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "scene": "scene 8192x4096 complex64 speckle, smooth-field secondary +15 dB, 120x60 grid "
             "(6844 valid pairs), chip 64, window 128, refine x2 margin 8, zoom 8",
    "large": "8192x8192 float32 lognormal(0.5) texture * 5x5 Gaussian, smooth-field secondary "
             "(<= 12 px) x (1 + 0.1 N), 64x32 grid (1984 valid pairs), chip 128, window 256, "
             "refine x2 margin 8, zoom 16",
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="bf16",
                    help="form whose throughput is the headline value (fp32, bf16, fp16, fp8); "
                         "the FP32 drop-in form is always timed beside it")
    ap.add_argument("--extra-forms", default="fp16,bf16x3,fp8",
                    help="further forms timed beside FP32 and the headline (comma list, or '')")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--workload", choices=("scene", "large"), default="scene",
                    help="scene: 8192x4096 complex64, chip 64 / window 128, zoom 8 (default); "
                         "large: 8192x8192 float32 amplitude, chip 128 / window 256, zoom 16")
    ap.add_argument("--serve-jobs", type=int, default=256,
                    help="jobs in the serving stream (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=100.0,
                    help="Poisson arrival rate of the serving stream, jobs/s")
    return ap.parse_args()


def cpu_baseline(scene, seconds):
    """FP64 oracle on the host cores over a bounded sample of the same pairs."""
    from oracle import pyoracle

    affinity = len(os.sched_getaffinity(0))
    # the host's CPU share: OMP_NUM_THREADS where the environment sets it (the
    # GPU hosts give each GPU a fixed share of cores), else every core this
    # process may run on
    cores = int(os.environ.get("OMP_NUM_THREADS") or affinity)
    args = (scene.reference, scene.target)
    geo = (scene.chip, scene.window, scene.refine_factor, scene.refine_margin, scene.zoom_factor)

    def run(n, threads, form="fp64"):
        idx = np.unique(np.linspace(0, scene.pairs - 1, n).astype(int))
        t0 = time.perf_counter()
        out = pyoracle.adjust(*args, scene.ref_begin[idx], scene.tgt_begin[idx], *geo,
                              form=form, threads=threads)
        return time.perf_counter() - t0, idx, out

    n = 2 * cores
    dt, _, _ = run(n, cores)
    n2 = int(min(scene.pairs, max(n, n * seconds / max(dt, 1e-3))))
    n2 = max(cores, (n2 // cores) * cores)
    dt, idx, exact = run(n2, cores)
    # the same oracle on one core (the reference's CPU path is single-threaded),
    # on a sample of about a third of the multi-core run's duration
    per_pair_1 = dt * cores / len(idx)
    n1 = int(max(4, min(len(idx), seconds / 3 / max(per_pair_1, 1e-6))))
    dt1, idx1, _ = run(n1, 1)
    # the FP32 twin (the GPU coarse stage's rounding, FP64 refinement) on a
    # quarter of the sample: the CPU-FP32 all-core rate
    dt_t, idx_t, twin = run(max(cores, len(idx) // 4), cores, form="fp32")
    return ({"value": len(idx) / dt, "unit": "tile pairs/sec", "cores": cores, "kind": "port",
             "affinity_cores": affinity,
             "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
             "sample": f"FP64 oracle (oracle/ampcor_oracle.c, OpenMP over pairs) on {len(idx)} "
                       f"of the {scene.pairs} pairs, evenly spaced, {dt:.1f} s, {cores} threads",
             "single_core": {"value": len(idx1) / dt1, "cores": 1,
                             "sample": f"the same oracle on one core, {len(idx1)} pairs, "
                                       f"{dt1:.1f} s"},
             "fp32_all_core": {"value": len(idx_t) / dt_t, "cores": cores,
                               "sample": f"the FP32 twin (FP32 coarse stage, FP64 refinement) on "
                                         f"{len(idx_t)} pairs, {dt_t:.1f} s"}},
            idx, exact, idx_t, twin)


def covered_pixels(scene):
    """Raster pixels the gathers must read at least: the union of the chips on
    the reference and of the windows on the target (neighbouring windows
    overlap, and each pixel need come from HBM once)."""
    cached = getattr(scene, "_covered", None)
    if cached is None:
        def union(shape, begin, size):
            mask = np.zeros(shape[:2], dtype=bool)
            for line, sample in begin:
                mask[line:line + size, sample:sample + size] = True
            return int(mask.sum())
        cached = (union(scene.reference.shape, scene.ref_begin, scene.chip),
                  union(scene.target.shape, scene.tgt_begin, scene.window))
        scene._covered = cached
    return cached


def stage_work(precision, scene, dtype):
    """Algorithmic work of one launch of each stage: (bound, amount, unit).

    HBM stages count the bytes they must move at least (each operand read once,
    each product written once, per pair; the gathers read each raster pixel the
    chips and windows cover once); the contractions count 2 flops per
    multiply-add of the NCC numerator over all placements.  The fused stages
    carry the work of the reference kernels they absorb: gather_prep = gather +
    detect (detect.cu) + chip statistics (refStats.cu) + quantization;
    normalize = window sums (sat.cu, tgtStats.cu) + normalization; the
    upsamplers = migrate (migrate.cu) + x2 upsampling; coarse_peak = maxcor.cu
    + nudge.cu.
    """
    N, C, W = scene.pairs, scene.chip, scene.window
    rf, rm = scene.refine_factor, scene.refine_margin
    P, E, R = W - C + 1, C + 2 * rm, 2 * rf * rm + 1
    Cr, Er = rf * C, rf * (C + 2 * rm)
    eb = 8 if dtype == "complex64" else 4            # raster element bytes
    split = precision.endswith("x3")
    qc = {"fp32": 4, "fp8": 1}.get(precision, 2)      # coarse operand bytes
    qr = 4 if precision == "fp32" else 2              # refined operand bytes (fp8: FP16)
    hbm = lambda b: ("hbm", float(b) * N, "B")
    mfma = "valu" if precision == "fp32" else "mfma"
    ref_px, tgt_px = covered_pixels(scene)
    raster_read = eb * (ref_px + tgt_px)                # unique raster bytes, whole launch
    win_r = (4 * Er * Er + 2 * qr * Er * Er) if split else qr * Er * Er
    return {
        "gather_detect": ("hbm", raster_read + 4.0 * (C * C + W * W) * N, "B"),
        "chip_stats": hbm(8 * C * C),
        "window_sums": hbm(4 * W * W + 16 * P * P),
        "gather_prep": ("hbm", raster_read + float(qc) * (C * C + W * W) * N, "B"),
        "correlate": (mfma, 2.0 * C * C * P * P * N, "FLOP"),
        "normalize": hbm(24 * P * P if precision == "fp32" else qc * W * W + 8 * P * P),
        "coarse_peak": hbm(4 * P * P),
        "upsample_ref": hbm(eb * C * C + (2 if split else 1) * qr * Cr * Cr),
        "upsample_tgt": hbm(eb * E * E + (4 * Er * Er if precision == "fp32" else win_r)),
        "chip_stats_refined": hbm(8 * Cr * Cr),
        "correlate_refined": (mfma, 2.0 * Cr * Cr * R * R * N, "FLOP"),
        "normalize_refined": hbm((4 * Er * Er if precision == "fp32" or split else qr * Er * Er)
                                 + 8 * R * R),
        "zoom": hbm(4 * R * R + 16),
    }


def stage_roofline(precision, scene, dtype, stages):
    """Per-stage achieved rate against MI355X's published peak."""
    peaks = {"hbm": (HBM_PEAK_GBS, "GB/s", 1e9), "valu": (FP32_PEAK_TFLOPS, "TFLOP/s", 1e12),
             "mfma": (FP8_PEAK_TFLOPS if precision == "fp8" else MIXED_PEAK_TFLOPS,
                      "TFLOP/s", 1e12)}
    work = stage_work(precision, scene, dtype)
    out = {}
    for name, ms in stages.items():
        if name not in work:
            continue
        bound, amount, unit = work[name]
        peak, punit, scale = peaks[bound]
        if name == "correlate_refined" and bound == "mfma":
            peak = MIXED_PEAK_TFLOPS          # the FP8 form refines in FP16
        achieved = amount / (ms * 1e-3) / scale
        out[name] = {"ms": round(ms, 4), "bound": bound, "work": amount, "work_unit": unit,
                     "achieved": round(achieved, 2), "unit": punit, "peak": peak,
                     "frac": round(achieved / peak, 4)}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); pairs are independent, so "
              "several GPUs would each run this workload on their share of the tile grid",
              file=sys.stderr)
        sys.exit(2)

    from ampcor_amd import ext, synth

    lib = ext.require()
    if lib.device_count() < 1:
        raise RuntimeError("no HIP device visible")

    t0 = time.perf_counter()
    if a.workload == "large":
        scene = synth.amplitude_scene()
        if a.precision == "bf16":
            a.precision = "fp16"
    else:
        scene = synth.field_scene()
    print(f"[bench] scene ready in {time.perf_counter() - t0:.1f} s: {scene.pairs} pairs",
          file=sys.stderr, flush=True)
    N, C, W = scene.pairs, scene.chip, scene.window
    dtype = "complex64" if scene.reference.dtype == np.complex64 else "float32"
    ref = lib.raster(scene.reference, *scene.reference.shape, dtype=dtype)
    tgt = lib.raster(scene.target, *scene.target.shape, dtype=dtype)

    def timed(precision):
        worker = lib.sequential(N, (C, C), (W, W), scene.refine_margin, scene.refine_factor,
                                scene.zoom_factor, precision=precision)
        lib.addPairs(worker, ref, tgt, 0, scene.ref_begin.tobytes(), scene.tgt_begin.tobytes())
        for _ in range(a.warmup):
            lib.adjust(worker)
        # setup, outside the timed steps: the plan autotuner picks the tilings
        # of both contractions on this workload's data (it times them on the
        # data of an adjust, so one runs first even without warm-up)
        if a.warmup == 0:
            lib.adjust(worker)
        lib.autotune(worker, 3)
        lib.adjust(worker)
        lib.set_timing(worker, True)
        stage_sum = {}
        elapsed = 0.0
        for _ in range(a.steps):
            t = time.perf_counter()
            lib.adjust(worker, offsets=False)   # synchronizes the plan stream before returning
            elapsed += time.perf_counter() - t
            for k, v in lib.stage_times(worker).items():
                stage_sum[k] = stage_sum.get(k, 0.0) + v
        stages = {k: v / a.steps for k, v in stage_sum.items()}
        # the same steps with the offsets copied to the host and built into the
        # reference's tuple (not the headline: the products stay on the device there)
        lib.set_timing(worker, False)
        lib.adjust(worker)                      # warm-up: first use of the host staging
        t = time.perf_counter()
        for _ in range(a.steps):
            lib.adjust(worker)
        host_step = (time.perf_counter() - t) / a.steps
        prods = {k: np.frombuffer(v, np.float32).copy() for k, v in lib.products(worker).items()}
        prods["offsets"] = prods["offsets"].reshape(N, 2)
        prods["guard"] = lib.fp8_guard(worker)
        prods["zoom_screen"] = lib.zoom_stats(worker)
        prods["coarse_loc"] = np.frombuffer(lib.stage_locations(worker)["coarse"],
                                            np.int32).reshape(N, 2).copy()
        print(f"[bench] {precision} stage ms/step: " +
              ", ".join(f"{k} {v:.3f}" for k, v in stages.items()), file=sys.stderr, flush=True)
        prods["host_step"] = host_step
        return elapsed / a.steps, stages, prods

    forms = ["fp32"] if a.precision == "fp32" else ["fp32", a.precision]
    for f in (x.strip() for x in a.extra_forms.split(",")):
        if f and f not in forms:
            forms.append(f)
    runs = {f: timed(f) for f in forms}
    step_s, stages, prods = runs[a.precision]
    off = prods["offsets"]
    ms_per_step = 1e3 * step_s
    value = N / step_s
    fp32_value = N / runs["fp32"][0]
    err = np.abs(off - scene.truth)
    print(f"[bench] {a.precision}: {value:.0f} pairs/s ({ms_per_step:.3f} ms/step); FP32 form "
          f"{fp32_value:.0f} pairs/s; |offset - true field| median {np.median(err):.4f} px, "
          f"p99 {np.percentile(err, 99):.4f} px", file=sys.stderr, flush=True)

    # dominant kernel: the larger of the two contractions
    P, R = W - C + 1, 2 * scene.refine_factor * scene.refine_margin + 1
    Cr = scene.refine_factor * C
    flops = {"correlate": 2.0 * C * C * P * P * N, "correlate_refined": 2.0 * Cr * Cr * R * R * N}
    dom = max(flops, key=lambda k: stages[k])
    achieved = flops[dom] / (stages[dom] * 1e-3) / 1e12
    # HBM bytes per launch of that kernel, from the committed rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE passes (copy-calibrated) on this same workload
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_scene8192x4096.json" if a.workload == "scene"
                         else "traffic_large8192x8192.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch", {}).get(a.precision, {}).get(dom)
    peak = FP32_PEAK_TFLOPS if a.precision == "fp32" else MIXED_PEAK_TFLOPS
    if a.precision == "fp8" and dom == "correlate":
        peak = FP8_PEAK_TFLOPS
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                "traffic": traffic, "kernel": dom, "kernel_ms": round(stages[dom], 4)}

    cpu = None
    accuracy = None
    if not a.no_cpu:
        cpu, idx, exact, idx_t, twin = cpu_baseline(scene, a.cpu_seconds)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)
        accuracy = {f"{f}_max_offset_err_px_vs_fp64":
                    float(np.max(np.abs(runs[f][2]["offsets"][idx] - exact["offsets"])))
                    for f in forms}
        rel = lambda x, y: float(np.max(np.abs(x.astype(np.float64) - y) / np.abs(y)))
        for f in forms:
            p = runs[f][2]
            accuracy[f"{f}_coarse_loc_equal_fp64"] = int(
                np.all(p["coarse_loc"][idx] == exact["coarse_loc"], axis=1).sum())
            accuracy[f"{f}_coarse_peak_max_rel_err_vs_fp64"] = rel(p["coarse_peak"][idx],
                                                                   exact["coarse_peak"])
            accuracy[f"{f}_snr_max_rel_err_vs_fp64"] = rel(p["snr"][idx], exact["snr"])
            accuracy[f"{f}_refined_peak_max_rel_err_vs_fp64"] = rel(p["refined_peak"][idx],
                                                                    exact["refined_peak"])
        if "fp8" in runs:
            accuracy["fp8_peak_guard"] = runs["fp8"][2]["guard"]
        accuracy["zoom_screen_exact_cells"] = {f: runs[f][2]["zoom_screen"] for f in forms}
        accuracy["fp32_coarse_loc_equal_fp32_twin"] = int(
            np.all(runs["fp32"][2]["coarse_loc"][idx_t] == twin["coarse_loc"], axis=1).sum())
        accuracy["fp32_coarse_peak_bitwise_equal_fp32_twin"] = int(
            (runs["fp32"][2]["coarse_peak"][idx_t].view(np.uint32)
             == twin["coarse_peak"].view(np.uint32)).sum())
        accuracy["pairs_compared"] = int(len(idx))
        accuracy["pairs_compared_fp32_twin"] = int(len(idx_t))
        print(f"[bench] accuracy vs FP64 oracle: {accuracy}", file=sys.stderr, flush=True)

    serving = None
    if a.serve_jobs > 0 and a.workload == "scene":
        from ampcor_amd import serving as serve
        log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
        t = time.perf_counter()
        pool = serve.raster_pool()
        log(f"serving raster pool ready in {time.perf_counter() - t:.1f} s")
        serving = {}
        for f in forms[:2]:
            serving[f] = serve.serve_benchmark(precision=f, count=a.serve_jobs, rate=a.serve_rate,
                                               pool=pool, log=log)
            log(f"serving {f}: {serving[f]}")

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "tile pairs/sec", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if a.precision == "fp32" else a.precision, "data": "synthetic",
        "config": {"workload": WORKLOADS[a.workload],
                   "pairs": N, "chip": C, "window": W, "refine_factor": scene.refine_factor,
                   "refine_margin": scene.refine_margin, "zoom_factor": scene.zoom_factor,
                   "precision": a.precision, "fp32_pairs_per_sec": round(fp32_value, 1),
                   "forms_pairs_per_sec": {f: round(N / runs[f][0], 1) for f in forms},
                   "speedup_vs_fp32": round(value / fp32_value, 3),
                   "host_inclusive_pairs_per_sec": round(N / prods["host_step"], 1),
                   "accuracy": accuracy},
        "roofline": roofline,
        "stages": {f: stage_roofline(f, scene, dtype, runs[f][1]) for f in forms[:2]},
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
