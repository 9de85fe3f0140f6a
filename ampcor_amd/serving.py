"""A stream of offset-tracking jobs through the continuous-batching server.

Workload: jobs arrive as a Poisson process at a fixed rate.  Each job
correlates its own raster pair (drawn from a pool of speckle pairs from
1024 x 1024 to 8192 x 4096, each secondary resampled by its own smooth offset
field of up to 8 px plus noise) with a chip edge from {32, 64, 128}, a
search margin from {16, 32, 64}, a pair count log-uniform in [64, 4096],
refine factor 2, refine margin 8 and zoom 8; pair centers are uniform over
the raster area the window fits in.  Latency is measured by the server from
submission to the moment the job's last batch finishes.  The benchmark runs
the stream three times on one server: at the given rate, with every job
submitted at once (its throughput is the saturated rate), and at a fixed
fraction of that saturated rate.
"""

import threading
import time
from dataclasses import dataclass

import numpy as np

from . import ext, synth

CHIPS = (32, 64, 128)
MARGINS = (16, 32, 64)
POOL_SHAPES = ((1024, 1024), (2048, 2048), (4096, 2048), (8192, 4096), (2048, 4096), (1024, 2048))


@dataclass
class Job:
    raster: int
    chip: int
    margin: int
    ref_begin: np.ndarray
    tgt_begin: np.ndarray
    arrival: float = 0.0

    @property
    def pairs(self):
        return len(self.ref_begin)

    @property
    def window(self):
        return self.chip + 2 * self.margin


def raster_pool(seed=5, shapes=POOL_SHAPES, max_shift=8.0, taps=8):
    """Speckle pairs, each secondary resampled by its own smooth separable field."""
    rng = np.random.default_rng(seed)
    pool = []
    for lines, samples in shapes:
        ref = synth.speckle((lines, samples), rng)
        a, b = rng.uniform(-max_shift, max_shift, size=2)
        dl = a * (np.arange(lines) / (lines - 1)) ** 2
        ds = b * np.arange(samples) / (samples - 1)
        tgt = synth.resample_axis(synth.resample_axis(ref, dl, axis=0, taps=taps), ds, axis=1, taps=taps)
        pool.append((ref, synth.add_noise(tgt, rng.uniform(15.0, 20.0), rng)))
    return pool


def make_jobs(pool, count=256, rate=100.0, seed=5, chips=CHIPS, margins=MARGINS,
              min_pairs=64, max_pairs=4096):
    """Jobs with Poisson arrival times (seconds from the start of the stream)."""
    rng = np.random.default_rng(seed)
    jobs, t = [], 0.0
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        r = int(rng.integers(len(pool)))
        C, m = int(rng.choice(chips)), int(rng.choice(margins))
        n = int(round(np.exp(rng.uniform(np.log(min_pairs), np.log(max_pairs)))))
        lines, samples = pool[r][0].shape
        half = C // 2 + m
        cl = rng.integers(half, lines - half - (C % 2), size=n)
        cs = rng.integers(half, samples - half - (C % 2), size=n)
        ref_begin = np.stack([cl - C // 2, cs - C // 2], axis=1).astype(np.int64)
        tgt_begin = ref_begin - m
        jobs.append(Job(r, C, m, ref_begin, tgt_begin, t))
    return jobs


def upload(lib, pool):
    return [(lib.raster(ref, *ref.shape), lib.raster(tgt, *tgt.shape)) for ref, tgt in pool]


def submit(lib, server, rasters, job, precision, rm=8, rf=2, zf=8):
    ref, tgt = rasters[job.raster]
    return lib.submit(server, ref, tgt, job.ref_begin.tobytes(), job.tgt_begin.tobytes(), job.chip,
                      job.window, rm, rf, zf, precision=precision)


def warm(lib, server, rasters, precision, pool_jobs):
    """One small job per geometry: creates the lanes, tunes them, captures graphs."""
    ids = [(submit(lib, server, rasters, j, precision), j.pairs) for j in pool_jobs]
    for jid, n in ids:
        lib.wait(server, jid)


def run_stream(lib, server, rasters, jobs, precision):
    """Submit jobs at their arrival times; returns per-job latencies (ms) and wall time."""
    ids = [None] * len(jobs)
    lat = np.zeros(len(jobs))
    submitted = threading.Event()
    results = {}

    def waiter():
        submitted.wait()
        for k, j in enumerate(jobs):
            while ids[k] is None:
                time.sleep(1e-4)
            out = lib.wait(server, ids[k])
            lat[k] = out["latency_ms"]
            results[k] = out

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for k, j in enumerate(jobs):
        delay = j.arrival - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids[k] = submit(lib, server, rasters, j, precision)
        if k == 0:
            submitted.set()
    th.join()
    wall = time.perf_counter() - t0
    return lat, wall, results


def stream_stats(jobs, lat, wall):
    pairs = sum(j.pairs for j in jobs)
    return {"pairs": int(pairs), "wall_s": round(wall, 3), "pairs_per_sec": round(pairs / wall, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "max_ms": round(float(lat.max()), 3)}


def at_rate(jobs, rate):
    """The same jobs with their arrival times rescaled to `rate` jobs/s (0: all at once)."""
    first = jobs[-1].arrival / len(jobs) if jobs else 1.0
    scale = 0.0 if rate <= 0 else 1.0 / (rate * first)
    return [Job(j.raster, j.chip, j.margin, j.ref_begin, j.tgt_begin, j.arrival * scale)
            for j in jobs]


def serve_benchmark(precision="bf16", count=256, rate=100.0, seed=5, batch=2048, slots=2,
                    budget_gb=96.0, pool=None, log=None, load=0.75):
    """Run the stream at `rate`, then all at once (the saturated throughput),
    then at `load` x that throughput; returns latency percentiles and pairs/s
    of each, the first at the top level."""
    lib = ext.require()
    pool = pool if pool is not None else raster_pool(seed=seed)
    rasters = upload(lib, pool)
    jobs = make_jobs(pool, count=count, rate=rate, seed=seed)
    server = lib.server(batch=batch, slots=slots, budget_gb=budget_gb)
    warmup = []
    rng = np.random.default_rng(seed + 1)
    for C in CHIPS:
        for m in MARGINS:
            lines, samples = pool[0][0].shape
            half = C // 2 + m
            cl = rng.integers(half, lines - half, size=64)
            cs = rng.integers(half, samples - half, size=64)
            rb = np.stack([cl - C // 2, cs - C // 2], axis=1).astype(np.int64)
            warmup.append(Job(0, C, m, rb, rb - m))
    t = time.perf_counter()
    warm(lib, server, rasters, precision, warmup)
    if log:
        log(f"serving warm-up ({len(warmup)} geometries): {time.perf_counter() - t:.1f} s")
    lat, wall, _ = run_stream(lib, server, rasters, jobs, precision)
    out = {"precision": precision, "jobs": count, "arrival_rate_jobs_per_s": rate}
    out.update(stream_stats(jobs, lat, wall))
    # saturation: every job submitted at once
    lat, wall, _ = run_stream(lib, server, rasters, at_rate(jobs, 0), precision)
    sat = stream_stats(jobs, lat, wall)
    out["saturation"] = sat
    # a fixed rate at `load` x the saturated pairs/s
    mean_pairs = sat["pairs"] / count
    loaded_rate = load * sat["pairs_per_sec"] / mean_pairs
    lat, wall, _ = run_stream(lib, server, rasters, at_rate(jobs, loaded_rate), precision)
    out["loaded"] = {"load_of_saturation": load,
                     "arrival_rate_jobs_per_s": round(loaded_rate, 1)}
    out["loaded"].update(stream_stats(jobs, lat, wall))
    stats = lib.server_stats(server)
    out.update({"batch_pairs": batch, "slots_per_lane": slots, "batches": int(stats["batches"]),
                "graph_launches": int(stats["graph_launches"]), "lanes": int(stats["lanes"])})
    return out
