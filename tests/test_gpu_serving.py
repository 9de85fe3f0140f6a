"""The continuous-batching server: every job's result equals the job run alone
(through a plan, eagerly) within the form's tolerance, whatever it was batched
with; graph replays are checked one by one with eager work between them."""

import numpy as np
import pytest

from ampcor_amd import serving, synth
from oracle import pyoracle
from tests.gpu_helpers import run_hip

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-3, "bf16": 1e-2, "fp16": 1e-2}


@pytest.fixture(scope="module")
def pool():
    return serving.raster_pool(seed=9, shapes=((512, 512), (768, 512)), max_shift=4.0)


def jobs_for(pool, seed, count=6):
    return serving.make_jobs(pool, count=count, rate=1000.0, seed=seed, chips=(32, 64),
                             margins=(16, 32), min_pairs=20, max_pairs=300)


def alone(lib, pool, job, precision):
    ref, tgt = pool[job.raster]
    return run_hip(lib, ref, tgt, job.ref_begin, job.tgt_begin, job.chip, job.window, 2, 8, 8,
                   precision=precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_server_results_equal_jobs_alone(lib, pool, precision):
    rasters = serving.upload(lib, pool)
    jobs = jobs_for(pool, seed=3)
    want = [alone(lib, pool, j, precision) for j in jobs]
    server = lib.server(batch=256, slots=2, budget_gb=8.0, graphs=True)
    for rnd in range(3):
        ids = [serving.submit(lib, server, rasters, j, precision) for j in jobs]
        for j, jid, w in zip(jobs, ids, want):
            got = lib.wait(server, jid, j.pairs)
            off = np.frombuffer(got["offsets"], np.float32).reshape(-1, 2)
            print(f"round {rnd} job C{j.chip} m{j.margin} n{j.pairs}: max diff "
                  f"{np.max(np.abs(off - w['offsets'])):.2e}", flush=True)
            np.testing.assert_allclose(off, w["offsets"], rtol=0, atol=TOL[precision])
            np.testing.assert_allclose(np.frombuffer(got["coarse_peak"], np.float32),
                                       w["coarse_peak"], rtol=1e-5 if precision == "fp32" else 1e-2)
            assert got["latency_ms"] > 0
        # eager work on the device between graph replays
        alone(lib, pool, jobs[0], precision)
    stats = lib.server_stats(server)
    assert stats["graph_launches"] == stats["batches"] > 0
    # jobs of one geometry share batches: fewer batches than job submissions
    assert stats["batches"] <= 3 * len(jobs)
    assert stats["lanes"] >= 2


def test_alone_and_in_the_mix_agree(lib, pool):
    rasters = serving.upload(lib, pool)
    jobs = jobs_for(pool, seed=4, count=8)
    solo = lib.server(batch=512, slots=2, budget_gb=8.0)
    first = serving.submit(lib, solo, rasters, jobs[0], "fp32")
    solo_out = np.frombuffer(lib.wait(solo, first, jobs[0].pairs)["offsets"], np.float32)
    mixed = lib.server(batch=512, slots=2, budget_gb=8.0)
    ids = [serving.submit(lib, mixed, rasters, j, "fp32") for j in jobs]
    outs = [lib.wait(mixed, jid, j.pairs) for jid, j in zip(ids, jobs)]
    mix_out = np.frombuffer(outs[0]["offsets"], np.float32)
    np.testing.assert_allclose(mix_out, solo_out, rtol=0, atol=1e-3)


def test_server_mixed_form_against_fp64(lib, pool):
    rasters = serving.upload(lib, pool)
    jobs = jobs_for(pool, seed=5, count=4)
    server = lib.server(batch=256, slots=2, budget_gb=8.0)
    ids = [serving.submit(lib, server, rasters, j, "fp16") for j in jobs]
    for j, jid in zip(jobs, ids):
        got = np.frombuffer(lib.wait(server, jid, j.pairs)["offsets"], np.float32).reshape(-1, 2)
        ref, tgt = pool[j.raster]
        exact = pyoracle.adjust(ref, tgt, j.ref_begin, j.tgt_begin, j.chip, j.window, 2, 8, 8)
        assert np.max(np.abs(got - exact["offsets"])) <= 1e-2


def test_server_small_budget_evicts_idle_lanes(lib, pool):
    """A budget that holds one lane at a time still serves jobs of several geometries."""
    rasters = serving.upload(lib, pool)
    jobs = jobs_for(pool, seed=6, count=4)
    server = lib.server(batch=64, slots=1, budget_gb=0.05)
    for j in jobs:
        jid = serving.submit(lib, server, rasters, j, "fp32")
        lib.wait(server, jid, j.pairs)
    assert lib.server_stats(server)["lanes"] >= 1
