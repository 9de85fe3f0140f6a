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
            got = lib.wait(server, jid)
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
    solo_out = np.frombuffer(lib.wait(solo, first)["offsets"], np.float32)
    mixed = lib.server(batch=512, slots=2, budget_gb=8.0)
    ids = [serving.submit(lib, mixed, rasters, j, "fp32") for j in jobs]
    outs = [lib.wait(mixed, jid) for jid, j in zip(ids, jobs)]
    mix_out = np.frombuffer(outs[0]["offsets"], np.float32)
    np.testing.assert_allclose(mix_out, solo_out, rtol=0, atol=1e-3)


def test_server_mixed_form_against_fp64(lib, pool):
    rasters = serving.upload(lib, pool)
    jobs = jobs_for(pool, seed=5, count=4)
    server = lib.server(batch=256, slots=2, budget_gb=8.0)
    ids = [serving.submit(lib, server, rasters, j, "fp16") for j in jobs]
    for j, jid in zip(jobs, ids):
        got = np.frombuffer(lib.wait(server, jid)["offsets"], np.float32).reshape(-1, 2)
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
        lib.wait(server, jid)
    assert lib.server_stats(server)["lanes"] >= 1


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-3), ("fp8", 1e-2)])
def test_server_large_chip_lane(lib, pool, precision, tol):
    """Chip 128 / margin 64 jobs (the largest geometry of the serving stream) in
    their own lane, each against the same job run alone."""
    rasters = serving.upload(lib, pool)
    jobs = serving.make_jobs(pool, count=3, rate=1000.0, seed=8, chips=(128,), margins=(64,),
                             min_pairs=4, max_pairs=24)
    server = lib.server(batch=64, slots=2, budget_gb=8.0, graphs=True)
    for rnd in range(2):
        ids = [serving.submit(lib, server, rasters, j, precision) for j in jobs]
        for j, jid in zip(jobs, ids):
            got = np.frombuffer(lib.wait(server, jid)["offsets"], np.float32).reshape(-1, 2)
            want = alone(lib, pool, j, precision)["offsets"]
            print(f"round {rnd} C{j.chip} m{j.margin} n{j.pairs} {precision}: max diff "
                  f"{np.max(np.abs(got - want)):.2e}", flush=True)
            np.testing.assert_allclose(got, want, rtol=0, atol=tol)
    assert lib.server_stats(server)["tunes"] >= 1


def test_server_refuses_a_bad_job_and_goes_on(lib, pool):
    """A job the FP8 form cannot run (window - chip + 1 > 144) is refused at
    submit; the server keeps serving the others."""
    rasters = serving.upload(lib, pool)
    server = lib.server(batch=64, slots=1, budget_gb=2.0)
    rb = np.zeros((2, 2), np.int64)
    with pytest.raises(ValueError):
        lib.submit(server, rasters[0][0], rasters[0][1], rb.tobytes(), rb.tobytes(), 32, 320, 8, 2,
                   8, precision="fp8")
    j = jobs_for(pool, seed=7, count=1)[0]
    jid = serving.submit(lib, server, rasters, j, "fp32")
    got = np.frombuffer(lib.wait(server, jid)["offsets"], np.float32).reshape(-1, 2)
    np.testing.assert_allclose(got, alone(lib, pool, j, "fp32")["offsets"], rtol=0, atol=1e-3)


def test_server_wait_checks_the_output_capacity(pool):
    """Through the C ABI: a wait whose buffers are smaller than the job is
    refused (nothing is written), and the job is still there to wait for."""
    import ctypes

    from ampcor_amd import build

    L = ctypes.CDLL(build.library_path())
    L.ampcor_hip_last_error.restype = ctypes.c_char_p
    vp, sz, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64
    j = jobs_for(pool, seed=7, count=1)[0]
    ref_img, tgt_img = pool[j.raster]
    server, ref, tgt = vp(), vp(), vp()
    assert L.ampcor_hip_server_create(0, sz(1 << 30), sz(64), 1, 0, ctypes.byref(server)) == 0
    try:
        for img, h in ((ref_img, ref), (tgt_img, tgt)):
            assert L.ampcor_hip_raster_create(0, 0, vp(img.ctypes.data), sz(img.shape[0]),
                                              sz(img.shape[1]), ctypes.byref(h)) == 0
        rb = np.ascontiguousarray(j.ref_begin, np.int64)
        tb = np.ascontiguousarray(j.tgt_begin, np.int64)
        jid = u64()
        assert L.ampcor_hip_server_submit(server, ref, tgt, sz(j.pairs), vp(rb.ctypes.data),
                                          vp(tb.ctypes.data), sz(j.chip), sz(j.window), sz(8),
                                          sz(2), sz(8), 0, ctypes.byref(jid)) == 0
        n = sz()
        assert L.ampcor_hip_server_job_pairs(server, jid, ctypes.byref(n)) == 0 and n.value == j.pairs
        small = np.full(2, -7.0, np.float32)
        rc = L.ampcor_hip_server_wait(server, jid, -1, sz(1), vp(small.ctypes.data), None, None,
                                      None, None)
        assert rc == 1 and b"output buffers" in L.ampcor_hip_last_error()
        assert np.all(small == -7.0)
        off = np.empty((j.pairs, 2), np.float32)
        assert L.ampcor_hip_server_wait(server, jid, -1, sz(j.pairs), vp(off.ctypes.data), None,
                                        None, None, None) == 0
        np.testing.assert_allclose(off, alone_plain(j, pool), rtol=0, atol=1e-3)
    finally:
        L.ampcor_hip_server_destroy(server)
        L.ampcor_hip_raster_destroy(ref)
        L.ampcor_hip_raster_destroy(tgt)


def alone_plain(job, pool):
    ref, tgt = pool[job.raster]
    return pyoracle.adjust(ref, tgt, job.ref_begin, job.tgt_begin, job.chip, job.window, 2, 8,
                           8)["offsets"]
