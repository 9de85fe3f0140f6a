// server.hip -- continuous-batching request server for offset-tracking jobs.
//
// Jobs (a raster pair, a list of tile pairs and a geometry) are submitted at
// any time.  Each job is admitted under an HBM budget and queued on the lane
// of its geometry (chip, window, refine margin/factor, zoom, precision, pixel
// type).  A lane owns a few plan slots, each a plan of `batch` pairs with its
// own stream; a scheduler thread fills a free slot with the next pairs of the
// lane's queue -- several jobs, and several rasters, can share one batch, since
// every pair carries its own raster pointers -- and launches it.  Launches
// replay a HIP graph captured per batch-size bucket (descriptor upload, every
// stage, product download); slots of different lanes run concurrently on
// their own streams.  The first batch of a lane also runs the plan autotuner,
// whose tiling the lane's other slots adopt.
//
// No counterpart in the reference, which has no serving loop; the per-pair
// work is the worker's adjust (lib/libampcor_cuda/correlators/Sequential.icc:47-141).
#include "plan_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>

using namespace ampcor_hip;
using Clock = std::chrono::steady_clock;

namespace {

struct Key {
    int C, W, rm, rf, zf, precision, dtype;
    bool operator<(const Key &o) const
    {
        return std::tie(C, W, rm, rf, zf, precision, dtype) <
               std::tie(o.C, o.W, o.rm, o.rf, o.zf, o.precision, o.dtype);
    }
};

struct Job {
    uint64_t id;
    Key key;
    std::vector<PairDesc> desc;
    size_t dispatched = 0, completed = 0;
    std::vector<float> offsets, cpeak, rpeak, snr;
    Clock::time_point submitted, finished;
    bool done = false;
    bool waited = false;  // a thread waits for it (one at a time)
    int error = AMPCOR_HIP_OK;  // its lane could not be built: the job failed alone
    std::string error_text;
};

struct Segment {
    Job *job;
    size_t first, count, at;  // job pairs [first, first + count) sit at batch rows [at, ...)
};

struct Slot {
    ampcor_hip_plan *plan = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Segment> segments;
    int active = 0;
    PairDesc *h_desc = nullptr;  // pinned
    float *h_out = nullptr;      // pinned: offsets 2B, coarse peak B, refined peak B, snr B
    std::map<int, hipGraphExec_t> graphs;
};

struct Lane {
    Key key;
    std::vector<Slot> slots;
    std::deque<Job *> queue;
    size_t bytes = 0;
    bool tuned = false;
    Clock::time_point last_used;
};

}  // namespace

struct ampcor_hip_server {
    int device = 0;
    size_t budget = 0, used = 0;
    int batch = 0, nslots = 0;
    bool graphs = true;
    std::mutex mu;
    std::condition_variable work, finished;
    std::map<Key, std::unique_ptr<Lane>> lanes;
    std::deque<Job *> admission;
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    std::thread worker;
    bool stop = false;
    int error = AMPCOR_HIP_OK;
    std::string error_text;
    uint64_t next_id = 1;
    uint64_t pairs_done = 0, batches = 0, graph_launches = 0, tunes = 0;
};

namespace {

int bucket_for(int n, int capacity)
{
    int b = 64;
    while (b < n) b *= 2;
    return std::min(b, capacity);
}

void release_slot(Slot &s)
{
    for (auto &g : s.graphs) (void)hipGraphExecDestroy(g.second);
    s.graphs.clear();
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.h_desc) (void)hipHostFree(s.h_desc);
    if (s.h_out) (void)hipHostFree(s.h_out);
    if (s.plan) ampcor_hip_plan_destroy(s.plan);
    s = Slot{};
}

int make_lane(ampcor_hip_server *sv, const Key &k, std::unique_ptr<Lane> &out)
{
    auto lane = std::make_unique<Lane>();
    lane->key = k;
    lane->slots.resize(sv->nslots);
    for (auto &s : lane->slots) {
        int rc = ampcor_hip_plan_create(sv->device, sv->batch, k.C, k.C, k.W, k.W, k.rm, k.rf, k.zf,
                                        k.precision, &s.plan);
        if (!rc && hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
            rc = fail(AMPCOR_HIP_EDEVICE, "creating a slot event");
        if (!rc && hipHostMalloc((void **)&s.h_desc, sv->batch * sizeof(PairDesc)) != hipSuccess)
            rc = fail(AMPCOR_HIP_ENOMEM, "pinned descriptor buffer");
        if (!rc && hipHostMalloc((void **)&s.h_out, 5 * (size_t)sv->batch * sizeof(float)) != hipSuccess)
            rc = fail(AMPCOR_HIP_ENOMEM, "pinned result buffer");
        if (rc) {
            for (auto &t : lane->slots) release_slot(t);
            return rc;
        }
        s.plan->dtype = k.dtype;
    }
    lane->bytes = (size_t)sv->nslots * sv->batch * plan_bytes_per_pair(k.C, k.W, k.rf, k.rm, k.precision);
    lane->last_used = Clock::now();
    out = std::move(lane);
    return AMPCOR_HIP_OK;
}

// admit waiting jobs in arrival order while their lanes fit the budget
void admit(ampcor_hip_server *sv)
{
    while (!sv->admission.empty()) {
        Job *job = sv->admission.front();
        auto it = sv->lanes.find(job->key);
        if (it == sv->lanes.end()) {
            const Key &k = job->key;
            const size_t need = (size_t)sv->nslots * sv->batch *
                                plan_bytes_per_pair(k.C, k.W, k.rf, k.rm, k.precision);
            // evict idle lanes, least recently used first, until the new one fits
            while (sv->used + need > sv->budget) {
                Lane *victim = nullptr;
                for (auto &l : sv->lanes) {
                    bool idle = l.second->queue.empty();
                    for (auto &s : l.second->slots) idle = idle && !s.busy;
                    if (idle && (!victim || l.second->last_used < victim->last_used))
                        victim = l.second.get();
                }
                if (!victim) break;
                for (auto &s : victim->slots) release_slot(s);
                sv->used -= victim->bytes;
                sv->lanes.erase(victim->key);
            }
            if (sv->used + need > sv->budget && !sv->lanes.empty()) return;  // wait for room
            std::unique_ptr<Lane> lane;
            if (int rc = make_lane(sv, k, lane)) {
                // this job (and only it) fails; the server goes on with the others
                job->error = rc;
                job->error_text = ampcor_hip_last_error();
                job->done = true;
                job->finished = Clock::now();
                sv->admission.pop_front();
                sv->finished.notify_all();
                continue;
            }
            sv->used += lane->bytes;
            it = sv->lanes.emplace(k, std::move(lane)).first;
        }
        it->second->queue.push_back(job);
        sv->admission.pop_front();
    }
}

int capture(ampcor_hip_server *sv, Slot &s, int n, hipGraphExec_t *exec)
{
    ampcor_hip_plan *p = s.plan;
    hipStream_t st = p->stream;
    const bool timing = p->timing;
    p->timing = false;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "beginning a capture");
    hipError_t e = hipMemcpyAsync(p->d_desc.ptr, s.h_desc, n * sizeof(PairDesc),
                                  hipMemcpyHostToDevice, st);
    p->desc_dirty = false;
    int rc = e == hipSuccess ? plan_enqueue(p, n) : hip_fail(e, "capturing the descriptor upload");
    const size_t B = sv->batch;
    if (!rc) {
        const hipMemcpyKind d2h = hipMemcpyDeviceToHost;
        e = hipMemcpyAsync(s.h_out, p->offsets.ptr, 2 * n * sizeof(float), d2h, st);
        if (e == hipSuccess) e = hipMemcpyAsync(s.h_out + 2 * B, p->peak.ptr, n * sizeof(float), d2h, st);
        if (e == hipSuccess) e = hipMemcpyAsync(s.h_out + 3 * B, p->peak_r.ptr, n * sizeof(float), d2h, st);
        if (e == hipSuccess) e = hipMemcpyAsync(s.h_out + 4 * B, p->snr.ptr, n * sizeof(float), d2h, st);
        if (e != hipSuccess) rc = hip_fail(e, "capturing the product download");
    }
    hipGraph_t graph = nullptr;
    hipError_t ee = hipStreamEndCapture(st, &graph);
    p->timing = timing;
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    HIP_TRY(ee, "ending a capture");
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(e, "instantiating a batch graph");
    return AMPCOR_HIP_OK;
}

// fill a free slot with the next pairs of the lane's queue and launch it
int dispatch(ampcor_hip_server *sv, Lane &lane, Slot &s)
{
    int n = 0;
    s.segments.clear();
    while (n < sv->batch && !lane.queue.empty()) {
        Job *job = lane.queue.front();
        const size_t take = std::min<size_t>(sv->batch - n, job->desc.size() - job->dispatched);
        std::copy(job->desc.begin() + job->dispatched, job->desc.begin() + job->dispatched + take,
                  s.h_desc + n);
        s.segments.push_back({job, job->dispatched, take, (size_t)n});
        job->dispatched += take;
        n += (int)take;
        if (job->dispatched == job->desc.size()) lane.queue.pop_front();
    }
    if (n == 0) return AMPCOR_HIP_OK;
    ampcor_hip_plan *p = s.plan;
    hipStream_t st = p->stream;
    if (sv->graphs) {
        const int b = bucket_for(n, sv->batch);
        for (int k = n; k < b; ++k) s.h_desc[k] = s.h_desc[n - 1];  // padding rows, results unused
        auto it = s.graphs.find(b);
        if (it == s.graphs.end()) {
            hipGraphExec_t exec;
            if (int rc = capture(sv, s, b, &exec)) return rc;
            it = s.graphs.emplace(b, exec).first;
        }
        HIP_TRY(hipGraphLaunch(it->second, st), "launching a batch graph");
        ++sv->graph_launches;
    } else {
        const size_t B = sv->batch;
        HIP_TRY(hipMemcpyAsync(p->d_desc.ptr, s.h_desc, n * sizeof(PairDesc), hipMemcpyHostToDevice, st),
                "uploading descriptors");
        p->desc_dirty = false;
        if (int rc = plan_enqueue(p, n)) return rc;
        HIP_TRY(hipMemcpyAsync(s.h_out, p->offsets.ptr, 2 * n * sizeof(float), hipMemcpyDeviceToHost, st), "download");
        HIP_TRY(hipMemcpyAsync(s.h_out + 2 * B, p->peak.ptr, n * sizeof(float), hipMemcpyDeviceToHost, st), "download");
        HIP_TRY(hipMemcpyAsync(s.h_out + 3 * B, p->peak_r.ptr, n * sizeof(float), hipMemcpyDeviceToHost, st), "download");
        HIP_TRY(hipMemcpyAsync(s.h_out + 4 * B, p->snr.ptr, n * sizeof(float), hipMemcpyDeviceToHost, st), "download");
    }
    HIP_TRY(hipEventRecord(s.done, st), "recording a batch event");
    s.active = n;
    s.busy = true;
    lane.last_used = Clock::now();
    ++sv->batches;
    return AMPCOR_HIP_OK;
}

// copy a finished batch's products to its jobs
void harvest(ampcor_hip_server *sv, Slot &s)
{
    const size_t B = sv->batch;
    for (const Segment &g : s.segments) {
        Job *job = g.job;
        for (size_t k = 0; k < g.count; ++k) {
            const size_t src = g.at + k, dst = g.first + k;
            job->offsets[2 * dst] = s.h_out[2 * src];
            job->offsets[2 * dst + 1] = s.h_out[2 * src + 1];
            job->cpeak[dst] = s.h_out[2 * B + src];
            job->rpeak[dst] = s.h_out[3 * B + src];
            job->snr[dst] = s.h_out[4 * B + src];
        }
        job->completed += g.count;
        if (job->completed == job->desc.size()) {
            job->done = true;
            job->finished = Clock::now();
        }
    }
    sv->pairs_done += s.active;
    s.busy = false;
    s.segments.clear();
}

// autotune a lane on the data its first batch left in the plan, then share
// the chosen tiling with the lane's other slots
int tune(ampcor_hip_server *sv, Lane &lane, Slot &s)
{
    ampcor_hip_plan *p = s.plan;
    p->adjusted = true;
    int rc = ampcor_hip_plan_autotune(p, 2);
    if (rc) return rc;
    for (auto &o : lane.slots) {
        o.plan->coarse_shape = p->coarse_shape;
        o.plan->refined_shape = p->refined_shape;
        o.plan->mixed_coarse = p->mixed_coarse;
        o.plan->mixed_refined = p->mixed_refined;
        o.plan->fp8_coarse = p->fp8_coarse;
        for (auto &g : o.graphs) (void)hipGraphExecDestroy(g.second);
        o.graphs.clear();
    }
    lane.tuned = true;
    ++sv->tunes;
    return AMPCOR_HIP_OK;
}

void run(ampcor_hip_server *sv)
{
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lock(sv->mu);
    while (!sv->stop) {
        admit(sv);
        bool busy = false, progressed = false;
        for (auto &kv : sv->lanes) {
            Lane &lane = *kv.second;
            for (auto &s : lane.slots) {
                if (s.busy) {
                    const hipError_t q = hipEventQuery(s.done);
                    if (q == hipSuccess) {
                        harvest(sv, s);
                        progressed = true;
                        if (!lane.tuned && sv->error == AMPCOR_HIP_OK) {
                            if (tune(sv, lane, s)) {
                                sv->error = AMPCOR_HIP_EDEVICE;
                                sv->error_text = ampcor_hip_last_error();
                            }
                        }
                    } else if (q != hipErrorNotReady) {
                        sv->error = AMPCOR_HIP_EDEVICE;
                        sv->error_text = std::string("batch failed: ") + hipGetErrorString(q);
                        s.busy = false;
                    }
                }
                if (!s.busy && !lane.queue.empty() && sv->error == AMPCOR_HIP_OK) {
                    if (dispatch(sv, lane, s)) {
                        sv->error = AMPCOR_HIP_EDEVICE;
                        sv->error_text = ampcor_hip_last_error();
                    }
                    progressed = true;
                }
                busy = busy || s.busy;
            }
        }
        if (progressed) sv->finished.notify_all();
        if (sv->error != AMPCOR_HIP_OK) sv->finished.notify_all();
        if (!busy && sv->admission.empty()) {
            bool queued = false;
            for (auto &kv : sv->lanes) queued = queued || !kv.second->queue.empty();
            if (!queued) {
                sv->work.wait(lock);
                continue;
            }
        }
        lock.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lock.lock();
    }
}

}  // namespace

extern "C" {

int ampcor_hip_server_create(int device, size_t hbm_budget_bytes, size_t batch_pairs,
                             int slots_per_lane, int use_graphs, ampcor_hip_server **out)
{
    if (!out) return fail(AMPCOR_HIP_EINVAL, "null output pointer");
    *out = nullptr;
    if (batch_pairs < 1 || batch_pairs > (1u << 20)) return fail(AMPCOR_HIP_EINVAL, "batch size out of range");
    if (slots_per_lane < 1 || slots_per_lane > 8) return fail(AMPCOR_HIP_EINVAL, "slots per lane must be 1..8");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(AMPCOR_HIP_EDEVICE, "no HIP device available");
    if (device < 0 || device >= count) return fail(AMPCOR_HIP_EINVAL, "device index out of range");
    auto *sv = new ampcor_hip_server();
    sv->device = device;
    sv->budget = hbm_budget_bytes;
    sv->batch = (int)batch_pairs;
    sv->nslots = slots_per_lane;
    sv->graphs = use_graphs != 0;
    sv->worker = std::thread(run, sv);
    *out = sv;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_server_submit(ampcor_hip_server *sv, const ampcor_hip_raster *reference,
                             const ampcor_hip_raster *target, size_t pairs, const int64_t *ref_begin,
                             const int64_t *tgt_begin, size_t chip, size_t window,
                             size_t refine_margin, size_t refine_factor, size_t zoom_factor,
                             int precision, uint64_t *job_id)
{
    if (!sv || !reference || !target || !job_id || (pairs && (!ref_begin || !tgt_begin)))
        return fail(AMPCOR_HIP_EINVAL, "null argument");
    if (pairs == 0) return fail(AMPCOR_HIP_EINVAL, "a job needs at least one pair");
    if (reference->dtype != target->dtype) return fail(AMPCOR_HIP_EINVAL, "rasters differ in pixel type");
    if (reference->device != sv->device || target->device != sv->device)
        return fail(AMPCOR_HIP_EINVAL, "rasters live on another device than the server");
    // the checks a plan of this geometry applies (its lane's plans are made later)
    {
        ampcor_hip_plan probe;
        if (int rc = plan_configure(&probe, sv->device, 1, chip, chip, window, window, refine_margin,
                                    refine_factor, zoom_factor, precision))
            return rc;
    }
    auto job = std::make_unique<Job>();
    job->key = Key{(int)chip, (int)window, (int)refine_margin, (int)refine_factor, (int)zoom_factor,
                   precision, reference->dtype};
    job->desc.resize(pairs);
    for (size_t k = 0; k < pairs; ++k) {
        const int64_t rl = ref_begin[2 * k], rs = ref_begin[2 * k + 1];
        const int64_t tl = tgt_begin[2 * k], ts = tgt_begin[2 * k + 1];
        if (rl < 0 || rs < 0 || tl < 0 || ts < 0 || rl + (int64_t)chip > (int64_t)reference->lines ||
            rs + (int64_t)chip > (int64_t)reference->samples || tl + (int64_t)window > (int64_t)target->lines ||
            ts + (int64_t)window > (int64_t)target->samples)
            return fail(AMPCOR_HIP_EINVAL, "pair " + std::to_string(k) + " leaves its raster");
        PairDesc &d = job->desc[k];
        d.ref = reference->data;
        d.tgt = target->data;
        d.ref_stride = (int32_t)reference->samples;
        d.tgt_stride = (int32_t)target->samples;
        d.ref_line = (int32_t)rl;
        d.ref_sample = (int32_t)rs;
        d.tgt_line = (int32_t)tl;
        d.tgt_sample = (int32_t)ts;
    }
    job->offsets.assign(2 * pairs, 0.0f);
    job->cpeak.assign(pairs, 0.0f);
    job->rpeak.assign(pairs, 0.0f);
    job->snr.assign(pairs, 0.0f);
    std::lock_guard<std::mutex> lock(sv->mu);
    if (sv->error) return fail(sv->error, sv->error_text);
    job->id = sv->next_id++;
    job->submitted = Clock::now();
    *job_id = job->id;
    sv->admission.push_back(job.get());
    sv->jobs.emplace(job->id, std::move(job));
    sv->work.notify_one();
    return AMPCOR_HIP_OK;
}

int ampcor_hip_server_job_pairs(ampcor_hip_server *sv, uint64_t job_id, size_t *pairs)
{
    if (!sv || !pairs) return fail(AMPCOR_HIP_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(sv->mu);
    auto it = sv->jobs.find(job_id);
    if (it == sv->jobs.end()) return fail(AMPCOR_HIP_EINVAL, "unknown job");
    *pairs = it->second->desc.size();
    return AMPCOR_HIP_OK;
}

int ampcor_hip_server_wait(ampcor_hip_server *sv, uint64_t job_id, int timeout_ms, size_t capacity,
                           float *offsets, float *coarse_peak, float *refined_peak, float *snr,
                           double *latency_ms)
{
    if (!sv) return fail(AMPCOR_HIP_EINVAL, "null server");
    std::unique_lock<std::mutex> lock(sv->mu);
    auto it = sv->jobs.find(job_id);
    if (it == sv->jobs.end()) return fail(AMPCOR_HIP_EINVAL, "unknown job");
    Job *job = it->second.get();
    const size_t n = job->desc.size();
    if ((offsets || coarse_peak || refined_peak || snr) && capacity < n)
        return fail(AMPCOR_HIP_EINVAL, "the job has " + std::to_string(n) +
                                           " pairs, the output buffers hold " +
                                           std::to_string(capacity));
    if (job->waited) return fail(AMPCOR_HIP_EINVAL, "another thread waits for this job");
    job->waited = true;
    auto ready = [&] { return job->done || sv->error != AMPCOR_HIP_OK; };
    bool ok = true;
    if (timeout_ms < 0) sv->finished.wait(lock, ready);
    else ok = sv->finished.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready);
    job->waited = false;
    if (!ok) return fail(AMPCOR_HIP_ESTATE, "job not finished yet");
    if (!job->done) return fail(sv->error, sv->error_text);
    if (job->error) {
        const int rc = fail(job->error, job->error_text);
        sv->jobs.erase(it);
        return rc;
    }
    if (offsets) std::copy(job->offsets.begin(), job->offsets.end(), offsets);
    if (coarse_peak) std::copy(job->cpeak.begin(), job->cpeak.end(), coarse_peak);
    if (refined_peak) std::copy(job->rpeak.begin(), job->rpeak.end(), refined_peak);
    if (snr) std::copy(job->snr.begin(), job->snr.end(), snr);
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(job->finished - job->submitted).count();
    sv->jobs.erase(it);
    return AMPCOR_HIP_OK;
}

int ampcor_hip_server_stats(ampcor_hip_server *sv, uint64_t *pairs, uint64_t *batches,
                            uint64_t *graph_launches, uint64_t *lanes, uint64_t *tunes)
{
    if (!sv) return fail(AMPCOR_HIP_EINVAL, "null server");
    std::lock_guard<std::mutex> lock(sv->mu);
    if (pairs) *pairs = sv->pairs_done;
    if (batches) *batches = sv->batches;
    if (graph_launches) *graph_launches = sv->graph_launches;
    if (lanes) *lanes = sv->lanes.size();
    if (tunes) *tunes = sv->tunes;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_server_destroy(ampcor_hip_server *sv)
{
    if (!sv) return AMPCOR_HIP_OK;
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        sv->stop = true;
        sv->work.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipSetDevice(sv->device);
    for (auto &kv : sv->lanes)
        for (auto &s : kv.second->slots) {
            if (s.busy) (void)hipEventSynchronize(s.done);
            release_slot(s);
        }
    delete sv;
    return AMPCOR_HIP_OK;
}

}  // extern "C"
