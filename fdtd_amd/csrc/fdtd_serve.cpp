// Serving loop (fdtd_amd.h, fdtd_serve_*): a worker thread owns a batch and
// runs continuous batching over a request queue.
//
// Each worker iteration:
//   1. admits queued requests in arrival order while slots and the arena
//      allow (a request that does not fit lets later ones that fit pass);
//      their inputs are uploaded on the batch's upload stream while the
//      device runs the previous segment;
//   2. enqueues one segment (segment_steps batch steps, a graph replay) on
//      the compute stream, then the probe copies of the requests that reach
//      their step count inside it, into pinned buffers, and an event;
//   3. retires the segment before last: waits for its event, hands the
//      probe series to the requests and returns their arena regions.
// One segment stays queued behind the running one, so the device does not
// wait for the host between segments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fdtd_batch_internal.h"

namespace {

struct Request {
    long long id = 0;
    FdtdParams p;
    FdtdExt e;
    std::vector<unsigned char> ids;
    std::vector<float> table, rp[6], jz;
    std::vector<int> src, probes;
    int nmat = 0, nsrc = 0, steps = 0, nprobe = 0;
    bool has_ids = false, has_rp[6] = {false, false, false, false, false, false};
    size_t bytes = 0;
    int state = FDTD_REQ_QUEUED, error = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int slot = -1;
    long long finish_step = 0;
    double *pinned = nullptr;
    size_t pinned_bytes = 0;
    std::vector<double> result;

    FdtdInstance instance() const
    {
        FdtdInstance in;
        memset(&in, 0, sizeof in);
        in.ids = has_ids ? ids.data() : nullptr;
        in.nmat = nmat;
        in.table = table.data();
        for (int r = 0; r < 6; r++) in.rp[r] = has_rp[r] ? rp[r].data() : nullptr;
        in.nsrc = nsrc;
        in.src_xyz = src.data();
        in.jz_len = (int)jz.size();
        in.jz = jz.data();
        in.steps = steps;
        in.nprobe = nprobe;
        in.probes = probes.data();
        return in;
    }
};

struct Segment {
    hipEvent_t ev = nullptr;
    std::vector<Request *> finished;
};

}  // namespace

struct fdtd_serve {
    FdtdServeConfig cfg;
    int device = 0;
    fdtd_batch *b = nullptr;
    hipStream_t stream = nullptr;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::unordered_map<long long, std::unique_ptr<Request>> reqs;
    long long next_id = 1;
    bool stop = false;
    int running = 0;        // admitted, not retired
    int submitted = 0, completed = 0, failed = 0;
    double t_first_submit = -1, t_last_done = 0;
    long long segments = 0;
    size_t peak_bytes = 0;
    int peak_running = 0;
    double cell_updates = 0;
    std::vector<double> latencies;

    // worker-only state
    std::vector<Request *> active;
    std::deque<Segment> inflight;
    std::vector<hipEvent_t> spare_events;
    std::multimap<size_t, double *> pinned_pool;

    std::thread worker;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

double *take_pinned(fdtd_serve *s, size_t bytes)
{
    auto it = s->pinned_pool.lower_bound(bytes);
    if (it != s->pinned_pool.end()) {
        double *p = it->second;
        s->pinned_pool.erase(it);
        return p;
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(bytes, 64), hipHostMallocDefault) != hipSuccess) return nullptr;
    return (double *)p;
}

void fail_request(fdtd_serve *s, Request *r, int code)
{
    // caller holds s->mu
    r->state = FDTD_REQ_FAILED;
    r->error = code;
    r->t_done = s->now();
    s->failed++;
}

// Step 1: admission.  Returns the number admitted.
int admit(fdtd_serve *s)
{
    std::vector<Request *> cand;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return 0;
        const int room = s->cfg.max_slots - s->running;
        for (Request *r : s->queue) {
            if ((int)cand.size() >= room) break;
            cand.push_back(r);
        }
    }
    int n = 0;
    std::vector<std::pair<Request *, int>> done;
    for (Request *r : cand) {
        if (r->bytes > batch_largest_free(s->b)) continue;  // may fit later
        const FdtdInstance in = r->instance();
        int slot = -1;
        const int rc = fdtd_batch_add(s->b, &r->p, &r->e, &in, &slot);
        if (rc == FDTD_ENOMEM) continue;
        if (rc == 0) {
            r->slot = slot;
            r->finish_step = batch_step(s->b) + r->steps;
            r->pinned_bytes = sizeof(double) * (size_t)r->nprobe * r->steps;
            r->pinned = r->pinned_bytes ? take_pinned(s, r->pinned_bytes) : nullptr;
            if (r->pinned_bytes && !r->pinned) {
                batch_remove_nosync(s->b, slot);
                done.push_back({r, hipErrorOutOfMemory});
                continue;
            }
        }
        done.push_back({r, rc});
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &d : done) {
        Request *r = d.first;
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), r));
        if (d.second != 0) {
            fail_request(s, r, d.second);
            continue;
        }
        r->state = FDTD_REQ_RUNNING;
        r->t_admit = s->now();
        s->running++;
        s->active.push_back(r);
        n++;
    }
    s->peak_running = std::max(s->peak_running, s->running);
    s->peak_bytes = std::max(s->peak_bytes, batch_used_bytes(s->b));
    if (!done.empty()) s->cv_done.notify_all();
    return n;
}

// Step 2: one segment.
int run_segment(fdtd_serve *s)
{
    int rc = fdtd_batch_run(s->b, s->cfg.segment_steps, s->stream);
    if (rc) return rc;
    Segment seg;
    const long long step = batch_step(s->b);
    std::vector<Request *> keep;
    for (Request *r : s->active) {
        if (r->finish_step <= step) {
            if (r->pinned) rc = batch_probe_copy_async(s->b, r->slot, r->pinned, r->steps, s->stream);
            if (rc) return rc;
            seg.finished.push_back(r);
        } else {
            keep.push_back(r);
        }
    }
    s->active.swap(keep);
    if (!s->spare_events.empty()) {
        seg.ev = s->spare_events.back();
        s->spare_events.pop_back();
    } else {
        HIP_TRY(hipEventCreateWithFlags(&seg.ev, hipEventDisableTiming));
    }
    HIP_TRY(hipEventRecord(seg.ev, s->stream));
    s->inflight.push_back(std::move(seg));
    std::lock_guard<std::mutex> lk(s->mu);
    s->segments++;
    return 0;
}

// Step 3: retire the oldest segment.
int retire(fdtd_serve *s)
{
    Segment seg = std::move(s->inflight.front());
    s->inflight.pop_front();
    HIP_TRY(hipEventSynchronize(seg.ev));
    s->spare_events.push_back(seg.ev);
    const double t = s->now();
    for (Request *r : seg.finished) {
        if (r->pinned) {
            r->result.assign(r->pinned, r->pinned + (size_t)r->nprobe * r->steps);
            s->pinned_pool.emplace(r->pinned_bytes, r->pinned);
            r->pinned = nullptr;
        }
        batch_remove_nosync(s->b, r->slot);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (Request *r : seg.finished) {
        r->state = FDTD_REQ_DONE;
        r->t_done = t;
        s->running--;
        s->completed++;
        s->t_last_done = t;
        s->latencies.push_back(r->t_done - r->t_submit);
        s->cell_updates += (double)r->p.nx * r->p.ny * r->p.nz * r->steps;
    }
    if (!seg.finished.empty()) s->cv_done.notify_all();
    return 0;
}

void worker_main(fdtd_serve *s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || s->running > 0; });
            if (s->stop && s->running == 0) {
                for (Request *r : s->queue) fail_request(s, r, FDTD_ESTATE);
                s->queue.clear();
                break;
            }
        }
        admit(s);
        int rc = 0;
        if (!s->active.empty()) rc = run_segment(s);
        while (!rc && (s->inflight.size() > 1 || (s->active.empty() && !s->inflight.empty()))) rc = retire(s);
        if (rc) {
            // A device error ends the server: every admitted and queued request fails.
            (void)hipStreamSynchronize(s->stream);  // queued probe copies may still write the pinned buffers
            std::lock_guard<std::mutex> lk(s->mu);
            auto drop = [&](Request *r) {
                if (r->pinned) s->pinned_pool.emplace(r->pinned_bytes, r->pinned);  // freed with the server
                r->pinned = nullptr;
                fail_request(s, r, rc);
            };
            for (Request *r : s->active) drop(r);
            for (auto &seg : s->inflight) {
                for (Request *r : seg.finished) drop(r);
                s->spare_events.push_back(seg.ev);  // destroyed with the server
            }
            for (Request *r : s->queue) fail_request(s, r, rc);
            s->active.clear();
            s->inflight.clear();
            s->queue.clear();
            s->running = 0;
            s->stop = true;
            s->cv_done.notify_all();
            break;
        }
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stop && s->active.empty() && s->inflight.empty()) {
            for (Request *r : s->queue) fail_request(s, r, FDTD_ESTATE);
            s->queue.clear();
            s->cv_done.notify_all();
            break;
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->cv_done.notify_all();
}

template <typename T>
void take(std::vector<T> &v, const T *p, size_t n)
{
    if (p && n) v.assign(p, p + n);
    else v.clear();
}

}  // namespace

extern "C" int fdtd_serve_create(const FdtdServeConfig *cfg, fdtd_serve **out)
{
    if (!cfg || !out) return fdtd_fail(FDTD_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->max_slots < 1 || cfg->max_slots > kBatchMaxInst) return fdtd_fail(FDTD_ERANGE, "max_slots %d", cfg->max_slots);
    auto s = std::make_unique<fdtd_serve>();
    s->cfg = *cfg;
    if (s->cfg.segment_steps <= 0) s->cfg.segment_steps = 8;
    s->t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipGetDevice(&s->device));
    if (int rc = fdtd_batch_create(cfg->hbm_bytes, &s->b)) return rc;
    int rc = fdtd_batch_set_shape(s->b, cfg->tile_y, cfg->kchunk, s->cfg.segment_steps);
    hipError_t e = rc ? hipSuccess : hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (rc || e != hipSuccess) {
        fdtd_batch_destroy(s->b);
        return rc ? rc : fdtd_hip_fail(e, "hipStreamCreate");
    }
    fdtd_serve *raw = s.release();
    raw->worker = std::thread(worker_main, raw);
    *out = raw;
    return 0;
}

extern "C" int fdtd_serve_destroy(fdtd_serve *s)
{
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipStreamSynchronize(s->stream);
    for (auto &seg : s->inflight) (void)hipEventDestroy(seg.ev);
    for (auto ev : s->spare_events) (void)hipEventDestroy(ev);
    for (auto &r : s->reqs)
        if (r.second->pinned) s->pinned_pool.emplace(r.second->pinned_bytes, r.second->pinned);
    for (auto &p : s->pinned_pool) (void)hipHostFree(p.second);
    fdtd_batch_destroy(s->b);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return 0;
}

extern "C" int fdtd_serve_submit(fdtd_serve *s, const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *in,
                                 long long *request)
{
    if (!s || !p || !in || !request) return fdtd_fail(FDTD_EINVAL, "null argument");
    size_t bytes = 0;
    if (int rc = fdtd_batch_instance_bytes(p, ext, in, &bytes)) return rc;
    if (bytes > s->cfg.hbm_bytes) return fdtd_fail(FDTD_ENOMEM, "request needs %zu bytes; budget %zu", bytes, s->cfg.hbm_bytes);
    auto r = std::make_unique<Request>();
    r->p = *p;
    fdtd_ext_defaults(&r->e);
    if (ext) r->e = *ext;
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    r->has_ids = in->ids != nullptr;
    take(r->ids, in->ids, N);
    take(r->table, in->table, (size_t)4 * in->nmat);
    const size_t rps[6] = {(size_t)2 * p->ny * p->nz, (size_t)p->nx * 2 * p->nz, (size_t)p->nx * p->ny * 2,
                           (size_t)2 * p->ny * p->nz, (size_t)p->nx * 2 * p->nz, (size_t)p->nx * p->ny * 2};
    for (int q = 0; q < 6; q++) {
        r->has_rp[q] = in->rp[q] != nullptr;
        take(r->rp[q], in->rp[q], rps[q]);
    }
    take(r->src, in->src_xyz, (size_t)3 * in->nsrc);
    take(r->jz, in->jz, (size_t)in->jz_len);
    take(r->probes, in->probes, (size_t)4 * in->nprobe);
    r->nmat = in->nmat;
    r->nsrc = in->nsrc;
    r->steps = in->steps;
    r->nprobe = in->nprobe;
    r->bytes = bytes;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return fdtd_fail(FDTD_ESTATE, "server stopped");
        r->id = s->next_id++;
        r->t_submit = s->now();
        if (s->t_first_submit < 0) s->t_first_submit = r->t_submit;
        s->submitted++;
        *request = r->id;
        s->queue.push_back(r.get());
        s->reqs.emplace(r->id, std::move(r));
    }
    s->cv_work.notify_one();
    return 0;
}

extern "C" int fdtd_serve_wait(fdtd_serve *s, long long request, double timeout_s)
{
    if (!s) return fdtd_fail(FDTD_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->reqs.find(request) == s->reqs.end()) return fdtd_fail(FDTD_EINVAL, "unknown request %lld", request);
    // looked up on every wake-up: another thread may release the request
    // once it has finished
    auto done = [&] {
        auto it = s->reqs.find(request);
        if (it == s->reqs.end()) return true;
        const int st = it->second->state;
        return st == FDTD_REQ_DONE || st == FDTD_REQ_FAILED;
    };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, done);
        return 0;
    }
    return s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done) ? 0 : 1;
}

extern "C" int fdtd_serve_result(fdtd_serve *s, long long request, double *probes, int max_rows, FdtdRequestInfo *info)
{
    if (!s) return fdtd_fail(FDTD_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    if (it == s->reqs.end()) return fdtd_fail(FDTD_EINVAL, "unknown request %lld", request);
    Request *r = it->second.get();
    const int rows = r->state == FDTD_REQ_DONE && r->nprobe > 0 ? std::min(max_rows, r->steps) : 0;
    if (rows > 0 && probes) memcpy(probes, r->result.data(), sizeof(double) * (size_t)rows * r->nprobe);
    if (info) {
        info->state = r->state;
        info->error = r->error;
        info->rows = r->state == FDTD_REQ_DONE ? r->steps : 0;
        info->nprobe = r->nprobe;
        info->t_submit = r->t_submit;
        info->t_admit = r->t_admit;
        info->t_done = r->t_done;
        info->cells = (long long)r->p.nx * r->p.ny * r->p.nz;
        info->steps = r->steps;
    }
    return rows;
}

extern "C" int fdtd_serve_release(fdtd_serve *s, long long request)
{
    if (!s) return fdtd_fail(FDTD_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    if (it == s->reqs.end()) return fdtd_fail(FDTD_EINVAL, "unknown request %lld", request);
    const int st = it->second->state;
    if (st != FDTD_REQ_DONE && st != FDTD_REQ_FAILED) return fdtd_fail(FDTD_ESTATE, "request %lld not finished", request);
    s->reqs.erase(it);
    return 0;
}

extern "C" int fdtd_serve_drain(fdtd_serve *s, double timeout_s)
{
    if (!s) return fdtd_fail(FDTD_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto idle = [&] { return s->queue.empty() && s->running == 0; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, idle);
        return 0;
    }
    return s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), idle) ? 0 : 1;
}

extern "C" int fdtd_serve_stats(fdtd_serve *s, FdtdServeStats *st)
{
    if (!s || !st) return fdtd_fail(FDTD_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof *st);
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->queued = (int)s->queue.size();
    st->running = s->running;
    st->t_first_submit = s->t_first_submit < 0 ? 0 : s->t_first_submit;
    st->t_last_done = s->t_last_done;
    const double span = s->t_last_done - st->t_first_submit;
    st->solves_per_s = (s->completed > 0 && span > 0) ? s->completed / span : 0.0;
    if (!s->latencies.empty()) {
        std::vector<double> l(s->latencies);
        std::sort(l.begin(), l.end());
        // nearest-rank percentiles
        auto pct = [&](double q) {
            size_t k = (size_t)std::ceil(q * l.size());
            k = std::min(std::max<size_t>(k, 1), l.size());
            return l[k - 1];
        };
        st->p50_latency_s = pct(0.50);
        st->p99_latency_s = pct(0.99);
        double sum = 0;
        for (double v : l) sum += v;
        st->mean_latency_s = sum / l.size();
    }
    st->cell_updates = s->cell_updates;
    st->segments = s->segments;
    st->peak_bytes = s->peak_bytes;
    st->peak_running = s->peak_running;
    return 0;
}
