// server.hip -- continuous batching of host-buffer requests (ostft_server_*).
//
// Request path:
//   submit():    checks the request, queues it (FIFO) with its submit time.
//   scheduler:   while fewer than max_inflight batches are on the GPU, takes
//                the queued requests that fit, in order, into one contiguous
//                chunk of the device arena ([inputs][outputs]); copies their
//                inputs into the pinned mirror of that chunk, then
//                  h2d stream:      one copy of the input block
//                  compute stream:  ostft_run_batch_device over the batch
//                  d2h stream:      one copy of the output block
//                with events between the streams, so batch k+1's input copy
//                and batch k-1's output copy overlap batch k's kernels.
//   completion:  waits for each batch in order, copies every result from
//                the mirror to its caller's buffer, records latency, frees
//                the chunk (the arena is a FIFO ring) and wakes the waiters
//                and the scheduler.
// A request's result does not depend on the batch it lands in: the batch
// kernels compute each frame exactly as the one-request kernel does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/occastft_hip.h"

namespace {

using Clock = std::chrono::steady_clock;
constexpr size_t ALIGN = 256;
size_t align_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

struct Request {
    uint64_t id = 0;
    const float* in = nullptr;
    uint64_t L = 0;
    ostft_params p{};
    ostft_geometry g{};
    float* out = nullptr;
    size_t in_off = 0, out_off = 0;  // arena byte offsets
    Clock::time_point t_submit, t_done;
    int status = -1;                 // -1: not finished
};

struct Batch {
    std::vector<Request*> reqs;
    size_t start = 0, in_bytes = 0, out_bytes = 0;  // chunk [start, start + in_bytes + out_bytes)
    hipEvent_t ev_in = nullptr, ev_comp = nullptr, ev_done = nullptr;
    int status = OSTFT_OK;
    std::string err;
};

// FIFO ring of chunks over [0, cap): chunks are freed in allocation order.
struct Ring {
    size_t cap = 0;
    std::deque<std::pair<size_t, size_t>> chunks;  // [start, end)

    // the largest contiguous free region a new chunk may take
    void free_region(size_t* off, size_t* len) const
    {
        if (chunks.empty()) {
            *off = 0;
            *len = cap;
            return;
        }
        const size_t head = chunks.back().second, tail = chunks.front().first;
        const bool wrapped = chunks.back().first < tail;
        if (wrapped) {
            *off = head;
            *len = tail - head;
        } else if (cap - head >= tail) {
            *off = head;
            *len = cap - head;
        } else {
            *off = 0;
            *len = tail;
        }
    }
    void push(size_t start, size_t len) { chunks.emplace_back(start, start + len); }
    void pop() { chunks.pop_front(); }
};

// A few workers for the host copies between callers' buffers and the pinned
// mirror: one thread moves ~10 GB/s, less than the PCIe link carries.
class CopyPool {
public:
    explicit CopyPool(int n)
    {
        for (int i = 0; i < n; ++i)
            workers_.emplace_back([this] { loop(); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_)
            t.join();
    }
    // Copies every (dst, src, bytes) job, split into pieces of at most 4 MiB,
    // with the calling thread helping; returns when all are done.
    void copy(const std::vector<std::tuple<void*, const void*, size_t>>& jobs)
    {
        constexpr size_t PIECE = 4u << 20;
        std::vector<std::tuple<char*, const char*, size_t>> pieces;
        for (const auto& j : jobs) {
            char* d = static_cast<char*>(std::get<0>(j));
            const char* sp = static_cast<const char*>(std::get<1>(j));
            for (size_t o = 0; o < std::get<2>(j); o += PIECE)
                pieces.emplace_back(d + o, sp + o, std::min(PIECE, std::get<2>(j) - o));
        }
        if (pieces.empty())
            return;
        std::unique_lock<std::mutex> lk(mu_);
        pieces_ = &pieces;
        next_ = 0;
        left_ = pieces.size();
        ++gen_;
        cv_.notify_all();
        work(lk);
        done_.wait(lk, [&] { return left_ == 0; });
        pieces_ = nullptr;
    }

private:
    void work(std::unique_lock<std::mutex>& lk)
    {
        while (pieces_ && next_ < pieces_->size()) {
            const auto pc = (*pieces_)[next_++];
            lk.unlock();
            std::memcpy(std::get<0>(pc), std::get<1>(pc), std::get<2>(pc));
            lk.lock();
            if (--left_ == 0)
                done_.notify_all();
        }
    }
    void loop()
    {
        std::unique_lock<std::mutex> lk(mu_);
        uint64_t seen = 0;
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || (gen_ != seen && pieces_ && next_ < pieces_->size()); });
            if (stop_)
                return;
            seen = gen_;
            work(lk);
        }
    }
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> workers_;
    const std::vector<std::tuple<char*, const char*, size_t>>* pieces_ = nullptr;
    size_t next_ = 0, left_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

}  // namespace

struct ostft_server {
    int device = 0;
    ostft_ctx* ctx = nullptr;
    ostft_server_config cfg{};
    hipStream_t h2d = nullptr, compute = nullptr, d2h = nullptr;
    char* arena_d = nullptr;
    char* arena_h = nullptr;
    Ring ring;

    std::mutex mu;
    std::condition_variable cv_sched;  // queue grew, arena freed, batch retired, stop
    std::condition_variable cv_comp;   // batch submitted, stop
    std::condition_variable cv_done;   // a request finished
    std::deque<Request*> pending;
    std::deque<std::unique_ptr<Batch>> inflight;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<hipEvent_t> event_pool;
    uint64_t next_id = 1;
    bool stopping = false, sched_done = false;
    std::thread sched, comp;
    std::string last_error;

    // statistics
    ostft_server_stats st{};
    std::vector<double> lat_ms;
    std::unique_ptr<CopyPool> copy_in, copy_out;  // one pool per direction (two threads use them)
    Clock::time_point t_first, t_last;
    bool have_first = false;

    hipEvent_t take_event()
    {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
        return e;
    }

    void finish(Request* r, int status, Clock::time_point now)
    {
        r->status = status;
        r->t_done = now;
        if (status == OSTFT_OK) {
            ++st.completed;
            st.frames += r->g.compute_frames;
            st.bytes_in += sizeof(float) * r->L;
            st.bytes_out += sizeof(float) * r->g.out_floats;
            lat_ms.push_back(std::chrono::duration<double, std::milli>(now - r->t_submit).count());
        } else {
            ++st.failed;
        }
        t_last = now;
    }

    void scheduler_loop();
    void completion_loop();
    int launch(Batch* b);
};

int ostft_server::launch(Batch* b)
{
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess)
        return OSTFT_ERR_NO_DEVICE;
    {
        std::vector<std::tuple<void*, const void*, size_t>> jobs;
        for (Request* r : b->reqs)
            if (r->L)
                jobs.emplace_back(arena_h + r->in_off, r->in, sizeof(float) * r->L);
        copy_in->copy(jobs);
    }
    if (b->in_bytes)
        e = hipMemcpyAsync(arena_d + b->start, arena_h + b->start, b->in_bytes, hipMemcpyHostToDevice, h2d);
    if (e == hipSuccess)
        e = hipEventRecord(b->ev_in, h2d);
    if (e == hipSuccess)
        e = hipStreamWaitEvent(compute, b->ev_in, 0);
    if (e != hipSuccess) {
        b->err = std::string("input copy: ") + hipGetErrorString(e);
        return OSTFT_ERR_HIP;
    }
    std::vector<ostft_job> jobs(b->reqs.size());
    for (size_t i = 0; i < jobs.size(); ++i) {
        const Request* r = b->reqs[i];
        jobs[i].in = reinterpret_cast<const float*>(arena_d + r->in_off);
        jobs[i].full_size = r->L;
        jobs[i].out = reinterpret_cast<float*>(arena_d + r->out_off);
        jobs[i].out_cap = r->g.out_floats;
        jobs[i].params = r->p;
    }
    const int rc = ostft_run_batch_device(ctx, jobs.data(), (uint32_t)jobs.size(), compute);
    if (rc != OSTFT_OK) {
        b->err = ostft_last_error(ctx);
        // nothing of this batch may be left running when its chunk is freed
        (void)hipStreamSynchronize(compute);
        return rc;
    }
    e = hipEventRecord(b->ev_comp, compute);
    if (e == hipSuccess)
        e = hipStreamWaitEvent(d2h, b->ev_comp, 0);
    if (e == hipSuccess && b->out_bytes)
        e = hipMemcpyAsync(arena_h + b->start + b->in_bytes, arena_d + b->start + b->in_bytes, b->out_bytes,
                           hipMemcpyDeviceToHost, d2h);
    if (e != hipSuccess) {
        b->err = std::string("output copy: ") + hipGetErrorString(e);
        return OSTFT_ERR_HIP;
    }
    return OSTFT_OK;
}

void ostft_server::scheduler_loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_sched.wait(lk, [&] {
            return (stopping && pending.empty()) || (!pending.empty() && inflight.size() < cfg.max_inflight);
        });
        if (pending.empty())
            break;  // stopping, queue drained
        // admission: the queued requests that fit, in order
        size_t off = 0, len = 0;
        ring.free_region(&off, &len);
        std::unique_ptr<Batch> b(new Batch());
        size_t in_b = 0, out_b = 0;
        for (Request* r : pending) {
            if (b->reqs.size() >= cfg.max_batch)
                break;
            const size_t ib = align_up(sizeof(float) * r->L), ob = align_up(sizeof(float) * r->g.out_floats);
            if (in_b + ib + out_b + ob > len)
                break;
            b->reqs.push_back(r);
            in_b += ib;
            out_b += ob;
        }
        if (b->reqs.empty()) {
            // the head request waits for arena space; a retiring batch wakes us
            if (inflight.empty()) {
                // cannot happen (submit rejects requests larger than the arena);
                // fail it rather than wait forever
                Request* r = pending.front();
                pending.pop_front();
                finish(r, OSTFT_ERR_TOO_LARGE, Clock::now());
                cv_done.notify_all();
                continue;
            }
            cv_sched.wait(lk);
            continue;
        }
        pending.erase(pending.begin(), pending.begin() + (ptrdiff_t)b->reqs.size());
        b->start = off;
        b->in_bytes = in_b;
        b->out_bytes = out_b;
        size_t pi = off, po = off + in_b;
        for (Request* r : b->reqs) {
            r->in_off = pi;
            r->out_off = po;
            pi += align_up(sizeof(float) * r->L);
            po += align_up(sizeof(float) * r->g.out_floats);
        }
        ring.push(off, in_b + out_b);
        b->ev_in = take_event();
        b->ev_comp = take_event();
        b->ev_done = take_event();
        ++st.batches;
        Batch* raw = b.get();
        lk.unlock();
        int rc = (raw->ev_in && raw->ev_comp && raw->ev_done) ? launch(raw) : OSTFT_ERR_HIP;
        if (rc == OSTFT_OK) {
            const hipError_t e = hipEventRecord(raw->ev_done, d2h);
            if (e != hipSuccess) {
                raw->err = hipGetErrorString(e);
                rc = OSTFT_ERR_HIP;
            }
        }
        if (rc != OSTFT_OK) {
            // nothing of the batch may still run when its chunk is freed: wait
            // for this server's streams (not for other work on the device)
            for (hipStream_t st : {h2d, compute, d2h})
                (void)hipStreamSynchronize(st);
            raw->status = rc;
        }
        lk.lock();
        inflight.push_back(std::move(b));
        cv_comp.notify_one();
    }
    sched_done = true;
    cv_comp.notify_one();
}

void ostft_server::completion_loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_comp.wait(lk, [&] { return !inflight.empty() || sched_done; });
        if (inflight.empty())
            break;
        Batch* b = inflight.front().get();
        lk.unlock();
        int status = b->status;
        if (status == OSTFT_OK) {
            const hipError_t e = hipEventSynchronize(b->ev_done);
            if (e != hipSuccess) {
                status = OSTFT_ERR_HIP;
                b->err = hipGetErrorString(e);
            }
        }
        if (status == OSTFT_OK) {
            std::vector<std::tuple<void*, const void*, size_t>> jobs;
            for (Request* r : b->reqs)
                if (r->g.out_floats)
                    jobs.emplace_back(r->out, arena_h + r->out_off, sizeof(float) * r->g.out_floats);
            copy_out->copy(jobs);
        }
        const Clock::time_point now = Clock::now();
        lk.lock();
        for (Request* r : b->reqs)
            finish(r, status, now);
        if (status != OSTFT_OK)
            last_error = b->err;
        event_pool.push_back(b->ev_in);
        event_pool.push_back(b->ev_comp);
        event_pool.push_back(b->ev_done);
        ring.pop();
        inflight.pop_front();
        cv_done.notify_all();
        cv_sched.notify_one();
    }
}

extern "C" {

int ostft_server_create(int32_t device, const ostft_server_config* config, ostft_server** out)
{
    if (!out)
        return OSTFT_ERR_INVALID_ARG;
    *out = nullptr;
    ostft_server_config cfg{};
    if (config)
        cfg = *config;
    else
        cfg.use_graphs = 1;
    if (cfg.arena_bytes == 0)
        cfg.arena_bytes = 1ull << 30;
    if (cfg.max_batch == 0)
        cfg.max_batch = 512;
    if (cfg.max_inflight == 0)
        cfg.max_inflight = 3;
    cfg.arena_bytes = align_up(cfg.arena_bytes);
    std::unique_ptr<ostft_server> s(new ostft_server());
    s->device = device;
    s->cfg = cfg;
    s->cfg.autotune_like = nullptr;  // read at creation only
    int rc = ostft_create(device, &s->ctx);
    if (rc != OSTFT_OK)
        return rc;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&s->h2d, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&s->d2h, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipMalloc(&s->arena_d, cfg.arena_bytes);
    if (e == hipSuccess)
        e = hipHostMalloc(&s->arena_h, cfg.arena_bytes, hipHostMallocDefault);
    if (e == hipSuccess) {
        // touch the pinned mirror and move it both ways once, so the first
        // requests through each part of the ring do not pay first-use costs
        std::memset(s->arena_h, 0, cfg.arena_bytes);
        e = hipMemcpy(s->arena_d, s->arena_h, cfg.arena_bytes, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(s->arena_h, s->arena_d, cfg.arena_bytes, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess)
        rc = ostft_batch_config(s->ctx, 0, 0, cfg.use_graphs ? 1 : 0, nullptr);
    if (e == hipSuccess && rc == OSTFT_OK && cfg.autotune) {
        for (int exp = 6; exp <= 15 && rc == OSTFT_OK; ++exp)
            rc = ostft_autotune_batch(s->ctx, exp, cfg.autotune_samples ? cfg.autotune_samples : 1ull << 26,
                                      cfg.autotune_like, nullptr, nullptr, nullptr, nullptr);
    }
    if (e != hipSuccess || rc != OSTFT_OK) {
        if (s->arena_d)
            (void)hipFree(s->arena_d);
        if (s->arena_h)
            (void)hipHostFree(s->arena_h);
        for (hipStream_t st : {s->h2d, s->compute, s->d2h})
            if (st)
                (void)hipStreamDestroy(st);
        ostft_destroy(s->ctx);
        return e != hipSuccess ? OSTFT_ERR_HIP : rc;
    }
    s->ring.cap = cfg.arena_bytes;
    s->copy_in.reset(new CopyPool(3));
    s->copy_out.reset(new CopyPool(3));
    ostft_server* raw = s.get();
    raw->sched = std::thread([raw] { raw->scheduler_loop(); });
    raw->comp = std::thread([raw] { raw->completion_loop(); });
    *out = s.release();
    return OSTFT_OK;
}

void ostft_server_destroy(ostft_server* srv)
{
    if (!srv)
        return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true;
    }
    srv->cv_sched.notify_all();
    srv->sched.join();
    srv->comp.join();
    (void)hipSetDevice(srv->device);
    for (hipStream_t st : {srv->h2d, srv->compute, srv->d2h})
        (void)hipStreamSynchronize(st);
    for (hipEvent_t e : srv->event_pool)
        (void)hipEventDestroy(e);
    (void)hipFree(srv->arena_d);
    (void)hipHostFree(srv->arena_h);
    for (hipStream_t st : {srv->h2d, srv->compute, srv->d2h})
        (void)hipStreamDestroy(st);
    ostft_destroy(srv->ctx);
    delete srv;
}

int ostft_server_submit(ostft_server* srv, const float* in, uint64_t full_size, const ostft_params* params,
                        float* out, uint64_t out_cap, uint64_t* ticket)
{
    if (!srv || !params || !out || !ticket || (!in && full_size > 0))
        return OSTFT_ERR_INVALID_ARG;
    std::unique_ptr<Request> r(new Request());
    int rc = ostft_get_geometry(params, full_size, &r->g);
    std::lock_guard<std::mutex> lk(srv->mu);
    if (rc != OSTFT_OK) {
        srv->last_error = "invalid request parameters";
        return rc;
    }
    if (out_cap < r->g.out_floats) {
        srv->last_error = "output buffer too small";
        return OSTFT_ERR_CAPACITY;
    }
    if (align_up(sizeof(float) * full_size) + align_up(sizeof(float) * r->g.out_floats) > srv->cfg.arena_bytes) {
        srv->last_error = "request footprint exceeds the server arena";
        return OSTFT_ERR_TOO_LARGE;
    }
    if (srv->stopping)
        return OSTFT_ERR_SHUTDOWN;
    r->id = srv->next_id++;
    r->in = in;
    r->L = full_size;
    r->p = *params;
    r->out = out;
    r->t_submit = Clock::now();
    if (!srv->have_first) {
        srv->t_first = r->t_submit;
        srv->have_first = true;
    }
    ++srv->st.submitted;
    *ticket = r->id;
    srv->pending.push_back(r.get());
    srv->reqs.emplace(r->id, std::move(r));
    srv->cv_sched.notify_one();
    return OSTFT_OK;
}

int ostft_server_wait(ostft_server* srv, uint64_t ticket, int32_t timeout_ms, uint64_t* out_len)
{
    if (!srv)
        return OSTFT_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end())
        return OSTFT_ERR_INVALID_ARG;
    Request* r = it->second.get();
    auto done = [&] { return r->status >= 0; };
    if (timeout_ms < 0)
        srv->cv_done.wait(lk, done);
    else if (!srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
        return OSTFT_ERR_TIMEOUT;
    const int status = r->status;
    if (out_len)
        *out_len = status == OSTFT_OK ? r->g.out_floats : 0;
    srv->reqs.erase(it);
    return status;
}

int ostft_server_stats_get(ostft_server* srv, ostft_server_stats* out, int32_t reset)
{
    if (!srv || !out)
        return OSTFT_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    ostft_server_stats s = srv->st;
    std::vector<double> v = srv->lat_ms;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            // nearest-rank percentile
            size_t k = (size_t)std::ceil(q * (double)v.size());
            k = std::min(std::max<size_t>(k, 1), v.size());
            return v[k - 1];
        };
        s.latency_p50_ms = pct(0.50);
        s.latency_p99_ms = pct(0.99);
        s.latency_max_ms = v.back();
        double sum = 0.0;
        for (double x : v)
            sum += x;
        s.latency_mean_ms = sum / (double)v.size();
    }
    s.mean_batch = s.batches ? (double)(s.completed + s.failed) / (double)s.batches : 0.0;
    s.span_s = srv->have_first && srv->t_last > srv->t_first
                   ? std::chrono::duration<double>(srv->t_last - srv->t_first).count()
                   : 0.0;
    *out = s;
    if (reset) {
        srv->st = ostft_server_stats{};
        srv->lat_ms.clear();
        srv->have_first = false;
    }
    return OSTFT_OK;
}

const char* ostft_server_last_error(const ostft_server* srv) { return srv ? srv->last_error.c_str() : ""; }

}  // extern "C"
