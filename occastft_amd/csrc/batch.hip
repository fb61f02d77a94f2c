// batch.hip -- many requests of uneven length, window size and overlap in one
// launch sequence: one persistent launch per window-size class present, each
// pulling tiles (G frames of one request) from a device-side counter.
//
// The descriptor block (per-class headers, the requests' FusedArgs, the
// tile -> request table) is built on the host in pinned memory and uploaded
// with one copy before the launches.  The launches themselves only name the
// block, so for a given set of classes they are the same every time: they are
// captured once into a HIP graph and replayed.
//
// Requests of windows >= 2^15 (multi-pass path) run after the persistent
// launches, one request at a time, on the same stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <vector>

#include "batch.hpp"
#include "multipass.hpp"

namespace ostft {

namespace {

constexpr int LOG_M_MIN = 5, LOG_M_MAX = 13, NCLS = LOG_M_MAX - LOG_M_MIN + 1;
constexpr size_t ALIGN = 256;

size_t align_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

// One descriptor block: device copy, pinned host mirror, and the event after
// the last launch that reads it (the mirror and the block are reused only
// after it).
struct Slot {
    char* d = nullptr;
    char* h = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool used = false;
};

}  // namespace

struct BatchState {
    Slot slot[2];
    int next = 0;
    int cu_count = 0;
    int use_graphs = 1;
    int per_cu[NCLS] = {};  // persistent workgroups per CU; 0: the occupancy limit
    BatchClassInfo info[NCLS] = {};
    bool info_ok[NCLS] = {};
    std::map<std::pair<uint32_t, int>, hipGraphExec_t> graphs;  // (class mask, slot) -> replayable launches
    hipStream_t capture = nullptr;

    void drop_graphs()
    {
        for (auto& kv : graphs)
            (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
};

namespace {

int state(ostft_ctx* ctx, BatchState** out)
{
    BatchState*& b = ctx_batch(ctx);
    if (!b) {
        BatchState* s = new BatchState();
        hipDeviceProp_t prop;
        hipError_t e = hipGetDeviceProperties(&prop, ctx_device(ctx));
        if (e == hipSuccess)
            e = hipStreamCreateWithFlags(&s->capture, hipStreamNonBlocking);
        for (int i = 0; e == hipSuccess && i < 2; ++i)
            e = hipEventCreateWithFlags(&s->slot[i].done, hipEventDisableTiming);
        if (e != hipSuccess) {
            if (s->capture)
                (void)hipStreamDestroy(s->capture);
            for (auto& sl : s->slot)
                if (sl.done)
                    (void)hipEventDestroy(sl.done);
            delete s;
            return ctx_hip_fail(ctx, e, "batch state");
        }
        s->cu_count = prop.multiProcessorCount;
        b = s;
    }
    *out = b;
    return OSTFT_OK;
}

int class_info(ostft_ctx* ctx, BatchState* st, int c)
{
    if (!st->info_ok[c]) {
        const hipError_t e = batch_class_info(c + LOG_M_MIN, &st->info[c]);
        if (e != hipSuccess)
            return ctx_hip_fail(ctx, e, "batch kernel occupancy");
        if (st->info[c].max_blocks_per_cu < 1)
            return ctx_fail(ctx, OSTFT_ERR_HIP, "batch kernel does not fit on a CU");
        st->info_ok[c] = true;
    }
    return OSTFT_OK;
}

uint32_t class_grid(const BatchState* st, int c)
{
    const int mx = st->info[c].max_blocks_per_cu;
    const int pc = st->per_cu[c] > 0 ? std::min(st->per_cu[c], mx) : mx;
    return (uint32_t)pc * (uint32_t)st->cu_count;
}

hipError_t launch_classes(BatchState* st, uint32_t mask, void* base, hipStream_t s)
{
    for (int c = 0; c < NCLS; ++c) {
        if (!(mask >> c & 1u))
            continue;
        BatchClassArgs a;
        a.base = base;
        a.header_off = (uint64_t)c * sizeof(BatchClassHeader);
        const hipError_t e = launch_batch(c + LOG_M_MIN, a, class_grid(st, c), s);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

// The launches for `mask` reading slot `si`, captured once.
hipError_t graph_for(BatchState* st, uint32_t mask, int si, hipGraphExec_t* out)
{
    const auto key = std::make_pair(mask, si);
    auto it = st->graphs.find(key);
    if (it != st->graphs.end()) {
        *out = it->second;
        return hipSuccess;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(st->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
        return e;
    const hipError_t le = launch_classes(st, mask, st->slot[si].d, st->capture);
    e = hipStreamEndCapture(st->capture, &g);
    if (le != hipSuccess) {
        if (g)
            (void)hipGraphDestroy(g);
        return le;
    }
    if (e != hipSuccess)
        return e;
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
        return e;
    st->graphs[key] = ex;
    *out = ex;
    return hipSuccess;
}

hipError_t ensure_slot(BatchState* st, int si, size_t bytes)
{
    Slot& sl = st->slot[si];
    if (sl.cap >= bytes)
        return hipSuccess;
    // graphs name the old block: drop them (nothing in flight reads it: the
    // caller waited on the slot's event)
    st->drop_graphs();
    if (sl.d)
        (void)hipFree(sl.d);
    if (sl.h)
        (void)hipHostFree(sl.h);
    sl.d = nullptr;
    sl.h = nullptr;
    sl.cap = 0;
    const size_t cap = std::max<size_t>(align_up(bytes + bytes / 2), 64 * 1024);
    hipError_t e = hipMalloc(&sl.d, cap);
    if (e == hipSuccess)
        e = hipHostMalloc(&sl.h, cap, hipHostMallocDefault);
    if (e != hipSuccess) {
        if (sl.d)
            (void)hipFree(sl.d);
        sl.d = nullptr;
        sl.h = nullptr;
        return e;
    }
    sl.cap = cap;
    return hipSuccess;
}

}  // namespace

int run_batch(ostft_ctx* ctx, const ostft_job* jobs, uint32_t njobs, hipStream_t s)
{
    BatchState* st = nullptr;
    int rc = state(ctx, &st);
    if (rc != OSTFT_OK)
        return rc;
    // 1. validate every request first: a bad request fails the whole call
    //    before anything is launched
    std::vector<FusedArgs> args(njobs);
    std::vector<int> lm(njobs, 0);
    for (uint32_t i = 0; i < njobs; ++i) {
        const ostft_job& j = jobs[i];
        rc = ctx_prepare(ctx, &j.params, j.in, j.full_size, j.out, j.out_cap, &args[i], &lm[i]);
        if (rc != OSTFT_OK)
            return ctx_fail(ctx, rc, "job " + std::to_string(i) + ": " + ostft_last_error(ctx));
    }
    // 2. group by class; tiles per request
    std::vector<uint32_t> members[NCLS];
    uint64_t ntiles[NCLS] = {};
    std::vector<uint32_t> large;
    uint32_t mask = 0;
    for (uint32_t i = 0; i < njobs; ++i) {
        if (args[i].frames == 0)
            continue;
        if (!fused_supported(lm[i])) {
            large.push_back(i);
            continue;
        }
        const int c = lm[i] - LOG_M_MIN;
        rc = class_info(ctx, st, c);
        if (rc != OSTFT_OK)
            return rc;
        const uint64_t G = (uint64_t)st->info[c].frames_per_tile;
        members[c].push_back(i);
        ntiles[c] += (args[i].frames + G - 1) / G;
        mask |= 1u << c;
    }
    for (int c = 0; c < NCLS; ++c)
        if (ntiles[c] + class_grid(st, c) >= 0xffffffffull)
            return ctx_fail(ctx, OSTFT_ERR_TOO_LARGE, "more than 2^32 tiles of one window size in a batch");
    if (mask) {
        // 3. descriptor block layout
        size_t off = align_up(sizeof(BatchClassHeader) * NCLS);
        size_t jobs_off[NCLS] = {}, tiles_off[NCLS] = {};
        for (int c = 0; c < NCLS; ++c) {
            if (!(mask >> c & 1u))
                continue;
            jobs_off[c] = off;
            off = align_up(off + sizeof(BatchJob) * members[c].size());
            tiles_off[c] = off;
            off = align_up(off + sizeof(uint32_t) * ntiles[c]);
        }
        const int si = st->next;
        st->next ^= 1;
        Slot& sl = st->slot[si];
        if (sl.used) {
            const hipError_t e = hipEventSynchronize(sl.done);
            if (e != hipSuccess)
                return ctx_hip_fail(ctx, e, "batch slot wait");
        }
        hipError_t e = ensure_slot(st, si, off);
        if (e != hipSuccess)
            return ctx_hip_fail(ctx, e, "batch descriptor block");
        // 4. fill the pinned mirror
        BatchClassHeader* hdr = reinterpret_cast<BatchClassHeader*>(sl.h);
        std::memset(hdr, 0, sizeof(BatchClassHeader) * NCLS);
        for (int c = 0; c < NCLS; ++c) {
            if (!(mask >> c & 1u))
                continue;
            hdr[c].jobs_off = jobs_off[c];
            hdr[c].tiles_off = tiles_off[c];
            hdr[c].ntiles = (uint32_t)ntiles[c];
            hdr[c].counter = 0;
            BatchJob* bj = reinterpret_cast<BatchJob*>(sl.h + jobs_off[c]);
            uint32_t* tj = reinterpret_cast<uint32_t*>(sl.h + tiles_off[c]);
            const uint64_t G = (uint64_t)st->info[c].frames_per_tile;
            uint64_t t = 0;
            for (size_t k = 0; k < members[c].size(); ++k) {
                const FusedArgs& a = args[members[c][k]];
                bj[k].a = a;
                bj[k].tile_begin = t;
                const uint64_t nt = (a.frames + G - 1) / G;
                std::fill(tj + t, tj + t + nt, (uint32_t)k);
                t += nt;
            }
        }
        // 5. upload, then the launches (replayed graph or direct)
        e = hipMemcpyAsync(sl.d, sl.h, off, hipMemcpyHostToDevice, s);
        if (e != hipSuccess)
            return ctx_hip_fail(ctx, e, "batch descriptor upload");
        if (st->use_graphs) {
            hipGraphExec_t ex = nullptr;
            e = graph_for(st, mask, si, &ex);
            if (e == hipSuccess)
                e = hipGraphLaunch(ex, s);
        } else {
            e = launch_classes(st, mask, sl.d, s);
        }
        if (e == hipSuccess)
            e = hipEventRecord(sl.done, s);
        if (e != hipSuccess)
            return ctx_hip_fail(ctx, e, "batch launch");
        sl.used = true;
    }
    for (uint32_t i : large) {
        rc = run_multipass(ctx, lm[i], args[i], s);
        if (rc != OSTFT_OK)
            return rc;
    }
    return OSTFT_OK;
}

int batch_config(ostft_ctx* ctx, int exp, int blocks_per_cu, int use_graphs, int* current)
{
    BatchState* st = nullptr;
    int rc = state(ctx, &st);
    if (rc != OSTFT_OK)
        return rc;
    if (use_graphs >= 0)
        st->use_graphs = use_graphs ? 1 : 0;
    if (exp == 0)
        return OSTFT_OK;
    const int c = exp - 1 - LOG_M_MIN;
    if (c < 0 || c >= NCLS)
        return ctx_fail(ctx, OSTFT_ERR_UNSUPPORTED, "the batch kernels cover window exponents 6..14");
    rc = class_info(ctx, st, c);
    if (rc != OSTFT_OK)
        return rc;
    if (blocks_per_cu > 0 && blocks_per_cu != st->per_cu[c]) {
        // captured launches carry the grid: wait for work in flight, recapture
        for (auto& sl : st->slot)
            if (sl.used)
                (void)hipEventSynchronize(sl.done);
        st->drop_graphs();
        st->per_cu[c] = std::min(blocks_per_cu, st->info[c].max_blocks_per_cu);
    }
    if (current)
        *current = (int)(class_grid(st, c) / (uint32_t)st->cu_count);
    return OSTFT_OK;
}

// Times the persistent launch of one class over a synthetic ragged batch
// (`samples` samples in requests of 1x..16x the window, overlap 0.5, Hann,
// DC removal, compat power) for each workgroups-per-CU setting and keeps the
// fastest.
int autotune_batch(ostft_ctx* ctx, int exp, uint64_t samples, int* best_blocks_per_cu, double* best_ms)
{
    BatchState* st = nullptr;
    int rc = state(ctx, &st);
    if (rc != OSTFT_OK)
        return rc;
    const int c = exp - 1 - LOG_M_MIN;
    if (c < 0 || c >= NCLS)
        return ctx_fail(ctx, OSTFT_ERR_UNSUPPORTED, "the batch kernels cover window exponents 6..14");
    rc = class_info(ctx, st, c);
    if (rc != OSTFT_OK)
        return rc;
    const uint64_t N = 1ull << exp;
    if (samples < N)
        samples = N;
    // ragged request lengths from a fixed LCG
    std::vector<uint64_t> lens;
    uint64_t total = 0, lcg = 0x5EED0005ull;
    while (total < samples) {
        lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
        uint64_t L = N * (1 + (lcg >> 60)) + ((lcg >> 20) % N);
        L = std::min(L, samples - total < N ? N : samples - total);
        lens.push_back(L);
        total += L;
    }
    ostft_params p;
    std::memset(&p, 0, sizeof(p));
    p.exp = exp;
    p.overlap = 0.5f;
    p.window = OSTFT_WIN_HANNING;
    p.output = OSTFT_OUT_COMPAT_POWER;
    p.flags = OSTFT_FLAG_REMOVE_DC;
    std::vector<ostft_job> jobs(lens.size());
    uint64_t in_off = 0, out_off = 0;
    for (size_t i = 0; i < lens.size(); ++i) {
        ostft_geometry g;
        rc = ostft_get_geometry(&p, lens[i], &g);
        if (rc != OSTFT_OK)
            return ctx_fail(ctx, rc, "autotune geometry");
        jobs[i].full_size = lens[i];
        jobs[i].out_cap = g.out_floats;
        jobs[i].params = p;
        jobs[i].in = reinterpret_cast<const float*>(in_off);  // offsets until the buffers exist
        jobs[i].out = reinterpret_cast<float*>(out_off);
        in_off += lens[i];
        out_off += (g.out_floats + 3) & ~3ull;
    }
    float* din = nullptr;
    float* dout = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s = nullptr;
    hipError_t e = hipMalloc(&din, sizeof(float) * in_off);
    if (e == hipSuccess)
        e = hipMalloc(&dout, sizeof(float) * out_off);
    if (e == hipSuccess)
        e = hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(din), 0x3e800000u /* 0.25f */, in_off, nullptr);
    if (e == hipSuccess)
        e = hipEventCreate(&e0);
    if (e == hipSuccess)
        e = hipEventCreate(&e1);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    if (e == hipSuccess) {
        for (auto& j : jobs) {
            j.in = din + reinterpret_cast<uint64_t>(j.in);
            j.out = dout + reinterpret_cast<uint64_t>(j.out);
        }
        const int saved_graphs = st->use_graphs, saved_pc = st->per_cu[c];
        st->use_graphs = 0;
        const int mx = st->info[c].max_blocks_per_cu;
        std::vector<int> cand;
        for (int k = 1; k < mx; k *= 2)
            cand.push_back(k);
        cand.push_back(mx);
        int best = mx;
        float best_t = 1e30f;
        for (int k : cand) {
            st->per_cu[c] = k;
            float tmin = 1e30f;
            for (int r = 0; r < 4 && rc == OSTFT_OK && e == hipSuccess; ++r) {
                e = hipEventRecord(e0, s);
                if (e == hipSuccess)
                    rc = run_batch(ctx, jobs.data(), (uint32_t)jobs.size(), s);
                if (e == hipSuccess && rc == OSTFT_OK)
                    e = hipEventRecord(e1, s);
                if (e == hipSuccess && rc == OSTFT_OK)
                    e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess && rc == OSTFT_OK)
                    e = hipEventElapsedTime(&ms, e0, e1);
                if (r > 0)  // the first run warms caches and tables
                    tmin = std::min(tmin, ms);
            }
            if (rc != OSTFT_OK || e != hipSuccess)
                break;
            if (tmin < best_t) {
                best_t = tmin;
                best = k;
            }
        }
        st->use_graphs = saved_graphs;
        st->per_cu[c] = saved_pc;
        if (rc == OSTFT_OK && e == hipSuccess) {
            for (auto& sl : st->slot)
                if (sl.used)
                    (void)hipEventSynchronize(sl.done);
            st->drop_graphs();
            st->per_cu[c] = best;
            if (best_blocks_per_cu)
                *best_blocks_per_cu = best;
            if (best_ms)
                *best_ms = best_t;
        }
    }
    if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    if (din)
        (void)hipFree(din);
    if (dout)
        (void)hipFree(dout);
    if (e != hipSuccess)
        return ctx_hip_fail(ctx, e, "autotune");
    return rc;
}

void batch_release(ostft_ctx* ctx)
{
    BatchState*& b = ctx_batch(ctx);
    if (!b)
        return;
    for (auto& sl : b->slot) {
        if (sl.used)
            (void)hipEventSynchronize(sl.done);
        if (sl.d)
            (void)hipFree(sl.d);
        if (sl.h)
            (void)hipHostFree(sl.h);
        if (sl.done)
            (void)hipEventDestroy(sl.done);
    }
    b->drop_graphs();
    if (b->capture)
        (void)hipStreamDestroy(b->capture);
    delete b;
    b = nullptr;
}

}  // namespace ostft
