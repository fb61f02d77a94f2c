// batch.hip -- many requests of uneven length, window size and overlap in one
// launch sequence: one launch per window-size class present over that class's
// tile queue (tiles of G frames of one request), either one workgroup per tile
// or a persistent grid pulling tiles from a device-side counter.  The launch
// form, the workgroup shape (G) and the radix plan of each class are measured
// by the autotuner.
//
// The descriptor block (per-class headers, the requests' FusedArgs, the
// tile -> request table) is built on the host in pinned memory and uploaded
// with one copy before the launches.  The launches themselves only name the
// block, so for a given set of classes they are the same every time: they are
// captured once into a HIP graph and replayed.
//
// Requests of windows >= 2^16 (multi-pass path) and requests of the matrix-core
// form run after the tile-queue launches, one request at a time, on the same
// stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "batch.hpp"
#include "multipass.hpp"

namespace ostft {

namespace {

constexpr int LOG_M_MIN = 5, LOG_M_MAX = 14, NCLS = LOG_M_MAX - LOG_M_MIN + 1;
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
    int persistent[NCLS] = {};  // 1: persistent grid claiming tiles; 0: one workgroup per tile
    int per_cu[NCLS] = {};      // persistent workgroups per CU; 0: the occupancy limit
    int shape[NCLS] = {};       // workgroup shape (engine variant / 4); persistent needs 0
    int plan_override[NCLS];    // >= 0: radix plan for this class instead of the context's (autotuning)
    int variant[NCLS] = {};     // engine variant of each class in the current batch
    BatchClassInfo info[NCLS][FUSED_VARIANTS] = {};
    bool info_ok[NCLS][FUSED_VARIANTS] = {};
    // per slot: the launches of every class, captured once; a batch sets
    // each class node's grid (its persistent grid, or one workgroup when the
    // class has no tiles) and replays it
    hipGraph_t graph[2] = {};
    hipGraphExec_t exec[2] = {};
    hipGraphNode_t node[2][NCLS] = {};
    uint32_t node_grid[2][NCLS] = {};
    int node_persistent[2][NCLS] = {};
    int node_variant[2][NCLS] = {};
    hipStream_t capture = nullptr;

    BatchState()
    {
        for (int& p : plan_override)
            p = -1;
    }
    void drop_graph(int i)
    {
        if (exec[i])
            (void)hipGraphExecDestroy(exec[i]);
        if (graph[i])
            (void)hipGraphDestroy(graph[i]);
        exec[i] = nullptr;
        graph[i] = nullptr;
    }
    void drop_graphs()
    {
        for (int i = 0; i < 2; ++i)
            drop_graph(i);
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

int class_info(ostft_ctx* ctx, BatchState* st, int c, int v)
{
    if (!st->info_ok[c][v]) {
        const hipError_t e = batch_class_info(c + LOG_M_MIN, v, &st->info[c][v]);
        if (e != hipSuccess)
            return ctx_hip_fail(ctx, e, "batch kernel occupancy");
        if (st->info[c][v].max_blocks_per_cu < 1)
            return ctx_fail(ctx, OSTFT_ERR_HIP, "batch kernel does not fit on a CU");
        st->info_ok[c][v] = true;
    }
    return OSTFT_OK;
}

// Engine variant of class c: its workgroup shape with radix plan `plan`, or
// the default shape where that pair is not built.
int class_variant(const BatchState* st, int c, int plan)
{
    const int v = st->shape[c] * 4 + plan;
    return fused_variant_ok(c + LOG_M_MIN, v) ? v : plan;
}

// Grid of class c for a batch with `ntiles` tiles of it.
uint32_t class_grid(const BatchState* st, int c, uint64_t ntiles)
{
    if (!st->persistent[c])
        return (uint32_t)std::max<uint64_t>(ntiles, 1);
    const int mx = st->info[c][st->variant[c]].max_blocks_per_cu;
    const int pc = st->per_cu[c] > 0 ? std::min(st->per_cu[c], mx) : mx;
    return (uint32_t)pc * (uint32_t)st->cu_count;
}

hipError_t launch_classes(BatchState* st, uint32_t mask, void* base, hipStream_t s, const uint64_t* ntiles,
                          bool one_block = false)
{
    for (int c = 0; c < NCLS; ++c) {
        if (!(mask >> c & 1u))
            continue;
        BatchClassArgs a;
        a.base = base;
        a.header_off = (uint64_t)c * sizeof(BatchClassHeader);
        const uint32_t grid = one_block ? 1u : class_grid(st, c, ntiles[c]);
        const hipError_t e = launch_batch(c + LOG_M_MIN, st->variant[c], a, grid, st->persistent[c] != 0, s);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

// Captures the launches of every class for slot `si` (grids of one
// workgroup; set per batch) and finds each class's kernel node.
hipError_t capture_slot(BatchState* st, int si)
{
    const uint32_t all = (1u << NCLS) - 1u;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(st->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
        return e;
    const hipError_t le = launch_classes(st, all, st->slot[si].d, st->capture, nullptr, true);
    e = hipStreamEndCapture(st->capture, &g);
    if (le != hipSuccess || e != hipSuccess) {
        if (g)
            (void)hipGraphDestroy(g);
        return le != hipSuccess ? le : e;
    }
    // the captured graph is the chain class 0 -> class NCLS-1
    for (int c = 0; c < NCLS; ++c)
        st->node[si][c] = nullptr;
    size_t n = 0;
    e = hipGraphGetNodes(g, nullptr, &n);
    std::vector<hipGraphNode_t> nodes(n);
    if (e == hipSuccess && n == (size_t)NCLS)
        e = hipGraphGetNodes(g, nodes.data(), &n);
    else if (e == hipSuccess)
        e = hipErrorInvalidValue;
    for (size_t i = 0; e == hipSuccess && i < n; ++i) {
        hipKernelNodeParams kp;
        e = hipGraphKernelNodeGetParams(nodes[i], &kp);
        if (e != hipSuccess)
            break;
        int c = -1;
        for (int k = 0; k < NCLS; ++k)
            if (kp.func == batch_kernel_fn(k + LOG_M_MIN, st->variant[k], st->persistent[k] != 0))
                c = k;
        if (c < 0 || st->node[si][c]) {
            e = hipErrorInvalidValue;
            break;
        }
        st->node[si][c] = nodes[i];
        st->node_grid[si][c] = 1;
        st->node_persistent[si][c] = st->persistent[c];
        st->node_variant[si][c] = st->variant[c];
    }
    hipGraphExec_t ex = nullptr;
    if (e == hipSuccess)
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return e;
    }
    st->graph[si] = g;
    st->exec[si] = ex;
    return hipSuccess;
}

// Replays slot `si`'s launches with the grids this batch needs.
hipError_t replay_slot(BatchState* st, int si, uint32_t mask, const uint64_t* ntiles, hipStream_t s)
{
    for (int c = 0; st->exec[si] && c < NCLS; ++c)
        if (st->node_persistent[si][c] != st->persistent[c] || st->node_variant[si][c] != st->variant[c]) {
            // a class changed form or variant: its node names another kernel
            (void)hipGraphExecDestroy(st->exec[si]);
            (void)hipGraphDestroy(st->graph[si]);
            st->exec[si] = nullptr;
            st->graph[si] = nullptr;
        }
    if (!st->exec[si]) {
        const hipError_t e = capture_slot(st, si);
        if (e != hipSuccess)
            return e;
    }
    for (int c = 0; c < NCLS; ++c) {
        const uint32_t want = (mask >> c & 1u) ? class_grid(st, c, ntiles[c]) : 1u;
        if (want == st->node_grid[si][c])
            continue;
        hipKernelNodeParams kp;
        hipError_t e = hipGraphKernelNodeGetParams(st->node[si][c], &kp);
        if (e != hipSuccess)
            return e;
        kp.gridDim = dim3(want);
        e = hipGraphExecKernelNodeSetParams(st->exec[si], st->node[si][c], &kp);
        if (e != hipSuccess)
            return e;
        st->node_grid[si][c] = want;
    }
    return hipGraphLaunch(st->exec[si], s);
}

hipError_t ensure_slot(BatchState* st, int si, size_t bytes)
{
    Slot& sl = st->slot[si];
    if (sl.cap >= bytes)
        return hipSuccess;
    // this slot's captured launches name the old block: drop them (nothing
    // in flight reads it: the caller waited on the slot's event); the other
    // slot's graph names its own block and may still be running
    st->drop_graph(si);
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
    std::vector<uint32_t> large, mfma;
    uint32_t mask = 0;
    int plan[NCLS];  // radix plan of each class: the context's, read once per class per batch
    for (int c = 0; c < NCLS; ++c)
        plan[c] = -1;
    for (uint32_t i = 0; i < njobs; ++i) {
        if (args[i].frames == 0)
            continue;
        if (!fused_supported(lm[i])) {
            large.push_back(i);
            continue;
        }
        if (jobs[i].params.flags & (OSTFT_FLAG_MFMA_FP16 | OSTFT_FLAG_MFMA_BF16)) {
            // the matrix-core form has no tile-queue kernel: such a request
            // runs on its own launch after the queues, with the one-request
            // kernel (the same result as ostft_run_device)
            mfma.push_back(i);
            continue;
        }
        const int c = lm[i] - LOG_M_MIN;
        if (plan[c] < 0) {
            plan[c] = st->plan_override[c] >= 0 ? st->plan_override[c] : args[i].plan;
            if (!fused_plan_ok(lm[i], plan[c]))
                plan[c] = 0;
            st->variant[c] = class_variant(st, c, plan[c]);
            if (!fused_variant_ok(lm[i], st->variant[c]))
                return ctx_fail(ctx, OSTFT_ERR_INVALID_ARG, "batch variant not built");
            rc = class_info(ctx, st, c, st->variant[c]);
            if (rc != OSTFT_OK)
                return rc;
        }
        // every request of a class runs the class's plan, in tiles of the
        // variant's workgroup
        args[i].plan = plan[c];
        args[i].tile_frames = (int32_t)fused_tile_frames(
            (uint64_t)st->info[c][st->variant[c]].frames_per_tile, 2ull << lm[i], args[i].hop);
        const uint64_t G = (uint64_t)args[i].tile_frames;
        members[c].push_back(i);
        ntiles[c] += (args[i].frames + G - 1) / G;
        mask |= 1u << c;
    }
    for (int c = 0; c < NCLS; ++c)
        if (ntiles[c] >= 0x7fffffffull - 2ull * (uint64_t)st->cu_count * 64u)  // grid limit; claims overrun by <= 2/workgroup
            return ctx_fail(ctx, OSTFT_ERR_TOO_LARGE, "more than 2^31 tiles of one window size in a batch");
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
            uint64_t t = 0;
            for (size_t k = 0; k < members[c].size(); ++k) {
                const FusedArgs& a = args[members[c][k]];
                bj[k].a = a;
                bj[k].tile_begin = t;
                const uint64_t nt = (a.frames + a.tile_frames - 1) / a.tile_frames;
                std::fill(tj + t, tj + t + nt, (uint32_t)k);
                t += nt;
            }
        }
        // 5. upload, then the launches (replayed graph or direct)
        e = hipMemcpyAsync(sl.d, sl.h, off, hipMemcpyHostToDevice, s);
        if (e != hipSuccess)
            return ctx_hip_fail(ctx, e, "batch descriptor upload");
        // the captured sequence holds a node per window size (an absent size
        // still costs one exiting workgroup, ~3.5 us): replay it when at least
        // three sizes are present, launch directly otherwise
        if (st->use_graphs && __builtin_popcount(mask) >= 3) {
            e = replay_slot(st, si, mask, ntiles, s);
        } else {
            e = launch_classes(st, mask, sl.d, s, ntiles);
        }
        if (e == hipSuccess)
            e = hipEventRecord(sl.done, s);
        if (e != hipSuccess)
            return ctx_hip_fail(ctx, e, "batch launch");
        sl.used = true;
    }
    for (uint32_t i : mfma) {
        const int xf = (jobs[i].params.flags & OSTFT_FLAG_MFMA_FP16) ? XF_MFMA_FP16 : XF_MFMA_BF16;
        const hipError_t e = launch_fused(lm[i], xf, args[i], 1, s);
        if (e != hipSuccess)
            return ctx_hip_fail(ctx, e, "batch MFMA launch");
    }
    for (uint32_t i : large) {
        rc = run_multipass(ctx, lm[i], args[i], s);
        if (rc != OSTFT_OK)
            return rc;
    }
    return OSTFT_OK;
}

namespace {

constexpr float AUTOTUNE_MARGIN = 0.02f;

int class_of(ostft_ctx* ctx, int exp, int* c)
{
    *c = exp - 1 - LOG_M_MIN;
    if (*c < 0 || *c >= NCLS)
        return ctx_fail(ctx, OSTFT_ERR_UNSUPPORTED, "the batch kernels cover window exponents 6..15");
    return OSTFT_OK;
}

}  // namespace

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
    int c = 0;
    rc = class_of(ctx, exp, &c);
    if (rc != OSTFT_OK)
        return rc;
    // captured launches take a new setting at their next replay
    if (blocks_per_cu > 0) {
        const int v = ctx_radix_plan(ctx, c + LOG_M_MIN);  // shape 0
        rc = class_info(ctx, st, c, v);
        if (rc != OSTFT_OK)
            return rc;
        st->shape[c] = 0;
        st->persistent[c] = 1;
        st->per_cu[c] = std::min(blocks_per_cu, st->info[c][v].max_blocks_per_cu);
    } else if (blocks_per_cu < 0) {
        st->persistent[c] = 0;
    }
    if (current) {
        int pc = 0;
        if (st->persistent[c]) {
            const int v = ctx_radix_plan(ctx, c + LOG_M_MIN);
            rc = class_info(ctx, st, c, v);
            if (rc != OSTFT_OK)
                return rc;
            const int mx = st->info[c][v].max_blocks_per_cu;
            pc = st->per_cu[c] > 0 ? std::min(st->per_cu[c], mx) : mx;
        }
        *current = pc;
    }
    return OSTFT_OK;
}

int engine_variant(ostft_ctx* ctx, int exp, int shape, int plan, int* cur_shape, int* cur_plan)
{
    BatchState* st = nullptr;
    int rc = state(ctx, &st);
    if (rc != OSTFT_OK)
        return rc;
    int c = 0;
    rc = class_of(ctx, exp, &c);
    if (rc != OSTFT_OK)
        return rc;
    const int log_m = c + LOG_M_MIN;
    if (plan >= 0 && !fused_plan_ok(log_m, plan))
        return ctx_fail(ctx, OSTFT_ERR_UNSUPPORTED, "radix plan " + std::to_string(plan) +
                                                        " is not built for window exponent " + std::to_string(exp));
    const int new_shape = shape >= 0 ? shape : st->shape[c];
    const int new_plan = plan >= 0 ? plan : ctx_radix_plan(ctx, log_m);
    if (new_shape < 0 || new_shape > 2 || !fused_variant_ok(log_m, new_shape * 4 + new_plan))
        return ctx_fail(ctx, OSTFT_ERR_UNSUPPORTED,
                        "workgroup shape " + std::to_string(new_shape) + " with radix plan " + std::to_string(new_plan) +
                            " is not built for window exponent " + std::to_string(exp));
    if (plan >= 0)
        ctx_set_radix_plan(ctx, log_m, plan);
    st->shape[c] = new_shape;
    if (new_shape != 0)
        st->persistent[c] = 0;  // the persistent kernel has shape 0 only
    if (cur_shape)
        *cur_shape = st->shape[c];
    if (cur_plan)
        *cur_plan = ctx_radix_plan(ctx, log_m);
    return OSTFT_OK;
}

// Times one class over a synthetic ragged batch (`samples` samples in 32
// requests with the window, overlap, output layout and flags of `like`, or
// without it overlap 0.5, Hann, DC removal, compat power) under every built
// variant (radix plan x workgroup shape) in the dispatch form, and for
// shape 0 in the persistent form at each workgroups-per-CU setting.  The
// default (variant 0, one workgroup per tile) is timed first and stays unless
// a candidate beats it by more than AUTOTUNE_MARGIN, in the sweep and again
// when the two are timed alternately: most candidates land within a percent
// or two of each other, where one timing's noise would otherwise decide.  Which variant wins depends on the requests: the N/2+1
// layouts write half the bytes of the N-value ones per frame, so the same
// candidates rank differently under them.
int autotune_batch(ostft_ctx* ctx, int exp, uint64_t samples, const ostft_params* like, int* best_blocks_per_cu,
                   int* best_shape, int* best_plan, double* best_ms)
{
    BatchState* st = nullptr;
    int rc = state(ctx, &st);
    if (rc != OSTFT_OK)
        return rc;
    int c = 0;
    rc = class_of(ctx, exp, &c);
    if (rc != OSTFT_OK)
        return rc;
    const int log_m = c + LOG_M_MIN;
    // candidates: (variant, workgroups per CU); 0 per CU = one workgroup per tile
    std::vector<std::pair<int, int>> cand;
    for (int v = 0; v < FUSED_VARIANTS; ++v) {
        if (!fused_variant_ok(log_m, v))
            continue;
        rc = class_info(ctx, st, c, v);
        if (rc != OSTFT_OK)
            return rc;
        cand.push_back({v, 0});
        if (v < 4) {  // shape 0: the persistent form too
            const int mx = st->info[c][v].max_blocks_per_cu;
            for (int k = 1; k < mx; k *= 2)
                cand.push_back({v, k});
            cand.push_back({v, mx});
        }
    }
    const uint64_t N = 1ull << exp;
    if (samples < N)
        samples = N;
    // 32 requests of 1/2 .. 3/2 x samples/32 each (ragged, not window
    // multiples): few enough that the host's preparation of a batch stays
    // shorter than the GPU's work on the previous one
    std::vector<uint64_t> lens;
    const uint64_t per = std::max<uint64_t>(samples / 32, N);
    for (int i = 0; i < 32; ++i)
        lens.push_back(per / 2 + (per * (uint64_t)(i % 8)) / 8 + (uint64_t)(i * 37) % N);
    ostft_params p;
    if (like) {
        p = *like;
        p.exp = exp;
        if (p.flags & (OSTFT_FLAG_MFMA_FP16 | OSTFT_FLAG_MFMA_BF16))
            return ctx_fail(ctx, OSTFT_ERR_INVALID_ARG,
                            "autotune: matrix-core requests do not run the batch kernels");
    } else {
        std::memset(&p, 0, sizeof(p));
        p.exp = exp;
        p.overlap = 0.5f;
        p.window = OSTFT_WIN_HANNING;
        p.output = OSTFT_OUT_COMPAT_POWER;
        p.flags = OSTFT_FLAG_REMOVE_DC;
    }
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
        e = hipEventCreate(&e0);
    if (e == hipSuccess)
        e = hipEventCreate(&e1);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(din), 0x3e800000u /* 0.25f */, in_off, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e == hipSuccess) {
        for (auto& j : jobs) {
            j.in = din + reinterpret_cast<uint64_t>(j.in);
            j.out = dout + reinterpret_cast<uint64_t>(j.out);
        }
        const int saved_graphs = st->use_graphs, saved_pc = st->per_cu[c], saved_ps = st->persistent[c],
                  saved_shape = st->shape[c];
        st->use_graphs = 0;
        // the GPU time of one batch under candidate (variant, per CU): min
        // over `rounds` - 1 timed rounds after a warming one
        auto timed = [&](const std::pair<int, int>& cd, int rounds) {
            const int v = cd.first, k = cd.second;
            st->shape[c] = v >> 2;
            st->plan_override[c] = v & 3;
            st->persistent[c] = k > 0;
            st->per_cu[c] = k;
            // run; e0; run; e1: the second batch is prepared while the first
            // runs, so e0 -> e1 is the GPU's time for one batch
            float tmin = 1e30f;
            for (int r = 0; r < rounds && rc == OSTFT_OK && e == hipSuccess; ++r) {
                rc = run_batch(ctx, jobs.data(), (uint32_t)jobs.size(), s);
                if (rc == OSTFT_OK)
                    e = hipEventRecord(e0, s);
                if (rc == OSTFT_OK && e == hipSuccess)
                    rc = run_batch(ctx, jobs.data(), (uint32_t)jobs.size(), s);
                if (rc == OSTFT_OK && e == hipSuccess)
                    e = hipEventRecord(e1, s);
                if (rc == OSTFT_OK && e == hipSuccess)
                    e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (rc == OSTFT_OK && e == hipSuccess)
                    e = hipEventElapsedTime(&ms, e0, e1);
                if (r > 0)  // the first round warms caches and tables
                    tmin = std::min(tmin, ms);
            }
            return tmin;
        };
        std::pair<int, int> best{0, 0};
        float best_t = 1e30f;
        for (const auto& cd : cand) {
            const float tmin = timed(cd, 5);
            if (rc != OSTFT_OK || e != hipSuccess)
                break;
            // the first candidate is the default; later ones must win clearly
            if (best_t == 1e30f || tmin < best_t * (1.0f - AUTOTUNE_MARGIN)) {
                best_t = tmin;
                best = cd;
            }
        }
        // a winner other than the default is timed against the default
        // again, alternately, and kept only if it still wins by the margin:
        // one lucky timing in the sweep does not decide
        if (rc == OSTFT_OK && e == hipSuccess && best != cand.front()) {
            float td = 1e30f, tb = 1e30f;
            for (int r = 0; r < 3 && rc == OSTFT_OK && e == hipSuccess; ++r) {
                td = std::min(td, timed(cand.front(), 3));
                tb = std::min(tb, timed(best, 3));
            }
            if (rc == OSTFT_OK && e == hipSuccess && !(tb < td * (1.0f - AUTOTUNE_MARGIN))) {
                best = cand.front();
                best_t = td;
            } else {
                best_t = tb;
            }
        }
        st->plan_override[c] = -1;
        st->use_graphs = saved_graphs;
        st->per_cu[c] = saved_pc;
        st->persistent[c] = saved_ps;
        st->shape[c] = saved_shape;
        if (rc == OSTFT_OK && e == hipSuccess) {
            // captured launches take it at their next replay
            st->shape[c] = best.first >> 2;
            st->persistent[c] = best.second > 0;
            st->per_cu[c] = best.second;
            ctx_set_radix_plan(ctx, log_m, best.first & 3);
            if (best_blocks_per_cu)
                *best_blocks_per_cu = best.second;
            if (best_shape)
                *best_shape = best.first >> 2;
            if (best_plan)
                *best_plan = best.first & 3;
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
