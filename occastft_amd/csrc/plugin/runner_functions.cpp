// runner_functions.cpp -- record-and-fuse implementation of the reference's
// per-stage runner functions (declarations: occastft_runner_functions.hpp,
// signatures of StandAlone/cross_gpgpu/GLOBAL/RunnerInterface.hpp:38-83).
//
// The reference runs one kernel (or one pass) per stage over OFullSize-float
// buffers.  Here every stage before the output only records its parameters in
// the runner's environment; ToPower / HalfComplex run the recorded pipeline
// as one engine call (the fused kernel or the two-pass path) with the
// geometry Default_Pipeline computed (OMove, OFullSize).
#include "../../../include/occastft_runner_functions.hpp"

#include <memory>
#include <string>
#include <vector>

struct ostft_plugin_env {
    ostft_ctx* ctx = nullptr;
    // recorded by the stages of the current pipeline
    const std::vector<float>* in = nullptr;
    unsigned full = 0, exp = 0, omove = 0, ofull = 0;
    int window = OSTFT_WIN_NONE;
    float sigma = 0.0f;
    bool dc = false;
    bool transformed = false;
    std::string error;
};

ostft_plugin_env* ostft_plugin_env_create(int device)
{
    ostft_ctx* ctx = nullptr;
    if (ostft_create(device, &ctx) != OSTFT_OK)
        return nullptr;
    ostft_plugin_env* e = new ostft_plugin_env();
    e->ctx = ctx;
    return e;
}

void ostft_plugin_env_destroy(ostft_plugin_env* env)
{
    if (!env)
        return;
    if (env->ctx)
        ostft_destroy(env->ctx);
    delete env;
}

namespace {

struct EnvDeleter {
    void operator()(ostft_plugin_env* e) const { ostft_plugin_env_destroy(e); }
};

// the environment of the Serial runner's convention (userStruct == nullptr)
thread_local std::unique_ptr<ostft_plugin_env, EnvDeleter> tls_env;
thread_local std::string tls_error;

ostft_plugin_env* env_of(void* userStruct)
{
    if (userStruct)
        return static_cast<ostft_plugin_env*>(userStruct);
    if (!tls_env) {
        tls_env.reset(ostft_plugin_env_create(0));
        if (!tls_env)
            tls_error = "no HIP device";
    }
    return tls_env.get();
}

bool record_window(void* userStruct, int kind, float sigma)
{
    ostft_plugin_env* e = env_of(userStruct);
    if (!e)
        return false;
    e->window = kind;
    e->sigma = sigma;
    return true;
}

bool record_transform(void* userStruct, unsigned exp)
{
    ostft_plugin_env* e = env_of(userStruct);
    if (!e)
        return false;
    if (exp != 0 && exp != e->exp) {
        e->error = "FFT stage for 2^" + std::to_string(exp) + " after Overlap for 2^" + std::to_string(e->exp);
        return false;
    }
    e->transformed = true;
    return true;
}

// the output stage: run everything recorded since Overlap
bool run(void* userStruct, void* out, int layout)
{
    ostft_plugin_env* e = env_of(userStruct);
    if (!e)
        return false;
    std::vector<float>* o = static_cast<std::vector<float>*>(out);
    if (!e->in || !e->transformed || !o) {
        e->error = "ToPower / HalfComplex without Overlap and an FFT stage before it";
        return false;
    }
    ostft_params p = {};
    p.exp = (int32_t)e->exp;
    p.overlap = 0.0f;  // the geometry is explicit
    p.window = e->window;
    p.sigma = e->sigma;
    p.output = layout;
    p.flags = e->dc ? OSTFT_FLAG_REMOVE_DC : 0u;
    const uint64_t frames = (uint64_t)e->ofull >> e->exp;
    if (o->size() < e->ofull)
        o->resize(e->ofull);
    const uint64_t L = e->full < e->in->size() ? e->full : e->in->size();
    uint64_t n = 0;
    const int rc = ostft_activate_geometry(e->ctx, &p, L ? e->in->data() : nullptr, L, e->omove, frames, o->data(),
                                           o->size(), &n);
    if (rc != OSTFT_OK) {
        e->error = std::string(ostft_strerror(rc)) + ": " + ostft_last_error(e->ctx);
        return false;
    }
    return true;
}

}  // namespace

std::string ostft_plugin_last_error(const ostft_plugin_env* env)
{
    if (env)
        return env->error;
    return tls_env ? tls_env->error : tls_error;
}

namespace runnerFunction {

bool Overlap(void* userStruct, void* origin, const unsigned int OFullSize, const unsigned int FullSize,
             const unsigned int windowSizeEXP, const unsigned int OMove, void* /*Realout*/)
{
    ostft_plugin_env* e = env_of(userStruct);
    if (!e)
        return false;
    e->in = static_cast<const std::vector<float>*>(origin);
    e->full = FullSize;
    e->exp = windowSizeEXP;
    e->omove = OMove;
    e->ofull = OFullSize;
    e->window = OSTFT_WIN_NONE;
    e->sigma = 0.0f;
    e->dc = false;
    e->transformed = false;
    e->error.clear();
    return e->in != nullptr;
}

bool Hanning(void* us, void*, const unsigned int, const unsigned int) { return record_window(us, OSTFT_WIN_HANNING, 0); }
bool Hamming(void* us, void*, const unsigned int, const unsigned int) { return record_window(us, OSTFT_WIN_HAMMING, 0); }
bool Blackman(void* us, void*, const unsigned int, const unsigned int)
{
    return record_window(us, OSTFT_WIN_BLACKMAN, 0);
}
bool Nuttall(void* us, void*, const unsigned int, const unsigned int) { return record_window(us, OSTFT_WIN_NUTTALL, 0); }
bool Blackman_Nuttall(void* us, void*, const unsigned int, const unsigned int)
{
    return record_window(us, OSTFT_WIN_BLACKMAN_NUTTALL, 0);
}
bool Blackman_Harris(void* us, void*, const unsigned int, const unsigned int)
{
    return record_window(us, OSTFT_WIN_BLACKMAN_HARRIS, 0);
}
bool FlatTop(void* us, void*, const unsigned int, const unsigned int) { return record_window(us, OSTFT_WIN_FLATTOP, 0); }
bool Gaussian(void* us, void*, const unsigned int, const unsigned int, const float sigma)
{
    return record_window(us, OSTFT_WIN_GAUSSIAN, sigma);
}

bool RemoveDC(void* userStruct, void*, const unsigned int, const unsigned int, const unsigned int)
{
    ostft_plugin_env* e = env_of(userStruct);
    if (!e)
        return false;
    e->dc = true;
    return true;
}

bool EXP6(void* us, void*, void*, const unsigned int) { return record_transform(us, 6); }
bool EXP7(void* us, void*, void*, const unsigned int) { return record_transform(us, 7); }
bool EXP8(void* us, void*, void*, const unsigned int) { return record_transform(us, 8); }
bool EXP9(void* us, void*, void*, const unsigned int) { return record_transform(us, 9); }
bool EXP10(void* us, void*, void*, const unsigned int) { return record_transform(us, 10); }
bool EXP11(void* us, void*, void*, const unsigned int) { return record_transform(us, 11); }

bool EXPC(void* us, void*, void*, void*, void*, void*, const unsigned int, const unsigned int windowSizeEXP,
          const unsigned int, void*, void*)
{
    return record_transform(us, windowSizeEXP);
}

bool HalfComplex(void* us, void* out, void*, void*, const unsigned int, const unsigned int)
{
    return run(us, out, OSTFT_OUT_HALF_COMPLEX);
}

bool ToPower(void* us, void* out, void*, void*, const unsigned int) { return run(us, out, OSTFT_OUT_COMPAT_POWER); }

}  // namespace runnerFunction
