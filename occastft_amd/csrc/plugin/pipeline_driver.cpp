// pipeline_driver.cpp -- `runnerfn_pipeline IN.f32 EXP OVERLAP OPTIONS OUT.f32`
//
// Drives the record-and-fuse runner functions the way a runner built from the
// reference does: the geometry of Runner::ActivateSTFT
// (StandAlone/cross_gpgpu/Serial/src/SerialImpl.cpp:36-43) and the stage
// sequence of runnerFunction::Default_Pipeline (StandAlone/cross_gpgpu/
// GLOBAL/RunnerCommon.cpp:184-316: Overlap, the first matching window or the
// Gaussian window with its sigma parse, RemoveDC, the FFT stage, HalfComplex
// or ToPower), with userStruct = nullptr as the Serial runner passes.  Prints
// the pipeline's status string; writes OFullSize floats (zeros where the
// pipeline stopped early, as the Serial runner returns them).  Used by the
// tests to check the plugin layer against the engine's own entry point.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/occastft_runner_functions.hpp"

namespace {

std::string pipeline(std::vector<float>& in, std::vector<float>& out, unsigned FullSize, unsigned windowSize,
                     unsigned qtConst, unsigned OFullSize, unsigned OHalfSize, unsigned OMove,
                     const std::string& options, int exp)
{
    using namespace runnerFunction;
    std::vector<float> real(OFullSize), imag(OFullSize), subreal, subimag;
    auto has = [&](const char* o) { return options.find(o) != std::string::npos; };
    Overlap(nullptr, &in, OFullSize, FullSize, (unsigned)exp, OMove, &real);
    struct W {
        const char* opt;
        bool (*fn)(void*, void*, const unsigned, const unsigned);
    };
    static const W windows[] = {{"--hanning_window", Hanning},          {"--hamming_window", Hamming},
                                {"--blackman_window", Blackman},        {"--nuttall_window", Nuttall},
                                {"--blackman_nuttall_window", Blackman_Nuttall},
                                {"--blackman_harris_window", Blackman_Harris}, {"--flattop_window", FlatTop}};
    bool matched = false;
    for (const W& w : windows) {
        if (has(w.opt)) {
            if (!w.fn(nullptr, &real, OFullSize, windowSize))
                return w.opt;
            matched = true;
            break;
        }
    }
    if (!matched && has("--gaussian_window=") && has("<<sigma")) {
        const size_t p1 = options.find("--gaussian_window=") + 19;  // one past the first sigma character
        const size_t p2 = options.find("<<sigma");
        float sigma = -1.0f;
        try {
            sigma = std::stof(options.substr(p1, p2 - p1));
        } catch (const std::exception&) {
            return "--gaussian_window";
        }
        if (sigma > 0 && !Gaussian(nullptr, &real, OFullSize, windowSize, sigma))
            return "--gaussian_window";
    }
    if (has("--remove_dc") && !RemoveDC(nullptr, &real, qtConst, OFullSize, windowSize))
        return "--remove_dc";
    if (exp < 6)
        return "Operation not supported";
    bool ok = true;
    switch (exp) {
    case 6: ok = EXP6(nullptr, &real, &imag, OHalfSize); break;
    case 7: ok = EXP7(nullptr, &real, &imag, OHalfSize); break;
    case 8: ok = EXP8(nullptr, &real, &imag, OHalfSize); break;
    case 9: ok = EXP9(nullptr, &real, &imag, OHalfSize); break;
    case 10: ok = EXP10(nullptr, &real, &imag, OHalfSize); break;
    case 11: ok = EXP11(nullptr, &real, &imag, OHalfSize); break;
    default:
        ok = EXPC(nullptr, &real, &imag, &subreal, &subimag, &out, windowSize >> 1, (unsigned)exp, OFullSize, &real,
                  &imag);
        break;
    }
    if (!ok)
        return "Err On Stockham";
    if (has("--half_complex_return")) {
        if (!HalfComplex(nullptr, &out, &real, &imag, OHalfSize, (unsigned)exp))
            return "--half_complex_return";
    } else {
        ToPower(nullptr, &out, &real, &imag, OFullSize);
    }
    return "OK";
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s IN.f32 EXP OVERLAP OPTIONS OUT.f32\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f)
        return 2;
    std::vector<float> in;
    float buf[4096];
    size_t n;
    while ((n = std::fread(buf, sizeof(float), 4096, f)) > 0)
        in.insert(in.end(), buf, buf + n);
    std::fclose(f);
    const int exp = std::atoi(argv[2]);
    const float overlap = std::strtof(argv[3], nullptr);
    const std::string options = argv[4];
    // the reference's geometry (toQuot in float32), through the engine's own
    // restatement of it; windows below 2^6 only need the sizes
    ostft_params p = {};
    p.exp = exp < 6 ? 6 : exp;
    p.overlap = overlap;
    ostft_geometry g;
    if (ostft_get_geometry(&p, in.size(), &g) != OSTFT_OK)
        return 2;
    const unsigned windowSize = 1u << exp;
    const unsigned qt = exp < 6 ? (unsigned)(overlap == 0.0f ? (float)in.size() / windowSize + 1.0f
                                                             : (float)in.size() / (windowSize * (1.0f - overlap)) + 1.0f)
                                : (unsigned)g.frames;
    const unsigned OFullSize = qt * windowSize, OHalfSize = OFullSize / 2;
    const unsigned OMove = (unsigned)(windowSize * (1.0f - overlap));
    std::vector<float> out(OFullSize, 0.0f);
    const std::string status =
        pipeline(in, out, (unsigned)in.size(), windowSize, qt, OFullSize, OHalfSize, OMove, options, exp);
    std::printf("%s\n", status.c_str());
    const std::string err = ostft_plugin_last_error(nullptr);
    if (!err.empty())
        std::printf("plugin: %s\n", err.c_str());
    FILE* o = std::fopen(argv[5], "wb");
    if (!o || std::fwrite(out.data(), sizeof(float), out.size(), o) != out.size())
        return 2;
    std::fclose(o);
    return 0;
}
