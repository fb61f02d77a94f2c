// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own OCCA-translated CPU kernels (compiled from
// StandAlone/cross_gpgpu/{Serial,OpenMP}/kernel/compiled.hpp where they lie,
// by oracle/Makefile, into oracle/_ref/) through the Serial runner's pipeline:
// SerialImpl.cpp:30-75 (geometry, buffers), RunnerCommon.cpp:184-316
// (Default_Pipeline option dispatch and stage order) and
// Serial/src/functionImpl.cpp:115-304 (which kernel each runnerFunction calls,
// EXPC's by-value result pointers, ToPower's (Real, Real) arguments).
//
// The reference runner executable itself cannot be built here (it needs
// IXWebSocket and Cap'n Proto), so this harness replaces only the runner
// glue; every arithmetic operation comes from the reference kernels.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

using CUI = const unsigned int;

namespace {
struct complex_t {
    float real, imag;
};
}  // namespace

// symbols defined by compiled.hpp (extern "C" there, references by value)
extern "C" {
void Overlap_Common(float*, CUI&, CUI&, CUI&, CUI&, float*);
void Window_Hanning(float*, CUI&, CUI&);
void Window_Hamming(float*, CUI&, CUI&);
void Window_Blackman(float*, CUI&, CUI&);
void Window_Nuttall(float*, CUI&, CUI&);
void Window_Blackman_Nuttall(float*, CUI&, CUI&);
void Window_Blackman_harris(float*, CUI&, CUI&);
void Window_FlatTop(float*, CUI&, CUI&);
void Window_Gaussian(float*, CUI&, CUI&, const float&);
void DCRemove_Common(float*, CUI&, CUI&);
void Stockhoptimized6(float*, float*, CUI&);
void Stockhoptimized7(float*, float*, CUI&);
void Stockhoptimized8(float*, float*, CUI&);
void Stockhoptimized9(float*, float*, CUI&);
void Stockhoptimized10(float*, float*, CUI&);
void Stockhoptimized11(float*, float*, CUI&);
void StockHamCommon(float*, float*, float*, float*, CUI&, CUI&, CUI&, CUI&);
void toPower(float*, float*, float*, CUI&);
void toHalfComplexFormat(complex_t*, float*, float*, CUI&, const int&);
}

namespace {

int to_quot(float fullSize, float overlapRatio, float windowSize)
{
    // RunnerInterface.hpp:26-35
    if (overlapRatio == 0.0f)
        return fullSize / windowSize + 1;
    return ((fullSize) / (windowSize * (1.0f - overlapRatio))) + 1;
}

bool has(const std::string& s, const char* key) { return s.find(key) != std::string::npos; }

// Returns 0 "OK", 1 option error, 2 "Operation not supported".
int pipeline(std::vector<float>& in, std::vector<float>& real, std::vector<float>& imag,
             std::vector<float>& subreal, std::vector<float>& subimag, std::vector<float>& out,
             unsigned fullSize, unsigned windowSize, unsigned ofull, unsigned ohalf, unsigned omove,
             const std::string& options, int exp)
{
    CUI uexp = (unsigned)exp;
    Overlap_Common(in.data(), ofull, fullSize, uexp, omove, real.data());
    if (has(options, "--hanning_window"))
        Window_Hanning(real.data(), ofull, windowSize);
    else if (has(options, "--hamming_window"))
        Window_Hamming(real.data(), ofull, windowSize);
    else if (has(options, "--blackman_window"))
        Window_Blackman(real.data(), ofull, windowSize);
    else if (has(options, "--nuttall_window"))
        Window_Nuttall(real.data(), ofull, windowSize);
    else if (has(options, "--blackman_nuttall_window"))
        Window_Blackman_Nuttall(real.data(), ofull, windowSize);
    else if (has(options, "--blackman_harris_window"))
        Window_Blackman_harris(real.data(), ofull, windowSize);
    else if (has(options, "--flattop_window"))
        Window_FlatTop(real.data(), ofull, windowSize);
    else if (has(options, "--gaussian_window=")) {
        if (has(options, "<<sigma")) {
            auto p1 = options.find("--gaussian_window=") + 19;
            auto p2 = options.find("<<sigma");
            float sigma = -1.0f;
            std::string text;
            try {
                text = options.substr(p1, p2 - p1);
                sigma = std::stof(text);
            } catch (...) {
                return 1;
            }
            if (sigma > 0)
                Window_Gaussian(real.data(), ofull, windowSize, sigma);
        }
    }
    if (has(options, "--remove_dc"))
        DCRemove_Common(real.data(), ofull, windowSize);
    if (exp < 6)
        return 2;
    float* realResult = real.data();
    float* imagResult = imag.data();
    switch (exp) {
    case 6: Stockhoptimized6(real.data(), imag.data(), ohalf); break;
    case 7: Stockhoptimized7(real.data(), imag.data(), ohalf); break;
    case 8: Stockhoptimized8(real.data(), imag.data(), ohalf); break;
    case 9: Stockhoptimized9(real.data(), imag.data(), ohalf); break;
    case 10: Stockhoptimized10(real.data(), imag.data(), ohalf); break;
    case 11: Stockhoptimized11(real.data(), imag.data(), ohalf); break;
    default: {
        // EXPC: the result pointers it reassigns are its own copies, so the
        // caller keeps reading `real` / `imag`.
        subreal.resize(ofull);
        subimag.resize(ofull);
        CUI hw = windowSize >> 1;
        for (unsigned stage = 0; stage < (unsigned)exp; ++stage) {
            CUI st = stage;
            if (stage % 2 == 0)
                StockHamCommon(real.data(), imag.data(), subreal.data(), subimag.data(), hw, st, ohalf,
                               uexp);
            else
                StockHamCommon(subreal.data(), subimag.data(), real.data(), imag.data(), hw, st, ohalf,
                               uexp);
        }
        break;
    }
    }
    if (has(options, "--half_complex_return")) {
        const int e = exp;
        toHalfComplexFormat(reinterpret_cast<complex_t*>(out.data()), realResult, imagResult, ohalf, e);
    } else {
        toPower(out.data(), realResult, realResult, ofull);
    }
    return 0;
}

}  // namespace

extern "C" int ref_activate(const float* in, uint32_t fullSize, int exp, float overlap,
                            const char* options, float* out)
{
    if (exp < 1 || exp > 30)
        return 2;
    std::vector<float> inData(in, in + fullSize);
    const unsigned int FullSize = inData.size();
    const int windowSize = 1 << exp;
    const int qtConst = to_quot(FullSize, overlap, windowSize);
    if ((int64_t)qtConst * (int64_t)windowSize >= ((int64_t)1 << 32))
        return 3;
    const unsigned int OFullSize = (unsigned)((int64_t)qtConst * windowSize);
    const unsigned int OHalfSize = OFullSize / 2;
    const unsigned int OMove = windowSize * (1.0f - overlap);
    if (FullSize == 0)
        inData.push_back(0.0f);  // the reference would read in[0] of an empty vector
    std::vector<float> Real(OFullSize);
    std::vector<float> subReal;
    std::vector<float> Imag(OFullSize);
    std::vector<float> subImag;
    std::vector<float> outMem(OFullSize);
    int rc = pipeline(inData, Real, Imag, subReal, subImag, outMem, FullSize, windowSize, OFullSize,
                      OHalfSize, OMove, std::string(options ? options : ""), exp);
    std::memcpy(out, outMem.data(), sizeof(float) * OFullSize);
    return rc;
}

// the geometry, for tests that want the reference's own numbers
extern "C" void ref_geometry(uint32_t fullSize, int exp, float overlap, int32_t* qt, uint32_t* ofull,
                             uint32_t* omove)
{
    const int windowSize = 1 << exp;
    *qt = to_quot((float)fullSize, overlap, (float)windowSize);
    *ofull = (unsigned)((int64_t)*qt * windowSize);
    *omove = windowSize * (1.0f - overlap);
}
