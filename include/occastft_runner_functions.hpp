/*
 * occastft_runner_functions.hpp -- the reference runner's per-stage plugin
 * functions (namespace runnerFunction, StandAlone/cross_gpgpu/GLOBAL/
 * RunnerInterface.hpp:36-83) on the MI355X engine, as "record and fuse":
 * Overlap, the window functions, RemoveDC and the FFT stages only record what
 * Default_Pipeline (RunnerCommon.cpp:184-316) asks for; the output stage
 * (ToPower or HalfComplex) runs the whole recorded pipeline as one fused
 * engine call and writes the result.  Results are the engine's
 * (ostft_activate with the same options), bit for bit.
 *
 * Link libocca_stft_runnerfn.so (and libocca_stft_hip.so) into a runner
 * built from the reference's RunnerCommon.cpp; the signatures below are the
 * reference's, so its Default_Pipeline links against them unchanged.
 *
 * Buffer conventions (those of the Serial runner, SerialImpl.cpp:30-75):
 * `origin` is the input std::vector<float>*, `out` the output
 * std::vector<float>* (grown to OFullSize if smaller); Real, Imag and the sub
 * buffers are not touched.  `userStruct` is an ostft_plugin_env* made by
 * ostft_plugin_env_create, or nullptr (as the Serial runner passes), which
 * selects a per-thread environment on device 0.  A stage that fails returns
 * false, as the reference's do; the detail is in ostft_plugin_last_error.
 */
#ifndef OCCASTFT_RUNNER_FUNCTIONS_HPP
#define OCCASTFT_RUNNER_FUNCTIONS_HPP

#include <string>

#include "occastft_hip.h"

struct ostft_plugin_env;

/* Environment for one runner (one engine context on `device`); nullptr if the
 * device cannot be opened. */
ostft_plugin_env* ostft_plugin_env_create(int device);
void ostft_plugin_env_destroy(ostft_plugin_env* env);
/* Last failure recorded for `env` (nullptr: the calling thread's default). */
std::string ostft_plugin_last_error(const ostft_plugin_env* env);

namespace runnerFunction {

bool Overlap(void* userStruct, void* origin, const unsigned int OFullSize, const unsigned int FullSize,
             const unsigned int windowSizeEXP, const unsigned int OMove, void* Realout);
bool Hanning(void* userStruct, void* data, const unsigned int OFullSize, const unsigned int windowSize);
bool Hamming(void* userStruct, void* data, const unsigned int OFullSize, const unsigned int windowSize);
bool Blackman(void* userStruct, void* data, const unsigned int OFullSize, const unsigned int windowSize);
bool Nuttall(void* userStruct, void* data, const unsigned int OFullSize, const unsigned int windowSize);
bool Blackman_Nuttall(void* userStruct, void* data, const unsigned int OFullSize, const unsigned int windowSize);
bool Blackman_Harris(void* userStruct, void* data, const unsigned int OFullSize, const unsigned int windowSize);
bool FlatTop(void* userStruct, void* data, const unsigned int OFullSize, const unsigned int windowSize);
bool Gaussian(void* userStruct, void* data, const unsigned int OFullSize, const unsigned int windowSize,
              const float sigma);
bool RemoveDC(void* userStruct, void* data, const unsigned int qtConst, const unsigned int OFullSize,
              const unsigned int windowSize);
bool EXP6(void* userStruct, void* Real, void* Imag, const unsigned int OHalfSize);
bool EXP7(void* userStruct, void* Real, void* Imag, const unsigned int OHalfSize);
bool EXP8(void* userStruct, void* Real, void* Imag, const unsigned int OHalfSize);
bool EXP9(void* userStruct, void* Real, void* Imag, const unsigned int OHalfSize);
bool EXP10(void* userStruct, void* Real, void* Imag, const unsigned int OHalfSize);
bool EXP11(void* userStruct, void* Real, void* Imag, const unsigned int OHalfSize);
bool EXPC(void* userStruct, void* real, void* imag, void* subreal, void* subimag, void* out,
          const unsigned int HWindowSize, const unsigned int windowSizeEXP, const unsigned int OFullSize,
          void* realResult, void* imagResult);
bool HalfComplex(void* userStruct, void* out, void* realResult, void* imagResult, const unsigned int OHalfSize,
                 const unsigned int windowSizeEXP);
bool ToPower(void* userStruct, void* out, void* realResult, void* imagResult, const unsigned int OFullSize);

}  // namespace runnerFunction

#endif /* OCCASTFT_RUNNER_FUNCTIONS_HPP */
