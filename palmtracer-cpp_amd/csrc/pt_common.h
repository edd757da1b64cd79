// Shared definitions for the MI355X (gfx950) implementation of PALMTracer's detection + fitting path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace pt {

// Row layout of one localization (lib/Constants.hpp:31-51).
enum Col : int {
    kId = 0, kPlane, kIndex, kChannel, kX, kY, kZ, kIntensity, kSigmaX, kSigmaY, kTheta,
    kQualityXY, kQualityZ, kAmplitude, kOffset, kSegIntensity, kSurface, kCircularity, kCols
};
// Caller rows for the fit-only entry (lib/Constants.hpp:56-64).
enum InCol : int { kInId = 0, kInPlane, kInX, kInY, kInSegIntensity, kInSurface, kInCols };

constexpr double kEpsilon = 1e-7;       // lib/Constants.hpp:53 (watershed aging step)
constexpr double kVolumeMin = 4.0;      // lib/Constants.hpp:23
constexpr int kThresholdIterations = 5; // lib/Constants.hpp:24
constexpr int kMaxIterDefault = 75;     // lib/fit/Fitting.hpp:80

// The thresholded potential W1 is stored as the exact integer P = W1 * 2^16 (u32, 0 = below
// threshold). A trous of u16 data with the B3 taps {1,4,6,4,1}/16 produces dyadic rationals of at
// most 32 fractional bits below 2^16, so P (and the FP64 value P * 2^-16) is exact.
constexpr double kPScale = 1.0 / 65536.0;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char* what, const char* file, int line)
{
    if (e != hipSuccess)
        throw Error(std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")");
}
#define PT_CHECK(x) ::pt::check((x), #x, __FILE__, __LINE__)

inline unsigned ceil_div(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

} // namespace pt
