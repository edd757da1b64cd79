// Host-side exports of the reference library that stay on the CPU: drift alignment, tracking,
// blinking reconnection, track analysis and the astigmatism calibration
// (CPU/main.cpp:47-48, :159, :185-186, :231-233, :293). They are not part of the
// detection + fitting path; this library exports them with the reference's signatures so that it
// can replace the reference library as a whole, and forwards each call to the reference CPU library
// named by the environment variable PALMTRACER_HOST_LIBRARY (loaded once, on first use). Without
// it, a call prints an error on stderr and returns 0 / false, leaving the outputs untouched.
#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "palmtracer.h"

namespace {

void* host_library()
{
    static std::once_flag once;
    static void* handle = nullptr;
    std::call_once(once, [] {
        const char* path = std::getenv("PALMTRACER_HOST_LIBRARY");
        if (!path || !*path) return;
        handle = dlopen(path, RTLD_NOW | RTLD_LOCAL);
        if (!handle) std::fprintf(stderr, "[ERROR] PALMTRACER_HOST_LIBRARY: %s\n", dlerror());
    });
    return handle;
}

// The reference library's symbol `name`, or nullptr (with a message) when it is not available or
// resolves to this library's own forwarder.
template <typename Fn>
Fn host_symbol(const char* name, Fn self)
{
    void* h = host_library();
    void* f = h ? dlsym(h, name) : nullptr;
    if (f == reinterpret_cast<void*>(self)) f = nullptr;
    if (!f)
        std::fprintf(stderr,
                     "[ERROR] %s runs on the host in the reference CPU library; set PALMTRACER_HOST_LIBRARY to its path.\n",
                     name);
    return reinterpret_cast<Fn>(f);
}

} // namespace

PT_API void PT_CALL Alignment(const uint16_t* input, uint16_t* output, uint64_t height, uint64_t width, uint64_t planes,
                              const double* factors, uint64_t upsampling)
{
    if (auto f = host_symbol("Alignment", &Alignment)) f(input, output, height, width, planes, factors, upsampling);
}

PT_API void PT_CALL Tracking(const double* input, uint64_t* output, uint64_t nPoints, double maxDistance)
{
    if (auto f = host_symbol("Tracking", &Tracking)) f(input, output, nPoints, maxDistance);
}

PT_API uint64_t PT_CALL BlinkingReconnection(const double* input, double* output, uint64_t nRow, double pixelSize,
                                             uint64_t mode, uint64_t maxDuration, double maxSpeed)
{
    auto f = host_symbol("BlinkingReconnection", &BlinkingReconnection);
    return f ? f(input, output, nRow, pixelSize, mode, maxDuration, maxSpeed) : 0;
}

PT_API bool PT_CALL TrackAnalysis(const double* input, double* outputMSD, double* outputInd, double* outputFit, uint64_t nRow,
                                  bool isMSD, bool isInd, bool is3D, double pixelSize, double exposureTime, uint64_t fitMode,
                                  const double* fitParams)
{
    auto f = host_symbol("TrackAnalysis", &TrackAnalysis);
    return f ? f(input, outputMSD, outputInd, outputFit, nRow, isMSD, isInd, is3D, pixelSize, exposureTime, fitMode, fitParams)
             : false;
}

PT_API void PT_CALL Astigmatism3DCalibration(const double* input, double* output, uint64_t size, double pixelSize, bool center)
{
    if (auto f = host_symbol("Astigmatism3DCalibration", &Astigmatism3DCalibration)) f(input, output, size, pixelSize, center);
}
