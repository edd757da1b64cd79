// C ABI of the library: the reference's exported entry points (CPU/main.cpp, CPU/PALMTracer.def) with
// their exact signatures, plus device-pointer and diagnostic entry points declared in
// include/palmtracer.h. No exception crosses this boundary; failures print one line on stderr and
// return 0 (or NaN for AutoThreshold), as the reference's callers expect from a failed call.
#include "engine.h"
#include "../../include/palmtracer.h"

#include <cmath>
#include <cstdio>
#include <dlfcn.h>

using pt::Engine;

namespace {

template <typename F>
auto guarded(const char* name, F&& f, decltype(f()) onError) -> decltype(f())
{
    try {
        return f();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] %s: %s\n", name, e.what());
    } catch (...) {
        std::fprintf(stderr, "[palmtracer] %s: unknown error\n", name);
    }
    return onError;
}

struct HostToDevice {
    pt::DevBuf& buf;
    const uint16_t* dptr = nullptr;
    HostToDevice(pt::DevBuf& b, const uint16_t* host, size_t n, hipStream_t s) : buf(b)
    {
        buf.ensure(n * 2);
        PT_CHECK(hipMemcpyAsync(buf.p, host, n * 2, hipMemcpyHostToDevice, s));
        dptr = buf.as<uint16_t>();
    }
};

} // namespace

extern "C" {

PT_API uint64_t PT_CALL Localization(const uint16_t* input, double* output, uint64_t outputSize, uint64_t height,
                                     uint64_t width, uint64_t planes, double thresholdVal, double watershedRatio,
                                     uint64_t fit, const double* fit_params)
{
    return guarded("Localization", [&]() -> uint64_t {
        pt::Lease lease;
        Engine& e = *lease;
        if (height > 0x7FFFFFFFull || width > 0x7FFFFFFFull) throw pt::Error("frame too large");
        return e.localize_host(input, output, outputSize, (int)height, (int)width, planes, thresholdVal, watershedRatio, fit,
                               fit_params);
    }, 0);
}

PT_API uint64_t PT_CALL Fitting(const uint16_t* input, const double* positions, uint64_t nPoints, double* output,
                                uint64_t outputSize, uint64_t height, uint64_t width, uint64_t planes, uint64_t fit,
                                const double* fit_params)
{
    (void)outputSize;  // the reference ignores it too (lib/Palm.cpp:74)
    return guarded("Fitting", [&]() -> uint64_t {
        if (nPoints == 0) return 0;
        pt::Lease lease;
        Engine& e = *lease;
        if (height > 0x7FFFFFFFull || width > 0x7FFFFFFFull) throw pt::Error("frame too large");
        return e.fit_positions_host(input, planes, (int)height, (int)width, positions, nPoints, fit, fit_params, output);
    }, 0);
}

PT_API void PT_CALL Wavelett(const uint16_t* input, double* output, uint64_t height, uint64_t width, uint64_t planes,
                             uint64_t levelOutput)
{
    guarded("Wavelett", [&]() -> int {
        if (planes == 0 || height == 0 || width == 0) return 0;
        if (levelOutput == 0) levelOutput = 1;                   // CPU/main.cpp:262
        const int level = (int)std::min<uint64_t>(levelOutput - 1, 2);
        pt::Lease lease(false);   // no counters
        Engine& e = *lease;
        const size_t npx = (size_t)height * width;
        // Bounded device footprint: planes in chunks.
        const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(planes, (size_t(256) << 20) / (npx * 10)));
        for (uint64_t z0 = 0; z0 < planes; z0 += chunk) {
            const uint64_t nz = std::min(chunk, planes - z0);
            HostToDevice h2d(e.in, input + z0 * npx, nz * npx, e.stream);
            e.out.ensure(nz * npx * 8);
            e.wavelet(h2d.dptr, nz, (int)height, (int)width, level, e.out.as<double>());
            PT_CHECK(hipMemcpyAsync(output + z0 * npx, e.out.p, nz * npx * 8, hipMemcpyDeviceToHost, e.stream));
            PT_CHECK(hipStreamSynchronize(e.stream));
        }
        return 0;
    }, 0);
}

PT_API double PT_CALL AutoThreshold(const uint16_t* input, uint64_t height, uint64_t width, const double* fit_params)
{
    return guarded("AutoThreshold", [&]() -> double {
        pt::Lease lease;
        Engine& e = *lease;
        if (height == 0 || width == 0) return std::nan("");
        HostToDevice h2d(e.in, input, (size_t)height * width, e.stream);
        return e.autothreshold(h2d.dptr, (int)height, (int)width, fit_params);
    }, std::nan(""));
}

PT_API void PT_CALL Astigmatism3DEstimation(const double* input, double* output, uint64_t size, double pixelSize,
                                            const double* model, double zMax)
{
    guarded("Astigmatism3DEstimation", [&]() -> int {
        if (size == 0) return 0;
        pt::Lease lease(false);   // no counters
        Engine& e = *lease;
        e.in.ensure(size * 16);
        e.out.ensure(size * 16);
        PT_CHECK(hipMemcpyAsync(e.in.p, input, size * 16, hipMemcpyHostToDevice, e.stream));
        pt::launch_astig_estimate(e.in.as<double>(), e.out.as<double>(), size, pixelSize, model, zMax, e.stream);
        PT_CHECK(hipMemcpyAsync(output, e.out.p, size * 16, hipMemcpyDeviceToHost, e.stream));
        PT_CHECK(hipStreamSynchronize(e.stream));
        return 0;
    }, 0);
}

PT_API int PT_CALL pt_astig_estimate_device(const double* d_input, double* d_output, uint64_t size, double pixelSize,
                                            const double* model, double zMax)
{
    return guarded("pt_astig_estimate_device", [&]() -> int {
        pt::Lease lease(false);   // no counters
        Engine& e = *lease;
        pt::launch_astig_estimate(d_input, d_output, size, pixelSize, model, zMax, e.stream);
        PT_CHECK(hipStreamSynchronize(e.stream));
        return 1;
    }, 0);
}

PT_API double PT_CALL pt_autothreshold_device(const uint16_t* d_input, uint64_t height, uint64_t width,
                                              const double* fit_params)
{
    return guarded("pt_autothreshold_device", [&]() -> double {
        pt::Lease lease;
        Engine& e = *lease;
        return e.autothreshold(d_input, (int)height, (int)width, fit_params);
    }, std::nan(""));
}

// ----- device-pointer entry points (no counterpart in the reference) -----

PT_API uint64_t PT_CALL pt_localization_device(const uint16_t* d_input, double* d_output, uint64_t outputCapDoubles,
                                               uint64_t height, uint64_t width, uint64_t planes, double thresholdVal,
                                               double watershedRatio, uint64_t fit, const double* fit_params,
                                               uint64_t maxPerPlane)
{
    return guarded("pt_localization_device", [&]() -> uint64_t {
        pt::Lease lease;
        Engine& e = *lease;
        return e.localize(d_input, planes, (int)height, (int)width, thresholdVal, watershedRatio, fit, fit_params,
                          maxPerPlane, d_output, outputCapDoubles / pt::kCols);
    }, 0);
}

PT_API int PT_CALL pt_wavelet_device(const uint16_t* d_input, double* d_output, uint64_t height, uint64_t width,
                                     uint64_t planes, uint64_t levelOutput)
{
    return guarded("pt_wavelet_device", [&]() -> int {
        pt::Lease lease(false);   // no counters
        Engine& e = *lease;
        if (levelOutput == 0) levelOutput = 1;
        e.wavelet(d_input, planes, (int)height, (int)width, (int)std::min<uint64_t>(levelOutput - 1, 2), d_output);
        return 1;
    }, 0);
}

PT_API int PT_CALL pt_debug_labels(const uint16_t* image, uint64_t height, uint64_t width, double thresholdVal,
                                   double watershedRatio, uint32_t* labels)
{
    return guarded("pt_debug_labels", [&]() -> int {
        pt::Lease lease;
        Engine& e = *lease;
        const size_t npx = (size_t)height * width;
        HostToDevice h2d(e.in, image, npx, e.stream);
        e.out.ensure(npx * 4);
        e.debug_labels(h2d.dptr, (int)height, (int)width, thresholdVal, watershedRatio, e.out.as<uint32_t>());
        PT_CHECK(hipMemcpyAsync(labels, e.out.p, npx * 4, hipMemcpyDeviceToHost, e.stream));
        PT_CHECK(hipStreamSynchronize(e.stream));
        return 1;
    }, 0);
}

PT_API int PT_CALL pt_set_option(const char* name, double value)
{
    return guarded("pt_set_option", [&]() -> int {
        static std::mutex setMu;   // one read-modify-write of the shared options at a time
        std::lock_guard<std::mutex> lk(setMu);
        pt::Options o = Engine::shared_options();
        const std::string n(name ? name : "");
        if (n == "fit_fp64") o.fitFp64 = value != 0.0;
        else if (n == "max_iter") o.maxIter = std::max(1, (int)value);
        else if (n == "chunk_pixels") o.chunkPixels = value > 0.0 ? (size_t)value : pt::Options::kDefaultChunkPixels;
        else if (n == "profile") o.profile = value != 0.0;
        else if (n == "force_sequential_watershed") o.forceSequential = value != 0.0;
        else if (n == "force_heap_watershed") o.exactPath = value != 0.0 ? 2 : 0;
        else if (n == "force_wave_flood") o.exactPath = value != 0.0 ? 1 : 0;
        else if (n == "force_rank_flood") o.exactPath = value != 0.0 ? 3 : 0;
        else if (n == "fit_waves") o.fitWaves = (unsigned)std::max(0.0, std::min(1048576.0, value));
        else if (n == "fit_refill") o.fitRefill = (int)std::max(1.0, std::min(64.0, value));
        else if (n == "fit_handover") o.fitHandover = (int)std::max(0.0, std::min(1e6, value));
        else if (n == "batch_autotune") o.batchAutotune = value != 0.0;
        else if (n == "batch_graph") o.batchGraph = value != 0.0;
        else if (n == "fit_objective") o.fitObjective = value != 0.0 ? 1 : 0;
        else if (n == "camera_gain") o.camGain = value;
        else if (n == "camera_offset") o.camOffset = value;
        else if (n == "camera_read_var") o.camReadVar = std::max(0.0, value);
        else if (n == "stream_pixels") o.streamPixels = value >= 1.0 ? (size_t)value : (size_t(1) << 27);
        else if (n == "host_threads") o.hostThreads = (int)std::max(1.0, std::min(64.0, value));
        else if (n == "dense_flag_mask") o.denseFlagMask = (uint32_t)std::max(0.0, std::min(4294967295.0, value));
        else if (n == "defer_flagged") o.deferFlagged = (int)std::max(0.0, std::min(2.0, value));
        else if (n == "detect_rows") o.detectRows = (int)std::max(0.0, std::min(1048576.0, value));
        else if (n == "lanes") o.lanes = value >= 2.0 ? 2 : 1;
        else throw pt::Error("unknown option '" + n + "'");
        Engine::set_shared_options(o);
        return 1;
    }, 0);
}

PT_API int PT_CALL pt_last_stats(uint64_t* out, uint64_t n)
{
    const pt::RunStats last = Engine::published();
    const uint64_t v[24] = {last.frames, last.chunks, last.rois, last.flaggedFrames, last.multiComponents,
                            last.deferredFrames, last.status, last.flagOverflow, last.flagUnproven,
                            last.floodPops, last.floodMaxQueue, last.rfTooBig, last.rfTie,
                            last.rfUnsettled, last.fbOverflow, last.heapFrames, last.rfTwoOrders,
                            last.fitHanded, last.gridFrames, last.gridTies, last.gridUnsettled,
                            last.gridRejected, last.gridOther, last.maskPixels};
    for (uint64_t i = 0; i < n && i < 24; ++i) out[i] = v[i];
    return 24;
}

PT_API int PT_CALL pt_stage_times(double* out, uint64_t n)
{
    const pt::RunStats last = Engine::published();
    const double v[13] = {last.msAtrous, last.msSegment, last.msFit, last.msAssemble, (double)last.chunks,
                          (double)last.jacPasses, (double)last.trialPasses, last.msDeferred,
                          (double)last.groupJacPasses, (double)last.groupTrialPasses, last.msH2D, last.msD2H,
                          (double)last.streamPieces};
    for (uint64_t i = 0; i < n && i < 13; ++i) out[i] = v[i];
    return 13;
}

} // extern "C"
