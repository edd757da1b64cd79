// Fitting over device memory without any allocation (pt_fitting_noalloc).
//
// The work space of a fit-only call (the fit jobs, the fast fit's queue words and hand-over list,
// the per-plane index of every position and the per-plane counters) lives in one file-scope device
// array, `scratch`, of fixed size. Each GPU that runs the library holds its own copy of the array
// until the process ends; its device address is looked up once per device with
// hipGetSymbolAddress (the array's name is a host-side stand-in in host code, never a device
// address). The call then launches the same kernels as Fitting on the legacy default stream and
// calls no hipMalloc, hipFree or allocating library routine. Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU race.
//
// Reference: lib/Palm.cpp:73-119 (FitLocalizations: per-plane index in input order, rows in input
// order), CPU/main.cpp:141-148 (the Fitting export).
#include "engine.h"
#include "../../include/palmtracer.h"

#include <cstdio>
#include <mutex>

namespace pt {
namespace {

constexpr uint64_t kNoAllocMaxPoints = uint64_t(1) << 17;   // positions per call
constexpr uint64_t kNoAllocMaxPlanes = (uint64_t(1) << 16) - 1;
constexpr int kMaxDevices = 64;

// Layout of the scratch array. Word 4 of `words` is the bad-plane flag.
struct NoAllocScratch {
    Roi rois[kNoAllocMaxPoints];
    uint32_t hand[kNoAllocMaxPoints];
    uint32_t local[kNoAllocMaxPoints];
    uint32_t planeCount[kNoAllocMaxPlanes + 1];   // one counter per plane, then positions outside [1, planes]
    uint32_t words[16];
};

static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];

__global__ void noalloc_zero_kernel(uint32_t* a, size_t na, uint32_t* b, size_t nb)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < na + nb; i += (size_t)gridDim.x * blockDim.x) {
        if (i < na) a[i] = 0u;
        else b[i - na] = 0u;
    }
}

// Per-plane 1-based index of each position in input order (lib/Palm.cpp:79-84), by one workgroup of
// 1024 threads over blocks of 1024 positions: a block's (plane, position) keys are sorted in LDS
// (bitonic), a position's index is the running count of its plane plus its rank among the block's
// positions of that plane, and the last position of each plane's run adds the run to the count.
// Positions whose plane is outside [1, planes] are counted under the extra bucket `planes`.
__global__ void __launch_bounds__(1024) noalloc_local_index_kernel(const double* __restrict__ pos, uint64_t n,
                                                                   uint64_t planes, uint32_t* __restrict__ planeCount,
                                                                   uint32_t* __restrict__ local)
{
    __shared__ uint32_t key[1024];
    __shared__ uint32_t runStart[1024];
    const uint32_t t = threadIdx.x;
    constexpr uint32_t kNone = 0x3FFFFFu;   // bucket of the slots past n (sorts last)
    for (uint64_t base = 0; base < n; base += 1024) {
        const uint64_t i = base + t;
        uint32_t bucket = kNone;
        if (i < n) {
            const double pl = pos[i * kInCols + kInPlane];
            bucket = (pl >= 1.0 && pl < (double)planes + 1.0) ? (uint32_t)((uint64_t)pl - 1) : (uint32_t)planes;
        }
        key[t] = (bucket << 10) | t;
        __syncthreads();
        for (uint32_t k = 2; k <= 1024; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                const uint32_t q = t ^ j;
                if (q > t) {
                    const uint32_t a = key[t], b = key[q];
                    const bool up = (t & k) == 0;
                    if ((a > b) == up) {
                        key[t] = b;
                        key[q] = a;
                    }
                }
                __syncthreads();
            }
        }
        // Start of the run of equal buckets that sorted slot t belongs to (inclusive max scan).
        const uint32_t kt = key[t], bt = kt >> 10;
        runStart[t] = (t == 0 || (key[t - 1] >> 10) != bt) ? t : 0u;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            const uint32_t v = t >= o ? runStart[t - o] : 0u;
            __syncthreads();
            runStart[t] = max(runStart[t], v);
            __syncthreads();
        }
        const uint32_t rank = t - runStart[t];
        const uint32_t slot = kt & 1023u;
        if (bt != kNone) local[base + slot] = planeCount[bt] + rank + 1u;
        __syncthreads();
        const bool last = t == 1023 || (key[t + 1] >> 10) != bt;
        if (bt != kNone && last) planeCount[bt] += rank + 1u;
        __syncthreads();
    }
}

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of this device's scratch array, looked up once per device (with it current).
NoAllocScratch* device_scratch()
{
    int dev = 0;
    PT_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("pt_fitting_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        PT_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw Error("pt_fitting_noalloc: no device address for the scratch array");
        g_scratch[dev] = ptr;
    }
    return static_cast<NoAllocScratch*>(g_scratch[dev]);
}

} // namespace
} // namespace pt

extern "C" PT_API uint64_t PT_CALL pt_fitting_noalloc(const uint16_t* d_input, const double* d_positions, uint64_t nPoints,
                                                      double* d_output, uint64_t height, uint64_t width, uint64_t planes,
                                                      uint64_t fit, const double* fit_params)
{
    using namespace pt;
    try {
        if (nPoints == 0) return 0;
        if (nPoints > kNoAllocMaxPoints)
            throw Error("at most " + std::to_string(kNoAllocMaxPoints) + " positions per call (got " +
                        std::to_string(nPoints) + ")");
        if (planes == 0 || planes > kNoAllocMaxPlanes)
            throw Error("planes must be in [1, " + std::to_string(kNoAllocMaxPlanes) + "] (got " + std::to_string(planes) + ")");
        if (fit < 1 || fit > 4) throw Error("the Gaussian fit modes 1-4 only (the spline model of mode 5 is uploaded by Fitting)");
        if (height == 0 || width == 0 || height > 0x7FFFFFFFull || width > 0x7FFFFFFFull) throw Error("bad frame size");
        const Options opt = Engine::shared_options();
        const FitParams fp = gaussian_fit_params(fit, fit_params, opt);
        NoAllocScratch* ws = device_scratch();
        const hipStream_t s = nullptr;   // legacy default stream: no stream is created
        noalloc_zero_kernel<<<64, 256, 0, s>>>(ws->planeCount, planes + 1, ws->words, 16);
        PT_CHECK(hipGetLastError());
        launch_positions_to_rois(d_positions, nPoints, planes, ws->rois, ws->words + 4, s);
        noalloc_local_index_kernel<<<1, 1024, 0, s>>>(d_positions, nPoints, planes, ws->planeCount, ws->local);
        PT_CHECK(hipGetLastError());
        const size_t npx = (size_t)height * width;
        launch_fit(d_input, npx, (int)width, (int)height, ws->rois, (uint32_t)nPoints, fp, d_output, opt.fitFp64,
                   FitScratch{ws->words, ws->hand}, s, nullptr, nullptr, opt.fitWaves);
        launch_positions_ids(d_positions, ws->local, nPoints, d_output, s);
        uint32_t bad = 0;
        PT_CHECK(hipMemcpyAsync(&bad, ws->words + 4, 4, hipMemcpyDeviceToHost, s));
        PT_CHECK(hipStreamSynchronize(s));
        if (bad) std::fprintf(stderr, "[WARNING] Fitting: positions with a plane outside [1, %llu] got failure rows.\n",
                              (unsigned long long)planes);
        return nPoints * kCols;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_fitting_noalloc: %s\n", e.what());
    } catch (...) {
        std::fprintf(stderr, "[palmtracer] pt_fitting_noalloc: unknown error\n");
    }
    return 0;
}
