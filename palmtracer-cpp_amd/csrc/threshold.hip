// AutoThreshold on the device (lib/Palm.cpp:122-172, CPU/main.cpp:75-85).
//
// sigma of the whole plane, then up to THRESHOLD_ITERATION times: detect with the current
// threshold (watershed ratio 10, no fit), mask a ROI square around every detection, and take the
// standard deviation of the remaining pixels as the next threshold. Counts and sums of u16 pixels
// are exact integers (the reference's FP64 sums of integers are exact too); the sum of squared
// deviations is an FP64 reduction in a fixed order (per block, then across blocks).
#include "engine.h"

#include <cfloat>
#include <cmath>

namespace pt {
namespace {
constexpr int NT = 256;
constexpr int NB = 1024;

__global__ void at_zero_kernel(uint8_t* __restrict__ m, size_t n, unsigned long long* __restrict__ acc)
{
    for (size_t i = blockIdx.x * (size_t)NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) m[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x < 2) acc[threadIdx.x] = 0ull;
}

// ROI squares of the detections (Palm.cpp:141-147): x in [max(0, (int)(X - h)), min(W, (size_t)(X + h))).
__global__ void at_mask_kernel(const double* __restrict__ rows, uint32_t n, uint8_t* __restrict__ m, int W, int H,
                               double half)
{
    for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        const double X = rows[(size_t)i * kCols + kX], Y = rows[(size_t)i * kCols + kY];
        const int xm = (int)(X - half), ym = (int)(Y - half);
        const long xmin = xm > 0 ? xm : 0, ymin = ym > 0 ? ym : 0;
        const long xa = (long)(X + half), ya = (long)(Y + half);
        const long xmax = xa < W ? xa : W, ymax = ya < H ? ya : H;
        for (long y = ymin; y < ymax; ++y)
            for (long x = xmin; x < xmax; ++x) m[y * W + x] = 1;
    }
}

__global__ void at_moments_kernel(const uint16_t* __restrict__ img, const uint8_t* __restrict__ m, size_t n,
                                  unsigned long long* __restrict__ acc)
{
    unsigned long long c = 0, s = 0;
    for (size_t i = blockIdx.x * (size_t)NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT)
        if (!m || !m[i]) { ++c; s += img[i]; }
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o);
        s += __shfl_xor(s, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&acc[0], c);
        atomicAdd(&acc[1], s);
    }
}

__global__ void __launch_bounds__(NT) at_var_kernel(const uint16_t* __restrict__ img, const uint8_t* __restrict__ m,
                                                    size_t n, double mu, double* __restrict__ part)
{
    __shared__ double sh[NT];
    double v = 0.0;
    for (size_t i = blockIdx.x * (size_t)NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT)
        if (!m || !m[i]) {
            const double d = (double)img[i] - mu;
            v += d * d;
        }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void at_final_kernel(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < nb; ++i) s += part[i];
        out[0] = s;
    }
}
} // namespace

double Engine::autothreshold(const uint16_t* d_frame, int H, int W, const double* fitParams)
{
    const size_t npx = (size_t)W * H;
    const double fp0 = fitParams ? fitParams[0] : 7.0;
    const double half = fp0 / 2;
    atMask_.ensure(npx);
    atPart_.ensure((NB + 4) * 8 + 16);
    double* part = atPart_.as<double>();
    double* varOut = part + NB;
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(part + NB + 2);
    uint8_t* mask = atMask_.as<uint8_t>();
    const unsigned g = (unsigned)std::min<size_t>(NB, (npx + NT - 1) / NT);

    auto moments = [&](const uint8_t* m, uint64_t& count, double& mu, double& var) {
        at_zero_kernel<<<1, NT, 0, stream>>>(mask, 0, acc);
        at_moments_kernel<<<g, NT, 0, stream>>>(d_frame, m, npx, acc);
        unsigned long long cs[2];
        PT_CHECK(hipMemcpyAsync(cs, acc, sizeof(cs), hipMemcpyDeviceToHost, stream));
        PT_CHECK(hipStreamSynchronize(stream));
        count = cs[0];
        if (count == 0) return;
        mu = (double)cs[1] / (double)count;
        at_var_kernel<<<g, NT, 0, stream>>>(d_frame, m, npx, mu, part);
        at_final_kernel<<<1, 64, 0, stream>>>(part, (int)g, varOut);
        PT_CHECK(hipMemcpyAsync(&var, varOut, 8, hipMemcpyDeviceToHost, stream));
        PT_CHECK(hipStreamSynchronize(stream));
        var /= (double)count;
    };

    uint64_t count = 0;
    double mu = 0, var = 0;
    moments(nullptr, count, mu, var);
    double thr = std::sqrt(var);
    // Palm(..., pointsListSize = 0.2 * h * w * 18, ratio 10, fit 0) (CPU/main.cpp:79-83).
    const uint64_t listSize = (uint64_t)(0.2 * (double)npx * kCols);
    const uint64_t maxPer = listSize / kCols;
    const double fpx[4] = {fp0, 0, 0, 0};
    atRows_.ensure(std::max<uint64_t>(1, std::min<uint64_t>(maxPer, npx / 4 + 2)) * kCols * 8);
    for (int it = 0; it < kThresholdIterations; ++it) {
        uint64_t n = 0;
        if (maxPer > 0)
            n = localize(d_frame, 1, H, W, thr, 10.0, 0, fpx, maxPer, atRows_.as<double>(),
                         std::min<uint64_t>(maxPer, npx / 4 + 2)) / kCols;
        at_zero_kernel<<<(unsigned)std::min<size_t>(1024, (npx + NT - 1) / NT), NT, 0, stream>>>(mask, npx, acc);
        if (n) at_mask_kernel<<<(unsigned)std::min<uint64_t>(1024, (n + NT - 1) / NT), NT, 0, stream>>>(
            atRows_.as<double>(), (uint32_t)n, mask, W, H, half);
        PT_CHECK(hipGetLastError());
        moments(mask, count, mu, var);
        if (count == 0) break;
        const double sd = std::sqrt(var);
        if (std::fabs(thr - sd) <= DBL_EPSILON) break;
        thr = sd;
    }
    return thr;
}

} // namespace pt
