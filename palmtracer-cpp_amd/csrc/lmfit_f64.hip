// FP64 instantiation of the LM fit. Compiled without floating-point contraction so that every
// product and sum rounds separately, as in the reference's x86-64 build.
#pragma clang fp contract(off)
#include "lmfit_impl.h"
#include "../../include/palmtracer.h"

#include <cstdio>

namespace pt {
void launch_fit_f64(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, uint32_t nRoi,
                    const FitParams& fp, double* d_rows, hipStream_t s, unsigned long long* d_passes,
                    const uint32_t* frameMap, const uint32_t* d_count)
{
    launch_fit_t<double>(d_frames, npx, W, H, d_rois, d_count, nRoi, fp, d_rows, d_passes, frameMap, s);
}
} // namespace pt

namespace {
__global__ void libm_exp_kernel(const double* x, double* y, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        y[i] = pt::libm_exp(x[i]);
}
} // namespace

PT_API int PT_CALL pt_debug_exp(const double* d_x, double* d_y, uint64_t n)
{
    try {
        if (n == 0) return 1;
        const unsigned grid = (unsigned)std::min<uint64_t>(4096, (n + 255) / 256);
        libm_exp_kernel<<<grid, 256>>>(d_x, d_y, n);
        PT_CHECK(hipGetLastError());
        PT_CHECK(hipDeviceSynchronize());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_debug_exp: %s\n", e.what());
    }
    return 0;
}
