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
__global__ void libm_kernel(uint32_t fn, const double* x, double* y, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        double sn, cn;
        if (fn != 0) pt::libm_sincos(x[i], sn, cn);
        y[i] = fn == 0 ? pt::libm_exp(x[i]) : fn == 1 ? sn : cn;
    }
}
} // namespace

PT_API int PT_CALL pt_debug_libm(uint64_t fn, const double* d_x, double* d_y, uint64_t n)
{
    try {
        if (fn > 2) throw pt::Error("fn must be 0 (exp), 1 (sin of sincos) or 2 (cos of sincos)");
        if (n == 0) return 1;
        const unsigned grid = (unsigned)std::min<uint64_t>(4096, (n + 255) / 256);
        libm_kernel<<<grid, 256>>>((uint32_t)fn, d_x, d_y, n);
        PT_CHECK(hipGetLastError());
        PT_CHECK(hipDeviceSynchronize());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_debug_libm: %s\n", e.what());
    }
    return 0;
}
