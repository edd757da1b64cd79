// FP32 instantiation of the LM fit: model, Jacobian and sums in FP32 (FP64 parameters and solve).
#include "lmfit_impl.h"

namespace pt {
void launch_fit_f32(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, const uint32_t* d_count,
                    uint32_t nHost, const FitParams& fp, double* d_rows, hipStream_t s, unsigned long long* d_passes,
                    const uint32_t* frameMap)
{
    launch_fit_t<float>(d_frames, npx, W, H, d_rois, d_count, nHost, fp, d_rows, d_passes, frameMap, s);
}

void launch_fit_f64(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, const uint32_t* d_count,
                    uint32_t nHost, const FitParams& fp, double* d_rows, hipStream_t s, unsigned long long* d_passes,
                    const uint32_t* frameMap);

void launch_fit(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, const uint32_t* d_count,
                uint32_t nHost, const FitParams& fp, double* d_rows, bool fp64, hipStream_t s,
                unsigned long long* d_passes, const uint32_t* frameMap)
{
    if (fp64) launch_fit_f64(d_frames, npx, W, H, d_rois, d_count, nHost, fp, d_rows, s, d_passes, frameMap);
    else launch_fit_f32(d_frames, npx, W, H, d_rois, d_count, nHost, fp, d_rows, s, d_passes, frameMap);
}
} // namespace pt
