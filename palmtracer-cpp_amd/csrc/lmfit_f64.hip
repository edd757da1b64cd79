// FP64 instantiation of the LM fit. Compiled without floating-point contraction so that every
// product and sum rounds separately, as in the reference's x86-64 build.
#pragma clang fp contract(off)
#include "lmfit_impl.h"

namespace pt {
void launch_fit_f64(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, uint32_t nRoi,
                    const FitParams& fp, double* d_rows, hipStream_t s, unsigned long long* d_passes,
                    const uint32_t* frameMap, const uint32_t* d_count)
{
    launch_fit_t<double>(d_frames, npx, W, H, d_rois, d_count, nRoi, fp, d_rows, d_passes, frameMap, s);
}
} // namespace pt
