// Output assembly: per-frame (Y, X) row order and the Id / Plane / Index columns.
//
// Reference: lib/localization/FrameProcessor.cpp:35-39 (sort by Y then X with std::weak_order) and
// lib/Palm.cpp:24-43 (assembleLocalizations: global Id from 1, Plane from 1, Index per plane from 1).
#include "pt_common.h"
#include "pt_kernels.h"

namespace pt {
namespace {
constexpr int NT = 256;
constexpr int LDS_ROWS = 2048;

// std::weak_order on doubles as a signed integer order: all negative NaNs < -inf < ... < -0 == +0
// < ... < +inf < all positive NaNs.
__device__ inline int64_t weak_key(double v)
{
    if (v != v) return signbit(v) ? INT64_MIN : INT64_MAX;
    if (v == 0.0) v = 0.0;
    const int64_t b = __double_as_longlong(v);
    return b >= 0 ? b : (b ^ INT64_MAX);
}

__device__ inline bool before(int64_t ya, int64_t xa, uint32_t ia, int64_t yb, int64_t xb, uint32_t ib)
{
    return ya != yb ? ya < yb : (xa != xb ? xa < xb : ia < ib);
}

__global__ void __launch_bounds__(NT) assemble_kernel(const double* __restrict__ rows, const uint32_t* __restrict__ roiBase,
                                                      uint64_t idBase, uint64_t planeBase, double* __restrict__ out)
{
    __shared__ int64_t ky[LDS_ROWS], kx[LDS_ROWS];
    const int f = blockIdx.x;
    const uint32_t b0 = roiBase[f], n = roiBase[f + 1] - b0;
    const bool lds = n <= LDS_ROWS;
    if (lds)
        for (uint32_t i = threadIdx.x; i < n; i += NT) {
            ky[i] = weak_key(rows[(size_t)(b0 + i) * kCols + kY]);
            kx[i] = weak_key(rows[(size_t)(b0 + i) * kCols + kX]);
        }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
        const double* r = rows + (size_t)(b0 + i) * kCols;
        const int64_t yi = lds ? ky[i] : weak_key(r[kY]), xi = lds ? kx[i] : weak_key(r[kX]);
        uint32_t rank = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const int64_t yj = lds ? ky[j] : weak_key(rows[(size_t)(b0 + j) * kCols + kY]);
            const int64_t xj = lds ? kx[j] : weak_key(rows[(size_t)(b0 + j) * kCols + kX]);
            rank += before(yj, xj, j, yi, xi, i);
        }
        double* o = out + (size_t)(b0 + rank) * kCols;
        for (int c = 0; c < kCols; ++c) o[c] = r[c];
        o[kId] = (double)(idBase + b0 + rank + 1);
        o[kPlane] = (double)(planeBase + f + 1);
        o[kIndex] = (double)(rank + 1);
    }
}
} // namespace

void launch_assemble(const double* d_rows, const uint32_t* d_roiBase, int frames, uint64_t idBase, uint64_t planeBase,
                     double* d_out, hipStream_t s)
{
    if (frames <= 0) return;
    assemble_kernel<<<frames, NT, 0, s>>>(d_rows, d_roiBase, idBase, planeBase, d_out);
    PT_CHECK(hipGetLastError());
}

} // namespace pt
