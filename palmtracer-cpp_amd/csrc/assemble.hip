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

// One workgroup per frame: rank every row of the frame by (Y, X) (ties by position, which the
// reference's unstable sort leaves unspecified) and write it at its rank.
__global__ void __launch_bounds__(NT) sort_frames_kernel(const double* __restrict__ rows, const uint32_t* __restrict__ roiBase,
                                                         double* __restrict__ staged)
{
    __shared__ int64_t ky[LDS_ROWS], kx[LDS_ROWS];
    __shared__ uint16_t perm[LDS_ROWS];          // perm[rank] = the frame's row of that rank
    const int f = blockIdx.x;
    const uint32_t b0 = roiBase[f], n = roiBase[f + 1] - b0;
    const bool lds = n <= LDS_ROWS;
    if (lds) {
        for (uint32_t i = threadIdx.x; i < n; i += NT) {
            ky[i] = weak_key(rows[(size_t)(b0 + i) * kCols + kY]);
            kx[i] = weak_key(rows[(size_t)(b0 + i) * kCols + kX]);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += NT) {
            uint32_t rank = 0;
            for (uint32_t j = 0; j < n; ++j) rank += before(ky[j], kx[j], j, ky[i], kx[i], i);
            perm[rank] = (uint16_t)i;
        }
        __syncthreads();
        // Rows to their sorted place, one double per thread (consecutive threads, consecutive doubles).
        for (uint32_t k = threadIdx.x; k < n * kCols; k += NT) {
            const uint32_t r = k / kCols, c = k - r * kCols;
            staged[(size_t)b0 * kCols + k] = c == kIndex ? (double)(r + 1) : rows[(size_t)(b0 + perm[r]) * kCols + c];
        }
        return;
    }
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
        const double* r = rows + (size_t)(b0 + i) * kCols;
        const int64_t yi = lds ? ky[i] : weak_key(r[kY]), xi = lds ? kx[i] : weak_key(r[kX]);
        uint32_t rank = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const int64_t yj = lds ? ky[j] : weak_key(rows[(size_t)(b0 + j) * kCols + kY]);
            const int64_t xj = lds ? kx[j] : weak_key(rows[(size_t)(b0 + j) * kCols + kX]);
            rank += before(yj, xj, j, yi, xi, i);
        }
        double* o = staged + (size_t)(b0 + rank) * kCols;
        for (int c = 0; c < kCols; ++c) o[c] = r[c];
        o[kIndex] = (double)(rank + 1);
    }
}

// Frames with up to SORT_MAX rows, one workgroup per frame: (Y, X, position) keys in LDS, a bitonic
// sort of the row positions, rows written in sorted order. For callers whose frames can hold more
// rows than the ranking kernel above keeps in LDS.
constexpr int SORT_MAX = 8192;
constexpr size_t SORT_LDS = (size_t)SORT_MAX * (8 + 8 + 2);
__global__ void __launch_bounds__(NT) sort_frames_big_kernel(const double* __restrict__ rows,
                                                             const uint32_t* __restrict__ roiBase,
                                                             double* __restrict__ staged)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int64_t* ky = reinterpret_cast<int64_t*>(smem);
    int64_t* kx = ky + SORT_MAX;
    uint16_t* id = reinterpret_cast<uint16_t*>(kx + SORT_MAX);
    constexpr uint16_t kPad = 0xFFFF;
    const int f = blockIdx.x;
    const uint32_t b0 = roiBase[f], n = roiBase[f + 1] - b0;
    if (n > (uint32_t)SORT_MAX) {
        // Beyond the LDS: rank every row against all others (global memory).
        for (uint32_t i = threadIdx.x; i < n; i += NT) {
            const double* r = rows + (size_t)(b0 + i) * kCols;
            const int64_t yi = weak_key(r[kY]), xi = weak_key(r[kX]);
            uint32_t rank = 0;
            for (uint32_t j = 0; j < n; ++j)
                rank += before(weak_key(rows[(size_t)(b0 + j) * kCols + kY]), weak_key(rows[(size_t)(b0 + j) * kCols + kX]),
                               j, yi, xi, i);
            double* o = staged + (size_t)(b0 + rank) * kCols;
            for (int c = 0; c < kCols; ++c) o[c] = r[c];
            o[kIndex] = (double)(rank + 1);
        }
        return;
    }
    uint32_t n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (uint32_t i = threadIdx.x; i < n2; i += NT) {
        if (i < n) {
            ky[i] = weak_key(rows[(size_t)(b0 + i) * kCols + kY]);
            kx[i] = weak_key(rows[(size_t)(b0 + i) * kCols + kX]);
        }
        id[i] = i < n ? (uint16_t)i : kPad;
    }
    __syncthreads();
    auto less = [&](uint16_t a, uint16_t b) {
        if (a == kPad) return false;
        if (b == kPad) return true;
        return before(ky[a], kx[a], a, ky[b], kx[b], b);
    };
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += NT) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint16_t a = id[i], b = id[ixj];
                    if ((i & k) == 0 ? less(b, a) : less(a, b)) { id[i] = b; id[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t r = threadIdx.x; r < n; r += NT) {
        const double* src = rows + (size_t)(b0 + id[r]) * kCols;
        double* o = staged + (size_t)(b0 + r) * kCols;
        for (int c = 0; c < kCols; ++c) o[c] = src[c];
        o[kIndex] = (double)(r + 1);
    }
}

// Placement plan of a call, one workgroup: plane p's rows are the main rows [base[p], base[p + 1])
// of the staged array, unless p is one of the nDef deferred planes listed in def (planes, then
// their first side row, then their row counts), whose rows start at staged row sideOff + that side
// row. dst is the exclusive prefix of the counts in plane order (lib/Palm.cpp:24-43: planes one
// after another), *total their sum.
constexpr int PL_NT = 1024;
__global__ void __launch_bounds__(PL_NT) plan_kernel(const uint32_t* __restrict__ base, uint64_t planes,
                                                     const uint64_t* __restrict__ def, uint32_t nDef, uint64_t sideOff,
                                                     uint64_t* __restrict__ src, uint64_t* __restrict__ dst,
                                                     uint32_t* __restrict__ cnt, uint64_t* __restrict__ total)
{
    __shared__ uint64_t waveSum[PL_NT / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (uint64_t p = t; p < planes; p += PL_NT) {
        cnt[p] = base[p + 1] - base[p];
        src[p] = base[p];
    }
    __syncthreads();
    for (uint32_t k = t; k < nDef; k += PL_NT) {
        const uint64_t p = def[k];
        src[p] = sideOff + def[nDef + k];
        cnt[p] = (uint32_t)def[2 * (uint64_t)nDef + k];
    }
    __syncthreads();
    const uint64_t per = (planes + PL_NT - 1) / PL_NT, p0 = min(planes, t * per), p1 = min(planes, p0 + per);
    uint64_t sum = 0;
    for (uint64_t p = p0; p < p1; ++p) sum += cnt[p];
    uint64_t incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    if (lane == 63) waveSum[wave] = incl;
    __syncthreads();
    uint64_t run = incl - sum;
    for (int w = 0; w < wave; ++w) run += waveSum[w];
    for (uint64_t p = p0; p < p1; ++p) {
        dst[p] = run;
        run += cnt[p];
    }
    if (t == PL_NT - 1) *total = run;
}

// One workgroup per plane: rows to their final place with the global Id and the plane number
// (lib/Palm.cpp:24-43); output rows at or past cap are not written.
__global__ void __launch_bounds__(NT) scatter_kernel(const double* __restrict__ staged, const uint64_t* __restrict__ src,
                                                     const uint64_t* __restrict__ dst, const uint32_t* __restrict__ count,
                                                     double* __restrict__ out, uint64_t p0, uint64_t cap)
{
    const uint64_t p = p0 + blockIdx.x;
    const uint64_t d = dst[p];
    if (d >= cap) return;
    const uint32_t n = (uint32_t)min<uint64_t>(count[p], cap - d);
    const double* a = staged + src[p] * kCols;
    double* b = out + d * kCols;
    for (uint32_t i = threadIdx.x; i < n * kCols; i += NT) {
        const uint32_t r = i / kCols, c = i - r * kCols;
        double v = a[i];
        if (c == kId) v = (double)(d + r + 1);
        else if (c == kPlane) v = (double)(p + 1);
        b[i] = v;
    }
}
} // namespace

void launch_sort_frames(const double* d_rows, const uint32_t* d_roiBase, int frames, double* d_staged, hipStream_t s,
                        uint64_t maxRows)
{
    if (frames <= 0) return;
    if (maxRows <= (uint64_t)LDS_ROWS)
        sort_frames_kernel<<<frames, NT, 0, s>>>(d_rows, d_roiBase, d_staged);
    else
        sort_frames_big_kernel<<<frames, NT, SORT_LDS, s>>>(d_rows, d_roiBase, d_staged);
    PT_CHECK(hipGetLastError());
}

void launch_plan(const uint32_t* d_base, uint64_t planes, const uint64_t* d_def, uint32_t nDef, uint64_t sideOff,
                 uint64_t* d_src, uint64_t* d_dst, uint32_t* d_count, uint64_t* d_total, hipStream_t s)
{
    plan_kernel<<<1, PL_NT, 0, s>>>(d_base, planes, d_def, nDef, sideOff, d_src, d_dst, d_count, d_total);
    PT_CHECK(hipGetLastError());
}

void launch_scatter(const double* d_staged, const uint64_t* d_src, const uint64_t* d_dst, const uint32_t* d_count,
                    uint64_t planes, double* d_out, uint64_t cap, hipStream_t s)
{
    for (uint64_t p0 = 0; p0 < planes; p0 += 1u << 30) {
        const uint64_t n = std::min<uint64_t>(planes - p0, 1u << 30);
        scatter_kernel<<<(unsigned)n, NT, 0, s>>>(d_staged, d_src, d_dst, d_count, d_out, p0, cap);
    }
    PT_CHECK(hipGetLastError());
}

} // namespace pt
