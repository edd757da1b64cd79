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

// Frames that can hold more rows than the ranking kernel keeps in LDS: a sort spread over many
// workgroups. Keys (weak Y, weak X, frame-local row) live in a scratch of kSortScratchPerRow bytes per
// row as structure-of-arrays, two copies for ping-pong: ky[2][n], kx[2][n] (int64), id[2][n] (u32).
//   1. sort_tiles_kernel: tiles of TILE rows of a frame, bitonic-sorted in LDS (one workgroup each);
//   2. merge_pass_kernel: sorted runs of L rows merged pairwise; a row's place in the merged run is its
//      place in its own run plus the rows of the partner run before it (binary search). Keys are
//      distinct (the row number breaks ties), so the places form a permutation;
//   3. gather_rows_kernel: rows copied to their sorted place, consecutive threads on consecutive doubles.
constexpr int TILE = 2048;
constexpr int TILE_NT = 512;
constexpr int MERGE_NT = 256;
constexpr int GATHER_NT = 256;
constexpr int GATHER_PER_THREAD = 4;

struct SortKeys {
    int64_t* ky[2];
    int64_t* kx[2];
    uint32_t* id[2];
};

__host__ __device__ inline SortKeys sort_keys(void* scratch, uint64_t n)
{
    SortKeys k;
    int64_t* w = static_cast<int64_t*>(scratch);
    k.ky[0] = w;
    k.ky[1] = w + n;
    k.kx[0] = w + 2 * n;
    k.kx[1] = w + 3 * n;
    uint32_t* u = reinterpret_cast<uint32_t*>(w + 4 * n);
    k.id[0] = u;
    k.id[1] = u + n;
    return k;
}

__global__ void __launch_bounds__(TILE_NT) sort_tiles_kernel(const double* __restrict__ rows,
                                                             const uint32_t* __restrict__ roiBase, uint32_t tilesPerFrame,
                                                             SortKeys keys)
{
    __shared__ int64_t ky[TILE], kx[TILE];
    __shared__ uint32_t id[TILE];
    const uint32_t f = blockIdx.x / tilesPerFrame, t = blockIdx.x - f * tilesPerFrame;
    const uint32_t b0 = roiBase[f], n = roiBase[f + 1] - b0, start = t * TILE;
    if (start >= n) return;
    const uint32_t m = min<uint32_t>(TILE, n - start);
    uint32_t m2 = 2;
    while (m2 < m) m2 <<= 1;
    for (uint32_t i = threadIdx.x; i < m2; i += TILE_NT) {
        if (i < m) {
            const double* r = rows + (size_t)(b0 + start + i) * kCols;
            ky[i] = weak_key(r[kY]);
            kx[i] = weak_key(r[kX]);
            id[i] = start + i;
        } else {
            // Padding sorts after every row: largest keys and a row number no frame reaches.
            ky[i] = INT64_MAX;
            kx[i] = INT64_MAX;
            id[i] = 0xFFFFFFFFu;
        }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= m2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t p = threadIdx.x; p < m2 / 2; p += TILE_NT) {
                const uint32_t i = (p / j) * 2 * j + (p % j), o = i + j;
                const bool up = (i & k) == 0;
                const int64_t ya = ky[i], xa = kx[i], yb = ky[o], xb = kx[o];
                const uint32_t ia = id[i], ib = id[o];
                if (up ? before(yb, xb, ib, ya, xa, ia) : before(ya, xa, ia, yb, xb, ib)) {
                    ky[i] = yb; kx[i] = xb; id[i] = ib;
                    ky[o] = ya; kx[o] = xa; id[o] = ia;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < m; i += TILE_NT) {
        const size_t g = (size_t)b0 + start + i;
        keys.ky[0][g] = ky[i];
        keys.kx[0][g] = kx[i];
        keys.id[0][g] = id[i];
    }
}

// One pass: runs of L rows (sorted, in copy `from`) merged pairwise into copy 1 - from.
__global__ void __launch_bounds__(MERGE_NT) merge_pass_kernel(const uint32_t* __restrict__ roiBase, uint32_t blocksPerFrame,
                                                              uint32_t L, int from, SortKeys keys)
{
    const uint32_t f = blockIdx.x / blocksPerFrame, c = blockIdx.x - f * blocksPerFrame;
    const uint32_t b0 = roiBase[f], n = roiBase[f + 1] - b0, q = c * MERGE_NT + threadIdx.x;
    if (q >= n) return;
    const int64_t* sy = keys.ky[from] + b0;
    const int64_t* sx = keys.kx[from] + b0;
    const uint32_t* si = keys.id[from] + b0;
    const int64_t y = sy[q], x = sx[q];
    const uint32_t i = si[q];
    const uint32_t r = q / L, own = q - r * L, ps = (r ^ 1u) * L;
    uint32_t before_n = 0;
    if (ps < n) {
        // Rows of the partner run that precede (y, x, i): lower bound by binary search.
        uint32_t lo = ps, hi = min(ps + L, n);
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (before(sy[mid], sx[mid], si[mid], y, x, i)) lo = mid + 1;
            else hi = mid;
        }
        before_n = lo - ps;
    }
    const uint32_t d = (r & ~1u) * L + own + before_n;
    keys.ky[1 - from][b0 + d] = y;
    keys.kx[1 - from][b0 + d] = x;
    keys.id[1 - from][b0 + d] = i;
}

__global__ void __launch_bounds__(GATHER_NT) gather_rows_kernel(const double* __restrict__ rows,
                                                                const uint32_t* __restrict__ roiBase,
                                                                uint32_t blocksPerFrame, const uint32_t* __restrict__ order,
                                                                double* __restrict__ staged)
{
    const uint32_t f = blockIdx.x / blocksPerFrame, c = blockIdx.x - f * blocksPerFrame;
    const uint32_t b0 = roiBase[f], n = roiBase[f + 1] - b0;
    const uint32_t d0 = c * (GATHER_NT * GATHER_PER_THREAD);
    if (d0 >= n * kCols) return;
#pragma unroll
    for (int k = 0; k < GATHER_PER_THREAD; ++k) {
        const uint32_t d = d0 + k * GATHER_NT + threadIdx.x;
        if (d >= n * kCols) break;
        const uint32_t r = d / kCols, col = d - r * kCols;
        staged[(size_t)b0 * kCols + d] =
            col == kIndex ? (double)(r + 1) : rows[((size_t)b0 + order[b0 + r]) * kCols + col];
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
                        uint64_t maxRows, uint64_t rows, void* d_scratch)
{
    if (frames <= 0) return;
    if (maxRows <= (uint64_t)LDS_ROWS) {
        sort_frames_kernel<<<frames, NT, 0, s>>>(d_rows, d_roiBase, d_staged);
        PT_CHECK(hipGetLastError());
        return;
    }
    if (!d_scratch) throw Error("the per-frame sort needs a key scratch");
    if (maxRows > 0x7FFFFFFFull) throw Error("too many rows per frame for the per-frame sort");
    // Every row index the kernels use lies below roiBase[frames] <= rows, the rows at d_rows.
    const uint32_t tiles = (uint32_t)((maxRows + TILE - 1) / TILE);
    if ((uint64_t)frames * tiles > 0x7FFFFFFFull) throw Error("per-frame sort grid too large");
    const SortKeys k = sort_keys(d_scratch, rows);
    sort_tiles_kernel<<<(unsigned)(frames * tiles), TILE_NT, 0, s>>>(d_rows, d_roiBase, tiles, k);
    PT_CHECK(hipGetLastError());
    int from = 0;
    const uint32_t mBlocks = (uint32_t)((maxRows + MERGE_NT - 1) / MERGE_NT);
    if ((uint64_t)frames * mBlocks > 0x7FFFFFFFull) throw Error("per-frame sort grid too large");
    for (uint64_t L = TILE; L < maxRows; L <<= 1) {
        merge_pass_kernel<<<(unsigned)(frames * mBlocks), MERGE_NT, 0, s>>>(d_roiBase, mBlocks, (uint32_t)L, from, k);
        PT_CHECK(hipGetLastError());
        from ^= 1;
    }
    const uint64_t per = (uint64_t)GATHER_NT * GATHER_PER_THREAD;
    const uint32_t gBlocks = (uint32_t)((maxRows * kCols + per - 1) / per);
    if ((uint64_t)frames * gBlocks > 0x7FFFFFFFull) throw Error("per-frame sort grid too large");
    gather_rows_kernel<<<(unsigned)(frames * gBlocks), GATHER_NT, 0, s>>>(d_rows, d_roiBase, gBlocks, k.id[from], d_staged);
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
