// A trous B3-spline wavelet planes on gfx950, in exact fixed-point integer arithmetic.
//
// Reference: lib/localization/ATrous.cpp:23-80 (decomposition2D, b3SplineFiltering2D, offsetTest).
// V0 = image; V_{j+1} = B3_j(V_j), the separable taps {1,4,6,4,1}/16 at offsets {0, +-s, +-2s},
// s = 2^j, clamp-to-edge at every scale; W_j = V_j - V_{j+1}. The reference evaluates this in FP64
// and, because every value is a dyadic rational with at most 32 fractional bits below 2^16, its
// FP64 result is exact. Here the same numbers are carried as integers:
//   A   = 16    * rowpass_0(V0)         < 2^20
//   V1s = 256   * V1                    < 2^24
//   B   = 4096  * rowpass_1(V1)         < 2^28
//   V2s = 65536 * V2                    < 2^32   (u32, all terms non-negative)
//   P   = 65536 * W1 = 256 * V1s - V2s  (int64; the detection kernel keeps it as u32 when > thr)
// so every plane is bit-identical to the reference's doubles.
//
// One workgroup of 256 threads owns a TILE_W x TILE_H output tile. The halo chain is staged in
// LDS as three windows clipped to the frame (reads of clamped coordinates always land inside the
// window): R0 = input, R1 = V1, O = output (+1 ring in the detection kernel for the local-maximum
// test). Input rows are read coalesced; the separable passes ping-pong between two LDS buffers.
#include "pt_common.h"
#include "pt_kernels.h"

namespace pt {

namespace {
constexpr int TILE_W = 64;
constexpr int TILE_H = 32;
constexpr int NT = 256;
constexpr int RING = 1;                 // local-maximum ring around the detection tile
constexpr int H1 = 4, H0 = 2;           // tap reach of scale 1 and scale 0
constexpr int MAXW0 = TILE_W + 2 * (RING + H1 + H0);
constexpr int MAXH0 = TILE_H + 2 * (RING + H1 + H0);
constexpr int MAXW1 = TILE_W + 2 * (RING + H1);
constexpr int MAXWO = TILE_W + 2 * RING;
constexpr int MAXHO = TILE_H + 2 * RING;
constexpr int L0_SIZE = MAXW0 * MAXH0;  // V0, later V1s
constexpr int L1_SIZE = MAXH0 * MAXW1;  // A, later B
constexpr int L2_SIZE = MAXHO * MAXWO;  // P / output window

struct Win {
    int y0, y1, x0, x1;
    __device__ int w() const { return x1 - x0; }
    __device__ int h() const { return y1 - y0; }
};

__device__ inline int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// Computes V1s over R1 (kept in l0) and V2s over O (returned through the callback per pixel of O).
// ring = halo ring of the output window O around the tile.
template <int Ring, typename Out>
__device__ inline void atrous_tile(const uint16_t* __restrict__ img, int W, int H, int tx0, int ty0,
                                   uint32_t* l0, uint32_t* l1, Win& O, Win& R1, Out&& out)
{
    const int tid = threadIdx.x;
    const int Wm = W - 1, Hm = H - 1;
    O = {max(ty0 - Ring, 0), min(ty0 + TILE_H + Ring, H), max(tx0 - Ring, 0), min(tx0 + TILE_W + Ring, W)};
    R1 = {max(O.y0 - H1, 0), min(O.y1 + H1, H), max(O.x0 - H1, 0), min(O.x1 + H1, W)};
    const Win R0{max(R1.y0 - H0, 0), min(R1.y1 + H0, H), max(R1.x0 - H0, 0), min(R1.x1 + H0, W)};

    // 1. V0 over R0 (coalesced along rows).
    {
        const int w = R0.w(), n = w * R0.h();
        for (int i = tid; i < n; i += NT) {
            const int y = R0.y0 + i / w, x = R0.x0 + i % w;
            l0[i] = img[(size_t)y * W + x];
        }
    }
    __syncthreads();
    // 2. A = 16 * rowpass_0(V0) over rows(R0) x cols(R1), into l1.
    {
        const int w = R1.w(), n = w * R0.h(), w0 = R0.w();
        for (int i = tid; i < n; i += NT) {
            const int x = R1.x0 + i % w;
            const uint32_t* r = l0 + (i / w) * w0;
            const int c = x - R0.x0;
            l1[i] = 6u * r[c] + 4u * (r[clampi(x - 1, Wm) - R0.x0] + r[clampi(x + 1, Wm) - R0.x0]) +
                    r[clampi(x - 2, Wm) - R0.x0] + r[clampi(x + 2, Wm) - R0.x0];
        }
    }
    __syncthreads();
    // 3. V1s = 256 * V1 = colpass_0(A) over R1, into l0.
    {
        const int w = R1.w(), n = w * R1.h();
        for (int i = tid; i < n; i += NT) {
            const int y = R1.y0 + i / w;
            const uint32_t* c = l1 + i % w;
            l0[i] = 6u * c[(y - R0.y0) * w] + 4u * (c[(clampi(y - 1, Hm) - R0.y0) * w] + c[(clampi(y + 1, Hm) - R0.y0) * w]) +
                    c[(clampi(y - 2, Hm) - R0.y0) * w] + c[(clampi(y + 2, Hm) - R0.y0) * w];
        }
    }
    __syncthreads();
    // 4. B = 16 * rowpass_1(V1s) over rows(R1) x cols(O), into l1.
    {
        const int w = O.w(), n = w * R1.h(), w1 = R1.w();
        for (int i = tid; i < n; i += NT) {
            const int x = O.x0 + i % w;
            const uint32_t* r = l0 + (i / w) * w1;
            l1[i] = 6u * r[x - R1.x0] + 4u * (r[clampi(x - 2, Wm) - R1.x0] + r[clampi(x + 2, Wm) - R1.x0]) +
                    r[clampi(x - 4, Wm) - R1.x0] + r[clampi(x + 4, Wm) - R1.x0];
        }
    }
    __syncthreads();
    // 5. V2s = colpass_1(B) over O; hand V1s and V2s to the caller.
    {
        const int w = O.w(), n = w * O.h(), w1 = R1.w();
        for (int i = tid; i < n; i += NT) {
            const int y = O.y0 + i / w, xx = i % w;
            const uint32_t* c = l1 + xx;
            const uint32_t v2s = 6u * c[(y - R1.y0) * w] +
                                 4u * (c[(clampi(y - 2, Hm) - R1.y0) * w] + c[(clampi(y + 2, Hm) - R1.y0) * w]) +
                                 c[(clampi(y - 4, Hm) - R1.y0) * w] + c[(clampi(y + 4, Hm) - R1.y0) * w];
            const uint32_t v1s = l0[(y - R1.y0) * w1 + (O.x0 + xx - R1.x0)];
            out(i, y, O.x0 + xx, v1s, v2s);
        }
    }
}

// Selected plane as FP64 (the Wavelett export): level 0 = W0, 1 = W1, 2 = V2.
__global__ void __launch_bounds__(NT) atrous_plane_kernel(const uint16_t* __restrict__ in, double* __restrict__ out,
                                                          int W, int H, int level)
{
    __shared__ uint32_t l0[L0_SIZE];
    __shared__ uint32_t l1[L1_SIZE];
    const size_t npx = (size_t)W * H;
    const uint16_t* img = in + (size_t)blockIdx.z * npx;
    double* o = out + (size_t)blockIdx.z * npx;
    Win O, R1;
    atrous_tile<0>(img, W, H, blockIdx.x * TILE_W, blockIdx.y * TILE_H, l0, l1, O, R1,
                   [&](int, int y, int x, uint32_t v1s, uint32_t v2s) {
                       double v;
                       if (level == 0) v = (double)((int64_t)img[(size_t)y * W + x] * 256 - (int64_t)v1s) * (1.0 / 256.0);
                       else if (level == 1) v = (double)((int64_t)v1s * 256 - (int64_t)v2s) * kPScale;
                       else v = (double)v2s * kPScale;
                       o[(size_t)y * W + x] = v;
                   });
}

// Detection front end: P = 65536 * W1 where W1 > threshold (and > 0), else 0; local maxima.
//   P[frame*npx + i]           dense, every pixel
//   lab[frame*npx + i] = i     for pixels with P > 0 (union-find initialisation)
//   maskList / seedList        per-frame compact lists (capacity npx each) of pixel indices with
//                              P > 0 / local maxima (Segmentation2D.cpp:294-333 test: no 8-neighbour
//                              inside the frame strictly greater), counts in maskCnt / seedCnt.
__global__ void __launch_bounds__(NT) atrous_detect_kernel(const uint16_t* __restrict__ in, uint32_t* __restrict__ P,
                                                           uint32_t* __restrict__ lab, uint32_t* __restrict__ maskList,
                                                           uint32_t* __restrict__ maskCnt, uint32_t* __restrict__ seedList,
                                                           uint32_t* __restrict__ seedCnt, int W, int H, int64_t tmin,
                                                           const uint32_t* __restrict__ frameMap)
{
    __shared__ uint32_t l0[L0_SIZE];
    __shared__ uint32_t l1[L1_SIZE];
    __shared__ uint32_t lp[L2_SIZE];
    __shared__ uint32_t sCnt[2], sBase[2];
    const size_t npx = (size_t)W * H;
    const int f = blockIdx.z;
    const uint16_t* img = in + (size_t)(frameMap ? frameMap[f] : (uint32_t)f) * npx;
    const int tx0 = blockIdx.x * TILE_W, ty0 = blockIdx.y * TILE_H;
    if (threadIdx.x < 2) sCnt[threadIdx.x] = 0;
    Win O, R1;
    atrous_tile<RING>(img, W, H, tx0, ty0, l0, l1, O, R1, [&](int i, int, int, uint32_t v1s, uint32_t v2s) {
        const int64_t w1s = (int64_t)v1s * 256 - (int64_t)v2s;
        lp[i] = w1s >= tmin ? (uint32_t)w1s : 0u;
    });
    __syncthreads();

    // Each thread owns up to 8 tile pixels: a column x and rows r, r+4, ..., so rows stay coalesced.
    const int ow = O.w();
    const int lx = threadIdx.x % TILE_W, ly0 = threadIdx.x / TILE_W;
    uint32_t myMask = 0, mySeed = 0;  // bit k: pixel k of this thread
    uint32_t pv[TILE_H / 4];
    const int x = tx0 + lx;
#pragma unroll
    for (int k = 0; k < TILE_H / 4; ++k) {
        const int y = ty0 + ly0 + 4 * k;
        pv[k] = 0;
        if (x >= W || y >= H) continue;
        const int li = (y - O.y0) * ow + (x - O.x0);
        const uint32_t v = lp[li];
        pv[k] = v;
        if (v == 0) continue;
        myMask |= 1u << k;
        bool mx = true;
        for (int dy = -1; dy <= 1; ++dy) {
            const int yy = y + dy;
            if (yy < 0 || yy >= H) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = x + dx;
                if ((dx | dy) == 0 || xx < 0 || xx >= W) continue;
                if (v < lp[(yy - O.y0) * ow + (xx - O.x0)]) mx = false;
            }
        }
        if (mx) mySeed |= 1u << k;
    }
    // Dense P for the tile.
#pragma unroll
    for (int k = 0; k < TILE_H / 4; ++k) {
        const int y = ty0 + ly0 + 4 * k;
        if (x < W && y < H) P[(size_t)f * npx + (size_t)y * W + x] = pv[k];
    }
    const uint32_t nm = __popc(myMask), ns = __popc(mySeed);
    uint32_t om = 0, os = 0;
    if (nm) om = atomicAdd(&sCnt[0], nm);
    if (ns) os = atomicAdd(&sCnt[1], ns);
    __syncthreads();
    if (threadIdx.x == 0) {
        sBase[0] = sCnt[0] ? atomicAdd(&maskCnt[f], sCnt[0]) : 0;
        sBase[1] = sCnt[1] ? atomicAdd(&seedCnt[f], sCnt[1]) : 0;
    }
    __syncthreads();
    uint32_t* ml = maskList + (size_t)f * npx + sBase[0] + om;
    uint32_t* sl = seedList + (size_t)f * npx + sBase[1] + os;
    uint32_t* lb = lab + (size_t)f * npx;
#pragma unroll
    for (int k = 0; k < TILE_H / 4; ++k) {
        const uint32_t idx = (uint32_t)((ty0 + ly0 + 4 * k) * W + x);
        if (myMask & (1u << k)) { *ml++ = idx; lb[idx] = idx; }
        if (mySeed & (1u << k)) *sl++ = idx;
    }
}
} // namespace

void launch_atrous_plane(const uint16_t* d_in, double* d_out, int W, int H, int planes, int level, hipStream_t s)
{
    if (planes <= 0 || W <= 0 || H <= 0) return;
    for (int z0 = 0; z0 < planes; z0 += 65535) {
        const int nz = min(planes - z0, 65535);
        const dim3 grid(ceil_div(W, TILE_W), ceil_div(H, TILE_H), nz);
        atrous_plane_kernel<<<grid, NT, 0, s>>>(d_in + (size_t)z0 * W * H, d_out + (size_t)z0 * W * H, W, H, level);
    }
    PT_CHECK(hipGetLastError());
}

void launch_atrous_detect(const uint16_t* d_in, const DetectBuffers& b, int W, int H, int frames, int64_t tmin,
                          hipStream_t s, const uint32_t* frameMap)
{
    if (frames <= 0 || W <= 0 || H <= 0) return;
    if (frames > 65535) throw Error("launch_atrous_detect: at most 65535 frames per launch");
    const dim3 grid(ceil_div(W, TILE_W), ceil_div(H, TILE_H), frames);
    atrous_detect_kernel<<<grid, NT, 0, s>>>(d_in, b.P, b.lab, b.maskList, b.maskCnt, b.seedList, b.seedCnt, W, H, tmin, frameMap);
    PT_CHECK(hipGetLastError());
}

int64_t threshold_to_pmin(double thr)
{
    // W1 > thr  <=>  P > thr * 2^16  <=>  P >= floor(thr * 2^16) + 1 (P integer); P >= 1 keeps W1 > 0,
    // which the segmentation requires anyway (Segmentation2D.hpp:80, THRESHOLD = 1e-8 < 2^-16).
    if (!(thr == thr)) return INT64_MAX;           // NaN: no pixel passes W1 > thr
    const double t = thr * 65536.0;
    if (t >= 8589934592.0) return INT64_MAX;       // beyond any representable P
    if (t < 0) return 1;
    const int64_t f = (int64_t)floor(t) + 1;
    return f < 1 ? 1 : f;
}

} // namespace pt
