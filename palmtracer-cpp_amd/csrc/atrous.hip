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

#include <type_traits>

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

// ---------------------------------------------------------------------------------------------
// Detection front end, one wave per (column strip, row segment) of a frame.
//
// A strip is 64 lanes x SC columns (lane l holds columns 3l, 3l+1, 3l+2 of the strip); it covers
// SOUT output columns (lanes 3..60) plus 9 halo columns (3 lanes) on each side, more than the
// 7-column dependency reach of the local-maximum test through both scales. The wave walks its
// segment's rows top to bottom as a pipeline: input row t gives A(t) (row pass, scale 0), V1s(t-2)
// (column pass from a 5-row ring of A), B(t-2) (row pass, scale 1), V2s(t-6) (column pass from a
// 10-row ring of B), P(t-6) and the local-maximum test of row t-7 (running column maxima of P). Horizontal
// neighbours come from the adjacent lanes through DPP wave shifts; vertical neighbours from
// registers. Clamp-to-edge is applied per plane as in the reference: input rows and columns are
// read clamped (so A is right everywhere), V1s columns outside the frame are replaced by the edge
// column, and the B ring is filled with B(0) at the top and repeats B(H-1) below the frame.
// Rows whose every stage lies inside the frame and the segment run in a loop unrolled by a multiple
// of the ring sizes without those edge cases: within an iteration the rings are addressed circularly
// with compile-time slots, and after it they are back in their shifted layout, so rows never move
// between registers.
//
//   P[frame*npx + i]           65536 * W1 where it is >= tmin (the mask pixels), written there only:
//                              the caller keeps P zero at every other pixel (it clears the mask
//                              pixels of a chunk once the chunk's segmentation is done), so the
//                              kernel writes 4 bytes per mask pixel instead of per pixel.
//   maskList / seedList        per-frame compact lists (capacity npx each, any order) of pixel
//                              indices with P > 0 / local maxima (Segmentation2D.cpp:294-333
//                              test: no 8-neighbour inside the frame strictly greater), counts in
//                              maskCnt / seedCnt. Entries are staged per wave in LDS and appended
//                              with one atomic per batch.
//   seedBits                   per-frame bitmap of the local maxima (zeroed by the caller).
constexpr int SC = 3;                   // columns per lane
constexpr int SW = 64 * SC;             // strip width
constexpr int SHALO = 9;                // halo columns on each side (3 lanes)
constexpr int SOUT = SW - 2 * SHALO;    // output columns per strip
constexpr int MCAP = 1024;              // staged mask entries per wave
constexpr int SCAP = 256;               // staged seed entries per wave
constexpr int UNROLL = 10;              // rows per iteration of the interior loop (a multiple of the ring sizes)
constexpr int NA = 5, NV = 5, NB = 10;  // ring sizes: A rows, V1s rows, B rows (9 used)
#ifndef PT_ROW_PREFETCH
#define PT_ROW_PREFETCH 2
#endif
constexpr int PF = PT_ROW_PREFETCH;     // input rows loaded ahead of the row in use (divides UNROLL)
static_assert(UNROLL % NA == 0 && UNROLL % NV == 0 && UNROLL % NB == 0 && UNROLL % PF == 0, "ring sizes");
constexpr uint64_t kOutLanes = ((1ull << (64 - SHALO / SC)) - 1ull) & ~((1ull << (SHALO / SC)) - 1ull);
constexpr int kIcmpNe = 33, kIcmpUge = 35;  // LLVM integer predicates of __builtin_amdgcn_uicmp

// Value of lane (l - 1) / (l + 1) of the wave (DPP wave shifts; edge lanes read 0).
__device__ inline uint32_t from_left(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true); }
__device__ inline uint32_t from_right(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, true); }

// {1, 4, 6, 4, 1} taps with shifts and adds only (no 32-bit multiply on the VALU).
__device__ inline uint32_t b3(uint32_t m2, uint32_t m1, uint32_t c, uint32_t p1, uint32_t p2)
{
    return ((m1 + p1 + c) << 2) + (c << 1) + m2 + p2;
}
// The same sum when the centre tap is below 2^24 (the input, A and V1s; not B): 6 * c added by one
// 24-bit multiply-add, three VALU operations instead of four (the compiler splits c * 6 + t into a
// multiply and an add, so the multiply-add is written out).
__device__ inline uint32_t b3s(uint32_t m2, uint32_t m1, uint32_t c, uint32_t p1, uint32_t p2)
{
    const uint32_t t = m2 + p2 + ((m1 + p1) << 2);
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, 6, %2" : "=v"(r) : "v"(c), "v"(t));
    return r;
}

__device__ inline uint32_t max3u(uint32_t a, uint32_t b, uint32_t c) { return max(max(a, b), c); }


// Number of set bits of m in the lanes below this one.
__device__ inline uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline void wave_fence()
{
    __builtin_amdgcn_wave_barrier();
    __asm__ __volatile__("" ::: "memory");
}

// Buffer resource over `bytes` bytes at p (raw buffer, dword 3 of gfx9). Loads and stores through
// it take a per-lane 32-bit offset and a wave-uniform scalar offset, so no 64-bit address math.
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// Five waves per SIMD, input rows loaded two ahead.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) atrous_detect_kernel(const uint16_t* __restrict__ in, uint32_t* __restrict__ P,
                                                           uint32_t* __restrict__ maskList,
                                                           uint32_t* __restrict__ maskCnt, uint32_t* __restrict__ seedList,
                                                           uint32_t* __restrict__ seedCnt, uint32_t* __restrict__ seedBits,
                                                           int W, int H, int64_t tmin, int strips, int segH,
                                                           const uint32_t* __restrict__ frameMap)
{
    __shared__ uint32_t sMask[MCAP];
    __shared__ uint32_t sSeed[SCAP];
    const int lane = threadIdx.x;
    // Workgroups go to the 8 XCDs in turn (workgroup i to XCD i mod 8), each XCD with its own L2.
    // Each XCD takes a contiguous range of tiles instead (whole frames, tiles in row-segment and
    // strip order), so that neighbouring strips and segments, which read the same halo columns and
    // rows, run side by side on one XCD and share those lines through its L2.
    const uint32_t nTiles = gridDim.x * gridDim.y;
    const uint32_t lin = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t xcd = lin & 7u, per = nTiles >> 3, rem = nTiles & 7u;
    const uint32_t tile = xcd * per + min(xcd, rem) + (lin >> 3);
    const int f = (int)(tile / gridDim.x);
    const uint32_t bx = tile % gridDim.x;
    const int strip = (int)bx % strips, seg = (int)bx / strips;
    const uint32_t npx = (uint32_t)W * (uint32_t)H;     // < 2^30 (checked by the launcher)
    const __amdgpu_buffer_rsrc_t rsI = make_rsrc(in + (size_t)(frameMap ? frameMap[f] : (uint32_t)f) * npx, npx * 2u);
    const __amdgpu_buffer_rsrc_t rsP = make_rsrc(P + (size_t)f * npx, npx * 4u);
    uint32_t* ml = maskList + (size_t)f * npx;
    uint32_t* sl = seedList + (size_t)f * npx;
    uint32_t* sbf = seedBits + (size_t)f * seed_words(npx);
    const int xs = strip * SOUT - SHALO;    // frame column of strip column 0
    const int ys = seg * segH, ye = min(ys + segH, H);
    // P < 2^32 - 1 (256 * V1s <= 2^32 - 2^16), so a saturated threshold passes nothing.
    const uint32_t tm = tmin > (int64_t)0xFFFFFFFFu ? 0xFFFFFFFFu : (uint32_t)tmin;

    // Per column k of the lane: its frame column x0 + k, the byte offset of its (clamped) input
    // column, whether it lies in the frame (columns outside never pass the threshold) and whether
    // the lane writes it. An output column is in the frame, so its P offset is 4 * (x0 + k).
    const int x0 = xs + SC * lane;
    uint32_t inOff[SC];
    bool inCol[SC], outCol[SC], pastR[SC];
    const bool outLane = lane >= SHALO / SC && lane < 64 - SHALO / SC;
#pragma unroll
    for (int k = 0; k < SC; ++k) {
        const int x = x0 + k;
        const int xc = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
        inOff[k] = (uint32_t)xc * 2u;
        inCol[k] = x >= 0 && x < W;
        outCol[k] = outLane && inCol[k];
        pastR[k] = x > W - 1;
    }
    const uint32_t outOff0 = (uint32_t)x0 * 4u;   // used only for columns in the frame
    // The frame's first column is strip column SHALO of strip 0 (lane SHALO / SC, slot 0); its last
    // column is strip column oR of the last strip.
    const bool fixL = xs < 0;
    const int oR = W - 1 - xs;
    const bool fixR = oR < SW - 1;
    const int lR = oR / SC, kR = oR % SC;

    // Rings in shifted layout between steps (slot 0 oldest); Pp = P of the last row, Mp = maximum of
    // the last two rows of P, per column.
    uint32_t rA[NA][SC], rB[NB][SC], rV[NV][SC], Pp[SC], Mp[SC];
#pragma unroll
    for (int k = 0; k < SC; ++k) {
#pragma unroll
        for (int i = 0; i < NA; ++i) rA[i][k] = 0;
#pragma unroll
        for (int i = 0; i < NV; ++i) rV[i][k] = 0;
#pragma unroll
        for (int i = 0; i < NB; ++i) rB[i][k] = 0;
        Pp[k] = 0;
        Mp[k] = 0;
    }
    uint32_t mPend = 0, sPend = 0;      // staged list entries (wave-uniform)

    auto load_row = [&](int t, uint32_t (&v)[SC]) {
        const int y = t < 0 ? 0 : (t > H - 1 ? H - 1 : t);
        const int rowOff = (int)((uint32_t)y * (uint32_t)W * 2u);
#pragma unroll
        for (int k = 0; k < SC; ++k) v[k] = __builtin_amdgcn_raw_buffer_load_b16(rsI, (int)inOff[k], rowOff, 0);
    };
    auto flush = [&](const uint32_t* stage, uint32_t n, uint32_t* cnt, uint32_t* list) {
        wave_fence();
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(cnt, n);
        base = __builtin_amdgcn_readfirstlane(base);
        for (uint32_t i = lane; i < n; i += 64) list[base + i] = stage[i];
        wave_fence();
    };
    // Dense P, mask and local maxima of output row y3: v = its P, cm = the column maxima of rows
    // y3 - 1 .. y3 + 1.
    auto emit = [&](int y3, const uint32_t (&v)[SC], const uint32_t (&cm)[SC]) {
        const uint32_t rowBase = (uint32_t)y3 * (uint32_t)W;
        const int rowOff = (int)(rowBase * 4u);
        // Pixels with P > 0 (columns outside the frame have P = 0) and the local maxima among them.
        const uint32_t cL = from_left(cm[2]), cR = from_right(cm[0]);
        const bool m0 = outLane && v[0] != 0u, m1 = outLane && v[1] != 0u, m2 = outLane && v[2] != 0u;
        const uint32_t M0 = max3u(cL, cm[0], cm[1]), M1 = max3u(cm[0], cm[1], cm[2]), M2 = max3u(cm[1], cm[2], cR);
        const bool s0 = m0 && v[0] >= M0, s1 = m1 && v[1] >= M1, s2 = m2 && v[2] >= M2;
        // The same predicates as wave masks (compare results, no round trip through a VGPR).
        const uint64_t mb0 = __builtin_amdgcn_uicmp(v[0], 0u, kIcmpNe) & kOutLanes;
        const uint64_t mb1 = __builtin_amdgcn_uicmp(v[1], 0u, kIcmpNe) & kOutLanes;
        const uint64_t mb2 = __builtin_amdgcn_uicmp(v[2], 0u, kIcmpNe) & kOutLanes;
        const uint64_t sb0 = __builtin_amdgcn_uicmp(v[0], M0, kIcmpUge) & mb0;
        const uint64_t sb1 = __builtin_amdgcn_uicmp(v[1], M1, kIcmpUge) & mb1;
        const uint64_t sb2 = __builtin_amdgcn_uicmp(v[2], M2, kIcmpUge) & mb2;
        const uint32_t mc0 = (uint32_t)__popcll(mb0), mc1 = (uint32_t)__popcll(mb1);
        const uint32_t mRow = mc0 + mc1 + (uint32_t)__popcll(mb2);
        if (mRow == 0) return;
        // P of the mask pixels only: the rest of the frame's P is zero already.
        if (m0) __builtin_amdgcn_raw_buffer_store_b32(v[0], rsP, (int)outOff0, rowOff, 0);
        if (m1) __builtin_amdgcn_raw_buffer_store_b32(v[1], rsP, (int)(outOff0 + 4u), rowOff, 0);
        if (m2) __builtin_amdgcn_raw_buffer_store_b32(v[2], rsP, (int)(outOff0 + 8u), rowOff, 0);
        const uint32_t i0 = rowBase + (uint32_t)x0, i1 = i0 + 1u, i2 = i0 + 2u;
        if (mPend + mRow > (uint32_t)MCAP) { flush(sMask, mPend, &maskCnt[f], ml); mPend = 0; }
        if (m0) sMask[mPend + lanes_below(mb0)] = i0;
        if (m1) sMask[mPend + mc0 + lanes_below(mb1)] = i1;
        if (m2) sMask[mPend + mc0 + mc1 + lanes_below(mb2)] = i2;
        mPend += mRow;

        const uint32_t sc0 = (uint32_t)__popcll(sb0), sc1 = (uint32_t)__popcll(sb1);
        const uint32_t sRow = sc0 + sc1 + (uint32_t)__popcll(sb2);
        if (sRow == 0) return;
        if (s0) atomicOr(&sbf[i0 >> 5], 1u << (i0 & 31));
        if (s1) atomicOr(&sbf[i1 >> 5], 1u << (i1 & 31));
        if (s2) atomicOr(&sbf[i2 >> 5], 1u << (i2 & 31));
        if (sPend + sRow > (uint32_t)SCAP) { flush(sSeed, sPend, &seedCnt[f], sl); sPend = 0; }
        if (s0) sSeed[sPend + lanes_below(sb0)] = i0;
        if (s1) sSeed[sPend + sc0 + lanes_below(sb1)] = i1;
        if (s2) sSeed[sPend + sc0 + sc1 + lanes_below(sb2)] = i2;
        sPend += sRow;
    };

    uint32_t pre[PF][SC];                // input rows t + 1 .. t + PF
    // Next input row: the oldest prefetched one; the row PF ahead is loaded in its place.
    auto next_row = [&](int t, uint32_t (&v0)[SC]) {
#pragma unroll
        for (int k = 0; k < SC; ++k) v0[k] = pre[0][k];
#pragma unroll
        for (int i = 0; i < PF - 1; ++i)
#pragma unroll
            for (int k = 0; k < SC; ++k) pre[i][k] = pre[i + 1][k];
        load_row(t + PF, pre[PF - 1]);      // clamped: rows past the last one are the last one
    };
    // A(t) = 16 * rowpass_0(V0): taps at columns c-2..c+2.
    auto row_a = [&](const uint32_t (&v0)[SC], uint32_t (&a)[SC]) {
        const uint32_t l1 = from_left(v0[1]), l2 = from_left(v0[2]);
        const uint32_t r0 = from_right(v0[0]), r1 = from_right(v0[1]);
        a[0] = b3s(l1, l2, v0[0], v0[1], v0[2]);
        a[1] = b3s(l2, v0[0], v0[1], v0[2], r0);
        a[2] = b3s(v0[0], v0[1], v0[2], r0, r1);
    };
    // Columns outside the frame hold the edge column's V1s (clamp of the scale-1 taps).
    auto fix_edges = [&](uint32_t (&u)[SC]) {
        if (fixL) {
            const uint32_t e = __builtin_amdgcn_readlane(u[0], SHALO / SC);
#pragma unroll
            for (int k = 0; k < SC; ++k) u[k] = lane < SHALO / SC ? e : u[k];
        }
        if (fixR) {
            const uint32_t e = __builtin_amdgcn_readlane(kR == 0 ? u[0] : (kR == 1 ? u[1] : u[2]), lR);
#pragma unroll
            for (int k = 0; k < SC; ++k) u[k] = pastR[k] ? e : u[k];
        }
    };
    // B(y1) = 16 * rowpass_1(V1s): taps at columns c-4, c-2, c, c+2, c+4.
    auto row_b = [&](const uint32_t (&u)[SC], uint32_t (&bb)[SC]) {
        const uint32_t L0 = from_left(u[0]), L1 = from_left(u[1]), L2 = from_left(u[2]);
        const uint32_t R0 = from_right(u[0]), R1 = from_right(u[1]), R2 = from_right(u[2]);
        const uint32_t LL2 = from_left(L2), RR0 = from_right(R0);
        bb[0] = b3s(LL2, L1, u[0], u[2], R1);
        bb[1] = b3s(L0, L2, u[1], R0, R2);
        bb[2] = b3s(L1, u[0], u[2], R1, RR0);
    };
    // 65536 * W1 = 256 * V1s - V2s, kept when >= tmin (u32: both terms are below 2^32); then the
    // running column maxima and the output row.
    auto finish = [&](int y3, bool rowIn, const uint32_t (&v1)[SC], const uint32_t (&v2s)[SC], bool emitRow) {
        uint32_t v[SC], cm[SC];
#pragma unroll
        for (int k = 0; k < SC; ++k) {
            const uint32_t a = v1[k] << 8;
            const uint32_t d = a - v2s[k];
            const uint32_t pr = (rowIn && inCol[k] && a >= v2s[k] && d >= tm) ? d : 0u;
            v[k] = Pp[k];
            cm[k] = max(Mp[k], pr);
            Mp[k] = max(Pp[k], pr);
            Pp[k] = pr;
        }
        if (emitRow) emit(y3, v, cm);
    };

    // One pipeline step for input row t, rings in shifted layout. Interior steps (every stage inside
    // the frame, output row inside the segment) skip the edge handling.
    auto step = [&](int t, auto interior) {
        constexpr bool kInterior = decltype(interior)::value;
        uint32_t v0[SC];
        next_row(t, v0);
        uint32_t a[SC];
        row_a(v0, a);
#pragma unroll
        for (int k = 0; k < SC; ++k) {
#pragma unroll
            for (int i = 0; i < NA - 1; ++i) rA[i][k] = rA[i + 1][k];
            rA[NA - 1][k] = a[k];
        }
        // V1s(y1) = colpass_0(A), y1 = t - 2.
        const int y1 = t - 2;
        uint32_t u[SC];
#pragma unroll
        for (int k = 0; k < SC; ++k) u[k] = b3s(rA[0][k], rA[1][k], rA[2][k], rA[3][k], rA[4][k]);
        fix_edges(u);
#pragma unroll
        for (int k = 0; k < SC; ++k) {
#pragma unroll
            for (int i = 0; i < NV - 1; ++i) rV[i][k] = rV[i + 1][k];
            rV[NV - 1][k] = u[k];
        }
        if constexpr (kInterior) {
            uint32_t bb[SC];
            row_b(u, bb);
#pragma unroll
            for (int k = 0; k < SC; ++k) {
#pragma unroll
                for (int i = 0; i < NB - 1; ++i) rB[i][k] = rB[i + 1][k];
                rB[NB - 1][k] = bb[k];
            }
        } else if (y1 >= 0) {
            uint32_t bb[SC];
            if (y1 < H) {
                row_b(u, bb);
            } else {
#pragma unroll
                for (int k = 0; k < SC; ++k) bb[k] = rB[NB - 1][k];  // below the frame: B(H-1)
            }
#pragma unroll
            for (int k = 0; k < SC; ++k) {
                if (y1 == 0) {
#pragma unroll
                    for (int i = 0; i < NB - 1; ++i) rB[i][k] = bb[k];  // above the frame: B(0)
                } else {
#pragma unroll
                    for (int i = 0; i < NB - 1; ++i) rB[i][k] = rB[i + 1][k];
                }
                rB[NB - 1][k] = bb[k];
            }
        }
        // P(y2), y2 = t - 6: V2s = colpass_1(B) from rows y2-4, y2-2, y2, y2+2, y2+4.
        const int y2 = t - 6;
        uint32_t v2s[SC], v1[SC];
#pragma unroll
        for (int k = 0; k < SC; ++k) {
            v2s[k] = b3(rB[1][k], rB[3][k], rB[5][k], rB[7][k], rB[9][k]);
            v1[k] = rV[0][k];
        }
        // Output row y3 = t - 7.
        const int y3 = t - 7;
        finish(y3, kInterior || (y2 >= 0 && y2 < H), v1, v2s, kInterior || y3 >= ys);
    };

    // Interior step j (0..UNROLL-1) of an unrolled iteration: the rings are addressed circularly (the
    // newest row of a ring of N at slot j % N, the row of age d at slot (j - d) mod N); the iteration
    // starts and ends with them in shifted layout because UNROLL is a multiple of every N.
    auto step_c = [&](int t, auto J) {
        constexpr int j = decltype(J)::value;
        uint32_t v0[SC];
        next_row(t, v0);
        uint32_t a[SC];
        row_a(v0, a);
        constexpr int sA = j % NA;
#pragma unroll
        for (int k = 0; k < SC; ++k) rA[sA][k] = a[k];
        uint32_t u[SC];
#pragma unroll
        for (int k = 0; k < SC; ++k)
            u[k] = b3s(rA[(j + 1) % NA][k], rA[(j + 2) % NA][k], rA[(j + 3) % NA][k], rA[(j + 4) % NA][k], rA[sA][k]);
        fix_edges(u);
        uint32_t v1[SC];
#pragma unroll
        for (int k = 0; k < SC; ++k) {
            v1[k] = rV[(j + 1) % NV][k];          // V1s(t - 6), age 4
            rV[j % NV][k] = u[k];
        }
        uint32_t bb[SC];
        row_b(u, bb);
        constexpr int sB = j % NB;
#pragma unroll
        for (int k = 0; k < SC; ++k) rB[sB][k] = bb[k];
        uint32_t v2s[SC];
#pragma unroll
        for (int k = 0; k < SC; ++k)
            v2s[k] = b3(rB[(j + 2) % NB][k], rB[(j + 4) % NB][k], rB[(j + 6) % NB][k], rB[(j + 8) % NB][k], rB[sB][k]);
        finish(t - 7, true, v1, v2s, true);
    };

    const int t0 = ys - 7, t1 = ye + 6;
    // Interior steps: output row inside the segment (t >= ys + 7) and V1s row inside the frame
    // (t - 2 <= H - 1); the P row t - 6 is then inside the frame too.
    const int tI0 = ys + 7, tI1 = min(t1, H + 1);
#pragma unroll
    for (int i = 0; i < PF; ++i) load_row(t0 + i, pre[i]);
    int t = t0;
    for (; t < tI0; ++t) step(t, std::false_type{});
    for (; t + UNROLL - 1 <= tI1; t += UNROLL) {
        step_c(t + 0, std::integral_constant<int, 0>{});
        step_c(t + 1, std::integral_constant<int, 1>{});
        step_c(t + 2, std::integral_constant<int, 2>{});
        step_c(t + 3, std::integral_constant<int, 3>{});
        step_c(t + 4, std::integral_constant<int, 4>{});
        step_c(t + 5, std::integral_constant<int, 5>{});
        step_c(t + 6, std::integral_constant<int, 6>{});
        step_c(t + 7, std::integral_constant<int, 7>{});
        step_c(t + 8, std::integral_constant<int, 8>{});
        step_c(t + 9, std::integral_constant<int, 9>{});
    }
    for (; t <= tI1; ++t) step(t, std::true_type{});
    for (; t <= t1; ++t) step(t, std::false_type{});

    if (mPend) flush(sMask, mPend, &maskCnt[f], ml);
    if (sPend) flush(sSeed, sPend, &seedCnt[f], sl);
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
                          hipStream_t s, const uint32_t* frameMap, int segRows)
{
    if (frames <= 0 || W <= 0 || H <= 0) return;
    if (frames > 65535) throw Error("launch_atrous_detect: at most 65535 frames per launch");
    if ((size_t)W * H >= (size_t(1) << 30)) throw Error("launch_atrous_detect: frames of 2^30 pixels or more are not supported");
    const int strips = (int)ceil_div(W, SOUT);
    // Row segments: long enough to amortise the 13-row pipeline warm-up, short enough to give the
    // chip a few waves per SIMD when the chunk is small.
    int segH = 128;
    while (segH > 32 && (size_t)frames * strips * ceil_div(H, segH) < 8192) segH /= 2;
    if (segRows > 0) segH = std::min(segRows, 1 << 20);
    const int segs = (int)ceil_div(H, segH);
    if ((size_t)strips * segs > 0x7FFFFFFF) throw Error("launch_atrous_detect: frame too large");
    const dim3 grid((unsigned)(strips * segs), (unsigned)frames);
    atrous_detect_kernel<<<grid, 64, 0, s>>>(d_in, b.P, b.maskList, b.maskCnt, b.seedList, b.seedCnt, b.seedBits,
                                             W, H, tmin, strips, segH, frameMap);
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
