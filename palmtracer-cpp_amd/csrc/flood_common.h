// Device helpers shared by the watershed kernels (component floods, exact frame floods).
// Reference: lib/localization/Segmentation2D.cpp:336-385 (watershed), lib/Structs.hpp:26 (Pixel).
#pragma once
#include "pt_common.h"
#include "pt_kernels.h"

namespace pt {

// Local maximum of getLocalMax (Segmentation2D.cpp:294-333: P > 0 and no 8-neighbour inside the
// frame strictly greater), from the bitmap the detection kernel wrote.
__device__ inline bool seed_bit(const uint32_t* sbf, uint32_t idx) { return (sbf[idx >> 5] >> (idx & 31)) & 1u; }

// A pixel that the watershed would seed: labelled with some 8-neighbour (inside the frame) that is
// unlabelled (Segmentation2D.cpp:340-359). Unlabelled = below threshold or label 0.
__device__ inline bool has_unlabelled_nbr(const uint32_t* Pf, const uint32_t* finf, int W, int H, int idx)
{
    const int y = idx / W, x = idx - y * W;
    for (int dy = -1; dy <= 1; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= H) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int xx = x + dx;
            if ((dx | dy) == 0 || xx < 0 || xx >= W) continue;
            const int q = yy * W + xx;
            if (Pf[q] == 0 || finf[q] == 0) return true;
        }
    }
    return false;
}

// Neighbour j (0..7) of a pixel in the reference's scan order (rows dy, then columns dx).
__device__ inline void nb_offset(int j, int& dy, int& dx)
{
    const int k = j < 4 ? j : j + 1;  // skip the centre
    dy = k / 3 - 1;
    dx = k % 3 - 1;
}

// Reference key of a push at pop count `it`: -output + it * EPSILON, rounded per operation
// (no fused multiply-add), Segmentation2D.cpp:378.
__device__ inline double age_key(uint32_t p, uint32_t it)
{
    const double o = (double)p * kPScale;
    return it == 0 ? -o : __dadd_rn(-o, __dmul_rn((double)it, kEpsilon));
}

// Seed of the exact frame flood of mask pixel p of frame f: a local maximum (first watershed), or
// the maximum pixel of a region that survived the basin fusion (second watershed, fz set; its
// label is still in fn at p).
__device__ inline bool flood_seed(const uint32_t* seedBits, const uint32_t* fn, const Fz* fz, const PixAux* aux,
                                  size_t regionCap, size_t npx, uint32_t f, uint32_t p)
{
    if (!fz) return seed_bit(seedBits + (size_t)f * seed_words(npx), p);
    const uint32_t L = fn[p];
    if (L == 0 || L > npx) return false;
    const uint32_t sl = aux[(size_t)f * npx + (L - 1)].rslot;
    return sl < regionCap && fz[(size_t)f * regionCap + sl].alive &&
           0xFFFFFFFFu - (uint32_t)fz[(size_t)f * regionCap + sl].maxKey == p;
}

// Workgroups go to the 8 XCDs in turn (workgroup i to XCD i mod 8), each XCD with its own L2. The
// two maps below keep a frame's workgroups on one XCD, so that their scattered accesses into the
// frame's arrays share one L2; both are one-to-one maps of the grid.
//
// Grid-stride kernels ((x, y) of (nx, ny): x strides over a frame's items, y over frames): with at
// least 8 frames and a multiple of 8 frame slots, frame slot y goes to the workgroups of XCD y mod 8;
// with fewer frames the hardware order spreads each frame over every XCD.
struct FrameGrid {
    uint32_t x, y, nx, ny;
};
__device__ inline FrameGrid frame_grid(uint32_t frames)
{
    FrameGrid g{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y};
    if (frames >= 8u && (gridDim.y & 7u) == 0u) {
        const uint32_t lin = blockIdx.y * gridDim.x + blockIdx.x;   // dispatch order
        const uint32_t turn = lin >> 3;                              // this workgroup's turn on its XCD
        g.x = turn % gridDim.x;
        g.y = (lin & 7u) + 8u * (turn / gridDim.x);
    }
    return g;
}

// Grids of one workgroup per (x, frame) with no striding: each XCD takes a contiguous range of the
// (frame, x) pairs in frame order.
__device__ inline void frame_tile(uint32_t& x, uint32_t& frame)
{
    const uint32_t n = gridDim.x * gridDim.y;
    const uint32_t lin = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t xcd = lin & 7u, per = n >> 3, rem = n & 7u;
    const uint32_t t = xcd * per + min(xcd, rem) + (lin >> 3);
    frame = t / gridDim.x;
    x = t % gridDim.x;
}

} // namespace pt
