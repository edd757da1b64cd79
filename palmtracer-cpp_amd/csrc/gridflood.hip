// Grid rank flood: the exact frame flood of the flagged frames of a chunk, each frame spread over
// many workgroups (Segmentation2D.cpp:336-385).
//
// The reference floods a frame from one priority queue whose keys age with the global pop count:
// key = -W1 + T * 1e-7, T the pop count when the entry was pushed. Number the pops 1, 2, ... and let
// G(x) be the pop number of entry x. A seed with an unlabelled neighbour is pushed at T = 0; any
// other mask pixel is pushed by its first popped neighbour, T(x) = min over its flooded neighbours
// of G. If the keys are pairwise distinct and every entry is pushed before the pop at its key rank
// (T(x) < G(x)), the pops are the keys in increasing order: G(x) = 1 + #{y : K(y) < K(x)}. So the
// flood is the fixed point of T = min-neighbour(G), G = rank(K(T)), iterated from G = rank of the
// potential alone. A key moves by at most n * 1e-7 (n entries), so the rank of x only involves the
// entries whose potential lies within that distance ("the window"): rank = (entries above the
// window) + (window entries with a smaller key). The labels follow the first-popped-neighbour links.
//
// Every phase is a kernel over all entries of all flagged frames (grid-stride over frames in y and
// entries in x), so a 2048^2 frame runs on dozens of workgroups:
//   setup     per-frame work space from one pool (per-entry arrays, 84 bytes per mask pixel);
//   reset     labels back to the seeds;           collect  entries with their sort keys;
//   sort      2048-key tiles in LDS, then merge passes (each element finds its place by one
//             binary search in the sibling run);
//   finalize  entries in potential order and their windows; the label map holds each entry's
//             index (tagged) until the write, the seeds' labels are kept per entry; then a table of
//             each entry's neighbour entries (read coalesced by the later steps) and the
//             fixed-point steps;
//   checks    T < G, equal keys (below), labels by walking the first-popped-neighbour links;
//   write     labels of the frames that passed; the others get their seed labels back and go to
//             the single-workgroup floods.
// Frames that do not fit the pool go to the single-workgroup floods.
//
// Equal keys. The heap pops equal keys in an order set by its internal layout; the solution above
// breaks them by entry order. A frame with equal keys is accepted when no order can change a label,
// shown with intervals that hold for every order:
//   start   an entry in a group of k equal keys has a rank in [first, first + k - 1] of the group,
//           every other entry its rank;
//   step    push ages in [min over neighbours of the lowest rank, min of the highest rank], keys
//           from those (the key grows with the age), ranks in [1 + entries whose highest key is
//           below the lowest key, n - entries whose lowest key is above the highest key];
//   repeat  until the intervals stop growing (at most kGridIvalIter steps, else the frame is
//           handed over);
//   accept  every non-seed is pushed before its earliest pop (highest push age < lowest rank), and
//           every neighbour that can be the first popped one (lowest rank <= highest push age)
//           carries the label of the solution.
// By induction over the pops of the flood under any order of the equal keys: the entry popped at
// step m has m inside its interval and every entry whose highest rank is below m has been popped
// (an entry with a smaller certain key that is not present at m would need an unpushed entry with
// a rank interval starting above m, and following certainly-smaller keys down ends at a present
// one); so the first popped neighbour of every entry is one of the accepted candidates, and the
// labels are those of the solution.
#include "flood_common.h"

#include <algorithm>

namespace pt {
namespace {

constexpr int GF_NT = 256;
constexpr int GF_TILE = 2048;                 // keys per tile (8 per thread)
constexpr uint32_t GF_NONE = 0xFFFFFFFFu;
constexpr uint32_t GF_SEED = 0x80000000u;     // seed bit of spix
constexpr uint32_t GF_TAG = 0x80000000u;      // label map: entry index (labels are below 2^31)

enum : uint32_t { kGfPool = 1u, kGfUnsettled = 2u, kGfOrder = 4u, kGfTie = 8u, kGfLabel = 16u };

// Per flagged frame (kGridMetaWords words).
struct GfMeta {
    uint32_t n;        // entries
    uint32_t fail;     // kGf* bits
    uint32_t cap;      // entry capacity (mask pixels)
    uint32_t tie;      // the solution has equal keys
    uint32_t offLo, offHi;  // pool offset (words)
    uint32_t changed[kGridMaxIter];
    uint32_t ichanged[kGridIvalIter];   // interval steps that widened some interval
};
static_assert(sizeof(GfMeta) <= kGridMetaWords * 4, "grid flood meta");

// Entry capacity rounded to 4, so that every array (and each entry's 16-byte neighbour rows) stays
// 16-byte aligned.
__host__ __device__ inline size_t gf_round(size_t v) { return (v + 3) & ~size_t(3); }
// Pool words of a frame of cap mask pixels: 2 u64 arrays, 9 u32 arrays and the 8-word neighbour
// table of cap entries.
constexpr size_t kGfWordsPerEntry = 21;
__host__ __device__ inline size_t gf_words(size_t cap) { return kGfWordsPerEntry * gf_round(cap); }

struct GfView {
    uint64_t* skA;     // sort keys (ping); after the sort: group size / first rank per entry
    uint64_t* skB;     // sort keys (pong)
    double* K;         // keys (the pong buffer, once sorted)
    uint32_t *sp, *spix, *lo, *hi, *Ga, *Gb, *T, *par, *lbl;
    uint32_t* nbr;     // [entry][8] neighbour entries (none: GF_NONE)
};

__device__ inline GfView gf_view(uint32_t* pool, const GfMeta& m)
{
    const size_t off = ((size_t)m.offHi << 32) | m.offLo, c = gf_round(m.cap);
    GfView v;
    v.skA = reinterpret_cast<uint64_t*>(pool + off);
    v.skB = v.skA + c;
    v.K = reinterpret_cast<double*>(v.skB);
    uint32_t* w = reinterpret_cast<uint32_t*>(v.skB + c);
    v.sp = w;
    v.spix = w + c;
    v.lo = w + 2 * c;
    v.hi = w + 3 * c;
    v.Ga = w + 4 * c;
    v.Gb = w + 5 * c;
    v.T = w + 6 * c;
    v.par = w + 7 * c;
    v.lbl = w + 8 * c;
    v.nbr = w + 9 * c;
    return v;
}

// Entry of neighbour j of pixel (y, x), or none (outside the frame or the mask, or a seed that is
// never pushed).
__device__ inline uint32_t gf_nbr(const uint32_t* Pf, const uint32_t* fn, int W, int H, int y, int x, int j)
{
    int dy, dx;
    nb_offset(j, dy, dx);
    const int yy = y + dy, xx = x + dx;
    if (yy < 0 || yy >= H || xx < 0 || xx >= W) return GF_NONE;
    const uint32_t q = (uint32_t)(yy * W + xx);
    if (Pf[q] == 0u) return GF_NONE;
    const uint32_t t = fn[q];
    return (t & GF_TAG) ? (t & ~GF_TAG) : GF_NONE;
}

// Merge passes a frame of n entries needs after the tile sort.
__device__ inline int gf_passes(uint32_t n)
{
    int p = 0;
    for (uint64_t w = GF_TILE; w < n; w <<= 1) ++p;
    return p;
}

// Window distance in potential quanta: keys of entries further apart never cross (ages <= n).
__device__ inline uint64_t gf_window(uint32_t n) { return (uint64_t)floor(((double)n * kEpsilon + 1e-9) * 65536.0) + 1; }

struct GfArgs {
    const uint32_t* P;
    const uint32_t* seedBits;
    uint32_t* fin;
    const uint32_t* maskList;
    const uint32_t* maskCnt;
    const uint32_t* flagList;   // frames of this round (count in counters[countSlot])
    int countSlot;
    int round;                  // 0: frames that do not fit the pool go to retryList for round 1
    uint32_t* counters;
    uint32_t* gridList;
    uint32_t* retryList;
    GfMeta* meta;
    uint32_t* pool;
    size_t poolWords;
    const Fz* fz;
    const PixAux* aux;
    size_t regionCap;
    size_t npx;
    int W, H, F;
};

__device__ inline uint32_t gf_count(const GfArgs& a) { return min(a.counters[a.countSlot], (uint32_t)a.F); }

// Wave-aggregated atomic OR of a per-lane flag into *dst.
__device__ inline void gf_or(uint32_t* dst, uint32_t bits)
{
    const uint64_t m = __ballot(bits != 0u);
    if (m == 0) return;
    uint32_t v = bits;
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
    if ((threadIdx.x & 63) == (uint32_t)__builtin_ctzll(m)) atomicOr(dst, v);
}

// Work space of every flagged frame from the pool (exclusive scan of the sizes in list order).
__global__ void __launch_bounds__(1024) gf_setup_kernel(GfArgs a)
{
    __shared__ unsigned long long part[1024];
    const uint32_t cnt = gf_count(a);
    const uint32_t per = (cnt + 1023) / 1024, k0 = min(cnt, threadIdx.x * per), k1 = min(cnt, k0 + per);
    unsigned long long s = 0;
    for (uint32_t k = k0; k < k1; ++k) s += gf_words(a.maskCnt[a.flagList[k]]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const unsigned long long v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0ull;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long off = part[threadIdx.x] - s;
    for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t cap = a.maskCnt[a.flagList[k]];
        const size_t words = gf_words(cap);
        GfMeta m{};
        m.cap = cap;
        m.offLo = (uint32_t)off;
        m.offHi = (uint32_t)(off >> 32);
        m.fail = off + words > a.poolWords ? kGfPool : 0u;
        a.meta[k] = m;
        off += words;
    }
}

// Labels back to the seed state.
__global__ void __launch_bounds__(GF_NT) gf_reset_kernel(GfArgs a)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        if (a.meta[k].fail) continue;
        const uint32_t f = a.flagList[k];
        const uint32_t* ml = a.maskList + (size_t)f * a.npx;
        uint32_t* fn = a.fin + (size_t)f * a.npx;
        const uint32_t mc = a.maskCnt[f];
        for (uint32_t i = gg.x * GF_NT + threadIdx.x; i < mc; i += gg.nx * GF_NT) {
            const uint32_t p = ml[i];
            if (!flood_seed(a.seedBits, fn, a.fz, a.aux, a.regionCap, a.npx, f, p)) fn[p] = 0u;
        }
    }
}

// Flood entries (non-seeds, and seeds next to an unlabelled pixel) with their sort keys
// (potential descending, then raster order).
__global__ void __launch_bounds__(GF_NT) gf_collect_kernel(GfArgs a)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    const int lane = threadIdx.x & 63;
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (m.fail) continue;
        const uint32_t f = a.flagList[k];
        const GfView v = gf_view(a.pool, m);
        const uint32_t* ml = a.maskList + (size_t)f * a.npx;
        const uint32_t* Pf = a.P + (size_t)f * a.npx;
        const uint32_t* fn = a.fin + (size_t)f * a.npx;
        const uint32_t mc = a.maskCnt[f];
        const uint32_t stride = gg.nx * GF_NT;
        for (uint32_t b = gg.x * GF_NT + (threadIdx.x & ~63u); b < mc; b += stride) {
            const uint32_t i = b + lane;
            bool entry = false;
            uint32_t p = 0;
            if (i < mc) {
                p = ml[i];
                const bool seed = flood_seed(a.seedBits, fn, a.fz, a.aux, a.regionCap, a.npx, f, p);
                entry = !seed || has_unlabelled_nbr(Pf, fn, a.W, a.H, (int)p);
            }
            const uint64_t bal = __ballot(entry);
            if (bal == 0) continue;
            const int leader = __builtin_ctzll(bal);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&a.meta[k].n, (uint32_t)__popcll(bal));
            base = __shfl(base, leader);
            if (entry) v.skA[base + __popcll(bal & ((1ull << lane) - 1ull))] = ((uint64_t)(~Pf[p]) << 32) | p;
        }
    }
}

// Sorts each GF_TILE-key tile of every frame in LDS (bitonic network).
__global__ void __launch_bounds__(GF_NT) gf_tile_sort_kernel(GfArgs a)
{
    __shared__ uint64_t s[GF_TILE];
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (m.fail || m.n == 0) continue;
        const GfView v = gf_view(a.pool, m);
        const uint32_t tiles = (m.n + GF_TILE - 1) / GF_TILE;
        for (uint32_t t = gg.x; t < tiles; t += gg.nx) {
            const uint32_t t0 = t * GF_TILE, len = min((uint32_t)GF_TILE, m.n - t0);
            for (uint32_t i = threadIdx.x; i < GF_TILE; i += GF_NT) s[i] = i < len ? v.skA[t0 + i] : ~0ull;
            __syncthreads();
            for (uint32_t kk = 2; kk <= GF_TILE; kk <<= 1) {
                for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
                    for (uint32_t i = threadIdx.x; i < GF_TILE; i += GF_NT) {
                        const uint32_t ixj = i ^ j;
                        if (ixj > i) {
                            const uint64_t x = s[i], y = s[ixj];
                            if ((x > y) == ((i & kk) == 0)) { s[i] = y; s[ixj] = x; }
                        }
                    }
                    __syncthreads();
                }
            }
            for (uint32_t i = threadIdx.x; i < len; i += GF_NT) v.skA[t0 + i] = s[i];
            __syncthreads();
        }
    }
}

// Number of keys in sorted src[b, e) below key.
__device__ inline uint32_t gf_lower(const uint64_t* src, uint32_t b, uint32_t e, uint64_t key)
{
    while (b < e) {
        const uint32_t mid = (b + e) >> 1;
        if (src[mid] < key) b = mid + 1; else e = mid;
    }
    return b;
}

// Merge pass `pass`: runs of GF_TILE << pass keys merged in pairs (ping-pong between skA and skB).
// Keys are distinct (the pixel is part of the key), so each key's place is its index in its own
// run plus the number of smaller keys in the sibling run.
__global__ void __launch_bounds__(GF_NT) gf_merge_kernel(GfArgs a, int pass)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    const uint64_t w = (uint64_t)GF_TILE << pass;
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (m.fail || m.n <= w) continue;
        const GfView v = gf_view(a.pool, m);
        const uint64_t* src = (pass & 1) ? v.skB : v.skA;
        uint64_t* dst = (pass & 1) ? v.skA : v.skB;
        const uint32_t n = m.n;
        for (uint32_t i = gg.x * GF_NT + threadIdx.x; i < n; i += gg.nx * GF_NT) {
            const uint64_t key = src[i];
            const uint64_t r = i / w, base = r * w;
            uint64_t out;
            if ((r & 1) == 0) {
                const uint64_t s0 = base + w;
                const uint64_t s1 = s0 + w < n ? s0 + w : n;
                out = s0 >= n ? i : i + (gf_lower(src, (uint32_t)s0, (uint32_t)s1, key) - s0);
            } else {
                const uint64_t s0 = base - w;
                out = s0 + (i - base) + (gf_lower(src, (uint32_t)s0, (uint32_t)base, key) - s0);
            }
            dst[out] = key;
        }
    }
}

// Entries in potential order: potentials, pixels (seed bit), seed labels, windows, first ranks; the
// label map of every entry's pixel holds its index (tagged).
__global__ void __launch_bounds__(GF_NT) gf_finalize_kernel(GfArgs a)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (m.fail || m.n == 0) continue;
        const GfView v = gf_view(a.pool, m);
        const uint32_t f = a.flagList[k];
        uint32_t* fn = a.fin + (size_t)f * a.npx;
        const uint32_t n = m.n;
        const uint64_t* S = (gf_passes(n) & 1) ? v.skB : v.skA;
        const uint64_t D = gf_window(n);
        for (uint32_t i = gg.x * GF_NT + threadIdx.x; i < n; i += gg.nx * GF_NT) {
            const uint64_t key = S[i];
            const uint32_t P = ~(uint32_t)(key >> 32), p = (uint32_t)key;
            const bool seed = flood_seed(a.seedBits, fn, a.fz, a.aux, a.regionCap, a.npx, f, p);
            v.sp[i] = P;
            v.spix[i] = p | (seed ? GF_SEED : 0u);
            v.lbl[i] = seed ? fn[p] : 0u;
            // lo: first entry with potential <= P + D; hi: first entry with potential < P - D.
            uint32_t b = 0, e = n;
            while (b < e) {
                const uint32_t mid = (b + e) >> 1;
                if ((uint64_t)(~(uint32_t)(S[mid] >> 32)) > (uint64_t)P + D) b = mid + 1; else e = mid;
            }
            v.lo[i] = b;
            b = 0; e = n;
            while (b < e) {
                const uint32_t mid = (b + e) >> 1;
                if ((uint64_t)(~(uint32_t)(S[mid] >> 32)) + D >= (uint64_t)P) b = mid + 1; else e = mid;
            }
            v.hi[i] = b;
            fn[p] = i | GF_TAG;
            v.Ga[i] = i + 1;
        }
    }
}

// Neighbour table (entries among the 8 neighbours of each entry's pixel), built from the mask list.
__global__ void __launch_bounds__(GF_NT) gf_neighbours_kernel(GfArgs a)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (m.fail || m.n == 0) continue;
        const GfView v = gf_view(a.pool, m);
        const uint32_t f = a.flagList[k];
        const uint32_t* Pf = a.P + (size_t)f * a.npx;
        const uint32_t* fn = a.fin + (size_t)f * a.npx;
        const uint32_t* ml = a.maskList + (size_t)f * a.npx;
        const uint32_t mc = a.maskCnt[f];
        // Pixels in mask-list order (neighbouring pixels of one detection strip are adjacent in the
        // list), so that the label-map reads of neighbouring threads share cache lines; each entry's
        // row of the table is one 32-byte write.
        for (uint32_t q = gg.x * GF_NT + threadIdx.x; q < mc; q += gg.nx * GF_NT) {
            const uint32_t p = ml[q];
            const uint32_t t = fn[p];
            if (!(t & GF_TAG)) continue;   // not an entry (a seed with no unlabelled neighbour)
            const uint32_t i = t & ~GF_TAG;
            const int y = (int)(p / (uint32_t)a.W), x = (int)(p % (uint32_t)a.W);
            uint32_t e[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) e[j] = gf_nbr(Pf, fn, a.W, a.H, y, x, j);
            uint4* dst = reinterpret_cast<uint4*>(v.nbr + (size_t)i * 8);
            dst[0] = make_uint4(e[0], e[1], e[2], e[3]);
            dst[1] = make_uint4(e[4], e[5], e[6], e[7]);
        }
    }
}

// Neighbour entries of entry i from the table.
__device__ inline void gf_nbrs(const GfView& v, uint32_t i, uint32_t e[8])
{
    const uint4* src = reinterpret_cast<const uint4*>(v.nbr + (size_t)i * 8);
    const uint4 a = src[0], b = src[1];
    e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
    e[4] = b.x; e[5] = b.y; e[6] = b.z; e[7] = b.w;
}

// a[e[j]] for the 8 neighbour entries, every load issued before any is used (no per-neighbour
// branch, which would wait for each load in turn); a neighbour without an entry reads entry 0 (every
// frame that runs these steps has one) and gets GF_NONE.
__device__ inline void gf_gather8(const uint32_t* a, const uint32_t e[8], uint32_t out[8])
{
    uint32_t raw[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) raw[j] = a[e[j] == GF_NONE ? 0u : e[j]];
#pragma unroll
    for (int j = 0; j < 8; ++j) out[j] = e[j] == GF_NONE ? GF_NONE : raw[j];
}

// Frames still iterating at step `it` (the previous step changed some rank).
__device__ inline bool gf_active(const GfMeta& m, int it) { return !m.fail && m.n > 0 && (it == 0 || m.changed[it - 1]); }

// Fixed-point step, part 1: push ages and keys from the ranks of step `it`.
__global__ void __launch_bounds__(GF_NT) gf_age_kernel(GfArgs a, int it)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (!gf_active(m, it)) continue;
        const GfView v = gf_view(a.pool, m);
        const uint32_t* G = (it & 1) ? v.Gb : v.Ga;
        for (uint32_t i = gg.x * GF_NT + threadIdx.x; i < m.n; i += gg.nx * GF_NT) {
            uint32_t t = 0;
            if (!(v.spix[i] & GF_SEED)) {
                uint32_t e[8], g[8];
                gf_nbrs(v, i, e);
                gf_gather8(G, e, g);
                t = GF_NONE;
#pragma unroll
                for (int j = 0; j < 8; ++j) t = min(t, g[j]);
            }
            v.T[i] = t;
            v.K[i] = t == GF_NONE ? 1.79769313486231570815e308 : age_key(v.sp[i], t);
        }
    }
}

// Fixed-point step, part 2: ranks of the keys (ties in entry order) into the other rank buffer.
__global__ void __launch_bounds__(GF_NT) gf_rank_kernel(GfArgs a, int it)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (!gf_active(m, it)) continue;
        const GfView v = gf_view(a.pool, m);
        const uint32_t* G = (it & 1) ? v.Gb : v.Ga;
        uint32_t* Gn = (it & 1) ? v.Ga : v.Gb;
        const uint32_t stride = gg.nx * GF_NT;
        for (uint32_t b = gg.x * GF_NT + (threadIdx.x & ~63u); b < m.n; b += stride) {
            const uint32_t i = b + (threadIdx.x & 63);
            bool changed = false;
            if (i < m.n) {
                const double ki = v.K[i];
                const uint32_t lo = v.lo[i], hi = v.hi[i];
                uint32_t c = 0;
                for (uint32_t j = lo; j < hi; ++j) {
                    const double kj = v.K[j];
                    c += (j != i) && (kj < ki || (kj == ki && j < i));
                }
                const uint32_t g = 1 + lo + c;
                Gn[i] = g;
                changed = g != G[i];
            }
            gf_or(&a.meta[k].changed[it], changed ? 1u : 0u);
        }
    }
}

// Group of equal keys of entry i: size and first rank, packed in skA.
__device__ inline uint32_t gf_gsize(const GfView& v, uint32_t i) { return (uint32_t)v.skA[i]; }
__device__ inline uint32_t gf_gfirst(const GfView& v, uint32_t i) { return (uint32_t)(v.skA[i] >> 32); }
// Checks 1: the iteration settled, T < G, first-popped neighbours, groups of equal keys.
__global__ void __launch_bounds__(GF_NT) gf_check1_kernel(GfArgs a)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (m.fail || m.n == 0) continue;
        if (m.changed[kGridMaxIter - 1]) {
            if (gg.x == 0 && threadIdx.x == 0) atomicOr(&a.meta[k].fail, kGfUnsettled);
            continue;
        }
        const GfView v = gf_view(a.pool, m);
        const uint32_t* G = v.Ga;     // settled: both rank buffers hold the solution
        const uint32_t stride = gg.nx * GF_NT;
        for (uint32_t b = gg.x * GF_NT + (threadIdx.x & ~63u); b < m.n; b += stride) {
            const uint32_t i = b + (threadIdx.x & 63);
            uint32_t bad = 0, tie = 0;
            if (i < m.n) {
                const uint32_t sx = v.spix[i];
                const bool seed = sx & GF_SEED;
                const uint32_t g = G[i];
                uint32_t pa = i;
                if (!seed) {
                    if (!(v.T[i] < g)) bad = kGfOrder;
                    uint32_t bg = GF_NONE;
                    pa = GF_NONE;
                    uint32_t e[8], g[8];
                    gf_nbrs(v, i, e);
                    gf_gather8(G, e, g);
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        if (g[j] < bg) { bg = g[j]; pa = e[j]; }
                    if (pa == GF_NONE) bad |= kGfOrder;
                }
                v.par[i] = pa;
                const double ki = v.K[i];
                uint32_t eq = 0, before = 0;
                for (uint32_t j = v.lo[i]; j < v.hi[i]; ++j)
                    if (j != i && v.K[j] == ki) {
                        ++eq;
                        before += G[j] < g;
                    }
                v.skA[i] = ((uint64_t)(g - before) << 32) | (eq + 1);
                tie = eq ? 1u : 0u;
            }
            gf_or(&a.meta[k].fail, bad);
            gf_or(&a.meta[k].tie, tie);
        }
    }
}

// Checks 2: labels through the first-popped-neighbour links (each link goes to a smaller rank, so
// every walk ends at a seed).
__global__ void __launch_bounds__(GF_NT) gf_check2_kernel(GfArgs a)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (m.fail || m.n == 0) continue;
        const GfView v = gf_view(a.pool, m);
        const uint32_t stride = gg.nx * GF_NT;
        for (uint32_t b = gg.x * GF_NT + (threadIdx.x & ~63u); b < m.n; b += stride) {
            const uint32_t i = b + (threadIdx.x & 63);
            uint32_t bad = 0;
            if (i < m.n) {
                const uint32_t sx = v.spix[i];
                uint32_t L;
                if (sx & GF_SEED) {
                    L = v.lbl[i];
                } else {
                    uint32_t e = v.par[i];
                    for (uint32_t steps = 0; steps < m.n && !(v.spix[e] & GF_SEED); ++steps) e = v.par[e];
                    L = (v.spix[e] & GF_SEED) ? v.lbl[e] : 0u;
                    v.lbl[i] = L;
                }
                if (L == 0u) bad |= kGfLabel;
            }
            gf_or(&a.meta[k].fail, bad);
        }
    }
}

// Intervals (frames with equal keys). Rank intervals in Gb (lowest) and T (highest), key intervals
// in K (lowest) and skA (highest, as doubles).
__device__ inline double* gf_khi(const GfView& v) { return reinterpret_cast<double*>(v.skA); }

__global__ void __launch_bounds__(GF_NT) gf_ival_init_kernel(GfArgs a)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (m.fail || m.n == 0 || !m.tie) continue;
        const GfView v = gf_view(a.pool, m);
        for (uint32_t i = gg.x * GF_NT + threadIdx.x; i < m.n; i += gg.nx * GF_NT) {
            const uint32_t first = gf_gfirst(v, i), size = gf_gsize(v, i);
            v.Gb[i] = first;
            v.T[i] = first + size - 1;
        }
    }
}

__device__ inline bool gf_ival_active(const GfMeta& m, int it)
{
    return !m.fail && m.n > 0 && m.tie && (it == 0 || m.ichanged[it - 1]);
}

// Interval step, part 1: push-age and key intervals from the neighbours' rank intervals.
__global__ void __launch_bounds__(GF_NT) gf_ival_key_kernel(GfArgs a, int it)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (!gf_ival_active(m, it)) continue;
        const GfView v = gf_view(a.pool, m);
        double* khi = gf_khi(v);
        for (uint32_t i = gg.x * GF_NT + threadIdx.x; i < m.n; i += gg.nx * GF_NT) {
            uint32_t tlo = 0, thi = 0;
            if (!(v.spix[i] & GF_SEED)) {
                uint32_t e[8], glo[8], ghi[8];
                gf_nbrs(v, i, e);
                gf_gather8(v.Gb, e, glo);
                gf_gather8(v.T, e, ghi);
                tlo = thi = GF_NONE;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    tlo = min(tlo, glo[j]);
                    thi = min(thi, ghi[j]);
                }
            }
            v.K[i] = tlo == GF_NONE ? 1.79769313486231570815e308 : age_key(v.sp[i], tlo);
            khi[i] = thi == GF_NONE ? 1.79769313486231570815e308 : age_key(v.sp[i], thi);
        }
    }
}

// Interval step, part 2: rank intervals from the key intervals of the window (only ever widened).
__global__ void __launch_bounds__(GF_NT) gf_ival_rank_kernel(GfArgs a, int it)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (!gf_ival_active(m, it)) continue;
        const GfView v = gf_view(a.pool, m);
        const double* khi = gf_khi(v);
        const uint32_t stride = gg.nx * GF_NT;
        for (uint32_t b = gg.x * GF_NT + (threadIdx.x & ~63u); b < m.n; b += stride) {
            const uint32_t i = b + (threadIdx.x & 63);
            bool changed = false;
            if (i < m.n) {
                const double klo = v.K[i], kh = khi[i];
                const uint32_t lo = v.lo[i], hi = v.hi[i];
                uint32_t before = 0, after = 0;
                for (uint32_t j = lo; j < hi; ++j) {
                    if (j == i) continue;
                    before += khi[j] < klo;
                    after += v.K[j] > kh;
                }
                const uint32_t glo = min(v.Gb[i], 1 + lo + before), ghi = max(v.T[i], hi - after);
                changed = glo != v.Gb[i] || ghi != v.T[i];
                v.Gb[i] = glo;
                v.T[i] = ghi;
            }
            gf_or(&a.meta[k].ichanged[it], changed ? 1u : 0u);
        }
    }
}

// Acceptance: intervals settled; every non-seed pushed before its earliest pop; every possible
// first-popped neighbour carries the label of the solution.
__global__ void __launch_bounds__(GF_NT) gf_ival_check_kernel(GfArgs a)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        if (m.fail || m.n == 0 || !m.tie) continue;
        if (m.ichanged[kGridIvalIter - 1]) {
            if (gg.x == 0 && threadIdx.x == 0) atomicOr(&a.meta[k].fail, kGfTie);
            continue;
        }
        const GfView v = gf_view(a.pool, m);
        const uint32_t stride = gg.nx * GF_NT;
        for (uint32_t b = gg.x * GF_NT + (threadIdx.x & ~63u); b < m.n; b += stride) {
            const uint32_t i = b + (threadIdx.x & 63);
            uint32_t bad = 0;
            if (i < m.n && !(v.spix[i] & GF_SEED)) {
                uint32_t e[8], th[8], lo[8], lb[8];
                gf_nbrs(v, i, e);
                gf_gather8(v.T, e, th);
                gf_gather8(v.Gb, e, lo);
                gf_gather8(v.lbl, e, lb);
                uint32_t thi = GF_NONE;
#pragma unroll
                for (int j = 0; j < 8; ++j) thi = min(thi, th[j]);
                if (!(thi < v.Gb[i])) bad = kGfTie;
                const uint32_t L = v.lbl[i];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (e[j] != GF_NONE && lo[j] <= thi && lb[j] != L) bad = kGfTie;
            }
            gf_or(&a.meta[k].fail, bad);
        }
    }
}

// Labels of the frames that passed; the others get their seed labels back (non-seeds unlabelled)
// and are listed for the next round (work space) or the single-workgroup floods.
__global__ void __launch_bounds__(GF_NT) gf_write_kernel(GfArgs a)
{
    const uint32_t cnt = gf_count(a);
    const FrameGrid gg = frame_grid(cnt);
    for (uint32_t k = gg.y; k < cnt; k += gg.ny) {
        const GfMeta m = a.meta[k];
        const uint32_t f = a.flagList[k];
        if (gg.x == 0 && threadIdx.x == 0) {
            if (m.fail == kGfPool && a.round == 0) {
                a.retryList[atomicAdd(&a.counters[kGridRetryCnt], 1u)] = f;
            } else if (m.fail) {
                a.gridList[atomicAdd(&a.counters[kGridCnt], 1u)] = f;
                atomicAdd(&a.counters[(m.fail & kGfUnsettled) ? kGridUnsettled : (m.fail & kGfTie) ? kGridRejected : kGridOther], 1u);
            } else {
                atomicAdd(&a.counters[kGridOk], 1u);
                if (m.tie) atomicAdd(&a.counters[kGridTies], 1u);
            }
        }
        if ((m.fail & kGfPool) || m.n == 0) continue;
        const GfView v = gf_view(a.pool, m);
        uint32_t* fn = a.fin + (size_t)f * a.npx;
        for (uint32_t i = gg.x * GF_NT + threadIdx.x; i < m.n; i += gg.nx * GF_NT) {
            const uint32_t sx = v.spix[i];
            fn[sx & ~GF_SEED] = (m.fail && !(sx & GF_SEED)) ? 0u : v.lbl[i];
        }
    }
}

} // namespace

size_t grid_pool_words(size_t F, size_t npx)
{
    // About 10 bytes per pixel of the chunk (a frame needs 84 bytes per mask pixel: masks of up to
    // 12 % of every frame of a chunk), 64 MiB to 1.5 GiB.
    const size_t bytes = std::clamp<size_t>(F * npx * 10, size_t(64) << 20, size_t(3) << 29);
    return bytes / 4;
}

namespace {

// The launch sequence of the grid rank flood over the flagged frames of a chunk of up to F frames
// (the flagged count is read on the device).
void grid_flood_sequence(const DetectBuffers& b, int W, int H, int F, const Fz* fz, hipStream_t s)
{
    const size_t npx = b.npx;
    const unsigned bx = (unsigned)std::clamp<size_t>(npx / 16384, 2, 64);
    const unsigned by = (unsigned)std::clamp<size_t>(2048 / bx, 1, (size_t)F);
    const dim3 g(bx, by);
    // Merge passes for the largest frame that fits the pool (a frame of n entries needs
    // ceil(log2(n / GF_TILE)) of them; the kernels skip frames already merged).
    const size_t maxN = std::min<size_t>(npx, b.gridPoolWords / kGfWordsPerEntry);
    int passes = 0;
    for (size_t w = GF_TILE; w < maxN; w <<= 1) ++passes;
    GfArgs a{b.P, b.seedBits, b.fin, b.maskList, b.maskCnt, b.flagList, kFlagCnt, 1, b.counters, b.gridList, b.gridRetry,
             reinterpret_cast<GfMeta*>(b.gridMeta), b.gridPool, b.gridPoolWords, fz, b.aux, b.regionCap, npx, W, H, F};
    gf_setup_kernel<<<1, 1024, 0, s>>>(a);
    gf_reset_kernel<<<g, GF_NT, 0, s>>>(a);
    gf_collect_kernel<<<g, GF_NT, 0, s>>>(a);
    gf_tile_sort_kernel<<<g, GF_NT, 0, s>>>(a);
    for (int p = 0; p < passes; ++p) gf_merge_kernel<<<g, GF_NT, 0, s>>>(a, p);
    gf_finalize_kernel<<<g, GF_NT, 0, s>>>(a);
    gf_neighbours_kernel<<<g, GF_NT, 0, s>>>(a);
    for (int it = 0; it < kGridMaxIter; ++it) {
        gf_age_kernel<<<g, GF_NT, 0, s>>>(a, it);
        gf_rank_kernel<<<g, GF_NT, 0, s>>>(a, it);
    }
    gf_check1_kernel<<<g, GF_NT, 0, s>>>(a);
    gf_check2_kernel<<<g, GF_NT, 0, s>>>(a);
    gf_ival_init_kernel<<<g, GF_NT, 0, s>>>(a);
    for (int it = 0; it < kGridIvalIter; ++it) {
        gf_ival_key_kernel<<<g, GF_NT, 0, s>>>(a, it);
        gf_ival_rank_kernel<<<g, GF_NT, 0, s>>>(a, it);
    }
    gf_ival_check_kernel<<<g, GF_NT, 0, s>>>(a);
    gf_write_kernel<<<g, GF_NT, 0, s>>>(a);
}

} // namespace

void GridFloodGraph::release()
{
    for (hipGraphExec_t& e : exec) {
        if (e) (void)hipGraphExecDestroy(e);
        e = nullptr;
    }
}

void launch_grid_flood(const DetectBuffers& b, int W, int H, int F, const Fz* fz, hipStream_t s)
{
    // Sized by the lane's capacity, so that one graph serves chunks (and side batches) of any
    // number of frames.
    const int Fcap = (int)std::max<size_t>(b.frames, (size_t)F);
    GridFloodGraph* cache = b.gridGraph;
    if (!cache) {
        grid_flood_sequence(b, W, H, Fcap, fz, s);
        PT_CHECK(hipGetLastError());
        return;
    }
    const int pass = fz ? 1 : 0;
    const uint64_t key[24] = {(uint64_t)(uintptr_t)b.P, (uint64_t)(uintptr_t)b.seedBits, (uint64_t)(uintptr_t)b.fin,
                              (uint64_t)(uintptr_t)b.maskList, (uint64_t)(uintptr_t)b.maskCnt, (uint64_t)(uintptr_t)b.flagList,
                              (uint64_t)(uintptr_t)b.counters, (uint64_t)(uintptr_t)b.gridList, (uint64_t)(uintptr_t)b.gridRetry,
                              (uint64_t)(uintptr_t)b.gridMeta, (uint64_t)(uintptr_t)b.gridPool, b.gridPoolWords,
                              (uint64_t)(uintptr_t)fz, (uint64_t)(uintptr_t)b.aux, b.regionCap, b.npx, (uint64_t)W,
                              (uint64_t)H, (uint64_t)Fcap, (uint64_t)(uintptr_t)s, 1, 0, 0, 0};
    bool same = cache->exec[pass] != nullptr;
    for (int i = 0; i < 24 && same; ++i) same = cache->key[pass][i] == key[i];
    if (!same) {
        if (cache->exec[pass]) PT_CHECK(hipGraphExecDestroy(cache->exec[pass]));
        cache->exec[pass] = nullptr;
        hipGraph_t graph = nullptr;
        PT_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        grid_flood_sequence(b, W, H, Fcap, fz, s);
        PT_CHECK(hipStreamEndCapture(s, &graph));
        PT_CHECK(hipGraphInstantiate(&cache->exec[pass], graph, nullptr, nullptr, 0));
        PT_CHECK(hipGraphDestroy(graph));
        for (int i = 0; i < 24; ++i) cache->key[pass][i] = key[i];
    }
    PT_CHECK(hipGraphLaunch(cache->exec[pass], s));
}

} // namespace pt
