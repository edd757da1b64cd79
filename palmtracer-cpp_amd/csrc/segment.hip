// Segmentation2D on gfx950: local maxima, watershed labels, region statistics, ROI list.
//
// Reference: lib/localization/Segmentation2D.cpp (getLocalMax :294-333, watershed :336-385,
// ComputeStatistics :144-175, FilterData :88-141), lib/fit/Fitting.cpp:44-58 (per-frame capacity).
//
// The reference floods every frame from one std::priority_queue whose keys age with a global pop
// counter (key = -W1 + it * 1e-7, label claimed at push time). Labels can only differ between
// pixels of one 8-connected component of the mask, so the frame is split:
//   1. connected components of the mask by lock-free union-find (root = smallest raster index);
//   2. every local maximum gets its getLocalMax label, identified here by the raster index of the
//      first maximum of its plateau chain (+1), which orders labels exactly as the reference does;
//   3. a component whose maxima all share one label is labelled wholesale;
//   4. a component with several labels is flooded by one wave with an exact priority queue, using
//      the component-local pop count as the age. At every pop the wave proves that the popped entry
//      precedes every other queued entry of the component in the frame-global order too, whatever
//      the other components' pops did to the counter: the global gap between two pushes is at least
//      the local gap, and at most the number of mask pixels whose potential is at least the
//      earlier-pushed entry's minus n * 1e-7 (n the frame's mask size), since every entry popped
//      while that entry waited had a key no larger than its key. If the proof fails (near-equal
//      potentials, exact ties), the frame is flagged;
//   5. flagged frames are re-flooded exactly as a whole (rank flood, wave-queue flood, or a
//      sequential restatement of the reference's heap: raster-order seeds, one frame-wide heap,
//      global counter).
// Statistics are exact integer sums of P = 65536 * W1, so they equal the reference's FP64
// raster-order sums whenever those are exact (sums below 2^53 ulps); larger sums are recomputed in
// raster order in FP64.
#include "flood_common.h"

#include <algorithm>

namespace pt {
namespace {

constexpr uint32_t kCompRoot = 0x80000000u;   // lab bit of a component's first pixel
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int NT = 256;

__device__ inline uint32_t ld_relaxed(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint32_t uf_find(const uint32_t* lab, uint32_t x)
{
    uint32_t p = ld_relaxed(lab + x);
    while (p != x) {
        x = p;
        p = ld_relaxed(lab + x);
    }
    return x;
}

// ---------------------------------------------------------------------------------------------
// Heap with the move sequence of libstdc++'s std::priority_queue<Pixel> (bits/stl_heap.h), whose
// Pixel::operator< is reversed (lib/Structs.hpp:26): the top holds the smallest key.
struct DHeap {
    double* k;
    uint64_t* v;
    uint32_t n;

    __device__ void sift_up(uint32_t hole, double key, uint64_t val)
    {
        while (hole > 0) {
            const uint32_t parent = (hole - 1) / 2;
            if (!(k[parent] > key)) break;
            k[hole] = k[parent];
            v[hole] = v[parent];
            hole = parent;
        }
        k[hole] = key;
        v[hole] = val;
    }
    __device__ void push(double key, uint64_t val) { sift_up(n++, key, val); }
    __device__ uint64_t pop()
    {
        const uint64_t top = v[0];
        if (n > 1) {
            const uint32_t len = n - 1;
            const double vk = k[len];
            const uint64_t vv = v[len];
            uint32_t hole = 0, second = 0;
            while (second < (len - 1) / 2) {
                second = 2 * (second + 1);
                if (k[second] > k[second - 1]) second--;
                k[hole] = k[second];
                v[hole] = v[second];
                hole = second;
            }
            if ((len & 1) == 0 && second == (len - 2) / 2) {
                second = 2 * (second + 1);
                k[hole] = k[second - 1];
                v[hole] = v[second - 1];
                hole = second - 1;
            }
            sift_up(hole, vk, vv);
        }
        n--;
        return top;
    }
};

__device__ inline void flag_frame(uint32_t* frameFlag, uint32_t* counters, uint32_t* flagList, int f)
{
    if (atomicExch(&frameFlag[f], 1u) == 0u) {
        const uint32_t k = atomicAdd(&counters[kFlagCnt], 1u);
        flagList[k] = (uint32_t)f;   // k < F: each frame is listed once
    }
}

// ---------------------------------------------------------------------------------------------
// Second-pass labelling of the components: aux[component] holds the smallest and largest label of
// the component's surviving maxima.
__global__ void __launch_bounds__(NT) label_assign_kernel(const uint32_t* __restrict__ lab, uint32_t* __restrict__ fin,
                                                          const PixAux* __restrict__ aux, const uint32_t* __restrict__ maskList,
                                                          const uint32_t* __restrict__ maskCnt, uint64_t* __restrict__ multiList,
                                                          size_t multiCap, uint32_t* __restrict__ counters,
                                                          uint32_t* __restrict__ frameFlag, uint32_t* __restrict__ flagList,
                                                          size_t npx)
{
    const int f = blockIdx.y;
    const uint32_t cnt = maskCnt[f];
    for (uint32_t k = blockIdx.x * NT + threadIdx.x; k < cnt; k += gridDim.x * NT) {
        const uint32_t p = maskList[(size_t)f * npx + k];
        const uint32_t e = lab[(size_t)f * npx + p], c = e & ~kCompRoot;
        const PixAux a = aux[(size_t)f * npx + c];
        if (a.cmin == kNone) { flag_frame(frameFlag, counters, flagList, f); continue; }  // cannot happen
        if (a.cmin == a.cmax) { fin[(size_t)f * npx + p] = a.cmin; continue; }
        // Several labels: fin holds the surviving maxima only; list the component once.
        if (e & kCompRoot) {
            const uint32_t j = atomicAdd(&counters[kMultiCnt], 1u);
            if (j < multiCap) multiList[j] = ((uint64_t)f << 32) | c;
            else flag_frame(frameFlag, counters, flagList, f);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Component labels, getLocalMax labels and the wholesale labelling of single-label components for
// one frame per workgroup (steps 1-3 of the file header).
//
// Component identities: in frames whose mask fits in LDS, the union-find runs over the ranks of the
// mask pixels (raster-order index among the frame's mask pixels, from a mask bitmap with per-word
// prefix counts), so a component is identified by the rank of its first pixel; otherwise it runs
// over pixel indices in global memory and a component is its first pixel. Either way
// lab[p] = component | (kCompRoot when p is that first pixel) for every mask pixel p, and the
// multi-label components are listed once, as (frame << 32) | component.
//
// getLocalMax labels (Segmentation2D.cpp:309-327): a maximum inherits the label of the first earlier
// maximum among (y-1,x-1), (y-1,x), (y-1,x+1), (y,x-1); otherwise it opens a new label. Here the
// label of a maximum is the raster index of the first maximum of its chain, plus one, which orders
// labels as the reference numbers them. A component's label state is 0 (no maximum yet), its single
// label, or kNone once two maxima of different labels were seen.
constexpr int FL_NT = 1024;
constexpr int FL_PER = 8;                       // mask entries per thread in the LDS path
constexpr uint32_t FL_MAX_LDS_N = FL_NT * FL_PER;
constexpr uint32_t FL_SLOTS = 16;               // multi-label components flooded inside the workgroup

__device__ inline uint32_t chain_label(const uint32_t* sbf, uint32_t s, int W)
{
    uint32_t r = s;
    while (true) {
        const int y = (int)(r / (uint32_t)W), x = (int)r - y * W;
        uint32_t q = kNone;
        if (y > 0) {
            const uint32_t up = r - (uint32_t)W;
            if (x > 0 && seed_bit(sbf, up - 1)) q = up - 1;
            else if (seed_bit(sbf, up)) q = up;
            else if (x < W - 1 && seed_bit(sbf, up + 1)) q = up + 1;
        }
        if (q == kNone && x > 0 && seed_bit(sbf, r - 1)) q = r - 1;
        if (q == kNone) return r + 1;
        r = q;
    }
}

// Adds label L to a component's label state (0 -> L -> kNone on a second, different label).
template <typename Cas, typename Xchg>
__device__ inline void add_label(uint32_t* st, uint32_t L, Cas cas, Xchg xchg)
{
    const uint32_t old = cas(st, 0u, L);
    if (old != 0u && old != L && old != kNone) xchg(st, kNone);
}

// Upper bound of the global pops between the push of an entry of potential Pb and any later push
// while it waits: every entry popped meanwhile had a key no larger than its key, hence a potential
// above Pb - D (D quanta cover n * 1e-7 and the rounding of the keys, n the frame's mask size);
// cum[k] counts the frame's mask pixels at levels >= k.
__device__ inline double gap_bound(uint32_t Pb, const uint32_t* cum, uint32_t D, double nmax)
{
    return Pb > D ? (double)cum[hist_bin(Pb - D)] : nmax;
}

// Quanta of potential that the age term of n pops (and the FP64 rounding of the keys) can cover.
__host__ __device__ inline uint32_t age_quanta(uint32_t n)
{
    return (uint32_t)floor(((double)n * kEpsilon + 1e-9) * 65536.0) + 1u;
}

// Per-frame histogram of the mask potentials by level (hist_bin), as counts at or above each level.
constexpr int MH_NT = 512;
static_assert(MH_NT == (int)kHistBins, "one thread per level");
__global__ void __launch_bounds__(MH_NT) mask_hist_kernel(const uint32_t* __restrict__ P, const uint32_t* __restrict__ maskList,
                                                          const uint32_t* __restrict__ maskCnt, uint32_t* __restrict__ hist,
                                                          size_t npx)
{
    __shared__ uint32_t h[kHistBins];
    const int f = blockIdx.x, t = threadIdx.x;
    h[t] = 0u;
    __syncthreads();
    const uint32_t n = maskCnt[f];
    const uint32_t* ml = maskList + (size_t)f * npx;
    const uint32_t* Pf = P + (size_t)f * npx;
    for (uint32_t k = t; k < n; k += MH_NT) atomicAdd(&h[hist_bin(Pf[ml[k]])], 1u);
    __syncthreads();
    // Suffix sums (Hillis-Steele).
    for (int o = 1; o < MH_NT; o <<= 1) {
        const uint32_t v = t + o < MH_NT ? h[t + o] : 0u;
        __syncthreads();
        h[t] += v;
        __syncthreads();
    }
    hist[(size_t)f * kHistBins + t] = h[t];
}

// Frames of at least kHistSplitPx pixels: the frame's mask list is split over the workgroups of a
// grid row (x: slices, y: frames); each adds its LDS histogram into the frame's histogram (zeroed with
// the chunk counters) and mask_hist_suffix_kernel turns the counts into counts at or above each level.
__global__ void __launch_bounds__(MH_NT) mask_hist_partial_kernel(const uint32_t* __restrict__ P,
                                                                  const uint32_t* __restrict__ maskList,
                                                                  const uint32_t* __restrict__ maskCnt,
                                                                  uint32_t* __restrict__ hist, size_t npx)
{
    __shared__ uint32_t h[kHistBins];
    const int f = blockIdx.y, t = threadIdx.x;
    h[t] = 0u;
    __syncthreads();
    const uint32_t n = maskCnt[f];
    const uint32_t* ml = maskList + (size_t)f * npx;
    const uint32_t* Pf = P + (size_t)f * npx;
    for (uint32_t k = blockIdx.x * MH_NT + t; k < n; k += gridDim.x * MH_NT) atomicAdd(&h[hist_bin(Pf[ml[k]])], 1u);
    __syncthreads();
    if (h[t]) atomicAdd(&hist[(size_t)f * kHistBins + t], h[t]);
}

__global__ void __launch_bounds__(MH_NT) mask_hist_suffix_kernel(uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[kHistBins];
    const int f = blockIdx.x, t = threadIdx.x;
    h[t] = hist[(size_t)f * kHistBins + t];
    __syncthreads();
    for (int o = 1; o < MH_NT; o <<= 1) {
        const uint32_t v = t + o < MH_NT ? h[t + o] : 0u;
        __syncthreads();
        h[t] += v;
        __syncthreads();
    }
    hist[(size_t)f * kHistBins + t] = h[t];
}

// Value of lane l (wave-uniform l) in every lane, without a round trip through LDS.
__device__ inline uint32_t from_lane(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ inline int from_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline uint64_t from_lane(uint64_t v, int l)
{
    return ((uint64_t)from_lane((uint32_t)(v >> 32), l) << 32) | from_lane((uint32_t)v, l);
}

// One step of wave_min_f64: the minimum of x and x as moved by the DPP control CTRL (lanes of rows
// outside ROWS keep x).
template <int CTRL, int ROWS>
__device__ inline double dpp_min_step(double x)
{
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const int lo = (int)(uint32_t)b, hi = (int)(uint32_t)(b >> 32);
    const uint32_t lo2 = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xf, false);
    const uint32_t hi2 = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xf, false);
    const double y = __longlong_as_double((long long)(((uint64_t)hi2 << 32) | lo2));
    return y < x ? y : x;
}

// Minimum of x over the 64 lanes of a fully active wave, in every lane: DPP steps within quads,
// half rows and rows, then the row broadcasts, and lane 63 read back. Exact (a minimum of doubles
// does not depend on the order).
__device__ inline double wave_min_f64(double x)
{
    x = dpp_min_step<0xB1, 0xf>(x);    // quad_perm [1, 0, 3, 2]
    x = dpp_min_step<0x4E, 0xf>(x);    // quad_perm [2, 3, 0, 1]
    x = dpp_min_step<0x141, 0xf>(x);   // row_half_mirror
    x = dpp_min_step<0x140, 0xf>(x);   // row_mirror: every lane holds its row's minimum
    x = dpp_min_step<0x142, 0xa>(x);   // row_bcast15 into rows 1 and 3
    x = dpp_min_step<0x143, 0xc>(x);   // row_bcast31 into rows 2 and 3: lane 63 holds all
    return __longlong_as_double((long long)from_lane((uint64_t)__double_as_longlong(x), 63));
}

// Watershed of one multi-label component of at most 64 pixels by one wave, one pixel per lane, the
// queue held in registers: the same flood and the same certainty proof as flood_kernel (step 3 of
// the file header), with the component's adjacency as lane masks. list holds the component's
// pixels; seeds carry their labels in fin and every other pixel 0. Returns false when a pop meets
// equal keys or cannot be proven.
__device__ inline bool flood_small(const uint32_t* list, uint32_t cnt, const uint32_t* Pf, uint32_t* fn, int W, int H,
                                   double nmax, const uint32_t* cum, uint32_t D)
{
    const int lane = threadIdx.x & 63;
    const bool act = (uint32_t)lane < cnt;
    const uint32_t p = act ? list[lane] : 0u;
    const int y = (int)(p / (uint32_t)W), x = (int)p - y * W;
    const uint32_t P = act ? Pf[p] : 0u;
    const double gapB = gap_bound(P, cum, D, nmax);     // global pops while this pixel waits, at most
    uint32_t lbl = act ? ld_relaxed(fn + p) : 0u;
    uint64_t adj = 0;
    for (uint32_t j = 0; j < cnt; ++j) {
        const int qy = from_lane(y, (int)j), qx = from_lane(x, (int)j);
        const int dy = qy - y, dx = qx - x;
        if (act && (uint32_t)lane != j && dy >= -1 && dy <= 1 && dx >= -1 && dx <= 1) adj |= 1ull << j;
    }
    // Neighbours inside the frame; those outside the component are below the threshold.
    const int inFrame = (1 + (y > 0) + (y < H - 1)) * (1 + (x > 0) + (x < W - 1)) - 1;
    uint64_t unl = __ballot(act && lbl == 0u);
    // Seeds next to an unlabelled pixel are pushed at age 0 (Segmentation2D.cpp:340-359).
    uint64_t qm = __ballot(act && lbl != 0u && (__popcll(adj) < inFrame || (adj & unl) != 0));
    uint32_t age = 0;
    double key = age_key(P, 0);
    uint32_t it = 0;
    const double eps = kEpsilon, delta = 1e-9;
    while (qm) {
        const bool queued = (qm >> lane) & 1ull;
        const double km = wave_min_f64(queued ? key : 1.79769313486231570815e308);
        const uint64_t at = __ballot(queued && key == km);
        if (__popcll(at) != 1) return false;          // equal keys: the heap order would decide
        const int a = __ffsll((unsigned long long)at) - 1;
        const uint32_t pa = from_lane(P, a), ta = from_lane(age, a), la = from_lane(lbl, a);
        bool good = true;
        if (queued && lane != a) {
            const double oa = (double)pa * kPScale, ob = (double)P * kPScale;
            const uint32_t tb = age;
            if (ta == tb) good = oa != ob;
            else if (ta < tb) good = ob - oa < (double)(tb - ta) * eps - delta;
            else good = oa - ob > gapB * eps + delta;
        }
        if (__ballot(!good)) return false;
        ++it;
        const uint64_t claims = from_lane(adj, a) & unl;
        if ((claims >> lane) & 1ull) {
            lbl = la;
            age = it;
            key = age_key(P, it);
            fn[p] = la;
        }
        unl &= ~claims;
        qm = (qm & ~(1ull << a)) | claims;
    }
    return true;
}

// At most 80 SGPRs: MI355X then holds two of these 1024-thread workgroups per CU (8 waves per
// SIMD); from about 86 SGPRs on it holds one (measured with a residency probe), which halves the
// overlap of the workgroups' barrier-separated phases.
__global__ void __launch_bounds__(FL_NT) __attribute__((amdgpu_num_sgpr(80))) frame_label_kernel(
    const uint32_t* __restrict__ P, const uint32_t* __restrict__ seedBits, uint32_t* __restrict__ lab,
    uint32_t* __restrict__ fin, PixAux* __restrict__ aux, const uint32_t* __restrict__ maskList,
    const uint32_t* __restrict__ maskCnt, const uint32_t* __restrict__ seedList, const uint32_t* __restrict__ seedCnt,
    uint32_t* __restrict__ rootCnt, Region* __restrict__ regions, size_t regionCap, uint64_t* __restrict__ multiList,
    size_t multiCap, uint32_t* __restrict__ counters, uint32_t* __restrict__ frameFlag, uint32_t* __restrict__ flagList,
    int W, int H, size_t npx, uint32_t ldsCap, uint32_t parentOff, const uint32_t* __restrict__ hist)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t waveSum[FL_NT / 64];
    __shared__ uint32_t nRoots, nSlots;
    const int f = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n = maskCnt[f], ns = seedCnt[f];
    const uint32_t* ml = maskList + (size_t)f * npx;
    const uint32_t* sl = seedList + (size_t)f * npx;
    const uint32_t* sbf = seedBits + (size_t)f * seed_words(npx);
    const uint32_t* Pf = P + (size_t)f * npx;
    uint32_t* lb = lab + (size_t)f * npx;
    uint32_t* fn = fin + (size_t)f * npx;
    PixAux* ax = aux + (size_t)f * npx;
    if (tid == 0) { nRoots = 0u; nSlots = 0u; }

    // A maximum s with chain label L: its label, and its region when it opens one.
    auto label_seed = [&](uint32_t s) -> uint32_t {
        const uint32_t L = chain_label(sbf, s, W);
        fn[s] = L;
        if (L - 1u == s) {
            const uint32_t slot = atomicAdd(&nRoots, 1u);
            ax[s].rslot = slot;
            if (slot < regionCap) regions[(size_t)f * regionCap + slot] = Region{s, 0u, 0ull, 0ull, 0ull};
            else atomicOr(&counters[kStatus], 1u);
        }
        return L;
    };
    // Final labels of mask pixel p of component c with label state st.
    auto assign = [&](uint32_t p, uint32_t c, bool first, uint32_t st) {
        lb[p] = c | (first ? kCompRoot : 0u);
        if (st == 0u) { flag_frame(frameFlag, counters, flagList, f); return; }   // cannot happen
        if (st != kNone) { fn[p] = st; return; }
        if (!seed_bit(sbf, p)) fn[p] = 0u;     // several labels: the flood starts from the maxima
        if (first) {
            const uint32_t j = atomicAdd(&counters[kMultiCnt], 1u);
            if (j < multiCap) multiList[j] = ((uint64_t)f << 32) | c;
            else flag_frame(frameFlag, counters, flagList, f);
        }
    };

    if (ldsCap > 0 && n <= ldsCap) {
        const uint32_t nd = (uint32_t)((npx + 63) / 64), nw = 2 * nd;   // bitmap in 64-bit and 32-bit words
        uint32_t* mbits = reinterpret_cast<uint32_t*>(smem);
        const uint64_t* mb64 = reinterpret_cast<const uint64_t*>(smem);
        uint16_t* pre = reinterpret_cast<uint16_t*>(mbits + nw);      // mask pixels before each 64-bit word
        uint32_t* parent = mbits + parentOff;
        uint32_t* state = mbits;               // reuses the bitmap once every rank is known
        auto rank = [&](uint32_t q) {
            return (uint32_t)pre[q >> 6] + (uint32_t)__popcll(mb64[q >> 6] & ((1ull << (q & 63)) - 1ull));
        };
        auto in_mask = [&](uint32_t q) { return (mbits[q >> 5] >> (q & 31)) & 1u; };
        auto ld = [&](const uint32_t* a) { return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); };
        auto find = [&](uint32_t x) {
            uint32_t p = ld(parent + x);
            while (p != x) { x = p; p = ld(parent + x); }
            return x;
        };

        // 1. Mask bitmap and prefix counts.
        for (uint32_t w = tid; w < nw; w += FL_NT) mbits[w] = 0u;
        __syncthreads();
        uint32_t pp[FL_PER], rr[FL_PER];
#pragma unroll
        for (int j = 0; j < FL_PER; ++j) {
            const uint32_t k = tid + j * FL_NT;
            pp[j] = k < n ? ml[k] : 0u;
            if (k < n) atomicOr(&mbits[pp[j] >> 5], 1u << (pp[j] & 31));
        }
        __syncthreads();
        {
            const uint32_t per = (nd + FL_NT - 1) / FL_NT, w0 = min(nd, tid * per), w1 = min(nd, w0 + per);
            uint32_t s = 0;
            for (uint32_t w = w0; w < w1; ++w) s += (uint32_t)__popcll(mb64[w]);
            uint32_t incl = s;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o);
                if (lane >= o) incl += t;
            }
            if (lane == 63) waveSum[wave] = incl;
            __syncthreads();
            uint32_t base = 0;
            for (int i = 0; i < wave; ++i) base += waveSum[i];
            uint32_t run = base + incl - s;
            for (uint32_t w = w0; w < w1; ++w) {
                pre[w] = (uint16_t)run;
                run += (uint32_t)__popcll(mb64[w]);
            }
        }
        __syncthreads();
        // 2. Union-find over ranks. A run of mask pixels in a row is one set from the start: every
        //    pixel's parent is the run's first pixel (the ranks of a run are consecutive). Run R =
        //    [a, b] of row y touches (8-connectivity) each run U = [c, d] of row y-1 with c <= b+1
        //    and d >= a-1, and is united with it once, from one pixel: from a with NW when c < a,
        //    from a with N when c == a, from c-1 with NE when c > a (there N is not in the mask).
#pragma unroll
        for (int j = 0; j < FL_PER; ++j) {
            if (tid + j * FL_NT < n) {
                const uint32_t p = pp[j];
                const uint32_t rowStart = p - p % (uint32_t)W;
                uint32_t q = p;                       // [q, p] is in the mask
                while (q > rowStart) {
                    const uint32_t w = q >> 5, zeros = ~mbits[w] & ((1u << (q & 31)) - 1u);
                    if (zeros) { q = max(rowStart, (w << 5) + 32u - (uint32_t)__clz(zeros)); break; }
                    q = max(rowStart, w << 5);
                    if (q == rowStart || !in_mask(q - 1)) break;
                    --q;
                }
                rr[j] = rank(p);
                parent[rr[j]] = rr[j] - (p - q);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < FL_PER; ++j) {
            if (tid + j * FL_NT >= n) continue;
            const uint32_t p = pp[j];
            const int y = (int)(p / (uint32_t)W), x = (int)p - y * W;
            if (y == 0) continue;
            auto unite = [&](uint32_t q) {
                uint32_t a = rr[j], b = rank(q);
                while (true) {
                    a = find(a);
                    b = find(b);
                    if (a == b) return;
                    if (a > b) { const uint32_t t = a; a = b; b = t; }
                    const uint32_t old = atomicMin(&parent[b], a);
                    if (old == b) return;
                    b = old;
                }
            };
            const uint32_t up = p - (uint32_t)W;
            const bool nIn = in_mask(up);
            if (x == 0 || !in_mask(p - 1)) {          // p starts its run
                if (x > 0 && in_mask(up - 1)) unite(up - 1);
                else if (nIn) unite(up);
            }
            if (!nIn && x < W - 1 && in_mask(up + 1)) unite(up + 1);
        }
        __syncthreads();
        // 3. Flatten: every rank points at its component (the component's smallest rank).
        uint32_t firstBits = 0;
#pragma unroll
        for (int j = 0; j < FL_PER; ++j) {
            if (tid + j * FL_NT >= n) continue;
            const uint32_t c = find(rr[j]);
            if (c == rr[j]) firstBits |= 1u << j;
            else __hip_atomic_store(&parent[rr[j]], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            rr[j] = c;
        }
        __syncthreads();
        // 4. Maxima: labels, regions, and their components (ranks still readable).
        uint32_t sc[FL_PER], sL[FL_PER];
#pragma unroll
        for (int j = 0; j < FL_PER; ++j) {
            const uint32_t k = tid + j * FL_NT;
            sc[j] = kNone;
            if (k < ns) {
                const uint32_t s = sl[k];
                sc[j] = parent[rank(s)];
                sL[j] = label_seed(s);
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < n; i += FL_NT) state[i] = 0u;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < FL_PER; ++j)
            if (sc[j] != kNone)
                add_label(&state[sc[j]], sL[j], [](uint32_t* a, uint32_t c, uint32_t v) { return atomicCAS(a, c, v); },
                          [](uint32_t* a, uint32_t v) { atomicExch(a, v); });
        __syncthreads();
        // 5. Labels of every mask pixel. Multi-label components take a slot (their pixel list, in
        //    the parents' space, which is free now) for the flood below; the rest of them are listed
        //    for flood_kernel.
        uint32_t* slotComp = parent;
        uint32_t* slotCnt = parent + FL_SLOTS;
        uint32_t* slotPix = parent + 2 * FL_SLOTS;
        const uint32_t slotMax = ldsCap >= 2 * FL_SLOTS + 64 ? min((uint32_t)FL_SLOTS, (ldsCap - 2 * FL_SLOTS) / 64) : 0u;
        if (tid < FL_SLOTS) slotCnt[tid] = 0u;
        uint32_t multiBits = 0;
#pragma unroll
        for (int j = 0; j < FL_PER; ++j) {
            if (tid + j * FL_NT >= n) continue;
            const uint32_t st = state[rr[j]];
            const bool first = (firstBits >> j) & 1u;
            if (st != kNone || slotMax == 0) { assign(pp[j], rr[j], first, st); continue; }
            lb[pp[j]] = rr[j] | (first ? kCompRoot : 0u);
            if (!seed_bit(sbf, pp[j])) fn[pp[j]] = 0u;
            multiBits |= 1u << j;
            if (first) {
                const uint32_t sl = atomicAdd(&nSlots, 1u);
                if (sl < slotMax) {
                    slotComp[sl] = rr[j];
                } else {
                    const uint32_t k = atomicAdd(&counters[kMultiCnt], 1u);
                    if (k < multiCap) multiList[k] = ((uint64_t)f << 32) | rr[j];
                    else flag_frame(frameFlag, counters, flagList, f);
                }
            }
        }
        __syncthreads();
        // 6. Pixel lists of the slotted components, then one wave per component: the register
        //    flood for components of up to 64 pixels, flood_kernel for larger ones.
        const uint32_t nSl = min(nSlots, slotMax);
#pragma unroll
        for (int j = 0; j < FL_PER; ++j) {
            if (!((multiBits >> j) & 1u)) continue;
            for (uint32_t sl = 0; sl < nSl; ++sl) {
                if (slotComp[sl] != rr[j]) continue;
                const uint32_t pos = atomicAdd(&slotCnt[sl], 1u);
                if (pos < 64u) slotPix[sl * 64 + pos] = pp[j];
                break;
            }
        }
        __syncthreads();
        for (uint32_t sl = (uint32_t)wave; sl < nSl; sl += FL_NT / 64) {
            const uint32_t cnt = slotCnt[sl];
            bool ok = true;
            if (cnt > 64u) {
                if (lane == 0) {
                    const uint32_t k = atomicAdd(&counters[kMultiCnt], 1u);
                    if (k < multiCap) multiList[k] = ((uint64_t)f << 32) | slotComp[sl];
                    else flag_frame(frameFlag, counters, flagList, f);
                }
            } else {
                ok = flood_small(slotPix + sl * 64, cnt, Pf, fn, W, H, (double)n, hist + (size_t)f * kHistBins,
                                 age_quanta(n));
            }
            if (!ok && lane == 0) {
                atomicAdd(&counters[kFlagUnproven], 1u);
                flag_frame(frameFlag, counters, flagList, f);
            }
        }
    }
    __syncthreads();
    if (tid == 0 && ldsCap > 0 && n <= ldsCap) rootCnt[f] = nRoots;
}

// ---------------------------------------------------------------------------------------------
// The same labelling for frames whose mask exceeds the LDS capacity of frame_label_kernel, in
// global memory over pixel indices, spread over several workgroups per frame (y: frames, strided;
// x: mask pixels): union-find with device-scope atomics, label states in aux[component].cmin,
// region slots from rootCnt. Five passes: parents, unions, flattening, maxima, labels.
struct FlgArgs {
    const uint32_t* P;
    const uint32_t* seedBits;
    uint32_t* lab;
    uint32_t* fin;
    PixAux* aux;
    const uint32_t* maskList;
    const uint32_t* maskCnt;
    const uint32_t* seedList;
    const uint32_t* seedCnt;
    uint32_t* rootCnt;
    Region* regions;
    size_t regionCap;
    uint64_t* multiList;
    size_t multiCap;
    uint32_t* counters;
    uint32_t* frameFlag;
    uint32_t* flagList;
    int W, H, F;
    size_t npx;
    uint32_t ldsCap;
};

__device__ inline bool flg_global(const FlgArgs& a, uint32_t f) { return !(a.ldsCap > 0 && a.maskCnt[f] <= a.ldsCap); }

__global__ void __launch_bounds__(NT) flg_parents_kernel(FlgArgs a)
{
    const FrameGrid gg = frame_grid((uint32_t)a.F);
    for (uint32_t f = gg.y; f < (uint32_t)a.F; f += gg.ny) {
        if (!flg_global(a, f)) continue;
        const uint32_t n = a.maskCnt[f];
        const uint32_t* ml = a.maskList + (size_t)f * a.npx;
        uint32_t* lb = a.lab + (size_t)f * a.npx;
        for (uint32_t k = gg.x * NT + threadIdx.x; k < n; k += gg.nx * NT) lb[ml[k]] = ml[k];
    }
}

__global__ void __launch_bounds__(NT) flg_union_kernel(FlgArgs a)
{
    const FrameGrid gg = frame_grid((uint32_t)a.F);
    for (uint32_t f = gg.y; f < (uint32_t)a.F; f += gg.ny) {
        if (!flg_global(a, f)) continue;
        const uint32_t n = a.maskCnt[f];
        const uint32_t* ml = a.maskList + (size_t)f * a.npx;
        const uint32_t* Pf = a.P + (size_t)f * a.npx;
        uint32_t* lb = a.lab + (size_t)f * a.npx;
        for (uint32_t k = gg.x * NT + threadIdx.x; k < n; k += gg.nx * NT) {
            const uint32_t p = ml[k];
            const int y = (int)(p / (uint32_t)a.W), x = (int)p - y * a.W;
            const int dys[4] = {0, -1, -1, -1}, dxs[4] = {-1, -1, 0, 1};
            for (int j = 0; j < 4; ++j) {
                const int yy = y + dys[j], xx = x + dxs[j];
                if (yy < 0 || xx < 0 || xx >= a.W) continue;
                const uint32_t q = (uint32_t)(yy * a.W + xx);
                if (Pf[q] == 0) continue;
                uint32_t u = p, v = q;
                while (true) {
                    u = uf_find(lb, u);
                    v = uf_find(lb, v);
                    if (u == v) break;
                    if (u > v) { const uint32_t t = u; u = v; v = t; }
                    const uint32_t old = atomicMin(&lb[v], u);
                    if (old == v) break;
                    v = old;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(NT) flg_flatten_kernel(FlgArgs a)
{
    const FrameGrid gg = frame_grid((uint32_t)a.F);
    for (uint32_t f = gg.y; f < (uint32_t)a.F; f += gg.ny) {
        if (!flg_global(a, f)) continue;
        const uint32_t n = a.maskCnt[f];
        const uint32_t* ml = a.maskList + (size_t)f * a.npx;
        uint32_t* lb = a.lab + (size_t)f * a.npx;
        PixAux* ax = a.aux + (size_t)f * a.npx;
        for (uint32_t k = gg.x * NT + threadIdx.x; k < n; k += gg.nx * NT) {
            const uint32_t p = ml[k];
            const uint32_t c = uf_find(lb, p);
            if (c == p) ax[p].cmin = 0u;
            else __hip_atomic_store(&lb[p], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ void __launch_bounds__(NT) flg_maxima_kernel(FlgArgs a)
{
    const FrameGrid gg = frame_grid((uint32_t)a.F);
    for (uint32_t f = gg.y; f < (uint32_t)a.F; f += gg.ny) {
        if (!flg_global(a, f)) continue;
        const uint32_t ns = a.seedCnt[f];
        const uint32_t* sl = a.seedList + (size_t)f * a.npx;
        const uint32_t* sbf = a.seedBits + (size_t)f * seed_words(a.npx);
        uint32_t* lb = a.lab + (size_t)f * a.npx;
        uint32_t* fn = a.fin + (size_t)f * a.npx;
        PixAux* ax = a.aux + (size_t)f * a.npx;
        for (uint32_t k = gg.x * NT + threadIdx.x; k < ns; k += gg.nx * NT) {
            const uint32_t s = sl[k];
            const uint32_t c = ld_relaxed(&lb[s]);
            const uint32_t L = chain_label(sbf, s, a.W);
            fn[s] = L;
            if (L - 1u == s) {
                const uint32_t slot = atomicAdd(&a.rootCnt[f], 1u);
                ax[s].rslot = slot;
                if (slot < a.regionCap) a.regions[(size_t)f * a.regionCap + slot] = Region{s, 0u, 0ull, 0ull, 0ull};
                else atomicOr(&a.counters[kStatus], 1u);
            }
            add_label(&ax[c].cmin, L, [](uint32_t* q, uint32_t c0, uint32_t v) { return atomicCAS(q, c0, v); },
                      [](uint32_t* q, uint32_t v) { atomicExch(q, v); });
        }
    }
}

__global__ void __launch_bounds__(NT) flg_labels_kernel(FlgArgs a)
{
    const FrameGrid gg = frame_grid((uint32_t)a.F);
    for (uint32_t f = gg.y; f < (uint32_t)a.F; f += gg.ny) {
        if (!flg_global(a, f)) continue;
        const uint32_t n = a.maskCnt[f];
        const uint32_t* ml = a.maskList + (size_t)f * a.npx;
        const uint32_t* sbf = a.seedBits + (size_t)f * seed_words(a.npx);
        uint32_t* lb = a.lab + (size_t)f * a.npx;
        uint32_t* fn = a.fin + (size_t)f * a.npx;
        const PixAux* ax = a.aux + (size_t)f * a.npx;
        for (uint32_t k = gg.x * NT + threadIdx.x; k < n; k += gg.nx * NT) {
            const uint32_t p = ml[k];
            const uint32_t c = ld_relaxed(&lb[p]);
            const bool first = c == p;
            const uint32_t st = ld_relaxed(&ax[c].cmin);
            lb[p] = c | (first ? kCompRoot : 0u);
            if (st == 0u) { flag_frame(a.frameFlag, a.counters, a.flagList, (int)f); continue; }   // cannot happen
            if (st != kNone) { fn[p] = st; continue; }
            if (!seed_bit(sbf, p)) fn[p] = 0u;     // several labels: the flood starts from the maxima
            if (first) {
                const uint32_t j = atomicAdd(&a.counters[kMultiCnt], 1u);
                if (j < a.multiCap) a.multiList[j] = ((uint64_t)f << 32) | c;
                else flag_frame(a.frameFlag, a.counters, a.flagList, (int)f);
            }
        }
    }
}

// LDS of frame_label_kernel for frames of npx pixels: the bitmap and prefix counts (later the
// component label states), then the union-find parents. Sets the largest mask handled in LDS and
// the offset of the parents in words.
size_t frame_label_lds(size_t npx, uint32_t& ldsCap, uint32_t& parentOff)
{
    const size_t nd = (npx + 63) / 64;
    const size_t base = nd * 8 + (nd + 1) / 2 * 4;        // bitmap + u16 prefix counts per 64-bit word
    const size_t staticBytes = 256;
    auto total = [&](size_t cap) { return std::max(base, cap * 4) + cap * 4; };
    auto fit = [&](size_t budget) {
        size_t cap = FL_MAX_LDS_N;
        while (cap > 0 && total(cap) > budget) cap -= 64;
        return cap;
    };
    // Two workgroups per CU when that leaves room for a typical mask, else one.
    size_t budget = 81920 - staticBytes;
    size_t cap = fit(budget);
    if (cap < 2048) {
        budget = 163840 - staticBytes;
        cap = fit(budget);
    }
    if (cap < 256) { ldsCap = 0; parentOff = 0; return 0; }
    ldsCap = (uint32_t)cap;
    parentOff = (uint32_t)(std::max(base, cap * 4) / 4);
    return total(cap);
}
// ---------------------------------------------------------------------------------------------
// Wave-cooperative priority queue in LDS: an unsorted array of entries (pixel, label, potential,
// age); the key -P/2^16 + age * 1e-7 is recomputed from (potential, age) exactly as at push time,
// and a pop is a wave-wide minimum. Without equal keys every exact priority queue pops in the
// same order as the reference's heap, so equal keys at a pop are reported (and handled by the
// caller) instead of being ordered.
struct WaveQ {
    uint64_t* pa;      // (P << 32) | age
    double* key;       // reference key of the entry, computed once at push
    uint32_t* pix;
    uint32_t* lbl;
    uint32_t cap;
};


// Orders the LDS accesses of a one-wave workgroup. A wave's LDS instructions execute in order, so no
// hardware wait is needed (a workgroup barrier would also wait for every outstanding global store);
// this only keeps the compiler from moving memory accesses across the point.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_wave_barrier();
    __asm__ __volatile__("" ::: "memory");
}

__device__ inline void waveq_put(WaveQ& q, uint32_t pos, uint32_t pix, uint32_t lbl, uint32_t p, uint32_t age)
{
    q.pa[pos] = ((uint64_t)p << 32) | age;
    q.key[pos] = age_key(p, age);
    q.pix[pos] = pix;
    q.lbl[pos] = lbl;
}

// Pops the minimum. Returns false (and does not pop) when two entries share the minimal key.
// With kCheck, check(pa_popped, pa_other) == false marks the pop unproven (result in `proven`).
template <bool kCheck, typename Check>
__device__ inline bool waveq_pop(WaveQ& q, uint32_t& n, uint32_t& outPix, uint32_t& outLbl, bool& proven, Check&& check)
{
    const int lane = threadIdx.x & 63;
    double km = 1.79769313486231570815e308;
    uint32_t pm = 0xFFFFFFFFu, cm = 0;
    // One wave alone on its CU: issue a batch of LDS reads before comparing, so that their latency
    // is paid once per batch rather than once per entry.
    constexpr uint32_t U = 8;
    for (uint32_t b = 0; b < n; b += 64 * U) {
        double kk[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t i = b + u * 64 + lane;
            kk[u] = i < n ? q.key[i] : 1.79769313486231570815e308;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t i = b + u * 64 + lane;
            if (i >= n) continue;
            if (kk[u] < km) { km = kk[u]; pm = i; cm = 1; }
            else if (kk[u] == km) ++cm;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double k2 = __shfl_xor(km, o);
        const uint32_t p2 = __shfl_xor(pm, o), c2 = __shfl_xor(cm, o);
        if (k2 < km) { km = k2; pm = p2; cm = c2; }
        else if (k2 == km) { cm += c2; pm = p2 < pm ? p2 : pm; }
    }
    if (cm > 1) return false;  // equal keys: the heap order would decide
    const uint64_t apa = q.pa[pm];
    const uint32_t a = q.pix[pm], la = q.lbl[pm];
    proven = true;
    if constexpr (kCheck) {
        bool bad = false;
        for (uint32_t b = 0; b < n; b += 64 * U) {
            uint64_t pp[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t i = b + u * 64 + lane;
                pp[u] = i < n ? q.pa[i] : 0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t i = b + u * 64 + lane;
                if (i < n && i != pm && !check(apa, pp[u])) bad = true;
            }
        }
        proven = __ballot(bad) == 0;
    }
    wave_sync();
    if (lane == 0) {
        q.pa[pm] = q.pa[n - 1];
        q.key[pm] = q.key[n - 1];
        q.pix[pm] = q.pix[n - 1];
        q.lbl[pm] = q.lbl[n - 1];
    }
    wave_sync();
    n -= 1;
    outPix = a;
    outLbl = la;
    return true;
}

// Claims the unlabelled mask neighbours of `a` for label la and pushes them with age `it`
// (Segmentation2D.cpp:368-383). Returns false on queue overflow.
__device__ inline bool waveq_expand(WaveQ& q, uint32_t& n, const uint32_t* Pf, uint32_t* fn, int W, int H, uint32_t a,
                                    uint32_t la, uint32_t it)
{
    const int lane = threadIdx.x & 63;
    bool claim = false;
    uint32_t qpix = 0, qp = 0;
    if (lane < 8) {
        int dy, dx;
        nb_offset(lane, dy, dx);
        const int y = (int)(a / W) + dy, x = (int)(a % W) + dx;
        if (y >= 0 && y < H && x >= 0 && x < W) {
            qpix = (uint32_t)(y * W + x);
            qp = Pf[qpix];
            // Labels written by this wave in earlier pops: read at L2 (coherent), with the potential.
            const uint32_t ql = ld_relaxed(fn + qpix);
            claim = qp != 0 && ql == 0u;
        }
    }
    const uint64_t m = __ballot(claim);
    const uint32_t cnt = __popcll(m);
    if (n + cnt > q.cap) return false;
    if (claim) {
        waveq_put(q, n + __popcll(m & ((1ull << lane) - 1ull)), qpix, la, qp, it);
        __hip_atomic_store(fn + qpix, la, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    wave_sync();
    n += cnt;
    return true;
}

constexpr uint32_t QCAP_FLOOD = 1024;

// One wave floods one multi-label component with the component-local pop count as the age, and
// proves at every pop that the frame-global order would pop the same entry (file header, step 4).
// Seeds are the component's local maxima (first pass) or its surviving regions' maximum pixels
// (second pass, fz != nullptr).
__global__ void __launch_bounds__(64) flood_kernel(const uint32_t* __restrict__ P, const uint32_t* __restrict__ lab,
                                                   uint32_t* __restrict__ fin, const uint32_t* __restrict__ seedList,
                                                   const uint32_t* __restrict__ seedCnt, const uint32_t* __restrict__ maskCnt,
                                                   const uint64_t* __restrict__ multiList, size_t multiCap,
                                                   uint32_t* __restrict__ counters, uint32_t* __restrict__ frameFlag,
                                                   uint32_t* __restrict__ flagList, int W, int H, size_t npx,
                                                   const Fz* __restrict__ fz, const uint32_t* __restrict__ rootCnt,
                                                   size_t regionCap, const uint32_t* __restrict__ hist, uint32_t denseMask)
{
    __shared__ uint64_t sPa[QCAP_FLOOD];
    __shared__ double sKey[QCAP_FLOOD];
    __shared__ uint32_t sPix[QCAP_FLOOD], sLbl[QCAP_FLOOD];
    WaveQ q{sPa, sKey, sPix, sLbl, QCAP_FLOOD};
    const int lane = threadIdx.x;
    const uint32_t nMulti = (uint32_t)min((size_t)counters[kMultiCnt], multiCap);
    for (uint32_t t = blockIdx.x; t < nMulti; t += gridDim.x) {
        const uint64_t e = multiList[t];
        const int f = (int)(e >> 32);
        const uint32_t c = (uint32_t)e;
        if (__hip_atomic_load(&frameFlag[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;
        if (denseMask && maskCnt[f] > denseMask) {
            if (lane == 0) flag_frame(frameFlag, counters, flagList, f);
            continue;
        }
        const uint32_t* Pf = P + (size_t)f * npx;
        const uint32_t* lb = lab + (size_t)f * npx;
        uint32_t* fn = fin + (size_t)f * npx;
        // Seeds of the component that the reference would push (labelled, next to an unlabelled pixel).
        uint32_t n = 0;
        bool ok = true, overflow = false;
        const uint32_t ns = fz ? (uint32_t)min((size_t)rootCnt[f], regionCap) : seedCnt[f];
        for (uint32_t b0 = 0; b0 < ns && ok; b0 += 64) {
            const uint32_t k = b0 + lane;
            bool sel = false;
            uint32_t sp = 0;
            if (k < ns) {
                if (fz) {
                    const Fz z = fz[(size_t)f * regionCap + k];
                    if (z.alive && z.maxKey != 0) { sp = 0xFFFFFFFFu - (uint32_t)z.maxKey; sel = true; }
                } else {
                    sp = seedList[(size_t)f * npx + k];
                    sel = true;
                }
                sel = sel && (lb[sp] & ~kCompRoot) == c && fn[sp] != 0u && has_unlabelled_nbr(Pf, fn, W, H, (int)sp);
            }
            const uint64_t m = __ballot(sel);
            const uint32_t cnt = __popcll(m);
            if (n + cnt > q.cap) { ok = false; overflow = true; break; }
            if (sel) waveq_put(q, n + __popcll(m & ((1ull << lane) - 1ull)), sp, fn[sp], Pf[sp], 0u);
            n += cnt;
        }
        __syncthreads();
        const double eps = kEpsilon, delta = 1e-9;
        const double nmax = (double)maskCnt[f];
        const uint32_t D = age_quanta(maskCnt[f]);
        const uint32_t* cum = hist + (size_t)f * kHistBins;
        uint32_t it = 0;
        bool dropped = false;
        while (ok && n > 0) {
            // Another wave flagged the frame: its labels come from the exact frame flood.
            if ((it & 63u) == 63u && __hip_atomic_load(&frameFlag[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                dropped = true;
                break;
            }
            uint32_t a, la;
            bool proven = true;
            // b (potential ob, age tb) stays behind the popped a (oa, ta) in the frame-global order
            // for every admissible global age: the global gap between two pushes is at least the
            // local gap (ta < tb case) and at most gap_bound of b (ta > tb case).
            auto check = [&](uint64_t ap, uint64_t bp) {
                const double oa = (double)(uint32_t)(ap >> 32) * kPScale, ob = (double)(uint32_t)(bp >> 32) * kPScale;
                const uint32_t ta = (uint32_t)ap, tb = (uint32_t)bp;
                if (ta == tb) return oa != ob;
                if (ta < tb) return ob - oa < (double)(tb - ta) * eps - delta;
                return oa - ob > gap_bound((uint32_t)(bp >> 32), cum, D, nmax) * eps + delta;
            };
            if (!waveq_pop<true>(q, n, a, la, proven, check) || !proven) { ok = false; break; }
            ++it;
            if (!waveq_expand(q, n, Pf, fn, W, H, a, la, it)) { ok = false; overflow = true; break; }
        }
        if (!ok && !dropped && lane == 0) {
            atomicAdd(&counters[overflow ? kFlagOverflow : kFlagUnproven], 1u);
            flag_frame(frameFlag, counters, flagList, f);
        }
        __syncthreads();
    }
}

constexpr uint32_t FB_LDS_BYTES = 163840;   // all of a CU's LDS: one wave per CU, one frame per wave

// Frame state held in LDS by the exact frame flood: the mask as a bitmap with per-word prefix
// counts (rank of a mask pixel = its index among the frame's mask pixels in raster order), the
// potentials of the mask pixels by rank, and the labelled state by rank. Every neighbour test of
// the flood is then an LDS access.
struct FrameLds {
    uint32_t* mbits;   // npx bits
    uint32_t* pre;     // exclusive prefix of popc(mbits[w])
    uint32_t* pc;      // P by rank
    uint32_t* lbits;   // labelled, by rank
    __device__ bool in_mask(uint32_t q) const { return (mbits[q >> 5] >> (q & 31)) & 1u; }
    __device__ uint32_t rank(uint32_t q) const
    {
        const uint32_t w = mbits[q >> 5];
        return pre[q >> 5] + __popc(w & ((1u << (q & 31)) - 1u));
    }
    __device__ bool labelled(uint32_t r) const { return (lbits[r >> 5] >> (r & 31)) & 1u; }
};

// waveq_expand on the LDS frame state: claims read and set the labelled bits, labels are only
// written out to fin.
__device__ inline bool waveq_expand_lds(WaveQ& q, uint32_t& n, const FrameLds& fr, uint32_t* fn, int W, int H, uint32_t a,
                                        uint32_t la, uint32_t it)
{
    const int lane = threadIdx.x & 63;
    bool claim = false;
    uint32_t qpix = 0, qp = 0, r = 0;
    if (lane < 8) {
        int dy, dx;
        nb_offset(lane, dy, dx);
        const int y = (int)(a / W) + dy, x = (int)(a % W) + dx;
        if (y >= 0 && y < H && x >= 0 && x < W) {
            qpix = (uint32_t)(y * W + x);
            if (fr.in_mask(qpix)) {
                r = fr.rank(qpix);
                qp = fr.pc[r];
                claim = !fr.labelled(r);
            }
        }
    }
    const uint64_t m = __ballot(claim);
    const uint32_t cnt = __popcll(m);
    if (n + cnt > q.cap) return false;
    if (claim) {
        waveq_put(q, n + __popcll(m & ((1ull << lane) - 1ull)), qpix, la, qp, it);
        atomicOr(&fr.lbits[r >> 5], 1u << (r & 31));
        fn[qpix] = la;
    }
    wave_sync();
    n += cnt;
    return true;
}

// ---------------------------------------------------------------------------------------------
// Rank flood: the exact frame flood as a parallel fixed point.
//
// Number the pops of the reference's flood 1, 2, ... and let G(x) be the pop number of entry x.
// A seed with an unlabelled neighbour is pushed at age T = 0; any other mask pixel is claimed,
// labelled and pushed by its first popped neighbour, at age T(x) = min over its flooded neighbours
// of G. Its key is K(x) = -W1(x) + T(x) * 1e-7. If the keys are pairwise distinct and every entry is
// pushed before the pop at its key rank (T(x) < G(x) for every non-seed), the pop sequence is the
// keys in increasing order, i.e. G(x) = 1 + #{y : K(y) < K(x)} (induction on the pop number: the
// entry of rank m is present and is the smallest key present). Conversely the reference's G satisfies
// both equations. So the flood is the fixed point of T = min-neighbour(G), G = rank(K(T)), found by
// iterating from G = rank of W1 alone; keys only move by at most n * 1e-7 (n entries), so only
// entries whose potentials lie within that distance ("the window", contiguous in potential order)
// can swap ranks, and the rank of x is (entries with a potential above the window) + (window entries
// with a smaller key). The labels follow the first-popped-neighbour links. A frame whose iteration
// does not settle, meets two equal keys, or breaks T < G goes to the wave-queue flood.
constexpr int RF_NT = 1024;
constexpr uint32_t RF_NONE = 0xFFFFFFFFu;
constexpr int RF_MAX_ITER = 64;
constexpr uint32_t RF_PAIRS = 6;              // equal-key pairs per frame (2^6 combinations of orders)

__global__ void __launch_bounds__(RF_NT) rank_flood_kernel(const uint32_t* __restrict__ P, const uint32_t* __restrict__ seedBits, uint32_t* __restrict__ fin,
                                                           const uint32_t* __restrict__ maskList,
                                                           const uint32_t* __restrict__ maskCnt,
                                                           const uint32_t* __restrict__ flagList, int countSlot,
                                                           uint32_t* __restrict__ counters,
                                                           uint32_t* __restrict__ waveList, uint32_t* __restrict__ rfScratch,
                                                           int W, int H, size_t npx, const Fz* __restrict__ fz,
                                                           const PixAux* __restrict__ aux, size_t regionCap)
{
    __shared__ uint64_t sk[kRankLds];  // (~P << 32) | pixel, sorted ascending = potential descending
    __shared__ uint32_t sN, sFrame, sPairs, sPairI[RF_PAIRS], sPairJ[RF_PAIRS];
    const int tid = threadIdx.x;
    const size_t R = rf_stride(npx);                     // entry capacity (power of two)
    uint32_t* idxOf = rfScratch + (size_t)blockIdx.x * rf_slot_words(npx);
    uint32_t* sp = idxOf + ((npx + 1) & ~size_t(1));   // P by entry
    uint32_t* spix = sp + R;                           // pixel by entry
    uint32_t* lo = spix + R;                           // window [lo, hi) of the entry
    uint32_t* hi = lo + R;
    uint32_t* Ga = hi + R;                             // pop ranks (two buffers)
    uint32_t* Gb = Ga + R;
    uint32_t* T = Gb + R;                              // push ages
    uint32_t* par = T + R;                             // first popped neighbour (labels)
    uint32_t* lbl = par + R;
    uint32_t* seedf = lbl + R;                         // entry is a seed
    double* K = reinterpret_cast<double*>(seedf + R);
    uint32_t* nbr = reinterpret_cast<uint32_t*>(K + R);      // [entry][8] flooded neighbours
    uint32_t* lblA = nbr + (size_t)8 * R;                    // labels of the base solution
    uint32_t* G0 = lblA + R;                                 // ranks and ages of the base solution
    uint32_t* T0 = G0 + R;
    uint64_t* skg = reinterpret_cast<uint64_t*>(T0 + R);     // sort keys beyond kRankLds entries
    while (true) {
        __syncthreads();
        if (tid == 0) sFrame = atomicAdd(&counters[kRfNext], 1u);
        __syncthreads();
        const uint32_t fi = sFrame;
        if (fi >= counters[countSlot]) return;
        const uint32_t f = flagList[fi];
        const uint32_t* Pf = P + (size_t)f * npx;
        uint32_t* fn = fin + (size_t)f * npx;
        const uint32_t* ml = maskList + (size_t)f * npx;
        const uint32_t cnt = maskCnt[f];

        // 1. Labels back to the seed state; seed marks in idxOf for now.
        for (uint32_t k = tid; k < cnt; k += RF_NT) {
            const uint32_t p = ml[k];
            const bool seed = flood_seed(seedBits, fn, fz, aux, regionCap, npx, f, p);
            if (!seed) fn[p] = 0u;
            idxOf[p] = seed ? 1u : 0u;
        }
        if (tid == 0) sN = 0;
        __threadfence_block();
        __syncthreads();
        // 2. Flood entries: non-seeds, and seeds next to an unlabelled pixel.
        for (uint32_t k = tid; k < cnt; k += RF_NT) {
            const uint32_t p = ml[k];
            const bool seed = idxOf[p] != 0u;
            if (!seed || has_unlabelled_nbr(Pf, fn, W, H, (int)p)) {
                const uint32_t j = atomicAdd(&sN, 1u);
                if (j < R) skg[j] = ((uint64_t)(~Pf[p]) << 32) | p;
            }
        }
        __syncthreads();
        const uint32_t n = sN;
        bool ok = n <= R;
        if (!ok && tid == 0) atomicAdd(&counters[kRfTooBig], 1u);
        if (ok && n > 0) {
            // 3. Sort by potential (bitonic; in LDS up to kRankLds entries, else in the work space),
            //    then the entries in potential order.
            uint32_t n2 = 2;
            while (n2 < n) n2 <<= 1;
            auto sort_entries = [&](uint64_t* s) {
                for (uint32_t i = n + tid; i < n2; i += RF_NT) s[i] = ~0ull;
                __syncthreads();
                for (uint32_t kk = 2; kk <= n2; kk <<= 1) {
                    for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
                        for (uint32_t i = tid; i < n2; i += RF_NT) {
                            const uint32_t ixj = i ^ j;
                            if (ixj > i) {
                                const uint64_t a = s[i], b = s[ixj];
                                if ((a > b) == ((i & kk) == 0)) { s[i] = b; s[ixj] = a; }
                            }
                        }
                        __syncthreads();
                    }
                }
                for (uint32_t i = tid; i < n; i += RF_NT) {
                    const uint64_t v = s[i];
                    const uint32_t p = (uint32_t)v;
                    sp[i] = ~(uint32_t)(v >> 32);
                    spix[i] = p;
                    seedf[i] = idxOf[p];
                }
            };
            // Beyond the LDS: the bitonic network with every pass of stride below kRankLds run on
            // kRankLds-key chunks in LDS (each pair of such a pass lies in one aligned chunk; the
            // direction comes from the global index), the wider passes in the work space.
            auto sort_large = [&]() {
                constexpr uint32_t C = kRankLds;
                for (uint32_t i = n + tid; i < n2; i += RF_NT) skg[i] = ~0ull;
                __syncthreads();
                auto chunk_passes = [&](uint32_t c0, uint32_t kk0, uint32_t kk1, uint32_t j0) {
                    for (uint32_t i = tid; i < C; i += RF_NT) sk[i] = skg[c0 + i];
                    __syncthreads();
                    for (uint32_t kk = kk0; kk <= kk1; kk <<= 1) {
                        for (uint32_t j = (kk == kk0 ? j0 : kk >> 1); j > 0; j >>= 1) {
                            for (uint32_t i = tid; i < C; i += RF_NT) {
                                const uint32_t ixj = i ^ j;
                                if (ixj > i) {
                                    const uint64_t a = sk[i], b = sk[ixj];
                                    if ((a > b) == (((c0 + i) & kk) == 0)) { sk[i] = b; sk[ixj] = a; }
                                }
                            }
                            __syncthreads();
                        }
                    }
                    for (uint32_t i = tid; i < C; i += RF_NT) skg[c0 + i] = sk[i];
                    __syncthreads();
                };
                for (uint32_t c0 = 0; c0 < n2; c0 += C) chunk_passes(c0, 2, C, 1);
                for (uint32_t kk = 2 * C; kk <= n2; kk <<= 1) {
                    for (uint32_t j = kk >> 1; j >= C; j >>= 1) {
                        for (uint32_t i = tid; i < n2; i += RF_NT) {
                            const uint32_t ixj = i ^ j;
                            if (ixj > i) {
                                const uint64_t a = skg[i], b = skg[ixj];
                                if ((a > b) == ((i & kk) == 0)) { skg[i] = b; skg[ixj] = a; }
                            }
                        }
                        __syncthreads();
                    }
                    for (uint32_t c0 = 0; c0 < n2; c0 += C) chunk_passes(c0, kk, kk, C >> 1);
                }
                for (uint32_t i = tid; i < n; i += RF_NT) {
                    const uint64_t v = skg[i];
                    const uint32_t p = (uint32_t)v;
                    sp[i] = ~(uint32_t)(v >> 32);
                    spix[i] = p;
                    seedf[i] = idxOf[p];
                }
            };
            if (n <= kRankLds) {
                for (uint32_t i = tid; i < n; i += RF_NT) sk[i] = skg[i];
                __syncthreads();
                sort_entries(sk);
            } else {
                sort_large();
            }
            __syncthreads();
            // 4. Pixel -> entry map; windows; flooded neighbours.
            for (uint32_t k = tid; k < cnt; k += RF_NT) idxOf[ml[k]] = RF_NONE;  // unpushed seeds stay out
            __syncthreads();
            for (uint32_t i = tid; i < n; i += RF_NT) idxOf[spix[i]] = i;
            __syncthreads();
            // Potentials further apart than D quanta keep their order for every age <= n.
            const uint64_t D = (uint64_t)floor(((double)n * kEpsilon + 1e-9) * 65536.0) + 1;
            // lo[i]: first entry with P <= P_i + D; hi[i]: first entry with P < P_i - D. Both only grow
            // with i (P descending), so each thread binary-searches them for the first entry of a
            // contiguous block and sweeps forward over the rest of it.
            {
                const uint32_t per = (n + RF_NT - 1) / RF_NT, i0 = min(n, tid * per), i1 = min(n, i0 + per);
                uint32_t a = 0, b = n, lw = 0, hw = 0;
                if (i0 < i1) {
                    const uint64_t p0 = sp[i0];
                    while (a < b) {
                        const uint32_t m = (a + b) >> 1;
                        if ((uint64_t)sp[m] > p0 + D) a = m + 1; else b = m;
                    }
                    lw = a;
                    a = 0; b = n;
                    while (a < b) {
                        const uint32_t m = (a + b) >> 1;
                        if ((uint64_t)sp[m] + D >= p0) a = m + 1; else b = m;
                    }
                    hw = a;
                }
                for (uint32_t i = i0; i < i1; ++i) {
                    const uint64_t pi = sp[i];
                    while (lw < n && (uint64_t)sp[lw] > pi + D) ++lw;
                    while (hw < n && (uint64_t)sp[hw] + D >= pi) ++hw;
                    lo[i] = lw;
                    hi[i] = hw;
                }
            }
            for (uint32_t i = tid; i < n; i += RF_NT) {
                const uint32_t p = spix[i];
                const int y = (int)(p / W), x = (int)(p % W);
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    int dy, dx;
                    nb_offset(j, dy, dx);
                    const int yy = y + dy, xx = x + dx;
                    uint32_t e = RF_NONE;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                        const uint32_t q = (uint32_t)(yy * W + xx);
                        if (Pf[q] != 0u) e = idxOf[q];
                    }
                    nbr[(size_t)i * 8 + j] = e;
                }
                Ga[i] = i + 1;
            }
            __syncthreads();
            // 5.-7. Fixed point, checks and labels. Two equal keys belong to siblings pushed by one
            // pop with one potential; the heap orders them by its internal layout. The flood is solved
            // with every pair in index order (the base), then for every other combination of the
            // pairs' orders (up to RF_PAIRS pairs), each started from the base with the flipped pairs'
            // ranks exchanged. The heap's choice is one of the combinations; when every combination
            // settles, keeps the base's pairs and gives the base's labels, the labels are the
            // reference's whatever the heap chose. Otherwise the frame goes to the wave-queue flood.
            uint32_t nPairs = 0;
            auto flipped = [&](uint32_t mask, uint32_t a, uint32_t b) {
                for (uint32_t p = 0; p < nPairs; ++p)
                    if (((mask >> p) & 1u) && ((a == sPairI[p] && b == sPairJ[p]) || (a == sPairJ[p] && b == sPairI[p])))
                        return true;
                return false;
            };
            for (uint32_t trial = 0; ok && trial < (1u << nPairs); ++trial) {
                for (uint32_t i = tid; i < n; i += RF_NT) {
                    uint32_t g = i + 1;
                    if (trial) {
                        g = G0[i];
                        for (uint32_t p = 0; p < nPairs; ++p)
                            if ((trial >> p) & 1u) {
                                if (i == sPairI[p]) g = G0[sPairJ[p]];
                                else if (i == sPairJ[p]) g = G0[sPairI[p]];
                            }
                    }
                    Ga[i] = g;
                }
                if (tid == 0) sPairs = 0;
                __syncthreads();
                uint32_t* G = Ga;
                uint32_t* Gn = Gb;
                bool settled = false;
                for (int iter = 0; iter < RF_MAX_ITER; ++iter) {
                    for (uint32_t i = tid; i < n; i += RF_NT) {
                        uint32_t t = 0;
                        if (!seedf[i]) {
                            t = RF_NONE;
#pragma unroll
                            for (int j = 0; j < 8; ++j) {
                                const uint32_t e = nbr[(size_t)i * 8 + j];
                                if (e != RF_NONE) t = min(t, G[e]);
                            }
                        }
                        T[i] = t;
                        K[i] = t == RF_NONE ? 1.79769313486231570815e308 : age_key(sp[i], t);
                    }
                    __syncthreads();
                    bool changed = false;
                    for (uint32_t i = tid; i < n; i += RF_NT) {
                        const double ki = K[i];
                        uint32_t c = 0;
                        for (uint32_t j = lo[i]; j < hi[i]; ++j) {
                            if (j == i) continue;
                            const double kj = K[j];
                            c += kj < ki || (kj == ki && ((j < i) != (trial != 0 && flipped(trial, i, j))));
                        }
                        const uint32_t g = 1 + lo[i] + c;
                        Gn[i] = g;
                        changed |= g != G[i];
                    }
                    const bool anyChange = __syncthreads_or(changed);
                    uint32_t* t2 = G; G = Gn; Gn = t2;
                    if (!anyChange) { settled = true; break; }
                }
                // Checks: settled, every entry pushed before its pop, the equal-key pairs (listed by the
                // base; the same pairs in every other combination).
                bool viol = !settled, dep = false;
                for (uint32_t i = tid; i < n && settled; i += RF_NT) {
                    if (!seedf[i] && !(T[i] < G[i])) viol = true;
                    const double ki = K[i];
                    for (uint32_t j = i + 1; j < hi[i]; ++j)
                        if (K[j] == ki) {
                            const uint32_t k = atomicAdd(&sPairs, 1u);
                            if (trial == 0) {
                                if (k < RF_PAIRS) { sPairI[k] = i; sPairJ[k] = j; }
                            } else if (!flipped(~0u, i, j)) {
                                dep = true;
                            }
                        }
                }
                const bool bad = __syncthreads_or(viol);
                bool tie = __syncthreads_or(dep) || (trial != 0 && sPairs != nPairs);
                if (!bad && trial == 0) {
                    nPairs = sPairs;
                    tie = nPairs > RF_PAIRS;
                    // Each entry in at most one pair (no three equal keys).
                    for (uint32_t a = tid; a < nPairs * nPairs && !tie; a += RF_NT) {
                        const uint32_t p = a / nPairs, q = a % nPairs;
                        if (p < q && (sPairI[p] == sPairI[q] || sPairI[p] == sPairJ[q] || sPairJ[p] == sPairI[q] ||
                                      sPairJ[p] == sPairJ[q]))
                            tie = true;
                    }
                    tie = __syncthreads_or(tie);
                }
                if (bad || tie) {
                    if (tid == 0) atomicAdd(&counters[bad ? kRfUnsettled : kRfTie], 1u);
                    ok = false;
                    break;
                }
                // Labels through the first-popped-neighbour links (pointer jumping).
                for (uint32_t i = tid; i < n; i += RF_NT) {
                    if (trial == 0) {
                        G0[i] = G[i];
                        T0[i] = T[i];
                    }
                    if (seedf[i]) {
                        par[i] = i;
                        lbl[i] = fn[spix[i]];
                    } else {
                        uint32_t best = RF_NONE, bg = RF_NONE;
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const uint32_t e = nbr[(size_t)i * 8 + j];
                            if (e != RF_NONE && G[e] < bg) { bg = G[e]; best = e; }
                        }
                        par[i] = best;
                        lbl[i] = RF_NONE;
                    }
                }
                __syncthreads();
                for (int round = 0; round < 48; ++round) {
                    bool pending = false;
                    for (uint32_t i = tid; i < n; i += RF_NT) {
                        if (lbl[i] != RF_NONE) continue;
                        const uint32_t pp = par[i];
                        const uint32_t lp = lbl[pp];
                        if (lp != RF_NONE) lbl[i] = lp;
                        else { par[i] = par[pp]; pending = true; }
                    }
                    if (!__syncthreads_or(pending)) break;
                }
                bool missing = false, differ = false;
                for (uint32_t i = tid; i < n; i += RF_NT) {
                    const uint32_t L = lbl[i];
                    if (L == RF_NONE || L == 0u) missing = true;
                    if (trial == 0) lblA[i] = L;
                    else differ |= L != lblA[i];
                }
                if (__syncthreads_or(missing) || __syncthreads_or(differ)) {
                    if (tid == 0) atomicAdd(&counters[missing ? kRfUnsettled : kRfTie], 1u);
                    ok = false;
                    break;
                }
            }
            if (ok) {
                for (uint32_t i = tid; i < n; i += RF_NT)
                    if (!seedf[i]) fn[spix[i]] = lblA[i];
                if (tid == 0 && nPairs > 0) atomicAdd(&counters[kRfTwoOrders], 1u);
            }
            __syncthreads();
        }
        if (!ok && tid == 0) waveList[atomicAdd(&counters[kWaveCnt], 1u)] = f;
    }
}

__device__ inline size_t align8(size_t v) { return (v + 7) & ~size_t(7); }

// Exact frame flood for flagged frames, one wave per frame: the reference's process (seeds in any
// order, global pop count as the age) with the wave queue. A frame whose queue meets equal keys at
// a pop, or overflows, goes to the heap restatement below (fallback_kernel) through slowList.
// The frame's mask, potentials and labelled state live in LDS when they fit next to a queue of
// useful size; otherwise they are read from global memory.
__global__ void __launch_bounds__(64) fallback_fast_kernel(const uint32_t* __restrict__ P, const uint32_t* __restrict__ seedBits, uint32_t* __restrict__ fin,
                                                           const uint32_t* __restrict__ maskList,
                                                           const uint32_t* __restrict__ maskCnt,
                                                           const uint32_t* __restrict__ seedList,
                                                           const uint32_t* __restrict__ seedCnt,
                                                           const uint32_t* __restrict__ flagList, int countSlot,
                                                           uint32_t* __restrict__ counters,
                                                           uint32_t* __restrict__ slowList, int W, int H, size_t npx,
                                                           const Fz* __restrict__ fz, const Region* __restrict__ regions,
                                                           const uint32_t* __restrict__ rootCnt, const PixAux* __restrict__ aux,
                                                           size_t regionCap, int forceSlow)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const uint32_t nw = (uint32_t)((npx + 31) / 32);
    while (true) {
        uint32_t i = 0;
        if (lane == 0) i = atomicAdd(&counters[kFbNext], 1u);
        i = __shfl(i, 0);
        if (i >= counters[countSlot]) return;
        const uint32_t f = flagList[i];
        if (forceSlow) {
            if (lane == 0) slowList[atomicAdd(&counters[kSlowCnt], 1u)] = f;
            continue;
        }
        const uint32_t* Pf = P + (size_t)f * npx;
        uint32_t* fn = fin + (size_t)f * npx;
        const uint32_t* ml = maskList + (size_t)f * npx;
        const uint32_t cnt = maskCnt[f];
        // LDS layout: frame state (when it fits beside a queue of at least 1024 entries), then the queue.
        const size_t frameBytes = align8((size_t)nw * 8 + (size_t)cnt * 4 + ((size_t)cnt + 31) / 32 * 4);
        const bool inLds = frameBytes + 1024 * 24 <= FB_LDS_BYTES;
        FrameLds fr{};
        size_t off = 0;
        if (inLds) {
            fr.mbits = reinterpret_cast<uint32_t*>(smem);
            fr.pre = fr.mbits + nw;
            fr.pc = fr.pre + nw;
            fr.lbits = fr.pc + cnt;
            off = frameBytes;
        }
        const uint32_t cap = (uint32_t)((FB_LDS_BYTES - off) / 24);
        WaveQ q{reinterpret_cast<uint64_t*>(smem + off), reinterpret_cast<double*>(smem + off + (size_t)cap * 8),
                reinterpret_cast<uint32_t*>(smem + off + (size_t)cap * 16),
                reinterpret_cast<uint32_t*>(smem + off + (size_t)cap * 20), cap};

        if (inLds) {
            // Mask bitmap, prefix counts, labelled bits cleared.
            for (uint32_t w = lane; w < nw; w += 64) fr.mbits[w] = 0u;
            for (uint32_t w = lane; w < (cnt + 31) / 32; w += 64) fr.lbits[w] = 0u;
            __syncthreads();
            for (uint32_t k = lane; k < cnt; k += 64) {
                const uint32_t p = ml[k];
                atomicOr(&fr.mbits[p >> 5], 1u << (p & 31));
            }
            __syncthreads();
            // Exclusive scan of the word popcounts: each lane a contiguous block, then across lanes.
            const uint32_t per = (nw + 63) / 64, w0 = lane * per, w1 = min(nw, w0 + per);
            uint32_t s = 0;
            for (uint32_t w = w0; w < w1; ++w) s += __popc(fr.mbits[w]);
            uint32_t incl = s;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o);
                if (lane >= o) incl += t;
            }
            uint32_t run = incl - s;
            for (uint32_t w = w0; w < w1; ++w) {
                fr.pre[w] = run;
                run += __popc(fr.mbits[w]);
            }
            __syncthreads();
        }
        // 1. Labels back to the seed state (seeds marked labelled); potentials by rank.
        for (uint32_t k = lane; k < cnt; k += 64) {
            const uint32_t p = ml[k];
            const bool seed = flood_seed(seedBits, fn, fz, aux, regionCap, npx, f, p);
            if (!seed) fn[p] = 0u;
            if (inLds) {
                const uint32_t r = fr.rank(p);
                fr.pc[r] = Pf[p];
                if (seed) atomicOr(&fr.lbits[r >> 5], 1u << (r & 31));
            }
        }
        __threadfence_block();
        __syncthreads();
        // 2. Seeds next to an unlabelled pixel (the reference pushes them in raster order; order is
        //    irrelevant here without equal keys).
        uint32_t n = 0;
        bool ok = true;
        const uint32_t nsrc = fz ? (uint32_t)min((size_t)rootCnt[f], regionCap) : seedCnt[f];
        for (uint32_t b0 = 0; b0 < nsrc && ok; b0 += 64) {
            const uint32_t k = b0 + lane;
            bool sel = false;
            uint32_t sp = 0;
            if (k < nsrc) {
                if (fz) {
                    const Fz z = fz[(size_t)f * regionCap + k];
                    if (z.alive && z.maxKey != 0) { sp = 0xFFFFFFFFu - (uint32_t)z.maxKey; sel = true; }
                } else {
                    sp = seedList[(size_t)f * npx + k];
                    sel = true;
                }
                sel = sel && fn[sp] != 0u && has_unlabelled_nbr(Pf, fn, W, H, (int)sp);
            }
            const uint64_t m = __ballot(sel);
            const uint32_t c = __popcll(m);
            if (n + c > q.cap) { ok = false; break; }
            if (sel) waveq_put(q, n + __popcll(m & ((1ull << lane) - 1ull)), sp, fn[sp], Pf[sp], 0u);
            n += c;
        }
        __syncthreads();
        // 3. Flood in key order with the frame-global pop count.
        uint32_t it = 0, maxN = n;
        while (ok && n > 0) {
            maxN = n > maxN ? n : maxN;
            uint32_t a, la;
            bool proven = true;
            if (!waveq_pop<false>(q, n, a, la, proven, [](uint64_t, uint64_t) { return true; })) { ok = false; break; }
            ++it;
            const bool pushed = inLds ? waveq_expand_lds(q, n, fr, fn, W, H, a, la, it)
                                      : waveq_expand(q, n, Pf, fn, W, H, a, la, it);
            if (!pushed) {
                ok = false;
                if (lane == 0) atomicAdd(&counters[kFbOverflow], 1u);
                break;
            }
        }
        if (lane == 0) {
            atomicAdd(&counters[kFloodPops], it);
            atomicMax(&counters[kFloodMaxQueue], maxN);
        }
        if (!ok && lane == 0) {
            const uint32_t j = atomicAdd(&counters[kSlowCnt], 1u);
            slowList[j] = f;
        }
        __syncthreads();
    }
}

// Exact sequential restatement of Segmentation2D::watershed with the libstdc++ heap order, one
// frame per workgroup (frames with equal keys at a pop, taken from slowList through a shared
// cursor). Work space per slot: fb_slot_words(npx) words.
__global__ void __launch_bounds__(NT) fallback_kernel(const uint32_t* __restrict__ P, const uint32_t* __restrict__ seedBits, uint32_t* __restrict__ fin,
                                                      const uint32_t* __restrict__ maskList, const uint32_t* __restrict__ maskCnt,
                                                      const uint32_t* __restrict__ flagList, uint32_t* __restrict__ counters,
                                                      uint32_t* __restrict__ fbScratch, int W, int H, size_t npx,
                                                      const Fz* __restrict__ fz, const PixAux* __restrict__ aux,
                                                      size_t regionCap)
{
    __shared__ uint32_t sFrame, sWave[NT / 64], sTotal;
    const size_t npx2 = (npx + 1) & ~size_t(1);
    uint32_t* seeds = fbScratch + (size_t)blockIdx.x * fb_slot_words(npx);
    double* hk = (double*)(seeds + npx2);
    uint64_t* hv = (uint64_t*)(seeds + 3 * npx2);
    while (true) {
        if (threadIdx.x == 0) {
            const uint32_t i = atomicAdd(&counters[kSlowNext], 1u);
            sFrame = i < counters[kSlowCnt] ? flagList[i] : kNone;
        }
        __syncthreads();
        const uint32_t f = sFrame;
        if (f == kNone) return;
        const uint32_t* Pf = P + (size_t)f * npx;
        uint32_t* fn = fin + (size_t)f * npx;
        // 1. Labels back to the seed state: local maxima (first pass) or the one maximum pixel per
        //    surviving region (after fusion) keep theirs, every other mask pixel 0.
        const uint32_t cnt = maskCnt[f];
        for (uint32_t k = threadIdx.x; k < cnt; k += NT) {
            const uint32_t p = maskList[(size_t)f * npx + k];
            const bool seed = flood_seed(seedBits, fn, fz, aux, regionCap, npx, f, p);
            if (!seed) fn[p] = 0u;
        }
        __syncthreads();
        // 2. Seeds to push, in raster order (block-wide ordered compaction).
        uint32_t total = 0;
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        for (size_t base = 0; base < npx; base += NT) {
            const size_t i = base + threadIdx.x;
            const bool sel = i < npx && Pf[i] != 0 && fn[i] != 0 && has_unlabelled_nbr(Pf, fn, W, H, (int)i);
            const uint64_t m = __ballot(sel);
            if (lane == 0) sWave[wv] = __popcll(m);
            __syncthreads();
            uint32_t before = total;
            for (int j = 0; j < wv; ++j) before += sWave[j];
            before += __popcll(m & ((1ull << lane) - 1ull));
            if (sel) seeds[before] = (uint32_t)i;
            if (threadIdx.x == 0) {
                uint32_t s = 0;
                for (int j = 0; j < NT / 64; ++j) s += sWave[j];
                sTotal = s;
            }
            __syncthreads();
            total += sTotal;
            __syncthreads();
        }
        // 3. The flood itself, sequential as in the reference.
        if (threadIdx.x == 0) {
            DHeap h{hk, hv, 0u};
            for (uint32_t i = 0; i < total; ++i) h.push(age_key(Pf[seeds[i]], 0), (uint64_t)seeds[i]);
            uint32_t it = 0;
            while (h.n > 0) {
                const uint32_t a = (uint32_t)h.pop();
                ++it;
                const uint32_t la = fn[a];
                const int y = a / W, x = a - y * W;
                for (int dy = -1; dy <= 1; ++dy) {
                    const int yy = y + dy;
                    if (yy < 0 || yy >= H) continue;
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int xx = x + dx;
                        if ((dx | dy) == 0 || xx < 0 || xx >= W) continue;
                        const uint32_t q = (uint32_t)(yy * W + xx);
                        if (Pf[q] != 0 && fn[q] == 0 && h.n < cnt) {
                            h.push(age_key(Pf[q], it), (uint64_t)q);
                            fn[q] = la;
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
}


// ---------------------------------------------------------------------------------------------
// Basin fusion (Segmentation2D.cpp:181-291), one workgroup per frame. Regions are identified by
// their first local maximum; label order is the order of those pixels, as in the reference.
__device__ inline uint64_t pkey(uint32_t v, uint32_t p) { return ((uint64_t)v << 32) | (0xFFFFFFFFu - p); }
__device__ inline uint32_t kpix(uint64_t k) { return 0xFFFFFFFFu - (uint32_t)k; }

__global__ void __launch_bounds__(NT) fusion_kernel(const uint32_t* __restrict__ P, uint32_t* __restrict__ fin,
                                                    const PixAux* __restrict__ aux, const uint32_t* __restrict__ maskList,
                                                    const uint32_t* __restrict__ maskCnt, const Region* __restrict__ regions,
                                                    size_t regionCap, const uint32_t* __restrict__ rootCnt, Fz* __restrict__ fz,
                                                    uint32_t* __restrict__ order, double ratio, int W, int H, size_t npx,
                                                    const uint32_t* __restrict__ skip)
{
    __shared__ uint32_t sChange;
    const int f = blockIdx.x;
    const uint32_t n = (skip && skip[f]) ? 0u : (uint32_t)min((size_t)rootCnt[f], regionCap);
    if (n == 0) return;
    const size_t base = (size_t)f * regionCap;
    const uint32_t* Pf = P + (size_t)f * npx;
    uint32_t* fn = fin + (size_t)f * npx;
    const PixAux* ax = aux + (size_t)f * npx;
    Fz* z = fz + base;
    const Region* rg = regions + base;
    uint32_t* ord = order + base;
    const uint32_t cnt = maskCnt[f];
    const uint32_t* ml = maskList + (size_t)f * npx;
    for (uint32_t s = threadIdx.x; s < n; s += NT) {
        z[s].alive = 1u;
        uint32_t r = 0;
        const uint32_t root = rg[s].root;
        for (uint32_t t = 0; t < n; ++t) r += rg[t].root < root;
        ord[r] = s;
    }
    __syncthreads();
    auto slot_of = [&](uint32_t L) -> uint32_t { return (L != 0 && L <= npx) ? ax[L - 1].rslot : kNone; };
    while (true) {
        for (uint32_t s = threadIdx.x; s < n; s += NT)
            if (z[s].alive) {
                z[s].maxKey = 0;
                z[s].nbKey = 0;
                z[s].minP = kNone;
                z[s].target = s;
                z[s].locked = 0;
                z[s].nbLabel = 0;
            }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < cnt; k += NT) {
            const uint32_t p = ml[k];
            const uint32_t L = fn[p], s = slot_of(L);
            if (s >= n) continue;
            const uint32_t v = Pf[p];
            atomicMin(&z[s].minP, v);
            atomicMax((unsigned long long*)&z[s].maxKey, (unsigned long long)pkey(v, p));
            const int y = p / W, x = p - y * W;
            bool bnd = false;
            for (int dy = -1; dy <= 1 && !bnd; ++dy) {
                const int yy = y + dy;
                if (yy < 0 || yy >= H) continue;
                for (int dx = -1; dx <= 1 && !bnd; ++dx) {
                    const int xx = x + dx;
                    if ((dx | dy) == 0 || xx < 0 || xx >= W) continue;
                    const int q = yy * W + xx;
                    bnd = Pf[q] != 0 && fn[q] != 0 && fn[q] != L;
                }
            }
            if (bnd) atomicMax((unsigned long long*)&z[s].nbKey, (unsigned long long)pkey(v, p));
        }
        __syncthreads();
        // Neighbour label: first other-region neighbour, in the reference's scan order, of the first
        // pixel that carries the region's largest boundary potential.
        for (uint32_t s = threadIdx.x; s < n; s += NT) {
            if (!z[s].alive || z[s].nbKey == 0) continue;
            const uint32_t p = kpix(z[s].nbKey), L = fn[p];
            const int y = p / W, x = p - y * W;
            uint32_t nb = 0;
            for (int dy = -1; dy <= 1 && !nb; ++dy) {
                const int yy = y + dy;
                if (yy < 0 || yy >= H) continue;
                for (int dx = -1; dx <= 1 && !nb; ++dx) {
                    const int xx = x + dx;
                    if ((dx | dy) == 0 || xx < 0 || xx >= W) continue;
                    const int q = yy * W + xx;
                    if (Pf[q] != 0 && fn[q] != 0 && fn[q] != L) nb = fn[q];
                }
            }
            z[s].nbLabel = nb;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t change = 0;
            for (uint32_t r = 0; r < n; ++r) {
                const uint32_t s = ord[r];
                if (!z[s].alive || z[s].nbLabel == 0) continue;
                const uint32_t t = slot_of(z[s].nbLabel);
                if (t >= n || z[s].locked || z[t].locked) continue;
                const uint32_t A = (uint32_t)(z[s].maxKey >> 32), m = z[s].minP, B = (uint32_t)(z[s].nbKey >> 32);
                if (A == m) continue;                            // |Delta| <= DBL_EPSILON
                const double rr = (double)(A - B) / (double)(A - m);
                if (rr < ratio && A < (uint32_t)(z[t].maxKey >> 32)) {
                    z[s].target = t;
                    z[t].locked = 1;
                    z[s].locked = 1;
                    change = 1;
                }
            }
            sChange = change;
        }
        __syncthreads();
        if (!sChange) break;
        for (uint32_t k = threadIdx.x; k < cnt; k += NT) {
            const uint32_t p = ml[k];
            const uint32_t s = slot_of(fn[p]);
            if (s >= n) continue;
            const uint32_t t = z[s].target;
            if (t != s) fn[p] = rg[t].root + 1;
        }
        __syncthreads();
        for (uint32_t s = threadIdx.x; s < n; s += NT)
            if (z[s].alive && z[s].target != s) z[s].alive = 0;
        __syncthreads();
    }
    // One seed per surviving region, at its first maximum pixel, for the second watershed.
    for (uint32_t k = threadIdx.x; k < cnt; k += NT) fn[ml[k]] = 0u;
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < n; s += NT)
        if (z[s].alive && z[s].maxKey != 0) fn[kpix(z[s].maxKey)] = rg[s].root + 1;
}

// Second watershed: reset the per-component label range, then register the surviving seeds.
__global__ void __launch_bounds__(NT) pass2_reset_kernel(uint32_t* __restrict__ lab, PixAux* __restrict__ aux,
                                                         const uint32_t* __restrict__ maskList,
                                                         const uint32_t* __restrict__ maskCnt, size_t npx)
{
    const int f = blockIdx.y;
    const uint32_t cnt = maskCnt[f];
    for (uint32_t k = blockIdx.x * NT + threadIdx.x; k < cnt; k += gridDim.x * NT) {
        const uint32_t p = maskList[(size_t)f * npx + k];
        const uint32_t c = lab[(size_t)f * npx + p] & ~kCompRoot;
        aux[(size_t)f * npx + c].cmin = kNone;
        aux[(size_t)f * npx + c].cmax = 0u;
    }
}

__global__ void __launch_bounds__(NT) pass2_seed_kernel(const uint32_t* __restrict__ lab, PixAux* __restrict__ aux,
                                                        const Fz* __restrict__ fz, const Region* __restrict__ regions,
                                                        const uint32_t* __restrict__ rootCnt, size_t regionCap, size_t npx)
{
    const int f = blockIdx.y;
    const uint32_t n = (uint32_t)min((size_t)rootCnt[f], regionCap);
    for (uint32_t s = blockIdx.x * NT + threadIdx.x; s < n; s += gridDim.x * NT) {
        const Fz z = fz[(size_t)f * regionCap + s];
        if (!z.alive || z.maxKey == 0) continue;
        const uint32_t p = kpix(z.maxKey), L = regions[(size_t)f * regionCap + s].root + 1;
        const uint32_t c = lab[(size_t)f * npx + p] & ~kCompRoot;
        atomicMin(&aux[(size_t)f * npx + c].cmin, L);
        atomicMax(&aux[(size_t)f * npx + c].cmax, L);
    }
}

// Every frame through the exact sequential watershed (verification of that path).
__global__ void force_flag_kernel(uint32_t* __restrict__ counters, uint32_t* __restrict__ frameFlag,
                                  uint32_t* __restrict__ flagList, int F)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < F; i += gridDim.x * blockDim.x) {
        frameFlag[i] = 1u;
        flagList[i] = (uint32_t)i;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) counters[kFlagCnt] = (uint32_t)F;
}

// Between the two watershed passes: per-pass counters only. Frame flags stay set: a frame flagged in
// either pass is resolved (or deferred) as a whole.
__global__ void reset_pass_kernel(uint32_t* __restrict__ counters)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[kMultiCnt] = 0u;
        counters[kSlowCnt] = 0u;
        counters[kSlowNext] = 0u;
        counters[kFbNext] = 0u;
        counters[kRfNext] = 0u;
        counters[kWaveCnt] = 0u;
        counters[kGridCnt] = 0u;
        counters[kGridRetryCnt] = 0u;
    }
}

// ComputeStatistics sums (Segmentation2D.cpp:144-175) as exact integers: surface, sum P, sum P*x,
// sum P*y per region. The frame's regions (slots [0, rootCnt) in the order the labelling opened
// them) are accumulated in LDS by one workgroup per frame when they fit, else with global atomics
// by all the workgroups of the frame (grid: x workgroups per frame, y frames).
constexpr int RS_NT = 1024;
constexpr uint32_t RS_LDS_REGIONS = 2048;
constexpr uint32_t RS_RUN = 8;   // list entries per thread and run (global-atomic path)

__global__ void __launch_bounds__(RS_NT) region_stats_kernel(const uint32_t* __restrict__ P, const uint32_t* __restrict__ fin,
                                                             const PixAux* __restrict__ aux, const uint32_t* __restrict__ maskList,
                                                             const uint32_t* __restrict__ maskCnt, const uint32_t* __restrict__ rootCnt,
                                                             Region* __restrict__ regions, size_t regionCap,
                                                             uint32_t* __restrict__ counters, int W, size_t npx,
                                                             const uint32_t* __restrict__ skip)
{
    __shared__ unsigned long long sP[RS_LDS_REGIONS], sPx[RS_LDS_REGIONS], sPy[RS_LDS_REGIONS];
    __shared__ uint32_t sSurf[RS_LDS_REGIONS];
    uint32_t bx, fu;
    frame_tile(bx, fu);   // each XCD a contiguous range of frames
    const int f = (int)fu;
    if (skip && skip[f]) return;   // deferred frame: its labels are resolved in a later pass
    const uint32_t cnt = maskCnt[f];
    const uint32_t nReg = (uint32_t)min((size_t)rootCnt[f], regionCap);
    const bool inLds = nReg <= RS_LDS_REGIONS;
    // LDS accumulation: one workgroup per frame; global atomics: all of the frame's workgroups.
    if (inLds && bx != 0) return;
    Region* rg = regions + (size_t)f * regionCap;
    if (!inLds) {
        // Global atomics: each thread sums a run of RS_RUN consecutive list entries (neighbouring
        // pixels of one detection strip, mostly of one region) and adds a sum per change of region.
        uint32_t cur = 0xFFFFFFFFu, aS = 0;
        uint64_t aP = 0, aX = 0, aY = 0;
        auto flush = [&]() {
            if (cur == 0xFFFFFFFFu) return;
            Region* r = rg + cur;
            atomicAdd(&r->surf, aS);
            atomicAdd((unsigned long long*)&r->sP, (unsigned long long)aP);
            atomicAdd((unsigned long long*)&r->sPx, (unsigned long long)aX);
            atomicAdd((unsigned long long*)&r->sPy, (unsigned long long)aY);
        };
        const uint32_t* ml = maskList + (size_t)f * npx;
        for (uint32_t k0 = (bx * RS_NT + threadIdx.x) * RS_RUN; k0 < cnt; k0 += gridDim.x * RS_NT * RS_RUN) {
            const uint32_t k1 = min(k0 + RS_RUN, cnt);
            for (uint32_t k = k0; k < k1; ++k) {
                const uint32_t p = ml[k];
                const uint32_t L = fin[(size_t)f * npx + p];
                if (L == 0 || L > npx) { atomicOr(&counters[kStatus], 2u); continue; }
                const uint32_t slot = aux[(size_t)f * npx + (L - 1)].rslot;
                if (slot >= nReg) { atomicOr(&counters[kStatus], 2u); continue; }
                if (slot != cur) {
                    flush();
                    cur = slot;
                    aS = 0;
                    aP = aX = aY = 0;
                }
                const uint64_t v = P[(size_t)f * npx + p];
                const uint32_t y = p / W, x = p - y * W;
                aS += 1;
                aP += v;
                aX += v * x;
                aY += v * y;
            }
        }
        flush();
        return;
    }
    const uint32_t k0 = threadIdx.x;
    const uint32_t kStride = RS_NT;
    for (uint32_t r = threadIdx.x; r < nReg; r += RS_NT) { sP[r] = 0; sPx[r] = 0; sPy[r] = 0; sSurf[r] = 0; }
    __syncthreads();
    for (uint32_t k = k0; k < cnt; k += kStride) {
        const uint32_t p = maskList[(size_t)f * npx + k];
        const uint32_t L = fin[(size_t)f * npx + p];
        if (L == 0 || L > npx) { atomicOr(&counters[kStatus], 2u); continue; }
        const uint32_t slot = aux[(size_t)f * npx + (L - 1)].rslot;
        if (slot >= nReg) { atomicOr(&counters[kStatus], 2u); continue; }
        const uint64_t v = P[(size_t)f * npx + p];
        const uint32_t y = p / W, x = p - y * W;
        atomicAdd(&sSurf[slot], 1u);
        atomicAdd(&sP[slot], (unsigned long long)v);
        atomicAdd(&sPx[slot], (unsigned long long)(v * x));
        atomicAdd(&sPy[slot], (unsigned long long)(v * y));
    }
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < nReg; r += RS_NT) {
        rg[r].surf = sSurf[r];
        rg[r].sP = sP[r];
        rg[r].sPx = sPx[r];
        rg[r].sPy = sPy[r];
    }
}

__device__ inline uint32_t block_sum(uint32_t v, uint32_t* sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int j = 0; j < (int)(blockDim.x >> 6); ++j) t += sh[j];
    return t;
}

// FilterData (surface > VOLUME_MIN) and the per-frame capacity of Fitting::FitAll.
__global__ void __launch_bounds__(NT) region_count_kernel(const Region* __restrict__ regions, size_t regionCap,
                                                          const uint32_t* __restrict__ rootCnt, uint32_t* __restrict__ found,
                                                          uint32_t* __restrict__ keep, uint64_t maxPoints,
                                                          const uint32_t* __restrict__ skip)
{
    __shared__ uint32_t sh[NT / 64];
    const int f = blockIdx.x;
    const uint32_t n = (skip && skip[f]) ? 0u : (uint32_t)min((size_t)rootCnt[f], regionCap);
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < n; i += NT) c += (double)regions[(size_t)f * regionCap + i].surf > kVolumeMin;
    const uint32_t t = block_sum(c, sh);
    if (threadIdx.x == 0) {
        found[f] = t;
        keep[f] = (uint32_t)min((uint64_t)t, maxPoints);
    }
}

// Sum over one region of the frame in raster order, in FP64 as ComputeStatistics does (used when
// an exact integer sum reaches 2^53 and the reference's FP64 accumulation may have rounded).
__device__ void raster_sums(const uint32_t* Pf, const uint32_t* fn, int W, size_t npx, uint32_t L, double& sx,
                            double& sy, double& s)
{
    sx = sy = s = 0.0;
    for (size_t i = 0; i < npx; ++i) {
        if (fn[i] != L) continue;
        const double v = (double)Pf[i] * kPScale;
        const double y = (double)(i / W), x = (double)(i % W);
        sx = __dadd_rn(sx, __dmul_rn(v, x));
        sy = __dadd_rn(sy, __dmul_rn(v, y));
        s = __dadd_rn(s, v);
    }
}

__global__ void __launch_bounds__(NT) roi_emit_kernel(const uint32_t* __restrict__ P, const uint32_t* __restrict__ fin,
                                                      const Region* __restrict__ regions, size_t regionCap,
                                                      const uint32_t* __restrict__ rootCnt, const uint32_t* __restrict__ found,
                                                      const uint32_t* __restrict__ keep, const uint32_t* __restrict__ roiBase,
                                                      Roi* __restrict__ rois, int W, size_t npx)
{
    __shared__ uint32_t sWave[NT / 64];
    __shared__ uint32_t sBase;
    const int f = blockIdx.x;
    const uint32_t kf = keep[f];
    const uint32_t n = kf == 0 ? 0u : (uint32_t)min((size_t)rootCnt[f], regionCap);
    const bool truncate = found[f] > kf;
    const Region* rg = regions + (size_t)f * regionCap;
    if (threadIdx.x == 0) sBase = 0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t b0 = 0; b0 < n; b0 += NT) {
        const uint32_t i = b0 + threadIdx.x;
        Region r{};
        bool sel = false;
        if (i < n) {
            r = rg[i];
            sel = (double)r.surf > kVolumeMin;
        }
        uint32_t pos = 0;
        if (truncate) {
            // Keep the kf survivors with the smallest labels (first local maxima in raster order).
            if (sel) {
                uint32_t rank = 0;
                for (uint32_t j = 0; j < n; ++j) {
                    const Region q = rg[j];
                    rank += ((double)q.surf > kVolumeMin) && q.root < r.root;
                }
                pos = rank;
                sel = rank < kf;
            }
        } else {
            const uint64_t m = __ballot(sel);
            if (lane == 0) sWave[wv] = __popcll(m);
            __syncthreads();
            pos = sBase + __popcll(m & ((1ull << lane) - 1ull));
            for (int j = 0; j < wv; ++j) pos += sWave[j];
            __syncthreads();
            if (threadIdx.x == 0) for (int j = 0; j < NT / 64; ++j) sBase += sWave[j];
            __syncthreads();
        }
        if (sel) {
            double sx, sy, s;
            const double lim = 9007199254740992.0; // 2^53
            if ((double)r.sPx >= lim || (double)r.sPy >= lim || (double)r.sP >= lim) {
                raster_sums(P + (size_t)f * npx, fin + (size_t)f * npx, W, npx, r.root + 1, sx, sy, s);
            } else {
                sx = (double)r.sPx * kPScale;
                sy = (double)r.sPy * kPScale;
                s = (double)r.sP * kPScale;
            }
            Roi o;
            o.x = sx / s;
            o.y = sy / s;
            o.seg = s;
            o.surf = (double)r.surf;
            o.frame = (uint32_t)f;
            o.slot = pos;
            rois[roiBase[f] + pos] = o;
        }
    }
}

// Exclusive scan of n counts into out[0..n]; with run set, also *base = *run (the first ROI of this
// chunk in the call) and *run += total.
__global__ void __launch_bounds__(1024) exclusive_scan_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int n,
                                                              uint32_t* __restrict__ run, uint32_t* __restrict__ base)
{
    // Single workgroup: each thread scans a contiguous segment, then the segment totals.
    __shared__ uint32_t seg[1024];
    const int per = (n + 1023) / 1024;
    const int b = threadIdx.x * per, e = min(n, b + per);
    uint32_t s = 0;
    for (int i = b; i < e; ++i) s += in[i];
    seg[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < 1024; ++i) { const uint32_t t = seg[i]; seg[i] = acc; acc += t; }
        out[n] = acc;
        if (run) {
            *base = *run;
            *run += acc;
        }
    }
    __syncthreads();
    uint32_t acc = seg[threadIdx.x];
    for (int i = b; i < e; ++i) { const uint32_t t = in[i]; out[i] = acc; acc += t; }
}

} // namespace

void launch_exclusive_scan(const uint32_t* d_in, uint32_t* d_out, int n, hipStream_t s)
{
    exclusive_scan_kernel<<<1, 1024, 0, s>>>(d_in, d_out, n, nullptr, nullptr);
    PT_CHECK(hipGetLastError());
}

void launch_segment(const DetectBuffers& b, int W, int H, int F, uint64_t maxPoints, double ratio, Roi* d_rois,
                    hipStream_t s, bool forceSequential, int exactPath, bool resolve)
{
    const size_t npx = b.npx;
    const unsigned bpf = (unsigned)std::min<size_t>(64, std::max<size_t>(1, npx / 16384));
    const dim3 g(bpf, F);
    // Frames that failed the component-flood proof: resolved here (exact frame flood) or left to a
    // later pass that re-runs them with forceSequential (their rows are then skipped here).
    const uint32_t* skip = resolve ? nullptr : b.frameFlag;
    uint32_t ldsCap = 0, parentOff = 0;
    const size_t flLds = frame_label_lds(npx, ldsCap, parentOff);
    if (npx >= kHistSplitPx) {
        // One slice per 16384 mask-list entries the frame could hold at 1/16 of its pixels.
        const unsigned slices = (unsigned)std::clamp<size_t>(npx / (size_t(16384) * 16), 1, 64);
        mask_hist_partial_kernel<<<dim3(slices, F), MH_NT, 0, s>>>(b.P, b.maskList, b.maskCnt, b.hist, npx);
        mask_hist_suffix_kernel<<<F, MH_NT, 0, s>>>(b.hist);
    } else {
        mask_hist_kernel<<<F, MH_NT, 0, s>>>(b.P, b.maskList, b.maskCnt, b.hist, npx);
    }
    if (ldsCap > 0)
        frame_label_kernel<<<F, FL_NT, flLds, s>>>(b.P, b.seedBits, b.lab, b.fin, b.aux, b.maskList, b.maskCnt, b.seedList,
                                                   b.seedCnt, b.rootCnt, b.regions, b.regionCap, b.multiList, b.multiCap,
                                                   b.counters, b.frameFlag, b.flagList, W, H, npx, ldsCap, parentOff, b.hist);
    {
        // Frames beyond the LDS capacity (every frame when it is 0).
        const FlgArgs fa{b.P, b.seedBits, b.lab, b.fin, b.aux, b.maskList, b.maskCnt, b.seedList, b.seedCnt, b.rootCnt,
                         b.regions, b.regionCap, b.multiList, b.multiCap, b.counters, b.frameFlag, b.flagList, W, H, F,
                         npx, ldsCap};
        const unsigned bx = (unsigned)std::clamp<size_t>(npx / 32768, 1, 64);
        const dim3 gg(bx, (unsigned)std::clamp<size_t>(2048 / bx, 1, (size_t)F));
        flg_parents_kernel<<<gg, NT, 0, s>>>(fa);
        flg_union_kernel<<<gg, NT, 0, s>>>(fa);
        flg_flatten_kernel<<<gg, NT, 0, s>>>(fa);
        flg_maxima_kernel<<<gg, NT, 0, s>>>(fa);
        flg_labels_kernel<<<gg, NT, 0, s>>>(fa);
    }
    auto watershed = [&](int pass2) {
        if (forceSequential) force_flag_kernel<<<ceil_div(F, 256), 256, 0, s>>>(b.counters, b.frameFlag, b.flagList, F);
        if (pass2)
            label_assign_kernel<<<g, NT, 0, s>>>(b.lab, b.fin, b.aux, b.maskList, b.maskCnt, b.multiList, b.multiCap,
                                                 b.counters, b.frameFlag, b.flagList, npx);
        flood_kernel<<<4096, 64, 0, s>>>(b.P, b.lab, b.fin, b.seedList, b.seedCnt, b.maskCnt, b.multiList, b.multiCap,
                                         b.counters, b.frameFlag, b.flagList, W, H, npx, pass2 ? b.fz : nullptr,
                                         b.rootCnt, b.regionCap, b.hist, pass2 ? 0u : b.denseMask);
        if (!resolve) return;
        // Exact frame flood of the flagged frames: the grid rank flood, then the single-workgroup rank
        // flood for the frames it hands over, the wave-queue flood for the frames that one hands over,
        // and the heap restatement for frames with equal keys at a pop. exactPath 3 starts at the
        // single-workgroup rank flood.
        const bool rank = exactPath == 0 || exactPath == 3;
        if (exactPath == 0) launch_grid_flood(b, W, H, F, pass2 ? b.fz : nullptr, s);
        if (rank)
            rank_flood_kernel<<<(unsigned)b.rfSlots, RF_NT, 0, s>>>(b.P, b.seedBits, b.fin, b.maskList, b.maskCnt,
                                                                   exactPath == 0 ? b.gridList : b.flagList,
                                                                   exactPath == 0 ? kGridCnt : kFlagCnt, b.counters,
                                                                   b.waveList, b.rfScratch, W, H, npx,
                                                                   pass2 ? b.fz : nullptr, b.aux, b.regionCap);
        // Few frames reach the wave-queue flood when the rank flood runs first: a small persistent grid
        // (each workgroup takes a whole CU's LDS).
        const size_t fbGrid = std::max<size_t>(1, std::min<size_t>(F, rank ? 16 : 1024));
        fallback_fast_kernel<<<(unsigned)fbGrid, 64, FB_LDS_BYTES, s>>>(
            b.P, b.seedBits, b.fin, b.maskList, b.maskCnt, b.seedList, b.seedCnt, rank ? b.waveList : b.flagList,
            rank ? kWaveCnt : kFlagCnt, b.counters, b.slowList, W, H, npx, pass2 ? b.fz : nullptr, b.regions, b.rootCnt,
            b.aux, b.regionCap, exactPath == 2 ? 1 : 0);
        fallback_kernel<<<(unsigned)b.fbSlots, NT, 0, s>>>(b.P, b.seedBits, b.fin, b.maskList, b.maskCnt, b.slowList, b.counters,
                                                            b.fbScratch, W, H, npx, pass2 ? b.fz : nullptr, b.aux,
                                                            b.regionCap);
    };
    watershed(0);
    if ((int)ratio > 0) {   // Segmentation2D.cpp:82-84
        fusion_kernel<<<F, NT, 0, s>>>(b.P, b.fin, b.aux, b.maskList, b.maskCnt, b.regions, b.regionCap, b.rootCnt,
                                       b.fz, b.order, ratio, W, H, npx, skip);
        pass2_reset_kernel<<<g, NT, 0, s>>>(b.lab, b.aux, b.maskList, b.maskCnt, npx);
        pass2_seed_kernel<<<g, NT, 0, s>>>(b.lab, b.aux, b.fz, b.regions, b.rootCnt, b.regionCap, npx);
        reset_pass_kernel<<<1, 64, 0, s>>>(b.counters);
        watershed(1);
    }
    region_stats_kernel<<<dim3((unsigned)std::clamp<size_t>(npx / 131072, 1, 32), F), RS_NT, 0, s>>>(
        b.P, b.fin, b.aux, b.maskList, b.maskCnt, b.rootCnt, b.regions, b.regionCap, b.counters, W, npx, skip);
    region_count_kernel<<<F, NT, 0, s>>>(b.regions, b.regionCap, b.rootCnt, b.found, b.keep, maxPoints, skip);
    if (b.runTotal && b.scanWait) PT_CHECK(hipStreamWaitEvent(s, b.scanWait, 0));
    exclusive_scan_kernel<<<1, 1024, 0, s>>>(b.keep, b.roiBase, F, b.runTotal, b.chunkBase);
    if (b.runTotal && b.scanDone) PT_CHECK(hipEventRecord(b.scanDone, s));
    roi_emit_kernel<<<F, NT, 0, s>>>(b.P, b.fin, b.regions, b.regionCap, b.rootCnt, b.found, b.keep, b.roiBase, d_rois,
                                     W, npx);
    PT_CHECK(hipGetLastError());
}

} // namespace pt
