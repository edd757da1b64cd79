// Device buffers and kernel launchers of the detection + fitting path.
#pragma once
#include "pt_common.h"

namespace pt {

// Per-region accumulator, one slot per root seed of a frame (exact integer sums of P = 65536*W1).
struct Region {
    uint32_t root;     // raster index of the region's first local maximum (its label identity)
    uint32_t surf;     // pixel count
    uint64_t sP;       // sum P
    uint64_t sPx;      // sum P * column
    uint64_t sPy;      // sum P * row
};

// One fit job: the segmentation info the fit reads (SegmentationInfo.hpp:32-71) and its frame.
struct Roi {
    double x, y, seg, surf;
    uint32_t frame;
    uint32_t slot;     // output row position within the frame (before the (Y, X) sort)
};

// Per-region state of the basin fusion (Segmentation2D.cpp:181-291, WSFusion), by region slot.
// Keys pack (P << 32) | (0xFFFFFFFF - pixel) so that an atomic maximum keeps the first pixel in
// raster order among equal potentials, as the reference's strict '>' updates do.
struct Fz {
    uint64_t maxKey;   // maximum potential and its first pixel
    uint64_t nbKey;    // maximum potential among pixels touching another region, and its first pixel
    uint32_t minP;     // minimum potential
    uint32_t target;   // slot this region merges into (itself if none)
    uint32_t locked;
    uint32_t alive;
    uint32_t nbLabel;  // label of the first other-region neighbour of the nbKey pixel (0: none)
    uint32_t pad;
};

// Per-pixel side record, used sparsely: union-find component data at component roots and the
// region slot at root seeds.
struct PixAux {
    uint32_t cmin, cmax;  // smallest / largest seed label in the component (labels = root seed + 1)
    uint32_t csize;       // component pixel count
    uint32_t rslot;       // region slot of a root seed
};

// The grid rank flood's launch sequence of one lane, captured once per watershed pass as a hipGraph
// and replayed while the lane's buffers and shapes stay the same (key).
struct GridFloodGraph {
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    uint64_t key[2][24] = {};
    void release();
};

// Buffers of one chunk of F frames of npx pixels. Per-pixel arrays are F*npx long; per-frame
// lists live in frame segments of capacity npx.
struct DetectBuffers {
    uint32_t* P = nullptr;        // 65536 * W1 where above threshold, else 0 (dense)
    uint32_t* lab = nullptr;      // union-find parent / component root
    uint32_t* fin = nullptr;      // watershed label (root seed index + 1)
    PixAux* aux = nullptr;
    uint32_t* maskList = nullptr; // per frame: pixels with P > 0
    uint32_t* seedList = nullptr; // per frame: local maxima
    uint32_t* maskCnt = nullptr;  // [F]
    uint32_t* seedCnt = nullptr;  // [F]
    uint32_t* rootCnt = nullptr;  // [F]
    uint32_t* keep = nullptr;     // [F] regions kept for fitting
    uint32_t* found = nullptr;    // [F] regions with surface > VOLUME_MIN (before truncation)
    uint32_t* roiBase = nullptr;  // [F+1] exclusive scan of keep
    uint32_t* frameFlag = nullptr;// [F] frame needs the exact sequential watershed
    Region* regions = nullptr;    // F * regionCap
    Fz* fz = nullptr;             // F * regionCap (basin fusion)
    uint32_t* order = nullptr;    // F * regionCap: slots in label order (basin fusion)
    uint64_t* multiList = nullptr;// (frame << 32 | component root) of multi-label components
    uint32_t* fbScratch = nullptr;// exact-fallback work space: fbSlots * fb_slot_words(npx) u32
    uint32_t* counters = nullptr; // [kNumCounters], see Counter
    uint32_t* flagList = nullptr; // [F] frames for the exact frame flood
    uint32_t* slowList = nullptr; // [F] frames for the heap restatement (equal keys met)
    uint32_t* waveList = nullptr; // [F] frames the rank flood could not settle (wave-queue flood)
    uint32_t* rfScratch = nullptr;// rank-flood work space: rfSlots * rf_slot_words(npx) u32
    uint32_t* gridList = nullptr; // [F] frames the grid rank flood hands to the single-workgroup floods
    uint32_t* gridRetry = nullptr;// [F] frames left to the grid rank flood's second round
    uint32_t* gridMeta = nullptr; // [F * kGridMetaWords] grid rank flood state per flagged frame
    uint32_t* gridPool = nullptr; // grid rank flood work space (gridPoolWords u32)
    size_t gridPoolWords = 0;
    GridFloodGraph* gridGraph = nullptr;   // optional: replay the grid rank flood from a captured graph
    uint32_t* seedBits = nullptr; // per frame: bitmap of the local maxima, seed_words(npx) u32
    uint32_t* hist = nullptr;     // per frame: kHistBins counts of mask pixels at or above each level
    uint32_t* runTotal = nullptr; // optional: ROIs of the call so far (shared by the lanes of a call)
    uint32_t* chunkBase = nullptr;// with runTotal: receives the first ROI of this chunk in the call
    hipEvent_t scanWait = nullptr;// with runTotal: the ROI scan waits for it (previous chunk's scan)
    hipEvent_t scanDone = nullptr;// with runTotal: recorded after this chunk's ROI scan
    size_t frames = 0, npx = 0, regionCap = 0, multiCap = 0, fbSlots = 0, rfSlots = 0;
    // Frames with more mask pixels than this go to the exact frame flood without component floods
    // (their component-flood proofs fail: the age bound grows with the mask); 0: no such limit.
    uint32_t denseMask = 0;
};

// Device counters of one chunk (u32 words).
enum Counter : int {
    kMultiCnt = 0, kStatus = 1, kSlowCnt = 2, kSlowNext = 3, kFlagCnt = 4, kFbNext = 5,
    kFlagOverflow = 6,   // component floods given up on a full queue
    kFlagUnproven = 7,   // component floods given up on an unproven order or equal keys
    kFloodPops = 8,      // pops of the exact frame flood (all frames of the launch)
    kFloodMaxQueue = 9,  // largest queue of the exact frame flood
    kRfNext = 10,        // rank flood: next flagged frame
    kWaveCnt = 11,       // frames handed from the rank flood to the wave-queue flood
    kRfTooBig = 12,      // rank flood: frames with more flood entries than rf_stride(npx)
    kRfTie = 13,         // rank flood: frames with two equal keys
    kRfUnsettled = 14,   // rank flood: frames whose iteration did not settle (or broke T < G)
    kFbOverflow = 15,    // wave-queue flood: frames whose queue overflowed
    kRfTwoOrders = 16,   // rank flood: frames settled by solving their equal-key pairs in both orders
    kGridCnt = 17,       // frames handed from the grid rank flood to the single-workgroup floods
    kGridOk = 18,        // frames resolved by the grid rank flood
    kGridTies = 19,      // of which frames whose solution has equal keys (proven not to matter)
    kGridUnsettled = 20, // grid rank flood: frames handed over unsettled after kGridMaxIter steps
    kGridRejected = 21,  // grid rank flood: frames handed over on equal keys that might matter
    kGridOther = 22,     // grid rank flood: frames handed over for lack of work space or T >= G
    kGridRetryCnt = 23,  // grid rank flood: frames left to its second round (work space)
    kMaskPx = 24,        // mask pixels of the chunk (P > 0), added when they are cleared
    kNumCounters = 25
};
// kStatus bits: 1 = region slots exhausted, 2 = unlabelled mask pixel (internal error).

// Levels of the per-frame potential histogram: P < 16 one bin each, then 16 bins per octave.
constexpr uint32_t kHistBins = 512;
// Frames of at least this many pixels build their mask histogram over several workgroups (the
// histogram is zeroed with the chunk counters).
constexpr size_t kHistSplitPx = size_t(1) << 20;
__host__ __device__ inline uint32_t hist_bin(uint32_t v)
{
    if (v < 16u) return v;
    const uint32_t msb = 31u - (uint32_t)__builtin_clz(v);
    return 16u * (msb - 3u) + ((v >> (msb - 4u)) & 15u);
}
__host__ __device__ inline size_t fb_slot_words(size_t npx) { return 5 * ((npx + 1) & ~size_t(1)); }
__host__ __device__ inline size_t seed_words(size_t npx) { return (npx + 31) / 32; }
// Rank flood: frames with at most rf_stride(npx) pixels in the flood (mask pixels but the seeds that
// are never pushed; at most kRankMax); their sort keys are sorted in LDS up to kRankLds entries, in
// the work space beyond. Work space per slot: pixel -> entry map, then 25 words per entry. The stride
// is a power of two, so the padded sort of any admitted frame fits.
constexpr uint32_t kRankMax = 262144;
// Grid rank flood: fixed-point steps, interval steps (frames with equal keys), state words per
// flagged frame.
constexpr int kGridMaxIter = 6;
constexpr int kGridIvalIter = 4;
constexpr size_t kGridMetaWords = 32;
constexpr uint32_t kRankLds = 16384;
__host__ __device__ inline size_t rf_stride(size_t npx)
{
    size_t r = kRankLds;
    while (r < npx && r < kRankMax) r <<= 1;
    return r;
}
__host__ __device__ inline size_t rf_slot_words(size_t npx) { return ((npx + 1) & ~size_t(1)) + 25 * rf_stride(npx); }

// A trous
void launch_atrous_plane(const uint16_t* d_in, double* d_out, int W, int H, int planes, int level, hipStream_t s);
// frameMap (optional): chunk frame f reads input frame frameMap[f] (else f). segRows: rows per wave
// (0: chosen from the number of frames).
void launch_atrous_detect(const uint16_t* d_in, const DetectBuffers& b, int W, int H, int frames, int64_t tmin,
                          hipStream_t s, const uint32_t* frameMap = nullptr, int segRows = 0);
int64_t threshold_to_pmin(double thr);

// Segmentation (ratio <= 0: watershed without fusion). Produces ROIs for frames [0, F) at
// d_rois + roiBase[f], with per-frame truncation to maxPoints in label order.
// forceSequential sends every frame through the exact frame flood. That flood is the rank flood,
// then, for frames it cannot settle, the wave-queue flood, then, for frames with equal keys at a
// pop, the libstdc++-order heap restatement; exactPath = 1 starts at the wave-queue flood, 2 at the
// heap restatement. With resolve = false, frames whose component floods could not be proven are
// left flagged (frameFlag / flagList) without rows, for a later pass over them with forceSequential.
// Grid rank flood of the flagged frames (fz: second watershed); frames it cannot prove go to
// b.gridList (count in counters[kGridCnt]). grid_pool_words: pool size for chunks of F frames.
void launch_grid_flood(const DetectBuffers& b, int W, int H, int F, const Fz* fz, hipStream_t s);
size_t grid_pool_words(size_t F, size_t npx);
void launch_segment(const DetectBuffers& b, int W, int H, int frames, uint64_t maxPoints, double ratio, Roi* d_rois,
                    hipStream_t s, bool forceSequential = false, int exactPath = 0, bool resolve = true);

// Spline PSF model of fit mode 5 (SplineFit.cpp:36-49): sizes, offsets and the device copy of the
// coefficients in voxel-major order (cv[voxel * 64 + monomial]; the reference's array is
// monomial-major).
struct SplineParams {
    uint64_t roi = 0, half = 0, off = 0, sx = 0, sy = 0, sz = 0;
    double initZ = 0, deltaZ = 0, ratio = 1, roiParam = 0;
    const double* cv = nullptr;
};

// One stack of a batched fit: its frames (frame f at base + f * npx, rows of W pixels).
struct StackRef {
    const uint16_t* base;
    uint64_t npx;
    int W, H;
};

// Gaussian LM fit of nRoi ROIs (count read from device memory *d_count when non-null).
struct FitParams {
    int mode;        // 1..4 Gaussian, 5 spline (0: no fit, rows carry the segmentation only)
    int roi;         // odd ROI side
    int maxIter;     // Fitting::N_ITER (75 in the reference)
    double sigmaX, sigmaY, theta;
    int refill = 16; // FP32 fit: free lanes of a wave that trigger a refill round while others work
    // Fast fit: iterations after which a ROI still iterating leaves the FP32 kernel for the FP64 group
    // kernel (0: no hand-over, every ROI and every mode stays in FP32).
    int handover = 16;
    const SplineParams* spline = nullptr;  // mode 5 (host struct, read by the launcher)
    // Batched fits (device table, optional): ROI r reads stack stacks[r.slot], frame r.frame; the
    // launch's frames / npx / W / H are then unused.
    const StackRef* stacks = nullptr;
    // Objective of modes 1-4: 0 least squares (the reference's), 1 Poisson likelihood with the camera
    // model ADU = camOffset + camGain * photons (+ read noise of variance camReadVar, ADU^2).
    int objective = 0;
    double camGain = 1.0, camOffset = 0.0, camReadVar = 0.0;
};
// Device work space of a fast fit launch: kFitWords queue words ([0] FP32 queue head, [1] ROIs handed
// to the FP64 group kernel, [2] group queue head) and the hand-over list (one word per ROI).
constexpr int kFitWords = 4;
struct FitScratch {
    unsigned int* words = nullptr;
    uint32_t* hand = nullptr;
};
// LM fit of nRoi ROIs into d_rows (18 columns each). fp64: the reference's arithmetic and operation
// order, one ROI per lane; otherwise the fast fit: the FP32 persistent work-queue kernel, which hands
// ROIs that are still iterating after fp.handover iterations (or that meet a decision FP32 cannot
// take) to the cooperative FP64 group kernel, and the group kernel alone for fit mode 4. ws: the
// fast fit's work space (nRoi hand-over words). d_passes (optional, 4 words) accumulates {Jacobian
// passes, trial passes} of the per-lane kernel, then of the group kernel; frameMap maps ROI frames
// to stack frames. With d_count set, the ROI count is read on the device (nRoi is then an upper bound).
// Poisson-likelihood fit of modes 1-4 (FP64, one ROI per lane; fp.objective == 1).
void launch_fit_mle(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, uint32_t nRoi, const FitParams& fp,
                    double* d_rows, hipStream_t s, const uint32_t* frameMap, const uint32_t* d_count);
void launch_fit(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, uint32_t nRoi, const FitParams& fp,
                double* d_rows, bool fp64, const FitScratch& ws, hipStream_t s, unsigned long long* d_passes = nullptr,
                const uint32_t* frameMap = nullptr, unsigned waves = 0, const uint32_t* d_count = nullptr);
// Cooperative FP64 LM fit (eight lanes per ROI) of the ROIs d_list[0 .. *d_count) (d_list null: ROIs
// 0 .. *d_count, or 0 .. nMax without d_count); d_next: the group queue head (zeroed here).
void launch_fit_group(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, const uint32_t* d_list,
                      const uint32_t* d_count, uint32_t nMax, const FitParams& fp, double* d_rows, unsigned int* d_next,
                      hipStream_t s, unsigned long long* d_passes, const uint32_t* frameMap, unsigned waves);
void launch_fit_f64(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, uint32_t nRoi,
                    const FitParams& fp, double* d_rows, hipStream_t s, unsigned long long* d_passes,
                    const uint32_t* frameMap, const uint32_t* d_count = nullptr);

// Spline fit (mode 5) of nRoi ROIs, one per lane in FP64 (count read on the device when d_count is set).
void launch_spline_fit(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, uint32_t nRoi,
                       const SplineParams& sp, double* d_rows, hipStream_t s, const uint32_t* frameMap,
                       const uint32_t* d_count);
// Voxel-major copy of the reference's coefficient array (64 monomials x ssize voxels).
void launch_spline_transpose(const double* d_raw, double* d_cv, uint64_t ssize, hipStream_t s);

// Per-frame (Y, X) sort (rows of frame f at d_rows + roiBase[f]) into d_staged with the same layout;
// Index is set, Id and Plane are left for launch_scatter. maxRows: the most rows a frame can hold;
// rows: the rows at d_rows (roiBase[frames] <= rows). When maxRows is above 2048 the sort needs
// d_scratch of kSortScratchPerRow bytes per row (8-byte aligned); callers lend the ROI store, which
// the fit has read by then (a ROI record is at least that large).
constexpr size_t kSortScratchPerRow = 40;
void launch_sort_frames(const double* d_rows, const uint32_t* d_roiBase, int frames, double* d_staged, hipStream_t s,
                        uint64_t maxRows, uint64_t rows, void* d_scratch);
// Placement plan on the device: per plane p, its staged first row src[p] (main rows from the plane
// bases base[0..planes]; the nDef deferred planes from d_def = {planes, first side rows, counts},
// offset by sideOff), its row count and its output row dst[p] (prefix of the counts); *d_total the
// number of rows.
void launch_plan(const uint32_t* d_base, uint64_t planes, const uint64_t* d_def, uint32_t nDef, uint64_t sideOff,
                 uint64_t* d_src, uint64_t* d_dst, uint32_t* d_count, uint64_t* d_total, hipStream_t s);
// Final placement: for each plane p, count[p] rows from staged row src[p] to output row dst[p],
// with Id = dst[p] + k + 1 and Plane = p + 1; output rows from cap on are not written.
void launch_scatter(const double* d_staged, const uint64_t* d_src, const uint64_t* d_dst, const uint32_t* d_count,
                    uint64_t planes, double* d_out, uint64_t cap, hipStream_t s);

void launch_exclusive_scan(const uint32_t* d_in, uint32_t* d_out, int n, hipStream_t s);
// AstigFit::Estimate of n (sigma_x, sigma_y) pairs (d_in, 2n doubles) into (z, score) pairs (d_out);
// model: 10 host doubles.
void launch_astig_estimate(const double* d_in, double* d_out, uint64_t n, double pixelSize, const double* model,
                           double zMax, hipStream_t s);

} // namespace pt
