// Device-resident pipeline of the PALMTracer detection + fitting path (host side).
//
// Mirrors the reference's orchestration (lib/Palm.cpp Palm::Processing / FitLocalizations /
// AutoThreshold, lib/localization/FrameProcessor.cpp) over chunks of frames: every stage of a chunk
// is one launch over all its frames, and all intermediate state stays in HBM.
#pragma once
#include "pt_kernels.h"

#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace pt {

struct DevBuf;
void gather_rois(const Roi* src, uint32_t z0, Roi* dst, const uint32_t* roiBase, uint32_t frames, const uint32_t* run,
                 uint32_t cap, uint32_t* dBase, uint32_t* counters, hipStream_t s);
void set_u32(uint32_t* p, uint32_t v, hipStream_t s);

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n);
    void release();
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct RunStats {
    uint64_t frames = 0, chunks = 0, rois = 0, flaggedFrames = 0, multiComponents = 0, maskPixels = 0, deferredFrames = 0;
    uint64_t flagOverflow = 0, flagUnproven = 0;  // component floods given up: full queue / unproven order
    uint64_t floodPops = 0, floodMaxQueue = 0;    // wave-queue frame flood: pops, largest queue
    // Exact frame flood paths: frames handed by the rank flood to the wave-queue flood (too many
    // entries / equal keys / not settled), wave-queue overflows, frames for the heap restatement.
    uint64_t rfTooBig = 0, rfTie = 0, rfUnsettled = 0, fbOverflow = 0, heapFrames = 0;
    uint64_t rfTwoOrders = 0;  // frames the rank flood settled by solving both orders of an equal-key pair
    uint64_t gridFrames = 0, gridTies = 0;  // frames the grid rank flood resolved; of which with equal keys
    uint64_t gridUnsettled = 0, gridRejected = 0, gridOther = 0;  // frames it handed over, by reason
    uint32_t status = 0;
    // With Options::profile: device time per stage (HIP events on the engine stream, summed over
    // chunks), the device time of the deferred-frame batches on the side stream, and the LM passes
    // executed (model + Jacobian, trial model).
    double msAtrous = 0, msSegment = 0, msFit = 0, msAssemble = 0, msDeferred = 0;
    uint64_t jacPasses = 0, trialPasses = 0;            // per-lane fit kernel (FP32 or FP64)
    uint64_t groupJacPasses = 0, groupTrialPasses = 0;  // cooperative FP64 group kernel
    uint64_t fitHanded = 0;   // fast fit of the main lanes: ROIs handed from the FP32 kernel to the group kernel
    double msH2D = 0, msD2H = 0;   // host entry points: device time of the frame uploads and row downloads
    uint64_t streamPieces = 0;     // host entry points: pieces the stack went through the device in
};

struct Options {
    bool fitFp64 = true;           // FP64 model/Jacobian (reference arithmetic) or FP32
    int maxIter = kMaxIterDefault; // Fitting::N_ITER
    static constexpr size_t kDefaultChunkPixels = size_t(256) << 20;  // 1024 frames of 512 x 512
    size_t chunkPixels = kDefaultChunkPixels;
    bool profile = false;           // stage timing with HIP events + LM pass counters
    bool forceSequential = false;   // every frame through the exact frame flood
    int exactPath = 0;              // exact frame flood from: 0 grid rank flood, 1 wave-queue flood, 2 heap,
                                    // 3 single-workgroup rank flood
    unsigned fitWaves = 0;          // persistent FP32 fit: resident waves (0: 16 per CU)
    int fitRefill = 16;             // persistent FP32 fit: free lanes that trigger a refill round
    int fitHandover = 16;           // fast fit: FP32 iterations before a ROI moves to the FP64 group kernel (0: never)
    int fitObjective = 0;           // Gaussian fits: 0 least squares (reference), 1 Poisson likelihood
    double camGain = 1.0, camOffset = 0.0, camReadVar = 0.0;   // camera model of the Poisson likelihood
    int detectRows = 0;             // rows per detection wave (0: chosen from the chunk size)
    int lanes = 2;                  // main lanes the chunks alternate between (1: one stream, for timing)
    bool batchAutotune = true;      // batched localization: tune each new fit class's schedule
    bool batchGraph = true;         // batched localization: the fit stage as a captured hipGraph
    size_t streamPixels = size_t(1) << 27;   // host entry points: pixels per piece of the stack
    int hostThreads = 8;            // host entry points: threads of the copies to and from pinned staging
    // Frames with more mask pixels go to the exact frame flood without component floods (0: none).
    uint32_t denseFlagMask = 75000;
    int deferFlagged = 2;           // flagged frames of all but the last chunk on the side lane: 0 never,
                                    // 1 always, 2 when the last known chunk flagged at most 2 % of its frames
    // Defaults with the environment's overrides (PALMTRACER_FIT_PRECISION, PALMTRACER_MAX_ITER,
    // PALMTRACER_CHUNK_PIXELS).
    static Options defaults();
};

// Fit parameters of the Gaussian modes (fit 1-4, and 0 / above 5 as no fit) from fit_params
// ([roi, sigmaX, sigmaY, theta]; lib/fit/GaussianFit.cpp:39-45, the odd ROI side of
// lib/fit/Fitting.cpp:35-38) and the options.
FitParams gaussian_fit_params(uint64_t fit, const double* fp, const Options& opt);
// Per-plane 1-based index of each position in input order (lib/Palm.cpp:79-84) and the fit jobs of
// the positions (fit-only entry), on stream s; bad receives 1 when a position's plane is outside
// [1, planes].
void launch_positions_to_rois(const double* d_pos, uint64_t n, uint64_t planes, Roi* d_rois, uint32_t* d_bad,
                              hipStream_t s);
void launch_positions_ids(const double* d_pos, const uint32_t* d_local, uint64_t n, double* d_rows, hipStream_t s);

// One stack of a batched localization (the C ABI's pt_batch_stack).
struct BatchItem {
    const uint16_t* d_input;
    uint64_t planes, height, width;
    double threshold, watershedRatio;
    uint64_t fit;
    const double* fit_params;
    double* d_output;
    uint64_t outputCapDoubles, maxPerPlane;
    uint64_t doublesWritten;
};
// Schedule of one fit class in a batch: resident waves and refill threshold of the FP32 kernel,
// iterations before the FP64 hand-over; the autotuner's time for it.
struct BatchSchedule {
    unsigned waves = 4096;
    int refill = 16;
    int handover = 16;
    double ms = 0.0;
    bool tuned = false;
};
struct BatchClassInfo {
    int mode, roi;
    uint64_t jobs;
    BatchSchedule sched;
};

// One stream with its own detection buffers. The engine runs the chunks on the main lane and the
// frames deferred to the exact frame flood on the side lane, concurrently.
struct Lane {
    hipStream_t stream = nullptr;
    DetectBuffers b{};
    DevBuf P, lab, fin, aux, mask, seed, perFrame, regions, fz, order, multi, fb, counters, flag, rf, gpool, gmeta, sbits, hist;
    DevBuf rois, rows, staged, frameMap, fitNext, fitHand;
    GridFloodGraph gridGraph;    // the grid rank flood of this lane as a captured graph
    // P and the seed bitmap are zero everywhere between chunks: the detection kernel writes the mask
    // pixels and the local maxima only, and clear_mask() zeroes them once the chunk's segmentation is
    // queued. pDirty: a detection was queued whose pixels have not been cleared (then prepare() and
    // the next detection zero both buffers whole).
    bool pDirty = false;
    uint32_t* pinned = nullptr;  // host-visible: frame map, roiBase and counters of an asynchronous batch
    size_t pinnedWords = 0;
    void prepare(size_t F, size_t npx);
    void zero();
    void clear_mask(size_t F, size_t npx);   // zeroes what the last detection set in P and the seed bitmap
    // Work space of a fast fit of up to nRoi ROIs on this lane (queue words and hand-over list).
    FitScratch fit_scratch(size_t nRoi);
    size_t device_bytes() const;   // device memory held by the lane's buffers
    void release_buffers();        // frees the device buffers (the stream stays)
    uint32_t* host(size_t words);
    void release();   // device and pinned memory and the stream
};

class Engine {
public:
    // The process-wide engine of the C entry points (lives until the process ends).
    static Engine& get();
    // An independent engine (own streams and buffers), e.g. one per serving slot.
    Engine();
    // Options of the C entry points (pt_set_option; the environment's defaults until then). Every
    // engine a call leases starts from them.
    static Options shared_options();
    static void set_shared_options(const Options& o);
    // Counters of the call that finished last, whichever engine ran it (pt_last_stats, pt_stage_times).
    static void publish(const RunStats& s);
    static RunStats published();
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    std::mutex mu;
    Options opt;
    RunStats last;
    hipStream_t stream = nullptr;  // the main lane's stream

    // Localization over a device-resident stack; rows written to d_out (capacity outCapRows).
    // Returns the number of doubles written (a multiple of 18).
    uint64_t localize(const uint16_t* d_in, uint64_t planes, int H, int W, double thr, double ratio, uint64_t fit,
                      const double* fitParams, uint64_t maxPerPlane, double* d_out, uint64_t outCapRows);
    // Localization of many device stacks in one call: detection per stack, one fit launch per fit
    // class over the jobs of every stack (batch.hip).
    void batch_localize(BatchItem* items, uint64_t nItems);
    const std::vector<BatchClassInfo>& batch_classes() const { return batchLast_; }
    // Fit-only over caller positions (6 columns each), rows in input order.
    uint64_t fit_positions(const uint16_t* d_in, uint64_t planes, int H, int W, const double* hostPos, uint64_t n,
                           uint64_t fit, const double* fitParams, double* d_out);
    // The same two over host buffers, the stack streamed through the device in pieces of at most
    // opt.streamPixels pixels (pinned staging, uploads and downloads overlapping the computation).
    // Return doubles written, like localize / fit_positions.
    uint64_t localize_host(const uint16_t* input, double* output, uint64_t outputSize, int H, int W, uint64_t planes,
                           double thr, double ratio, uint64_t fit, const double* fitParams);
    uint64_t fit_positions_host(const uint16_t* input, uint64_t planes, int H, int W, const double* positions, uint64_t n,
                                uint64_t fit, const double* fitParams, double* output);
    void wavelet(const uint16_t* d_in, uint64_t planes, int H, int W, int level, double* d_out);
    // Label map of one frame after the watershed (labels = first-maximum raster index + 1).
    void debug_labels(const uint16_t* d_frame, int H, int W, double thr, double ratio, uint32_t* d_labels);
    // AutoThreshold of one device frame (lib/Palm.cpp:122-172).
    double autothreshold(const uint16_t* d_frame, int H, int W, const double* fitParams);
    // Frames per device chunk for frames of npx pixels.
    size_t chunk_frames(uint64_t planes, size_t npx) const;
    // Device bytes a Localization of `planes` frames of H x W allocates in this engine (its lanes'
    // chunk buffers, ROI stores and rows; the sizes Lane::prepare and localize use).
    size_t workspace_bytes(uint64_t planes, int H, int W, uint64_t maxPerPlane, uint64_t outCapRows) const;
    // Device memory the engine holds now, and a release of all of it (buffers grow again on demand).
    size_t device_bytes() const;
    void release_device_buffers();

    DevBuf in, out, planeInfo;

private:
    Lane main_, mainB_, side_;   // two alternating lanes for the chunks of a call, one for deferred frames
    hipEvent_t sideEv_[2] = {nullptr, nullptr};  // side batch start / done
    DevBuf atMask_, atPart_, atRows_, passes_, allRois_, allBase_, sideRois_, sideBase_;
    hipEvent_t fitEv_[3] = {nullptr, nullptr, nullptr};
    hipEvent_t chunkEv_[2] = {nullptr, nullptr};  // a main-lane chunk's counters are on the host
    hipEvent_t scanEv_[2] = {nullptr, nullptr};   // a chunk's ROI scan is done (orders the running total)
    DevBuf run_;                                    // ROI running total of the call (device)
    // Host entry points: staging of the streamed pieces (two each way), their copy stream and events.
    void stream_prepare(size_t inBytes, size_t outBytes);
    void stream_release();
    hipStream_t copyStream_ = nullptr;
    void* pinIn_[2] = {nullptr, nullptr};
    void* pinOut_[2] = {nullptr, nullptr};
    size_t pinInBytes_ = 0, pinOutBytes_ = 0;
    DevBuf streamIn_[2], streamOut_[2];
    hipEvent_t inEv_[2] = {}, outEv_[2] = {}, inStartEv_[2] = {}, outStartEv_[2] = {}, computeEv_[2] = {};
    std::vector<hipEvent_t> ev_;
    hipEvent_t event(size_t i);
    // Detection and segmentation of F frames queued on lane L's stream (chunk frame f = input frame
    // map[f], or the chunk's own frame f without a map); per-stage profiling events evBase..+2
    // unless evBase is kNoEvents.
    static constexpr size_t kNoEvents = ~size_t(0);
    void detect_chunk(Lane& L, const uint16_t* d_frames, size_t F, int H, int W, double thr, double ratio,
                      uint64_t maxPoints, size_t evBase, const uint32_t* d_map, bool resolve);
    void grow(DevBuf& buf, size_t need, size_t keep, hipStream_t s);
    // Fit parameters of a call; for mode 5 it parses the spline model and uploads its coefficients.
    FitParams fit_params(uint64_t fit, const double* fp);
    SplineParams spline_;
    DevBuf splineRaw_, splineCv_;
    // Batched localization: jobs, rows, stack table, per-plane bases, placement, counts, fit work space;
    // tuned schedules by class signature; the captured fit stage and the layout it was captured for.
    DevBuf bRois_, bRows_, bStaged_, bStacks_, bJobs_, bBases_, bPlan_, bCounts_, bHand_, bWords_;
    std::map<std::string, BatchSchedule> tuned_;
    hipGraphExec_t batchGraphExec_ = nullptr;
    std::string batchGraphKey_;
    hipEvent_t forkEv_[2] = {nullptr, nullptr};
    std::vector<BatchClassInfo> batchLast_;
};

// The engine of one call of a C entry point. Concurrent calls run independently: a call takes the
// process-wide engine when it is free, else a free pooled engine (own streams and buffers, created
// on first need, at most kPoolEngines of them, kept until the process ends), else it waits for the
// first of them to come free. The leased engine starts from the shared options; the counters of a
// call that computes them (stats = true) are published when the lease ends.
class Lease {
public:
    static constexpr int kPoolEngines = 4;
    explicit Lease(bool stats = true);
    ~Lease();
    Lease(const Lease&) = delete;
    Lease& operator=(const Lease&) = delete;
    Engine& operator*() const { return *e_; }
    Engine* operator->() const { return e_; }

private:
    Engine* e_ = nullptr;
    int slot_ = -1;   // pool slot, or -1 for the process-wide engine (its mutex is held)
    bool stats_ = true;
};

} // namespace pt
