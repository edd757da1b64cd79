// Host orchestration of the device pipeline. Reference: lib/Palm.cpp, lib/localization/FrameProcessor.cpp.
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace pt {

namespace {
__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Zeroes three word ranges in one launch.
__global__ void zero3_kernel(uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t* c, size_t nc)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < na + nb + nc; i += stride) {
        if (i < na) a[i] = 0u;
        else if (i < na + nb) b[i - na] = 0u;
        else c[i - na - nb] = 0u;
    }
}

// Zeroes P at the mask pixels and the seed-bitmap words of the local maxima of frames 0..F-1 (frames
// strided in y, list entries in x), so that both are zero everywhere for the lane's next detection.
// The number of mask pixels is added to counters[kMaskPx].
__global__ void clear_mask_kernel(uint32_t* __restrict__ P, const uint32_t* __restrict__ maskList,
                                  const uint32_t* __restrict__ maskCnt, uint32_t* __restrict__ seedBits,
                                  const uint32_t* __restrict__ seedList, const uint32_t* __restrict__ seedCnt, uint32_t F,
                                  size_t npx, uint32_t* __restrict__ counters)
{
    const size_t sw = seed_words(npx);
    for (uint32_t f = blockIdx.y; f < F; f += gridDim.y) {
        const uint32_t n = maskCnt[f], ns = seedCnt[f];
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counters[kMaskPx], n);
        const uint32_t* ml = maskList + (size_t)f * npx;
        uint32_t* Pf = P + (size_t)f * npx;
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) Pf[ml[k]] = 0u;
        const uint32_t* sl = seedList + (size_t)f * npx;
        uint32_t* sbf = seedBits + (size_t)f * sw;
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ns; k += gridDim.x * blockDim.x) sbf[sl[k] >> 5] = 0u;
    }
}

// Fit jobs from caller rows [Id, Plane, X, Y, SegIntensity, Surface] (lib/Palm.cpp:105-109).
__global__ void positions_to_rois(const double* __restrict__ pos, uint64_t n, uint64_t planes, Roi* __restrict__ rois,
                                  uint32_t* __restrict__ bad)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double* p = pos + i * kInCols;
        Roi r;
        r.x = p[kInX];
        r.y = p[kInY];
        r.seg = p[kInSegIntensity];
        r.surf = p[kInSurface];
        const double pl = p[kInPlane];
        const bool ok = pl >= 1.0 && pl < (double)planes + 1.0;
        r.frame = ok ? (uint32_t)((uint64_t)pl - 1) : 0u;
        if (!ok) { r.x = -2.0; atomicOr(bad, 1u); }  // outside the ROI border test: sentinel row
        r.slot = 0;
        rois[i] = r;
    }
}

// ROIs of a chunk appended to the call's ROI store with absolute frame numbers, and the chunk's
// per-frame ROI offsets shifted to the store. *base is the chunk's first ROI in the store (set by
// the chunk's ROI scan); ROIs beyond the store's capacity set status bit 4.
__global__ void gather_rois_kernel(const Roi* __restrict__ src, uint32_t z0, Roi* __restrict__ dst,
                                   const uint32_t* __restrict__ roiBase, uint32_t frames, const uint32_t* __restrict__ base,
                                   uint32_t cap, uint32_t* __restrict__ dBase, uint32_t* __restrict__ counters)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    const uint32_t off = *base, n = roiBase[frames];
    if (t == 0 && (uint64_t)off + n > cap) atomicOr(&counters[kStatus], 4u);
    const uint32_t m = off >= cap ? 0u : min(n, cap - off);
    for (uint32_t i = t; i < m; i += stride) {
        Roi r = src[i];
        r.frame += z0;
        dst[off + i] = r;
    }
    for (uint32_t k = t; k < frames; k += stride) dBase[k] = off + roiBase[k];
}

// ROIs of a side-lane batch appended to the side ROI store with their stack frame numbers.
__global__ void side_gather_kernel(const Roi* __restrict__ src, const uint32_t* __restrict__ n, const uint32_t* __restrict__ map,
                                   Roi* __restrict__ dst)
{
    const uint32_t m = *n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        Roi r = src[i];
        r.frame = map[r.frame];
        dst[i] = r;
    }
}

__global__ void masked_labels(const uint32_t* __restrict__ fin, const uint32_t* __restrict__ P, uint32_t* __restrict__ out,
                              size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = P[i] ? fin[i] : 0u;
}

__global__ void positions_ids(const double* __restrict__ pos, const uint32_t* __restrict__ local, uint64_t n,
                              double* __restrict__ rows)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        rows[i * kCols + kId] = pos[i * kInCols + kInId];
        rows[i * kCols + kPlane] = pos[i * kInCols + kInPlane];
        rows[i * kCols + kIndex] = (double)local[i];
    }
}
} // namespace

void launch_positions_to_rois(const double* d_pos, uint64_t n, uint64_t planes, Roi* d_rois, uint32_t* d_bad,
                              hipStream_t s)
{
    const unsigned g = std::max<uint64_t>(1, std::min<uint64_t>(4096, (n + 255) / 256));
    positions_to_rois<<<g, 256, 0, s>>>(d_pos, n, planes, d_rois, d_bad);
    PT_CHECK(hipGetLastError());
}

void launch_positions_ids(const double* d_pos, const uint32_t* d_local, uint64_t n, double* d_rows, hipStream_t s)
{
    const unsigned g = std::max<uint64_t>(1, std::min<uint64_t>(4096, (n + 255) / 256));
    positions_ids<<<g, 256, 0, s>>>(d_pos, d_local, n, d_rows);
    PT_CHECK(hipGetLastError());
}

void gather_rois(const Roi* src, uint32_t z0, Roi* dst, const uint32_t* roiBase, uint32_t frames, const uint32_t* base,
                 uint32_t cap, uint32_t* dBase, uint32_t* counters, hipStream_t s)
{
    gather_rois_kernel<<<1024, 256, 0, s>>>(src, z0, dst, roiBase, frames, base, cap, dBase, counters);
    PT_CHECK(hipGetLastError());
}

void set_u32(uint32_t* p, uint32_t v, hipStream_t s)
{
    fill_u32_kernel<<<1, 64, 0, s>>>(p, 1, v);
    PT_CHECK(hipGetLastError());
}

void DevBuf::ensure(size_t n)
{
    if (n <= bytes) return;
    release();
    n = std::max<size_t>(n, 256);
    PT_CHECK(hipMalloc(&p, n));
    bytes = n;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

Engine& Engine::get()
{
    static Engine* e = new Engine();  // lives until process exit
    return *e;
}

namespace {
struct Shared {
    std::mutex optMu, statMu, poolMu;
    Options opt = Options::defaults();
    RunStats stat;
    std::condition_variable poolCv;
    Engine* pool[Lease::kPoolEngines] = {};
    bool busy[Lease::kPoolEngines] = {};
};
Shared& shared()
{
    static Shared* s = new Shared();   // lives until process exit, like the engines it lists
    return *s;
}
} // namespace

Options Engine::shared_options()
{
    Shared& sh = shared();
    std::lock_guard<std::mutex> lk(sh.optMu);
    return sh.opt;
}

void Engine::set_shared_options(const Options& o)
{
    Shared& sh = shared();
    std::lock_guard<std::mutex> lk(sh.optMu);
    sh.opt = o;
}

void Engine::publish(const RunStats& s)
{
    Shared& sh = shared();
    std::lock_guard<std::mutex> lk(sh.statMu);
    sh.stat = s;
}

RunStats Engine::published()
{
    Shared& sh = shared();
    std::lock_guard<std::mutex> lk(sh.statMu);
    return sh.stat;
}

Lease::Lease(bool stats) : stats_(stats)
{
    Engine& g = Engine::get();
    Shared& sh = shared();
    bool create = false;
    for (;;) {
        if (g.mu.try_lock()) {
            e_ = &g;
            break;
        }
        std::unique_lock<std::mutex> lk(sh.poolMu);
        int pick = -1;
        for (int i = 0; i < Lease::kPoolEngines && pick < 0; ++i)   // an engine that exists already first
            if (!sh.busy[i] && sh.pool[i]) pick = i;
        for (int i = 0; i < Lease::kPoolEngines && pick < 0; ++i)
            if (!sh.busy[i]) pick = i;
        if (pick >= 0) {
            sh.busy[pick] = true;
            slot_ = pick;
            e_ = sh.pool[pick];
            create = e_ == nullptr;
            break;
        }
        // The process-wide engine's mutex is also taken by calls that do not lease (options, batches),
        // which do not notify: wake now and then to try it again.
        sh.poolCv.wait_for(lk, std::chrono::milliseconds(2));
    }
    if (create) {
        try {
            e_ = new Engine();
        } catch (...) {
            std::lock_guard<std::mutex> lk(sh.poolMu);
            sh.busy[slot_] = false;
            sh.poolCv.notify_all();
            throw;
        }
        std::lock_guard<std::mutex> lk(sh.poolMu);
        sh.pool[slot_] = e_;
    }
    e_->opt = Engine::shared_options();
}

Lease::~Lease()
{
    if (stats_) Engine::publish(e_->last);
    Shared& sh = shared();
    if (slot_ < 0) {
        e_->mu.unlock();
    } else {
        std::lock_guard<std::mutex> lk(sh.poolMu);
        sh.busy[slot_] = false;
    }
    sh.poolCv.notify_all();
}

Options Options::defaults()
{
    Options o;
    if (const char* v = std::getenv("PALMTRACER_FIT_PRECISION")) o.fitFp64 = std::strcmp(v, "fp32") != 0;
    if (const char* v = std::getenv("PALMTRACER_MAX_ITER")) o.maxIter = std::max(1, std::atoi(v));
    if (const char* v = std::getenv("PALMTRACER_CHUNK_PIXELS")) o.chunkPixels = std::max<size_t>(1, std::strtoull(v, nullptr, 10));
    return o;
}

Engine::Engine()
{
    PT_CHECK(hipStreamCreateWithFlags(&main_.stream, hipStreamNonBlocking));
    PT_CHECK(hipStreamCreateWithFlags(&mainB_.stream, hipStreamNonBlocking));
    for (hipEvent_t& e : scanEv_) PT_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    PT_CHECK(hipStreamCreateWithFlags(&side_.stream, hipStreamNonBlocking));
    stream = main_.stream;
    for (hipEvent_t& e : sideEv_) PT_CHECK(hipEventCreate(&e));
    for (hipEvent_t& e : chunkEv_) PT_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (hipEvent_t& e : fitEv_) PT_CHECK(hipEventCreate(&e));
    for (hipEvent_t& e : forkEv_) PT_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    opt = Options::defaults();
}

Engine::~Engine()
{
    for (Lane* L : {&main_, &mainB_, &side_})
        if (L->stream) (void)hipStreamSynchronize(L->stream);
    for (Lane* L : {&main_, &mainB_, &side_}) L->release();
    stream_release();
    for (DevBuf* b : {&in, &out, &planeInfo, &atMask_, &atPart_, &atRows_, &passes_, &allRois_, &allBase_, &sideRois_, &sideBase_,
                      &run_, &splineRaw_, &splineCv_})
        b->release();
    for (hipEvent_t e : sideEv_) (void)hipEventDestroy(e);
    for (hipEvent_t e : fitEv_) (void)hipEventDestroy(e);
    for (hipEvent_t e : chunkEv_) (void)hipEventDestroy(e);
    for (hipEvent_t e : scanEv_) (void)hipEventDestroy(e);
    for (hipEvent_t e : forkEv_) (void)hipEventDestroy(e);
    if (batchGraphExec_) (void)hipGraphExecDestroy(batchGraphExec_);
    for (DevBuf* b : {&bRois_, &bRows_, &bStaged_, &bStacks_, &bJobs_, &bBases_, &bPlan_, &bCounts_, &bHand_, &bWords_})
        b->release();
    for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
}

void Lane::release()
{
    gridGraph.release();
    for (DevBuf* d : {&P, &lab, &fin, &aux, &mask, &seed, &perFrame, &regions, &fz, &order, &multi, &fb, &counters, &flag,
                      &rf, &gpool, &gmeta, &sbits, &hist, &rois, &rows, &staged, &frameMap, &fitNext, &fitHand})
        d->release();
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    pinnedWords = 0;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

namespace {
// Device bytes of an allocation of v bytes: whole 2 MiB pages.
size_t alloc_pages(size_t v)
{
    constexpr size_t kPage = size_t(2) << 20;
    return (std::max<size_t>(v, 1) + kPage - 1) / kPage * kPage;
}

// Sizes of a lane's buffers for chunks of F frames of npx pixels (Lane::prepare allocates them;
// Engine::workspace_bytes adds them up for the serving loop's HBM budget).
struct LaneSizes {
    size_t n, regionCap, multiCap, fbSlots, rfSlots, gridWords;
    LaneSizes(size_t F, size_t npx)
    {
        n = F * npx;
        regionCap = npx / 4 + 2;  // roots are pairwise non-adjacent: at most ceil(H/2)*ceil(W/2)
        multiCap = n / 64 + 1024;
        // Heap restatement slots (frames the other exact floods hand over, rare): about 16M pixels.
        fbSlots = std::clamp<size_t>((size_t(1) << 24) / std::max<size_t>(npx, 1), 2, 16);
        // Single-workgroup rank-flood slots (frames the grid rank flood hands over): up to 128 frames at
        // a time, within 256 MiB of work space.
        rfSlots = std::clamp<size_t>(std::min<size_t>(F, (size_t(1) << 28) / (rf_slot_words(npx) * 4)), 1, 128);
        gridWords = grid_pool_words(F, npx);
    }
    // Device bytes of a prepared lane plus its ROI buffer, each buffer counted in whole 2 MiB pages
    // (the granularity of device allocations).
    size_t bytes(size_t F, size_t npx) const
    {
        auto b = alloc_pages;
        return b(n * 4) * 5 + b(n * sizeof(PixAux)) + b((F + 1) * 4 * 9) + b(F * regionCap * sizeof(Region)) +
               b(F * regionCap * sizeof(Fz)) + b(F * regionCap * 4) + b(multiCap * 8) + b(fbSlots * fb_slot_words(npx) * 4) +
               b(kNumCounters * 4) + b((F + 1) * 20) + b(rfSlots * rf_slot_words(npx) * 4) + b(F * seed_words(npx) * 4) +
               b(F * kHistBins * 4) + b(F * regionCap * sizeof(Roi)) + b(gridWords * 4) + b(F * kGridMetaWords * 4);
    }
};
} // namespace

void Lane::prepare(size_t F, size_t npx)
{
    const LaneSizes z(F, npx);
    const size_t n = z.n;
    void* oldP = P.p;
    void* oldBits = sbits.p;
    P.ensure(n * 4);
    sbits.ensure(F * seed_words(npx) * 4);
    if (P.p != oldP || sbits.p != oldBits || pDirty) {
        // New buffers, or buffers whose last detection was not cleared: zero them whole.
        PT_CHECK(hipMemsetAsync(P.p, 0, P.bytes, stream));
        PT_CHECK(hipMemsetAsync(sbits.p, 0, sbits.bytes, stream));
        pDirty = false;
    }
    lab.ensure(n * 4);
    fin.ensure(n * 4);
    aux.ensure(n * sizeof(PixAux));
    mask.ensure(n * 4);
    seed.ensure(n * 4);
    perFrame.ensure((F + 1) * 4 * 9);
    const size_t regionCap = z.regionCap;
    regions.ensure(F * regionCap * sizeof(Region));
    fz.ensure(F * regionCap * sizeof(Fz));
    order.ensure(F * regionCap * 4);
    const size_t multiCap = z.multiCap;
    multi.ensure(multiCap * 8);
    const size_t fbSlots = z.fbSlots;
    fb.ensure(fbSlots * fb_slot_words(npx) * 4);
    counters.ensure(kNumCounters * 4);
    flag.ensure((F + 1) * 20);
    const size_t rfSlots = z.rfSlots;
    rf.ensure(rfSlots * rf_slot_words(npx) * 4);
    gpool.ensure(z.gridWords * 4);
    gmeta.ensure(F * kGridMetaWords * 4);
    hist.ensure(F * kHistBins * 4);
    uint32_t* pf = perFrame.as<uint32_t>();
    const size_t s = F + 1;
    b.P = P.as<uint32_t>();
    b.lab = lab.as<uint32_t>();
    b.fin = fin.as<uint32_t>();
    b.aux = aux.as<PixAux>();
    b.maskList = mask.as<uint32_t>();
    b.seedList = seed.as<uint32_t>();
    b.maskCnt = pf;
    b.seedCnt = pf + s;
    b.rootCnt = pf + 2 * s;
    b.keep = pf + 3 * s;
    b.found = pf + 4 * s;
    b.frameFlag = pf + 5 * s;
    b.roiBase = pf + 6 * s;
    b.chunkBase = pf + 7 * s;
    b.regions = regions.as<Region>();
    b.fz = fz.as<Fz>();
    b.order = order.as<uint32_t>();
    b.multiList = multi.as<uint64_t>();
    b.fbScratch = fb.as<uint32_t>();
    b.counters = counters.as<uint32_t>();
    b.flagList = flag.as<uint32_t>();
    b.slowList = flag.as<uint32_t>() + (F + 1);
    b.waveList = flag.as<uint32_t>() + 2 * (F + 1);
    b.gridList = flag.as<uint32_t>() + 3 * (F + 1);
    b.gridRetry = flag.as<uint32_t>() + 4 * (F + 1);
    b.gridMeta = gmeta.as<uint32_t>();
    b.gridPool = gpool.as<uint32_t>();
    b.gridPoolWords = z.gridWords;
    b.gridGraph = &gridGraph;
    b.rfScratch = rf.as<uint32_t>();
    b.rfSlots = rfSlots;
    b.seedBits = sbits.as<uint32_t>();
    b.hist = hist.as<uint32_t>();
    b.frames = F;
    b.npx = npx;
    b.regionCap = regionCap;
    b.multiCap = multiCap;
    b.fbSlots = fbSlots;
}

void Lane::zero()
{
    // Per-frame counters (6 arrays laid out with a stride of frames + 1 words, frames being the
    // chunk capacity, whatever the number of frames in this chunk) and the chunk counters; a fill
    // kernel rather than a memset so that the sequence can be captured in a graph. (P and the seed
    // bitmap are zero already: clear_mask() zeroes what a detection set.)
    // Frames of at least kHistSplitPx pixels also start with a zero mask histogram (several workgroups
    // add into it).
    const bool hz = b.npx >= kHistSplitPx;
    zero3_kernel<<<64, 256, 0, stream>>>(b.maskCnt, 6 * (b.frames + 1), b.counters, kNumCounters, hz ? b.hist : nullptr,
                                         hz ? b.frames * kHistBins : 0);
    PT_CHECK(hipGetLastError());
}

void Lane::clear_mask(size_t F, size_t npx)
{
    if (F == 0) return;
    const unsigned gx = (unsigned)std::max<size_t>(1, std::min<size_t>(64, npx / 65536));
    const unsigned gy = (unsigned)std::min<size_t>(F, 4096);
    clear_mask_kernel<<<dim3(gx, gy), 256, 0, stream>>>(b.P, b.maskList, b.maskCnt, b.seedBits, b.seedList, b.seedCnt,
                                                        (uint32_t)F, npx, b.counters);
    PT_CHECK(hipGetLastError());
    pDirty = false;
}

size_t Lane::device_bytes() const
{
    size_t t = 0;
    for (const DevBuf* d : {&P, &lab, &fin, &aux, &mask, &seed, &perFrame, &regions, &fz, &order, &multi, &fb, &counters,
                            &flag, &rf, &gpool, &gmeta, &sbits, &hist, &rois, &rows, &staged, &frameMap, &fitNext, &fitHand})
        t += d->bytes;
    return t;
}

void Lane::release_buffers()
{
    gridGraph.release();
    for (DevBuf* d : {&P, &lab, &fin, &aux, &mask, &seed, &perFrame, &regions, &fz, &order, &multi, &fb, &counters, &flag,
                      &rf, &gpool, &gmeta, &sbits, &hist, &rois, &rows, &staged, &frameMap, &fitNext, &fitHand})
        d->release();
}

size_t Engine::chunk_frames(uint64_t planes, size_t npx) const
{
    return std::max<size_t>(1, std::min<size_t>(planes, std::min<size_t>(65535, opt.chunkPixels / std::max<size_t>(npx, 1))));
}

size_t Engine::workspace_bytes(uint64_t planes, int H, int W, uint64_t maxPerPlane, uint64_t outCapRows) const
{
    if (planes == 0 || H <= 0 || W <= 0) return 0;
    const size_t npx = (size_t)W * H;
    const size_t F = chunk_frames(planes, npx);
    const size_t SF = std::max<size_t>(1, std::min<size_t>(F, std::max<size_t>(64, (size_t(1) << 24) / npx)));
    size_t t = LaneSizes(F, npx).bytes(F, npx);                                   // first main lane
    if (planes > F && opt.lanes > 1) t += LaneSizes(F, npx).bytes(F, npx);       // second main lane
    if (opt.deferFlagged && planes > F) t += LaneSizes(SF, npx).bytes(SF, npx) + SF * 4;   // side lane + frame map
    const uint64_t roiCap = std::min<uint64_t>(outCapRows, planes * std::min<uint64_t>(maxPerPlane, npx / 4 + 2));
    // ROI store, side ROI store (grown by half again), fit rows, staged rows (grown by half again),
    // hand-over list, plane bases and placement plan, the call's small buffers (counters, fit words).
    t += alloc_pages(roiCap * sizeof(Roi)) + alloc_pages(roiCap * sizeof(Roi) * 3 / 2) + alloc_pages(roiCap * kCols * 8) +
         alloc_pages(roiCap * kCols * 8 * 3 / 2) + alloc_pages(roiCap * 4) + alloc_pages((planes + 1) * 4) +
         alloc_pages(planes * 28) + 4 * alloc_pages(4096);
    return t;
}

size_t Engine::device_bytes() const
{
    size_t t = 0;
    for (const Lane* L : {&main_, &mainB_, &side_}) t += L->device_bytes();
    for (const DevBuf* b : {&in, &out, &planeInfo, &atMask_, &atPart_, &atRows_, &passes_, &allRois_, &allBase_, &sideRois_,
                            &sideBase_, &run_, &splineRaw_, &splineCv_, &bRois_, &bRows_, &bStaged_, &bStacks_, &bJobs_,
                            &bBases_, &bPlan_, &bCounts_, &bHand_, &bWords_, &streamIn_[0], &streamIn_[1], &streamOut_[0],
                            &streamOut_[1]})
        t += b->bytes;
    return t;
}

void Engine::release_device_buffers()
{
    for (Lane* L : {&main_, &mainB_, &side_}) {
        if (L->stream) PT_CHECK(hipStreamSynchronize(L->stream));
        L->release_buffers();
    }
    for (DevBuf* b : {&in, &out, &planeInfo, &atMask_, &atPart_, &atRows_, &passes_, &allRois_, &allBase_, &sideRois_,
                      &sideBase_, &run_, &splineRaw_, &splineCv_, &bRois_, &bRows_, &bStaged_, &bStacks_, &bJobs_,
                      &bBases_, &bPlan_, &bCounts_, &bHand_, &bWords_, &streamIn_[0], &streamIn_[1], &streamOut_[0],
                      &streamOut_[1]})
        b->release();
    if (batchGraphExec_) PT_CHECK(hipGraphExecDestroy(batchGraphExec_));
    batchGraphExec_ = nullptr;
    batchGraphKey_.clear();
}

FitScratch Lane::fit_scratch(size_t nRoi)
{
    fitNext.ensure(kFitWords * 4);
    fitHand.ensure(std::max<size_t>(nRoi, 1) * 4);
    return FitScratch{fitNext.as<unsigned int>(), fitHand.as<uint32_t>()};
}

uint32_t* Lane::host(size_t words)
{
    if (words > pinnedWords) {
        if (pinned) PT_CHECK(hipHostFree(pinned));
        pinned = nullptr;
        pinnedWords = 0;
        PT_CHECK(hipHostMalloc(reinterpret_cast<void**>(&pinned), words * 4, hipHostMallocDefault));
        pinnedWords = words;
    }
    return pinned;
}

namespace {
// Double to size_t as the reference's x86-64 build converts (truncation; wraps for negatives;
// 2^63 for NaN and out-of-range values).
uint64_t host_to_size(double v)
{
    if (!(v == v) || v >= 9223372036854775808.0 || v <= -9223372036854775808.0) return 1ull << 63;
    return (uint64_t)(long long)v;
}
} // namespace

FitParams Engine::fit_params(uint64_t fit, const double* fp)
{
    FitParams p{};
    if (fit == 5) {
        // SplineFit::parseParams (SplineFit.cpp:36-49), then the odd ROI side (Fitting.cpp:35-38).
        if (!fp) throw Error("fit mode 5 needs the spline parameters");
        SplineParams& sp = spline_;
        sp = SplineParams{};
        sp.roiParam = fp[0];
        const uint64_t roiRaw = host_to_size(fp[0]);
        sp.sx = host_to_size(fp[1]);
        sp.sy = host_to_size(fp[2]);
        sp.sz = host_to_size(fp[3]);
        if (roiRaw < 1 || roiRaw > 4095) throw Error("fit mode 5: the ROI side must be in [1, 4095]");
        const uint64_t kMaxSide = 1u << 16;
        if (sp.sx < 1 || sp.sy < 1 || sp.sz < 1 || sp.sx > kMaxSide || sp.sy > kMaxSide || sp.sz > kMaxSide ||
            sp.sx * sp.sy > (1ull << 26) / sp.sz)
            throw Error("fit mode 5: spline sizes must be positive and hold at most 2^26 voxels");
        const uint64_t ssize = sp.sx * sp.sy * sp.sz, sxy = sp.sx * sp.sy;
        sp.initZ = (double)sp.sz / 2.0;
        sp.deltaZ = fp[4];
        const double* c = fp + 5;
        const double center = c[(32 * ssize) + (sp.sz / 2 * sxy) + (sp.sy / 2 * sp.sx) + (sp.sx / 2)];
        sp.ratio = std::fabs(center) < 2.220446049250313e-16 ? 1.0 : 4.0 / center;
        sp.off = host_to_size(std::floor(((double)sp.sx + 1.0 - (double)roiRaw) / 2));
        sp.roi = roiRaw % 2 == 0 ? roiRaw - 1 : roiRaw;
        sp.half = sp.roi / 2;
        const size_t bytes = (size_t)ssize * 64 * 8;
        splineRaw_.ensure(bytes);
        splineCv_.ensure(bytes);
        PT_CHECK(hipMemcpyAsync(splineRaw_.p, c, bytes, hipMemcpyHostToDevice, stream));
        launch_spline_transpose(splineRaw_.as<double>(), splineCv_.as<double>(), ssize, stream);
        PT_CHECK(hipStreamSynchronize(stream));
        sp.cv = splineCv_.as<double>();
        p.mode = 5;
        p.roi = (int)sp.roi;
        p.maxIter = opt.maxIter;
        p.spline = &spline_;
        return p;
    }
    return gaussian_fit_params(fit, fp, opt);
}

FitParams gaussian_fit_params(uint64_t fit, const double* fp, const Options& opt)
{
    FitParams p{};
    p.mode = (fit >= 1 && fit <= 4) ? (int)fit : 0;
    long roi = 7;
    if (fp) {
        const double r = fp[0];
        roi = (r == r && r >= 1 && r < 1e6) ? (long)r : 1;  // size_t conversion of fitParams[0]
        p.sigmaX = fp[1];
        p.sigmaY = fp[2];
        p.theta = fp[3];
    }
    if (roi % 2 == 0) roi--;                                // Fitting.cpp:37
    p.roi = (int)std::max<long>(roi, 1);
    p.maxIter = opt.maxIter;
    p.refill = opt.fitRefill;
    p.handover = opt.fitHandover;
    p.objective = opt.fitObjective;
    p.camGain = opt.camGain;
    p.camOffset = opt.camOffset;
    p.camReadVar = opt.camReadVar;
    return p;
}

hipEvent_t Engine::event(size_t i)
{
    while (ev_.size() <= i) {
        hipEvent_t e;
        PT_CHECK(hipEventCreate(&e));
        ev_.push_back(e);
    }
    return ev_[i];
}

void Engine::detect_chunk(Lane& L, const uint16_t* d_frames, size_t F, int H, int W, double thr, double ratio,
                          uint64_t maxPoints, size_t evBase, const uint32_t* d_map, bool resolve)
{
    const size_t npx = (size_t)W * H;
    const hipStream_t st = L.stream;
    L.zero();
    const bool prof = opt.profile && evBase != kNoEvents;
    if (L.pDirty) {
        PT_CHECK(hipMemsetAsync(L.P.p, 0, L.P.bytes, st));
        PT_CHECK(hipMemsetAsync(L.sbits.p, 0, L.sbits.bytes, st));
        L.pDirty = false;
    }
    if (prof) PT_CHECK(hipEventRecord(event(evBase + 0), st));
    L.pDirty = true;
    launch_atrous_detect(d_frames, L.b, W, H, (int)F, threshold_to_pmin(thr), st, d_map, opt.detectRows);
    if (prof) PT_CHECK(hipEventRecord(event(evBase + 1), st));
    L.rois.ensure(F * (npx / 4 + 2) * sizeof(Roi));
    // Deferred frames (d_map set) go through the exact frame flood.
    const bool forceSeq = opt.forceSequential || d_map != nullptr;
    L.b.denseMask = opt.denseFlagMask;
    launch_segment(L.b, W, H, (int)F, maxPoints, ratio, L.rois.as<Roi>(), st, forceSeq, opt.exactPath, resolve);
    // The segmentation kernels were the last readers of this chunk's P.
    L.clear_mask(F, npx);
    if (prof) PT_CHECK(hipEventRecord(event(evBase + 2), st));
}

// Grows buf to hold at least `need` bytes, keeping its first `keep` bytes (copied on stream s).
void Engine::grow(DevBuf& buf, size_t need, size_t keep, hipStream_t s)
{
    if (need <= buf.bytes) return;
    DevBuf nb;
    nb.ensure(std::max(need, buf.bytes * 3 / 2));
    if (keep) PT_CHECK(hipMemcpyAsync(nb.p, buf.p, keep, hipMemcpyDeviceToDevice, s));
    PT_CHECK(hipStreamSynchronize(s));
    buf.release();
    buf = nb;
}

uint64_t Engine::localize(const uint16_t* d_in, uint64_t planes, int H, int W, double thr, double ratio, uint64_t fit,
                          const double* fitParams, uint64_t maxPerPlane, double* d_out, uint64_t outCapRows)
{
    last = RunStats{};
    if (planes == 0 || maxPerPlane == 0 || H <= 0 || W <= 0) return 0;
    if (planes >= 0xFFFFFFFFull) throw Error("too many planes");
    const size_t npx = (size_t)W * H;
    const size_t F = chunk_frames(planes, npx);
    main_.prepare(F, npx);
    if (planes > F && opt.lanes > 1) mainB_.prepare(F, npx);
    const FitParams fp = fit_params(fit, fitParams);
    const uint64_t frameRows = std::min<uint64_t>(maxPerPlane, npx / 4 + 2);   // rows a plane can hold
    unsigned long long* passes = nullptr;
    if (opt.profile) {
        passes_.ensure(32);
        passes = passes_.as<unsigned long long>();
        fill_u32_kernel<<<1, 64, 0, main_.stream>>>(reinterpret_cast<uint32_t*>(passes), 8, 0u);   // stream-ordered
    }
    // The main lane detects chunk after chunk and gathers the ROIs of all chunks; they are fitted in
    // one launch at the end, which gives every lane of the persistent fit many ROIs, so that the
    // uneven Levenberg-Marquardt iteration counts even out. ROIs are in plane order, each plane's
    // in label order; allBase[p] is the first ROI of plane p.
    std::vector<uint32_t> deferred;
    uint64_t totalRois = 0;
    allBase_.ensure((planes + 1) * 4);
    uint32_t* dBase = allBase_.as<uint32_t>();

    // Side lane: frames whose component floods could not be proven are re-detected with the exact
    // frame flood on the side stream while the main lanes go on. Each batch's ROIs are appended, with
    // their stack frame numbers, to the side ROI store (in the order of `deferred`), and fitted with the
    // main lanes' ROIs in the call's one fit launch.
    const size_t SF = std::max<size_t>(1, std::min<size_t>(F, std::max<size_t>(64, (size_t(1) << 24) / npx)));
    std::vector<uint64_t> sideSrc;      // per deferred entry: first ROI in the side store
    std::vector<uint32_t> sideCnt;
    uint64_t sideRoiTotal = 0;
    size_t deferredNext = 0;            // next entry of `deferred` not yet sent
    size_t batchFirst = 0, batchCount = 0;
    hipEvent_t sideStart = sideEv_[0], sideDone = sideEv_[1];
    auto side_launch = [&]() {
        batchCount = std::min(SF, deferred.size() - deferredNext);
        batchFirst = deferredNext;
        deferredNext += batchCount;
        side_.prepare(SF, npx);
        uint32_t* h = side_.host(SF + (SF + 1) + kNumCounters);
        std::memcpy(h, deferred.data() + batchFirst, batchCount * 4);
        side_.frameMap.ensure(SF * 4);
        if (opt.profile) PT_CHECK(hipEventRecord(sideStart, side_.stream));
        PT_CHECK(hipMemcpyAsync(side_.frameMap.p, h, batchCount * 4, hipMemcpyHostToDevice, side_.stream));
        const uint32_t* dmap = side_.frameMap.as<uint32_t>();
        detect_chunk(side_, d_in, batchCount, H, W, thr, ratio, maxPerPlane, kNoEvents, dmap, true);
        PT_CHECK(hipMemcpyAsync(h + SF, side_.b.roiBase, (batchCount + 1) * 4, hipMemcpyDeviceToHost, side_.stream));
        PT_CHECK(hipMemcpyAsync(h + SF + SF + 1, side_.b.counters, kNumCounters * 4, hipMemcpyDeviceToHost, side_.stream));
        PT_CHECK(hipEventRecord(sideDone, side_.stream));
    };
    // Waits for the batch in flight and appends its ROIs to the side store (on the side stream,
    // before the next batch reuses the side lane's buffers).
    auto side_collect = [&]() {
        if (!batchCount) return;
        PT_CHECK(hipEventSynchronize(sideDone));
        const uint32_t* h = side_.pinned;
        const uint32_t* sbase = h + SF;
        const uint32_t* ctr = h + SF + SF + 1;
        last.status |= ctr[kStatus];
        last.floodPops += ctr[kFloodPops];
        last.rfTooBig += ctr[kRfTooBig];
        last.rfTie += ctr[kRfTie];
        last.rfUnsettled += ctr[kRfUnsettled];
        last.fbOverflow += ctr[kFbOverflow];
        last.heapFrames += ctr[kSlowCnt];
        last.rfTwoOrders += ctr[kRfTwoOrders];
        last.gridFrames += ctr[kGridOk];
        last.gridTies += ctr[kGridTies];
        last.gridUnsettled += ctr[kGridUnsettled];
        last.gridRejected += ctr[kGridRejected];
        last.gridOther += ctr[kGridOther];
        last.maskPixels += ctr[kMaskPx];
        last.floodMaxQueue = std::max<uint64_t>(last.floodMaxQueue, ctr[kFloodMaxQueue]);
        if (ctr[kStatus]) throw Error("segmentation status " + std::to_string(ctr[kStatus]));
        if (opt.profile) {
            float ms = 0;
            PT_CHECK(hipEventElapsedTime(&ms, sideStart, sideDone));
            last.msDeferred += ms;
        }
        const uint32_t n = sbase[batchCount];
        grow(sideRois_, (sideRoiTotal + n) * sizeof(Roi), sideRoiTotal * sizeof(Roi), side_.stream);
        if (n) {
            side_gather_kernel<<<std::max(1u, std::min(1024u, ceil_div(n, 256))), 256, 0, side_.stream>>>(
                side_.rois.as<Roi>(), side_.b.roiBase + batchCount, side_.frameMap.as<uint32_t>(),
                sideRois_.as<Roi>() + sideRoiTotal);
            PT_CHECK(hipGetLastError());
        }
        for (size_t k = 0; k < batchCount; ++k) {
            sideSrc[batchFirst + k] = sideRoiTotal + sbase[k];
            sideCnt[batchFirst + k] = sbase[k + 1] - sbase[k];
        }
        sideRoiTotal += n;
        batchCount = 0;
    };
    // A batch is sent whenever the side lane is idle, so that the deferred frames of a chunk are
    // resolved while the next chunks run.
    auto side_poll = [&](bool wait) {
        while (true) {
            if (batchCount) {
                if (!wait && hipEventQuery(sideDone) == hipErrorNotReady) return;
                side_collect();
            }
            if (deferred.size() == deferredNext) return;
            sideSrc.resize(deferred.size());
            sideCnt.resize(deferred.size());
            side_launch();
            if (!wait) return;
        }
    };

    // Chunks are queued back to back; the host reads a chunk's counters and deferred frames (copied
    // to pinned memory behind the chunk) once the next chunk is queued, so the main stream does not
    // wait for the host between chunks. ROI offsets in the store are carried on the device.
    const size_t metaWords = kNumCounters + 1 + F;      // counters, ROIs of the call so far, flag list
    uint32_t* meta = main_.host(2 * metaWords);
    size_t chunkNo = 0;
    std::vector<uint64_t> chunkZ0;
    std::vector<bool> chunkInLine;
    double recentFlagged = 0.0;    // flagged fraction of the last chunk whose counters were read
    auto read_meta = [&](size_t c) {
        PT_CHECK(hipEventSynchronize(chunkEv_[c & 1]));
        const uint32_t* m = meta + (c & 1) * metaWords;
        const uint32_t* ctr = m;
        last.flaggedFrames += ctr[kFlagCnt];
        last.multiComponents += ctr[kMultiCnt];
        last.flagOverflow += ctr[kFlagOverflow];
        last.flagUnproven += ctr[kFlagUnproven];
        // Exact-flood counters: non-zero only for a chunk that resolved its flagged frames itself.
        last.floodPops += ctr[kFloodPops];
        last.rfTooBig += ctr[kRfTooBig];
        last.rfTie += ctr[kRfTie];
        last.rfUnsettled += ctr[kRfUnsettled];
        last.fbOverflow += ctr[kFbOverflow];
        last.heapFrames += ctr[kSlowCnt];
        last.rfTwoOrders += ctr[kRfTwoOrders];
        last.gridFrames += ctr[kGridOk];
        last.gridTies += ctr[kGridTies];
        last.gridUnsettled += ctr[kGridUnsettled];
        last.gridRejected += ctr[kGridRejected];
        last.gridOther += ctr[kGridOther];
        last.maskPixels += ctr[kMaskPx];
        last.floodMaxQueue = std::max<uint64_t>(last.floodMaxQueue, ctr[kFloodMaxQueue]);
        last.status |= ctr[kStatus];
        if (ctr[kStatus] & 4u) throw Error("output capacity exceeded");
        if (ctr[kStatus]) throw Error("segmentation status " + std::to_string(ctr[kStatus]));
        totalRois = m[kNumCounters];
        const uint32_t* fl = m + kNumCounters + 1;
        const uint64_t fc = std::min<uint64_t>(F, planes - chunkZ0[c]);
        if (!chunkInLine[c])
            for (uint32_t k = 0; k < std::min<uint32_t>(ctr[kFlagCnt], (uint32_t)F); ++k)
                deferred.push_back((uint32_t)(chunkZ0[c] + fl[k]));
        recentFlagged = (double)ctr[kFlagCnt] / (double)std::max<uint64_t>(fc, 1);
    };
    run_.ensure(16);
    uint32_t* dRun = run_.as<uint32_t>();
    fill_u32_kernel<<<1, 64, 0, main_.stream>>>(dRun, 2, 0u);
    PT_CHECK(hipEventRecord(scanEv_[1], main_.stream));   // the first chunk's scan waits for the reset
    const uint64_t roiCap = std::min<uint64_t>(std::min<uint64_t>(outCapRows, planes * std::min<uint64_t>(maxPerPlane, npx / 4 + 2)),
                                               0xFFFFFFF0ull);
    allRois_.ensure(std::max<uint64_t>(roiCap, 1) * sizeof(Roi));
    // Consecutive chunks alternate between two lanes (buffers and streams), so that the latency-bound
    // end of one chunk (component floods, statistics, ROI list) overlaps the next chunk's detection.
    // Only the ROI scans are ordered across the lanes (the running ROI total).
    for (uint64_t z0 = 0; z0 < planes; z0 += F, ++chunkNo) {
        Lane& L = (opt.lanes > 1 && (chunkNo & 1)) ? mainB_ : main_;
        const size_t f = std::min<uint64_t>(F, planes - z0);
        chunkZ0.push_back(z0);
        L.b.runTotal = dRun;
        L.b.scanWait = scanEv_[(chunkNo + 1) & 1];
        L.b.scanDone = scanEv_[chunkNo & 1];
        // The flagged frames of a chunk are resolved in line (exact frame flood on the chunk's lane) or
        // deferred to the side lane. The last chunk resolves its own (a side batch started after it
        // would only end after the fit of the call); the others follow defer_flagged: never, always,
        // or (2) when the last chunk whose counters are known flagged at most 2 % of its frames, so
        // that sparse stacks pay for the exact flood once per side batch and dense ones do not
        // detect their frames twice.
        const bool lastChunk = z0 + F >= planes;
        const bool inLine = lastChunk || opt.deferFlagged == 0 || (opt.deferFlagged == 2 && recentFlagged > 0.02);
        chunkInLine.push_back(inLine);
        detect_chunk(L, d_in + z0 * npx, f, H, W, thr, ratio, maxPerPlane, 3 * chunkNo, nullptr, inLine);
        L.b.runTotal = nullptr;
        L.b.scanWait = L.b.scanDone = nullptr;
        gather_rois(L.rois.as<Roi>(), (uint32_t)z0, allRois_.as<Roi>(), L.b.roiBase, (uint32_t)f, L.b.chunkBase,
                    (uint32_t)roiCap, dBase + z0, L.b.counters, L.stream);
        uint32_t* m = meta + (chunkNo & 1) * metaWords;
        PT_CHECK(hipMemcpyAsync(m, L.b.counters, kNumCounters * 4, hipMemcpyDeviceToHost, L.stream));
        PT_CHECK(hipMemcpyAsync(m + kNumCounters, dRun, 4, hipMemcpyDeviceToHost, L.stream));
        PT_CHECK(hipMemcpyAsync(m + kNumCounters + 1, L.b.flagList, f * 4, hipMemcpyDeviceToHost, L.stream));
        PT_CHECK(hipEventRecord(chunkEv_[chunkNo & 1], L.stream));
        if (chunkNo > 0) read_meta(chunkNo - 1);
        side_poll(false);
        last.chunks++;
    }
    read_meta(chunkNo - 1);
    // Every deferred frame detected and its ROIs in the side store.
    side_poll(true);
    PT_CHECK(hipStreamSynchronize(side_.stream));
    last.deferredFrames = deferred.size();
    set_u32(dBase + planes, (uint32_t)totalRois, main_.stream);
    // The fit covers the main lanes' ROIs (plane order) followed by the side store's (deferred order).
    const uint64_t fitRois = totalRois + sideRoiTotal;
    if (fitRois > 0xFFFFFFF0ull) throw Error("too many ROIs in one call");
    const uint32_t nDef = (uint32_t)deferred.size();
    // Host block: deferred planes, their first side ROI and count (u64 each), then the row total, the
    // LM pass counters and the hand-over count, then the side rows' per-frame bases (u32).
    uint32_t* hw = main_.host(2 * (3 * (size_t)nDef + 6) + nDef + 2);
    uint64_t* hp = reinterpret_cast<uint64_t*>(hw);
    uint32_t* hSideBase = hw + 2 * (3 * (size_t)nDef + 6);
    for (uint32_t k = 0; k < nDef; ++k) {
        hp[k] = deferred[k];
        hp[nDef + k] = sideSrc[k];
        hp[2 * (size_t)nDef + k] = sideCnt[k];
        hSideBase[k] = (uint32_t)sideSrc[k];
    }
    hSideBase[nDef] = (uint32_t)sideRoiTotal;
    if (sideRoiTotal) {
        grow(allRois_, fitRois * sizeof(Roi), totalRois * sizeof(Roi), main_.stream);
        PT_CHECK(hipMemcpyAsync(allRois_.as<Roi>() + totalRois, sideRois_.p, sideRoiTotal * sizeof(Roi),
                                hipMemcpyDeviceToDevice, main_.stream));
    }
    // Fit of every ROI, then each plane's rows in (Y, X) order: the main planes by their plane bases,
    // the deferred planes' rows (after them) by their side bases.
    main_.rows.ensure(std::max<uint64_t>(fitRois, 1) * kCols * 8);
    grow(main_.staged, std::max<uint64_t>(fitRois, 1) * kCols * 8, 0, main_.stream);
    if (opt.profile) PT_CHECK(hipEventRecord(fitEv_[0], main_.stream));
    if (fitRois)
        launch_fit(d_in, npx, W, H, allRois_.as<Roi>(), (uint32_t)fitRois, fp, main_.rows.as<double>(), opt.fitFp64,
                   main_.fit_scratch(fitRois), main_.stream, passes, nullptr, opt.fitWaves);
    if (opt.profile) PT_CHECK(hipEventRecord(fitEv_[1], main_.stream));
    // The sort's key scratch is the ROI store, free once the fit (earlier on this stream) has read it.
    static_assert(sizeof(Roi) >= kSortScratchPerRow, "the ROI store holds the sort keys");
    if (totalRois)
        launch_sort_frames(main_.rows.as<double>(), dBase, (int)planes, main_.staged.as<double>(), main_.stream, frameRows,
                           totalRois, allRois_.as<Roi>());
    if (sideRoiTotal) {
        sideBase_.ensure((nDef + 1) * 4);
        PT_CHECK(hipMemcpyAsync(sideBase_.p, hSideBase, (nDef + 1) * 4, hipMemcpyHostToDevice, main_.stream));
        launch_sort_frames(main_.rows.as<double>() + totalRois * kCols, sideBase_.as<uint32_t>(), (int)nDef,
                           main_.staged.as<double>() + totalRois * kCols, main_.stream, frameRows, sideRoiTotal,
                           allRois_.as<Roi>() + totalRois);
    }
    if (opt.profile) PT_CHECK(hipEventRecord(fitEv_[2], main_.stream));

    // Placement in plane order, planned on the device from the plane bases of the main rows; the
    // deferred planes take the side store's rows, placed after the main lanes' in the staged array,
    // and are the only host input. The host reads the row total once, after the last copy.
    planeInfo.ensure(planes * 20 + (3 * (uint64_t)nDef + 1) * 8 + 64);
    uint64_t* dSrc = planeInfo.as<uint64_t>();
    uint64_t* dDst = dSrc + planes;
    uint32_t* dCnt = reinterpret_cast<uint32_t*>(dDst + planes);
    uint64_t* dDef = dDst + planes + (planes + 1) / 2;      // after the counts, 8-byte aligned
    uint64_t* dTotal = dDef + 3 * (uint64_t)nDef;
    if (nDef) PT_CHECK(hipMemcpyAsync(dDef, hp, 3 * (size_t)nDef * 8, hipMemcpyHostToDevice, main_.stream));
    launch_plan(dBase, planes, dDef, nDef, totalRois, dSrc, dDst, dCnt, dTotal, main_.stream);
    launch_scatter(main_.staged.as<double>(), dSrc, dDst, dCnt, planes, d_out, outCapRows, main_.stream);
    PT_CHECK(hipMemcpyAsync(hp + 3 * (size_t)nDef, dTotal, 8, hipMemcpyDeviceToHost, main_.stream));
    if (passes) PT_CHECK(hipMemcpyAsync(hp + 3 * (size_t)nDef + 1, passes, 32, hipMemcpyDeviceToHost, main_.stream));
    hp[3 * (size_t)nDef + 5] = 0;
    const bool fastFit = fitRois && !opt.fitFp64 && fp.mode >= 1 && fp.mode <= 4;
    if (fastFit)
        PT_CHECK(hipMemcpyAsync(hp + 3 * (size_t)nDef + 5, main_.fitNext.as<unsigned int>() + 1, 4, hipMemcpyDeviceToHost,
                                main_.stream));
    PT_CHECK(hipStreamSynchronize(main_.stream));
    const uint64_t total = hp[3 * (size_t)nDef];
    if (total > outCapRows) throw Error("output capacity exceeded");
    if (opt.profile) {
        for (size_t c = 0; c < last.chunks; ++c) {
            float ms[2] = {0, 0};
            PT_CHECK(hipEventElapsedTime(&ms[0], event(3 * c + 0), event(3 * c + 1)));
            PT_CHECK(hipEventElapsedTime(&ms[1], event(3 * c + 1), event(3 * c + 2)));
            last.msAtrous += ms[0];
            last.msSegment += ms[1];
        }
        float ms[2] = {0, 0};
        PT_CHECK(hipEventElapsedTime(&ms[0], fitEv_[0], fitEv_[1]));
        PT_CHECK(hipEventElapsedTime(&ms[1], fitEv_[1], fitEv_[2]));
        last.msFit = ms[0];
        last.msAssemble = ms[1];
        last.jacPasses = hp[3 * (size_t)nDef + 1];
        last.trialPasses = hp[3 * (size_t)nDef + 2];
        last.groupJacPasses = hp[3 * (size_t)nDef + 3];
        last.groupTrialPasses = hp[3 * (size_t)nDef + 4];
    }
    last.fitHanded = fastFit ? (fp.mode == 4 && fp.handover > 0 ? fitRois : hp[3 * (size_t)nDef + 5]) : 0;
    last.frames = planes;
    last.rois = total;
    return total * kCols;
}

uint64_t Engine::fit_positions(const uint16_t* d_in, uint64_t planes, int H, int W, const double* hostPos, uint64_t n,
                               uint64_t fit, const double* fitParams, double* d_out)
{
    last = RunStats{};
    if (n == 0) return 0;
    if (fit < 1 || fit > 5) throw Error("Fitting supports the Gaussian modes 1-4 and the spline mode 5");
    if (n > 0xFFFFFFFFull) throw Error("too many positions");
    // Per-plane local index in input order (lib/Palm.cpp:79-84).
    std::vector<uint32_t> local(n);
    std::vector<uint64_t> counts(planes + 1, 0);
    for (uint64_t i = 0; i < n; ++i) {
        const double pl = hostPos[i * kInCols + kInPlane];
        const uint64_t k = (pl >= 1.0 && pl < (double)planes + 1.0) ? (uint64_t)pl - 1 : planes;
        local[i] = (uint32_t)++counts[k];
    }
    DevBuf dpos, dloc, dbad;
    dpos.ensure(n * kInCols * 8);
    dloc.ensure(n * 4);
    dbad.ensure(4);
    main_.rois.ensure(n * sizeof(Roi));
    PT_CHECK(hipMemcpyAsync(dpos.p, hostPos, n * kInCols * 8, hipMemcpyHostToDevice, stream));
    PT_CHECK(hipMemcpyAsync(dloc.p, local.data(), n * 4, hipMemcpyHostToDevice, stream));
    fill_u32_kernel<<<1, 64, 0, stream>>>(dbad.as<uint32_t>(), 1, 0u);
    launch_positions_to_rois(dpos.as<double>(), n, planes, main_.rois.as<Roi>(), dbad.as<uint32_t>(), stream);
    const FitParams fp = fit_params(fit, fitParams);
    launch_fit(d_in, (size_t)W * H, W, H, main_.rois.as<Roi>(), (uint32_t)n, fp, d_out, opt.fitFp64, main_.fit_scratch(n),
               stream, nullptr, nullptr, opt.fitWaves);
    launch_positions_ids(dpos.as<double>(), dloc.as<uint32_t>(), n, d_out, stream);
    uint32_t bad = 0;
    PT_CHECK(hipMemcpyAsync(&bad, dbad.p, 4, hipMemcpyDeviceToHost, stream));
    PT_CHECK(hipStreamSynchronize(stream));
    if (bad) std::fprintf(stderr, "[WARNING] Fitting: positions with a plane outside [1, %llu] got failure rows.\n",
                          (unsigned long long)planes);
    dpos.release();
    dloc.release();
    dbad.release();
    last.frames = planes;
    last.rois = n;
    return n * kCols;
}

void Engine::wavelet(const uint16_t* d_in, uint64_t planes, int H, int W, int level, double* d_out)
{
    launch_atrous_plane(d_in, d_out, W, H, (int)planes, level, stream);
    PT_CHECK(hipStreamSynchronize(stream));
}

void Engine::debug_labels(const uint16_t* d_frame, int H, int W, double thr, double ratio, uint32_t* d_labels)
{
    const size_t npx = (size_t)W * H;
    main_.prepare(1, npx);
    main_.zero();
    main_.pDirty = true;
    launch_atrous_detect(d_frame, main_.b, W, H, 1, threshold_to_pmin(thr), stream);
    main_.rois.ensure((npx / 4 + 2) * sizeof(Roi));
    main_.b.denseMask = opt.denseFlagMask;
    launch_segment(main_.b, W, H, 1, ~0ull, ratio, main_.rois.as<Roi>(), stream, opt.forceSequential, opt.exactPath);
    // Labels only where P > 0 (fin is not written elsewhere).
    masked_labels<<<256, 256, 0, stream>>>(main_.b.fin, main_.b.P, d_labels, npx);
    PT_CHECK(hipGetLastError());
    main_.clear_mask(1, npx);
    uint32_t ctr[kNumCounters];
    PT_CHECK(hipMemcpyAsync(ctr, main_.b.counters, sizeof(ctr), hipMemcpyDeviceToHost, stream));
    PT_CHECK(hipStreamSynchronize(stream));
    last = RunStats{};
    last.flaggedFrames = ctr[kFlagCnt];
    last.multiComponents = ctr[kMultiCnt];
    last.rfTooBig = ctr[kRfTooBig];
    last.rfTie = ctr[kRfTie];
    last.rfUnsettled = ctr[kRfUnsettled];
    last.fbOverflow = ctr[kFbOverflow];
    last.heapFrames = ctr[kSlowCnt];
    last.rfTwoOrders = ctr[kRfTwoOrders];
    last.gridFrames = ctr[kGridOk];
    last.gridTies = ctr[kGridTies];
    last.gridUnsettled = ctr[kGridUnsettled];
    last.gridRejected = ctr[kGridRejected];
    last.gridOther = ctr[kGridOther];
    last.status = ctr[kStatus];
}

} // namespace pt
