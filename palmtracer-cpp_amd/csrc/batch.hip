// Batched localization of many stacks in one call (no counterpart in the reference: its callers run
// Localization once per stack, lib/Palm.cpp:48-70).
//
// 1. Every stack is detected and segmented on its own (the chunked pipeline of Engine::localize with
//    fit mode 0): its region list becomes rows in the caller's output, in plane order. A region row
//    holds what a fit job needs (plane, X, Y, segmentation intensity, surface; the rows the reference's
//    FitLocalizations takes, lib/Palm.cpp:73-119).
// 2. The fit jobs of all stacks are grouped by fit class (mode, ROI side, initial widths and angle)
//    and fitted by one persistent work-queue launch per class (the fast fit: FP32 kernel, FP64
//    hand-over), the classes side by side on two streams; the launches read each job's frame through
//    a table of the stacks (StackRef). The fit stage is captured once in a hipGraph and replayed while
//    the class layout and the buffers stay the same; ROI counts are read from device memory, so one
//    graph serves batches of any size within the buffers.
// 3. Each class's schedule (resident waves, refill threshold of the FP32 kernel, iterations before the
//    FP64 hand-over) is picked by a small autotuner on the class's own jobs the first time the class
//    signature (fit parameters, job count to within a factor of two) is seen: a coordinate search over
//    the candidates, each timed with HIP events on the fit stream.
// 4. Each stack's fitted rows are sorted per plane by (Y, X) and placed with their Id, Plane and Index
//    (lib/localization/FrameProcessor.cpp:28-82, lib/Palm.cpp:24-43), so a stack's rows equal those of
//    Localization on that stack alone with the same schedule.
#include "engine.h"
#include "../../include/palmtracer.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

namespace pt {
namespace {

struct RowsJob {
    const double* rows;   // region rows of one stack (18 doubles each)
    uint64_t n;           // rows
    uint64_t dst;         // first fit job in the batch
    uint32_t stack;
    uint32_t pad;
};

// Region rows to fit jobs (Roi.slot = stack index, Roi.frame = plane - 1); one y-block per stack.
__global__ void rows_to_rois_kernel(const RowsJob* __restrict__ jobs, Roi* __restrict__ rois)
{
    const RowsJob j = jobs[blockIdx.y];
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < j.n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double* r = j.rows + i * kCols;
        Roi o;
        o.x = r[kX];
        o.y = r[kY];
        o.seg = r[kSegIntensity];
        o.surf = r[kSurface];
        o.frame = (uint32_t)(r[kPlane] - 1.0);
        o.slot = j.stack;
        rois[j.dst + i] = o;
    }
}

// First row of each plane p in plane-ordered rows (lower bound of Plane = p + 1), p = 0 .. planes.
__global__ void plane_bases_kernel(const double* __restrict__ rows, uint64_t n, uint64_t planes, uint32_t* __restrict__ base)
{
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p <= planes; p += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t a = 0, b = n;
        const double key = (double)(p + 1);
        while (a < b) {
            const uint64_t m = (a + b) >> 1;
            if (rows[m * kCols + kPlane] < key) a = m + 1; else b = m;
        }
        base[p] = (uint32_t)a;
    }
}

std::string class_key(const FitParams& fp)
{
    char b[256];
    std::snprintf(b, sizeof(b), "%d/%d/%.17g/%.17g/%.17g/%d/%d/%.17g/%.17g/%.17g", fp.mode, fp.roi, fp.sigmaX, fp.sigmaY,
                  fp.theta, fp.maxIter, fp.objective, fp.camGain, fp.camOffset, fp.camReadVar);
    return b;
}

} // namespace

void Engine::batch_localize(BatchItem* items, uint64_t nItems)
{
    last = RunStats{};
    batchLast_.clear();
    if (nItems == 0) return;
    if (nItems > 0xFFFFFFFFull) throw Error("too many stacks");
    // 1. Detection and segmentation of every stack; region rows in the caller's output.
    std::vector<uint64_t> nRows(nItems, 0);
    RunStats total{};
    for (uint64_t s = 0; s < nItems; ++s) {
        BatchItem& it = items[s];
        it.doublesWritten = 0;
        if (it.fit == 5) throw Error("batched localization: Gaussian fit modes 0-4 only (mode 5 runs through Localization)");
        if (it.height == 0 || it.width == 0 || it.height > 0x7FFFFFFFull || it.width > 0x7FFFFFFFull)
            throw Error("batched localization: bad frame size in stack " + std::to_string(s));
        const uint64_t d = localize(it.d_input, it.planes, (int)it.height, (int)it.width, it.threshold, it.watershedRatio, 0,
                                    it.fit_params, it.maxPerPlane, it.d_output, it.outputCapDoubles / kCols);
        nRows[s] = d / kCols;
        it.doublesWritten = d;
        total.frames += last.frames;
        total.chunks += last.chunks;
        total.flaggedFrames += last.flaggedFrames;
        total.deferredFrames += last.deferredFrames;
        total.msAtrous += last.msAtrous;
        total.msSegment += last.msSegment;
    }
    // 2. Fit classes: stacks with the same fit parameters, in order of first appearance.
    struct Class {
        FitParams fp;
        std::string key;
        std::vector<uint64_t> stacks;
        uint64_t first = 0, count = 0;
    };
    std::vector<Class> classes;
    std::vector<uint64_t> jobOff(nItems, 0), classOf(nItems, ~0ull);
    for (uint64_t s = 0; s < nItems; ++s) {
        if (items[s].fit < 1 || items[s].fit > 4 || nRows[s] == 0) continue;
        FitParams fp = gaussian_fit_params(items[s].fit, items[s].fit_params, opt);
        const std::string k = class_key(fp);
        size_t c = 0;
        while (c < classes.size() && classes[c].key != k) ++c;
        if (c == classes.size()) classes.push_back(Class{fp, k, {}, 0, 0});
        classes[c].stacks.push_back(s);
        classOf[s] = c;
    }
    uint64_t nJobs = 0;
    for (Class& c : classes) {
        c.first = nJobs;
        for (uint64_t s : c.stacks) {
            jobOff[s] = nJobs;
            nJobs += nRows[s];
        }
        c.count = nJobs - c.first;
    }
    last = total;
    if (nJobs == 0) return;
    if (nJobs > 0xFFFFFFF0ull) throw Error("batched localization: too many fit jobs");
    const hipStream_t s0 = main_.stream;
    // Buffers (they only grow, so that a captured fit graph stays valid across batches).
    auto reserve = [&](DevBuf& b, size_t bytes) {
        if (bytes > b.bytes) {
            PT_CHECK(hipStreamSynchronize(s0));
            b.ensure(bytes + bytes / 2);
            batchGraphKey_.clear();   // buffers moved: the graph is captured again
        }
    };
    uint64_t basesWords = 0, planWords = 0;
    for (uint64_t s = 0; s < nItems; ++s)
        if (classOf[s] != ~0ull) {
            basesWords += items[s].planes + 1;
            planWords += items[s].planes * 5 + 2;
        }
    reserve(bRois_, nJobs * sizeof(Roi));
    reserve(bRows_, nJobs * kCols * 8);
    reserve(bStaged_, nJobs * kCols * 8);
    reserve(bStacks_, nItems * sizeof(StackRef));
    reserve(bJobs_, nItems * sizeof(RowsJob));
    reserve(bBases_, basesWords * 4);
    reserve(bPlan_, planWords * 4 * 2 + 64);
    reserve(bCounts_, std::max<size_t>(classes.size(), 1) * 4 + 64);
    reserve(bHand_, nJobs * 4);
    reserve(bWords_, std::max<size_t>(classes.size(), 1) * kFitWords * 4 * 16);
    // Host tables (pinned): stacks, jobs, class counts.
    const size_t hostWords = nItems * (sizeof(StackRef) + sizeof(RowsJob)) / 4 + classes.size() + 16;
    uint32_t* hw = side_.host(hostWords);
    StackRef* hStacks = reinterpret_cast<StackRef*>(hw);
    RowsJob* hJobs = reinterpret_cast<RowsJob*>(hStacks + nItems);
    uint32_t* hCounts = reinterpret_cast<uint32_t*>(hJobs + nItems);
    uint32_t nJobTables = 0;
    for (uint64_t s = 0; s < nItems; ++s) {
        hStacks[s] = StackRef{items[s].d_input, items[s].height * items[s].width, (int)items[s].width, (int)items[s].height};
        if (classOf[s] == ~0ull) continue;
        hJobs[nJobTables++] = RowsJob{items[s].d_output, nRows[s], jobOff[s], (uint32_t)s, 0};
    }
    for (size_t c = 0; c < classes.size(); ++c) hCounts[c] = (uint32_t)classes[c].count;
    PT_CHECK(hipMemcpyAsync(bStacks_.p, hStacks, nItems * sizeof(StackRef), hipMemcpyHostToDevice, s0));
    PT_CHECK(hipMemcpyAsync(bJobs_.p, hJobs, nJobTables * sizeof(RowsJob), hipMemcpyHostToDevice, s0));
    PT_CHECK(hipMemcpyAsync(bCounts_.p, hCounts, classes.size() * 4, hipMemcpyHostToDevice, s0));
    {
        const uint64_t mx = *std::max_element(nRows.begin(), nRows.end());
        dim3 g((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(64, (mx + 255) / 256)), nJobTables);
        rows_to_rois_kernel<<<g, 256, 0, s0>>>(bJobs_.as<RowsJob>(), bRois_.as<Roi>());
        PT_CHECK(hipGetLastError());
    }
    // Per-plane bases of every fitted stack, from its region rows (before they are overwritten).
    std::vector<uint64_t> baseOff(nItems, 0), planOff(nItems, 0);
    {
        uint64_t bo = 0, po = 0;
        for (uint64_t s = 0; s < nItems; ++s) {
            if (classOf[s] == ~0ull) continue;
            baseOff[s] = bo;
            planOff[s] = po;
            const uint64_t P = items[s].planes;
            plane_bases_kernel<<<(unsigned)std::max<uint64_t>(1, std::min<uint64_t>(1024, (P + 256) / 256)), 256, 0, s0>>>(
                items[s].d_output, nRows[s], P, bBases_.as<uint32_t>() + bo);
            PT_CHECK(hipGetLastError());
            bo += P + 1;
            po += P * 5 + 2;
        }
    }
    // Class launch parameters: the fast fit through the stack table, the count on the device.
    auto class_fp = [&](const Class& c, const BatchSchedule& sc) {
        FitParams fp = c.fp;
        fp.stacks = bStacks_.as<StackRef>();
        fp.handover = sc.handover;
        fp.refill = sc.refill;
        return fp;
    };
    auto launch_class = [&](size_t ci, const BatchSchedule& sc, hipStream_t st) {
        const Class& c = classes[ci];
        const FitScratch ws{bWords_.as<unsigned int>() + ci * kFitWords * 16, bHand_.as<uint32_t>() + c.first};
        launch_fit(nullptr, 0, 0, 0, bRois_.as<Roi>() + c.first, (uint32_t)c.count, class_fp(c, sc),
                   bRows_.as<double>() + c.first * kCols, false, ws, st, nullptr, nullptr, sc.waves,
                   bCounts_.as<uint32_t>() + ci);
    };
    // 3. Schedules: cached per class signature, tuned on first sight.
    std::vector<BatchSchedule> sched(classes.size());
    for (size_t ci = 0; ci < classes.size(); ++ci) {
        const Class& c = classes[ci];
        uint64_t bucket = 0;
        while ((2ull << bucket) <= c.count) ++bucket;
        const std::string sig = c.key + "#" + std::to_string(bucket);
        auto found = tuned_.find(sig);
        if (found != tuned_.end()) {
            sched[ci] = found->second;
            continue;
        }
        BatchSchedule best{opt.fitWaves ? opt.fitWaves : 4096u, opt.fitRefill, opt.fitHandover > 0 ? opt.fitHandover : 16, 0.0};
        if (opt.batchAutotune) {
            hipEvent_t e0 = event(0), e1 = event(1);
            auto time_it = [&](const BatchSchedule& sc) {
                launch_class(ci, sc, s0);   // warm (first launch of a kernel variant)
                PT_CHECK(hipEventRecord(e0, s0));
                launch_class(ci, sc, s0);
                PT_CHECK(hipEventRecord(e1, s0));
                PT_CHECK(hipEventSynchronize(e1));
                float ms = 0;
                PT_CHECK(hipEventElapsedTime(&ms, e0, e1));
                return (double)ms;
            };
            best.ms = time_it(best);
            const int ks[3] = {12, 16, 24};
            const unsigned ws[3] = {1024, 2048, 4096};
            const int rs[3] = {8, 16, 32};
            for (int k : ks) {
                BatchSchedule t = best;
                t.handover = k;
                if (t.handover == best.handover) continue;
                t.ms = time_it(t);
                if (t.ms < best.ms) best = t;
            }
            for (unsigned w : ws) {
                BatchSchedule t = best;
                t.waves = w;
                if (t.waves == best.waves) continue;
                t.ms = time_it(t);
                if (t.ms < best.ms) best = t;
            }
            for (int r : rs) {
                BatchSchedule t = best;
                t.refill = r;
                if (t.refill == best.refill) continue;
                t.ms = time_it(t);
                if (t.ms < best.ms) best = t;
            }
            best.tuned = true;
        }
        tuned_[sig] = best;
        sched[ci] = best;
    }
    // The fit stage: one graph, the classes on two streams (fork and join through events).
    std::string gkey;
    for (size_t ci = 0; ci < classes.size(); ++ci) {
        const Class& c = classes[ci];
        gkey += class_key(c.fp) + ":" + std::to_string(sched[ci].waves) + "," + std::to_string(sched[ci].refill) + "," +
                std::to_string(sched[ci].handover) + "," + std::to_string(c.first) + "," + std::to_string(c.count) + ";";
    }
    if (opt.profile) PT_CHECK(hipEventRecord(fitEv_[0], s0));
    if (opt.batchGraph) {
        if (gkey != batchGraphKey_) {
            if (batchGraphExec_) PT_CHECK(hipGraphExecDestroy(batchGraphExec_));
            batchGraphExec_ = nullptr;
            batchGraphKey_.clear();
            hipStream_t s1 = mainB_.stream;
            hipGraph_t g = nullptr;
            PT_CHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
            PT_CHECK(hipEventRecord(forkEv_[0], s0));
            PT_CHECK(hipStreamWaitEvent(s1, forkEv_[0], 0));
            for (size_t ci = 0; ci < classes.size(); ++ci) launch_class(ci, sched[ci], (ci & 1) ? s1 : s0);
            PT_CHECK(hipEventRecord(forkEv_[1], s1));
            PT_CHECK(hipStreamWaitEvent(s0, forkEv_[1], 0));
            PT_CHECK(hipStreamEndCapture(s0, &g));
            PT_CHECK(hipGraphInstantiate(&batchGraphExec_, g, nullptr, nullptr, 0));
            PT_CHECK(hipGraphDestroy(g));
            batchGraphKey_ = gkey;
        }
        PT_CHECK(hipGraphLaunch(batchGraphExec_, s0));
    } else {
        for (size_t ci = 0; ci < classes.size(); ++ci) launch_class(ci, sched[ci], s0);
    }
    if (opt.profile) PT_CHECK(hipEventRecord(fitEv_[1], s0));
    // 4. Per stack: rows sorted per plane by (Y, X), then placed with Id and Plane.
    uint64_t* plan = bPlan_.as<uint64_t>();
    for (uint64_t s = 0; s < nItems; ++s) {
        if (classOf[s] == ~0ull) continue;
        const BatchItem& it = items[s];
        const uint64_t P = it.planes, npx = it.height * it.width;
        const uint32_t* base = bBases_.as<uint32_t>() + baseOff[s];
        double* rows = bRows_.as<double>() + jobOff[s] * kCols;
        double* staged = bStaged_.as<double>() + jobOff[s] * kCols;
        // Key scratch of the sort: the stack's ROIs, which the fit has read by now.
        launch_sort_frames(rows, base, (int)P, staged, s0, std::min<uint64_t>(it.maxPerPlane, npx / 4 + 2), nRows[s],
                           bRois_.as<Roi>() + jobOff[s]);
        uint64_t* src = plan + planOff[s];
        uint64_t* dst = src + P;
        uint32_t* cnt = reinterpret_cast<uint32_t*>(dst + P);
        uint64_t* tot = reinterpret_cast<uint64_t*>(cnt + ((P + 1) & ~uint64_t(1)));
        launch_plan(base, P, nullptr, 0, 0, src, dst, cnt, tot, s0);
        launch_scatter(staged, src, dst, cnt, P, it.d_output, it.outputCapDoubles / kCols, s0);
    }
    if (opt.profile) PT_CHECK(hipEventRecord(fitEv_[2], s0));
    PT_CHECK(hipStreamSynchronize(s0));
    if (opt.profile) {
        float ms[2] = {0, 0};
        PT_CHECK(hipEventElapsedTime(&ms[0], fitEv_[0], fitEv_[1]));
        PT_CHECK(hipEventElapsedTime(&ms[1], fitEv_[1], fitEv_[2]));
        last.msFit = ms[0];
        last.msAssemble = ms[1];
    }
    for (size_t ci = 0; ci < classes.size(); ++ci)
        batchLast_.push_back(BatchClassInfo{classes[ci].fp.mode, classes[ci].fp.roi, classes[ci].count, sched[ci]});
    last.rois = nJobs;
}

} // namespace pt

extern "C" PT_API int PT_CALL pt_batch_localization(pt_batch_stack* stacks, uint64_t nStacks)
{
    try {
        pt::Engine& e = pt::Engine::get();   // batches keep their tuned schedules and captured graph here
        std::lock_guard<std::mutex> lk(e.mu);
        e.opt = pt::Engine::shared_options();
        std::vector<pt::BatchItem> items(nStacks);
        for (uint64_t s = 0; s < nStacks; ++s) {
            const pt_batch_stack& b = stacks[s];
            items[s] = pt::BatchItem{b.d_input, b.planes, b.height, b.width, b.thresholdVal, b.watershedRatio, b.fit,
                                     b.fit_params, b.d_output, b.outputCapDoubles, b.maxPerPlane, 0};
        }
        struct Publish {
            pt::Engine& e;
            ~Publish() { pt::Engine::publish(e.last); }
        } publish{e};
        e.batch_localize(items.data(), nStacks);
        for (uint64_t s = 0; s < nStacks; ++s) stacks[s].doublesWritten = items[s].doublesWritten;
        return 1;
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "[palmtracer] pt_batch_localization: %s\n", ex.what());
    } catch (...) {
        std::fprintf(stderr, "[palmtracer] pt_batch_localization: unknown error\n");
    }
    for (uint64_t s = 0; s < nStacks; ++s) stacks[s].doublesWritten = 0;
    return 0;
}

extern "C" PT_API int PT_CALL pt_batch_schedules(double* out, uint64_t n)
{
    pt::Engine& e = pt::Engine::get();
    std::lock_guard<std::mutex> lk(e.mu);
    const auto& v = e.batch_classes();
    uint64_t k = 0;
    for (const auto& c : v) {
        const double row[8] = {(double)c.mode, (double)c.roi, (double)c.jobs, (double)c.sched.waves, (double)c.sched.refill,
                               (double)c.sched.handover, c.sched.ms, c.sched.tuned ? 1.0 : 0.0};
        for (int j = 0; j < 8 && k < n; ++j) out[k++] = row[j];
    }
    return (int)v.size();
}
