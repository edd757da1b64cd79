// Localization and Fitting over host buffers (the reference's exported entry points, CPU/main.cpp:
// 112-121, 141-148), streamed: the stack goes through the device in pieces of at most
// Options::streamPixels pixels, so that stacks larger than HBM work and the transfers overlap the
// computation.
//
// Per piece p: host frames -> pinned staging (host threads) -> device (copy stream), the device
// pipeline over the piece (Engine::localize / fit_positions, unchanged), rows -> pinned staging
// (copy stream) -> caller's output (host threads). Two staging buffers each way: while piece p is
// computed, piece p + 1 is uploaded and the rows of piece p - 1 are copied out.
//
// A piece's rows equal the rows of the same planes in a whole-stack call but for the global Id and
// the plane number (lib/Palm.cpp:24-43: Id counts rows over the whole stack, Plane is the stack
// plane), which are shifted on the device before the download.
#include "engine.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace pt {
namespace {

__global__ void shift_rows_kernel(double* __restrict__ rows, uint64_t n, double idOffset, double planeOffset)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        rows[i * kCols + kId] += idOffset;
        rows[i * kCols + kPlane] += planeOffset;
    }
}

// Joins its thread when it goes out of scope (also when an exception leaves the scope).
struct Joiner {
    std::thread t;
    explicit Joiner(std::thread&& th) : t(std::move(th)) {}
    void join()
    {
        if (t.joinable()) t.join();
    }
    ~Joiner() { join(); }
};

// memcpy split over up to `threads` host threads (blocks of at least 4 MiB).
void parallel_copy(void* dst, const void* src, size_t bytes, int threads)
{
    const size_t minBlock = size_t(4) << 20;
    const int t = (int)std::max<size_t>(1, std::min<size_t>((size_t)threads, bytes / minBlock));
    if (t <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    std::vector<std::thread> pool;
    const size_t per = (bytes + t - 1) / t;
    for (int i = 0; i < t; ++i) {
        const size_t a = (size_t)i * per, b = std::min(bytes, a + per);
        if (a >= b) break;
        pool.emplace_back([=] { std::memcpy((char*)dst + a, (const char*)src + a, b - a); });
    }
    for (std::thread& th : pool) th.join();
}

} // namespace

void Engine::stream_prepare(size_t inBytes, size_t outBytes)
{
    if (!copyStream_) {
        PT_CHECK(hipStreamCreateWithFlags(&copyStream_, hipStreamNonBlocking));
        for (int i = 0; i < 2; ++i) {
            PT_CHECK(hipEventCreate(&inEv_[i]));
            PT_CHECK(hipEventCreate(&outEv_[i]));
            PT_CHECK(hipEventCreate(&inStartEv_[i]));
            PT_CHECK(hipEventCreate(&outStartEv_[i]));
            PT_CHECK(hipEventCreateWithFlags(&computeEv_[i], hipEventDisableTiming));
        }
    }
    for (int i = 0; i < 2; ++i) {
        if (inBytes > pinInBytes_) {
            if (pinIn_[i]) PT_CHECK(hipHostFree(pinIn_[i]));
            pinIn_[i] = nullptr;
        }
        if (!pinIn_[i]) PT_CHECK(hipHostMalloc(&pinIn_[i], std::max<size_t>(inBytes, 256), hipHostMallocDefault));
        if (outBytes > pinOutBytes_) {
            if (pinOut_[i]) PT_CHECK(hipHostFree(pinOut_[i]));
            pinOut_[i] = nullptr;
        }
        if (!pinOut_[i]) PT_CHECK(hipHostMalloc(&pinOut_[i], std::max<size_t>(outBytes, 256), hipHostMallocDefault));
        streamIn_[i].ensure(inBytes);
        streamOut_[i].ensure(outBytes);
    }
    pinInBytes_ = std::max(pinInBytes_, inBytes);
    pinOutBytes_ = std::max(pinOutBytes_, outBytes);
}

void Engine::stream_release()
{
    for (int i = 0; i < 2; ++i) {
        if (pinIn_[i]) (void)hipHostFree(pinIn_[i]);
        if (pinOut_[i]) (void)hipHostFree(pinOut_[i]);
        pinIn_[i] = nullptr;
        pinOut_[i] = nullptr;
        streamIn_[i].release();
        streamOut_[i].release();
    }
    pinInBytes_ = pinOutBytes_ = 0;
    if (copyStream_) {
        (void)hipStreamSynchronize(copyStream_);
        for (int i = 0; i < 2; ++i) {
            (void)hipEventDestroy(inEv_[i]);
            (void)hipEventDestroy(outEv_[i]);
            (void)hipEventDestroy(inStartEv_[i]);
            (void)hipEventDestroy(outStartEv_[i]);
            (void)hipEventDestroy(computeEv_[i]);
        }
        (void)hipStreamDestroy(copyStream_);
        copyStream_ = nullptr;
    }
}

uint64_t Engine::localize_host(const uint16_t* input, double* output, uint64_t outputSize, int H, int W, uint64_t planes,
                               double thr, double ratio, uint64_t fit, const double* fitParams)
{
    const uint64_t maxPer = planes == 0 ? 0 : (outputSize / kCols) / planes;   // Palm.cpp:51
    if (maxPer == 0 || H <= 0 || W <= 0) return 0;
    const size_t npx = (size_t)W * H;
    const uint64_t S = std::max<uint64_t>(1, std::min<uint64_t>(planes, opt.streamPixels / npx));
    const uint64_t pieces = (planes + S - 1) / S;
    const uint64_t rowsPerPiece = S * std::min<uint64_t>(maxPer, npx / 4 + 2);
    stream_prepare(S * npx * 2, rowsPerPiece * kCols * 8);
    const int threads = opt.hostThreads;
    double msH2D = 0, msD2H = 0;
    RunStats total{};
    // Host copy into staging (host threads) and device copy of piece p, in blocks of 32 MiB: the
    // upload of a block runs while the host copies the next.
    auto upload = [&](uint64_t p) {
        const uint64_t z0 = p * S, nz = std::min<uint64_t>(S, planes - z0);
        const int b = (int)(p & 1);
        const size_t bytes = nz * npx * 2, block = size_t(32) << 20;
        const char* src = reinterpret_cast<const char*>(input + z0 * npx);
        char* pin = static_cast<char*>(pinIn_[b]);
        char* dev = static_cast<char*>(streamIn_[b].p);
        for (size_t o = 0; o < bytes; o += block) {
            const size_t m = std::min(block, bytes - o);
            parallel_copy(pin + o, src + o, m, threads);
            if (o == 0) PT_CHECK(hipEventRecord(inStartEv_[b], copyStream_));
            PT_CHECK(hipMemcpyAsync(dev + o, pin + o, m, hipMemcpyHostToDevice, copyStream_));
        }
        PT_CHECK(hipEventRecord(inEv_[b], copyStream_));
    };
    std::vector<uint64_t> rowsOf(pieces, 0), offsetOf(pieces, 0);
    std::string helperError;
    auto copy_out = [&](uint64_t p) {   // rows of piece p from staging to the caller's output
        const int b = (int)(p & 1);
        PT_CHECK(hipEventSynchronize(outEv_[b]));
        float ms = 0;
        PT_CHECK(hipEventElapsedTime(&ms, outStartEv_[b], outEv_[b]));
        msD2H += ms;
        if (rowsOf[p]) parallel_copy(output + offsetOf[p] * kCols, pinOut_[b], rowsOf[p] * kCols * 8, threads);
    };
    upload(0);
    uint64_t rows = 0;
    for (uint64_t p = 0; p < pieces; ++p) {
        const int b = (int)(p & 1);
        const uint64_t z0 = p * S, nz = std::min<uint64_t>(S, planes - z0);
        PT_CHECK(hipEventSynchronize(inEv_[b]));
        float ms = 0;
        PT_CHECK(hipEventElapsedTime(&ms, inStartEv_[b], inEv_[b]));
        msH2D += ms;
        // Host side of the neighbours while the device computes this piece: the rows of piece p - 1
        // out of their staging (their download was queued after that piece's computation), the
        // frames of piece p + 1 into theirs (free: the upload of piece p - 1 is complete).
        Joiner helper(std::thread([&, p] {
            try {
                if (p > 0) copy_out(p - 1);
                if (p + 1 < pieces) upload(p + 1);
            } catch (const std::exception& ex) {
                helperError = ex.what();
            }
        }));
        const uint64_t n = localize(streamIn_[b].as<uint16_t>(), nz, H, W, thr, ratio, fit, fitParams, maxPer,
                                    streamOut_[b].as<double>(), rowsPerPiece) / kCols;
        const RunStats& l = last;
        total.frames += l.frames; total.chunks += l.chunks; total.rois += l.rois; total.flaggedFrames += l.flaggedFrames;
        total.multiComponents += l.multiComponents; total.deferredFrames += l.deferredFrames; total.status |= l.status;
        total.gridFrames += l.gridFrames; total.fitHanded += l.fitHanded; total.maskPixels += l.maskPixels;
        total.msAtrous += l.msAtrous; total.msSegment += l.msSegment; total.msFit += l.msFit; total.msAssemble += l.msAssemble;
        total.jacPasses += l.jacPasses; total.trialPasses += l.trialPasses;
        total.groupJacPasses += l.groupJacPasses; total.groupTrialPasses += l.groupTrialPasses;
        if (n) {
            shift_rows_kernel<<<(unsigned)std::min<uint64_t>(4096, (n + 255) / 256), 256, 0, stream>>>(
                streamOut_[b].as<double>(), n, (double)rows, (double)z0);
            PT_CHECK(hipGetLastError());
        }
        PT_CHECK(hipEventRecord(computeEv_[b], stream));
        helper.join();                       // piece p - 1's staging is read out before it is reused
        if (!helperError.empty()) throw Error(helperError);
        PT_CHECK(hipStreamWaitEvent(copyStream_, computeEv_[b], 0));
        PT_CHECK(hipEventRecord(outStartEv_[b], copyStream_));
        if (n) PT_CHECK(hipMemcpyAsync(pinOut_[b], streamOut_[b].p, n * kCols * 8, hipMemcpyDeviceToHost, copyStream_));
        PT_CHECK(hipEventRecord(outEv_[b], copyStream_));
        rowsOf[p] = n;
        offsetOf[p] = rows;
        rows += n;
    }
    copy_out(pieces - 1);
    last = total;
    last.msH2D = msH2D;
    last.msD2H = msD2H;
    last.streamPieces = pieces;
    return rows * kCols;
}

uint64_t Engine::fit_positions_host(const uint16_t* input, uint64_t planes, int H, int W, const double* positions,
                                    uint64_t n, uint64_t fit, const double* fitParams, double* output)
{
    if (n == 0) return 0;
    if (H <= 0 || W <= 0 || planes == 0) throw Error("Fitting: empty stack");
    const size_t npx = (size_t)W * H;
    const uint64_t S = std::max<uint64_t>(1, std::min<uint64_t>(planes, opt.streamPixels / npx));
    const uint64_t pieces = (planes + S - 1) / S;
    // Positions by piece (input order kept inside each piece); positions outside the stack go with
    // the first piece and get failure rows there.
    std::vector<std::vector<uint64_t>> idx(pieces);
    for (uint64_t i = 0; i < n; ++i) {
        const double pl = positions[i * kInCols + kInPlane];
        const bool ok = pl >= 1.0 && pl < (double)planes + 1.0;
        idx[ok ? ((uint64_t)pl - 1) / S : 0].push_back(i);
    }
    uint64_t maxN = 0;
    for (const auto& v : idx) maxN = std::max<uint64_t>(maxN, v.size());
    stream_prepare(S * npx * 2, maxN * kCols * 8);
    const int threads = opt.hostThreads;
    double msH2D = 0, msD2H = 0;
    std::vector<double> rel(maxN * kInCols);
    for (uint64_t p = 0; p < pieces; ++p) {
        if (idx[p].empty()) continue;
        const uint64_t z0 = p * S, nz = std::min<uint64_t>(S, planes - z0);
        parallel_copy(pinIn_[0], input + z0 * npx, nz * npx * 2, threads);
        PT_CHECK(hipEventRecord(inStartEv_[0], copyStream_));
        PT_CHECK(hipMemcpyAsync(streamIn_[0].p, pinIn_[0], nz * npx * 2, hipMemcpyHostToDevice, copyStream_));
        PT_CHECK(hipEventRecord(inEv_[0], copyStream_));
        const uint64_t m = idx[p].size();
        for (uint64_t k = 0; k < m; ++k) {
            const double* src = positions + idx[p][k] * kInCols;
            double* dst = rel.data() + k * kInCols;
            std::memcpy(dst, src, kInCols * 8);
            const double pl = src[kInPlane];
            if (pl >= 1.0 && pl < (double)planes + 1.0) dst[kInPlane] = pl - (double)z0;   // plane in the piece
        }
        PT_CHECK(hipEventSynchronize(inEv_[0]));
        float ms = 0;
        PT_CHECK(hipEventElapsedTime(&ms, inStartEv_[0], inEv_[0]));
        msH2D += ms;
        fit_positions(streamIn_[0].as<uint16_t>(), nz, H, W, rel.data(), m, fit, fitParams, streamOut_[0].as<double>());
        PT_CHECK(hipEventRecord(outStartEv_[0], stream));
        PT_CHECK(hipMemcpyAsync(pinOut_[0], streamOut_[0].p, m * kCols * 8, hipMemcpyDeviceToHost, stream));
        PT_CHECK(hipEventRecord(outEv_[0], stream));
        PT_CHECK(hipEventSynchronize(outEv_[0]));
        PT_CHECK(hipEventElapsedTime(&ms, outStartEv_[0], outEv_[0]));
        msD2H += ms;
        const double* rowsP = static_cast<const double*>(pinOut_[0]);
        for (uint64_t k = 0; k < m; ++k) {
            double* dst = output + idx[p][k] * kCols;
            std::memcpy(dst, rowsP + k * kCols, kCols * 8);
            dst[kPlane] = positions[idx[p][k] * kInCols + kInPlane];   // the caller's plane
        }
    }
    last = RunStats{};
    last.frames = planes;
    last.rois = n;
    last.msH2D = msH2D;
    last.msD2H = msD2H;
    last.streamPieces = pieces;
    return n * kCols;
}

} // namespace pt
