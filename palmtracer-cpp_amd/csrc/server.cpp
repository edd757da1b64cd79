// Serving loop: Localization requests from several streams (cameras) admitted onto a fixed set of
// engine slots under an HBM budget, with per-request timing.
//
// The reference processes one stack per call (lib/Palm.cpp:55-70); a caller with several cameras
// serialises their chunks. Here each slot is an independent Engine (its own streams and chunk
// buffers) driven by one host thread, so requests on different slots run on the device at the same
// time. A request leaves the queue, in arrival order, as soon as a slot is free and the HBM it needs
// fits the budget beside what is already committed: the input frames and output rows of the requests
// in flight, and every slot's engine work space (chunk buffers, ROI stores, rows), which a slot keeps
// between requests and which grows with the largest request it has run. A running slot counts the
// larger of what it holds and the request's estimated work space (Engine::workspace_bytes, from the
// sizes the engine allocates); an idle slot counts what it holds, and idle slots release their buffers
// when the head request needs the room. Each request is computed exactly as a Localization call over
// the same frames: the slots share no device state, so a result does not depend on what else is in
// flight.
#include "engine.h"
#include "../../include/palmtracer.h"

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

namespace pt {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    const uint16_t* d_in = nullptr;
    uint64_t planes = 0;
    int H = 0, W = 0;
    double thr = 0, ratio = 0;
    uint64_t fit = 0;
    std::vector<double> fitParams;
    double* d_out = nullptr;
    uint64_t outCapDoubles = 0, maxPerPlane = 0;
    uint64_t io = 0;                 // input frames and output rows (caller's HBM)
    uint64_t ws = 0;                 // estimated engine work space of the request
    Clock::time_point tSubmit;
};

struct Result {
    bool done = false, ok = false;
    uint64_t doubles = 0;
    double tSubmit = 0, tStart = 0, tDone = 0;   // ms since the server was created
    int slot = -1;
};

class Server {
public:
    Server(int slots, uint64_t budget, uint64_t chunkPixels, const Options& base) : budget_(budget), t0_(Clock::now())
    {
        for (int i = 0; i < slots; ++i) {
            ctx_.push_back(std::make_unique<Engine>());
            ctx_.back()->opt = base;
            ctx_.back()->opt.profile = false;
            // A request is a chunk of at most a few hundred frames: deferring its flagged frames to the
            // side lane (which detects them again) costs more than it overlaps, so slots resolve them
            // in line.
            ctx_.back()->opt.deferFlagged = 0;
            if (chunkPixels) ctx_.back()->opt.chunkPixels = chunkPixels;
        }
        held_.assign(slots, 0);
        reserved_.assign(slots, 0);
        busy_.assign(slots, false);
        for (int i = 0; i < slots; ++i) threads_.emplace_back([this, i] { worker(i); });
    }

    // HBM admission counts for a request: its input and output plus its estimated work space.
    uint64_t request_bytes(uint64_t planes, int H, int W, uint64_t outCapDoubles, uint64_t maxPerPlane) const
    {
        return planes * (uint64_t)H * (uint64_t)W * 2 + outCapDoubles * 8 +
               ctx_[0]->workspace_bytes(planes, H, W, maxPerPlane, outCapDoubles / kCols);
    }
    uint64_t workspace(const Request& r) const
    {
        return ctx_[0]->workspace_bytes(r.planes, r.H, r.W, r.maxPerPlane, r.outCapDoubles / kCols);
    }

    ~Server()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cvWork_.notify_all();
        for (std::thread& t : threads_) t.join();   // the queue is drained first
    }

    int64_t submit(Request r)
    {
        r.ws = workspace(r);
        if (r.io + r.ws > budget_)
            throw Error("request needs " + std::to_string(r.io + r.ws) + " bytes of HBM (frames, rows and work space), above "
                        "the server's budget of " + std::to_string(budget_));
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) throw Error("server is shutting down");
        r.id = nextId_++;
        r.tSubmit = Clock::now();
        results_[r.id] = Result{};
        results_[r.id].tSubmit = ms(r.tSubmit);
        const int64_t id = r.id;
        queue_.push_back(std::move(r));
        cvWork_.notify_one();
        return id;
    }

    bool wait(int64_t id, Result& out)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = results_.find(id);
        if (it == results_.end()) throw Error("unknown request");
        cvDone_.wait(lk, [&] { return results_[id].done; });
        out = results_[id];
        results_.erase(id);
        return out.ok;
    }

    double now() const { return ms(Clock::now()); }

private:
    double ms(Clock::time_point t) const { return std::chrono::duration<double, std::milli>(t - t0_).count(); }

    // HBM committed by the slots other than s: running slots their reservation, idle slots what they hold.
    uint64_t committed_except(int s) const
    {
        uint64_t t = 0;
        for (size_t k = 0; k < held_.size(); ++k)
            if ((int)k != s) t += busy_[k] ? reserved_[k] : held_[k];
        return t;
    }

    // Admission (FIFO, under mu_): the head request goes to the free slot s when the requests in
    // flight, the other slots' work space and its own fit the budget; idle slots release their
    // buffers (and then s its own) when that makes room. The submission check guarantees that a
    // request fits once nothing else holds HBM.
    bool try_admit(int s)
    {
        if (queue_.empty()) return false;
        const Request& r = queue_.front();
        auto need = [&] { return inFlightIO_ + r.io + committed_except(s) + std::max(held_[s], r.ws); };
        if (need() <= budget_) return true;
        for (size_t k = 0; k < held_.size(); ++k) {
            if ((int)k == s || busy_[k] || held_[k] == 0) continue;
            ctx_[k]->release_device_buffers();   // that slot's thread is idle and waits for mu_
            held_[k] = 0;
            if (need() <= budget_) return true;
        }
        if (held_[s]) {
            ctx_[s]->release_device_buffers();
            held_[s] = 0;
        }
        return need() <= budget_;
    }

    void worker(int slot)
    {
        Engine& e = *ctx_[slot];
        for (;;) {
            Request r;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cvWork_.wait(lk, [&] { return try_admit(slot) || (stop_ && queue_.empty()); });
                if (queue_.empty()) return;
                r = std::move(queue_.front());
                queue_.pop_front();
                inFlightIO_ += r.io;
                busy_[slot] = true;
                reserved_[slot] = std::max(held_[slot], r.ws);
            }
            cvWork_.notify_one();   // another slot may take the next request
            const Clock::time_point tStart = Clock::now();
            bool ok = true;
            uint64_t n = 0;
            try {
                n = e.localize(r.d_in, r.planes, r.H, r.W, r.thr, r.ratio, r.fit,
                               r.fitParams.empty() ? nullptr : r.fitParams.data(), r.maxPerPlane, r.d_out,
                               r.outCapDoubles / kCols);
            } catch (const std::exception& ex) {
                std::fprintf(stderr, "[palmtracer] server request %lld: %s\n", (long long)r.id, ex.what());
                ok = false;
            }
            const Clock::time_point tDone = Clock::now();
            const uint64_t holds = e.device_bytes();
            {
                std::lock_guard<std::mutex> lk(mu_);
                inFlightIO_ -= r.io;
                busy_[slot] = false;
                held_[slot] = holds;
                reserved_[slot] = 0;
                Result& res = results_[r.id];
                res.done = true;
                res.ok = ok;
                res.doubles = n;
                res.tStart = ms(tStart);
                res.tDone = ms(tDone);
                res.slot = slot;
            }
            cvDone_.notify_all();
            cvWork_.notify_all();
        }
    }

    std::mutex mu_;
    std::condition_variable cvWork_, cvDone_;
    std::deque<Request> queue_;
    std::unordered_map<int64_t, Result> results_;
    std::vector<std::unique_ptr<Engine>> ctx_;
    std::vector<std::thread> threads_;
    uint64_t budget_, inFlightIO_ = 0;
    std::vector<uint64_t> held_, reserved_;   // per slot: HBM its engine holds; reservation while running
    std::vector<bool> busy_;
    int64_t nextId_ = 0;
    bool stop_ = false;
    Clock::time_point t0_;
};

} // namespace
} // namespace pt

using pt::Server;

extern "C" {

PT_API void* PT_CALL pt_server_create(uint64_t slots, uint64_t budgetBytes, uint64_t chunkPixels)
{
    try {
        if (slots < 1 || slots > 64) throw pt::Error("slots must be in [1, 64]");
        const pt::Options base = pt::Engine::shared_options();
        return new Server((int)slots, budgetBytes, chunkPixels, base);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_server_create: %s\n", e.what());
        return nullptr;
    }
}

PT_API int64_t PT_CALL pt_server_submit(void* server, const uint16_t* d_input, uint64_t planes, uint64_t height,
                                        uint64_t width, double thresholdVal, double watershedRatio, uint64_t fit,
                                        const double* fit_params, uint64_t nFitParams, double* d_output,
                                        uint64_t outputCapDoubles, uint64_t maxPerPlane)
{
    try {
        if (!server) throw pt::Error("no server");
        if (height == 0 || width == 0 || height > 0x7FFFFFFF || width > 0x7FFFFFFF) throw pt::Error("bad frame size");
        pt::Request r;
        r.d_in = d_input;
        r.planes = planes;
        r.H = (int)height;
        r.W = (int)width;
        r.thr = thresholdVal;
        r.ratio = watershedRatio;
        r.fit = fit;
        if (fit_params && nFitParams) r.fitParams.assign(fit_params, fit_params + nFitParams);
        if (!r.fitParams.empty() && r.fitParams.size() < 4) r.fitParams.resize(4, 0.0);
        r.d_out = d_output;
        r.outCapDoubles = outputCapDoubles;
        r.maxPerPlane = maxPerPlane;
        r.io = planes * height * width * 2 + outputCapDoubles * 8;
        return static_cast<Server*>(server)->submit(std::move(r));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_server_submit: %s\n", e.what());
        return -1;
    }
}

PT_API int PT_CALL pt_server_wait(void* server, int64_t request, uint64_t* doublesWritten, double* times)
{
    try {
        if (!server) throw pt::Error("no server");
        pt::Result res;
        const bool ok = static_cast<Server*>(server)->wait(request, res);
        if (doublesWritten) *doublesWritten = res.doubles;
        if (times) {
            times[0] = res.tSubmit;
            times[1] = res.tStart;
            times[2] = res.tDone;
            times[3] = (double)res.slot;
        }
        return ok ? 1 : 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_server_wait: %s\n", e.what());
        return 0;
    }
}

PT_API uint64_t PT_CALL pt_server_request_bytes(void* server, uint64_t planes, uint64_t height, uint64_t width,
                                                uint64_t outputCapDoubles, uint64_t maxPerPlane)
{
    try {
        if (!server) throw pt::Error("no server");
        if (height == 0 || width == 0 || height > 0x7FFFFFFF || width > 0x7FFFFFFF) throw pt::Error("bad frame size");
        return static_cast<Server*>(server)->request_bytes(planes, (int)height, (int)width, outputCapDoubles, maxPerPlane);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_server_request_bytes: %s\n", e.what());
        return 0;
    }
}

PT_API double PT_CALL pt_server_now(void* server)
{
    return server ? static_cast<Server*>(server)->now() : 0.0;
}

PT_API void PT_CALL pt_server_destroy(void* server)
{
    try {
        delete static_cast<Server*>(server);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_server_destroy: %s\n", e.what());
    }
}

} // extern "C"
