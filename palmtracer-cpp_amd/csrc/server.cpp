// Serving loop: Localization requests from several streams (cameras) admitted onto a fixed set of
// engine slots under an HBM budget, with per-request timing.
//
// The reference processes one stack per call (lib/Palm.cpp:55-70); a caller with several cameras
// serialises their chunks. Here each slot is an independent Engine (its own streams and chunk
// buffers) driven by one host thread, so requests on different slots run on the device at the same
// time. A request leaves the queue, in arrival order, as soon as a slot is free and the bytes of the
// requests in flight (their input frames and output rows, resident in HBM) plus its own stay within
// the budget. Each request is computed exactly as a Localization call over the same frames: the
// slots share no device state, so a result does not depend on what else is in flight.
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
    uint64_t bytes = 0;              // HBM held by the request while queued or in flight
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
            if (chunkPixels) ctx_.back()->opt.chunkPixels = chunkPixels;
        }
        for (int i = 0; i < slots; ++i) threads_.emplace_back([this, i] { worker(i); });
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
        if (r.bytes > budget_) throw Error("request larger than the server's HBM budget");
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

    // Admission: FIFO; the head request goes to a free slot when the budget allows it (or nothing
    // else is in flight).
    bool admissible() const
    {
        return !queue_.empty() && (inFlight_ == 0 || inFlight_ + queue_.front().bytes <= budget_);
    }

    void worker(int slot)
    {
        Engine& e = *ctx_[slot];
        for (;;) {
            Request r;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cvWork_.wait(lk, [&] { return admissible() || (stop_ && queue_.empty()); });
                if (queue_.empty()) return;
                r = std::move(queue_.front());
                queue_.pop_front();
                inFlight_ += r.bytes;
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
            {
                std::lock_guard<std::mutex> lk(mu_);
                inFlight_ -= r.bytes;
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
    uint64_t budget_, inFlight_ = 0;
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
        pt::Options base;
        {
            pt::Engine& g = pt::Engine::get();
            std::lock_guard<std::mutex> lk(g.mu);
            base = g.opt;
        }
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
        r.bytes = planes * height * width * 2 + outputCapDoubles * 8;
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
