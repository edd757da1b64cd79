// Persistent FP32 Levenberg-Marquardt fit: one ROI per lane, lanes refill from a work queue.
//
// Reference algorithm: lib/fit/GaussianFit.cpp:48-295 (same parameters, bounds, lambda schedule,
// clamps, gain ratio, stopping rules and sentinels; lib/fit/Fitting.cpp:61-175 for initP / fill).
//
// Schedule. Each lane runs the reference's LM for its ROI as a state machine whose states are
// separated by passes over the ROI's pixels: chi^2 with the Jacobian products (at the initial
// parameters, or at a trial step), or the correlation sums of R^2. Every loop turn each lane
// performs the pass its state asks for, then advances to its next pass (damped solves that fail or
// give no gain run in between, as in the reference). A lane whose ROI is finished takes the next
// ROI from a global counter (one atomic per wave), so lanes do not idle behind a slow neighbour
// until the queue runs dry. A lane stages its ROI's pixels in LDS, laid out [pixel][lane], when it
// takes the ROI; every pass then reads LDS instead of the frame.
//
// Precision. Jacobian, alpha and beta in FP32. The model value, residual and chi^2 are FP64: the
// gain ratio and the stopping rules compare chi^2 differences that fall below FP32 resolution near
// convergence (an FP32 exp is good to ~1e-7 relative, i.e. ~5e-4 ADU at a 4000 ADU peak, which
// makes late trial steps accept or reject at random). An axis-aligned Gaussian needs only
// 2 * roi FP64 exps per pass (column factors kept in LDS, one row factor per row). Parameters, the
// damped normal equations (LDL^T of the symmetric matrix), the step and R^2 are FP64.
#include "pt_common.h"
#include "pt_kernels.h"
#include "lm_core.h"

namespace pt {
namespace {

using lm::Tri;
using lm::tri;
using lm::ldl_solve;
using lm::static_for;
using lm::Model;
using lm::NP;

constexpr double kSqrt2d = 1.41421356237309504880;
constexpr double kPid = 3.14159265358979323846;


__device__ inline uint64_t to_size_f(double v)
{
    if (!(v == v) || v >= 9223372036854775808.0 || v <= -9223372036854775808.0) return 1ull << 63;
    return (uint64_t)(long long)v;
}

__device__ inline void row_fill(double* pt, int mode, double x, double y, double seg, double surf, double off, double inten,
                                double sx, double sy, double th, double rss)
{
    pt[kId] = 0; pt[kPlane] = 0; pt[kIndex] = 0; pt[kChannel] = -1;
    pt[kX] = x; pt[kY] = y; pt[kZ] = 0.0;
    pt[kIntensity] = (inten > kEpsilon && sx > kEpsilon && sy > kEpsilon) ? 2.0 * kPid * inten * sx * sy : -1;
    pt[kSigmaX] = sx; pt[kSigmaY] = sy; pt[kTheta] = th;
    pt[kQualityXY] = rss; pt[kQualityZ] = -1;
    pt[kAmplitude] = inten; pt[kOffset] = off;
    pt[kSegIntensity] = seg; pt[kSurface] = surf;
    pt[kCircularity] = (sx > kEpsilon && sy > kEpsilon) ? (sy < sx ? sy : sx) / (sx < sy ? sy : sx) : 1;
    (void)mode;
}

__device__ inline void row_fail(double* pt, int mode, double x, double y, double seg, double surf)
{
    row_fill(pt, mode, x, y, seg, surf, -1, -1, -1, -1, -1, -1);
}

// kPassRestore: a pass at p after a rejected step, which restores alpha, beta and the model sums at p.
enum : int { kFetch = 0, kDone = 1, kPassInit = 2, kPassTrial = 3, kPassRestore = 4 };

// Largest ROI side staged in LDS: pixels (15 x 15 x 64 lanes x 2 bytes = 28.8 KB per wave) and the
// column factors of the separable exponential (15 x 64 x 8 bytes). Larger ROIs read the frame and
// evaluate the exponential per pixel in every pass.
constexpr int kMaxStagedRoi = 15;
// Squared correlation below which a converged FP32 fit is handed to the FP64 group kernel.
constexpr double kLowCorrelation = 0.05;

__host__ __device__ constexpr size_t staged_lds_bytes(int roi)
{
    return (size_t)roi * roi * 64 * sizeof(uint16_t) + (size_t)roi * 64 * sizeof(double);
}

// Every pass also accumulates the Jacobian products: the reference evaluates the Jacobian at pTry
// right after accepting pTry (GaussianFit.cpp:147-151) and chi^2 at pTry is the same sum either way,
// so an accepted trial pass already holds the next alpha, beta and chi^2. A lane keeps one set of
// sums (alpha, beta, the model sums of R^2); after a rejected step (under 1 % of the steps) a pass
// at p restores them.
// R > 0: the ROI side is the compile-time constant R (pixel loops unrolled, column factors in
// registers); R = 0: the side is read from fp.roi.
template <int MODE, bool STAGED, int R, typename G>
__global__ void __launch_bounds__(64) fit_persistent_kernel(const uint16_t* __restrict__ frames, size_t npx, int W, int H,
                                                            const Roi* __restrict__ rois, uint32_t nRoiMax, FitParams fp,
                                                            double* __restrict__ rows, unsigned int* __restrict__ next,
                                                            unsigned long long* __restrict__ passes,
                                                            const uint32_t* __restrict__ frameMap,
                                                            const uint32_t* __restrict__ count,
                                                            uint32_t* __restrict__ hand, unsigned int* __restrict__ handCount)
{
    const uint32_t nRoi = count ? *count : nRoiMax;
    // With a hand-over list, a ROI leaves this kernel (its index appended to the list, its row left to
    // the FP64 group kernel) when it reaches fp.handover iterations or the cap, or when its FP32
    // arithmetic meets a decision that exact arithmetic does not take (see the call sites below).
    const int limit = hand ? min(fp.maxIter, max(fp.handover, 1)) : fp.maxIter;
    constexpr int N = NP<MODE>::v;
    constexpr int T = Tri<N>::v;
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x;
    const int roi = R > 0 ? R : fp.roi, half = roi / 2, nPixI = roi * roi;
    // STAGED: column factors [column][lane] (FP64), then the ROI's pixels [pixel][lane].
    double* colE = reinterpret_cast<double*>(lds);
    uint16_t* patch = reinterpret_cast<uint16_t*>(lds + (size_t)roi * 64 * sizeof(double));
    const double nPix = (double)nPixI;

    int state = kFetch, it = 0;
    uint32_t idx = 0, ux = 0, uy = 0;
    const uint16_t* img = nullptr;     // ROI's top-left pixel in the frame (unstaged reads)
    size_t stride = (size_t)W;         // row stride of the ROI's frame
    double x0 = 0, y0 = 0;             // detection position (bounds, failure rows)
    double p[N], pTry[N];
    double relAcc = 0, relRej = 0;     // max |dP| / max(|p|, 1e-12) with p after acceptance / rejection
    double lambda = 0, err = 0, mq = 0;
    double sd = 0, sdd = 0;            // data sums of the ROI (exact integers)
    double cm = 0, cmm = 0, cdm = 0;   // model sums of the last pass: sum m, sum m^2, sum d*m (R^2)
    G al[T], be[N];                    // alpha (packed upper triangle) and beta of the last pass
    bool satStop = false;              // the rejected step being restored had saturated the damping
    uint32_t nJac = 0, nTrial = 0;

    auto pix = [&](int k, int i, int j) -> float {
        if constexpr (STAGED) return (float)patch[k * 64 + lane];
        else return (float)img[(size_t)i * stride + j];
    };
    // Visit the ROI's pixels in raster order with the model's FP64 exponential E at each:
    // body(k, i, j, X, Y, E) with X, Y the FP64 offsets from the centre (cx, cy).
    auto for_pixels = [&](const Model<MODE, G>& m, double cx, double cy, auto&& body) {
        const double bx = (double)(ux - half) - cx, by = (double)(uy - half) - cy;
        if constexpr (R > 0) {
            const bool sep = m.separable();
            double ce[R];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const double X = bx + j;
                ce[j] = sep ? exp(-(m.ax * X * X)) : 0.0;
            }
#pragma unroll 1
            for (int i = 0; i < R; ++i) {
                const double Y = by + i;
                const double rowE = sep ? exp(-(m.ay * Y * Y)) : 0.0;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const double X = bx + j;
                    const double E = sep ? ce[j] * rowE : m.E(X, Y);
                    body(i * R + j, i, j, X, Y, E);
                }
            }
        } else {
            const bool sep = STAGED && m.separable();
            if (sep)
                for (int j = 0; j < roi; ++j) {
                    const double X = bx + j;
                    colE[j * 64 + lane] = exp(-(m.ax * X * X));
                }
            for (int i = 0, k = 0; i < roi; ++i) {
                const double Y = by + i;
                const double rowE = sep ? exp(-(m.ay * Y * Y)) : 0.0;
                for (int j = 0; j < roi; ++j, ++k) {
                    const double X = bx + j;
                    const double E = sep ? colE[j * 64 + lane] * rowE : m.E(X, Y);
                    body(k, i, j, X, Y, E);
                }
            }
        }
    };

    // Job queue. The wave claims blocks of 64 ROI indices (one atomic per block, the next block
    // claimed one block ahead), and every lane holds its next job prefetched (index and Roi record,
    // loaded one refill ahead), so a refill round waits only for the new ROIs' pixels.
    uint32_t poolBase = 0, poolUsed = 64, claim = 0;
    if (lane == 0) claim = atomicAdd(next, 64u);
    auto next_block = [&]() {
        const uint32_t b = __builtin_amdgcn_readfirstlane(claim);
        claim = 0;
        if (lane == 0) claim = atomicAdd(next, 64u);
        return b;
    };
    uint32_t nIdx = 0;
    Roi nR{};
    double seg = 0, surf = 0;
    {
        poolBase = next_block();
        poolUsed = 64;
        nIdx = poolBase + (uint32_t)lane;
        if (nIdx < nRoi) nR = rois[nIdx];
    }

    while (true) {
        // ---- lanes without a ROI take their prefetched jobs ----
        while (true) {
            const uint64_t want = __ballot(state == kFetch);
            if (!want) break;
            // Refills are batched: a refill round costs the whole wave a memory round trip, so
            // lanes wait for one until enough of them are free or none can run a pass.
            if (__popcll(want) < fp.refill && __ballot(state >= kPassInit) != 0) break;
            // Next prefetch indices from the block pool, in lane order.
            const uint32_t c = (uint32_t)__popcll(want), rank = (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
            uint32_t newIdx;
            if (poolUsed + c <= 64u) {
                newIdx = poolBase + poolUsed + rank;
                poolUsed += c;
            } else {
                const uint32_t k0 = 64u - poolUsed, nb = next_block();
                newIdx = rank < k0 ? poolBase + poolUsed + rank : nb + (rank - k0);
                poolBase = nb;
                poolUsed = c - k0;
            }
            if (state != kFetch) continue;
            idx = nIdx;
            const Roi r = nR;
            nIdx = newIdx;
            if (nIdx < nRoi) nR = rois[nIdx];
            if (idx >= nRoi) {
                state = kDone;
                continue;
            }
            x0 = r.x;
            y0 = r.y;
            seg = r.seg;
            surf = r.surf;
            const uint64_t sx = to_size_f(r.x), sy = to_size_f(r.y), h64 = (uint64_t)half;
            const uint16_t* fbase = frames;
            size_t fnpx = npx;
            int Wl = W, Hl = H;
            if (fp.stacks) {
                const StackRef sr = fp.stacks[r.slot];
                fbase = sr.base;
                fnpx = sr.npx;
                Wl = sr.W;
                Hl = sr.H;
            }
            // Border check with the reference's unsigned arithmetic (GaussianFit.cpp:51-56).
            // A half side beyond the frame (where the reference's unsigned bound wraps and it reads out
            // of the image) gets the border row.
            if (h64 > (uint64_t)Wl || h64 > (uint64_t)Hl || h64 > sx || sx >= (uint64_t)Wl - h64 || h64 > sy ||
                sy >= (uint64_t)Hl - h64) {
                row_fail(rows + (size_t)idx * kCols, MODE, r.x, r.y, r.seg, r.surf);
                continue;  // still kFetch
            }
            ux = (uint32_t)sx;
            uy = (uint32_t)sy;
            const size_t fr = frameMap ? frameMap[r.frame] : r.frame;
            stride = (size_t)Wl;
            img = fbase + fr * fnpx + (size_t)(uy - half) * stride + (ux - half);
            float mn = 3.0e38f, mx = -3.0e38f;
            uint32_t s1 = 0;
            uint64_t s2 = 0;
            for (int i = 0, k = 0; i < roi; ++i)
                for (int j = 0; j < roi; ++j, ++k) {
                    const uint16_t u = img[(size_t)i * stride + j];
                    if constexpr (STAGED) patch[k * 64 + lane] = u;
                    const float v = (float)u;
                    mn = v < mn ? v : mn;
                    mx = mx < v ? v : mx;
                    s1 += u;
                    s2 += (uint64_t)u * u;
                }
            sd = (double)s1;
            sdd = (double)s2;
            // initP (Fitting.cpp:126-175)
            p[0] = mn; p[1] = (double)mx - (double)mn; p[2] = r.x; p[3] = r.y;
            if constexpr (MODE == 2) p[4] = fp.sigmaX;
            if constexpr (MODE >= 3) { p[4] = kSqrt2d * fp.sigmaX; p[5] = kSqrt2d * fp.sigmaY; }
            if constexpr (MODE == 4) p[6] = fp.theta;
            lambda = 1e-2;
            state = kPassInit;
        }
        if (__ballot(state != kDone) == 0) break;
        if (state < kPassInit) continue;

        double* o = rows + (size_t)idx * kCols;
        // ---- chi^2 and Jacobian products at p (initial pass) or pTry (trial pass) ----
        const bool trial = state == kPassTrial;
        double e = 0.0;
        cm = cmm = cdm = 0.0;
#pragma unroll
        for (int a = 0; a < T; ++a) al[a] = G(0);
#pragma unroll
        for (int a = 0; a < N; ++a) be[a] = G(0);
        {
            double q[N];
#pragma unroll
            for (int a = 0; a < N; ++a) q[a] = trial ? pTry[a] : p[a];
            Model<MODE, G> m;
            m.set(q, fp.sigmaX, fp.theta);
            bool circ = false;
            if constexpr (MODE == 4) circ = fabs(q[4] - q[5]) <= 2.220446049250313e-16;
            for_pixels(m, q[2], q[3], [&](int k, int i, int j, double X, double Y, double E) {
                const double d = (double)pix(k, i, j), mm = m.off + m.amp * E;
                const double resd = d - mm;
                e += resd * resd;
                cm += mm;
                cmm += mm * mm;
                cdm += d * mm;
                const G res = (G)resd;
                G g[N];
                m.grad((G)X, (G)Y, (G)E, g);
                if constexpr (MODE == 4)
                    if (circ) g[6] = G(0);
#pragma unroll
                for (int a = 0; a < N; ++a) {
                    be[a] += g[a] * res;
#pragma unroll
                    for (int b = a; b < N; ++b) al[tri<N>(a, b)] += g[a] * g[b];
                }
            });
        }

        // ---- advance the LM state machine to its next pass (GaussianFit.cpp:75-165) ----
        auto max_abs_beta = [&]() {
            double mb = fabs((double)be[0]);
#pragma unroll
            for (int a = 1; a < N; ++a) { const double v = fabs((double)be[a]); mb = mb < v ? v : mb; }
            return mb;
        };
        bool finish = false, failed = false, handOff = false;
        // End-of-iteration stopping rules (GaussianFit.cpp:155-162); otherwise the next iteration.
        auto stop_after_step = [&](double rel) {
            if (it >= 2 && (rel < 1e-6 || err / nPix < 1e-5 || max_abs_beta() < 1e-6)) return true;
            ++it;
            return false;
        };
        if (state == kPassInit) {
            ++nJac;
            err = e;
            bool bad = !isfinite(err);
#pragma unroll
            for (int a = 0; a < N; ++a) bad |= !isfinite(be[a]);
            if (bad) failed = true;
            else it = (max_abs_beta() < 1e-6) ? fp.maxIter : 0;
        } else if (state == kPassRestore) {
            // The sums are at p again (chi^2 there is err): the end of the rejected iteration.
            finish = satStop || stop_after_step(relRej);
        } else {
            ++nTrial;
            const double rho = (err - e) / (2.0 * mq);
            if (rho > 1e-2) {
                ++nJac;
#pragma unroll
                for (int a = 0; a < N; ++a) p[a] = pTry[a];
                err = e;
                lambda = fmax(lambda / 9, 1e-7);
                finish = stop_after_step(relAcc);
            } else if (lambda >= 1e7 && hand) {
                finish = true;
                handOff = true;   // damping saturated: rejected steps all the way up
            } else {
                // Rejected: the stopping rules and the next solve read alpha and beta at p, which this
                // pass at pTry overwrote; the restore pass recomputes them first.
                satStop = lambda >= 1e7;
                lambda = fmin(lambda * 11, 1e7);
                state = kPassRestore;
                continue;
            }
        }
        // Damped solves until a trial pass is needed (GaussianFit.cpp:83-144).
        while (!finish && !failed) {
            if (it >= limit) {
                finish = true;
                handOff = hand != nullptr;
                break;
            }
            double d[N];
            const int rc = ldl_solve<N, G>(al, be, lambda, d);
            if (rc != 0 && hand) { handOff = true; break; }   // singular damped matrix or non-finite step
            if (rc == 1) {
                if (lambda < 1e7) { lambda = fmin(lambda * 11, 1e7); ++it; continue; }
                failed = true;
                break;
            }
            if (rc == 2) { failed = true; break; }
            bool tiny = true;
            relAcc = 0.0;
            relRej = 0.0;
            double lin = 0.0;
#pragma unroll
            for (int a = 0; a < N; ++a) {
                const double v = p[a] + d[a];
                const double lo = a <= 1 ? 0.0 : (a == 2 ? x0 - 1 : (a == 3 ? y0 - 1 : (a == 6 ? -150.0 : 0.1)));
                const double hi = a <= 1 ? 65535.0 : (a == 2 ? x0 + 1 : (a == 3 ? y0 + 1 : (a == 6 ? 150.0 : 20.0)));
                pTry[a] = v < lo ? lo : (hi < v ? hi : v);
                d[a] = pTry[a] - p[a];  // dP
                tiny &= fabs(d[a]) < 1e-15;
                const double ad = fabs(d[a]);
                const double ra = ad / fmax(fabs(pTry[a]), 1e-12), rr = ad / fmax(fabs(p[a]), 1e-12);
                relAcc = relAcc < ra ? ra : relAcc;
                relRej = relRej < rr ? rr : relRej;
                lin += d[a] * (double)be[a];
            }
            if (tiny && hand) { handOff = true; break; }
            if (tiny) {
                if (lambda >= 1e7) { finish = true; break; }
                lambda = fmin(lambda * 11, 1e7);
                ++it;
                continue;
            }
            // gain = 2 dP.beta - dP.alpha.dP (GaussianFit.cpp:248-259)
            double quad = 0.0;
#pragma unroll
            for (int a = 0; a < N; ++a) {
                double ti = 0.0;
#pragma unroll
                for (int b = 0; b < N; ++b) ti += (double)al[tri<N>(a, b)] * d[b];
                quad += d[a] * ti;
            }
            // A non-positive gain needs a clamped step or rounding: exact arithmetic gives
            // dP.(beta + lambda D dP) > 0 for the unclamped step.
            if ((2.0 * lin) - quad <= 0.0 && hand) { handOff = true; break; }
            if ((2.0 * lin) - quad <= 0.0) {
                if (lambda >= 1e7) { finish = true; break; }
                lambda = fmin(lambda * 11, 1e7);
                if (relRej < 1e-6 && it >= 2) { finish = true; break; }
                ++it;
                continue;
            }
            mq = 0.0;
#pragma unroll
            for (int a = 0; a < N; ++a) mq += d[a] * (lambda * d[a] + (double)be[a]);
            state = kPassTrial;
            break;
        }
        if (failed && hand) handOff = true;
        if (finish && !handOff && hand) {
            // A final parameter at (or within FP32 reach of) one of its bounds: whether the FP64 fit
            // ends on the bound is a clamping decision of the last steps, so the ROI is handed over.
            // Bands: 0.5 ADU for offset and amplitude, 0.01 px for the centre, 2 % for the widths.
#pragma unroll
            for (int a = 0; a < N; ++a) {
                if (a == 6) continue;   // theta (mode 4 is fitted by the group kernel)
                const double lo = a <= 1 ? 0.0 : (a == 2 ? x0 - 1 : (a == 3 ? y0 - 1 : 0.1));
                const double hi = a <= 1 ? 65535.0 : (a == 2 ? x0 + 1 : (a == 3 ? y0 + 1 : 20.0));
                const double bl = a <= 1 ? 0.5 : (a <= 3 ? 0.01 : 0.02 * lo), bh = a <= 1 ? 0.5 : (a <= 3 ? 0.01 : 0.02 * hi);
                if (p[a] - lo <= bl || hi - p[a] <= bh) handOff = true;
            }
        }
        double rss = 0.0;
        if (finish && !handOff) {
            // computeR2 at p (GaussianFit.cpp:270-295): squared Pearson correlation of data and model,
            // NaN -> 0, from the sums of the pass that evaluated p.
            const double vd = sdd - sd * sd / nPix, vm = cmm - cm * cm / nPix, cv = cdm - sd * cm / nPix;
            const double rr = cv / sqrt(vd * vm);
            rss = isnan(rr) ? 0.0 : rr * rr;
            // A correlation near the failure threshold (1e-7) is decided in FP64, and so are fits that
            // barely correlate with the data (noise spikes, tangled spots): their minima are shallow
            // and ill-conditioned, where FP32 sums end elsewhere in the valley.
            if (rss < kLowCorrelation && hand) handOff = true;
        }
        if (handOff) {
            hand[atomicAdd(handCount, 1u)] = idx;
            state = kFetch;
        } else if (failed) {
            row_fail(o, MODE, x0, y0, seg, surf);
            state = kFetch;
        } else if (finish) {
            if (rss < kEpsilon) {
                row_fail(o, MODE, p[2], p[3], seg, surf);
            } else {
                double sx = fp.sigmaX, sy = fp.sigmaY, th = fp.theta;
                if constexpr (MODE == 1) { sx = sy = fp.sigmaX; th = 0; }
                if constexpr (MODE == 2) { sx = sy = p[4]; th = 0; }
                if constexpr (MODE >= 3) { sx = p[4] / kSqrt2d; sy = p[5] / kSqrt2d; }
                if constexpr (MODE == 4) th = p[6];
                row_fill(o, MODE, p[2], p[3], seg, surf, p[0], p[1], sx, sy, th, rss);
            }
            state = kFetch;
        }
    }
    if (passes) {
        // Passes of the reference's schedule (Jacobian passes, trial passes) for the FLOP count.
        unsigned long long a = nJac, b = nTrial;
        for (int o2 = 32; o2 > 0; o2 >>= 1) {
            a += __shfl_xor(a, o2);
            b += __shfl_xor(b, o2);
        }
        if (lane == 0) {
            atomicAdd(&passes[0], a);
            atomicAdd(&passes[1], b);
        }
    }
}

// Queue heads and the hand-over count of a fit launch start at zero (a fill kernel, so that the
// sequence can be captured in a graph).
__global__ void reset_words_kernel(unsigned int* w) { if (threadIdx.x < kFitWords) w[threadIdx.x] = 0u; }

} // namespace

void launch_fit(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, uint32_t nRoi, const FitParams& fp,
                double* d_rows, bool fp64, const FitScratch& ws, hipStream_t s, unsigned long long* d_passes,
                const uint32_t* frameMap, unsigned waves, const uint32_t* d_count)
{
    if (nRoi == 0) return;
    if (fp.mode == 5) {
        if (!fp.spline) throw Error("launch_fit: fit mode 5 needs the spline model");
        launch_spline_fit(d_frames, npx, W, H, d_rois, nRoi, *fp.spline, d_rows, s, frameMap, d_count);
        return;
    }
    if (fp.objective == 1 && fp.mode >= 1 && fp.mode <= 4) {
        launch_fit_mle(d_frames, npx, W, H, d_rois, nRoi, fp, d_rows, s, frameMap, d_count);
        return;
    }
    if (fp64 || fp.mode < 1 || fp.mode > 4) {
        launch_fit_f64(d_frames, npx, W, H, d_rois, nRoi, fp, d_rows, s, d_passes, frameMap, d_count);
        return;
    }
    if (!ws.words) throw Error("launch_fit: the fast fit needs its queue words");
    reset_words_kernel<<<1, 64, 0, s>>>(ws.words);
    unsigned long long* groupPasses = d_passes ? d_passes + 2 : nullptr;
    // Fit mode 4 (theta free) is fitted whole in FP64 by the group kernel (see lmfit_group.hip).
    if (fp.mode == 4 && fp.handover > 0) {
        launch_fit_group(d_frames, npx, W, H, d_rois, nullptr, d_count, nRoi, fp, d_rows, ws.words + 2, s, groupPasses,
                         frameMap, 0);
        return;
    }
    const bool handing = fp.handover > 0;
    if (handing && !ws.hand) throw Error("launch_fit: the hand-over needs a list of nRoi words");
    // Persistent waves: by default 16 per CU of the 256-CU part, fewer when there are fewer ROIs.
    const unsigned want = waves ? waves : 4096u;
    const unsigned grid = std::max(1u, std::min(want, ceil_div(nRoi, 64)));
    const bool staged = fp.roi <= kMaxStagedRoi;
    const size_t lds = staged ? staged_lds_bytes(fp.roi) : 0;
    uint32_t* hand = handing ? ws.hand : nullptr;
    auto go = [&](auto kern) {
        kern<<<grid, 64, lds, s>>>(d_frames, npx, W, H, d_rois, nRoi, fp, d_rows, ws.words, d_passes, frameMap, d_count,
                                   hand, ws.words + 1);
    };
    // ROI sides 7 (the reference default) and 9 (astigmatic ROIs) are compiled out.
    if (fp.roi == 7) {
        switch (fp.mode) {
        case 1: go(fit_persistent_kernel<1, true, 7, float>); break;
        case 2: go(fit_persistent_kernel<2, true, 7, float>); break;
        case 3: go(fit_persistent_kernel<3, true, 7, float>); break;
        default: go(fit_persistent_kernel<4, true, 7, float>); break;
        }
    } else if (fp.roi == 9) {
        switch (fp.mode) {
        case 1: go(fit_persistent_kernel<1, true, 9, float>); break;
        case 2: go(fit_persistent_kernel<2, true, 9, float>); break;
        case 3: go(fit_persistent_kernel<3, true, 9, float>); break;
        default: go(fit_persistent_kernel<4, true, 9, float>); break;
        }
    } else {
        switch (fp.mode * 2 + (staged ? 1 : 0)) {
        case 2: go(fit_persistent_kernel<1, false, 0, float>); break;
        case 3: go(fit_persistent_kernel<1, true, 0, float>); break;
        case 4: go(fit_persistent_kernel<2, false, 0, float>); break;
        case 5: go(fit_persistent_kernel<2, true, 0, float>); break;
        case 6: go(fit_persistent_kernel<3, false, 0, float>); break;
        case 7: go(fit_persistent_kernel<3, true, 0, float>); break;
        case 8: go(fit_persistent_kernel<4, false, 0, float>); break;
        default: go(fit_persistent_kernel<4, true, 0, float>); break;
        }
    }
    PT_CHECK(hipGetLastError());
    // The handed-over ROIs, fitted again from the start in FP64 (count read on the device).
    if (handing)
        launch_fit_group(d_frames, npx, W, H, d_rois, ws.hand, ws.words + 1, nRoi, fp, d_rows, ws.words + 2, s,
                         groupPasses, frameMap, 0);
}

} // namespace pt
