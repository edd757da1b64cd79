// Cooperative FP64 Levenberg-Marquardt fit: eight lanes per ROI, eight ROIs per wave.
//
// Reference algorithm: lib/fit/GaussianFit.cpp:48-295 (Fit, lm, gain, getError, computeR2),
// lib/fit/GaussianModel.cpp:19-52 (the model and its gradient in FP64, with the reciprocal widths
// precomputed) and lib/fit/Fitting.cpp:61-175 (initP, fill). The damped normal equations
// are solved by an FP64 LDL^T factorisation instead of the reference's Gauss-Jordan inverse
// (lib/signal_processing/Matrix.cpp:297-355): the two differ in the last bits of the step, and their
// singular-matrix tests (pivot against n * DBL_EPSILON after row equilibration; LDL^T pivot against
// n * DBL_EPSILON * a_jj) only part on matrices singular to working precision, which the
// reference's fits do not produce (no Gauss-Jordan failure in 576 000 reference fits of synthetic
// ROIs, modes 2-4).
//
// Role. The persistent FP32 kernel fits most ROIs in a few iterations; the ROIs it hands over (those
// still iterating after a set number of iterations, and those whose FP32 arithmetic met a decision
// that exact arithmetic cannot take: a non-positive gain, a singular damped matrix, a non-finite
// step, a correlation at the failure threshold) are fitted again here from the start in FP64. Fit
// mode 4 (theta free) runs here whole: its theta direction is nearly undetermined on round spots,
// where an FP32 Jacobian steers theta to another point of the flat valley.
//
// Schedule. A group of eight lanes holds one ROI; lane s of a group evaluates the pixels k with
// k % 8 == s (raster index k), so a pass over a 7 x 7 ROI is 7 pixels per lane instead of 49 and a
// ROI that runs to the iteration cap takes a fraction of the time it takes on one lane. The
// per-lane partial sums (chi^2, alpha, beta) are added across the group by three DPP butterfly
// steps (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror); every lane adds the same two
// partial sums in each step, so all eight lanes hold bit-identical sums and take the same control
// path. Every lane of a group runs the LM state machine on the same values. A lane's pixels are
// staged in LDS ([pixel slot][lane]) when the group takes the ROI. As in the FP32 kernel, a trial
// pass also accumulates the Jacobian products at pTry, which the reference evaluates right after
// accepting pTry (GaussianFit.cpp:147-151), so an accepted step needs no extra pass; the group
// keeps one set of sums, and after a rejected step (under 1 % of the steps) a pass at p restores
// the sums there.
// Groups of a wave take ROIs from a per-wave pool of eight indices (one atomic per eight ROIs).
#pragma clang fp contract(off)
#include "lm_core.h"
#include "lmfit_impl.h"

namespace pt {
namespace {

constexpr int kGroup = 8;

__device__ inline double dpp_move(double v, int sel)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    int l2, h2;
    if (sel == 0) {
        l2 = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xF, 0xF, false);     // quad_perm [1,0,3,2]
        h2 = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xF, 0xF, false);
    } else if (sel == 1) {
        l2 = __builtin_amdgcn_mov_dpp(lo, 0x4E, 0xF, 0xF, false);     // quad_perm [2,3,0,1]
        h2 = __builtin_amdgcn_mov_dpp(hi, 0x4E, 0xF, 0xF, false);
    } else {
        l2 = __builtin_amdgcn_mov_dpp(lo, 0x141, 0xF, 0xF, false);    // row_half_mirror
        h2 = __builtin_amdgcn_mov_dpp(hi, 0x141, 0xF, 0xF, false);
    }
    return __hiloint2double(h2, l2);
}

// Sum of v over the eight lanes of the aligned group, identical in all eight lanes.
__device__ inline double group_sum(double v)
{
    v = v + dpp_move(v, 0);
    v = v + dpp_move(v, 1);
    return v + dpp_move(v, 2);
}
__device__ inline double group_min(double v)
{
    v = fmin(v, dpp_move(v, 0));
    v = fmin(v, dpp_move(v, 1));
    return fmin(v, dpp_move(v, 2));
}
__device__ inline double group_max(double v)
{
    v = fmax(v, dpp_move(v, 0));
    v = fmax(v, dpp_move(v, 1));
    return fmax(v, dpp_move(v, 2));
}

// Pixel slots per lane staged in LDS: ROI sides up to 15 (29 slots of 4 bytes per lane, 7.4 KB
// per wave); larger ROIs read the frame in every pass.
constexpr int kGroupMaxStagedRoi = 15;
constexpr int kGroupSlots = (kGroupMaxStagedRoi * kGroupMaxStagedRoi + kGroup - 1) / kGroup;

// gPassInit: first pass at p; gPassTrial: pass at pTry; gPassRestore: pass at p after a rejected
// step (restores alpha and beta at p); gPassR2: correlation of data and model at the final p.
enum : int { gFetch = 0, gDone = 1, gPassInit = 2, gPassTrial = 3, gPassRestore = 4, gPassR2 = 5 };

// R > 0: the ROI side is the compile-time constant R; R = 0: the side is fp.roi. ROIs up to
// kGroupMaxStagedRoi are staged in LDS, larger ones read from the frame in every pass.
// list (optional): ROI indices to fit, *count of them; without it ROIs 0 .. *count (or nHost).
template <int MODE, int R>
__global__ void __launch_bounds__(64) fit_group_kernel(const uint16_t* __restrict__ frames, size_t npx, int W, int H,
                                                       const Roi* __restrict__ rois, const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ count, uint32_t nHost, FitParams fp,
                                                       double* __restrict__ rows, unsigned int* __restrict__ next,
                                                       unsigned long long* __restrict__ passes,
                                                       const uint32_t* __restrict__ frameMap)
{
    constexpr int N = lm::NP<MODE>::v;
    constexpr int T = lm::Tri<N>::v;
    const uint32_t nItems = count ? *count : nHost;
    const int lane = threadIdx.x, sub = lane & (kGroup - 1), grp = lane / kGroup;
    const int roi = R > 0 ? R : fp.roi, half = roi / 2, nPixI = roi * roi;
    const double nPix = (double)nPixI;
    const bool staged = roi <= kGroupMaxStagedRoi;
    __shared__ float patch[kGroupSlots * 64];   // [slot][lane]: pixel sub + 8 * slot of the lane's ROI

    int state = gFetch, it = 0;
    uint32_t idx = 0;
    uint64_t ux = 0, uy = 0;
    const uint16_t* img = nullptr;
    size_t stride = (size_t)W;         // row stride of the ROI's frame
    double x0 = 0, y0 = 0, seg = 0, surf = 0;
    double p[N], pTry[N];
    double al[T], be[N];   // alpha (packed upper triangle) and beta of the last pass
    double lambda = 0, err = 0, mq = 0, relAcc = 0, relRej = 0;
    uint32_t nJac = 0, nTrial = 0;

    // Visits this lane's pixels: f(k, value) for raster indices k = sub, sub + 8, ... below roi^2.
    auto for_pixels = [&](auto&& f) {
        if (staged) {
            for (int k = sub, m2 = 0; k < nPixI; k += kGroup, ++m2) f(k, (double)patch[m2 * 64 + lane]);
        } else {
            for (int k = sub; k < nPixI; k += kGroup) f(k, (double)img[(size_t)(k / roi) * stride + (size_t)(k % roi)]);
        }
    };
    // The group's first lane stores the row.
    auto store_row = [&](const double (&row)[kCols]) {
        if (sub == 0) {
#pragma unroll
            for (int c2 = 0; c2 < kCols; ++c2) rows[(size_t)idx * kCols + c2] = row[c2];
        }
    };
    auto fail_row = [&](double x, double y) {
        double row[kCols];
        fill_row(MODE, row, x, y, 0.0, seg, surf, -1, -1, -1, -1, -1, -1, -1);
        store_row(row);
    };

    uint32_t poolBase = 0, poolUsed = kGroup;
    while (true) {
        // ---- groups without a ROI take the next indices of the wave's pool ----
        while (true) {
            const uint64_t want = __ballot(state == gFetch && sub == 0);
            if (!want) break;
            const uint32_t c = (uint32_t)__popcll(want);
            const uint32_t rank = (uint32_t)__popcll(want & ((1ull << (grp * kGroup)) - 1ull));
            uint32_t item;
            if (poolUsed + c <= (uint32_t)kGroup) {
                item = poolBase + poolUsed + rank;
                poolUsed += c;
            } else {
                uint32_t claim = 0;
                if (lane == 0) claim = atomicAdd(next, (unsigned)kGroup);
                const uint32_t nb = __builtin_amdgcn_readfirstlane(claim);
                const uint32_t k0 = (uint32_t)kGroup - poolUsed;
                item = rank < k0 ? poolBase + poolUsed + rank : nb + (rank - k0);
                poolBase = nb;
                poolUsed = c - k0;
            }
            if (state != gFetch) continue;
            if (item >= nItems) {
                state = gDone;
                continue;
            }
            idx = list ? list[item] : item;
            const Roi r = rois[idx];
            x0 = r.x;
            y0 = r.y;
            seg = r.seg;
            surf = r.surf;
            ux = to_size(r.x);
            uy = to_size(r.y);
            const uint64_t h64 = (uint64_t)half;
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
            if (h64 > (uint64_t)Wl || h64 > (uint64_t)Hl || h64 > ux || ux >= (uint64_t)Wl - h64 || h64 > uy ||
                uy >= (uint64_t)Hl - h64) {
                fail_row(r.x, r.y);
                continue;   // still gFetch
            }
            const size_t fr = frameMap ? frameMap[r.frame] : r.frame;
            stride = (size_t)Wl;
            img = fbase + fr * fnpx + (size_t)(uy - h64) * stride + (ux - h64);
            double mn = 1.79769313486231570815e308, mx = -1.79769313486231570815e308;
            for (int k = sub, m2 = 0; k < nPixI; k += kGroup, ++m2) {
                const uint16_t u = img[(size_t)(k / roi) * stride + (size_t)(k % roi)];
                if (staged) patch[m2 * 64 + lane] = (float)u;
                mn = fmin(mn, (double)u);
                mx = fmax(mx, (double)u);
            }
            mn = group_min(mn);
            mx = group_max(mx);
            // initP (Fitting.cpp:126-175)
            p[0] = mn; p[1] = mx - mn; p[2] = r.x; p[3] = r.y;
            if constexpr (MODE == 2) p[4] = fp.sigmaX;
            if constexpr (MODE >= 3) { p[4] = kSqrt2 * fp.sigmaX; p[5] = kSqrt2 * fp.sigmaY; }
            if constexpr (MODE == 4) p[6] = fp.theta;
            lambda = 1e-2;
            state = gPassInit;
        }
        if (__ballot(state != gDone) == 0) break;
        if (state < gPassInit) continue;

        if (state == gPassR2) {
            // computeR2 at p (GaussianFit.cpp:274-295): squared Pearson correlation of the data and
            // the model, centred sums; NaN -> 0.
            lm::Model<MODE, double> m;
            m.set(p, fp.sigmaX, fp.theta);
            const double bx = (double)(ux - half) - p[2], by = (double)(uy - half) - p[3];
            double sd = 0, sm = 0;
            for_pixels([&](int k, double d) {
                sd = sd + d;
                sm = sm + (m.off + m.amp * m.E(bx + (double)(k % roi), by + (double)(k / roi)));
            });
            sd = group_sum(sd);
            sm = group_sum(sm);
            const double muD = sd / nPix, muM = sm / nPix;
            double t1 = 0, t2 = 0, t3 = 0;
            for_pixels([&](int k, double d) {
                const double dd = d - muD, hh = (m.off + m.amp * m.E(bx + (double)(k % roi), by + (double)(k / roi))) - muM;
                t1 = t1 + dd * hh;
                t2 = t2 + dd * dd;
                t3 = t3 + hh * hh;
            });
            t1 = group_sum(t1);
            t2 = group_sum(t2);
            t3 = group_sum(t3);
            const double rr = t1 / sqrt(t2 * t3);
            const double rss = isnan(rr) ? 0.0 : rr * rr;
            if (rss < kEpsilon) {
                fail_row(p[2], p[3]);
            } else {
                double sx = fp.sigmaX, sy = fp.sigmaY, th = fp.theta;
                if constexpr (MODE == 1) { sx = sy = fp.sigmaX; th = 0; }
                if constexpr (MODE == 2) { sx = sy = p[4]; th = 0; }
                if constexpr (MODE >= 3) { sx = p[4] / kSqrt2; sy = p[5] / kSqrt2; }
                if constexpr (MODE == 4) th = p[6];
                double row[kCols];
                fill_row(MODE, row, p[2], p[3], 0.0, seg, surf, p[0], p[1], sx, sy, th, rss, -1);
                store_row(row);
            }
            state = gFetch;
            continue;
        }

        // ---- chi^2 and Jacobian products at p or pTry (GaussianFit::lm) ----
        const bool trial = state == gPassTrial;
        double e = 0.0;
        {
            double q[N];
#pragma unroll
            for (int a = 0; a < N; ++a) q[a] = trial ? pTry[a] : p[a];
#pragma unroll
            for (int a = 0; a < T; ++a) al[a] = 0.0;
#pragma unroll
            for (int a = 0; a < N; ++a) be[a] = 0.0;
            lm::Model<MODE, double> m;
            m.set(q, fp.sigmaX, fp.theta);
            bool circ = false;
            if constexpr (MODE == 4) circ = fabs(q[4] - q[5]) <= 2.220446049250313e-16;
            const double bx = (double)(ux - half) - q[2], by = (double)(uy - half) - q[3];
            // An axis-aligned Gaussian (axy == 0) is a column factor times a row factor: lane s of
            // the group evaluates the factors of columns and rows s and s + 8, and each pixel takes
            // its two factors from the lanes that hold them (ROI sides up to 16).
            const int gbase = grp * kGroup;
            const bool sep = roi <= 2 * kGroup && m.separable();
            double cf0 = 0.0, cf1 = 0.0, rf0 = 0.0, rf1 = 0.0;
            if (sep) {
                const double X0 = bx + (double)sub, Y0 = by + (double)sub;
                cf0 = exp(-(m.ax * X0 * X0));
                rf0 = exp(-(m.ay * Y0 * Y0));
                if (roi > kGroup) {
                    const double X1 = X0 + (double)kGroup, Y1 = Y0 + (double)kGroup;
                    cf1 = exp(-(m.ax * X1 * X1));
                    rf1 = exp(-(m.ay * Y1 * Y1));
                }
            }
            // Every lane of the group runs the same number of pixel slots, so that the shuffles read
            // factors from active lanes; a slot past the ROI contributes nothing.
            const int slots = (nPixI + kGroup - 1) / kGroup;
            for (int m2 = 0; m2 < slots; ++m2) {
                const int k = sub + m2 * kGroup;
                const bool valid = k < nPixI;
                const int kk = valid ? k : 0;
                const int i = kk / roi, j = kk % roi;
                const double X = bx + (double)j, Y = by + (double)i;
                double E;
                if (sep) {
                    const double c0 = __shfl(cf0, gbase + (j & (kGroup - 1))), r0 = __shfl(rf0, gbase + (i & (kGroup - 1)));
                    if (roi > kGroup) {
                        const double c1 = __shfl(cf1, gbase + (j & (kGroup - 1))), r1 = __shfl(rf1, gbase + (i & (kGroup - 1)));
                        E = (j < kGroup ? c0 : c1) * (i < kGroup ? r0 : r1);
                    } else {
                        E = c0 * r0;
                    }
                } else {
                    E = m.E(X, Y);
                }
                if (!valid) continue;
                const double d = staged ? (double)patch[m2 * 64 + lane]
                                        : (double)img[(size_t)i * stride + (size_t)j];
                const double res = d - (m.off + m.amp * E);
                e += res * res;
                double g[N];
                m.grad(X, Y, E, g);
                if constexpr (MODE == 4)
                    if (circ) g[N - 1] = 0.0;
#pragma unroll
                for (int a = 0; a < N; ++a) {
                    be[a] += g[a] * res;
#pragma unroll
                    for (int b = a; b < N; ++b) al[lm::tri<N>(a, b)] += g[a] * g[b];
                }
            }
        }
        e = group_sum(e);
#pragma unroll
        for (int a = 0; a < T; ++a) al[a] = group_sum(al[a]);
#pragma unroll
        for (int a = 0; a < N; ++a) be[a] = group_sum(be[a]);

        // ---- advance the LM state machine to its next pass (GaussianFit.cpp:75-165) ----
        auto max_abs_beta = [&]() {
            double mb = fabs(be[0]);
#pragma unroll
            for (int a = 1; a < N; ++a) { const double v = fabs(be[a]); mb = mb < v ? v : mb; }
            return mb;
        };
        auto stop_after_step = [&](double rel) {
            if (it >= 2 && (rel < 1e-6 || err / nPix < 1e-5 || max_abs_beta() < 1e-6)) return true;
            ++it;
            return false;
        };
        bool finish = false, failed = false, restore = false;
        if (state == gPassInit) {
            ++nJac;
            err = e;
            bool bad = !isfinite(err);
#pragma unroll
            for (int a = 0; a < N; ++a) bad |= !isfinite(be[a]);
            if (bad) failed = true;
            else it = (max_abs_beta() < 1e-6) ? fp.maxIter : 0;
        } else if (state == gPassRestore) {
            // alpha and beta at p again (err is unchanged: the same pass at the same p).
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
            } else if (lambda >= 1e7) {
                finish = true;
            } else {
                // Rejected: the stopping rules read beta at p, which the pass at pTry overwrote, so
                // the restore pass comes first; it then continues with the damped solves.
                lambda = fmin(lambda * 11, 1e7);
                restore = true;
            }
        }
        if (restore) {
            state = gPassRestore;
            continue;
        }
        if (state == gPassRestore) finish = stop_after_step(relRej);
        // Damped solves until a trial pass is needed (GaussianFit.cpp:83-144).
        while (!finish && !failed) {
            if (it >= fp.maxIter) { finish = true; break; }
            double dP[N];
            const int rc = lm::ldl_solve<N, double>(al, be, lambda, dP);
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
                const double v = p[a] + dP[a];
                const double lo = a <= 1 ? 0.0 : (a == 2 ? x0 - 1 : (a == 3 ? y0 - 1 : (a == 6 ? -150.0 : 0.1)));
                const double hi = a <= 1 ? 65535.0 : (a == 2 ? x0 + 1 : (a == 3 ? y0 + 1 : (a == 6 ? 150.0 : 20.0)));
                pTry[a] = v < lo ? lo : (hi < v ? hi : v);
                dP[a] = pTry[a] - p[a];
                tiny &= fabs(dP[a]) < 1e-15;
                const double ad = fabs(dP[a]);
                const double ra = ad / fmax(fabs(pTry[a]), 1e-12), rr = ad / fmax(fabs(p[a]), 1e-12);
                relAcc = relAcc < ra ? ra : relAcc;
                relRej = relRej < rr ? rr : relRej;
                lin = lin + dP[a] * be[a];
            }
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
                for (int b = 0; b < N; ++b) ti += al[lm::tri<N>(a, b)] * dP[b];
                quad += dP[a] * ti;
            }
            if ((2.0 * lin) - quad <= 0.0) {
                if (lambda >= 1e7) { finish = true; break; }
                lambda = fmin(lambda * 11, 1e7);
                if (relRej < 1e-6 && it >= 2) { finish = true; break; }
                ++it;
                continue;
            }
            mq = 0.0;
#pragma unroll
            for (int a = 0; a < N; ++a) mq = mq + dP[a] * ((lambda * dP[a]) + be[a]);
            state = gPassTrial;
            break;
        }
        if (failed) {
            fail_row(x0, y0);
            state = gFetch;
        } else if (finish) {
            state = gPassR2;
        }
    }
    if (passes) {
        // Passes of the reference's schedule, counted once per group.
        unsigned long long a = sub == 0 ? nJac : 0, b = sub == 0 ? nTrial : 0;
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

__global__ void zero_words_kernel(unsigned int* w, int n)
{
    if ((int)threadIdx.x < n) w[threadIdx.x] = 0u;
}

} // namespace

void launch_fit_group(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, const uint32_t* d_list,
                      const uint32_t* d_count, uint32_t nMax, const FitParams& fp, double* d_rows, unsigned int* d_next,
                      hipStream_t s, unsigned long long* d_passes, const uint32_t* frameMap, unsigned waves)
{
    if (nMax == 0) return;
    if (fp.mode < 1 || fp.mode > 4) throw Error("launch_fit_group: Gaussian fit modes 1-4 only");
    zero_words_kernel<<<1, 64, 0, s>>>(d_next, 1);
    const unsigned want = waves ? waves : 4096u;
    const unsigned grid = std::max(1u, std::min(want, ceil_div(nMax, kGroup)));
    auto go = [&](auto kern) {
        kern<<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_list, d_count, nMax, fp, d_rows, d_next, d_passes, frameMap);
    };
    if (fp.roi == 7) {
        switch (fp.mode) {
        case 1: go(fit_group_kernel<1, 7>); break;
        case 2: go(fit_group_kernel<2, 7>); break;
        case 3: go(fit_group_kernel<3, 7>); break;
        default: go(fit_group_kernel<4, 7>); break;
        }
    } else if (fp.roi == 9) {
        switch (fp.mode) {
        case 1: go(fit_group_kernel<1, 9>); break;
        case 2: go(fit_group_kernel<2, 9>); break;
        case 3: go(fit_group_kernel<3, 9>); break;
        default: go(fit_group_kernel<4, 9>); break;
        }
    } else {
        switch (fp.mode) {
        case 1: go(fit_group_kernel<1, 0>); break;
        case 2: go(fit_group_kernel<2, 0>); break;
        case 3: go(fit_group_kernel<3, 0>); break;
        default: go(fit_group_kernel<4, 0>); break;
        }
    }
    PT_CHECK(hipGetLastError());
}

} // namespace pt
