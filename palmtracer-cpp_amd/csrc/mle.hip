// Poisson maximum-likelihood fit of the Gaussian model (fit modes 1-4), an optional objective beside
// the reference's least squares (lib/fit/GaussianFit.cpp:48-295). FP64, one ROI per lane.
//
// Same ROI, border rule, initial values and bounds as the reference's fit (GaussianFit.cpp:51-56,
// Fitting.cpp:126-175) and the same model and parameters (GaussianModel.cpp:19-52); the objective is
// the Poisson negative log-likelihood of the pixels in photons, with the camera model
// ADU = offset + gain * photons (+ read noise, taken as a Poisson term readVar / gain^2 added to data
// and model, the usual sCMOS treatment):
//     L(p) = sum_i mu_i - d_i log mu_i,  mu_i = (m_i(p) - offset) / gain + rv,  d_i = max(0, (v_i - offset) / gain) + rv.
// Levenberg-Marquardt on Fisher scoring (the published method of Laurence and Chromy, Nat. Methods 7,
// 338 (2010)): step (J + lambda diag J) dp = -grad L with J = sum dmu dmu^T / mu (the expected
// information) and grad L = sum (1 - d / mu) dmu, with parameters held at a bound by an outward
// gradient left out of the step; a step is kept when it lowers L (lambda / 10), else lambda * 10. Stops when L changes by less than 1e-12 relative or the step by less than 1e-9, after
// maxIter iterations, or when lambda passes 1e12. Rows as the reference's (Fitting.cpp:61-106), the
// XY quality column being the reference's R^2 of the final model.
#include "engine.h"
#include "lm_core.h"
#include "lmfit_impl.h"

#include <algorithm>

namespace pt {
namespace {

template <int MODE>
__device__ double mle_nll(const Patch& pa, const double* p, const FitParams& fp, double rv, bool& ok)
{
    const Model<double, MODE> m(p, fp.sigmaX, fp.theta);
    const double gain = fp.camGain, c0 = fp.camOffset;
    double L = 0.0;
    ok = true;
    for (int i = -pa.half; i <= pa.half; ++i)
        for (int j = -pa.half; j <= pa.half; ++j) {
            double g[NPar<MODE>::v];
            const double X = (double)(pa.ux + j), Y = (double)(pa.uy + i);
            const double mu = (m.eval(X - p[2], Y - p[3], g) - c0) / gain + rv;
            const double d = fmax(0.0, (pa.v(i, j) - c0) / gain) + rv;
            if (!(mu > 0.0)) { ok = false; return 0.0; }
            L += mu - (d > 0.0 ? d * log(mu) : 0.0);
        }
    return L;
}

template <int MODE>
__device__ void mle_one(const uint16_t* img, int W, int H, const Roi& r, const FitParams& fp, double* row)
{
    constexpr int NP = NPar<MODE>::v;
    const int roi = fp.roi, half = roi / 2;
    const uint64_t ux = to_size(r.x), uy = to_size(r.y), h64 = (uint64_t)half;
    if (h64 > (uint64_t)W || h64 > (uint64_t)H || h64 > ux || ux >= (uint64_t)W - h64 || h64 > uy ||
        uy >= (uint64_t)H - h64) {
        fill_row(MODE, row, r.x, r.y, 0.0, r.seg, r.surf, -1, -1, -1, -1, -1, -1, -1);
        return;
    }
    const Patch pa = frame_patch(img, W, half, roi, ux, uy);
    double minI = 1.79769313486231570815e308, maxI = -1.79769313486231570815e308;
    for (int i = -half; i <= half; ++i)
        for (int j = -half; j <= half; ++j) {
            const double v = pa.v(i, j);
            minI = fmin(minI, v);
            maxI = fmax(maxI, v);
        }
    double p[NP], lo[NP], hi[NP];
    p[0] = minI; p[1] = maxI - minI; p[2] = r.x; p[3] = r.y;
    lo[0] = 0; lo[1] = 0; lo[2] = r.x - 1; lo[3] = r.y - 1;
    hi[0] = 65535.0; hi[1] = 65535.0; hi[2] = r.x + 1; hi[3] = r.y + 1;
    if constexpr (MODE == 2) { p[4] = fp.sigmaX; lo[4] = 0.1; hi[4] = 20.0; }
    if constexpr (MODE >= 3) {
        p[4] = kSqrt2 * fp.sigmaX; p[5] = kSqrt2 * fp.sigmaY;
        lo[4] = lo[5] = 0.1; hi[4] = hi[5] = 20.0;
    }
    if constexpr (MODE == 4) { p[6] = fp.theta; lo[6] = -150; hi[6] = 150; }
    // The model must stay above the camera offset: the offset parameter starts above it.
    const double gain = fp.camGain, c0 = fp.camOffset, rv = fp.camReadVar / (gain * gain);
    if (p[0] <= c0) p[0] = c0 + 1e-3 * gain;
    bool ok = true;
    double L = mle_nll<MODE>(pa, p, fp, rv, ok);
    double lambda = 1e-3;
    bool failed = !ok;
    for (int it = 0; it < fp.maxIter && !failed; ++it) {
        // Gradient and expected information at p.
        const Model<double, MODE> m(p, fp.sigmaX, fp.theta);
        constexpr int T = lm::Tri<NP>::v;
        double J[T] = {}, gr[NP] = {};
        const bool circular = MODE == 4 && fabs(p[4] - p[5]) <= 2.220446049250313e-16;
        for (int i = -half; i <= half; ++i)
            for (int j = -half; j <= half; ++j) {
                double g[NP];
                const double X = (double)(ux + j), Y = (double)(uy + i);
                const double mu = (m.eval(X - p[2], Y - p[3], g) - c0) / gain + rv;
                const double d = fmax(0.0, (pa.v(i, j) - c0) / gain) + rv;
                if (circular) g[NP - 1] = 0.0;
                const double w = 1.0 / mu, f = 1.0 - d * w;
                for (int a = 0; a < NP; ++a) {
                    const double ga = g[a] / gain;
                    gr[a] += f * ga;
                    for (int b = a; b < NP; ++b) J[lm::tri<NP>(a, b)] += ga * (g[b] / gain) * w;
                }
            }
        // Parameters held at a bound by a gradient that points out of the box are left out of the
        // step (projected Levenberg-Marquardt): a clamped full step would otherwise stall the rest.
        double be[NP], dp[NP];
        for (int a = 0; a < NP; ++a) {
            be[a] = -gr[a];
            if ((p[a] <= lo[a] && gr[a] > 0.0) || (p[a] >= hi[a] && gr[a] < 0.0)) {
                be[a] = 0.0;
                for (int b = 0; b < NP; ++b) J[lm::tri<NP>(a, b)] = a == b ? 1.0 : 0.0;
            }
        }
        bool accepted = false, tiny = false;
        while (lambda <= 1e12) {
            if (lm::ldl_solve<NP, double>(J, be, lambda, dp) != 0) { lambda *= 10.0; continue; }
            double pt[NP], mx = 0.0;
            for (int a = 0; a < NP; ++a) {
                pt[a] = fmin(fmax(p[a] + dp[a], lo[a]), hi[a]);
                mx = fmax(mx, fabs(pt[a] - p[a]));
            }
            bool okT = true;
            const double Lt = mle_nll<MODE>(pa, pt, fp, rv, okT);
            if (okT && Lt < L) {
                const double rel = (L - Lt) / fmax(fabs(L), 1.0);
                for (int a = 0; a < NP; ++a) p[a] = pt[a];
                L = Lt;
                lambda = fmax(lambda * 0.1, 1e-12);
                accepted = true;
                tiny = rel < 1e-12 || mx < 1e-9;
                break;
            }
            if (mx < 1e-9) { tiny = true; break; }
            lambda *= 10.0;
        }
        if (!accepted || tiny) break;
    }
    const double rss = failed ? 0.0 : lm_r2<double, MODE>(pa, p, fp.sigmaX, fp.theta);
    if (failed || rss < kEpsilon) {
        fill_row(MODE, row, p[2], p[3], 0.0, r.seg, r.surf, -1, -1, -1, -1, -1, -1, -1);
        return;
    }
    double sx = fp.sigmaX, sy = fp.sigmaY, th = fp.theta;
    if constexpr (MODE == 1) { sx = sy = fp.sigmaX; th = 0; }
    if constexpr (MODE == 2) { sx = sy = p[4]; th = 0; }
    if constexpr (MODE >= 3) { sx = p[4] / kSqrt2; sy = p[5] / kSqrt2; }
    if constexpr (MODE == 4) th = p[6];
    fill_row(MODE, row, p[2], p[3], 0.0, r.seg, r.surf, p[0], p[1], sx, sy, th, rss, -1);
}

template <int MODE>
__global__ void __launch_bounds__(64) mle_fit_kernel(const uint16_t* __restrict__ frames, size_t npx, int W, int H,
                                                     const Roi* __restrict__ rois, const uint32_t* __restrict__ count,
                                                     uint32_t nHost, FitParams fp, double* __restrict__ rows,
                                                     const uint32_t* __restrict__ frameMap)
{
    const uint32_t n = count ? *count : nHost;
    for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) {
        const Roi r = rois[i];
        const uint16_t* img;
        int w = W, h = H;
        if (fp.stacks) {
            const StackRef sr = fp.stacks[r.slot];
            img = sr.base + (size_t)r.frame * sr.npx;
            w = sr.W;
            h = sr.H;
        } else {
            img = frames + (size_t)(frameMap ? frameMap[r.frame] : r.frame) * npx;
        }
        double row[kCols];
        mle_one<MODE>(img, w, h, r, fp, row);
        double* o = rows + (size_t)i * kCols;
        for (int c = 0; c < kCols; ++c) o[c] = row[c];
    }
}

} // namespace

void launch_fit_mle(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, uint32_t nRoi, const FitParams& fp,
                    double* d_rows, hipStream_t s, const uint32_t* frameMap, const uint32_t* d_count)
{
    if (!(fp.camGain > 0.0)) throw Error("Poisson fit: the camera gain must be positive");
    const unsigned grid = d_count ? 4096u : std::max(1u, std::min(65535u, ceil_div(nRoi, 64)));
    switch (fp.mode) {
    case 1: mle_fit_kernel<1><<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nRoi, fp, d_rows, frameMap); break;
    case 2: mle_fit_kernel<2><<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nRoi, fp, d_rows, frameMap); break;
    case 3: mle_fit_kernel<3><<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nRoi, fp, d_rows, frameMap); break;
    case 4: mle_fit_kernel<4><<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nRoi, fp, d_rows, frameMap); break;
    default: throw Error("Poisson fit: Gaussian fit modes 1-4 only");
    }
    PT_CHECK(hipGetLastError());
}

} // namespace pt
