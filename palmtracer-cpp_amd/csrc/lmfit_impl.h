// FP64 Gaussian Levenberg-Marquardt fit in the reference's operation order (one ROI per lane), and the
// model, Gauss-Jordan inverse and row fill that the cooperative FP64 kernel (lmfit_group.hip) shares.
//
// Reference: lib/fit/GaussianFit.cpp:48-295 (Fit, lm, gain, getError, computeR2),
// lib/fit/GaussianModel.cpp:19-52 (Evaluate), lib/fit/Fitting.cpp:61-175 (fill, preCompute, initP),
// lib/signal_processing/Matrix.cpp:297-355 (InverseGaussJordan).
//
// Schedule: one ROI per lane. A lane walks its ROI's pixels in the reference's raster order, so
// every sum is accumulated in the reference's order, and exp and sincos are bit-equal to the x86-64
// glibc functions the reference's Linux build calls (pt_libm.h). The control flow (lambda
// schedule, clamps, stopping rules, sentinels) and the damped normal equations (Gauss-Jordan with
// row equilibration and partial pivoting) are the reference's, step for step.
#pragma once
#include "pt_common.h"
#include "pt_libm.h"
#include "pt_kernels.h"

namespace pt {
namespace {

constexpr double kSqrt2 = 1.41421356237309504880;
constexpr double kPi = 3.14159265358979323846;

template <int MODE> struct NPar { static constexpr int v = MODE == 1 ? 4 : MODE == 2 ? 5 : MODE == 3 ? 6 : 7; };

// Double to size_t as the reference's x86-64 build converts (truncation; wraps for negatives;
// 2^63 for NaN and out-of-range values).
__device__ inline uint64_t to_size(double v)
{
    if (!(v == v) || v >= 9223372036854775808.0 || v <= -9223372036854775808.0) return 1ull << 63;
    return (uint64_t)(long long)v;
}

// GaussianModel: isotropic (modes 1, 2) or rotated anisotropic (modes 3, 4), FP64.
template <typename T, int MODE>
struct Model {
    static_assert(sizeof(T) == 8, "the reference-order model is FP64");
    T off, amp, sx, sy, sx2, sy2, c, s, c2, s2, cs;
    __device__ Model(const double* p, double sigma, double theta)
    {
        off = (T)p[0];
        amp = (T)p[1];
        if constexpr (MODE <= 2) {
            sx = sy = (T)(MODE == 1 ? sigma : p[MODE == 1 ? 0 : 4]);
            sx2 = sy2 = sx * sx;
            c = 1; s = 0; c2 = 1; s2 = 0; cs = 0;
        } else {
            const double th = MODE == 3 ? theta : p[MODE == 3 ? 0 : 6];
            sx = (T)p[4];
            sy = (T)p[5];
            sx2 = sx * sx;
            sy2 = sy * sy;
            double sn, cn;
            libm_sincos(th, sn, cn);
            c = (T)cn;
            s = (T)sn;
            c2 = c * c; s2 = s * s; cs = c * s;
        }
    }
    __device__ T eval(T x, T y, T* g) const
    {
        const T X2 = x * x, Y2 = y * y;
        if constexpr (MODE <= 2) {
            const T E = libm_exp(-(X2 + Y2) / (T(2) * sx2));
            const T P = amp * E;
            if (g) {
                g[0] = 1;
                g[1] = E;
                g[2] = P * x / sx2;
                g[3] = P * y / sx2;
                if constexpr (MODE == 2) g[4] = P / (sx * sx2) * (X2 + Y2);
            }
            return P + off;
        } else {
        const T XC = x * c, XS = x * s, YC = y * c, YS = y * s;
        const T F1 = (XC - YS) / sx, F2 = (XS + YC) / sy;
        const T E = libm_exp(-((F1 * F1) + (F2 * F2)));
        const T P = amp * E;
        if (g) {
            const T XY = x * y, F3 = T(2) * XY * cs, F4 = (T(2) * XY / sx2) - (T(2) * XY / sy2);
            g[0] = 1;
            g[1] = E;
            g[2] = P * (-(((T(-2) * c2 * x) + (T(2) * y * cs)) / sx2) - (((T(-2) * s2 * x) - (T(2) * y * cs)) / sy2));
            g[3] = P * ((((T(2) * s2 * y) - (T(2) * x * cs)) / sx2) - (((T(-2) * c2 * y) - (T(2) * x * cs)) / sy2));
            g[4] = P * (T(2) * ((XC * XC) - F3 + (YS * YS)) / (sx * sx2));
            g[5] = P * (T(2) * ((XS * XS) + F3 + (YC * YC)) / (sy * sy2));
            if constexpr (MODE == 4) g[6] = P * ((T(2) * cs * (((X2 - Y2) / sx2) + ((Y2 - X2) / sy2))) - (s2 * F4) + (c2 * F4));
        }
        return P + off;
        }
    }
};

// InverseGaussJordan (Matrix.cpp:297-355) on an NP x NP matrix in registers. Returns false on the
// reference's domain_error.
template <int N>
__device__ inline bool gauss_jordan(double (&a)[N][N], double (&o)[N][N])
{
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) o[i][j] = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double scale = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (!isfinite(a[i][j])) return false;
            const double v = fabs(a[i][j]);
            if (scale < v) scale = v;
        }
        if (scale == 0.0) return false;
#pragma unroll
        for (int j = 0; j < N; ++j) a[i][j] /= scale;
        o[i][i] = 1.0 / scale;
    }
    const double tol = 2.220446049250313e-16 * (double)N;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        int pr = i;
#pragma unroll
        for (int j = i + 1; j < N; ++j)
            if (fabs(a[j][i]) > fabs(a[pr][i])) pr = j;
        // Bring the pivot row to row i (register arrays: select by index, no dynamic addressing).
        double piv = 0.0;
#pragma unroll
        for (int j = i; j < N; ++j) if (j == pr) piv = a[j][i];
        if (!isfinite(piv) || fabs(piv) <= tol) return false;
        if (pr != i) {
#pragma unroll
            for (int j = i + 1; j < N; ++j) {
                if (j != pr) continue;
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    double t = a[i][k]; a[i][k] = a[j][k]; a[j][k] = t;
                    t = o[i][k]; o[i][k] = o[j][k]; o[j][k] = t;
                }
            }
        }
        const double pinv = 1.0 / a[i][i];
#pragma unroll
        for (int k = 0; k < N; ++k) { a[i][k] *= pinv; o[i][k] *= pinv; }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            const double fct = a[j][i];
#pragma unroll
            for (int k = 0; k < N; ++k) {
                a[j][k] -= fct * a[i][k];
                o[j][k] -= fct * o[i][k];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (!isfinite(a[i][j]) || !isfinite(o[i][j])) return false;
    return true;
}

// Fitting::fill (Fitting.cpp:61-106).
__device__ inline void fill_row(int mode, double* pt, double x, double y, double z, double seg, double surf, double offset,
                                double intensity, double sx, double sy, double theta, double rssXY, double rssZ)
{
    pt[kId] = 0; pt[kPlane] = 0; pt[kIndex] = 0; pt[kChannel] = -1;
    pt[kX] = x; pt[kY] = y; pt[kZ] = z;
    if (mode == 0) pt[kIntensity] = seg;
    else pt[kIntensity] = (intensity > kEpsilon && sx > kEpsilon && sy > kEpsilon) ? 2.0 * kPi * intensity * sx * sy : -1;
    pt[kSigmaX] = sx; pt[kSigmaY] = sy; pt[kTheta] = theta;
    pt[kQualityXY] = rssXY; pt[kQualityZ] = rssZ;
    pt[kAmplitude] = intensity; pt[kOffset] = offset;
    pt[kSegIntensity] = seg; pt[kSurface] = surf;
    pt[kCircularity] = (sx > kEpsilon && sy > kEpsilon) ? (sy < sx ? sy : sx) / (sx < sy ? sy : sx) : 1;
}

// A ROI's pixels: pixel (i, j), i and j in [-half, half] around the centre pixel (ux, uy), at
// base[(i + half) * rs + (j + half) * cs] (in the frame: rs = W, cs = 1; staged: see kStageMax).
struct Patch {
    const uint16_t* base;
    uint32_t rs, cs;
    int half, roi;
    uint64_t ux, uy;
    __device__ double v(int i, int j) const { return (double)base[(size_t)(i + half) * rs + (size_t)(j + half) * cs]; }
};

__device__ inline Patch frame_patch(const uint16_t* img, int W, int half, int roi, uint64_t ux, uint64_t uy)
{
    return Patch{img + (uy - (uint64_t)half) * (uint64_t)W + (ux - (uint64_t)half), (uint32_t)W, 1u, half, roi, ux, uy};
}

// ROIs of up to kStageMax pixels are copied once per ROI into LDS, pixel k of lane l at
// stage[k * 64 + l] (consecutive lanes in consecutive halfwords), so that the passes over the ROI
// read LDS rather than the frame.
constexpr int kStageMax = 121;   // 11 x 11

// GaussianFit::lm: model, residuals, alpha (upper triangle), beta and chi^2 at p.
template <typename T, int MODE>
__device__ inline T lm_pass(const Patch& pa, const double* p, double sigma, double theta, T (&al)[NPar<MODE>::v][NPar<MODE>::v],
                            T (&be)[NPar<MODE>::v])
{
    constexpr int NP = NPar<MODE>::v;
#pragma unroll
    for (int a = 0; a < NP; ++a) {
        be[a] = 0;
#pragma unroll
        for (int b = 0; b < NP; ++b) al[a][b] = 0;
    }
    T err = 0;
    const Model<T, MODE> m(p, sigma, theta);
    bool circular = false;
    if constexpr (MODE == 4) circular = fabs(p[4] - p[5]) <= 2.220446049250313e-16;
    const T px = (T)p[2], py = (T)p[3];
    for (int i = -pa.half; i <= pa.half; ++i) {
        for (int j = -pa.half; j <= pa.half; ++j) {
            T g[NP];
            const T X = (T)(double)(pa.ux + j), Y = (T)(double)(pa.uy + i);
            const T hat = m.eval(X - px, Y - py, g);
            const T d = (T)pa.v(i, j) - hat;
            err += d * d;
            if constexpr (MODE == 4) if (circular) g[NP - 1] = 0;
#pragma unroll
            for (int a = 0; a < NP; ++a) {
                be[a] += g[a] * d;
#pragma unroll
                for (int b = a; b < NP; ++b) al[a][b] += g[a] * g[b];
            }
        }
    }
    return err;
}

// GaussianFit::getError at pTry.
template <typename T, int MODE>
__device__ inline T lm_error(const Patch& pa, const double* p, double sigma, double theta)
{
    const Model<T, MODE> m(p, sigma, theta);
    const T px = (T)p[2], py = (T)p[3];
    T res = 0;
    for (int i = -pa.half; i <= pa.half; ++i)
        for (int j = -pa.half; j <= pa.half; ++j) {
            const T X = (T)(double)(pa.ux + j), Y = (T)(double)(pa.uy + i);
            const T dx = (T)pa.v(i, j) - m.eval(X - px, Y - py, nullptr);
            res += dx * dx;
        }
    return res;
}

// GaussianFit::computeR2 at p (squared Pearson correlation of data and model; NaN -> 0).
template <typename T, int MODE>
__device__ inline double lm_r2(const Patch& pa, const double* p, double sigma, double theta)
{
    const Model<T, MODE> m(p, sigma, theta);
    const T px = (T)p[2], py = (T)p[3];
    const double n = (double)(pa.roi * pa.roi);
    double sd = 0, sm = 0;
    for (int i = -pa.half; i <= pa.half; ++i)
        for (int j = -pa.half; j <= pa.half; ++j) {
            const T X = (T)(double)(pa.ux + j), Y = (T)(double)(pa.uy + i);
            sd = sd + pa.v(i, j);
            sm = sm + (double)m.eval(X - px, Y - py, nullptr);
        }
    const double muD = sd / n, muM = sm / n;
    double t1 = 0, t2 = 0, t3 = 0;
    for (int i = -pa.half; i <= pa.half; ++i)
        for (int j = -pa.half; j <= pa.half; ++j) {
            const T X = (T)(double)(pa.ux + j), Y = (T)(double)(pa.uy + i);
            const double dd = pa.v(i, j) - muD;
            const double hh = (double)m.eval(X - px, Y - py, nullptr) - muM;
            t1 = t1 + dd * hh;
            t2 = t2 + dd * dd;
            t3 = t3 + hh * hh;
        }
    const double r = t1 / sqrt(t2 * t3);
    return isnan(r) ? 0.0 : r * r;
}

// GaussianFit::Fit (GaussianFit.cpp:48-215) as three steps over a per-ROI state, so that a lane can
// run its ROIs one after the other at its own pace: lm_start (border check, initP, the first model
// and Jacobian pass), lm_iterate (one pass of the LM loop body) and lm_finish (R^2 and the row).
template <typename T, int MODE>
struct LmState {
    static constexpr int NP = NPar<MODE>::v;
    Patch pa;
    double p[NP], lo[NP], hi[NP];
    T al[NP][NP], be[NP];
    double dP[NP];
    double err, lambda, m;
    int it;
};

// Outcomes of a step: the loop goes on at s.it + 1; the loop ends (lm_finish writes the row); the
// row is written (failure); the trial point was accepted and the model and Jacobian pass at the new
// point is due (lm_pass, then lm_accepted).
enum : int { kLmContinue = 0, kLmFinish = 1, kLmDone = 2, kLmAccept = 3 };

__device__ inline void fail_row(int mode, const Roi& r, double* row)
{
    fill_row(mode, row, r.x, r.y, 0.0, r.seg, r.surf, -1, -1, -1, -1, -1, -1, -1);
}

// Border check and initP: kLmDone when the row is written already (border), else kLmContinue with
// the state ready for the first model and Jacobian pass (lm_pass), after which lm_started runs. With
// a stage (this lane's slot of the workgroup's LDS, see kStageMax) a small ROI is copied there.
template <typename T, int MODE>
__device__ inline int lm_prepare(const uint16_t* img, int W, int H, const Roi& r, const FitParams& fp, double* row,
                                 LmState<T, MODE>& s, uint16_t* stage)
{
    const int roi = fp.roi, half = roi / 2;
    const uint64_t ux = to_size(r.x), uy = to_size(r.y);
    const uint64_t h64 = (uint64_t)half;
    // Border check (GaussianFit.cpp:51-56), with the reference's unsigned arithmetic.
    // A half side beyond the frame (where the reference's unsigned bound wraps and it reads out of the
    // image) gets the border row.
    if (h64 > (uint64_t)W || h64 > (uint64_t)H || h64 > ux || ux >= (uint64_t)W - h64 || h64 > uy ||
        uy >= (uint64_t)H - h64) {
        fail_row(MODE, r, row);
        return kLmDone;
    }
    s.pa = frame_patch(img, W, half, roi, ux, uy);
    if (stage && roi * roi <= kStageMax) {
        const Patch src = s.pa;
        s.pa.base = stage;
        s.pa.rs = (uint32_t)roi * 64u;
        s.pa.cs = 64u;
        for (int i = 0; i < roi; ++i)
            for (int j = 0; j < roi; ++j) stage[(i * roi + j) * 64] = src.base[(size_t)i * src.rs + j];
    }
    double minI = 1.79769313486231570815e308, maxI = -1.79769313486231570815e308;
    for (int i = -half; i <= half; ++i)
        for (int j = -half; j <= half; ++j) {
            const double v = s.pa.v(i, j);
            if (v < minI) minI = v;
            if (maxI < v) maxI = v;
        }
    // initP (Fitting.cpp:126-175)
    double* p = s.p;
    double* lo = s.lo;
    double* hi = s.hi;
    p[0] = minI; p[1] = maxI - minI; p[2] = r.x; p[3] = r.y;
    lo[0] = 0; lo[1] = 0; lo[2] = r.x - 1; lo[3] = r.y - 1;
    hi[0] = 65535.0; hi[1] = 65535.0; hi[2] = r.x + 1; hi[3] = r.y + 1;
    if constexpr (MODE == 2) { p[4] = fp.sigmaX; lo[4] = 0.1; hi[4] = 20.0; }
    if constexpr (MODE >= 3) {
        p[4] = kSqrt2 * fp.sigmaX; p[5] = kSqrt2 * fp.sigmaY;
        lo[4] = lo[5] = 0.1; hi[4] = hi[5] = 20.0;
    }
    if constexpr (MODE == 4) { p[6] = fp.theta; lo[6] = -150; hi[6] = 150; }
    return kLmContinue;
}

// After the first pass (s.err, s.al, s.be at the initial point): kLmDone when the row is written
// (non-finite), else kLmContinue with s.it at the loop's first value.
template <typename T, int MODE>
__device__ inline int lm_started(LmState<T, MODE>& s, const Roi& r, const FitParams& fp, double* row)
{
    constexpr int NP = NPar<MODE>::v;
    bool bad = !isfinite(s.err);
#pragma unroll
    for (int a = 0; a < NP; ++a) bad |= !isfinite((double)s.be[a]);
    if (bad) {
        fail_row(MODE, r, row);
        return kLmDone;
    }
    double mb = fabs((double)s.be[0]);
#pragma unroll
    for (int a = 1; a < NP; ++a) { const double v = fabs((double)s.be[a]); if (mb < v) mb = v; }
    s.lambda = 1e-2;
    s.m = mb;
    s.it = (mb < 1e-6) ? fp.maxIter : 0;
    return kLmContinue;
}

template <typename T, int MODE>
__device__ inline double lm_rel_step(const LmState<T, MODE>& s)
{
    constexpr int NP = NPar<MODE>::v;
    double m = 0.0;
#pragma unroll
    for (int a = 0; a < NP; ++a) {
        const double ap = fabs(s.p[a]);
        const double q = fabs(s.dP[a]) / (ap < 1e-12 ? 1e-12 : ap);
        if (m < q) m = q;
    }
    return m;
}

// The end of the loop body (after the accept / reject): the stopping rules.
template <typename T, int MODE>
__device__ inline int lm_tail(LmState<T, MODE>& s, const FitParams& fp)
{
    constexpr int NP = NPar<MODE>::v;
    s.m = lm_rel_step<T, MODE>(s);
    if (s.it >= 2) {
        if (s.m < 1e-6) return kLmFinish;
        if (s.err / (double)(fp.roi * fp.roi) < 1e-5) return kLmFinish;
        double mb = fabs((double)s.be[0]);
#pragma unroll
        for (int a = 1; a < NP; ++a) { const double v = fabs((double)s.be[a]); if (mb < v) mb = v; }
        if (mb < 1e-6) return kLmFinish;
    }
    return kLmContinue;
}

// The loop body at s.it up to the trial decision: kLmContinue, kLmFinish, kLmDone, or kLmAccept
// (s.p moved to the trial point; the pass there and lm_accepted follow).
template <typename T, int MODE>
__device__ inline int lm_iterate(LmState<T, MODE>& s, const Roi& r, const FitParams& fp, double* row, uint32_t& nTrial)
{
    constexpr int NP = NPar<MODE>::v;
    double* p = s.p;
    double* dP = s.dP;
    double pTry[NP];
    double A[NP][NP], Ai[NP][NP];
#pragma unroll
    for (int a = 0; a < NP; ++a)
#pragma unroll
        for (int b = 0; b < NP; ++b) A[a][b] = (double)(a <= b ? s.al[a][b] : s.al[b][a]);
#pragma unroll
    for (int a = 0; a < NP; ++a) A[a][a] = A[a][a] == 0.0 ? 1.0 : (1.0 + s.lambda) * A[a][a];
    if (!gauss_jordan<NP>(A, Ai)) {
        if (s.lambda < 1e7) { s.lambda = fmin(s.lambda * 11, 1e7); return kLmContinue; }
        fail_row(MODE, r, row);
        return kLmDone;
    }
    bool nonfinite = false;
#pragma unroll
    for (int a = 0; a < NP; ++a) {
        double d = 0.0;
#pragma unroll
        for (int b = 0; b < NP; ++b) d += Ai[a][b] * (double)s.be[b];
        dP[a] = d;
        nonfinite |= !isfinite(d);
        const double v = p[a] + d;
        pTry[a] = v < s.lo[a] ? s.lo[a] : (s.hi[a] < v ? s.hi[a] : v);
    }
    if (nonfinite) {
        fail_row(MODE, r, row);
        return kLmDone;
    }
    bool tiny = true;
#pragma unroll
    for (int a = 0; a < NP; ++a) { dP[a] = pTry[a] - p[a]; tiny &= fabs(dP[a]) < 1e-15; }
    if (tiny) {
        if (s.lambda >= 1e7) return kLmFinish;
        s.lambda = fmin(s.lambda * 11, 1e7);
        return kLmContinue;
    }
    // gain = 2 dP.beta - dP.alpha.dP (GaussianFit.cpp:248-259)
    double quad = 0.0, lin = 0.0;
#pragma unroll
    for (int a = 0; a < NP; ++a) {
        double ti = 0.0;
#pragma unroll
        for (int b = 0; b < NP; ++b) ti += (double)(a <= b ? s.al[a][b] : s.al[b][a]) * dP[b];
        quad += dP[a] * ti;
        lin += dP[a] * (double)s.be[a];
    }
    const double gain = (2.0 * lin) - quad;
    if (gain <= 0.0) {
        if (s.lambda >= 1e7) return kLmFinish;
        s.lambda = fmin(s.lambda * 11, 1e7);
        s.m = lm_rel_step<T, MODE>(s);
        if (s.m < 1e-6 && s.it >= 2) return kLmFinish;
        return kLmContinue;
    }
    double m = 0.0;
#pragma unroll
    for (int a = 0; a < NP; ++a) m = m + dP[a] * ((s.lambda * dP[a]) + (double)s.be[a]);
    s.m = m;
    const double errTry = (double)lm_error<T, MODE>(s.pa, pTry, fp.sigmaX, fp.theta);
    ++nTrial;
    const double rho = (s.err - errTry) / (2.0 * s.m);
    if (rho > 1e-2) {
#pragma unroll
        for (int a = 0; a < NP; ++a) p[a] = pTry[a];
        return kLmAccept;
    }
    if (s.lambda >= 1e7) return kLmFinish;
    s.lambda = fmin(s.lambda * 11, 1e7);
    return lm_tail<T, MODE>(s, fp);
}

// After the pass at an accepted point (s.err, s.al, s.be there).
template <typename T, int MODE>
__device__ inline int lm_accepted(LmState<T, MODE>& s, const FitParams& fp)
{
    s.lambda = fmax(s.lambda / 9, 1e-7);
    return lm_tail<T, MODE>(s, fp);
}

template <typename T, int MODE>
__device__ inline void lm_finish(const LmState<T, MODE>& s, const Roi& r, const FitParams& fp, double* row)
{
    const double* p = s.p;
    const double off = p[0], inten = p[1], x = p[2], y = p[3];
    const double rss = lm_r2<T, MODE>(s.pa, p, fp.sigmaX, fp.theta);
    if (rss < kEpsilon) {
        fill_row(MODE, row, x, y, 0.0, r.seg, r.surf, -1, -1, -1, -1, -1, -1, -1);
        return;
    }
    double sx = fp.sigmaX, sy = fp.sigmaY, th = fp.theta;
    if constexpr (MODE == 1) { sx = sy = fp.sigmaX; th = 0; }
    if constexpr (MODE == 2) { sx = sy = p[4]; th = 0; }
    if constexpr (MODE >= 3) { sx = p[4] / kSqrt2; sy = p[5] / kSqrt2; }
    if constexpr (MODE == 4) th = p[6];
    fill_row(MODE, row, x, y, 0.0, r.seg, r.surf, off, inten, sx, sy, th, rss, -1);
}

__device__ inline void store_row(double* rows, uint32_t i, const double* row)
{
    double* o = rows + (size_t)i * kCols;
#pragma unroll
    for (int c = 0; c < kCols; ++c) o[c] = row[c];
}

// One ROI per lane, each lane running its ROIs (i, i + stride, ...) one after the other at its own
// pace: a lane whose ROI ends takes its next one while the other lanes of the wave go on iterating,
// so a wave is not held by its slowest ROI of every round. Every ROI's arithmetic is that of the
// reference, in its order, whatever the schedule.
template <typename T, int MODE>
__global__ void __launch_bounds__(64) fit_kernel(const uint16_t* __restrict__ frames, size_t npx, int W, int H,
                                                 const Roi* __restrict__ rois, const uint32_t* __restrict__ count,
                                                 uint32_t nHost, FitParams fp, double* __restrict__ rows,
                                                 unsigned long long* __restrict__ passes, const uint32_t* __restrict__ frameMap)
{
    const uint32_t n = count ? *count : nHost;
    const uint32_t stride = gridDim.x * 64;
    uint32_t nJac = 0, nTrial = 0;
    uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if constexpr (MODE == 0) {
        for (; i < n; i += stride) {
            const Roi r = rois[i];
            double row[kCols];
            fill_row(0, row, r.x, r.y, 0.0, r.seg, r.surf, 0, 0, -1, -1, -1, -1, -1);
            store_row(rows, i, row);
        }
    } else {
        __shared__ uint16_t sStage[kStageMax * 64];
        uint16_t* stage = sStage + threadIdx.x;
        LmState<T, MODE> s;
        Roi r{};
        int phase = 0;   // 0: needs a ROI, 1: in the LM loop, 2: loop ended (its row is due)
        while (true) {
            double row[kCols];
            bool pass = false;    // the model and Jacobian pass at s.p is due in this step
            bool fresh = false;   // ... for a ROI started in this step
            // One loop body for the lanes in the loop.
            if (phase == 1) {
                const int st = lm_iterate<T, MODE>(s, r, fp, row, nTrial);
                if (st == kLmAccept) {
                    pass = true;
                } else if (st == kLmDone) {
                    store_row(rows, i, row);
                    i += stride;
                    phase = 0;
                } else if (st == kLmFinish || ++s.it >= fp.maxIter) {
                    phase = 2;
                }
            }
            // Rows of the loops that ended.
            if (phase == 2) {
                lm_finish<T, MODE>(s, r, fp, row);
                store_row(rows, i, row);
                i += stride;
                phase = 0;
            }
            // Next ROI: border check and initP (border rows written on the way).
            while (phase == 0 && i < n) {
                r = rois[i];
                const size_t fr = frameMap ? frameMap[r.frame] : r.frame;
                if (lm_prepare<T, MODE>(frames + fr * npx, W, H, r, fp, row, s, stage) == kLmDone) {
                    store_row(rows, i, row);
                    i += stride;
                } else {
                    pass = fresh = true;
                    phase = 1;
                }
            }
            if (__ballot(phase != 0) == 0) break;
            // The model and Jacobian passes of this step, of accepted points and new ROIs together.
            if (pass) {
                s.err = (double)lm_pass<T, MODE>(s.pa, s.p, fp.sigmaX, fp.theta, s.al, s.be);
                ++nJac;
            }
            if (fresh) {
                if (lm_started<T, MODE>(s, r, fp, row) == kLmDone) {
                    store_row(rows, i, row);
                    i += stride;
                    phase = 0;
                } else if (s.it >= fp.maxIter) {
                    phase = 2;
                }
            } else if (pass && (lm_accepted<T, MODE>(s, fp) == kLmFinish || ++s.it >= fp.maxIter)) {
                phase = 2;
            }
        }
    }
    if (passes) {
        // Passes over the ROI actually executed (model + Jacobian, and trial model), for the FLOP count.
        unsigned long long a = nJac, b = nTrial;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
        }
        if (threadIdx.x == 0) {
            atomicAdd(&passes[0], a);
            atomicAdd(&passes[1], b);
        }
    }
}

template <typename T>
void launch_fit_t(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, const uint32_t* d_count,
                  uint32_t nHost, const FitParams& fp, double* d_rows, unsigned long long* d_passes,
                  const uint32_t* frameMap, hipStream_t s)
{
    const unsigned grid = d_count ? 4096u : std::max(1u, std::min(65535u, ceil_div(nHost, 64)));
    switch (fp.mode) {
    case 1: fit_kernel<T, 1><<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nHost, fp, d_rows, d_passes, frameMap); break;
    case 2: fit_kernel<T, 2><<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nHost, fp, d_rows, d_passes, frameMap); break;
    case 3: fit_kernel<T, 3><<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nHost, fp, d_rows, d_passes, frameMap); break;
    case 4: fit_kernel<T, 4><<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nHost, fp, d_rows, d_passes, frameMap); break;
    default: fit_kernel<T, 0><<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nHost, fp, d_rows, d_passes, frameMap); break;
    }
    PT_CHECK(hipGetLastError());
}

} // namespace
} // namespace pt
