// Spline PSF fit with the EMCCD Poisson likelihood (fit mode 5) on gfx950.
//
// Reference: lib/fit/SplineFit.cpp:36-318 (parseParams, Fit, derivativeSpline,
// updateJacobianHessian, updateMuLambda, luEvaluate, delta3D), lib/fit/Fitting.cpp:60-175 (fill,
// preCompute, initP), lib/signal_processing/Matrix.cpp:245-296 (CenterOfMass, Cholesky) and
// :358-411 (InverseLU). One ROI per lane in FP64, every operation in the reference's order and
// without contraction: the Levenberg-Marquardt iterations on the Poisson deviance, then the CRLB as
// the square roots of the diagonal of the inverse Fisher information. The tricubic patch monomials
// of a pass are kept in registers (their derivatives are multiples of them); the coefficients are
// read voxel by voxel from a voxel-major copy of the reference's array (same values and order of
// summation).
#include "pt_common.h"
#include "pt_kernels.h"

#pragma clang fp contract(off)

namespace pt {
namespace {

constexpr int SN = 5;              // offset, intensity, x, y, z
constexpr double kEps2 = 1e-10;    // SplineFit.hpp: EPSILON_2
constexpr int kIter = 75;          // Fitting.hpp: N_ITER

__device__ inline size_t to_size_d(double v)
{
    if (!(v == v) || v >= 9223372036854775808.0 || v <= -9223372036854775808.0) return (size_t)1 << 63;
    return (size_t)(long long)v;
}
__device__ inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

// SqrtReal (SplineFit.cpp:30-34): real part of the complex square root of x + 0i.
__device__ inline double sqrt_real(double x)
{
    if (!(x == x)) return x;
    return x > 0 ? __dsqrt_rn(x) : 0.0;
}

// Fitting::fill (Fitting.cpp:60-101) for the spline modes.
__device__ inline void spline_row(double* o, double x, double y, double z, double seg, double surf, double offset,
                                  double inten, double sxy, double theta, double rssXY, double rssZ)
{
    o[kId] = 0; o[kPlane] = 0; o[kIndex] = 0; o[kChannel] = -1;
    o[kX] = x; o[kY] = y; o[kZ] = z;
    o[kIntensity] = inten;
    o[kSigmaX] = sxy; o[kSigmaY] = sxy; o[kTheta] = theta;
    o[kQualityXY] = rssXY; o[kQualityZ] = rssZ;
    o[kAmplitude] = inten; o[kOffset] = offset;
    o[kSegIntensity] = seg; o[kSurface] = surf;
    o[kCircularity] = 1.0;   // equal sigmas: min / max is 1, and 1 when they are not positive
}

struct Patch {
    double f[64];   // monomials z^i y^j x^k at index 16 i + 4 j + k (delta3D)
};

// delta3D (SplineFit.cpp:296-317): f; the derivatives are k f[q - 1], j f[q - 4], i f[q - 16].
__device__ inline void delta3d(double x, double y, double z, Patch& P)
{
    double cz = 1.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double cy = 1.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double cx = 1.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                P.f[(i * 16) + (j * 4) + k] = cz * cy * cx;
                cx *= x;
            }
            cy *= y;
        }
        cz *= z;
    }
}

// derivativeSpline (SplineFit.cpp:219-240) at the (clamped) voxel v: dudt and the model value.
__device__ inline double deriv(const SplineParams& sp, size_t v, const Patch& P, const double* p, double (&dudt)[SN])
{
    const double* c = sp.cv + v * 64;
    double d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0;
#pragma unroll
    for (int q = 0; q < 64; ++q) {
        const int i = q / 16, j = (q / 4) % 4, k = q % 4;
        const double cq = c[q];
        d1 += P.f[q] * cq;
        d2 += (k > 0 ? (double)k * P.f[q - (k > 0 ? 1 : 0)] : 0.0) * cq;
        d3 += (j > 0 ? (double)j * P.f[q - (j > 0 ? 4 : 0)] : 0.0) * cq;
        d4 += (i > 0 ? (double)i * P.f[q - (i > 0 ? 16 : 0)] : 0.0) * cq;
    }
    dudt[0] = 1;
    dudt[1] = d1;
    dudt[2] = d2 * (-1 * p[1]);
    dudt[3] = d3 * (-1 * p[1]);
    dudt[4] = d4 * p[1];
    return p[0] + (p[1] * dudt[1]);
}

// The voxel index of ROI pixel (i, j) for patch origin (xi, yi, zi) (SplineFit.cpp:115-121 with the
// unsigned clamp of Vec3I::Clamp: negative indices wrap and clamp to the far edge).
__device__ inline size_t voxel(const SplineParams& sp, int xi, int yi, size_t zc, size_t i, size_t j)
{
    size_t px = (size_t)(long long)xi + i + sp.off, py = (size_t)(long long)yi + j + sp.off;
    px = sp.sx - 1 < px ? sp.sx - 1 : px;
    py = sp.sy - 1 < py ? sp.sy - 1 : py;
    return (((zc * sp.sy) + py) * sp.sx) + px;
}

__device__ void spline_fit_one(const uint16_t* img, int W, int H, const Roi& r, const SplineParams& sp, double* o)
{
    const size_t ux = to_size_d(r.x), uy = to_size_d(r.y), half = (size_t)sp.half, roi = (size_t)sp.roi;
    // Border check (SplineFit.cpp:55-60); a half side beyond the frame (where the reference's unsigned
    // bound wraps and it reads out of the image) gets the border row.
    if (half > (size_t)W || half > (size_t)H || half > ux || ux >= (size_t)W - half || half > uy ||
        uy >= (size_t)H - half) {
        spline_row(o, r.x, r.y, 0.0, r.seg, r.surf, -1, -1, -1, -1, -1, -1);
        return;
    }
    const uint16_t* g = img + (uy - half) * (size_t)W + (ux - half);
    auto data = [&](size_t i, size_t j) { return (double)g[i * (size_t)W + j]; };
    // preCompute (Fitting.cpp:109-123) and CenterOfMass (Matrix.cpp:245-269), ROI-local.
    double minI = 1.79769313486231570815e308, maxI = -1.79769313486231570815e308;
    double x = 0.0, y = 0.0, sum = 0.0;
    for (size_t i = 0; i < roi; ++i)
        for (size_t j = 0; j < roi; ++j) {
            const double v = data(i, j);
            minI = minI < v ? minI : v;
            maxI = maxI < v ? v : maxI;
            y += v * (double)i;
            x += v * (double)j;
            sum += v;
        }
    if (sum > 0) { x /= sum; y /= sum; }
    else { x = (double)roi / 2.0; y = (double)roi / 2.0; }
    // initP (Fitting.cpp:159-171) and the spline adjustments (SplineFit.cpp:73-77).
    const double roiC = (sp.roiParam - 1) / 2.0, roiD = sp.roiParam / 4.0;
    const double dmin = roiC - roiD, dmax = roiC + roiD;
    const double lo[SN] = {-0.01, -1, dmin, dmin, 0};
    const double hi[SN] = {0.01, 1, dmax, dmax, (double)sp.sz};
    double p[SN] = {minI, maxI - minI, x, y, sp.initZ};
    p[1] *= sp.ratio;
#pragma unroll
    for (int i = 0; i < SN; ++i) p[i] = clampd(p[i], lo[i], hi[i]);
    double jump[SN];
    jump[0] = 20.0 < p[1] ? p[1] : 20.0;
    jump[1] = 100.0 < p[1] ? p[1] : 100.0;
    jump[2] = 1.0;
    jump[3] = 1.0;
    jump[4] = (double)sp.sz / 3.0 < 2.0 ? 2.0 : (double)sp.sz / 3.0;
    double lambda = 0.1, err = 1.79769313486231570815e308, mu = 1 + lambda;
    double upd[SN] = {0, 0, 0, 0, 0};

    // One pass of the model over the ROI at p: with fisher, the Fisher information; otherwise the
    // deviance, gradient and Hessian (updateJacobianHessian, SplineFit.cpp:243-254).
    auto pass = [&](double* jac, double* hes, double* fisher) {
        const double xr = -1 * (p[2] - (double)half + 0.5), yr = -1 * (p[3] - (double)half + 0.5);
        const int xi = (int)floor(xr), yi = (int)floor(yr), zi = (int)floor(p[4]);
        Patch P;
        delta3d(xr - xi, yr - yi, p[4] - zi, P);
        size_t zc = (size_t)(long long)zi;
        zc = sp.sz - 1 < zc ? sp.sz - 1 : zc;
        for (size_t i = 0; i < roi; ++i)
            for (size_t j = 0; j < roi; ++j) {
                double dudt[SN];
                const double model = deriv(sp, voxel(sp, xi, yi, zc, i, j), P, p, dudt);
                if (fisher) {
#pragma unroll
                    for (int a = 0; a < SN; ++a)
#pragma unroll
                        for (int b = a; b < SN; ++b) {
                            fisher[(a * SN) + b] += dudt[b] * dudt[a] / model;
                            fisher[(b * SN) + a] = fisher[(a * SN) + b];
                        }
                    continue;
                }
                const double d = data(i, j);
                if (d > 0) err = err + (2 * (model - d - (d * log(model / d))));
                else err = err + (2 * model);
                const double t1 = 1 - (d / model);
#pragma unroll
                for (int a = 0; a < SN; ++a) jac[a] += t1 * dudt[a];
                const double t2 = d / (model * model);
#pragma unroll
                for (int a = 0; a < SN; ++a)
#pragma unroll
                    for (int b = a; b < SN; ++b) {
                        hes[(a * SN) + b] += t2 * dudt[a] * dudt[b];
                        hes[(b * SN) + a] = hes[(a * SN) + b];
                    }
            }
    };
    auto mu_lambda = [&](bool up) {   // updateMuLambda (SplineFit.cpp:257-267)
        if (up) {
            const double a = (1 + (lambda * 10)) / (1 + lambda);
            mu = a < 1.3 ? 1.3 : a;
            lambda = 10 * lambda;
        } else {
            lambda = 0.1 * lambda;
            mu = 1 + lambda;
        }
    };

    for (int iteration = 0;;) {
        iteration++;
        const double oldErr = err, oldLambda = lambda;
        double oldP[SN], oldUpd[SN], jac[SN], hes[SN * SN], l[SN * SN], u[SN * SN];
#pragma unroll
        for (int a = 0; a < SN; ++a) { oldP[a] = p[a]; oldUpd[a] = upd[a]; jac[a] = 0.0; upd[a] = 0.0; }
#pragma unroll
        for (int a = 0; a < SN * SN; ++a) { hes[a] = 0.0; l[a] = 0.0; u[a] = 0.0; }
        pass(jac, hes, nullptr);
#pragma unroll
        for (int a = 0; a < SN; ++a) hes[(a * SN) + a] = hes[(a * SN) + a] * mu;
        // Cholesky (Matrix.cpp:272-294)
        bool ok = true;
#pragma unroll
        for (int i = 0; i < SN; ++i)
#pragma unroll
            for (int j = 0; j < i + 1; ++j) {
                if (!ok) continue;
                double s = 0;
#pragma unroll
                for (int k = 0; k < j; ++k) s += u[(i * SN) + k] * u[(j * SN) + k];
                if (i == j) {
                    if (hes[(i * SN) + i] - s >= 0) {
                        u[(i * SN) + j] = __dsqrt_rn(hes[(i * SN) + i] - s);
                        l[(j * SN) + i] = u[(i * SN) + j];
                    } else {
                        ok = false;
                    }
                } else {
                    u[(i * SN) + j] = 1.0 / u[(j * SN) + j] * (hes[(i * SN) + j] - s);
                    l[(j * SN) + i] = u[(i * SN) + j];
                }
            }
        if (ok) {
            // luEvaluate (SplineFit.cpp:270-293)
            double yv[SN];
#pragma unroll
            for (int i = 0; i < SN; ++i) {
                yv[i] = jac[i];
#pragma unroll
                for (int j = 0; j < i; ++j) yv[i] -= l[(j * SN) + i] * yv[j];
                yv[i] /= l[(i * SN) + i];
            }
#pragma unroll
            for (int i = SN - 1; i >= 0; --i) {
                upd[i] = yv[i];
#pragma unroll
                for (int j = i + 1; j < SN; j++) upd[i] -= u[(j * SN) + i] * upd[j];
                upd[i] /= u[(i * SN) + i];
            }
#pragma unroll
            for (int i = 0; i < SN; ++i) {
                if (fabs(oldUpd[i]) > kEpsilon && upd[i] / oldUpd[i] < -0.5) jump[i] *= 0.5;
                const double scale = upd[i] / (1.0 + (fabs(upd[i]) / jump[i]));
                p[i] -= scale;
                p[i] = clampd(p[i], lo[i], hi[i]);
                upd[i] = scale;
            }
            if (err > oldErr) {
                err = oldErr;
#pragma unroll
                for (int a = 0; a < SN; ++a) { p[a] = oldP[a]; upd[a] = oldUpd[a]; }
                lambda = oldLambda;
                mu_lambda(true);
            } else {
                mu_lambda(false);
            }
        } else {
            err = oldErr;
#pragma unroll
            for (int a = 0; a < SN; ++a) { p[a] = oldP[a]; upd[a] = oldUpd[a]; }
            lambda = oldLambda;
            mu_lambda(true);
        }
        if (iteration >= kIter) break;
        if (iteration > 2 && (err < 1e-6 || fabs(err - oldErr) / err < 1e-6)) break;
    }

    // CRLB (SplineFit.cpp:166-203): InverseLU (Matrix.cpp:358-411) of the Fisher information.
    double m[SN * SN], mInv[SN * SN];
#pragma unroll
    for (int a = 0; a < SN * SN; ++a) { m[a] = 0.0; mInv[a] = 0.0; }
    pass(nullptr, nullptr, m);
#pragma unroll
    for (int i = 0; i < SN; ++i)
        if (fabs(m[(i * SN) + i]) < 2.220446049250313e-16) m[(i * SN) + i] = kEps2;
    {
        double tmp1 = 0, yv[SN];
#pragma unroll
        for (int j = 0; j < SN; ++j) {
#pragma unroll
            for (int i = 1; i <= j; ++i) {
#pragma unroll
                for (int k = 0; k <= i - 1; ++k) tmp1 += m[i + (k * SN)] * m[k + (j * SN)];
                m[i + (j * SN)] -= tmp1;
                tmp1 = 0;
            }
#pragma unroll
            for (int i = j + 1; i < SN; ++i) {
                if (j > 0) {
#pragma unroll
                    for (int k = 0; k <= j - 1; ++k) tmp1 += m[i + (k * SN)] * m[k + (j * SN)];
                    m[i + (j * SN)] = (1 / m[j + (j * SN)]) * (m[i + (j * SN)] - tmp1);
                    tmp1 = 0;
                } else {
                    m[i + (j * SN)] = (1 / m[j + (j * SN)]) * m[i + (j * SN)];
                }
            }
        }
#pragma unroll
        for (int num = 0; num < SN; ++num) {
            yv[0] = num == 0 ? 1 : 0;
#pragma unroll
            for (int i = 1; i < SN; ++i) {
                const double b = i == num ? 1 : 0;
#pragma unroll
                for (int j = 0; j <= i - 1; ++j) tmp1 += m[i + (j * SN)] * yv[j];
                yv[i] = b - tmp1;
                tmp1 = 0;
            }
            mInv[SN - 1 + (num * SN)] = yv[SN - 1] / m[(SN - 1) + ((SN - 1) * SN)];
#pragma unroll
            for (int i = SN - 2; i >= 0; --i) {
#pragma unroll
                for (int j = i + 1; j < SN; ++j) tmp1 += m[i + (j * SN)] * mInv[j + (num * SN)];
                mInv[i + (num * SN)] = (1 / m[i + (i * SN)]) * (yv[i] - tmp1);
                tmp1 = 0;
            }
        }
    }
    double crlb[SN];
#pragma unroll
    for (int i = 0; i < SN; ++i) crlb[i] = sqrt_real(mInv[(i * SN) + i]);
    const double rssXY = (crlb[2] + crlb[3]) / 2, rssZ = sp.deltaZ * crlb[4];
    const double fx = p[2] - (double)half + r.x, fy = p[3] - (double)half + r.y;
    const double fz = sp.deltaZ * (p[4] - sp.initZ);
    spline_row(o, fx, fy, fz, r.seg, r.surf, crlb[0], crlb[1], 0, 0, rssXY, rssZ);
}

__global__ void spline_transpose_kernel(const double* __restrict__ raw, double* __restrict__ cv, uint64_t ssize)
{
    const uint64_t n = ssize * 64;
    for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (uint64_t)gridDim.x * blockDim.x)
        cv[o] = raw[(o % 64) * ssize + o / 64];
}

__global__ void __launch_bounds__(64) spline_fit_kernel(const uint16_t* __restrict__ frames, size_t npx, int W, int H,
                                                        const Roi* __restrict__ rois, const uint32_t* __restrict__ count,
                                                        uint32_t nHost, SplineParams sp, double* __restrict__ rows,
                                                        const uint32_t* __restrict__ frameMap)
{
    const uint32_t n = count ? *count : nHost;
    for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) {
        const Roi r = rois[i];
        const size_t fr = frameMap ? frameMap[r.frame] : r.frame;
        double row[kCols];
        spline_fit_one(frames + fr * npx, W, H, r, sp, row);
        double* o = rows + (size_t)i * kCols;
#pragma unroll
        for (int c = 0; c < kCols; ++c) o[c] = row[c];
    }
}

} // namespace

void launch_spline_transpose(const double* d_raw, double* d_cv, uint64_t ssize, hipStream_t s)
{
    if (ssize == 0) return;
    const unsigned grid = (unsigned)std::min<uint64_t>(8192, (ssize * 64 + 255) / 256);
    spline_transpose_kernel<<<grid, 256, 0, s>>>(d_raw, d_cv, ssize);
    PT_CHECK(hipGetLastError());
}

void launch_spline_fit(const uint16_t* d_frames, size_t npx, int W, int H, const Roi* d_rois, uint32_t nRoi,
                       const SplineParams& sp, double* d_rows, hipStream_t s, const uint32_t* frameMap,
                       const uint32_t* d_count)
{
    if (nRoi == 0) return;
    if (!sp.cv) throw Error("launch_spline_fit: spline coefficients are not on the device");
    const unsigned grid = d_count ? 4096u : std::max(1u, std::min(65535u, ceil_div(nRoi, 64)));
    spline_fit_kernel<<<grid, 64, 0, s>>>(d_frames, npx, W, H, d_rois, d_count, nRoi, sp, d_rows, frameMap);
    PT_CHECK(hipGetLastError());
}

} // namespace pt
