// Astigmatic z estimation on gfx950: the Astigmatism3DEstimation export.
//
// Reference: lib/fit/AstigFit.cpp:36-72 (Estimate), :74-108 (getSigma, getError, getQuality),
// :340-409 (brent); CPU/main.cpp:324-353. For every observed (sigma_x, sigma_y) pair: the z of the
// smallest squared sigma distance on a 201-point grid over [-zMax, zMax], refined by Brent's method
// on the two grid steps around it (kept only when it lowers the distance), then the axial score at
// that z. One thread per pair, FP64 in the reference's operation order with no contraction, so the
// results equal the reference's. The grid's model sigmas do not depend on the pair: each workgroup
// computes them once into LDS.
#include "pt_common.h"
#include "pt_kernels.h"

#pragma clang fp contract(off)

namespace pt {
namespace {

constexpr int kZStep = 201;          // AstigFit.hpp:83
constexpr int kBrentIter = 80;       // AstigFit.hpp:110
constexpr double kBrentTol = 1e-3;   // AstigFit.hpp:111
constexpr double kZTol = 1e-9;       // AstigFit.hpp:112
constexpr double kQualityScale = 0.2;  // AstigFit.hpp:114
constexpr int AST_NT = 256;

struct AstModel {
    double c[10];  // [z0, w, c3, c4, A] of the x axis, then of the y axis
};

// std::min / std::max argument order (ties and NaN keep the first argument).
__device__ inline double smin(double a, double b) { return b < a ? b : a; }
__device__ inline double smax(double a, double b) { return a < b ? b : a; }

// getSigma (AstigFit.cpp:76-88); scale = sampling / pixelSize with sampling 1.
__device__ inline double ast_sigma(const double* c, double z, double scale)
{
    const double u = (z - c[0]) / c[1];
    const double u2 = u * u, u3 = u2 * u, u4 = u3 * u;
    const double poly = 1.0 + u2 + (c[2] * u3) + (c[3] * u4);
    const double sigma = c[4] * __dsqrt_rn(smax(0.0, poly));
    return scale * sigma;
}

__device__ inline double ast_error(const AstModel& m, double z, double sx, double sy, double scale)
{
    const double dx = ast_sigma(m.c, z, scale) - sx, dy = ast_sigma(m.c + 5, z, scale) - sy;
    return (dx * dx) + (dy * dy);
}

__device__ inline double ast_quality(const AstModel& m, double z, double sx, double sy, double scale)
{
    const double px = ast_sigma(m.c, z, scale), py = ast_sigma(m.c + 5, z, scale);
    if (!isfinite(px) || !isfinite(py) || px <= 0 || py <= 0 || !isfinite(sx) || !isfinite(sy) || sx <= 0 || sy <= 0) return 0;
    const double rx = (sx - px) / px / kQualityScale, ry = (sy - py) / py / kQualityScale;
    return 1.0 / (1.0 + (0.5 * ((rx * rx) + (ry * ry))));
}

// brent (AstigFit.cpp:340-409): minimum of the squared sigma distance on [a, b]; returns x, f(x).
__device__ inline double ast_brent(double a, double b, const AstModel& m, double sx, double sy, double scale, double& fxOut)
{
    if (b < a) { const double t = a; a = b; b = t; }
    constexpr double invPhi2 = 1.0 - 0.6180339887498948482;
    double x = a + (invPhi2 * (b - a));
    double w = x, v = x;
    double fx = ast_error(m, x, sx, sy, scale);
    double fw = fx, fv = fx;
    double d = 0.0, e = 0.0;
    for (int iter = 0; iter < kBrentIter; ++iter) {
        const double mid = 0.5 * (a + b);
        const double tol1 = (kBrentTol * fabs(x)) + 1e-12, tol2 = 2.0 * tol1;
        if (fabs(x - mid) <= tol2 - (0.5 * (b - a))) break;
        bool parabolic = false;
        if (fabs(e) > tol1) {
            const double r = (x - w) * (fx - fv);
            double q = (x - v) * (fx - fw);
            double p = ((x - v) * q) - ((x - w) * r);
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            q = fabs(q);
            const double etmp = e;
            e = d;
            if (fabs(p) < fabs(0.5 * q * etmp) && p > q * (a - x) && p < q * (b - x)) {
                parabolic = true;
                d = p / q;
                const double u = x + d;
                if (u - a < tol2 || b - u < tol2) d = x < mid ? tol1 : -tol1;
            }
        }
        if (!parabolic) {
            e = x < mid ? b - x : a - x;
            d = invPhi2 * e;
        }
        const double minStep = d > 0.0 ? tol1 : -tol1;
        const double u = x + (fabs(d) >= tol1 ? d : minStep);
        const double fu = ast_error(m, u, sx, sy, scale);
        if (fu <= fx) {
            if (u < x) b = x;
            else a = x;
            v = w; fv = fw;
            w = x; fw = fx;
            x = u; fx = fu;
        } else {
            if (u < x) a = u;
            else b = u;
            if (fu <= fw || fabs(x - w) <= kZTol) {
                v = w; fv = fw;
                w = u; fw = fu;
            } else if (fu <= fv || fabs(v - x) <= kZTol || fabs(v - w) <= kZTol) {
                v = u; fv = fu;
            }
        }
    }
    fxOut = fx;
    return x;
}

__global__ void __launch_bounds__(AST_NT) astig_estimate_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                                uint64_t n, double pixelSize, AstModel m, double zMax)
{
    __shared__ double gz[kZStep], gsx[kZStep], gsy[kZStep];
    const double scale = 1.0 / pixelSize;
    const double z0 = -zMax, z1 = zMax, deltaZ = z1 - z0;
    for (int i = threadIdx.x; i < kZStep; i += AST_NT) {
        const double t = (double)i / (double)(kZStep - 1);
        const double z = z0 + (t * deltaZ);
        gz[i] = z;
        gsx[i] = ast_sigma(m.c, z, scale);
        gsy[i] = ast_sigma(m.c + 5, z, scale);
    }
    __syncthreads();
    for (uint64_t k = blockIdx.x * (uint64_t)AST_NT + threadIdx.x; k < n; k += (uint64_t)gridDim.x * AST_NT) {
        const double sx = in[2 * k], sy = in[2 * k + 1];
        double bestZ = z0, bestE = __builtin_inf();
        for (int i = 0; i < kZStep; ++i) {
            const double dx = gsx[i] - sx, dy = gsy[i] - sy;
            const double e = (dx * dx) + (dy * dy);
            if (e < bestE) { bestE = e; bestZ = gz[i]; }
        }
        const double dz = deltaZ / (double)(kZStep - 1);
        const double a = smax(z0, smin(bestZ - (2.0 * dz), z1));
        const double b = smax(z0, smin(bestZ + (2.0 * dz), z1));
        double fr = 0.0;
        const double zr = ast_brent(a, b, m, sx, sy, scale, fr);
        const double z = fr < bestE ? zr : bestZ;
        out[2 * k] = z;
        out[2 * k + 1] = ast_quality(m, z, sx, sy, scale);
    }
}

} // namespace

void launch_astig_estimate(const double* d_in, double* d_out, uint64_t n, double pixelSize, const double* model,
                           double zMax, hipStream_t s)
{
    if (n == 0) return;
    AstModel m;
    for (int i = 0; i < 10; ++i) m.c[i] = model[i];
    const uint64_t blocks = std::min<uint64_t>((n + AST_NT - 1) / AST_NT, 65536);
    astig_estimate_kernel<<<(unsigned)blocks, AST_NT, 0, s>>>(d_in, d_out, n, pixelSize, m, zMax);
    PT_CHECK(hipGetLastError());
}

} // namespace pt
