// Cramer-Rao lower bounds of the Gaussian fits (fit modes 1-4). The reference states them for its
// spline fit only (lib/fit/SplineFit.cpp:168-205: Fisher information of the model over the ROI,
// inverted, square roots of the diagonal); here the same quantities for the Gaussian model of
// lib/fit/GaussianModel.cpp:19-52 at the fitted parameters of each row, in FP64.
//
// Noise model: each pixel is camera offset + gain x Poisson(photons) + read noise, so a pixel whose
// model value is m (ADU) has variance gain * (m - offset) + readVar (ADU^2). The Fisher information
// is F = sum over the ROI of g g^T / var, g the gradient of the model with respect to the fit's
// parameters [offset, amplitude, x, y, (sigma | sqrt2 sigmaX, sqrt2 sigmaY, (theta))]
// (Fitting.cpp:126-175 parameterisation), and the bound of parameter i is sqrt((F^-1)_ii). The ROI
// is the fit's side centred on the pixel of the fitted position. Intensity (2 pi A sigmaX sigmaY,
// Fitting.cpp:61-106) gets the delta-method bound sqrt(d^T F^-1 d), d its gradient.
//
// The bounds depend on the model at the fitted parameters only (expected information), not on the
// pixel values. Output per row, 8 values in the units of the row's columns: X, Y, SigmaX, SigmaY, Theta,
// Amplitude, Offset, Intensity; NaN for a parameter the mode does not fit, for rows whose fit
// failed (amplitude <= 0), whose ROI leaves the frame, or whose Fisher matrix is not positive
// definite.
#include "pt_common.h"
#include "pt_kernels.h"
#include "engine.h"
#include "lm_core.h"
#include "../../include/palmtracer.h"

#include <algorithm>
#include <cmath>

namespace pt {
namespace {

struct Camera {
    double gain, offset, readVar;
};

constexpr double kPiD = 3.14159265358979323846;
constexpr double kSqrt2D = 1.41421356237309504880;

// In-place inverse of the symmetric positive definite n x n matrix a (row-major) by Cholesky;
// false when it is not numerically positive definite.
template <int N>
__device__ bool spd_inverse(double (&a)[N][N])
{
    double L[N][N] = {};
    for (int j = 0; j < N; ++j) {
        double d = a[j][j];
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        if (!(d > 1e-300) || !isfinite(d)) return false;
        L[j][j] = sqrt(d);
        for (int i = j + 1; i < N; ++i) {
            double v = a[i][j];
            for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
            L[i][j] = v / L[j][j];
        }
    }
    // inv(L), lower triangular
    double M[N][N] = {};
    for (int i = 0; i < N; ++i) {
        M[i][i] = 1.0 / L[i][i];
        for (int j = 0; j < i; ++j) {
            double v = 0.0;
            for (int k = j; k < i; ++k) v -= L[i][k] * M[k][j];
            M[i][j] = v / L[i][i];
        }
    }
    // a^-1 = inv(L)^T inv(L)
    for (int i = 0; i < N; ++i)
        for (int j = i; j < N; ++j) {
            double v = 0.0;
            for (int k = j; k < N; ++k) v += M[k][i] * M[k][j];
            a[i][j] = a[j][i] = v;
        }
    return true;
}

template <int MODE>
__global__ void __launch_bounds__(64) crlb_kernel(int W, int H, const double* __restrict__ rows, uint64_t n, FitParams fp,
                                                  Camera cam, double* __restrict__ out)
{
    constexpr int N = lm::NP<MODE>::v;
    const double nan = __builtin_nan("");
    for (uint64_t r = blockIdx.x * 64ull + threadIdx.x; r < n; r += (uint64_t)gridDim.x * 64) {
        const double* row = rows + r * kCols;
        double* o = out + r * 8;
        double res[8] = {nan, nan, nan, nan, nan, nan, nan, nan};
        const double X = row[kX], Y = row[kY], amp = row[kAmplitude], off = row[kOffset];
        const int half = fp.roi / 2;
        bool ok = amp > 0.0 && isfinite(X) && isfinite(Y);
        const double fx = ok ? floor(X) : 0.0, fy = ok ? floor(Y) : 0.0;
        ok = ok && fx - half >= 0.0 && fx + half < (double)W && fy - half >= 0.0 && fy + half < (double)H;
        if (ok) {
            double p[7] = {off, amp, X, Y, 0.0, 0.0, 0.0};
            if constexpr (MODE == 2) p[4] = row[kSigmaX];
            if constexpr (MODE >= 3) { p[4] = kSqrt2D * row[kSigmaX]; p[5] = kSqrt2D * row[kSigmaY]; }
            if constexpr (MODE == 4) p[6] = row[kTheta];
            lm::Model<MODE, double> m;
            m.set(p, fp.sigmaX, fp.theta);
            // Theta is not estimable from a circular spot (the reference's Jacobian drops it then).
            const bool circ = MODE == 4 && fabs(p[4] - p[5]) <= 2.220446049250313e-16;
            double F[N][N] = {};
            const int cx = (int)fx, cy = (int)fy;
            for (int i = -half; i <= half; ++i)
                for (int j = -half; j <= half; ++j) {
                    const double dx = (double)(cx + j) - X, dy = (double)(cy + i) - Y;
                    const double E = m.E(dx, dy);
                    const double mu = off + amp * E;
                    double g[N];
                    m.grad(dx, dy, E, g);
                    if (circ) g[N - 1] = 0.0;
                    const double var = cam.gain * fmax(mu - cam.offset, 0.0) + cam.readVar;
                    const double w = 1.0 / fmax(var, 1e-12);
                    for (int a = 0; a < N; ++a)
                        for (int b = a; b < N; ++b) F[a][b] += g[a] * g[b] * w;
                }
            for (int a = 0; a < N; ++a)
                for (int b = 0; b < a; ++b) F[a][b] = F[b][a];
            if (circ) F[N - 1][N - 1] = 1.0;
            if (spd_inverse<N>(F)) {
                res[0] = sqrt(F[2][2]);
                res[1] = sqrt(F[3][3]);
                res[5] = sqrt(F[1][1]);
                res[6] = sqrt(F[0][0]);
                double d[N] = {};
                if constexpr (MODE == 1) d[1] = 2.0 * kPiD * fp.sigmaX * fp.sigmaX;
                if constexpr (MODE == 2) {
                    res[2] = res[3] = sqrt(F[4][4]);
                    d[1] = 2.0 * kPiD * p[4] * p[4];
                    d[4] = 4.0 * kPiD * amp * p[4];
                }
                if constexpr (MODE >= 3) {
                    res[2] = sqrt(F[4][4]) / kSqrt2D;
                    res[3] = sqrt(F[5][5]) / kSqrt2D;
                    d[1] = kPiD * p[4] * p[5];
                    d[4] = kPiD * amp * p[5];
                    d[5] = kPiD * amp * p[4];
                }
                if constexpr (MODE == 4) res[4] = circ ? nan : sqrt(F[6][6]);
                double vi = 0.0;
                for (int a = 0; a < N; ++a)
                    for (int b = 0; b < N; ++b) vi += d[a] * F[a][b] * d[b];
                res[7] = sqrt(fmax(vi, 0.0));
            }
        }
        for (int c = 0; c < 8; ++c) o[c] = res[c];
    }
}

} // namespace
} // namespace pt

extern "C" PT_API int PT_CALL pt_crlb_device(const double* d_rows, uint64_t nRows, uint64_t height, uint64_t width,
                                             uint64_t fit, const double* fit_params, const double* camera, double* d_out)
{
    try {
        if (fit < 1 || fit > 4) throw pt::Error("CRLB: Gaussian fit modes 1-4 only");
        if (height == 0 || width == 0 || height > 0x7FFFFFFFull || width > 0x7FFFFFFFull) throw pt::Error("bad frame size");
        if (nRows == 0) return 1;
        pt::FitParams fp;
        fp.mode = (int)fit;
        fp.roi = fit_params ? (int)fit_params[0] : 7;
        fp.sigmaX = fit_params ? fit_params[1] : 1.3;
        fp.sigmaY = fit_params ? fit_params[2] : 1.3;
        fp.theta = fit_params ? fit_params[3] : 0.0;
        if (fp.roi < 3 || fp.roi > 31) throw pt::Error("CRLB: roi side must be in [3, 31]");
        const pt::Camera cam{camera ? camera[0] : 1.0, camera ? camera[1] : 0.0, camera ? camera[2] : 0.0};
        if (!(cam.gain > 0.0)) throw pt::Error("CRLB: camera gain must be positive");
        pt::Lease lease(false);   // no counters
        pt::Engine& e = *lease;
        const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(16384, (nRows + 63) / 64));
        switch (fit) {
        case 1: pt::crlb_kernel<1><<<grid, 64, 0, e.stream>>>((int)width, (int)height, d_rows, nRows, fp, cam, d_out); break;
        case 2: pt::crlb_kernel<2><<<grid, 64, 0, e.stream>>>((int)width, (int)height, d_rows, nRows, fp, cam, d_out); break;
        case 3: pt::crlb_kernel<3><<<grid, 64, 0, e.stream>>>((int)width, (int)height, d_rows, nRows, fp, cam, d_out); break;
        default: pt::crlb_kernel<4><<<grid, 64, 0, e.stream>>>((int)width, (int)height, d_rows, nRows, fp, cam, d_out); break;
        }
        PT_CHECK(hipGetLastError());
        PT_CHECK(hipStreamSynchronize(e.stream));
        return 1;
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "[palmtracer] pt_crlb_device: %s\n", ex.what());
    }
    return 0;
}
