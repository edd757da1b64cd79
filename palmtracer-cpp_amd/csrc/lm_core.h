// Gaussian model, gradient and damped normal-equation solve shared by the fast fit kernels (device
// code, header only).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace pt {
namespace lm {

template <int N> struct Tri { static constexpr int v = N * (N + 1) / 2; };
// Packed upper triangle of a symmetric N x N matrix: element (i, j) at tri(i, j) == tri(j, i).
template <int N> __host__ __device__ constexpr int tri(int i, int j)
{
    return i <= j ? i * N - i * (i - 1) / 2 + (j - i) : j * N - j * (j - 1) / 2 + (i - j);
}

template <int B, int E, typename F>
__device__ inline void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Solve (alpha_R) dp = beta, alpha_R = alpha with diagonal (1 + lambda) * a_ii (1 where a_ii == 0),
// by LDL^T in FP64. Returns 0 on success, 1 when the damped matrix is not numerically positive
// definite (the reference's singular-matrix case), 2 when the step is not finite.
template <int N, typename G>
__device__ inline int ldl_solve(const G (&al)[Tri<N>::v], const G (&be)[N], double lambda, double (&dp)[N])
{
    double L[Tri<N>::v], D[N], z[N];
    static_for<0, Tri<N>::v>([&](auto i) { L[i] = (double)al[i]; });
    static_for<0, N>([&](auto i) {
        constexpr int d = tri<N>(i, i);
        L[d] = L[d] == 0.0 ? 1.0 : (1.0 + lambda) * L[d];
    });
    bool pd = true;
    static_for<0, N>([&](auto j) {
        const double ajj = L[tri<N>(j, j)];
        double dj = ajj;
        static_for<0, j>([&](auto k) { dj -= L[tri<N>(j, k)] * L[tri<N>(j, k)] * D[k]; });
        pd &= (dj > 2.220446049250313e-16 * N * fabs(ajj)) && isfinite(dj);
        D[j] = dj;
        const double inv = 1.0 / dj;
        static_for<j + 1, N>([&](auto i) {
            double v = L[tri<N>(i, j)];
            static_for<0, j>([&](auto k) { v -= L[tri<N>(i, k)] * L[tri<N>(j, k)] * D[k]; });
            L[tri<N>(i, j)] = v * inv;
        });
    });
    if (!pd) return 1;
    static_for<0, N>([&](auto i) {
        double v = (double)be[i];
        static_for<0, i>([&](auto k) { v -= L[tri<N>(i, k)] * z[k]; });
        z[i] = v;
    });
    static_for<0, N>([&](auto ii) {
        constexpr int i = N - 1 - ii;
        double v = z[i] / D[i];
        static_for<i + 1, N>([&](auto k) { v -= L[tri<N>(k, i)] * dp[k]; });
        dp[i] = v;
    });
    bool ok = true;
    static_for<0, N>([&](auto i) { ok &= isfinite(dp[i]); });
    return ok ? 0 : 2;
}

template <int MODE> struct NP { static constexpr int v = MODE == 1 ? 4 : MODE == 2 ? 5 : MODE == 3 ? 6 : 7; };

// GaussianModel (lib/fit/GaussianModel.cpp:19-52) at one parameter vector. The value is FP64:
// E = exp(-(ax X^2 + ay Y^2 + axy X Y)) with X, Y the pixel offsets from the centre, which for an
// axis-aligned Gaussian (axy == 0) is the product of a column factor and a row factor. The
// gradient is of type G (FP32 in the persistent kernel, FP64 in the group kernel), formed from E
// with the reciprocal widths precomputed.
template <int MODE, typename G>
struct Model {
    double off, amp, ax, ay, axy;
    G ampf, isx2, isy2, isx, isy, c, s, cs, c2ms2, k;  // k: 1/sigma^3 (mode 2)
    __device__ void set(const double* p, double sigma, double theta)
    {
        off = p[0];
        amp = p[1];
        ampf = (G)amp;
        if constexpr (MODE <= 2) {
            const double sg = MODE == 1 ? sigma : p[MODE == 1 ? 0 : 4];
            ax = ay = 1.0 / (2.0 * sg * sg);
            axy = 0.0;
            const G sf = (G)sg;
            isx2 = G(1) / (sf * sf);
            k = isx2 / sf;
        } else {
            const double th = MODE == 3 ? theta : p[MODE == 3 ? 0 : 6];
            const double sx = p[4], sy = p[5];
            const double cd = cos(th), sd = sin(th);
            const double ix2 = 1.0 / (sx * sx), iy2 = 1.0 / (sy * sy);
            ax = cd * cd * ix2 + sd * sd * iy2;
            ay = sd * sd * ix2 + cd * cd * iy2;
            axy = 2.0 * cd * sd * (iy2 - ix2);
            isx = (G)(1.0 / sx);
            isy = (G)(1.0 / sy);
            isx2 = (G)ix2;
            isy2 = (G)iy2;
            c = (G)cd;
            s = (G)sd;
            cs = c * s;
            c2ms2 = c * c - s * s;
        }
    }
    __device__ bool separable() const { return axy == 0.0; }
    __device__ double E(double X, double Y) const { return exp(-(ax * X * X + ay * Y * Y + axy * X * Y)); }
    // Gradient at (x, y) given E (same components as GaussianModel::Evaluate).
    __device__ void grad(G x, G y, G Ef, G* g) const
    {
        const G P = ampf * Ef;
        g[0] = G(1);
        g[1] = Ef;
        if constexpr (MODE <= 2) {
            g[2] = P * x * isx2;
            g[3] = P * y * isx2;
            if constexpr (MODE == 2) g[4] = P * (x * x + y * y) * k;
        } else {
            const G F1 = (x * c - y * s) * isx, F2 = (x * s + y * c) * isy;
            g[2] = P * (G(2) * (c * c * x - y * cs) * isx2 + G(2) * (s * s * x + y * cs) * isy2);
            g[3] = P * (G(2) * (s * s * y - x * cs) * isx2 + G(2) * (c * c * y + x * cs) * isy2);
            g[4] = P * G(2) * F1 * F1 * isx;
            g[5] = P * G(2) * F2 * F2 * isy;
            if constexpr (MODE == 4) g[6] = P * G(2) * (isx2 - isy2) * (cs * (x * x - y * y) + c2ms2 * x * y);
        }
    }
};

} // namespace lm
} // namespace pt
