/* CPU oracle for the PALMTracer detection + fitting path.
 *
 * TEST INFRASTRUCTURE ONLY: an FP64 restatement in plain C of the reference algorithm
 * (tmonseigne/PALMTracer-cpp, lib/), used by tests/, the smoke check and the bench's CPU
 * baseline leg, to compare results against. The product library never links, loads or calls it.
 *
 * Every function names the reference file:line it restates. Pinning (how each part is
 * checked independently of this restatement) is listed in oracle.c's header.
 */
#ifndef PALMTRACER_ORACLE_H
#define PALMTRACER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One segmented region, as SegmentationInfo (lib/localization/SegmentationInfo.hpp:32-71). */
typedef struct {
    double x, y, z, intensity, surf;
} or_region;

/* A trous planes W0, W1, V2 of one frame (lib/localization/ATrous.cpp:23-80). Any output may be NULL. */
void or_atrous(const uint16_t* img, size_t width, size_t height, double* w0, double* w1, double* v2);

/* Segmentation2D::Process on a given W1 plane (lib/localization/Segmentation2D.cpp:50-74).
 * labels: width*height u32 (final map after FilterData). regions: capacity cap, label order.
 * Returns the region count (may exceed cap; only cap are written). */
size_t or_segment(const double* w1, size_t width, size_t height, double threshold, double ratio,
                  uint32_t* labels, or_region* regions, size_t cap);

/* Watershed-only entry: runs getLocalMax + watershed (Segmentation2D.cpp:294-385) on the
 * thresholded potential and writes the label map before any fusion / filtering. Returns the
 * number of seed labels. */
size_t or_watershed_labels(const double* w1, size_t width, size_t height, double threshold, uint32_t* labels);

/* GaussianFit::Fit / Fitting::Fit for one region (lib/fit/GaussianFit.cpp:48-215,
 * lib/fit/Fitting.hpp:63-66). mode 0..4, fitParams = [roi, sigmaX, sigmaY, theta].
 * maxIter is Fitting::N_ITER (75 in the reference). Writes 18 doubles. */
void or_fit_region(int mode, const uint16_t* img, size_t width, size_t height, const double* fitParams,
                   int maxIter, const or_region* info, double* row);

/* Gauss-Jordan inverse with row equilibration (lib/signal_processing/Matrix.cpp:297-355).
 * in is overwritten. Returns 0 on success, -1 on the reference's domain_error. */
int or_inverse_gauss_jordan(double* in, double* out, size_t n);

/* Export-level entry points (CPU/main.cpp). Same argument meaning as the reference. */
uint64_t or_localization(const uint16_t* input, double* output, uint64_t outputSize, uint64_t height, uint64_t width,
                         uint64_t planes, double threshold, double ratio, uint64_t fit, const double* fitParams,
                         int maxIter);
uint64_t or_fitting(const uint16_t* input, const double* positions, uint64_t nPoints, double* output,
                    uint64_t outputSize, uint64_t height, uint64_t width, uint64_t planes, uint64_t fit,
                    const double* fitParams, int maxIter);
void or_wavelett(const uint16_t* input, double* output, uint64_t height, uint64_t width, uint64_t planes,
                 uint64_t levelOutput);
double or_autothreshold(const uint16_t* input, uint64_t height, uint64_t width, const double* fitParams);

/* AstigFit::Estimate (lib/fit/AstigFit.cpp:36-72) over size (sigma_x, sigma_y) pairs in pixels:
 * output (z in nm, axial score) pairs; model = 2 x [z0, w, c3, c4, A]; search on [-zMax, zMax]. */
void or_astig_estimate(const double* input, double* output, uint64_t size, double pixelSize, const double* model,
                       double zMax);

/* AstigFit::getQuality (lib/fit/AstigFit.cpp:99-108) of one observed pair at z. */
double or_astig_quality(const double* model, double z, double sx, double sy, double pixelSize);

/* The C library's exp (fn 0) or the sine (1) or cosine (2) of its sincos over n doubles: std::exp
 * and the std::cos / std::sin pair of the reference's Gaussian model (lib/fit/GaussianModel.cpp:
 * 19-52), which gcc turns into one sincos call, in the reference's build and in this oracle's. */
void or_libm(int fn, const double* x, uint64_t n, double* y);

#ifdef __cplusplus
}
#endif
#endif
