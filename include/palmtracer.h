/* C ABI of libpalmtracer_amd.so: the PALMTracer detection + fitting path on one MI355X (gfx950).
 *
 * The first block is a drop-in for the reference's exported hot-path entry points
 * (tmonseigne/PALMTracer-cpp, CPU/main.cpp and CPU/PALMTracer.def): same names, argument types,
 * argument meaning, return values and failure behaviour (0 / no rows on failure, never an
 * exception across the boundary). All buffers are caller-owned host memory, borrowed for the call.
 *
 * The second block has no counterpart in the reference: entry points over device pointers (for
 * callers that keep their stacks in HBM), options and diagnostics.
 */
#ifndef PALMTRACER_AMD_H
#define PALMTRACER_AMD_H

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#if defined(_WIN32) || defined(_WIN64)
#define PT_EXPORT __declspec(dllexport)
#else
#define PT_EXPORT __attribute__((visibility("default")))
#endif
#define PT_CALL
#ifdef __cplusplus
#define PT_API extern "C" PT_EXPORT
#else
#define PT_API PT_EXPORT
#endif

/* ---- reference entry points --------------------------------------------------------------- */

/* Replaces Localization (CPU/main.cpp:112-121 -> lib/Palm.cpp:48-70).
 * input: planes x height x width u16, x fastest. output: outputSize doubles; rows of 18 columns
 * [Id, Plane, Index, Channel, X, Y, Z, Intensity, SigmaX, SigmaY, Theta, QualityXY, QualityZ,
 *  Amplitude, Offset, SegIntensity, Surface, Circularity] (lib/Constants.hpp:31-51), plane-major,
 * each plane sorted by (Y, X). At most (outputSize / 18) / planes rows per plane, kept in label
 * order (lib/fit/Fitting.cpp:44-58). fit: 0 none, 1 sigma fixed, 2 sigma free, 3 sigmaX/sigmaY,
 * 4 sigmaX/sigmaY/theta, 5 spline PSF with the EMCCD Poisson likelihood (lib/fit/SplineFit.cpp:36-216);
 * fit_params = [roi, sigmaX, sigmaY, theta] for 1-4, [roi, spline sizes x, y, z, deltaZ (nm), then
 * 64 * x * y * z coefficients, monomial-major] for 5. watershedRatio <= 0 keeps the
 * maxima separate (reference value 0); (int)ratio >= 1 enables basin fusion.
 * The stack goes through the device in pieces of at most "stream_pixels" pixels (pinned staging,
 * uploads and downloads overlapping the computation), so it may be larger than HBM.
 * Returns the number of doubles written. */
PT_API uint64_t PT_CALL Localization(const uint16_t* input, double* output, uint64_t outputSize, uint64_t height,
                                     uint64_t width, uint64_t planes, double thresholdVal, double watershedRatio,
                                     uint64_t fit, const double* fit_params);

/* Replaces Fitting (CPU/main.cpp:141-148 -> lib/Palm.cpp:73-119). positions: nPoints rows
 * [Id, Plane(1-based), X, Y, SegIntensity, Surface] (lib/Constants.hpp:56-64). Rows are written in
 * input order; outputSize is ignored as in the reference. fit in 1..5, fit_params as for
 * Localization. The stack is streamed through the device in pieces as in Localization (each piece
 * with the positions of its planes). Returns 18 * nPoints. */
PT_API uint64_t PT_CALL Fitting(const uint16_t* input, const double* positions, uint64_t nPoints, double* output,
                                uint64_t outputSize, uint64_t height, uint64_t width, uint64_t planes, uint64_t fit,
                                const double* fit_params);

/* Replaces Wavelett (CPU/main.cpp:257-284): level 1 = W0, 2 = W1, 3 (or more) = residual V2,
 * 0 is read as 1. output: planes x height x width doubles. */
PT_API void PT_CALL Wavelett(const uint16_t* input, double* output, uint64_t height, uint64_t width, uint64_t planes,
                             uint64_t levelOutput);

/* Replaces AutoThreshold (CPU/main.cpp:75-85 -> lib/Palm.cpp:122-172): threshold for one plane,
 * from the standard deviation of the pixels outside the ROI squares of the current detections
 * (watershed ratio 10, no fit), iterated at most 5 times. fit_params[0] is the ROI side.
 * Returns NaN on failure. */
PT_API double PT_CALL AutoThreshold(const uint16_t* input, uint64_t height, uint64_t width, const double* fit_params);

/* Astigmatism3DEstimation (CPU/main.cpp:324-353, lib/fit/AstigFit.cpp:36-72): for each of the size
 * (sigmaX, sigmaY) pairs in pixels, the z (nm) of the model's closest sigmas on [-zMax, zMax]
 * (201-point grid, then Brent's method) and the axial score Q_Z; output holds 2 * size doubles.
 * model = 2 axes x [z0, w, c3, c4, A] as Astigmatism3DCalibration writes it. Runs on the GPU. */
PT_API void PT_CALL Astigmatism3DEstimation(const double* input, double* output, uint64_t size, double pixelSize,
                                            const double* model, double zMax);

/* ---- reference host-side entry points (forwarded) ----------------------------------------- */
/* Exported with the reference's signatures so that this library replaces the reference library
 * as a whole. They are not part of the detection + fitting path and run on the host: each call is
 * forwarded to the reference CPU library named by the environment variable PALMTRACER_HOST_LIBRARY.
 * Without it the call prints an error on stderr and returns 0 / false, outputs untouched. */

/* Alignment (CPU/main.cpp:47-48): drift correction of a stack; factors = 2 x 10 coefficients. */
PT_API void PT_CALL Alignment(const uint16_t* input, uint16_t* output, uint64_t height, uint64_t width, uint64_t planes,
                              const double* factors, uint64_t upsampling);
/* Tracking (CPU/main.cpp:159): writes 2 * nPoints values. */
PT_API void PT_CALL Tracking(const double* input, uint64_t* output, uint64_t nPoints, double maxDistance);
/* BlinkingReconnection (CPU/main.cpp:185-186): returns the number of doubles written. */
PT_API uint64_t PT_CALL BlinkingReconnection(const double* input, double* output, uint64_t nRow, double pixelSize,
                                             uint64_t mode, uint64_t maxDuration, double maxSpeed);
/* TrackAnalysis (CPU/main.cpp:231-233). */
PT_API bool PT_CALL TrackAnalysis(const double* input, double* outputMSD, double* outputInd, double* outputFit, uint64_t nRow,
                                  bool isMSD, bool isInd, bool is3D, double pixelSize, double exposureTime, uint64_t fitMode,
                                  const double* fitParams);
/* Astigmatism3DCalibration (CPU/main.cpp:293): 10 model parameters from (sigmaX, sigmaY, z) points. */
PT_API void PT_CALL Astigmatism3DCalibration(const double* input, double* output, uint64_t size, double pixelSize, bool center);

/* ---- no counterpart in the reference ------------------------------------------------------ */

/* Astigmatism3DEstimation over device memory (d_input: 2 * size doubles, d_output: 2 * size doubles);
 * model (10 doubles) in host memory. Returns 1 on success, 0 on failure. */
PT_API int PT_CALL pt_astig_estimate_device(const double* d_input, double* d_output, uint64_t size, double pixelSize,
                                            const double* model, double zMax);

/* Serving loop for several streams of frame chunks (cameras). `slots` independent engines (own
 * streams and chunk buffers; chunkPixels frames x pixels per device chunk, 0: the default) each run
 * one request at a time, so requests on different slots overlap on the GPU. Requests leave the queue
 * in arrival order when a slot is free and their HBM (input frames, output capacity and the engine
 * work space they need, see pt_server_request_bytes) fits in budgetBytes beside the requests in flight
 * and the work space the slots hold (idle slots release theirs when a request needs the room). A request is Localization over device memory
 * (as pt_localization_device; nFitParams doubles of fit_params are copied at submission, so the
 * caller may reuse them); its result equals that call's whatever else is in flight. Options
 * (pt_set_option) in force at creation apply to every slot. Returns NULL on failure. */
PT_API void* PT_CALL pt_server_create(uint64_t slots, uint64_t budgetBytes, uint64_t chunkPixels);
/* Queues a request; d_input and d_output must stay valid until pt_server_wait returns for it.
 * Returns the request id (>= 0) or -1 (bad arguments, request larger than the budget). */
PT_API int64_t PT_CALL pt_server_submit(void* server, const uint16_t* d_input, uint64_t planes, uint64_t height,
                                        uint64_t width, double thresholdVal, double watershedRatio, uint64_t fit,
                                        const double* fit_params, uint64_t nFitParams, double* d_output,
                                        uint64_t outputCapDoubles, uint64_t maxPerPlane);
/* Blocks until the request is done. doublesWritten: doubles of rows in d_output; times (4 doubles):
 * submission, start and completion in ms on the server's clock, and the slot that ran it. Returns 1
 * on success, 0 if the request failed (message on stderr). Each id is waited for once. */
PT_API int PT_CALL pt_server_wait(void* server, int64_t request, uint64_t* doublesWritten, double* times);
/* HBM the server's admission counts for a request of these sizes: its input frames and output rows
 * plus the engine work space it needs (chunk buffers, ROI stores, rows; the sizes the engine
 * allocates). A submission above budgetBytes is refused. Returns 0 on bad arguments. */
PT_API uint64_t PT_CALL pt_server_request_bytes(void* server, uint64_t planes, uint64_t height, uint64_t width,
                                                uint64_t outputCapDoubles, uint64_t maxPerPlane);
/* The server's clock (ms since creation). */
PT_API double PT_CALL pt_server_now(void* server);
/* Finishes the queued requests, then releases the slots. */
PT_API void PT_CALL pt_server_destroy(void* server);

/* AutoThreshold of one frame in device memory. */
PT_API double PT_CALL pt_autothreshold_device(const uint16_t* d_input, uint64_t height, uint64_t width,
                                              const double* fit_params);

/* Localization over device memory: d_input (planes x height x width u16) and d_output
 * (outputCapDoubles doubles) are HBM pointers; maxPerPlane is the per-plane row capacity. */
PT_API uint64_t PT_CALL pt_localization_device(const uint16_t* d_input, double* d_output, uint64_t outputCapDoubles,
                                               uint64_t height, uint64_t width, uint64_t planes, double thresholdVal,
                                               double watershedRatio, uint64_t fit, const double* fit_params,
                                               uint64_t maxPerPlane);

/* Wavelett over device memory. Returns 1 on success, 0 on failure. */
PT_API int PT_CALL pt_wavelet_device(const uint16_t* d_input, double* d_output, uint64_t height, uint64_t width,
                                     uint64_t planes, uint64_t levelOutput);

/* The fit model's exp (fn 0) or the sine (1) or cosine (2) of its sincos over n device doubles,
 * d_y[i] = f(d_x[i]), bit-equal to the x86-64 glibc exp and sincos of the reference's Linux build;
 * runs on the legacy default stream and returns after it completes. Returns 1 on success. */
PT_API int PT_CALL pt_debug_libm(uint64_t fn, const double* d_x, double* d_y, uint64_t n);

/* Watershed label map of one host frame: labels[i] = raster index + 1 of the first local maximum
 * of the pixel's label (its region's identity), 0 below threshold. With (int)watershedRatio >= 1
 * the map is the one after basin fusion and the second watershed. Returns 1 on success. */
PT_API int PT_CALL pt_debug_labels(const uint16_t* image, uint64_t height, uint64_t width, double thresholdVal,
                                   double watershedRatio, uint32_t* labels);

/* Options: "fit_fp64" (1, the default: FP64 fit in the reference's operation order, one ROI per lane;
 * 0: the fast fit, an FP32 persistent kernel whose ROIs still iterating after "fit_handover"
 * iterations, or meeting a decision FP32 cannot take, are fitted again in FP64 by a cooperative
 * kernel of eight lanes per ROI; fit mode 4 runs in that FP64 kernel whole), "fit_handover" (FP32
 * iterations of the fast fit, 16; 0: no hand-over, every ROI and mode in FP32, errors as stated for
 * that precision), "max_iter" (LM iteration cap, 75),
 * "chunk_pixels" (frames per device chunk x pixels, 0: default 2^28), "profile" (1: stage timing),
 * "force_sequential_watershed" (1: every frame through the exact frame flood), "force_rank_flood"
 * (1: that flood from the single-workgroup rank flood, without the grid rank flood), "force_wave_flood"
 * (1: that flood without the rank floods), "force_heap_watershed" (1: that flood with the heap
 * restatement of the reference), "fit_waves" (resident waves of the
 * persistent FP32 fit, 0: default), "fit_refill" (free lanes of an FP32 fit wave that start a refill
 * round, 1-64, default 16), "detect_rows" (rows per detection wave, 0: chosen from the chunk),
 * "lanes" (2: consecutive chunks alternate between two streams; 1: one stream), "batch_autotune" (1: tune
 * new fit classes of pt_batch_localization), "batch_graph" (1: its fit stage as a hipGraph), "stream_pixels"
 * (host Localization / Fitting: pixels per piece of the stack, default 2^27), "host_threads" (host
 * threads of the copies to and from pinned staging, default 8), "defer_flagged"
 * (flagged frames of all but the last chunk re-detected and resolved on the side stream: 0 never, 1
 * always, 2 (default) when the last chunk whose counters are known flagged at most 2 % of its frames),
 * "fit_objective" (0: the reference's least squares; 1: Poisson maximum likelihood of fit modes 1-4,
 * FP64, an extension with no counterpart in the reference), "camera_gain" (ADU per photon, 1),
 * "camera_offset" (ADU, 0) and "camera_read_var" (ADU^2, 0): the camera model of that likelihood.
 * Options apply to every call that starts after this one returns, whichever engine serves it
 * (concurrent calls run on separate engines). Returns 1 on success. */
PT_API int PT_CALL pt_set_option(const char* name, double value);

/* Counters of the Localization, Fitting, AutoThreshold, label-map or batch call that finished last
 * (of concurrent calls, the last to return; Fitting fills frames and rows only): frames, chunks,
 * rows, frames flagged for the exact frame flood, multi-label components, frames deferred to the final pass, status, component floods given up on a
 * full queue, component floods given up on an unproven order, pops of the wave-queue frame floods,
 * their largest queue, frames the rank flood handed over (too many entries, equal keys, not
 * settled), wave-queue overflows, frames of the heap restatement, frames the rank flood settled by
 * solving both orders of an equal-key pair, ROIs of the fast fit fitted by the FP64 group kernel
 * (main lanes), frames the grid rank flood resolved, of which frames whose flood has equal keys that
 * were shown not to change a label, and the frames it handed to the single-workgroup floods: not
 * settled, equal keys that might change a label, other (work space, push order), and the mask pixels
 * (thresholded potential above zero) of the main lanes' detections. Writes min(n, 24) values,
 * returns 24. */
PT_API int PT_CALL pt_last_stats(uint64_t* out, uint64_t n);

/* With the "profile" option, device milliseconds of the last Localization per stage (a trous,
 * segmentation, fit, assembly; HIP events on the library's main stream, summed over chunks), the
 * chunk count, the LM passes of the reference's schedule (model + Jacobian, trial) of the per-lane
 * fit kernel, the device milliseconds of the deferred-frame batches on the side stream (overlapped
 * with the main stream), the LM passes (model + Jacobian, trial) of the FP64 group kernel, and for
 * the host entry points the device milliseconds of the frame uploads and row downloads and the
 * number of pieces the stack went through the device in (whatever the "profile" option).
 * Writes min(n, 13) values, returns 13. */
PT_API int PT_CALL pt_stage_times(double* out, uint64_t n);

/* Cramer-Rao lower bounds of Gaussian fits (fit 1-4; no counterpart in the reference's Gaussian path,
 * whose spline fit states them: lib/fit/SplineFit.cpp:168-205). d_rows: nRows rows of 18 doubles on
 * the device (Localization / Fitting output of frames height x width with this fit and fit_params).
 * camera (host, 3 doubles; NULL: 1, 0, 0): gain (ADU per photon), offset (ADU) and read-noise
 * variance (ADU^2) of the noise model var = gain * (model - offset) + readVar. d_out: nRows x 8
 * doubles on the device, the bounds of X, Y, SigmaX, SigmaY, Theta, Amplitude, Offset, Intensity in
 * the units of those columns (FP64 Fisher information of the fitted model over the ROI centred on
 * the fitted pixel, inverted by Cholesky); NaN where the mode does not fit the parameter or the row's
 * fit failed. Returns 1 on success, 0 on error. */
PT_API int PT_CALL pt_crlb_device(const double* d_rows, uint64_t nRows, uint64_t height, uint64_t width, uint64_t fit,
                                  const double* fit_params, const double* camera, double* d_out);

/* Batched localization of many stacks (no counterpart in the reference, whose callers run Localization
 * once per stack). Each stack: device input (planes x height x width u16), threshold, watershed ratio,
 * fit mode 0-4 with its fit_params (host, [roi, sigmaX, sigmaY, theta]), device output of
 * outputCapDoubles doubles with at most maxPerPlane rows per plane; doublesWritten is set by the call.
 * Every stack is detected and segmented on its own; the fit jobs of all stacks are fitted by one
 * persistent launch per fit class (equal fit parameters), the classes side by side in one captured
 * hipGraph, with schedules (resident waves, refill threshold, FP32 iterations before the FP64
 * hand-over) picked by an autotuner the first time a class is seen ("batch_autotune"). A stack's rows
 * equal pt_localization_device on that stack with the fast fit ("fit_fp64" 0) and the same
 * hand-over iterations (pt_batch_schedules). Returns 1 on success, 0 on failure. */
typedef struct {
    const uint16_t* d_input;
    uint64_t planes, height, width;
    double thresholdVal, watershedRatio;
    uint64_t fit;
    const double* fit_params;
    double* d_output;
    uint64_t outputCapDoubles, maxPerPlane;
    uint64_t doublesWritten;
} pt_batch_stack;
PT_API int PT_CALL pt_batch_localization(pt_batch_stack* stacks, uint64_t nStacks);
/* The fit classes of the last pt_batch_localization: per class 8 doubles [mode, roi, jobs, waves,
 * refill, hand-over iterations, autotuner ms of the chosen schedule, 1 if tuned]. Writes min(n, 8 *
 * classes) values, returns the number of classes. */
PT_API int PT_CALL pt_batch_schedules(double* out, uint64_t n);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Fitting (as above, same
 * arguments and rows, outputSize dropped) over device memory: d_input planes x height x width u16,
 * d_positions nPoints x 6 doubles, d_output 18 * nPoints doubles. Its work space is a fixed device
 * array of the library (about 6.6 MB per GPU, held until the process ends), so the call allocates
 * nothing; at most 131072 positions and 65535 planes per call, Gaussian fit modes 1-4 only (mode 5
 * uploads its spline model); larger inputs are refused with a message on stderr. Runs on the legacy
 * default stream and returns when the rows are written. Calls that overlap on one GPU share that
 * array and race. Returns 18 * nPoints, or 0 on failure. */
PT_API uint64_t PT_CALL pt_fitting_noalloc(const uint16_t* d_input, const double* d_positions, uint64_t nPoints,
                                           double* d_output, uint64_t height, uint64_t width, uint64_t planes,
                                           uint64_t fit, const double* fit_params);

/* Seeded synthetic stack on the device (benchmark / test input): d_out receives frames x height x
 * width u16. params (15 doubles): bg photons, PSF sigma (px), photon median, photon sigma_log, camera
 * offset (ADU), gain (ADU/photon), astigmatic flag, s0 (px), c (nm), d (nm), zmin, zmax (nm),
 * sCMOS flag, offset sd, gain sd. Returns 1 on success. */
PT_API int PT_CALL pt_synth_device(uint16_t* d_out, uint64_t seed, int width, int height, int frames, int emitters,
                                   const double* params);

/* Ground truth of pt_synth_device: frames x emitters x [x, y, sigma_x, sigma_y, photons] floats. */
PT_API int PT_CALL pt_synth_emitters_device(float* d_out, uint64_t seed, int width, int height, int frames,
                                            int emitters, const double* params);

#endif
