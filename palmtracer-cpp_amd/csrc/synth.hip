// Seeded synthetic single-molecule stacks, generated on the device (benchmark and test input).
//
// Camera model: ADU = offset + gain * Poisson(bg + sum_k N_k * PSF_k(pixel)), clipped to u16, where
// PSF_k is a 2D Gaussian integrated over the pixel (erf differences; pixel centres at integer
// coordinates, the convention of the reference's model, lib/fit/GaussianModel.cpp) with sigma_x,
// sigma_y per emitter, N_k ~ LogNormal(ln(median), sigma_log), positions uniform over the frame.
// Astigmatic stacks draw z ~ U[zmin, zmax] and set sigma_{x,y} = s0 * sqrt(1 + ((z -/+ c) / d)^2).
// sCMOS stacks use per-pixel offset ~ N(offset, sdOffset) and gain ~ N(gain, sdGain), fixed per seed.
//
// Randomness is counter based (Philox4x32-10, Salmon et al., SC'11), so any frame can be produced
// independently of the others: emitter k of frame f uses counters (k, f, stream, 0..1), pixel
// (x, y) of frame f uses (x, y, f, stream + chunk). Poisson variates are drawn by inversion in
// chunks of mean <= 64 (sum of independent Poissons).
#include "pt_common.h"
#include "../../include/palmtracer.h"

namespace pt {

struct SynthConfig {
    uint64_t seed;
    int W, H, frames, emitters;       // emitters per frame
    double bg, sigma, median, sigmaLog; // background photons, PSF sigma (px), photon log-normal
    double offset, gain;                // camera offset (ADU) and gain (ADU / photon)
    int astig;                          // 1: z-dependent sigma_x / sigma_y
    double s0, c, d, zmin, zmax;        // astigmatism (nm units for c, d, z)
    int scmos;                          // 1: per-pixel offset / gain maps
    double sdOffset, sdGain;
};

namespace {

__device__ __host__ inline void philox(uint32_t c[4], const uint32_t k0in, const uint32_t k1in)
{
    uint32_t k0 = k0in, k1 = k1in;
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ inline double u53(uint32_t a, uint32_t b)
{
    return (double)(((uint64_t)a << 21) ^ (b >> 11)) * (1.0 / 9007199254740992.0);
}

struct Emitter {
    float x, y, sx, sy, n;
};

__device__ inline Emitter make_emitter(const SynthConfig& cfg, int f, int k)
{
    const uint32_t key0 = (uint32_t)cfg.seed, key1 = (uint32_t)(cfg.seed >> 32);
    uint32_t c0[4] = {(uint32_t)k, (uint32_t)f, 0xE5u, 0u};
    philox(c0, key0, key1);
    uint32_t c1[4] = {(uint32_t)k, (uint32_t)f, 0xE5u, 1u};
    philox(c1, key0, key1);
    Emitter e;
    e.x = (float)(u53(c0[0], c0[1]) * cfg.W);
    e.y = (float)(u53(c0[2], c0[3]) * cfg.H);
    const double u1 = 1.0 - u53(c1[0], c1[1]), u2 = u53(c1[2], c1[3]);
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);  // Box-Muller
    e.n = (float)exp(log(cfg.median) + cfg.sigmaLog * z);
    if (cfg.astig) {
        uint32_t c2[4] = {(uint32_t)k, (uint32_t)f, 0xE5u, 2u};
        philox(c2, key0, key1);
        const double zz = cfg.zmin + (cfg.zmax - cfg.zmin) * u53(c2[0], c2[1]);
        const double a = (zz - cfg.c) / cfg.d, b = (zz + cfg.c) / cfg.d;
        e.sx = (float)(cfg.s0 * sqrt(1.0 + a * a));
        e.sy = (float)(cfg.s0 * sqrt(1.0 + b * b));
    } else {
        e.sx = e.sy = (float)cfg.sigma;
    }
    return e;
}

__device__ inline uint32_t poisson(double lam, uint32_t x, uint32_t y, uint32_t f, const SynthConfig& cfg)
{
    const uint32_t key0 = (uint32_t)cfg.seed, key1 = (uint32_t)(cfg.seed >> 32);
    uint32_t total = 0;
    for (uint32_t chunk = 0; lam > 0.0 && chunk < 4096; ++chunk) {
        const double l = lam > 64.0 ? 64.0 : lam;
        lam -= l;
        uint32_t c[4] = {x, y, f, 0x90150000u + chunk};
        philox(c, key0, key1);
        const double u = u53(c[0], c[1]);
        double p = exp(-l), F = p;
        uint32_t k = 0;
        while (u > F && k < 1024) {
            ++k;
            p *= l / (double)k;
            F += p;
        }
        total += k;
    }
    return total;
}

constexpr int TS = 32;          // tile side
constexpr int MAXE = 1024;      // emitters kept per tile (dense stacks beyond this are clipped)
constexpr int NT = 256;

__global__ void __launch_bounds__(NT) synth_kernel(SynthConfig cfg, uint16_t* __restrict__ out)
{
    __shared__ Emitter se[MAXE];
    __shared__ uint32_t sWave[NT / 64];
    __shared__ uint32_t sN;
    const int f = blockIdx.z;
    const int tx0 = blockIdx.x * TS, ty0 = blockIdx.y * TS;
    const float R = 7.0f;
    if (threadIdx.x == 0) sN = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // Ordered compaction of the emitters that reach this tile (deterministic summation order).
    for (int b = 0; b < cfg.emitters; b += NT) {
        const int k = b + threadIdx.x;
        Emitter e{};
        bool sel = false;
        if (k < cfg.emitters) {
            e = make_emitter(cfg, f, k);
            const float r = R * fmaxf(1.0f, fmaxf(e.sx, e.sy) / 1.3f);
            sel = e.x + r >= tx0 && e.x - r < tx0 + TS && e.y + r >= ty0 && e.y - r < ty0 + TS;
        }
        const uint64_t m = __ballot(sel);
        if (lane == 0) sWave[wv] = __popcll(m);
        __syncthreads();
        uint32_t pos = sN + __popcll(m & ((1ull << lane) - 1ull));
        for (int j = 0; j < wv; ++j) pos += sWave[j];
        if (sel && pos < MAXE) se[pos] = e;
        __syncthreads();
        if (threadIdx.x == 0) for (int j = 0; j < NT / 64; ++j) sN += sWave[j];
        __syncthreads();
    }
    const uint32_t ne = min(sN, (uint32_t)MAXE);
    const uint32_t key0 = (uint32_t)cfg.seed, key1 = (uint32_t)(cfg.seed >> 32);
    for (int i = threadIdx.x; i < TS * TS; i += NT) {
        const int x = tx0 + i % TS, y = ty0 + i / TS;
        if (x >= cfg.W || y >= cfg.H) continue;
        double lam = cfg.bg;
        for (uint32_t j = 0; j < ne; ++j) {
            const Emitter e = se[j];
            const double dx = x - e.x, dy = y - e.y;
            if (fabs(dx) > 7.0 * fmax(1.0, e.sx / 1.3) || fabs(dy) > 7.0 * fmax(1.0, e.sy / 1.3)) continue;
            const double ax = 1.0 / (1.4142135623730951 * e.sx), ay = 1.0 / (1.4142135623730951 * e.sy);
            const double gx = 0.5 * (erf((dx + 0.5) * ax) - erf((dx - 0.5) * ax));
            const double gy = 0.5 * (erf((dy + 0.5) * ay) - erf((dy - 0.5) * ay));
            lam += (double)e.n * gx * gy;
        }
        double off = cfg.offset, gain = cfg.gain;
        if (cfg.scmos) {
            uint32_t c[4] = {(uint32_t)x, (uint32_t)y, 0x5C305u, 0u};  // frame independent
            philox(c, key0, key1);
            const double u1 = 1.0 - u53(c[0], c[1]), u2 = u53(c[2], c[3]);
            const double rr = sqrt(-2.0 * log(u1));
            off += cfg.sdOffset * rr * cos(6.283185307179586 * u2);
            gain += cfg.sdGain * rr * sin(6.283185307179586 * u2);
        }
        const double adu = rint(off + gain * (double)poisson(lam, (uint32_t)x, (uint32_t)y, (uint32_t)f, cfg));
        out[(size_t)f * cfg.W * cfg.H + (size_t)y * cfg.W + x] = (uint16_t)fmin(fmax(adu, 0.0), 65535.0);
    }
}

// Emitter table (x, y, sigma_x, sigma_y, photons) of frames [0, frames), for ground-truth checks.
__global__ void emitters_kernel(SynthConfig cfg, float* __restrict__ out)
{
    const int f = blockIdx.y;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < cfg.emitters; k += gridDim.x * blockDim.x) {
        const Emitter e = make_emitter(cfg, f, k);
        float* o = out + ((size_t)f * cfg.emitters + k) * 5;
        o[0] = e.x; o[1] = e.y; o[2] = e.sx; o[3] = e.sy; o[4] = e.n;
    }
}
} // namespace

void launch_synth(const SynthConfig& cfg, uint16_t* d_out, hipStream_t s)
{
    if (cfg.frames <= 0 || cfg.W <= 0 || cfg.H <= 0) return;
    const dim3 grid(ceil_div(cfg.W, TS), ceil_div(cfg.H, TS), cfg.frames);
    synth_kernel<<<grid, NT, 0, s>>>(cfg, d_out);
    PT_CHECK(hipGetLastError());
}

void launch_synth_emitters(const SynthConfig& cfg, float* d_out, hipStream_t s)
{
    emitters_kernel<<<dim3(ceil_div(cfg.emitters, 256), cfg.frames), 256, 0, s>>>(cfg, d_out);
    PT_CHECK(hipGetLastError());
}

} // namespace pt

extern "C" PT_EXPORT int pt_synth_device(uint16_t* d_out, uint64_t seed, int width, int height,
                                                                      int frames, int emitters, const double* params)
{
    // params: bg, sigma, median, sigmaLog, offset, gain, astig, s0, c, d, zmin, zmax, scmos, sdOffset, sdGain
    try {
        if (frames > 65535) throw pt::Error("at most 65535 frames per call");
        pt::SynthConfig c{};
        c.seed = seed; c.W = width; c.H = height; c.frames = frames; c.emitters = emitters;
        c.bg = params[0]; c.sigma = params[1]; c.median = params[2]; c.sigmaLog = params[3];
        c.offset = params[4]; c.gain = params[5]; c.astig = (int)params[6]; c.s0 = params[7]; c.c = params[8];
        c.d = params[9]; c.zmin = params[10]; c.zmax = params[11]; c.scmos = (int)params[12];
        c.sdOffset = params[13]; c.sdGain = params[14];
        pt::launch_synth(c, d_out, nullptr);
        PT_CHECK(hipDeviceSynchronize());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_synth_device: %s\n", e.what());
        return 0;
    }
}

extern "C" PT_EXPORT int pt_synth_emitters_device(float* d_out, uint64_t seed, int width, int height,
                                                                               int frames, int emitters, const double* params)
{
    try {
        pt::SynthConfig c{};
        c.seed = seed; c.W = width; c.H = height; c.frames = frames; c.emitters = emitters;
        c.bg = params[0]; c.sigma = params[1]; c.median = params[2]; c.sigmaLog = params[3];
        c.offset = params[4]; c.gain = params[5]; c.astig = (int)params[6]; c.s0 = params[7]; c.c = params[8];
        c.d = params[9]; c.zmin = params[10]; c.zmax = params[11]; c.scmos = (int)params[12];
        c.sdOffset = params[13]; c.sdGain = params[14];
        pt::launch_synth_emitters(c, d_out, nullptr);
        PT_CHECK(hipDeviceSynchronize());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[palmtracer] pt_synth_emitters_device: %s\n", e.what());
        return 0;
    }
}
