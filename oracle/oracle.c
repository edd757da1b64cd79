/* CPU oracle: FP64 restatement of the PALMTracer detection + fitting path.
 *
 * TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg, as the reference against which results are compared. The product library
 * never links or calls it.
 *
 * Reference: tmonseigne/PALMTracer-cpp (paths relative to its root).
 *
 * How each part is pinned independently of this file (tests/test_oracle.py):
 *   or_atrous            reference's own invariants (sum of planes == image, W0 == image - direct
 *                        separable [1,4,6,4,1]/16 convolution with clamp, constant image -> zero
 *                        details; Tests_CPU/test_loc_ATrous.cpp:27-107) + bit equality with the
 *                        compiled reference (oracle/_ref) on stack.tif and synthetic frames.
 *   or_segment           golden CSVs (Tests_CPU/Input/ref/stack-localizations-*_0_*.csv: fit mode 0
 *                        rows are the region list) for watershed on and off, the hand-built
 *                        potentials of Tests_CPU/test_loc_Segmentation2D.cpp:120-287, and
 *                        label-map equality with oracle/_ref (committed as fixtures).
 *   or_fit_region        golden CSVs modes 1-4 (atol 1e-3 + rtol 1e-5, Tests_CPU/test_Palm.cpp:55-76),
 *                        synthetic recovery (Tests_CPU/test_fit_Fitting.cpp:236-256), R^2 equal to an
 *                        independently computed squared Pearson r, sentinel cases.
 *   or_inverse_gauss_jordan  A * inv(A) == I on random well-conditioned matrices, numpy.linalg.inv,
 *                        the failure cases of Tests_CPU/test_sp_Matrix.cpp:230-298.
 *   or_autothreshold     the 10 known answers of Tests_CPU/test_Palm.cpp:307-310.
 *   or_localization / or_fitting / or_wavelett   golden CSVs + FitLocalizations == Localization
 *                        (Tests_CPU/test_Palm.cpp:269-295) + per-plane capacity semantics.
 *   or_astig_estimate    the axial cases of Tests_CPU/test_fit_AstigFit.cpp:41-54, 241-345 (bounds,
 *                        between-grid refinement, score formula and scale invariance, invalid widths)
 *                        + bit equality with the compiled reference on random and degenerate pairs.
 *
 * Floating-point: built with -ffp-contract=off so that every product and sum rounds as in
 * the reference's x86-64 build (which has no FMA at its default target).
 */
#define _GNU_SOURCE   /* sincos */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_PLANE_ID 1        /* lib/Constants.hpp:20 */
#define OR_VOLUME_MIN 4.0    /* lib/Constants.hpp:23 */
#define OR_THRESHOLD_IT 5    /* lib/Constants.hpp:24 */
#define OR_EPSILON 1e-7      /* lib/Constants.hpp:53 */
#define OR_SEG_THRESHOLD 1e-8 /* lib/localization/Segmentation2D.hpp:80 */
#define OR_COLS 18           /* lib/Constants.hpp:50 */
#define OR_PI 3.14159265358979323846
#define OR_SQRT2 1.41421356237309504880

/* ------------------------------------------------------------------------------------------ */
/* A trous (lib/localization/ATrous.cpp:23-80)                                                   */
/* ------------------------------------------------------------------------------------------ */

static inline size_t clampi(long v, long hi) { return (size_t)(v < 0 ? 0 : (v > hi ? hi : v)); }

/* b3SplineFiltering2D (ATrous.cpp:49-80): rows pass then columns pass, taps at s = 2^plane. */
static void b3_filter(const double* in, double* tmp, double* out, size_t w, size_t h, int plane)
{
    const double c0 = 3.0 / 8.0, c1 = 1.0 / 4.0, c2 = 1.0 / 16.0;
    const long s = 1L << plane, s2 = 2 * s, W = (long)w - 1, H = (long)h - 1;
    for (size_t i = 0; i < h; ++i) {
        const double* r = in + i * w;
        for (size_t j = 0; j < w; ++j) {
            const long J = (long)j;
            tmp[i * w + j] = (c0 * r[j]) + (c1 * (r[clampi(J - s, W)] + r[clampi(J + s, W)])) +
                             (c2 * (r[clampi(J - s2, W)] + r[clampi(J + s2, W)]));
        }
    }
    for (size_t i = 0; i < h; ++i) {
        const long I = (long)i;
        const double* a = tmp + clampi(I - s, H) * w;
        const double* b = tmp + clampi(I + s, H) * w;
        const double* c = tmp + clampi(I - s2, H) * w;
        const double* d = tmp + clampi(I + s2, H) * w;
        for (size_t j = 0; j < w; ++j) {
            out[i * w + j] = (c0 * tmp[i * w + j]) + (c1 * (a[j] + b[j])) + (c2 * (c[j] + d[j]));
        }
    }
}

void or_atrous(const uint16_t* img, size_t w, size_t h, double* w0, double* w1, double* v2)
{
    const size_t n = w * h;
    double* p[3];
    double* own = (double*)malloc(sizeof(double) * n * 4);
    p[0] = own; p[1] = own + n; p[2] = own + 2 * n;
    double* tmp = own + 3 * n;
    for (size_t i = 0; i < n; ++i) p[0][i] = (double)img[i];          /* ATrous.cpp:111 */
    for (int s = 0; s < 2; ++s) {                                      /* ATrous.cpp:114-119 */
        b3_filter(p[s], tmp, p[s + 1], w, h, s);
        for (size_t i = 0; i < n; ++i) p[s][i] -= p[s + 1][i];
    }
    if (w0) memcpy(w0, p[0], sizeof(double) * n);
    if (w1) memcpy(w1, p[1], sizeof(double) * n);
    if (v2) memcpy(v2, p[2], sizeof(double) * n);
    free(own);
}

/* ------------------------------------------------------------------------------------------ */
/* Segmentation (lib/localization/Segmentation2D.cpp)                                            */
/* ------------------------------------------------------------------------------------------ */

/* Binary max-heap with the element order and move sequence of libstdc++'s std::priority_queue
 * (bits/stl_heap.h: __push_heap / __adjust_heap / __pop_heap) over Pixel, whose operator< is
 * reversed (lib/Structs.hpp:26), so the top is the smallest key. Equal keys therefore leave the
 * heap in the same order as the reference binary built with GCC's libstdc++. */
typedef struct { double v; uint32_t idx; } heap_el;
typedef struct { heap_el* a; size_t n, cap; } heap_t;

/* comp(a, b) == (a < b) == (a.v > b.v)  (lib/Structs.hpp:26) */
static inline int hcomp(const heap_el* a, const heap_el* b) { return a->v > b->v; }

static void heap_sift_up(heap_el* a, size_t hole, size_t top, heap_el val)
{
    size_t parent = (hole - 1) / 2;
    while (hole > top && hcomp(&a[parent], &val)) {
        a[hole] = a[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[hole] = val;
}

static void heap_push(heap_t* h, double v, uint32_t idx)
{
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 1024;
        h->a = (heap_el*)realloc(h->a, sizeof(heap_el) * h->cap);
    }
    heap_el e = {v, idx};
    h->a[h->n++] = e;
    heap_sift_up(h->a, h->n - 1, 0, e);
}

static heap_el heap_pop(heap_t* h)
{
    heap_el* a = h->a;
    heap_el top = a[0];
    if (h->n > 1) {
        const size_t len = h->n - 1;      /* __pop_heap over [first, last-1) */
        heap_el val = a[len];
        a[len] = a[0];
        size_t hole = 0, second = 0;
        while (second < (len - 1) / 2) {  /* __adjust_heap */
            second = 2 * (second + 1);
            if (hcomp(&a[second], &a[second - 1])) second--;
            a[hole] = a[second];
            hole = second;
        }
        if ((len & 1) == 0 && second == (len - 2) / 2) {
            second = 2 * (second + 1);
            a[hole] = a[second - 1];
            hole = second - 1;
        }
        heap_sift_up(a, hole, 0, val);
    }
    h->n--;
    return top;
}

typedef struct {
    size_t w, h, n;
    double* out;       /* output_ (thresholded potential) */
    uint32_t* map;     /* map_ */
    heap_t heap;
} seg_t;

/* getLocalMax (Segmentation2D.cpp:294-333). Returns label count. */
static size_t seg_local_max(seg_t* s)
{
    const size_t w = s->w, h = s->h;
    memset(s->map, 0, sizeof(uint32_t) * s->n);
    uint32_t label = 1;
    for (size_t i = 0, idx = 0; i < h; ++i) {
        for (size_t j = 0; j < w; ++j, ++idx) {
            if (!(s->out[idx] > OR_SEG_THRESHOLD)) continue;
            const size_t i0 = i == 0 ? 0 : i - 1, i1 = (i + 1 < h - 1) ? i + 1 : h - 1;
            const size_t j0 = j == 0 ? 0 : j - 1, j1 = (j + 1 < w - 1) ? j + 1 : w - 1;
            int err = 0;
            for (size_t ii = i0; ii <= i1 && !err; ++ii)
                for (size_t jj = j0; jj <= j1 && !err; ++jj)
                    if ((ii != i || jj != j) && s->out[idx] < s->out[ii * w + jj]) err = 1;
            if (err) continue;
            uint32_t maxi = 0;
            int found = 0;
            for (size_t ii = i0; ii <= i1 && !found; ++ii)
                for (size_t jj = j0; jj <= j1 && !found; ++jj)
                    if ((ii != i || jj != j) && maxi < s->map[ii * w + jj]) { maxi = s->map[ii * w + jj]; found = 1; }
            s->map[idx] = maxi > 0 ? maxi : label++;
        }
    }
    return label - 1;
}

/* watershed (Segmentation2D.cpp:336-385): seeds in raster order, global pop counter 'it' in the
 * key of every later push, label claimed at push time. */
static void seg_watershed(seg_t* s)
{
    const long w = (long)s->w, h = (long)s->h;
    s->heap.n = 0;
    for (long x = 0, idx = 0; x < h; ++x) {
        for (long y = 0; y < w; ++y, ++idx) {
            if (s->map[idx] == 0) continue;
            int vois = 0;
            for (int dx = -1; dx <= 1 && !vois; ++dx) {
                const long nx = x + dx;
                if (nx < 0 || nx >= h) continue;
                for (int dy = -1; dy <= 1 && !vois; ++dy) {
                    const long ny = y + dy;
                    if ((dx == 0 && dy == 0) || ny < 0 || ny >= w) continue;
                    if (s->map[nx * w + ny] == 0) vois = 1;
                }
            }
            if (vois) heap_push(&s->heap, -s->out[idx], (uint32_t)idx);
        }
    }
    size_t it = 0;
    while (s->heap.n) {
        const heap_el p = heap_pop(&s->heap);
        ++it;
        const long px = p.idx / w, py = p.idx % w;
        for (int dx = -1; dx <= 1; ++dx) {
            const long nx = px + dx;
            if (nx < 0 || nx >= h) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const long ny = py + dy;
                if ((dx == 0 && dy == 0) || ny < 0 || ny >= w) continue;
                const long nid = nx * w + ny;
                if (s->out[nid] > OR_SEG_THRESHOLD && s->map[nid] == 0) {
                    const double key = -s->out[nid] + ((double)it * OR_EPSILON);
                    heap_push(&s->heap, key, (uint32_t)nid);
                    s->map[nid] = s->map[p.idx];
                }
            }
        }
    }
}


typedef struct {
    double maxPotWithNeighbor, minPot, maxPot;
    uint32_t x, y, neighborLabel, structToBecome, newIndex;
    int locked;
} ws_fusion; /* WSFusion (lib/localization/SegmentationInfo.hpp:19-29) */

/* filterStructureWS (Segmentation2D.cpp:181-291): iterative basin fusion, then one seed pixel per
 * surviving region. *nregions is the label count in and out. */
static void seg_fusion(seg_t* s, double ratio, size_t* nregions)
{
    const size_t nb = *nregions;
    if (nb == 0) return;
    const size_t w = s->w, h = s->h;
    ws_fusion* f = (ws_fusion*)calloc(nb, sizeof(ws_fusion));
    size_t nf = nb;
    int change = 1;
    uint32_t label = 0;
    while (change) {
        change = 0;
        for (size_t i = 0; i < nf; ++i) {
            f[i].maxPotWithNeighbor = -1e10;
            f[i].minPot = +1e10;
            f[i].maxPot = -1e10;
            f[i].neighborLabel = 0;
            f[i].locked = 0;
            f[i].structToBecome = (uint32_t)i;
            f[i].newIndex = 0;
        }
        for (size_t i = 0, idx = 0; i < h; ++i) {
            for (size_t j = 0; j < w; ++j, ++idx) {
                const uint32_t sid = s->map[idx];
                if (sid == 0) continue;
                ws_fusion* st = &f[sid - 1];
                const double o = s->out[idx];
                if (o < st->minPot) st->minPot = o;                 /* std::min */
                if (o > st->maxPot) { st->maxPot = o; st->x = (uint32_t)i; st->y = (uint32_t)j; }
                const size_t i0 = (i > 1 ? i : 1) - 1, i1 = (i + 1 < h - 1) ? i + 1 : h - 1;
                const size_t j0 = (j > 1 ? j : 1) - 1, j1 = (j + 1 < w - 1) ? j + 1 : w - 1;
                for (size_t ii = i0; ii <= i1; ++ii)
                    for (size_t jj = j0; jj <= j1; ++jj) {
                        if (ii == i && jj == j) continue;
                        const uint32_t nsid = s->map[ii * w + jj];
                        if (nsid != 0 && nsid != sid && o > st->maxPotWithNeighbor) {
                            st->maxPotWithNeighbor = o;
                            st->neighborLabel = nsid;
                        }
                    }
            }
        }
        for (size_t i = 0; i < nf; ++i) {
            ws_fusion* st = &f[i];
            if (st->neighborLabel == 0) continue;
            const size_t ni = st->neighborLabel - 1;
            if (!st->locked && !f[ni].locked) {
                const double delta = st->maxPot - st->minPot;
                if (fabs(delta) <= DBL_EPSILON) continue;
                const double r = (st->maxPot - st->maxPotWithNeighbor) / delta;
                if (r < ratio && st->maxPot < f[ni].maxPot) {
                    st->structToBecome = (uint32_t)ni;
                    f[ni].locked = 1;
                    st->locked = 1;
                    change = 1;
                }
            }
        }
        label = 0;
        for (size_t i = 0; i < nf; ++i)
            if (f[i].structToBecome == i) f[i].newIndex = ++label;
        for (size_t i = 0; i < s->n; ++i)
            if (s->map[i] != 0) s->map[i] = f[f[s->map[i] - 1].structToBecome].newIndex;
        nf = label; /* wsFusions.resize(label) */
    }
    /* infos_ are re-sized to `label`; they are recomputed by ComputeStatistics afterwards. */
    memset(s->map, 0, sizeof(uint32_t) * s->n);
    for (size_t i = 0; i < nf; ++i)
        if (f[i].structToBecome == i) s->map[(size_t)f[i].x * w + f[i].y] = f[i].newIndex;
    *nregions = label;
    free(f);
}

/* ComputeStatistics (Segmentation2D.cpp:144-175) + FilterData (:88-141; only the surface filter
 * surf <= VOLUME_MIN is enabled, :23-31). Returns the kept region count. */
static size_t seg_stats_filter(seg_t* s, size_t nregions, or_region* out, size_t cap)
{
    const size_t w = s->w, h = s->h;
    double* acc = (double*)calloc(nregions * 4 + 1, sizeof(double)); /* surf, sx, sy, sum */
    for (size_t i = 0, idx = 0; i < h; ++i) {
        for (size_t j = 0; j < w; ++j, ++idx) {
            const uint32_t sl = s->map[idx];
            if (sl == 0) continue;
            double* a = acc + 4 * (sl - 1);
            const double v = s->out[idx];
            a[0] += 1.0;
            a[1] += v * (double)j;
            a[2] += v * (double)i;
            a[3] += v;
        }
    }
    uint32_t* remap = (uint32_t*)calloc(nregions + 1, sizeof(uint32_t));
    size_t kept = 0;
    for (size_t r = 0; r < nregions; ++r) {
        double* a = acc + 4 * r;
        if (a[0] <= OR_VOLUME_MIN) continue; /* CheckFilter(surf, lessThan): surf <= 4 rejected */
        remap[r] = (uint32_t)(++kept);
        if (kept <= cap) {
            or_region* o = &out[kept - 1];
            o->x = a[1] / a[3];
            o->y = a[2] / a[3];
            o->z = 0.0;
            o->intensity = a[3];
            o->surf = a[0];
        }
    }
    for (size_t i = 0; i < s->n; ++i)
        if (s->map[i] > 0) s->map[i] = remap[s->map[i] - 1];
    free(remap);
    free(acc);
    return kept;
}

static void seg_init(seg_t* s, const double* w1, size_t w, size_t h, double thr)
{
    s->w = w; s->h = h; s->n = w * h;
    s->out = (double*)malloc(sizeof(double) * (s->n + 1));
    s->map = (uint32_t*)malloc(sizeof(uint32_t) * (s->n + 1));
    s->heap.a = NULL; s->heap.n = s->heap.cap = 0;
    for (size_t i = 0; i < s->n; ++i) {                 /* Segmentation2D.cpp:60-70 */
        if (w1[i] > thr) { s->out[i] = w1[i]; s->map[i] = 1; }
        else { s->out[i] = 0; s->map[i] = 0; }
    }
}

static void seg_free(seg_t* s) { free(s->out); free(s->map); free(s->heap.a); }

size_t or_watershed_labels(const double* w1, size_t w, size_t h, double thr, uint32_t* labels)
{
    if (w == 0 || h == 0) return 0;
    seg_t s;
    seg_init(&s, w1, w, h, thr);
    const size_t n = seg_local_max(&s);
    seg_watershed(&s);
    memcpy(labels, s.map, sizeof(uint32_t) * s.n);
    seg_free(&s);
    return n;
}

size_t or_segment(const double* w1, size_t w, size_t h, double thr, double ratio, uint32_t* labels,
                  or_region* regions, size_t cap)
{
    if (w == 0 || h == 0) return 0;
    seg_t s;
    seg_init(&s, w1, w, h, thr);
    size_t n = seg_local_max(&s);                  /* WatershedIntelligent (:77-85) */
    seg_watershed(&s);
    if ((int)ratio > 0) {
        seg_fusion(&s, ratio, &n);
        seg_watershed(&s);
    }
    const size_t kept = seg_stats_filter(&s, n, regions, cap);
    if (labels) memcpy(labels, s.map, sizeof(uint32_t) * s.n);
    seg_free(&s);
    return kept;
}

/* ------------------------------------------------------------------------------------------ */
/* Gauss-Jordan inverse (lib/signal_processing/Matrix.cpp:297-355)                              */
/* ------------------------------------------------------------------------------------------ */

int or_inverse_gauss_jordan(double* in, double* out, size_t n)
{
    for (size_t i = 0; i < n * n; ++i) out[i] = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double scale = 0.0;
        for (size_t j = 0; j < n; ++j) {
            const double v = in[i * n + j];
            if (!isfinite(v)) goto fail;
            const double a = fabs(v);
            if (scale < a) scale = a;                   /* std::max(scale, a) */
        }
        if (scale == 0.0) goto fail;
        for (size_t j = 0; j < n; ++j) in[i * n + j] /= scale;
        out[i * n + i] = 1.0 / scale;
    }
    {
        const double tol = DBL_EPSILON * (double)n;
        for (size_t i = 0; i < n; ++i) {
            size_t pr = i;
            for (size_t j = i + 1; j < n; ++j)
                if (fabs(in[j * n + i]) > fabs(in[pr * n + i])) pr = j;
            if (!isfinite(in[pr * n + i]) || fabs(in[pr * n + i]) <= tol) goto fail;
            if (pr != i) {
                for (size_t j = 0; j < n; ++j) {
                    double t = in[i * n + j]; in[i * n + j] = in[pr * n + j]; in[pr * n + j] = t;
                    t = out[i * n + j]; out[i * n + j] = out[pr * n + j]; out[pr * n + j] = t;
                }
            }
            const double pinv = 1.0 / in[i * n + i];
            for (size_t j = 0; j < n; ++j) { in[i * n + j] *= pinv; out[i * n + j] *= pinv; }
            for (size_t j = 0; j < n; ++j) {
                if (j == i) continue;
                const double fct = in[j * n + i];
                for (size_t k = 0; k < n; ++k) {
                    in[j * n + k] -= fct * in[i * n + k];
                    out[j * n + k] -= fct * out[i * n + k];
                }
            }
        }
        for (size_t i = 0; i < n * n; ++i)
            if (!isfinite(in[i]) || !isfinite(out[i])) goto fail;
    }
    return 0;
fail:
    for (size_t i = 0; i < n * n; ++i) out[i] = 0.0;
    return -1;
}

/* ------------------------------------------------------------------------------------------ */
/* Gaussian model + LM fit (lib/fit/GaussianModel.cpp, GaussianFit.cpp, Fitting.cpp)              */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    double offset, amp, sx, sy, sx2, sy2, c, s, c2, s2, cs;
    int iso, hasSigma, hasTheta;
} gmodel;

/* PrepareModel (GaussianFit.cpp:22-34) + GaussianModel ctors (GaussianModel.cpp:8-16). */
static gmodel prepare_model(int mode, const double* p, double sigma, double theta)
{
    gmodel m;
    memset(&m, 0, sizeof(m));
    m.offset = p[0];
    m.amp = p[1];
    if (mode == 1 || mode == 2) {
        const double sg = mode == 1 ? sigma : p[4];
        m.sx = m.sy = sg;
        m.sx2 = m.sy2 = sg * sg;
        m.c = 1; m.s = 0; m.c2 = 1; m.s2 = 0; m.cs = 0;
        m.iso = 1; m.hasSigma = mode == 2; m.hasTheta = 0;
    } else {
        const double th = mode == 3 ? theta : p[6];
        m.sx = p[4]; m.sy = p[5];
        m.sx2 = p[4] * p[4]; m.sy2 = p[5] * p[5];
        /* The reference's std::cos / std::sin of one angle, which gcc merges into one sincos call
         * (the compiled reference imports sincos only); glibc's sincos can differ from its sin and
         * cos in the last bit. */
        sincos(th, &m.s, &m.c);
        m.c2 = m.c * m.c; m.s2 = m.s * m.s; m.cs = m.c * m.s;
        m.iso = 0; m.hasSigma = 0; m.hasTheta = mode == 4;
    }
    return m;
}

/* GaussianModel::Evaluate (GaussianModel.cpp:19-52). g may be NULL. */
static double model_eval(const gmodel* m, double x, double y, double* g)
{
    const double X2 = x * x, Y2 = y * y;
    if (m->iso) {
        const double E = exp(-(X2 + Y2) / (2 * m->sx2));
        const double P = m->amp * E;
        if (g) {
            g[0] = 1;
            g[1] = E;
            g[2] = P * x / m->sx2;
            g[3] = P * y / m->sx2;
            if (m->hasSigma) g[4] = P / (m->sx * m->sx2) * (X2 + Y2);
        }
        return P + m->offset;
    }
    const double XC = x * m->c, XS = x * m->s, YC = y * m->c, YS = y * m->s;
    const double F1 = (XC - YS) / m->sx, F2 = (XS + YC) / m->sy;
    const double E = exp(-((F1 * F1) + (F2 * F2)));
    const double P = m->amp * E;
    if (g) {
        const double XY = x * y, F3 = 2 * XY * m->cs, F4 = (2 * XY / m->sx2) - (2 * XY / m->sy2);
        g[0] = 1;
        g[1] = E;
        g[2] = P * (-(((-2 * m->c2 * x) + (2 * y * m->cs)) / m->sx2) - (((-2 * m->s2 * x) - (2 * y * m->cs)) / m->sy2));
        g[3] = P * ((((2 * m->s2 * y) - (2 * x * m->cs)) / m->sx2) - (((-2 * m->c2 * y) - (2 * x * m->cs)) / m->sy2));
        g[4] = P * (2 * ((XC * XC) - F3 + (YS * YS)) / (m->sx * m->sx2));
        g[5] = P * (2 * ((XS * XS) + F3 + (YC * YC)) / (m->sy * m->sy2));
        if (m->hasTheta)
            g[6] = P * ((2 * m->cs * (((X2 - Y2) / m->sx2) + ((Y2 - X2) / m->sy2))) - (m->s2 * F4) + (m->c2 * F4));
    }
    return P + m->offset;
}

/* Fitting::fill (lib/fit/Fitting.cpp:61-106). */
static void fill_row(int mode, const or_region* info, double* pt, double x, double y, double z, double offset,
                     double intensity, double sx, double sy, double theta, double rssXY, double rssZ)
{
    pt[0] = 0; pt[1] = 0; pt[2] = 0; pt[3] = -1;
    pt[4] = x; pt[5] = y; pt[6] = z;
    if (mode == 0) pt[7] = info->intensity;
    else if (mode == 5) pt[7] = intensity;                              /* Fitting.cpp:86-88 */
    else pt[7] = (intensity > OR_EPSILON && sx > OR_EPSILON && sy > OR_EPSILON) ? 2.0 * OR_PI * intensity * sx * sy : -1;
    pt[8] = sx; pt[9] = sy; pt[10] = theta; pt[11] = rssXY; pt[12] = rssZ;
    pt[13] = intensity; pt[14] = offset; pt[15] = info->intensity; pt[16] = info->surf;
    if (sx > OR_EPSILON && sy > OR_EPSILON) pt[17] = (sy < sx ? sy : sx) / (sx < sy ? sy : sx);
    else pt[17] = 1;
}

static inline void fill_fail(int mode, const or_region* info, double* pt, double x, double y, double z)
{
    fill_row(mode, info, pt, x, y, z, -1, -1, -1, -1, -1, -1, -1);
}

/* size_t conversion of a double as GCC/x86-64 performs it (cvttsd2si): truncation toward zero,
 * negative values wrap, NaN gives 2^63. */
static inline size_t to_size(double v)
{
    if (v != v) return (size_t)1 << 63;
    if (v >= 9223372036854775808.0) return (size_t)1 << 63;
    if (v <= -9223372036854775808.0) return (size_t)1 << 63;
    return (size_t)(long long)v;
}

typedef struct {
    int mode, np;
    size_t roi, half, n;
    double sigmaX, sigmaY, theta;
    double alpha[49], alphaR[49], alphaRi[49], beta[7], dP[7], lbdp[7], grad[7], pTry[7];
    double lo[7], p[7], hi[7];
    double* gx; double* gy; double* gv; double* hat;
} lm_ws;

/* GaussianFit::lm (GaussianFit.cpp:219-245) */
static double lm_pass(lm_ws* ws)
{
    const int np = ws->np;
    for (int i = 0; i < np * np; ++i) ws->alpha[i] = 0.0;
    for (int i = 0; i < np; ++i) ws->beta[i] = 0.0;
    double err = 0.0;
    const gmodel m = prepare_model(ws->mode, ws->p, ws->sigmaX, ws->theta);
    const int circular = ws->mode == 4 && fabs(ws->p[4] - ws->p[5]) <= DBL_EPSILON;
    for (size_t i = 0; i < ws->n; ++i) {
        ws->hat[i] = model_eval(&m, ws->gx[i] - ws->p[2], ws->gy[i] - ws->p[3], ws->grad);
        const double d = ws->gv[i] - ws->hat[i];
        const double dw = 1.0 * d;
        err += d * dw;
        if (circular) ws->grad[6] = 0.0;
        for (int a = 0; a < np; ++a) {
            ws->beta[a] += ws->grad[a] * dw;
            for (int b = 0; b < np; ++b) ws->alpha[a * np + b] += 1.0 * ws->grad[a] * ws->grad[b];
        }
    }
    return err;
}

/* GaussianFit::gain (GaussianFit.cpp:248-259) */
static double lm_gain(const lm_ws* ws)
{
    double quad = 0.0, lin = 0.0;
    for (int i = 0; i < ws->np; ++i) {
        double ti = 0.0;
        for (int j = 0; j < ws->np; ++j) ti += ws->alpha[i * ws->np + j] * ws->dP[j];
        quad += ws->dP[i] * ti;
        lin += ws->dP[i] * ws->beta[i];
    }
    return (2.0 * lin) - quad;
}

/* GaussianFit::getError (GaussianFit.cpp:262-271) */
static double lm_error(const lm_ws* ws)
{
    double res = 0.0;
    const gmodel m = prepare_model(ws->mode, ws->pTry, ws->sigmaX, ws->theta);
    for (size_t i = 0; i < ws->n; ++i) {
        const double dx = ws->gv[i] - model_eval(&m, ws->gx[i] - ws->pTry[2], ws->gy[i] - ws->pTry[3], NULL);
        res += 1.0 * dx * dx;
    }
    return res;
}

/* GaussianFit::computeR2 (GaussianFit.cpp:274-295); mutates gv and hat as the reference does. */
static double lm_r2(lm_ws* ws)
{
    double sd = 0.0, sm = 0.0;
    for (size_t i = 0; i < ws->n; ++i) sd = sd + ws->gv[i];
    for (size_t i = 0; i < ws->n; ++i) sm = sm + ws->hat[i];
    const double muD = sd / (double)ws->n, muM = sm / (double)ws->n;
    for (size_t i = 0; i < ws->n; ++i) { ws->gv[i] -= muD; ws->hat[i] -= muM; }
    double t1 = 0.0, t2 = 0.0, t3 = 0.0;
    for (size_t i = 0; i < ws->n; ++i) t1 = t1 + ws->gv[i] * ws->hat[i];
    for (size_t i = 0; i < ws->n; ++i) t2 = t2 + ws->gv[i] * ws->gv[i];
    for (size_t i = 0; i < ws->n; ++i) t3 = t3 + ws->hat[i] * ws->hat[i];
    const double r = t1 / sqrt(t2 * t3);
    return isnan(r) ? 0 : r * r;
}

static double max_abs(const double* v, int n)
{
    double m = fabs(v[0]);
    for (int i = 1; i < n; ++i) { const double a = fabs(v[i]); if (m < a) m = a; }
    return m;
}

static double rel_step(const lm_ws* ws)
{
    double m = 0.0;
    for (int i = 0; i < ws->np; ++i) {
        const double ap = fabs(ws->p[i]);
        const double r = fabs(ws->dP[i]) / (ap < 1e-12 ? 1e-12 : ap);
        if (m < r) m = r;
    }
    return m;
}

/* ------------------------------------------------------------------------------------------ */
/* Spline PSF fit, EMCCD Poisson likelihood (lib/fit/SplineFit.cpp:36-318, Fitting.cpp:60-175,  */
/* Matrix.cpp:245-296, :358-411). fp = [roi, spline x, y, z sizes, deltaZ, 64 x size coefficients]. */
/* ------------------------------------------------------------------------------------------ */

#define OR_SPL_N 5
#define OR_SPL_EPS2 1e-10        /* SplineFit.hpp: EPSILON_2 */

typedef struct {
    size_t roi, half, n, sx, sy, sz, sxy, ssize, off;
    double initZ, deltaZ, ratio, roiParam;
    const double* c;
} spl_ws;

/* parseParams (SplineFit.cpp:36-49), then the odd ROI side of Fitting::initialize (:35-39). */
static void spl_parse(const double* fp, spl_ws* s)
{
    s->roiParam = fp[0];
    s->roi = to_size(fp[0]);
    s->sx = to_size(fp[1]); s->sy = to_size(fp[2]); s->sz = to_size(fp[3]);
    s->sxy = s->sx * s->sy;
    s->ssize = s->sx * s->sy * s->sz;
    s->initZ = (double)s->sz / 2.0;
    s->deltaZ = fp[4];
    s->c = fp + 5;
    const double center = s->c[(32 * s->ssize) + (s->sz / 2 * s->sxy) + (s->sy / 2 * s->sx) + (s->sx / 2)];
    s->ratio = fabs(center) < DBL_EPSILON ? 1.0 : 4.0 / center;
    s->off = to_size(floor(((double)s->sx + 1.0 - (double)s->roi) / 2));
    if (s->roi % 2 == 0) s->roi--;
    s->half = s->roi / 2;
    s->n = s->roi * s->roi;
}

/* delta3D (SplineFit.cpp:296-317): monomials x^k y^j z^i of the 4x4x4 patch and their derivatives. */
static void spl_delta3d(double x, double y, double z, double* f, double* dfx, double* dfy, double* dfz)
{
    for (int i = 0; i < 64; ++i) { f[i] = 0.0; dfx[i] = 0.0; dfy[i] = 0.0; dfz[i] = 0.0; }
    double cz = 1.0;
    for (int i = 0; i < 4; ++i) {
        double cy = 1.0;
        for (int j = 0; j < 4; ++j) {
            double cx = 1.0;
            for (int k = 0; k < 4; ++k) {
                const int id = (i * 16) + (j * 4) + k;
                const double coef = cz * cy * cx;
                f[id] = coef;
                if (k + 1 < 4) dfx[id + 1] = (double)(k + 1) * coef;
                if (j + 1 < 4) dfy[id + 4] = (double)(j + 1) * coef;
                if (i + 1 < 4) dfz[id + 16] = (double)(i + 1) * coef;
                cx *= x;
            }
            cy *= y;
        }
        cz *= z;
    }
}

static inline size_t clamp_sz(size_t v, size_t hi) { return hi < v ? hi : v; }   /* std::clamp<size_t>(v, 0, hi) */
static inline double clamp_d(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

/* derivativeSpline (SplineFit.cpp:219-240): pos clamped in place (unsigned: negative indices wrap
 * and clamp to the far edge). */
static double spl_deriv(const spl_ws* s, size_t* px, size_t* py, size_t* pz, const double* f, const double* dfx,
                        const double* dfy, const double* dfz, const double* p, double* dudt)
{
    *px = clamp_sz(*px, s->sx - 1); *py = clamp_sz(*py, s->sy - 1); *pz = clamp_sz(*pz, s->sz - 1);
    const size_t pos = (((*pz * s->sy) + *py) * s->sx) + *px;
    for (int i = 0; i < OR_SPL_N; ++i) dudt[i] = 0.0;
    for (size_t i = 0; i < 64; ++i) {
        const size_t id = (i * s->ssize) + pos;
        dudt[1] += f[i] * s->c[id];
        dudt[2] += dfx[i] * s->c[id];
        dudt[3] += dfy[i] * s->c[id];
        dudt[4] += dfz[i] * s->c[id];
    }
    dudt[0] = 1;
    dudt[2] *= -1 * p[1];
    dudt[3] *= -1 * p[1];
    dudt[4] *= p[1];
    return p[0] + (p[1] * dudt[1]);
}

/* Cholesky (Matrix.cpp:272-294). */
static int spl_cholesky(const double* a, double* l, double* u)
{
    const int n = OR_SPL_N;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i + 1; ++j) {
            double sum = 0;
            for (int k = 0; k < j; ++k) sum += u[(i * n) + k] * u[(j * n) + k];
            if (i == j) {
                if (a[(i * n) + i] - sum >= 0) {
                    u[(i * n) + j] = sqrt(a[(i * n) + i] - sum);
                    l[(j * n) + i] = u[(i * n) + j];
                } else {
                    return 0;
                }
            } else {
                u[(i * n) + j] = 1.0 / u[(j * n) + j] * (a[(i * n) + j] - sum);
                l[(j * n) + i] = u[(i * n) + j];
            }
        }
    return 1;
}

/* luEvaluate (SplineFit.cpp:270-293). */
static void spl_lu_eval(const double* l, const double* u, const double* b, double* x)
{
    const int n = OR_SPL_N;
    double y[OR_SPL_N];
    for (int i = 0; i < n; ++i) {
        y[i] = b[i];
        for (int j = 0; j < i; ++j) y[i] -= l[(j * n) + i] * y[j];
        y[i] /= l[(i * n) + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        x[i] = y[i];
        for (int j = i + 1; j < n; j++) x[i] -= u[(j * n) + i] * x[j];
        x[i] /= u[(i * n) + i];
    }
}

/* InverseLU (Matrix.cpp:358-411), in place on in. */
static void spl_inverse_lu(double* in, double* out)
{
    const size_t n = OR_SPL_N;
    double tmp1 = 0, y[OR_SPL_N * OR_SPL_N];
    for (size_t j = 0; j < n; ++j) {
        for (size_t i = 0; i <= j; ++i) {
            if (i > 0) {
                for (size_t k = 0; k <= i - 1; ++k) tmp1 += in[i + (k * n)] * in[k + (j * n)];
                in[i + (j * n)] -= tmp1;
                tmp1 = 0;
            }
        }
        for (size_t i = j + 1; i < n; ++i) {
            if (j > 0) {
                for (size_t k = 0; k <= j - 1; ++k) tmp1 += in[i + (k * n)] * in[k + (j * n)];
                in[i + (j * n)] = (1 / in[j + (j * n)]) * (in[i + (j * n)] - tmp1);
                tmp1 = 0;
            } else {
                in[i + (j * n)] = (1 / in[j + (j * n)]) * in[i + (j * n)];
            }
        }
    }
    for (size_t num = 0; num < n; ++num) {
        y[0] = num == 0 ? 1 : 0;
        for (size_t i = 1; i < n; ++i) {
            const size_t b = i == num ? 1 : 0;
            for (size_t j = 0; j <= i - 1; ++j) tmp1 += in[i + (j * n)] * y[j];
            y[i] = (double)b - tmp1;
            tmp1 = 0;
        }
        out[n - 1 + (num * n)] = y[n - 1] / in[(n - 1) + ((n - 1) * n)];
        for (int i = (int)n - 2; i >= 0; --i) {
            for (size_t j = (size_t)i + 1; j < n; ++j) tmp1 += in[(size_t)i + (j * n)] * out[j + (num * n)];
            out[(size_t)i + (num * n)] = (1 / in[(size_t)i + ((size_t)i * n)]) * (y[i] - tmp1);
            tmp1 = 0;
        }
    }
}

/* SqrtReal (SplineFit.cpp:30-34): real part of the complex square root of x + 0i (glibc csqrt:
 * sqrt(x) for x > 0, +0 for x <= 0 and x = -inf, NaN for NaN). */
static double spl_sqrt_real(double x)
{
    if (x != x) return x;
    return x > 0 ? sqrt(x) : 0.0;
}

/* updateMuLambda (SplineFit.cpp:257-267). */
static void spl_mu_lambda(double* mu, double* lambda, int up)
{
    if (up) {
        const double a = (1 + (*lambda * 10)) / (1 + *lambda);
        *mu = a < 1.3 ? 1.3 : a;
        *lambda = 10 * *lambda;
    } else {
        *lambda = 0.1 * *lambda;
        *mu = 1 + *lambda;
    }
}

/* One pass of the model over the ROI at p: derivatives, and the callback per pixel. */
static void spl_pass(const spl_ws* s, const double* p, const double* gv, double* err, double* jac, double* hes,
                     double* fisher)
{
    double f[64], dfx[64], dfy[64], dfz[64], dudt[OR_SPL_N];
    const double xr = -1 * (p[2] - (double)s->half + 0.5), yr = -1 * (p[3] - (double)s->half + 0.5);
    const int xi = (int)floor(xr), yi = (int)floor(yr), zi = (int)floor(p[4]);
    const double xc = xr - xi, yc = yr - yi, zc = p[4] - zi;
    spl_delta3d(xc, yc, zc, f, dfx, dfy, dfz);
    size_t px = 0, py = 0, pz = (size_t)zi;
    for (size_t i = 0, idx = 0; i < s->roi; ++i) {
        px = (size_t)xi + i + s->off;
        for (size_t j = 0; j < s->roi; ++j, ++idx) {
            py = (size_t)yi + j + s->off;
            const double model = spl_deriv(s, &px, &py, &pz, f, dfx, dfy, dfz, p, dudt);
            if (fisher) {
                for (int i2 = 0; i2 < OR_SPL_N; ++i2)
                    for (int j2 = i2; j2 < OR_SPL_N; ++j2) {
                        fisher[(i2 * OR_SPL_N) + j2] += dudt[j2] * dudt[i2] / model;
                        fisher[(j2 * OR_SPL_N) + i2] = fisher[(i2 * OR_SPL_N) + j2];
                    }
                continue;
            }
            /* updateJacobianHessian (SplineFit.cpp:243-254) */
            const double data = gv[idx];
            if (data > 0) *err = *err + (2 * (model - data - (data * log(model / data))));
            else *err = *err + (2 * model);
            const double t1 = 1 - (data / model);
            for (int a = 0; a < OR_SPL_N; ++a) jac[a] += t1 * dudt[a];
            const double t2 = data / (model * model);
            for (int a = 0; a < OR_SPL_N; ++a)
                for (int b = a; b < OR_SPL_N; ++b) {
                    hes[(a * OR_SPL_N) + b] += t2 * dudt[a] * dudt[b];
                    hes[(b * OR_SPL_N) + a] = hes[(a * OR_SPL_N) + b];
                }
        }
    }
}

/* SplineFit::Fit (SplineFit.cpp:52-216). */
static void spline_fit_region(const uint16_t* img, size_t w, size_t h, const double* fp, const or_region* info,
                              double* pt)
{
    spl_ws s;
    spl_parse(fp, &s);
    const size_t ux = to_size(info->x), uy = to_size(info->y);
    if (s.half > ux || ux >= w - s.half || s.half > uy || uy >= h - s.half) {
        fill_row(5, info, pt, info->x, info->y, info->z, -1, -1, -1, -1, -1, -1, -1);
        return;
    }
    double z = s.initZ;
    double lambda = 0.1, err = DBL_MAX, mu = 1 + lambda;
    double hes[25], l[25], u[25], jac[OR_SPL_N], jump[OR_SPL_N], upd[OR_SPL_N] = {0, 0, 0, 0, 0};
    double gvbuf[225];
    double* gv = s.n <= 225 ? gvbuf : (double*)malloc(sizeof(double) * s.n);
    /* preCompute (Fitting.cpp:109-123) */
    double minI = DBL_MAX, maxI = -DBL_MAX;
    const long half = (long)s.half;
    for (long i = -half; i <= half; ++i)
        for (long j = -half; j <= half; ++j) {
            const size_t k = (size_t)(half + ((i + half) * (long)s.roi) + j);
            gv[k] = (double)img[ux + j + ((uy + i) * w)];
            if (gv[k] < minI) minI = gv[k];
            if (maxI < gv[k]) maxI = gv[k];
        }
    /* CenterOfMass (Matrix.cpp:245-269): ROI-local centre */
    double x = 0.0, y = 0.0, sum = 0.0;
    for (size_t i = 0, idx = 0; i < s.roi; ++i)
        for (size_t j = 0; j < s.roi; ++j, ++idx) {
            const double v = gv[idx];
            y += v * (double)i;
            x += v * (double)j;
            sum += v;
        }
    if (sum > 0) { x /= sum; y /= sum; }
    else { x = (double)s.roi / 2.0; y = (double)s.roi / 2.0; }
    /* initP (Fitting.cpp:159-171) and the spline adjustments (SplineFit.cpp:73-77) */
    const double roiC = (s.roiParam - 1) / 2.0, roiD = s.roiParam / 4.0;
    const double dmin = roiC - roiD, dmax = roiC + roiD;
    const double lo[OR_SPL_N] = {-0.01, -1, dmin, dmin, 0};
    double hi[OR_SPL_N] = {0.01, 1, dmax, dmax, 1};
    double p[OR_SPL_N] = {minI, maxI - minI, x, y, z};
    p[1] *= s.ratio;
    hi[4] = (double)s.sz;
    for (int i = 0; i < OR_SPL_N; ++i) p[i] = clamp_d(p[i], lo[i], hi[i]);
    jump[0] = 20.0 < p[1] ? p[1] : 20.0;
    jump[1] = 100.0 < p[1] ? p[1] : 100.0;
    jump[2] = 1.0;
    jump[3] = 1.0;
    jump[4] = (double)s.sz / 3.0 < 2.0 ? 2.0 : (double)s.sz / 3.0;

    for (size_t iteration = 0;;) {
        iteration++;
        const double oldErr = err, oldLambda = lambda;
        double oldP[OR_SPL_N], oldUpd[OR_SPL_N];
        memcpy(oldP, p, sizeof(p));
        memcpy(oldUpd, upd, sizeof(upd));
        for (int i = 0; i < OR_SPL_N; ++i) { jac[i] = 0.0; upd[i] = 0.0; }
        for (int i = 0; i < 25; ++i) { hes[i] = 0.0; l[i] = 0.0; u[i] = 0.0; }
        spl_pass(&s, p, gv, &err, jac, hes, NULL);
        for (int i = 0; i < OR_SPL_N; ++i) hes[(i * OR_SPL_N) + i] = hes[(i * OR_SPL_N) + i] * mu;
        if (spl_cholesky(hes, l, u)) {
            spl_lu_eval(l, u, jac, upd);
            for (int i = 0; i < OR_SPL_N; ++i) {
                if (fabs(oldUpd[i]) > OR_EPSILON && upd[i] / oldUpd[i] < -0.5) jump[i] *= 0.5;
                const double scale = upd[i] / (1.0 + (fabs(upd[i]) / jump[i]));
                p[i] -= scale;
                p[i] = clamp_d(p[i], lo[i], hi[i]);
                upd[i] = scale;
            }
            if (err > oldErr) {
                err = oldErr;
                memcpy(p, oldP, sizeof(p));
                memcpy(upd, oldUpd, sizeof(upd));
                lambda = oldLambda;
                spl_mu_lambda(&mu, &lambda, 1);
            } else {
                spl_mu_lambda(&mu, &lambda, 0);
            }
        } else {
            err = oldErr;
            memcpy(p, oldP, sizeof(p));
            memcpy(upd, oldUpd, sizeof(upd));
            lambda = oldLambda;
            spl_mu_lambda(&mu, &lambda, 1);
        }
        if (iteration >= 75) break;                                       /* Fitting.hpp: N_ITER */
        if (iteration > 2 && (err < 1e-6 || fabs(err - oldErr) / err < 1e-6)) break;
    }
    /* CRLB: inverse of the Fisher information at p (SplineFit.cpp:166-203) */
    double m[25], mInv[25], crlb[OR_SPL_N];
    for (int i = 0; i < 25; ++i) { m[i] = 0.0; mInv[i] = 0.0; }
    spl_pass(&s, p, gv, NULL, NULL, NULL, m);
    for (int i = 0; i < OR_SPL_N; ++i)
        if (fabs(m[(i * OR_SPL_N) + i]) < DBL_EPSILON) m[(i * OR_SPL_N) + i] = OR_SPL_EPS2;
    spl_inverse_lu(m, mInv);
    for (int i = 0; i < OR_SPL_N; ++i) crlb[i] = spl_sqrt_real(mInv[(i * OR_SPL_N) + i]);
    const double rssXY = (crlb[2] + crlb[3]) / 2, rssZ = s.deltaZ * crlb[4];
    const double fx = p[2] - (double)s.half + info->x;
    const double fy = p[3] - (double)s.half + info->y;
    const double fz = s.deltaZ * (p[4] - s.initZ);
    fill_row(5, info, pt, fx, fy, fz, crlb[0], crlb[1], 0, 0, 0, rssXY, rssZ);
    if (gv != gvbuf) free(gv);
}

void or_fit_region(int mode, const uint16_t* img, size_t w, size_t h, const double* fp, int maxIter,
                   const or_region* info, double* pt)
{
    if (mode == 0) { fill_row(0, info, pt, info->x, info->y, info->z, 0, 0, -1, -1, -1, -1, -1); return; }
    if (mode == 5) { spline_fit_region(img, w, h, fp, info, pt); return; }
    lm_ws ws;
    ws.mode = mode;
    ws.np = mode == 1 ? 4 : mode == 2 ? 5 : mode == 3 ? 6 : 7;      /* Fitting.cpp:14-40 */
    ws.roi = to_size(fp[0]);
    if (ws.roi % 2 == 0) ws.roi--;
    ws.half = ws.roi / 2;
    ws.n = ws.roi * ws.roi;
    ws.sigmaX = fp[1]; ws.sigmaY = fp[2]; ws.theta = fp[3];           /* GaussianFit.cpp:39-45 */
    const int np = ws.np;

    const size_t ux = to_size(info->x), uy = to_size(info->y);        /* GaussianFit.cpp:51-56 */
    if (ws.half > ux || ux >= w - ws.half || ws.half > uy || uy >= h - ws.half) {
        fill_fail(mode, info, pt, info->x, info->y, info->z);
        return;
    }
    double sx = ws.sigmaX, sy = ws.sigmaY, th = ws.theta;
    double lambda = 1e-2;
    double buf[4 * 128];
    double* heapbuf = NULL;
    if (ws.n <= 128) { ws.gx = buf; } else { heapbuf = (double*)malloc(sizeof(double) * 4 * ws.n); ws.gx = heapbuf; }
    ws.gy = ws.gx + ws.n; ws.gv = ws.gy + ws.n; ws.hat = ws.gv + ws.n;

    /* preCompute (Fitting.cpp:109-123) */
    double minI = DBL_MAX, maxI = -DBL_MAX;
    const long half = (long)ws.half;
    for (long i = -half; i <= half; ++i) {
        for (long j = -half; j <= half; ++j) {
            const size_t k = (size_t)((j + half) + (i + half) * (long)ws.roi);
            const size_t px = ux + j, py = uy + i;
            ws.gx[k] = (double)px; ws.gy[k] = (double)py;
            ws.gv[k] = (double)img[px + py * w];
            if (ws.gv[k] < minI) minI = ws.gv[k];
            if (maxI < ws.gv[k]) maxI = ws.gv[k];
        }
    }
    /* initP (Fitting.cpp:126-175) */
    const double x0 = info->x, y0 = info->y;
    ws.lo[0] = 0; ws.lo[1] = 0; ws.lo[2] = x0 - 1; ws.lo[3] = y0 - 1;
    ws.p[0] = minI; ws.p[1] = maxI - minI; ws.p[2] = x0; ws.p[3] = y0;
    ws.hi[0] = 65535.0; ws.hi[1] = 65535.0; ws.hi[2] = x0 + 1; ws.hi[3] = y0 + 1;
    if (mode == 2) { ws.lo[4] = 0.1; ws.p[4] = fp[1]; ws.hi[4] = 20.0; }
    if (mode >= 3) {
        ws.lo[4] = 0.1; ws.lo[5] = 0.1; ws.p[4] = OR_SQRT2 * fp[1]; ws.p[5] = OR_SQRT2 * fp[2];
        ws.hi[4] = 20.0; ws.hi[5] = 20.0;
    }
    if (mode == 4) { ws.lo[6] = -150; ws.p[6] = fp[3]; ws.hi[6] = 150; }

    double err = lm_pass(&ws);                                      /* GaussianFit.cpp:74-78 */
    int bad = !isfinite(err);
    for (int i = 0; i < np; ++i) bad |= !isfinite(ws.beta[i]);
    if (bad) { fill_fail(mode, info, pt, info->x, info->y, info->z); goto done; }

    double m = max_abs(ws.beta, np);
    for (int it = (m < 1e-6) ? maxIter : 0; it < maxIter; ++it) {   /* GaussianFit.cpp:83-163 */
        for (int i = 0; i < np * np; ++i) ws.alphaR[i] = ws.alpha[i];
        for (int i = 0; i < np; ++i) {
            const int d = i * np + i;
            ws.alphaR[d] = ws.alpha[d] == 0.0 ? 1.0 : (1.0 + lambda) * ws.alpha[d];
        }
        if (or_inverse_gauss_jordan(ws.alphaR, ws.alphaRi, (size_t)np) != 0) {
            if (lambda < 1e7) { lambda = fmin(lambda * 11, 1e7); continue; }
            fill_fail(mode, info, pt, info->x, info->y, info->z);
            goto done;
        }
        for (int i = 0; i < np; ++i) ws.dP[i] = 0.0;
        for (int i = 0; i < np; ++i) {
            for (int j = 0; j < np; ++j) ws.dP[i] += ws.alphaRi[i * np + j] * ws.beta[j];
            if (!isfinite(ws.dP[i])) { fill_fail(mode, info, pt, info->x, info->y, info->z); goto done; }
            const double v = ws.p[i] + ws.dP[i];
            ws.pTry[i] = v < ws.lo[i] ? ws.lo[i] : (ws.hi[i] < v ? ws.hi[i] : v);  /* std::clamp */
        }
        for (int i = 0; i < np; ++i) ws.dP[i] = ws.pTry[i] - ws.p[i];
        int tiny = 1;
        for (int i = 0; i < np; ++i) tiny &= fabs(ws.dP[i]) < 1e-15;
        if (tiny) {
            if (lambda >= 1e7) break;
            lambda = fmin(lambda * 11, 1e7);
            continue;
        }
        const double g = lm_gain(&ws);
        if (g <= 0.0) {
            if (lambda >= 1e7) break;
            lambda = fmin(lambda * 11, 1e7);
            m = rel_step(&ws);
            if (m < 1e-6 && it >= 2) break;
            continue;
        }
        for (int i = 0; i < np; ++i) ws.lbdp[i] = (lambda * ws.dP[i]) + ws.beta[i];
        m = 0.0;
        for (int i = 0; i < np; ++i) m = m + ws.dP[i] * ws.lbdp[i];
        const double errTry = lm_error(&ws);
        const double rho = (err - errTry) / (2.0 * m);
        if (rho > 1e-2) {
            for (int i = 0; i < np; ++i) ws.p[i] = ws.pTry[i];
            err = lm_pass(&ws);
            lambda = fmax(lambda / 9, 1e-7);
        } else {
            if (lambda >= 1e7) break;
            lambda = fmin(lambda * 11, 1e7);
        }
        m = rel_step(&ws);
        if (it >= 2) {
            if (m < 1e-6) break;
            if (err / (double)ws.n < 1e-5) break;
            if (max_abs(ws.beta, np) < 1e-6) break;
        }
    }
    {
        const double off = ws.p[0], inten = ws.p[1], x = ws.p[2], y = ws.p[3];
        const double rss = lm_r2(&ws);
        if (rss < OR_EPSILON) { fill_fail(mode, info, pt, x, y, info->z); goto done; }
        if (mode == 2) sx = ws.p[4];
        if (mode >= 3) { sx = ws.p[4] / OR_SQRT2; sy = ws.p[5] / OR_SQRT2; }
        if (mode == 4) th = ws.p[6];
        if (mode == 1) { sx = sy = ws.sigmaX; th = 0; }
        if (mode == 2) { sy = sx; th = 0; }
        fill_row(mode, info, pt, x, y, info->z, off, inten, sx, sy, th, rss, -1);
    }
done:
    free(heapbuf);
}

/* ------------------------------------------------------------------------------------------ */
/* Per-frame pipeline + stack orchestration                                                      */
/* (lib/localization/FrameProcessor.cpp:28-82, lib/Palm.cpp:24-172)                              */
/* ------------------------------------------------------------------------------------------ */

/* std::weak_order on doubles for the (Y, X) row sort (FrameProcessor.cpp:35-39): NaNs at the ends
 * by sign, +0 and -0 equivalent, otherwise numeric order. */
static int weak_cmp(double a, double b)
{
    const int an = a != a, bn = b != b;
    if (an || bn) {
        const int ra = an ? (signbit(a) ? -2 : 2) : 0, rb = bn ? (signbit(b) ? -2 : 2) : 0;
        if (an && bn) return (ra > rb) - (ra < rb);
        if (an) return ra < 0 ? -1 : 1;
        return rb < 0 ? 1 : -1;
    }
    return (a > b) - (a < b);
}

static int row_cmp(const void* pa, const void* pb)
{
    const double* a = (const double*)pa;
    const double* b = (const double*)pb;
    const int c = weak_cmp(a[5], b[5]);
    return c != 0 ? c : weak_cmp(a[4], b[4]);
}

static int fit_mode_of(uint64_t fit) { return (fit >= 1 && fit <= 5) ? (int)fit : 0; }

/* FrameProcessor::Process for one plane: rows (capacity-truncated in label order, then sorted). */
static size_t process_frame(const uint16_t* img, size_t w, size_t h, double thr, double ratio, int mode,
                            const double* fp, int maxIter, size_t maxPoints, double* rows, double* w1,
                            or_region* regions)
{
    or_atrous(img, w, h, NULL, w1, NULL);
    const size_t n = or_segment(w1, w, h, thr, ratio, NULL, regions, maxPoints);
    if (n > maxPoints) fprintf(stderr, "[WARNING] Density is too high; some points are missed (output capacity exceeded).\n");
    const size_t cnt = n < maxPoints ? n : maxPoints;          /* Fitting.cpp:44-58 */
    for (size_t i = 0; i < cnt; ++i) or_fit_region(mode, img, w, h, fp, maxIter, &regions[i], rows + i * OR_COLS);
    qsort(rows, cnt, sizeof(double) * OR_COLS, row_cmp);
    return cnt;
}

uint64_t or_localization(const uint16_t* input, double* output, uint64_t outputSize, uint64_t height, uint64_t width,
                         uint64_t planes, double thr, double ratio, uint64_t fit, const double* fp, int maxIter)
{
    const size_t maxPts = planes == 0 ? 0 : (outputSize / OR_COLS) / planes;   /* Palm.cpp:51 */
    if (maxPts == 0 || width == 0 || height == 0) return 0;                   /* Palm.cpp:57 */
    const size_t npx = (size_t)width * height;
    const int mode = fit_mode_of(fit);
    size_t* counts = (size_t*)calloc(planes, sizeof(size_t));
    /* per-plane row blocks: the regions of one plane never exceed its pixel count */
    const size_t rowCap = maxPts < npx ? maxPts : npx;
    double* rows = (double*)malloc(sizeof(double) * OR_COLS * rowCap * planes);
    #pragma omp parallel
    {
        double* w1 = (double*)malloc(sizeof(double) * npx);
        or_region* reg = (or_region*)malloc(sizeof(or_region) * (rowCap + 1));
        #pragma omp for schedule(dynamic, 1)
        for (long z = 0; z < (long)planes; ++z)
            counts[z] = process_frame(input + (size_t)z * npx, width, height, thr, ratio, mode, fp, maxIter, rowCap,
                                      rows + (size_t)z * rowCap * OR_COLS, w1, reg);
        free(w1);
        free(reg);
    }
    size_t count = 0, global = 1;                                              /* Palm.cpp:24-43 */
    for (size_t z = 0; z < planes; ++z) {
        for (size_t k = 0; k < counts[z]; ++k) {
            double* o = output + count;
            memcpy(o, rows + (z * rowCap + k) * OR_COLS, sizeof(double) * OR_COLS);
            o[0] = (double)global++;
            o[1] = (double)(z + 1);
            o[2] = (double)(k + 1);
            count += OR_COLS;
        }
    }
    free(rows);
    free(counts);
    return count;
}

uint64_t or_fitting(const uint16_t* input, const double* pos, uint64_t nPoints, double* output, uint64_t outputSize,
                    uint64_t height, uint64_t width, uint64_t planes, uint64_t fit, const double* fp, int maxIter)
{
    (void)outputSize;                                                           /* Palm.cpp:74 */
    if (nPoints == 0) return 0;
    if (fit < 1 || fit > 5) { fprintf(stderr, "oracle: Fitting supports modes 1-5\n"); return 0; }
    const size_t npx = (size_t)width * height;
    size_t* local = (size_t*)malloc(sizeof(size_t) * nPoints);
    size_t* pc = (size_t*)calloc(planes + 1, sizeof(size_t));
    for (size_t i = 0; i < nPoints; ++i) local[i] = ++pc[to_size(pos[i * 6 + 1]) - 1];  /* Palm.cpp:79-84 */
    #pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)nPoints; ++i) {
        const double* P = pos + (size_t)i * 6;
        const size_t plane = to_size(P[1]) - 1;
        or_region info = {P[2], P[3], 0.0, P[4], P[5]};
        double* o = output + (size_t)i * OR_COLS;
        or_fit_region((int)fit, input + plane * npx, width, height, fp, maxIter, &info, o);
        o[0] = P[0];
        o[1] = P[1];
        o[2] = (double)local[i];
    }
    free(pc);
    free(local);
    return nPoints * OR_COLS;
}

void or_wavelett(const uint16_t* input, double* output, uint64_t height, uint64_t width, uint64_t planes, uint64_t level)
{
    if (planes == 0) return;                                                    /* CPU/main.cpp:257-284 */
    if (level == 0) level = 1;
    const uint64_t lv = (level - 1) < 2 ? level - 1 : 2;
    const size_t npx = (size_t)width * height;
    #pragma omp parallel for schedule(static)
    for (long z = 0; z < (long)planes; ++z) {
        double* o = output + (size_t)z * npx;
        or_atrous(input + (size_t)z * npx, width, height, lv == 0 ? o : NULL, lv == 1 ? o : NULL, lv == 2 ? o : NULL);
    }
}

double or_autothreshold(const uint16_t* input, uint64_t height, uint64_t width, const double* fp)
{
    /* CPU/main.cpp:75-85 then Palm::AutoThreshold (Palm.cpp:122-172): ratio 10, fit none. */
    const size_t w = width, h = height, n = w * h;
    const uint64_t cap = (uint64_t)(0.2 * (double)(h * w) * OR_COLS);
    const size_t maxPts = (cap / OR_COLS) / 1;
    const double halfRoi = fp[0] / 2;
    double mu = 0.0;
    for (size_t i = 0; i < n; ++i) mu = mu + input[i];
    mu /= (double)n;
    double var = 0.0;
    for (size_t i = 0; i < n; ++i) var += (input[i] - mu) * (input[i] - mu);
    var /= (double)n;
    double thr = sqrt(var);
    const size_t rowCap = maxPts < n ? maxPts : n;
    double* rows = (double*)malloc(sizeof(double) * OR_COLS * (rowCap + 1));
    double* w1 = (double*)malloc(sizeof(double) * n);
    or_region* reg = (or_region*)malloc(sizeof(or_region) * (rowCap + 1));
    unsigned char* mask = (unsigned char*)malloc(n);
    for (int iter = 0; iter < OR_THRESHOLD_IT; ++iter) {
        const size_t cnt = maxPts == 0 ? 0 : process_frame(input, w, h, thr, 10.0, 0, fp, 75, rowCap, rows, w1, reg);
        memset(mask, 0, n);
        for (size_t k = 0; k < cnt; ++k) {
            const double X = rows[k * OR_COLS + 4], Y = rows[k * OR_COLS + 5];
            const int xm = (int)(X - halfRoi), ym = (int)(Y - halfRoi);
            const size_t xmin = (size_t)(xm > 0 ? xm : 0), ymin = (size_t)(ym > 0 ? ym : 0);
            const size_t xa = to_size(X + halfRoi), ya = to_size(Y + halfRoi);
            const size_t xmax = xa < w ? xa : w, ymax = ya < h ? ya : h;
            for (size_t y = ymin; y < ymax; ++y)
                for (size_t x = xmin; x < xmax; ++x) mask[y * w + x] = 1;
        }
        size_t count = 0;
        double sum = 0;
        for (size_t j = 0; j < n; ++j)
            if (!mask[j]) { sum += input[j]; ++count; }
        if (count == 0) break;
        mu = sum / (double)count;
        var = 0.0;
        for (size_t j = 0; j < n; ++j)
            if (!mask[j]) var += (input[j] - mu) * (input[j] - mu);
        var /= (double)count;
        var = sqrt(var);
        if (fabs(thr - var) <= DBL_EPSILON) break;
        thr = var;
    }
    free(mask);
    free(reg);
    free(w1);
    free(rows);
    return thr;
}

/* ------------------------------------------------------------------------------------------ */
/* Astigmatic z estimation (lib/fit/AstigFit.cpp:36-72 Estimate, :74-108 model / cost / score,   */
/* :340-409 brent; CPU/main.cpp:324-353 Astigmatism3DEstimation).                               */
/* ------------------------------------------------------------------------------------------ */

#define OR_AST_ZSTEP 201          /* AstigFit.hpp:83 default zStep */
#define OR_AST_BRENT_ITER 80      /* AstigFit.hpp:110 */
#define OR_AST_BRENT_TOL 1e-3     /* AstigFit.hpp:111 */
#define OR_AST_Z_TOL 1e-9         /* AstigFit.hpp:112 */
#define OR_AST_QUALITY_SCALE 0.2  /* AstigFit.hpp:114 */

/* std::min / std::max argument order (ties and NaN keep the first argument). */
static inline double std_min(double a, double b) { return b < a ? b : a; }
static inline double std_max(double a, double b) { return a < b ? b : a; }

/* getSigma (AstigFit.cpp:76-88): sigma in pixels of one axis at z; sampling is 1. */
static double ast_sigma(const double* c, double z, double pixelSize)
{
    const double u = (z - c[0]) / c[1];
    const double u2 = u * u, u3 = u2 * u, u4 = u3 * u;
    const double poly = 1.0 + u2 + (c[2] * u3) + (c[3] * u4);
    const double sigma = c[4] * sqrt(std_max(0.0, poly));
    return 1.0 / pixelSize * sigma;
}

/* getError (AstigFit.cpp:91-96): squared distance of the model sigmas at z to the observed pair. */
static double ast_error(const double* m, double z, double sx, double sy, double pixelSize)
{
    const double dx = ast_sigma(m, z, pixelSize) - sx, dy = ast_sigma(m + 5, z, pixelSize) - sy;
    return (dx * dx) + (dy * dy);
}

/* getQuality (AstigFit.cpp:99-108). */
static double ast_quality(const double* m, double z, double sx, double sy, double pixelSize)
{
    const double px = ast_sigma(m, z, pixelSize), py = ast_sigma(m + 5, z, pixelSize);
    if (!isfinite(px) || !isfinite(py) || px <= 0 || py <= 0 || !isfinite(sx) || !isfinite(sy) || sx <= 0 || sy <= 0) return 0;
    const double rx = (sx - px) / px / OR_AST_QUALITY_SCALE, ry = (sy - py) / py / OR_AST_QUALITY_SCALE;
    return 1.0 / (1.0 + (0.5 * ((rx * rx) + (ry * ry))));
}

/* brent (AstigFit.cpp:340-409): Brent's minimisation on [a, b]; returns x, writes f(x). */
static double ast_brent(double a, double b, const double* m, double sx, double sy, double pixelSize, double* fxOut)
{
    if (b < a) { const double t = a; a = b; b = t; }
    const double invPhi2 = 1.0 - 0.6180339887498948482;
    double x = a + (invPhi2 * (b - a));
    double w = x, v = x;
    double fx = ast_error(m, x, sx, sy, pixelSize);
    double fw = fx, fv = fx;
    double d = 0.0, e = 0.0;
    for (int iter = 0; iter < OR_AST_BRENT_ITER; ++iter) {
        const double mid = 0.5 * (a + b);
        const double tol1 = (OR_AST_BRENT_TOL * fabs(x)) + 1e-12, tol2 = 2.0 * tol1;
        if (fabs(x - mid) <= tol2 - (0.5 * (b - a))) break;
        int parabolic = 0;
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
                parabolic = 1;
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
        const double fu = ast_error(m, u, sx, sy, pixelSize);
        if (fu <= fx) {
            if (u < x) b = x;
            else a = x;
            v = w; fv = fw;
            w = x; fw = fx;
            x = u; fx = fu;
        } else {
            if (u < x) a = u;
            else b = u;
            if (fu <= fw || fabs(x - w) <= OR_AST_Z_TOL) {
                v = w; fv = fw;
                w = u; fw = fu;
            } else if (fu <= fv || fabs(v - x) <= OR_AST_Z_TOL || fabs(v - w) <= OR_AST_Z_TOL) {
                v = u; fv = fu;
            }
        }
    }
    *fxOut = fx;
    return x;
}

void or_astig_estimate(const double* input, double* output, uint64_t size, double pixelSize, const double* model,
                       double zMax)
{
    const double z0 = -zMax, z1 = zMax, deltaZ = z1 - z0;
    for (uint64_t n = 0; n < size; ++n) {
        const double sx = input[2 * n], sy = input[2 * n + 1];
        double bestZ = z0, bestE = INFINITY;
        for (int i = 0; i < OR_AST_ZSTEP; ++i) {
            const double t = (double)i / (double)(OR_AST_ZSTEP - 1);
            const double z = z0 + (t * deltaZ);
            const double e = ast_error(model, z, sx, sy, pixelSize);
            if (e < bestE) { bestE = e; bestZ = z; }
        }
        const double dz = deltaZ / (double)(OR_AST_ZSTEP - 1);
        const double a = std_max(z0, std_min(bestZ - (2.0 * dz), z1));
        const double b = std_max(z0, std_min(bestZ + (2.0 * dz), z1));
        double fr = 0.0;
        const double zr = ast_brent(a, b, model, sx, sy, pixelSize, &fr);
        const double z = fr < bestE ? zr : bestZ;
        output[2 * n] = z;
        output[2 * n + 1] = ast_quality(model, z, sx, sy, pixelSize);
    }
}

double or_astig_quality(const double* model, double z, double sx, double sy, double pixelSize)
{
    return ast_quality(model, z, sx, sy, pixelSize);
}

void or_libm(int fn, const double* x, uint64_t n, double* y)
{
    for (uint64_t i = 0; i < n; ++i) {
        if (fn == 0) {
            y[i] = exp(x[i]);
        } else {
            double s, c;
            sincos(x[i], &s, &c);
            y[i] = fn == 1 ? s : c;
        }
    }
}
