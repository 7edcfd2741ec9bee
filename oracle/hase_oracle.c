/*
 * hase_oracle.c - FP64 CPU restatement of the importance-sampled ASE ray
 * tracer.  TEST INFRASTRUCTURE ONLY (see hase_oracle.h): the checker of the
 * HIP library, never part of the product.
 *
 * Estimator (the retired legacy calcPhiAse, restated from the surviving
 * physics; the point flux at a sample point p is
 *     phi(p) = (nTot/tau) * integral beta(x) G(x->p) R(x->p) / (4 pi |x'-p|^2) dV
 * with G = exp(sum_j g_j l_j) along the straight (or mirror-unfolded) path,
 * g_j = nTot (beta_j (sigma_e + sigma_a) - sigma_a) in gain prisms and
 * -claddingAbsorption in cladding prisms (rayWalk.hpp:170-192), R the product
 * of surface factors of reflected paths (reflection.hpp:21-47), and the
 * nTot/tau scaling of docs/source/theoryAndModel.rst:155-163; the sphere
 * closed form of tests/python/phiAse/test_analyticalSphere.py:19-27 is the
 * same quantity):
 *   1. importance of every (reflection slice k, prism j): the integrand at the
 *      prism centroid times the prism volume (spectrally averaged gain);
 *   2. proposal q = (1-eps) imp/sum(imp) + eps beta V/(R sum(beta V));
 *   3. rays mapped to (k, j) by systematic sampling on the FP64 prefix of q
 *      (one uniform offset per sample and launch), start point uniform in the
 *      prism, walk to the (image of the) sample point;
 *   4. per ray X = mean_lambda(G_lambda) R beta_j V_j / (4 pi D^2 q), FP64 sums
 *      of X and X^2; relative standard error as calcForwardPhiAse.cpp:81-98,
 *      mse = its square (the relative variance of the mean);
 *   5. adaptive schedule of types.hpp:333-354 with pooled launches.
 */
#include "hase_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Ray geometry type of the walk.  The oracle proper is FP64 throughout; the
 * build with ORC_GEOMETRY_FP32 (oracle/build/libhase_oracle_f32.so) walks in
 * FP32 with FP64 sums, the CPU restatement the benchmark's CPU baseline is
 * timed on. */
#ifdef ORC_GEOMETRY_FP32
typedef float orc_g;
#define ORC_SQRT sqrtf
#else
typedef double orc_g;
#define ORC_SQRT sqrt
#endif

#define ORC_MAX_STEPS 10000u
#define ORC_KEY_TAG 0x48415345u /* key word 1 of every stream */
#define ORC_TAG_RAY 0u
#define ORC_TAG_OFFSET 1u

static const double ORC_PI = 3.14159265358979323846;

/* ---------------------------------------------------------------- Philox */
/* Philox4x32-10 as specified by Salmon et al. (SC'11): two 32x32->64
 * multiplications per round, Weyl key schedule, ten rounds. */
static inline void orc_mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint32_t hi0, lo0, hi1, lo1;
        orc_mulhilo(0xD2511F53u, c0, &hi0, &lo0);
        orc_mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* (2k+1) * 2^-24 with k the top 23 bits: strictly inside (0,1) and exact in
 * FP32 and FP64, so the HIP kernels see the same numbers. */
double orc_u32_to_unit(uint32_t x)
{
    return (double)(((x >> 9) << 1) | 1u) * (1.0 / 16777216.0);
}

/* ------------------------------------------------------------- geometry */
static inline uint32_t orc_real_level(int64_t layer, uint32_t layers)
{
    int64_t n = (int64_t)layers;
    int64_t c = layer >= 0 ? layer / n : -((-layer + n - 1) / n);
    int64_t r = layer - c * n;
    return (uint32_t)((c & 1) ? (n - 1 - r) : r);
}

static inline int64_t orc_copy_of(int64_t layer, uint32_t layers)
{
    int64_t n = (int64_t)layers;
    return layer >= 0 ? layer / n : -((-layer + n - 1) / n);
}

/* Mirror image of p after `reflections` alternating reflections at the
 * planes z = 0 (bottom, 0) and z = Z = (L-1) h (top, 1), the first at
 * `first_plane`.  The legacy plane height is fixed at (L-1) h
 * (tests/data/laserPumpCladding/fixed_legacy_reflection_pump3_wedge_reference/README.md). */
void orc_image_point(const orc_mesh* m, const double p[3], uint32_t reflections,
                     uint32_t first_plane, double out[3])
{
    double Z = (double)(m->nl - 1) * m->h;
    double mr = (double)reflections;
    out[0] = p[0];
    out[1] = p[1];
    if (reflections == 0) {
        out[2] = p[2];
    } else if (first_plane == 1) {
        out[2] = (reflections & 1) ? mr * Z + (Z - p[2]) : mr * Z + p[2];
    } else {
        out[2] = (reflections & 1) ? -(mr - 1.0) * Z - p[2] : -mr * Z + p[2];
    }
}

static double orc_surface_factor(const orc_mesh* m, uint32_t t, uint32_t plane, double sin2)
{
    if (m->nidx) {
        double nin = m->nidx[2 * plane], nout = m->nidx[2 * plane + 1];
        if (nin > 0.0 && nout > 0.0) {
            double ratio = nin / nout;
            if (ratio * ratio * sin2 > 1.0) return 1.0;
        }
    }
    if (!m->refl) return 0.0;
    double r = m->refl[2 * (size_t)t + plane];
    return r > 0.0 ? r : 0.0;
}

/* Prism walk from a (inside prism `prism`) to the unfolded end point e.
 * Exit edge: smallest crossing parameter among edges whose outward normal
 * faces the ray (the entry edge excluded); exit plane: the next level plane
 * of the unfolded layer; parameters are measured from a, not accumulated. */
int orc_walk_ray(const orc_mesh* m, const double a[3], uint32_t prism, const double e[3],
                 uint32_t reflections, uint32_t first_plane, orc_walk* w)
{
    memset(w, 0, sizeof(*w));
    w->refl = 1.0;
    const uint32_t T = m->nt;
    const uint32_t layers = m->nl - 1;
    if (T == 0 || layers == 0 || prism >= T * layers) {
        w->status = 3;
        return 3;
    }
    uint32_t t = prism % T;
    int64_t u = (int64_t)(prism / T);
    const orc_g h = (orc_g)m->h;
    /* heights in layer units; an end point on a level plane sits on it exactly */
    const orc_g ax = (orc_g)a[0], ay = (orc_g)a[1], ex0 = (orc_g)e[0], ey0 = (orc_g)e[1];
    orc_g za = (orc_g)a[2] / h, ze = (orc_g)e[2] / h, zr = (orc_g)nearbyint((double)ze);
    if (fabs((double)(ze - zr)) < 1e-5) ze = zr;
    orc_g dx = ex0 - ax, dy = ey0 - ay, dzeta = ze - za, dz = dzeta * h;
    orc_g D = ORC_SQRT(dx * dx + dy * dy + dz * dz);
    w->length = D;
    if (!(D > 0.0)) return 0;
    orc_g vx = dx / D, vy = dy / D, vz = dz / D;
    orc_g sin2 = 1 - vz * vz;
    if (sin2 < 0) sin2 = 0;
    const orc_g tol = (orc_g)(sizeof(orc_g) == 4 ? 2e-5 : 1e-10) * (1 + D);
    int entry = -1;
    orc_g s = 0;
    uint32_t applied = 0;
    for (uint32_t step = 0; step < ORC_MAX_STEPS; ++step) {
        w->steps++;
        orc_g best = INFINITY;
        int bestk = -1;
        for (int k = 0; k < 3; ++k) {
            if (k == entry) continue;
            uint32_t i0 = m->tri[t + (size_t)k * T];
            uint32_t i1 = m->tri[t + (size_t)((k + 1) % 3) * T];
            const orc_g x0 = (orc_g)m->px[i0], y0 = (orc_g)m->py[i0];
            const orc_g x1 = (orc_g)m->px[i1], y1 = (orc_g)m->py[i1];
            orc_g ex = x1 - x0, ey = y1 - y0;
            orc_g nx = ey, ny = -ex;
            orc_g nv = nx * vx + ny * vy;
            if (!(nv > 0.0)) continue;
            /* an edge through the end point's (x, y) meets the ray exactly at the end */
            int touches = (x0 == ex0 && y0 == ey0) || (x1 == ex0 && y1 == ey0);
            orc_g sk = touches ? D : (nx * (x0 - ax) + ny * (y0 - ay)) / nv;
            if (sk < best) { best = sk; bestk = k; }
        }
        orc_g sz = INFINITY;
        int dir = 0;
        if (dzeta > 0) { sz = D * (((orc_g)(u + 1) - za) / dzeta); dir = 1; }
        else if (dzeta < 0) { sz = D * (((orc_g)u - za) / dzeta); dir = -1; }
        orc_g snext = best < sz ? best : sz;
        if (D < snext) snext = D;
        if (snext < s) snext = s;
        orc_g seg = snext - s;
        uint32_t level = orc_real_level(u, layers);
        int cladding = m->clad && m->clad[t] == m->clad_number;
        if (cladding) {
            w->clad_len += seg;
        } else {
            w->beta_len += m->beta[t + (size_t)level * T] * seg;
            w->gain_len += seg;
        }
        s = snext;
        int done = s >= D;
        if (!done && sz > best && m->nbr[t + (size_t)bestk * T] < 0) {
            /* leaving the mesh sideways: only rounding short of the end counts */
            if (D - s <= tol) done = 1;
            else { w->status = 1; return 1; }
        }
        if (done) {
            /* an end point on a reflecting plane takes its last factor here */
            while (applied < reflections) {
                uint32_t plane = (applied & 1) ? 1u - first_plane : first_plane;
                w->refl *= orc_surface_factor(m, t, plane, sin2);
                applied++;
            }
            return 0;
        }
        if (sz <= best) {
            int64_t nu = u + dir;
            int64_t c0 = orc_copy_of(u, layers), c1 = orc_copy_of(nu, layers);
            if (c0 != c1 && applied < reflections) {
                uint32_t plane;
                if (dir > 0) plane = (c0 & 1) ? 0u : 1u;
                else plane = (c0 & 1) ? 1u : 0u;
                w->refl *= orc_surface_factor(m, t, plane, sin2);
                applied++;
            }
            u = nu;
        } else {
            entry = m->fe[t + (size_t)bestk * T];
            t = (uint32_t)m->nbr[t + (size_t)bestk * T];
        }
    }
    w->status = 2;
    return 2;
}

/* ------------------------------------------------------------ schedule */
uint32_t orc_adaptive_target(uint32_t min_rays, uint32_t max_rays, uint32_t steps, uint32_t k)
{
    if (max_rays <= min_rays || steps == 0) return min_rays;
    if (k >= steps) return max_rays;
    double growth = pow((double)max_rays / (double)min_rays, 1.0 / (double)steps);
    double target = (double)min_rays * pow(growth, (double)k);
    double r = ceil(target);
    if (r < (double)min_rays) r = (double)min_rays;
    if (r > (double)max_rays) r = (double)max_rays;
    return (uint32_t)r;
}

double orc_relative_standard_error(double s1, double s2, double n)
{
    if (n < 2.0 || !isfinite(s1) || !isfinite(s2)) return DBL_MAX;
    if (s1 == 0.0) return NAN;
    double rv = (n * s2 / (s1 * s1) - 1.0) / n;
    return sqrt(rv > 0.0 ? rv : 0.0);
}

/* -------------------------------------------------------- importance */
static inline double orc_tri_area(const orc_mesh* m, uint32_t t)
{
    uint32_t T = m->nt;
    uint32_t i0 = m->tri[t], i1 = m->tri[t + T], i2 = m->tri[t + 2 * (size_t)T];
    double cr = (m->px[i1] - m->px[i0]) * (m->py[i2] - m->py[i0])
              - (m->py[i1] - m->py[i0]) * (m->px[i2] - m->px[i0]);
    return 0.5 * fabs(cr);
}

static inline uint32_t orc_slices(const orc_params* prm)
{
    uint32_t K = prm->use_reflections ? (prm->max_reflections ? prm->max_reflections : 1u) : 0u;
    return 1u + 2u * K;
}

static inline void orc_slice(uint32_t k, uint32_t* refl, uint32_t* first)
{
    *refl = (k + 1u) / 2u;
    *first = (k & 1u) ? 1u : 0u;
}

/* gain of a walked path averaged over the spectral samples: the gain
 * exponent is linear in the walk's three length sums, so one walk serves every
 * wavelength (g = nTot (beta (sigma_e + sigma_a) - sigma_a), rayWalk.hpp:170-182) */
static double orc_spectral_gain(const orc_mesh* m, const orc_params* prm, const orc_walk* w)
{
    double g = 0.0;
    for (uint32_t lam = 0; lam < prm->n_lambda; ++lam) {
        double c1 = prm->n_tot * (prm->sigma_e[lam] + prm->sigma_a[lam]);
        double c2 = prm->n_tot * prm->sigma_a[lam];
        g += exp(c1 * w->beta_len - c2 * w->gain_len - m->clad_alpha * w->clad_len);
    }
    return g / (double)prm->n_lambda;
}

static inline void orc_sample_point(const orc_mesh* m, uint32_t sample, double p[3])
{
    uint32_t pt = sample % m->np, lv = sample / m->np;
    p[0] = m->px[pt];
    p[1] = m->py[pt];
    p[2] = (double)lv * m->h;
}

int orc_importance(const orc_mesh* m, const orc_params* prm, uint32_t sample, double* imp)
{
    const uint32_t T = m->nt, P = T * (m->nl - 1), R = orc_slices(prm);
    double p[3];
    orc_sample_point(m, sample, p);
    for (uint32_t k = 0; k < R; ++k) {
        uint32_t nr, fp;
        orc_slice(k, &nr, &fp);
        double e[3];
        orc_image_point(m, p, nr, fp, e);
        for (uint32_t j = 0; j < P; ++j) {
            uint32_t t = j % T, l = j / T;
            double b = m->beta[j];
            double value = 0.0;
            if (b > 0.0) {
                uint32_t i0 = m->tri[t], i1 = m->tri[t + T], i2 = m->tri[t + 2 * (size_t)T];
                double c[3] = {(m->px[i0] + m->px[i1] + m->px[i2]) / 3.0,
                               (m->py[i0] + m->py[i1] + m->py[i2]) / 3.0,
                               ((double)l + 0.5) * m->h};
                orc_walk w;
                if (orc_walk_ray(m, c, j, e, nr, fp, &w) == 0 && w.length > 0.0) {
                    double V = orc_tri_area(m, t) * m->h;
                    value = b * V * orc_spectral_gain(m, prm, &w) * w.refl
                          / (4.0 * ORC_PI * w.length * w.length);
                    if (!isfinite(value) || value < 0.0) value = 0.0;
                }
            }
            imp[(size_t)k * P + j] = value;
        }
    }
    return 0;
}

/* Index of (x, y) on a 2^16 x 2^16 Hilbert curve. */
static uint64_t orc_hilbert_index(uint32_t x, uint32_t y)
{
    const uint32_t n = 1u << 16;
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) { x = n - 1 - x; y = n - 1 - y; }
            uint32_t t = x; x = y; y = t;
        }
    }
    return d;
}

typedef struct { uint64_t key; uint32_t index; } orc_keyed;

static int orc_keyed_cmp(const void* a, const void* b)
{
    const orc_keyed* x = (const orc_keyed*)a;
    const orc_keyed* y = (const orc_keyed*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->index < y->index ? -1 : (x->index > y->index);
}

static uint32_t orc_quant(double v, double lo, double hi)
{
    if (!(hi > lo)) return 0u;
    double q = floor((v - lo) / (hi - lo) * 65535.0);
    if (q < 0.0) q = 0.0;
    if (q > 65535.0) q = 65535.0;
    return (uint32_t)q;
}

/* Order in which the proposal prefix visits the triangles of a level:
 * ascending Hilbert index of the centroid (16 bits per axis over the
 * centroids' bounding box), ties by index. */
void orc_triangle_order(const orc_mesh* m, uint32_t* order)
{
    const uint32_t T = m->nt;
    double* cx = (double*)malloc(T * sizeof(double));
    double* cy = (double*)malloc(T * sizeof(double));
    orc_keyed* k = (orc_keyed*)malloc(T * sizeof(orc_keyed));
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (uint32_t t = 0; t < T; ++t) {
        uint32_t a = m->tri[t], b = m->tri[t + T], c = m->tri[t + 2 * (size_t)T];
        cx[t] = (m->px[a] + m->px[b] + m->px[c]) / 3.0;
        cy[t] = (m->py[a] + m->py[b] + m->py[c]) / 3.0;
        if (cx[t] < x0) x0 = cx[t];
        if (cx[t] > x1) x1 = cx[t];
        if (cy[t] < y0) y0 = cy[t];
        if (cy[t] > y1) y1 = cy[t];
    }
    for (uint32_t t = 0; t < T; ++t) {
        k[t].key = orc_hilbert_index(orc_quant(cx[t], x0, x1), orc_quant(cy[t], y0, y1));
        k[t].index = t;
    }
    qsort(k, T, sizeof(orc_keyed), orc_keyed_cmp);
    for (uint32_t t = 0; t < T; ++t) order[t] = k[t].index;
    free(cx);
    free(cy);
    free(k);
}

/* Inclusive FP64 prefix of the defensive mixture proposal over entries
 * (slice k, level l, triangle order[i]) in that nesting. */
int orc_proposal_prefix(const orc_mesh* m, const orc_params* prm, const double* imp,
                        const uint32_t* order, double* prefix)
{
    const uint32_t T = m->nt, P = T * (m->nl - 1), R = orc_slices(prm);
    double I = 0.0, B = 0.0;
    for (size_t e = 0; e < (size_t)R * P; ++e) I += imp[e];
    for (uint32_t j = 0; j < P; ++j) {
        double b = m->beta[j];
        if (b > 0.0) B += b * orc_tri_area(m, j % T) * m->h;
    }
    if (!(B > 0.0)) return 1;
    double eps = prm->defensive;
    int use_imp = I > 0.0 && isfinite(I);
    double acc = 0.0;
    for (uint32_t k = 0; k < R; ++k) {
        for (uint32_t i = 0; i < P; ++i) {
            uint32_t j = order[i % T] + (i / T) * T;
            double b = m->beta[j];
            double bv = b > 0.0 ? b * orc_tri_area(m, j % T) * m->h : 0.0;
            double q = use_imp ? (1.0 - eps) * imp[(size_t)k * P + j] / I + eps * bv / (B * (double)R)
                               : bv / (B * (double)R);
            acc += q;
            prefix[(size_t)k * P + i] = acc;
        }
    }
    return 0;
}

void orc_ray_start(const orc_mesh* m, uint32_t seed, uint32_t sample, uint32_t launch,
                   uint32_t ray, uint32_t prism, double out[3])
{
    const uint32_t T = m->nt;
    uint32_t ctr[4] = {ray, sample, launch, ORC_TAG_RAY};
    uint32_t key[2] = {seed, ORC_KEY_TAG};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    double u1 = orc_u32_to_unit(r[0]), u2 = orc_u32_to_unit(r[1]), u3 = orc_u32_to_unit(r[2]);
    if (u1 + u2 > 1.0) { u1 = 1.0 - u1; u2 = 1.0 - u2; }
    uint32_t t = prism % T, l = prism / T;
    uint32_t i0 = m->tri[t], i1 = m->tri[t + T], i2 = m->tri[t + 2 * (size_t)T];
    out[0] = m->px[i0] + u1 * (m->px[i1] - m->px[i0]) + u2 * (m->px[i2] - m->px[i0]);
    out[1] = m->py[i0] + u1 * (m->py[i1] - m->py[i0]) + u2 * (m->py[i2] - m->py[i0]);
    out[2] = ((double)l + u3) * m->h;
}

double orc_systematic_offset(uint32_t seed, uint32_t sample, uint32_t launch)
{
    uint32_t ctr[4] = {0xFFFFFFFFu, sample, launch, ORC_TAG_OFFSET};
    uint32_t key[2] = {seed, ORC_KEY_TAG};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    return ((double)(r[0] >> 5) * 67108864.0 + (double)(r[1] >> 6)) * (1.0 / 9007199254740992.0);
}

/* first index e with prefix[e] > target */
static size_t orc_upper_bound(const double* prefix, size_t n, double target)
{
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (prefix[mid] > target) hi = mid; else lo = mid + 1;
    }
    return lo < n ? lo : n - 1;
}

static void orc_trace_launch(const orc_mesh* m, const orc_params* prm, uint32_t sample,
                             uint32_t launch, uint32_t nrays, const double* prefix,
                             const uint32_t* order, double* s1, double* s2, uint32_t* dropped)
{
    const uint32_t T = m->nt, P = T * (m->nl - 1), R = orc_slices(prm);
    const size_t n = (size_t)R * P;
    const double Q = prefix[n - 1];
    const double U = orc_systematic_offset(prm->seed, sample, launch);
    double p[3];
    orc_sample_point(m, sample, p);
    double acc1 = 0.0, acc2 = 0.0;
    uint32_t drop = 0;
    for (uint32_t r = 0; r < nrays; ++r) {
        double target = ((double)r + U) / (double)nrays * Q;
        size_t ent = orc_upper_bound(prefix, n, target);
        double q = prefix[ent] - (ent ? prefix[ent - 1] : 0.0);
        uint32_t k = (uint32_t)(ent / P), i = (uint32_t)(ent % P);
        uint32_t j = order[i % T] + (i / T) * T;
        uint32_t nr, fp;
        orc_slice(k, &nr, &fp);
        double e[3], a[3];
        orc_image_point(m, p, nr, fp, e);
        orc_ray_start(m, prm->seed, sample, launch, r, j, a);
        orc_walk w;
        if (orc_walk_ray(m, a, j, e, nr, fp, &w) != 0 || !(w.length > 0.0)) {
            drop++;
            continue;
        }
        double g = orc_spectral_gain(m, prm, &w);
        double V = orc_tri_area(m, j % T) * m->h;
        double X = g * w.refl * m->beta[j] * V * Q / (4.0 * ORC_PI * w.length * w.length * q);
        if (!isfinite(X)) {
            drop++;
            continue;
        }
        acc1 += X;
        acc2 += X * X;
    }
    *s1 += acc1;
    *s2 += acc2;
    *dropped += drop;
}

int orc_calc_phi_ase(const orc_mesh* m, const orc_params* prm, double* phi, double* mse,
                     uint32_t* total_rays, double* dndt, uint32_t* dropped, double* s1_out,
                     double* s2_out)
{
    const uint32_t T = m->nt, P = T * (m->nl - 1), R = orc_slices(prm);
    const uint32_t S = m->np * m->nl;
    if (prm->sample_max >= S || prm->sample_min > prm->sample_max) return 3;
    if (prm->min_rays == 0 || prm->max_rays < prm->min_rays) return 3;
    const size_t n = (size_t)R * P;
    const double scale = prm->n_tot / prm->tau;
    for (uint32_t s = 0; s < S; ++s) {
        phi[s] = 0.0; mse[s] = 0.0; total_rays[s] = 0; dndt[s] = 0.0;
        if (dropped) dropped[s] = 0;
        if (s1_out) s1_out[s] = 0.0;
        if (s2_out) s2_out[s] = 0.0;
    }
    const uint32_t reps = prm->max_repetitions ? prm->max_repetitions : 1u;
#ifdef _OPENMP
    int nthreads = prm->threads > 0 ? prm->threads : omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double* imp = (double*)malloc(n * sizeof(double));
        double* prefix = (double*)malloc(n * sizeof(double));
        uint32_t* order = (uint32_t*)malloc(T * sizeof(uint32_t));
        orc_triangle_order(m, order);
        const int64_t count = prm->sample_list ? (int64_t)prm->n_list
                                               : (int64_t)prm->sample_max - (int64_t)prm->sample_min + 1;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t si = 0; si < count; ++si) {
            uint32_t s = prm->sample_list ? prm->sample_list[si] : prm->sample_min + (uint32_t)si;
            if (s < prm->sample_min || s > prm->sample_max) continue;
            if (prm->importance) {
                const double* src = prm->importance + (size_t)(s - prm->sample_min) * n;
                for (size_t e = 0; e < n; ++e) imp[e] = src[e];
            } else {
                orc_importance(m, prm, s, imp);
            }
            if (orc_proposal_prefix(m, prm, imp, order, prefix) != 0) continue; /* no source */
            double s1 = 0.0, s2 = 0.0;
            uint32_t N = 0, drop = 0, launch = 0, k = 0, extra = 0;
            double rse = DBL_MAX;
            for (;;) {
                uint32_t target = orc_adaptive_target(prm->min_rays, prm->max_rays,
                                                      prm->adaptive_steps, k);
                uint32_t nr;
                if (target > N) nr = target - N;
                else if (N >= prm->max_rays || target >= prm->max_rays) {
                    if (extra + 1 >= reps) break;
                    extra++;
                    nr = prm->max_rays;
                } else if (k >= prm->adaptive_steps) break;   /* flat schedule: nothing to add */
                else { k++; continue; }
                orc_trace_launch(m, prm, s, launch, nr, prefix, order, &s1, &s2, &drop);
                N += nr;
                launch++;
                rse = orc_relative_standard_error(s1, s2, (double)N);
                /* mse: squared relative standard error (relative variance of the mean) */
                if (isfinite(rse) && rse < DBL_MAX) rse = rse * rse;
                int ok = drop == 0 && isfinite(rse) && rse <= prm->mse_threshold;
                if (ok) break;
                if (target < prm->max_rays) k++;
                if (prm->importance && prm->fallback && launch == 1) {
                    /* external proposal missed the threshold: restart on the walked map */
                    orc_importance(m, prm, s, imp);
                    if (orc_proposal_prefix(m, prm, imp, order, prefix) != 0) break;
                    s1 = 0.0; s2 = 0.0; N = 0; drop = 0; k = 0; extra = 0;
                    rse = DBL_MAX;
                }
            }
            double ph = scale * s1 / (double)N;
            phi[s] = ph;
            mse[s] = rse;
            total_rays[s] = N;
            double bs = prm->beta_cells ? prm->beta_cells[s] : 0.0;
            dndt[s] = (bs * (prm->max_sigma_e + prm->max_sigma_a) - prm->max_sigma_a) * ph;
            if (dropped) dropped[s] = drop;
            if (s1_out) s1_out[s] = s1;
            if (s2_out) s2_out[s] = s2;
        }
        free(imp);
        free(prefix);
        free(order);
    }
    return 0;
}
