/*
 * forward_oracle.c - FP64 CPU restatement of the reference's forward Tet4
 * estimator.  TEST INFRASTRUCTURE ONLY (see forward_oracle.h for the
 * reference files followed and the choices made here).
 */
#include "forward_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u32_to_unit(uint32_t x);
uint32_t orc_adaptive_target(uint32_t min_rays, uint32_t max_rays, uint32_t steps, uint32_t k);

#define ORF_KEY_TAG 0x46574431u /* key word 1 of the forward streams */
#define ORF_TOL (64.0 * DBL_EPSILON) /* rayWalk.hpp:17 */
#define ORF_PROBE 1.0e-12            /* rayTransition.hpp:18 */

/* ------------------------------------------------------------ mesh set-up */
typedef struct {
    uint32_t v[3];
    uint32_t slot; /* cell * 4 + local face */
} orf_face_key;

static int orf_face_cmp(const void* a, const void* b)
{
    const orf_face_key* x = (const orf_face_key*)a;
    const orf_face_key* y = (const orf_face_key*)b;
    for (int i = 0; i < 3; ++i)
        if (x->v[i] != y->v[i]) return x->v[i] < y->v[i] ? -1 : 1;
    return x->slot < y->slot ? -1 : (x->slot > y->slot);
}

static void orf_sort3(uint32_t v[3])
{
    uint32_t t;
    if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
    if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
    if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
}

/* Face k of a cell is the face opposite its vertex k; its plane is the
 * affine barycentric coordinate of that vertex (mesh.hpp:657-725): zero on
 * the face, one at the vertex. */
int orf_prepare(orf_mesh* m)
{
    const uint32_t n = m->n_cells;
    orf_face_key* keys = (orf_face_key*)malloc((size_t)n * 4 * sizeof(orf_face_key));
    if (!keys) return -3;
    memset(m->lumped, 0, 2 * (size_t)m->n_points * sizeof(double));
    for (uint32_t c = 0; c < n; ++c) {
        const uint32_t* cv = m->cells + 4 * (size_t)c;
        for (uint32_t k = 0; k < 4; ++k) {
            uint32_t f[3], j = 0;
            for (uint32_t q = 0; q < 4; ++q)
                if (q != k) f[j++] = cv[q];
            orf_sort3(f);
            const uint32_t o = cv[k];
            const double ax = m->px[f[0]], ay = m->py[f[0]], az = m->pz[f[0]];
            const double bx = m->px[f[1]] - ax, by = m->py[f[1]] - ay, bz = m->pz[f[1]] - az;
            const double cx = m->px[f[2]] - ax, cy = m->py[f[2]] - ay, cz = m->pz[f[2]] - az;
            const double nx = by * cz - bz * cy, ny = bz * cx - bx * cz, nz = bx * cy - by * cx;
            const double den = nx * (m->px[o] - ax) + ny * (m->py[o] - ay) + nz * (m->pz[o] - az);
            if (fabs(den) <= DBL_EPSILON) { free(keys); return -1; }
            double* pl = m->planes + 16 * (size_t)c + 4 * k;
            const double s = 1.0 / den;
            pl[0] = nx * s;
            pl[1] = ny * s;
            pl[2] = nz * s;
            pl[3] = -(pl[0] * ax + pl[1] * ay + pl[2] * az);
            keys[4 * (size_t)c + k].v[0] = f[0];
            keys[4 * (size_t)c + k].v[1] = f[1];
            keys[4 * (size_t)c + k].v[2] = f[2];
            keys[4 * (size_t)c + k].slot = 4 * c + k;
            m->nbr[4 * (size_t)c + k] = -1;
            m->nbr_face[4 * (size_t)c + k] = -1;
        }
        const double ax = m->px[cv[0]], ay = m->py[cv[0]], az = m->pz[cv[0]];
        const double bx = m->px[cv[1]] - ax, by = m->py[cv[1]] - ay, bz = m->pz[cv[1]] - az;
        const double cx = m->px[cv[2]] - ax, cy = m->py[cv[2]] - ay, cz = m->pz[cv[2]] - az;
        const double dx = m->px[cv[3]] - ax, dy = m->py[cv[3]] - ay, dz = m->pz[cv[3]] - az;
        const double vol = fabs((by * cz - bz * cy) * dx + (bz * cx - bx * cz) * dy + (bx * cy - by * cx) * dz) / 6.0;
        m->volume[c] = vol;
        const uint32_t off = (m->cell_types && m->cell_types[c] == m->cladding_number) ? m->n_points : 0u;
        for (uint32_t k = 0; k < 4; ++k) m->lumped[off + cv[k]] += vol / 4.0;
    }
    qsort(keys, (size_t)n * 4, sizeof(orf_face_key), orf_face_cmp);
    for (size_t i = 0; i + 1 < (size_t)n * 4; ++i) {
        if (memcmp(keys[i].v, keys[i + 1].v, sizeof(keys[i].v)) != 0) continue;
        if (i + 2 < (size_t)n * 4 && memcmp(keys[i].v, keys[i + 2].v, sizeof(keys[i].v)) == 0) {
            free(keys);
            return -2; /* a face shared by more than two cells */
        }
        const uint32_t a = keys[i].slot, b = keys[i + 1].slot;
        m->nbr[a] = (int32_t)(b / 4);
        m->nbr_face[a] = (int32_t)(b % 4);
        m->nbr[b] = (int32_t)(a / 4);
        m->nbr_face[b] = (int32_t)(a % 4);
        ++i;
    }
    free(keys);
    orf_set_beta(m, m->beta);
    return 0;
}

void orf_set_beta(orf_mesh* m, const double* beta)
{
    m->beta = beta;
    double run = 0.0;
    for (uint32_t c = 0; c < m->n_cells; ++c) {
        run += beta[c] * m->volume[c];
        m->bv_prefix[c] = run;
    }
}

/* --------------------------------------------------------------- walking */
static inline double orf_lambda(const double* pl, const double p[3])
{
    return pl[0] * p[0] + pl[1] * p[1] + pl[2] * p[2] + pl[3];
}

static inline double orf_dlambda(const double* pl, const double u[3])
{
    return pl[0] * u[0] + pl[1] * u[1] + pl[2] * u[2];
}

/* Nearest exit face ahead of p among the faces other than `forbidden`
 * (rayWalk.hpp:106-163): faces whose barycentric coordinate falls along u;
 * faces that reach zero within the tolerance at the same distance are tied,
 * and the lowest tied index is taken. */
orf_hit orf_next_face(const orf_mesh* m, uint32_t cell, const double p[3], const double u[3], int32_t forbidden)
{
    orf_hit h = {-1, DBL_MAX, 0u};
    double cand[4] = {0.0, 0.0, 0.0, 0.0};
    const double* pl = m->planes + 16 * (size_t)cell;
    for (int k = 0; k < 4; ++k) {
        if (k == forbidden) continue;
        const double c = orf_lambda(pl + 4 * k, p), d = orf_dlambda(pl + 4 * k, u);
        if (d >= 0.0) continue;
        const double len = -c / d;
        cand[k] = (len > 0.0 && len <= DBL_MAX) ? len : 0.0;
        if (cand[k] > 0.0 && cand[k] < h.length) {
            h.length = cand[k];
            h.face = k;
        }
    }
    if (h.face < 0) return h;
    for (int k = 0; k < 4; ++k) {
        if (k == forbidden || cand[k] <= 0.0) continue;
        const double c = orf_lambda(pl + 4 * k, p), d = orf_dlambda(pl + 4 * k, u);
        if (fabs(c + h.length * d) <= ORF_TOL) h.tied_mask |= 1u << k;
    }
    h.tied_mask |= 1u << h.face;
    for (int k = 0; k < 4; ++k)
        if (h.tied_mask & (1u << k)) { h.face = k; break; }
    return h;
}

/* Crossing at a point where several faces meet (rayTransition.hpp:74-160):
 * step through neighbours until a cell owns the ray (no face whose
 * coordinate, probed a little along u, turns negative), preferring exits
 * into cells not among the last three visited. */
orf_transition orf_recover(const orf_mesh* m, uint32_t cell, int32_t exit_face, const double p[3],
                           const double u[3])
{
    orf_transition r = {cell, -1, -1, ORF_ENTERED};
    const uint32_t none = 0xFFFFFFFFu;
    uint32_t recent0 = cell, recent1 = none, recent2 = none;
    for (uint32_t step = 0; step < ORF_MAX_RECOVERY; ++step) {
        const int32_t nb = m->nbr[4 * (size_t)cell + exit_face];
        if (nb < 0) {
            r.cell = cell;
            r.boundary_face = exit_face;
            r.status = ORF_BOUNDARY;
            return r;
        }
        const int32_t nf = m->nbr_face[4 * (size_t)cell + exit_face];
        cell = (uint32_t)nb;
        r.cell = cell;
        r.forbidden = nf;
        recent2 = recent1;
        recent1 = recent0;
        recent0 = cell;
        int32_t fallback = -1, preferred = -1;
        const double* pl = m->planes + 16 * (size_t)cell;
        for (int k = 0; k < 4; ++k) {
            double c = orf_lambda(pl + 4 * k, p);
            if (fabs(c) <= ORF_TOL) c = 0.0;
            const double d = orf_dlambda(pl + 4 * k, u);
            if (c + ORF_PROBE * d >= 0.0) continue;
            if (fallback < 0) fallback = k;
            const int32_t cn = m->nbr[4 * (size_t)cell + k];
            if (cn < 0) continue;
            if ((uint32_t)cn != recent0 && (uint32_t)cn != recent1 && (uint32_t)cn != recent2) {
                preferred = k;
                break;
            }
        }
        exit_face = preferred >= 0 ? preferred : fallback;
        if (exit_face < 0) return r;
    }
    r.status = ORF_FAILED;
    return r;
}

/* Transition across the exit face found by orf_next_face
 * (rayTransition.hpp:162-188). */
orf_transition orf_cross(const orf_mesh* m, uint32_t cell, orf_hit hit, const double p[3], const double u[3])
{
    orf_transition r = {cell, -1, -1, ORF_FAILED};
    if (hit.face < 0) return r;
    if (hit.tied_mask & (hit.tied_mask - 1u)) return orf_recover(m, cell, hit.face, p, u);
    const int32_t nb = m->nbr[4 * (size_t)cell + hit.face];
    if (nb < 0) {
        r.boundary_face = hit.face;
        r.status = ORF_BOUNDARY;
        return r;
    }
    r.cell = (uint32_t)nb;
    r.forbidden = m->nbr_face[4 * (size_t)cell + hit.face];
    r.status = ORF_ENTERED;
    return r;
}

static inline double orf_gain_coefficient(const orf_mesh* m, uint32_t cell, double sa, double se)
{
    if (m->cell_types && m->cell_types[cell] == m->cladding_number) return -m->cladding_absorption;
    return m->n_tot * (m->beta[cell] * (se + sa) - sa);
}

int orf_walk_from(const orf_mesh* m, double sigma_a, double sigma_e, uint32_t cell, const double p0[3],
                  const double u[3], int32_t forbidden, double weight, double* scores, uint32_t* visits,
                  uint32_t* dropped, uint64_t* segments, orf_exit* out)
{
    double p[3] = {p0[0], p0[1], p0[2]};
    double gain = 1.0;
    if (out) {
        out->cell = cell;
        out->face = -1;
        out->weight = 0.0;
    }
    for (uint32_t step = 0; step < ORF_MAX_STEPS; ++step) {
        orf_hit h = orf_next_face(m, cell, p, u, forbidden);
        if (h.face < 0) {
            /* stuck on a face: leave through a face the ray is already on
             * (rayTransition.hpp:41-72, policyRay.hpp:394-425) */
            int32_t face = -1;
            const double* pl = m->planes + 16 * (size_t)cell;
            int near = 0;
            for (int k = 0; k < 4; ++k)
                if (fabs(orf_lambda(pl + 4 * k, p)) <= ORF_TOL) near = 1;
            if (near)
                for (int k = 0; k < 4; ++k) {
                    if (k == forbidden) continue;
                    if (fabs(orf_lambda(pl + 4 * k, p)) <= ORF_TOL && orf_dlambda(pl + 4 * k, u) < 0.0) {
                        face = k;
                        break;
                    }
                }
            if (face < 0) { dropped[cell] += 1; return 1; }
            orf_transition t = orf_recover(m, cell, face, p, u);
            if (t.status == ORF_FAILED) { dropped[t.cell] += 1; return 1; }
            if (t.status == ORF_BOUNDARY) {
                if (out) {
                    out->cell = t.cell;
                    out->face = t.boundary_face;
                    out->weight = weight * gain;
                }
                return 0;
            }
            cell = t.cell;
            forbidden = t.forbidden;
            continue;
        }
        /* score the segment (accumulation.hpp:310-345): the track-length
         * integral of exp(g s) over the segment (rayWalk.hpp:194-208, with
         * expm1), then the gain carried to the next cell */
        const double g = orf_gain_coefficient(m, cell, sigma_a, sigma_e);
        const double gl = g * h.length;
        const double em1 = expm1(gl); /* track integral and gain factor */
        const double contribution = weight * gain * (gl == 0.0 ? h.length : em1 / g);
        if (isfinite(contribution)) {
            visits[cell] += 1;
            *segments += 1;
            const double half = 0.5 * h.length;
            const double mid[3] = {p[0] + u[0] * half, p[1] + u[1] * half, p[2] + u[2] * half};
            const double* pl = m->planes + 16 * (size_t)cell;
            double w[4], sum = 0.0;
            for (int k = 0; k < 4; ++k) {
                w[k] = fmax(0.0, orf_lambda(pl + 4 * k, mid));
                sum += w[k];
            }
            const uint32_t* cv = m->cells + 4 * (size_t)cell;
            const uint32_t off = (m->cell_types && m->cell_types[cell] == m->cladding_number) ? m->n_points : 0u;
            if (sum <= DBL_EPSILON) {
                for (int k = 0; k < 4; ++k) scores[off + cv[k]] += contribution * 0.25;
            } else {
                const double inv = 1.0 / sum;
                for (int k = 0; k < 4; ++k) scores[off + cv[k]] += contribution * (w[k] * inv);
            }
        } else {
            dropped[cell] += 1;
        }
        gain *= 1.0 + em1;
        p[0] = p[0] + u[0] * h.length;
        p[1] = p[1] + u[1] * h.length;
        p[2] = p[2] + u[2] * h.length;
        orf_transition t = orf_cross(m, cell, h, p, u);
        if (t.status == ORF_FAILED) { dropped[t.cell] += 1; return 1; }
        if (t.status == ORF_BOUNDARY) {
            if (out) {
                out->cell = t.cell;
                out->face = t.boundary_face;
                out->weight = weight * gain;
            }
            return 0;
        }
        cell = t.cell;
        forbidden = t.forbidden;
    }
    dropped[cell] += 1;
    return 1;
}

int orf_walk(const orf_mesh* m, double sigma_a, double sigma_e, uint32_t cell, const double p0[3],
             const double u[3], double* scores, uint32_t* visits, uint32_t* dropped, uint64_t* segments)
{
    return orf_walk_from(m, sigma_a, sigma_e, cell, p0, u, -1, 1.0, scores, visits, dropped, segments, NULL);
}

/* ------------------------------------------------------------ ray starts */
void orf_batch_range(uint32_t rays, uint32_t batch, uint32_t* first, uint32_t* count)
{
    const uint32_t base = rays / ORF_BATCHES, rem = rays % ORF_BATCHES;
    *count = base + (batch < rem ? 1u : 0u);
    *first = batch * base + (batch < rem ? batch : rem);
}

void orf_batch_controls(uint32_t seed, uint32_t launch, uint32_t batch, uint32_t n_lambda, double* shift,
                        uint32_t* phase)
{
    const uint32_t key[2] = {seed, ORF_KEY_TAG};
    const uint32_t ctr[4] = {0xFFFFFFFFu, batch, launch, 2u};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    *shift = orc_u32_to_unit(o[0]);
    *phase = n_lambda ? o[1] % n_lambda : 0u;
}

/* Cube folded into the unit tetrahedron (Rocchini & Cignoni 2000): returns
 * the barycentric weights of vertices 0..3. */
void orf_point_in_tet(const double u3[3], double bary[4])
{
    double s = u3[0], t = u3[1], v = u3[2];
    if (s + t > 1.0) {
        s = 1.0 - s;
        t = 1.0 - t;
    }
    if (t + v > 1.0) {
        const double tmp = v;
        v = 1.0 - s - t;
        t = 1.0 - tmp;
    } else if (s + t + v > 1.0) {
        const double tmp = v;
        v = s + t + v - 1.0;
        s = 1.0 - t - tmp;
    }
    bary[0] = 1.0 - s - t - v;
    bary[1] = s;
    bary[2] = t;
    bary[3] = v;
}

/* Marsaglia (1972): (x1, x2) uniform in the unit disk gives the unit vector
 * (2 x1 r, 2 x2 r, 1 - 2 s), s = x1^2 + x2^2, r = sqrt(1 - s).  Returns the
 * number of Philox calls used (0 when none of the 64 pairs was accepted and
 * +z was taken). */
int orf_direction(uint32_t seed, uint32_t launch, uint32_t ray, double u[3])
{
    const uint32_t key[2] = {seed, ORF_KEY_TAG};
    for (uint32_t call = 0; call < ORF_DIRECTION_CALLS; ++call) {
        const uint32_t ctr[4] = {ray, launch, call + 1u, 0u};
        uint32_t o[4];
        orc_philox4x32_10(ctr, key, o);
        for (int pair = 0; pair < 2; ++pair) {
            const double x1 = 2.0 * orc_u32_to_unit(o[2 * pair]) - 1.0;
            const double x2 = 2.0 * orc_u32_to_unit(o[2 * pair + 1]) - 1.0;
            const double s = x1 * x1 + x2 * x2;
            if (s >= 1.0) continue;
            const double r = sqrt(1.0 - s);
            u[0] = 2.0 * x1 * r;
            u[1] = 2.0 * x2 * r;
            u[2] = 1.0 - 2.0 * s;
            return (int)call + 1;
        }
    }
    u[0] = 0.0;
    u[1] = 0.0;
    u[2] = 1.0;
    return 0;
}

void orf_ray_start(const orf_mesh* m, uint32_t seed, uint32_t launch, uint32_t rays, uint32_t ray,
                   uint32_t n_lambda, orf_ray* out)
{
    const uint32_t base = rays / ORF_BATCHES, rem = rays % ORF_BATCHES;
    const uint32_t split = rem * (base + 1u);
    const uint32_t b = ray < split ? ray / (base + 1u) : rem + (ray - split) / base;
    uint32_t first, count;
    orf_batch_range(rays, b, &first, &count);
    const uint32_t i = ray - first;
    double shift;
    uint32_t phase;
    orf_batch_controls(seed, launch, b, n_lambda, &shift, &phase);
    /* stratified beta*V source (volumeSampling.hpp:65-79, 118-137) */
    const uint32_t n = m->n_cells;
    const double total = m->bv_prefix[n - 1];
    const double target = ((double)i + shift) / (double)count * total;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (target < m->bv_prefix[mid]) hi = mid;
        else lo = mid + 1u;
    }
    const uint32_t cell = lo < n ? lo : n - 1u;
    const uint32_t key[2] = {seed, ORF_KEY_TAG};
    const uint32_t ctr[4] = {ray, launch, 0u, 0u};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const double u3[3] = {orc_u32_to_unit(o[0]), orc_u32_to_unit(o[1]), orc_u32_to_unit(o[2])};
    double w[4];
    orf_point_in_tet(u3, w);
    const uint32_t* cv = m->cells + 4 * (size_t)cell;
    out->p[0] = w[0] * m->px[cv[0]] + w[1] * m->px[cv[1]] + w[2] * m->px[cv[2]] + w[3] * m->px[cv[3]];
    out->p[1] = w[0] * m->py[cv[0]] + w[1] * m->py[cv[1]] + w[2] * m->py[cv[2]] + w[3] * m->py[cv[3]];
    out->p[2] = w[0] * m->pz[cv[0]] + w[1] * m->pz[cv[1]] + w[2] * m->pz[cv[2]] + w[3] * m->pz[cv[3]];
    orf_direction(seed, launch, ray, out->u);
    out->cell = cell;
    /* stratified wavelength (volumeSampling.hpp:19-32) */
    out->lambda = n_lambda ? (uint32_t)(((uint64_t)i * n_lambda / count + phase) % n_lambda) : 0u;
    out->batch = b;
    out->batch_index = i;
    out->batch_count = count;
}

/* ---------------------------------------------------------------- launch */
void orf_trace(const orf_mesh* m, const double* sigma_a, const double* sigma_e, uint32_t n_lambda, uint32_t seed,
               uint32_t launch, uint32_t rays, double* scores, uint32_t* visits, uint32_t* dropped,
               uint32_t* batch_counts, uint64_t* segments, int threads)
{
    const size_t nv = 2 * (size_t)m->n_points;
    for (uint32_t b = 0; b < ORF_BATCHES; ++b) {
        uint32_t first, count;
        orf_batch_range(rays, b, &first, &count);
        batch_counts[b] += count;
    }
    if (rays == 0 || m->bv_prefix[m->n_cells - 1] <= 0.0) return;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    uint64_t seg_total = 0;
#pragma omp parallel reduction(+ : seg_total)
    {
        double* my_scores = (double*)calloc(ORF_BATCHES * nv, sizeof(double));
        uint32_t* my_visits = (uint32_t*)calloc(m->n_cells, sizeof(uint32_t));
        uint32_t* my_dropped = (uint32_t*)calloc(m->n_cells, sizeof(uint32_t));
        uint64_t seg = 0;
#pragma omp for schedule(static)
        for (int64_t r = 0; r < (int64_t)rays; ++r) {
            orf_ray ray;
            orf_ray_start(m, seed, launch, rays, (uint32_t)r, n_lambda, &ray);
            orf_walk(m, sigma_a[ray.lambda], sigma_e[ray.lambda], ray.cell, ray.p, ray.u, my_scores + ray.batch * nv,
                     my_visits, my_dropped, &seg);
        }
#pragma omp critical
        {
            for (size_t i = 0; i < ORF_BATCHES * nv; ++i) scores[i] += my_scores[i];
            for (uint32_t c = 0; c < m->n_cells; ++c) {
                visits[c] += my_visits[c];
                dropped[c] += my_dropped[c];
            }
        }
        seg_total += seg;
        free(my_scores);
        free(my_visits);
        free(my_dropped);
    }
    if (segments) *segments += seg_total;
}

void orf_finalize(const orf_mesh* m, const double* scores, const uint32_t* batch_counts, uint64_t rays,
                  const uint32_t* dropped, double sigma_a, double sigma_e, float* phi, double* se, double* rse,
                  double* dndt)
{
    const size_t nv = 2 * (size_t)m->n_points;
    const double total = m->bv_prefix[m->n_cells - 1];
    const double rate = m->n_tot / m->tau;
    for (uint32_t c = 0; c < m->n_cells; ++c) {
        const double vol = m->volume[c];
        double rel = DBL_MAX, abs_err = DBL_MAX, estimate = 0.0;
        if (rays > 0 && vol > 0.0) {
            const uint32_t* cv = m->cells + 4 * (size_t)c;
            const uint32_t off = (m->cell_types && m->cell_types[c] == m->cladding_number) ? m->n_points : 0u;
            double score_sum = 0.0, mean_sum = 0.0, mean_sq = 0.0;
            uint32_t active = 0;
            for (uint32_t b = 0; b < ORF_BATCHES; ++b) {
                double dens = 0.0;
                for (int k = 0; k < 4; ++k) {
                    const double lv = m->lumped[off + cv[k]];
                    dens += lv > 0.0 ? scores[b * nv + off + cv[k]] / lv : 0.0;
                }
                dens /= 4.0;
                const double bs = dens * vol;
                score_sum += bs;
                if (batch_counts[b] == 0) continue;
                const double mean = bs / (double)batch_counts[b];
                mean_sum += mean;
                mean_sq += mean * mean;
                ++active;
            }
            estimate = score_sum * total / ((double)rays * vol);
            if (dropped[c] == 0 && active >= 2) {
                const double n = (double)active;
                const double mean = mean_sum / n;
                if (mean == 0.0) {
                    rel = NAN;
                    abs_err = 0.0;
                } else {
                    const double var = fmax(0.0, (mean_sq - mean_sum * mean_sum / n) / (n - 1.0));
                    rel = sqrt(var / n) / fabs(mean);
                    abs_err = rel * fabs(estimate) * rate;
                }
            }
        }
        const float f = (float)(estimate * rate);
        phi[c] = f;
        se[c] = abs_err;
        rse[c] = rel;
        dndt[c] = (m->beta[c] * (sigma_e + sigma_a) - sigma_a) * (double)f;
    }
}

static int orf_evaluate_impl(const orf_mesh* m, const orf_surfaces* sf, orf_srm* srm, const double* sigma_a,
                             const double* sigma_e, uint32_t n_lambda, double max_sigma_a, double max_sigma_e,
                             uint32_t min_rays, uint32_t max_rays, uint32_t steps, double threshold, uint32_t seed,
                             float* phi, double* se, double* rse, double* dndt, uint32_t* visits, uint32_t* dropped,
                             uint32_t* convergence_rays, uint64_t* total_rays, int threads);

int orf_evaluate(const orf_mesh* m, const double* sigma_a, const double* sigma_e, uint32_t n_lambda,
                 double max_sigma_a, double max_sigma_e, uint32_t min_rays, uint32_t max_rays, uint32_t steps,
                 double threshold, uint32_t seed, float* phi, double* se, double* rse, double* dndt,
                 uint32_t* visits, uint32_t* dropped, uint32_t* convergence_rays, uint64_t* total_rays, int threads)
{
    return orf_evaluate_impl(m, NULL, NULL, sigma_a, sigma_e, n_lambda, max_sigma_a, max_sigma_e, min_rays,
                             max_rays, steps, threshold, seed, phi, se, rse, dndt, visits, dropped,
                             convergence_rays, total_rays, threads);
}

int orf_evaluate_srm(const orf_mesh* m, const orf_surfaces* sf, orf_srm* srm, const double* sigma_a,
                     const double* sigma_e, uint32_t n_lambda, double max_sigma_a, double max_sigma_e,
                     uint32_t min_rays, uint32_t max_rays, uint32_t steps, double threshold, uint32_t seed,
                     float* phi, double* se, double* rse, double* dndt, uint32_t* visits, uint32_t* dropped,
                     uint32_t* convergence_rays, uint64_t* total_rays)
{
    return orf_evaluate_impl(m, sf, srm, sigma_a, sigma_e, n_lambda, max_sigma_a, max_sigma_e, min_rays, max_rays,
                             steps, threshold, seed, phi, se, rse, dndt, visits, dropped, convergence_rays,
                             total_rays, 1);
}

static int orf_evaluate_impl(const orf_mesh* m, const orf_surfaces* sf, orf_srm* srm, const double* sigma_a,
                             const double* sigma_e, uint32_t n_lambda, double max_sigma_a, double max_sigma_e,
                             uint32_t min_rays, uint32_t max_rays, uint32_t steps, double threshold, uint32_t seed,
                             float* phi, double* se, double* rse, double* dndt, uint32_t* visits, uint32_t* dropped,
                             uint32_t* convergence_rays, uint64_t* total_rays, int threads)
{
    const size_t nv = 2 * (size_t)m->n_points;
    double* scores = (double*)calloc(ORF_BATCHES * nv, sizeof(double));
    if (!scores) return -3;
    uint32_t counts[ORF_BATCHES] = {0};
    memset(visits, 0, m->n_cells * sizeof(uint32_t));
    memset(dropped, 0, m->n_cells * sizeof(uint32_t));
    memset(convergence_rays, 0, m->n_cells * sizeof(uint32_t));
    uint64_t done = 0;
    int launches = 0;
    for (uint32_t k = 0;; ++k) {
        const uint32_t target = orc_adaptive_target(min_rays, max_rays, steps, k);
        const uint32_t launch_rays = (uint32_t)(target - done);
        if (sf)
            orf_trace_srm(m, sf, srm, sigma_a, sigma_e, n_lambda, seed, (uint32_t)launches, launch_rays, scores,
                          visits, dropped, counts, NULL);
        else
            orf_trace(m, sigma_a, sigma_e, n_lambda, seed, (uint32_t)launches, launch_rays, scores, visits, dropped,
                      counts, NULL, threads);
        done = target;
        ++launches;
        orf_finalize(m, scores, counts, done, dropped, max_sigma_a, max_sigma_e, phi, se, rse, dndt);
        int all = 1;
        for (uint32_t c = 0; c < m->n_cells; ++c) {
            const int ok = dropped[c] == 0 && isfinite(rse[c]) && rse[c] * rse[c] <= threshold;
            if (ok && convergence_rays[c] == 0) convergence_rays[c] = target;
            if (!ok) all = 0;
        }
        if (all || target >= max_rays || steps == 0 || max_rays <= min_rays) break;
    }
    *total_rays = done;
    free(scores);
    return launches;
}

/* ------------------------------------------------------------ reflections */
static void orf_normalize(double v[3])
{
    const double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len <= DBL_EPSILON) {
        v[0] = v[1] = v[2] = 0.0;
        return;
    }
    const double inv = 1.0 / len;
    v[0] = v[0] * inv;
    v[1] = v[1] * inv;
    v[2] = v[2] * inv;
}

int orf_surfaces_build(const orf_mesh* m, const int32_t* face_surface, uint32_t n_surfaces,
                       const float* reflectivity, const float* n_in, const float* n_out, orf_surfaces* out)
{
    memset(out, 0, sizeof(*out));
    const uint32_t n = m->n_cells;
    out->face_index = (int32_t*)malloc(4 * (size_t)n * sizeof(int32_t));
    if (!out->face_index) return -3;
    uint32_t count = 0;
    for (size_t i = 0; i < 4 * (size_t)n; ++i) {
        const int32_t id = face_surface ? face_surface[i] : 0;
        const int ok = m->nbr[i] < 0 && id > 0 && (uint32_t)id < n_surfaces;
        out->face_index[i] = ok ? (int32_t)count++ : -1;
    }
    out->n_faces = count;
    out->face = (uint32_t*)malloc((count ? count : 1) * sizeof(uint32_t));
    out->centroid = (double*)malloc((count ? count : 1) * 3 * sizeof(double));
    out->normal = (double*)malloc((count ? count : 1) * 3 * sizeof(double));
    out->reflectivity = (double*)malloc((count ? count : 1) * sizeof(double));
    out->n_in = (double*)malloc((count ? count : 1) * sizeof(double));
    out->n_out = (double*)malloc((count ? count : 1) * sizeof(double));
    if (!out->face || !out->centroid || !out->normal || !out->reflectivity || !out->n_in || !out->n_out) return -3;
    for (size_t i = 0; i < 4 * (size_t)n; ++i) {
        const int32_t f = out->face_index[i];
        if (f < 0) continue;
        const uint32_t c = (uint32_t)(i / 4), k = (uint32_t)(i % 4);
        const uint32_t* cv = m->cells + 4 * (size_t)c;
        uint32_t fv[3], j = 0;
        for (uint32_t q = 0; q < 4; ++q)
            if (q != k) fv[j++] = cv[q];
        orf_sort3(fv);
        /* face centroid and outward normal (rayWalk.hpp:42-84) */
        double cen[3], ctr[3];
        cen[0] = (m->px[fv[0]] + m->px[fv[1]] + m->px[fv[2]]) * (1.0 / 3.0);
        cen[1] = (m->py[fv[0]] + m->py[fv[1]] + m->py[fv[2]]) * (1.0 / 3.0);
        cen[2] = (m->pz[fv[0]] + m->pz[fv[1]] + m->pz[fv[2]]) * (1.0 / 3.0);
        ctr[0] = (m->px[cv[0]] + m->px[cv[1]] + m->px[cv[2]] + m->px[cv[3]]) * 0.25;
        ctr[1] = (m->py[cv[0]] + m->py[cv[1]] + m->py[cv[2]] + m->py[cv[3]]) * 0.25;
        ctr[2] = (m->pz[cv[0]] + m->pz[cv[1]] + m->pz[cv[2]] + m->pz[cv[3]]) * 0.25;
        const double bx = m->px[fv[1]] - m->px[fv[0]], by = m->py[fv[1]] - m->py[fv[0]],
                     bz = m->pz[fv[1]] - m->pz[fv[0]];
        const double cx = m->px[fv[2]] - m->px[fv[0]], cy = m->py[fv[2]] - m->py[fv[0]],
                     cz = m->pz[fv[2]] - m->pz[fv[0]];
        double nrm[3] = {by * cz - bz * cy, bz * cx - bx * cz, bx * cy - by * cx};
        orf_normalize(nrm);
        if (nrm[0] * (ctr[0] - cen[0]) + nrm[1] * (ctr[1] - cen[1]) + nrm[2] * (ctr[2] - cen[2]) > 0.0) {
            nrm[0] = -nrm[0];
            nrm[1] = -nrm[1];
            nrm[2] = -nrm[2];
        }
        const int32_t id = face_surface[i];
        out->face[f] = (uint32_t)i;
        memcpy(out->centroid + 3 * (size_t)f, cen, sizeof(cen));
        memcpy(out->normal + 3 * (size_t)f, nrm, sizeof(nrm));
        out->reflectivity[f] = reflectivity ? (double)reflectivity[id] : 0.0;
        out->n_in[f] = n_in ? (double)n_in[id] : 0.0;
        out->n_out[f] = n_out ? (double)n_out[id] : 0.0;
    }
    return 0;
}

void orf_surfaces_free(orf_surfaces* s)
{
    free(s->face);
    free(s->face_index);
    free(s->centroid);
    free(s->normal);
    free(s->reflectivity);
    free(s->n_in);
    free(s->n_out);
    memset(s, 0, sizeof(*s));
}

/* boundaryReflectance (include/kernels/reflection.hpp:21-47): total internal
 * reflection when (n_in / n_out)^2 sin^2(theta) > 1, else the reflectivity. */
double orf_reflectance(const orf_surfaces* s, uint32_t f, const double dir[3])
{
    const double nin = s->n_in[f], nout = s->n_out[f];
    if (nin > 0.0 && nout > 0.0) {
        double d[3] = {dir[0], dir[1], dir[2]};
        orf_normalize(d);
        const double* nr = s->normal + 3 * (size_t)f;
        const double cosi = fmin(1.0, fabs(d[0] * nr[0] + d[1] * nr[1] + d[2] * nr[2]));
        const double sin2 = fmax(0.0, 1.0 - cosi * cosi);
        const double ratio = nin / nout;
        if (ratio * ratio * sin2 > 1.0) return 1.0;
    }
    return fmax(0.0, s->reflectivity[f]);
}

typedef struct {
    uint32_t* count;       /* per face */
    double* face_weight;   /* per face */
    double* dir;           /* 3 per slot */
    double* weight;        /* per slot */
    uint32_t* sigma;       /* per slot */
} orf_reservoir;

static int orf_reservoir_alloc(orf_reservoir* r, size_t faces, size_t slots)
{
    const size_t fs = faces * slots > 0 ? faces * slots : 1, nf = faces > 0 ? faces : 1;
    r->count = (uint32_t*)calloc(nf, sizeof(uint32_t));
    r->face_weight = (double*)calloc(nf, sizeof(double));
    r->dir = (double*)calloc(fs * 3, sizeof(double));
    r->weight = (double*)calloc(fs, sizeof(double));
    r->sigma = (uint32_t*)calloc(fs, sizeof(uint32_t));
    return (r->count && r->face_weight && r->dir && r->weight && r->sigma) ? 0 : -3;
}

static void orf_reservoir_free(orf_reservoir* r)
{
    free(r->count);
    free(r->face_weight);
    free(r->dir);
    free(r->weight);
    free(r->sigma);
}

/* depositReflection (accumulation.hpp:409-455): weighted reservoir sampling
 * of the reflected direction; u0 decides a replacement, u1 picks the slot. */
static void orf_deposit(const orf_surfaces* sf, orf_reservoir* r, uint32_t slots, uint32_t f, const double dir[3],
                        double w, uint32_t sigma, double u0, double u1)
{
    if (slots == 0 || !(w > 0.0) || !isfinite(w)) return;
    const double rw = w * orf_reflectance(sf, f, dir);
    if (!(rw > 0.0) || !isfinite(rw)) return;
    const uint32_t seen = r->count[f]++;
    const double prior = r->face_weight[f];
    r->face_weight[f] = prior + rw;
    uint32_t slot = seen;
    if (slot >= slots) {
        const double keep = fmin(1.0, (double)slots * rw / (prior + rw));
        if (u0 >= keep) return;
        slot = (uint32_t)(u1 * (double)slots);
        if (slot >= slots) slot = slots - 1;
    }
    const size_t idx = (size_t)f * slots + slot;
    const double* nr = sf->normal + 3 * (size_t)f;
    const double dn = 2.0 * (dir[0] * nr[0] + dir[1] * nr[1] + dir[2] * nr[2]);
    double out[3] = {dir[0] - nr[0] * dn, dir[1] - nr[1] * dn, dir[2] - nr[2] * dn};
    orf_normalize(out);
    r->dir[3 * idx] = out[0];
    r->dir[3 * idx + 1] = out[1];
    r->dir[3 * idx + 2] = out[2];
    r->weight[idx] = rw;
    r->sigma[idx] = sigma;
}

static void orf_srm_uniforms(uint32_t seed, uint32_t launch, uint32_t ray, uint32_t pass, uint32_t use,
                             double* u0, double* u1)
{
    const uint32_t key[2] = {seed, ORF_KEY_TAG};
    const uint32_t ctr[4] = {ray, launch, 0x10000u + pass, use};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    *u0 = orc_u32_to_unit(o[0]);
    *u1 = orc_u32_to_unit(o[1]);
}

static double orf_srm_shift(uint32_t seed, uint32_t launch, uint32_t batch, uint32_t pass)
{
    const uint32_t key[2] = {seed, ORF_KEY_TAG};
    const uint32_t ctr[4] = {0xFFFFFFFEu, batch, launch, pass};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return orc_u32_to_unit(o[0]);
}

/* position k of n -> ray index; a permutation only when the stride is coprime
 * to n, else the ray order is kept */
static uint32_t orf_deposit_order(const orf_srm* srm, uint32_t k, uint32_t n)
{
    uint32_t a = srm->order_stride, b = n;
    if (a <= 1u || n == 0u) return k;
    while (b) {
        const uint32_t r = a % b;
        a = b;
        b = r;
    }
    return a == 1u ? (uint32_t)(((uint64_t)k * srm->order_stride) % n) : k;
}

void orf_trace_srm(const orf_mesh* m, const orf_surfaces* sf, orf_srm* srm, const double* sigma_a,
                   const double* sigma_e, uint32_t n_lambda, uint32_t seed, uint32_t launch, uint32_t rays,
                   double* scores, uint32_t* visits, uint32_t* dropped, uint32_t* batch_counts, uint64_t* segments)
{
    const size_t nv = 2 * (size_t)m->n_points, nb = sf->n_faces, S = srm->slots;
    for (uint32_t b = 0; b < ORF_BATCHES; ++b) {
        uint32_t first, count;
        orf_batch_range(rays, b, &first, &count);
        batch_counts[b] += count;
        srm->status[b] = ORF_SRM_CONVERGED;
        srm->passes[b] = 0;
        srm->remaining[b] = 0.0;
    }
    if (rays == 0 || m->bv_prefix[m->n_cells - 1] <= 0.0) return;
    orf_reservoir res[2][ORF_BATCHES];
    for (int side = 0; side < 2; ++side)
        for (uint32_t b = 0; b < ORF_BATCHES; ++b) orf_reservoir_alloc(&res[side][b], nb, S);
    uint64_t seg = 0;
    /* pass 0: primary rays, reflections into side 0 */
    for (uint32_t k = 0; k < rays; ++k) {
        const uint32_t r = orf_deposit_order(srm, k, rays);
        orf_ray ray;
        orf_ray_start(m, seed, launch, rays, r, n_lambda, &ray);
        orf_exit ex;
        orf_walk_from(m, sigma_a[ray.lambda], sigma_e[ray.lambda], ray.cell, ray.p, ray.u, -1, 1.0,
                      scores + ray.batch * nv, visits, dropped, &seg, &ex);
        if (ex.face >= 0) {
            const int32_t f = sf->face_index[4 * (size_t)ex.cell + ex.face];
            if (f >= 0) {
                double u0, u1;
                orf_srm_uniforms(seed, launch, r, 0, 1, &u0, &u1);
                orf_deposit(sf, &res[0][ray.batch], (uint32_t)S, (uint32_t)f, ray.u, ex.weight, ray.lambda, u0, u1);
            }
        }
    }
    double initial[ORF_BATCHES], previous[ORF_BATCHES];
    uint32_t grow[ORF_BATCHES];
    int active[ORF_BATCHES];
    for (uint32_t b = 0; b < ORF_BATCHES; ++b) {
        double w = 0.0;
        for (size_t f = 0; f < nb; ++f) w += res[0][b].face_weight[f];
        initial[b] = previous[b] = w;
        grow[b] = 0;
        active[b] = w > 0.0;
        srm->status[b] = active[b] ? ORF_SRM_MAX_PASSES : ORF_SRM_CONVERGED;
        srm->remaining[b] = active[b] ? 1.0 : 0.0;
    }
    double* cdf = (double*)malloc((nb ? nb : 1) * sizeof(double));
    int in = 0;
    for (uint32_t pass = 1; pass <= srm->max_passes; ++pass) {
        int any = 0;
        for (uint32_t b = 0; b < ORF_BATCHES; ++b) any |= active[b];
        if (!any) break;
        const int out = 1 - in;
        for (uint32_t b = 0; b < ORF_BATCHES; ++b) {
            memset(res[out][b].count, 0, (nb ? nb : 1) * sizeof(uint32_t));
            memset(res[out][b].face_weight, 0, (nb ? nb : 1) * sizeof(double));
        }
        for (uint32_t b = 0; b < ORF_BATCHES; ++b) {
            if (!active[b]) continue;
            uint32_t first, count;
            orf_batch_range(rays, b, &first, &count);
            if (count == 0) continue;
            const orf_reservoir* src = &res[in][b];
            double run = 0.0;
            for (size_t f = 0; f < nb; ++f) {
                run += src->face_weight[f];
                cdf[f] = run;
            }
            for (size_t f = 0; f < nb; ++f) cdf[f] = run > 0.0 ? cdf[f] / run : 0.0;
            const double source = previous[b] / (double)count;
            const double shift = orf_srm_shift(seed, launch, b, pass);
            for (uint32_t k = 0; k < count; ++k) {
                const uint32_t i = orf_deposit_order(srm, k, count);
                const uint32_t r = first + i;
                double u0, u1;
                orf_srm_uniforms(seed, launch, r, pass, 2, &u0, &u1);
                /* face by reservoir weight, stratified when every face can get a ray
                 * (accumulation.hpp:706-746) */
                const double target = nb <= count ? ((double)i + shift) / (double)count : u0;
                uint32_t lo = 0, hi = (uint32_t)nb;
                while (lo < hi) {
                    const uint32_t mid = lo + (hi - lo) / 2u;
                    if (cdf[mid] <= target) lo = mid + 1u;
                    else hi = mid;
                }
                const uint32_t f = lo < nb ? lo : (uint32_t)nb - 1u;
                const uint32_t filled = src->count[f] < S ? src->count[f] : (uint32_t)S;
                if (filled == 0) continue;
                double slot_weight = 0.0;
                for (uint32_t k = 0; k < filled; ++k) slot_weight += src->weight[(size_t)f * S + k];
                if (!(slot_weight > 0.0)) continue;
                const double t2 = u1 * slot_weight;
                double cum = 0.0;
                uint32_t pick = filled - 1u;
                for (uint32_t k = 0; k < filled; ++k) {
                    cum += src->weight[(size_t)f * S + k];
                    if (cum >= t2) {
                        pick = k;
                        break;
                    }
                }
                const size_t idx = (size_t)f * S + pick;
                double dir[3] = {src->dir[3 * idx], src->dir[3 * idx + 1], src->dir[3 * idx + 2]};
                orf_normalize(dir);
                const uint32_t lam = src->sigma[idx];
                const uint32_t face = sf->face[f];
                orf_exit ex;
                orf_walk_from(m, sigma_a[lam], sigma_e[lam], face / 4, sf->centroid + 3 * (size_t)f, dir,
                              (int32_t)(face % 4), source, scores + b * nv, visits, dropped, &seg, &ex);
                if (ex.face >= 0) {
                    const int32_t g = sf->face_index[4 * (size_t)ex.cell + ex.face];
                    if (g >= 0) {
                        double v0, v1;
                        orf_srm_uniforms(seed, launch, r, pass, 1, &v0, &v1);
                        orf_deposit(sf, &res[out][b], (uint32_t)S, (uint32_t)g, dir, ex.weight, lam, v0, v1);
                    }
                }
            }
        }
        /* stop rules (forwardSrm.hpp:322-361) */
        for (uint32_t b = 0; b < ORF_BATCHES; ++b) {
            if (!active[b]) continue;
            double w = 0.0;
            for (size_t f = 0; f < nb; ++f) w += res[out][b].face_weight[f];
            srm->passes[b] = pass;
            srm->remaining[b] = w / initial[b];
            if (w > previous[b]) {
                if (++grow[b] >= srm->divergence_streak) {
                    srm->status[b] = ORF_SRM_DIVERGED;
                    active[b] = 0;
                    continue;
                }
            } else {
                grow[b] = 0;
                if (fabs(w - previous[b]) / fmax(w, 1.0e-30) < srm->tolerance) {
                    srm->status[b] = ORF_SRM_STABLE;
                    active[b] = 0;
                    continue;
                }
            }
            if (srm->remaining[b] < srm->tolerance) {
                srm->status[b] = ORF_SRM_CONVERGED;
                active[b] = 0;
                continue;
            }
            previous[b] = w;
        }
        in = out;
    }
    free(cdf);
    for (int side = 0; side < 2; ++side)
        for (uint32_t b = 0; b < ORF_BATCHES; ++b) orf_reservoir_free(&res[side][b]);
    if (segments) *segments += seg;
}
