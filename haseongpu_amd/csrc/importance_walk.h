// importance_walk.h - centroid walks of the walked importance map, shared by
// the importance kernels and host-side checks (the functions are
// __host__ __device__).
//
// The importance of prism j for a sample point is the emission integrand at
// the prism centroid times the prism volume: beta_j V_j G(c_j -> p) R / (4 pi
// d^2), G the spectral gain along the straight path from the centroid to the
// sample (or its mirror image), R the surface factors of the reflections on
// the way (importanceSamplingPropagation's centroid walk; the reference's
// backward path is retired, its present equivalents are the walk rules of
// include/kernels/forward/rayWalk.hpp:170-208 and
// include/kernels/reflection.hpp:21-47).
#pragma once

#include <cmath>

#include "prism_walk.h"

namespace hase {

// Levels whose centroid walks share one xy walk in the importance kernels:
// more for the direct slice than for the mirror slices, whose per-level
// copy and surface-factor state costs more registers (both groups keep the
// kernels within 128 VGPRs, four wavefronts per SIMD).
constexpr int kGroupDirect = 8;
constexpr int kGroupMirror = 4;
constexpr int kLevelGroup = kGroupDirect;   // the larger of the two

// Column items of one sample: for each reflection slice k and each group of
// that slice's levels, one item per triangle column t0 (items of slice 0
// first, then the mirror slices; columns fastest, so the lanes of a
// wavefront walk neighbouring columns).
struct ColumnItems {
    uint32_t T, layers, R, g0, g1;
    HASE_HD ColumnItems(uint32_t T_, uint32_t layers_, uint32_t R_)
        : T(T_), layers(layers_), R(R_), g0((layers_ + kGroupDirect - 1) / kGroupDirect),
          g1((layers_ + kGroupMirror - 1) / kGroupMirror)
    {
    }
    HASE_HD uint32_t per_sample() const { return T * (g0 + (R - 1u) * g1); }
    HASE_HD void decode(uint32_t it, uint32_t& t0, uint32_t& k, int& lv0, int& nlev) const
    {
        t0 = it % T;
        uint32_t rest = it / T;
        int group;
        if (rest < g0) {
            k = 0u;
            group = kGroupDirect;
        } else {
            rest -= g0;
            k = 1u + rest / g1;
            rest %= g1;
            group = kGroupMirror;
        }
        lv0 = (int)rest * group;
        nlev = (int)layers - lv0 < group ? (int)layers - lv0 : group;
    }
};

// Importance of the prisms (t0, lv0 .. lv0+nlev-1) for one reflection slice.
// Every level (and every mirror slice) of a triangle column projects onto the
// same xy segment from the triangle's centroid to the sample point, so the
// column is walked once in xy and each level's ray is advanced along it.
// All of them are parametrised by the same sigma in [0, 1] (xy and height are
// both linear in sigma, lengths are sigma * D of each level), so what a
// triangle step adds is shared except for beta: the cladding share of sigma
// is one sum for the column, and per level only sum(beta dsigma) is kept,
// split at that level's own plane crossings (a crossing is where the level's
// ray enters the next layer; the next one is tracked as its sigma).  The
// per-level results are those of walk_ray on the same ray: sum(beta l),
// sum(l) over gain prisms, sum(l) over cladding prisms, the surface factors
// at the copy boundaries, and the prisms visited.  The xy walk is the
// side-test walk of prism_walk.h (one new vertex and one side test per
// triangle).
template <int G, bool kReflect, class Fetch>
HASE_HD inline void column_importance(const MeshDev& m, const Fetch& fetch, uint32_t t0, int lv0, int nlev, float px,
                                      float py, int end_plane, uint32_t nrefl, uint32_t first,
                                      float out_bl[G], float out_gl[G],
                                      float out_cl[G], float out_rf[G], float D[G],
                                      unsigned long long& steps, bool& ok)
{
    const int layers = (int)m.layers;
    const float h = m.h;
    const float2 p0 = fetch.vertex(m, t0, 0u), p1 = fetch.vertex(m, t0, 1u), p2 = fetch.vertex(m, t0, 2u);
    const float cx = (p0.x + p1.x + p2.x) * (1.0f / 3.0f);
    const float cy = (p0.y + p1.y + p2.y) * (1.0f / 3.0f);
    const float dx = px - cx, dy = py - cy;
    const float d2 = dx * dx + dy * dy;
    const float Dxy = sqrtf(d2);
    const float inv_d2 = d2 > 0.0f ? 1.0f / d2 : 0.0f;
    const float ze = (float)end_plane;
    // per level: unfolded layer u = cp * layers + r (the prism it is in has
    // folded level r, or layers - 1 - r in odd mirror copies), the sigma of
    // its next plane crossing (measured back from the end, so the end's
    // plane sits at exactly 1 and is never crossed), sum(beta dsigma) and the
    // surface factors so far (one per copy boundary crossed: |cp| of them)
    int r[G], cp[G];
    float snp[G], bsum[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const float dz = ze - ((float)(lv0 + g) + 0.5f);
        r[g] = lv0 + g;
        cp[g] = 0;
        const int plane = lv0 + g + (dz > 0.0f ? 1 : 0);
        snp[g] = fmaf((float)plane - ze, fast_rcp(dz), 1.0f);
        bsum[g] = 0.0f;
        out_rf[g] = 1.0f;
    }
    const float tol = 2e-5f * (1.0f + Dxy);
    // exit of the start triangle: the edge whose start vertex is right of the
    // line and whose end vertex is left of it (counter-clockwise order)
    const float c0 = side_of(dx, dy, p0.x - cx, p0.y - cy), c1 = side_of(dx, dy, p1.x - cx, p1.y - cy);
    const float c2 = side_of(dx, dy, p2.x - cx, p2.y - cy);
    const bool b0 = c0 >= 0.0f, b1 = c1 >= 0.0f, b2 = c2 >= 0.0f;
    int k = (!b0 && b1) ? 0 : (!b1 && b2) ? 1 : (!b2 && b0) ? 2 : -1;
    if (k < 0) {
        if (b0) {
            k = (c0 <= c1 && c0 <= c2) ? 0 : (c1 <= c2 ? 1 : 2);
        } else {
            const int km = (c0 >= c1 && c0 >= c2) ? 0 : (c1 >= c2 ? 1 : 2);
            k = km == 0 ? 2 : km - 1;
        }
    }
    const float2 pr = k == 0 ? p0 : (k == 1 ? p1 : p2);
    const float2 pl = k == 0 ? p1 : (k == 1 ? p2 : p0);
    float cr = k == 0 ? c0 : (k == 1 ? c1 : c2);
    float cl = k == 0 ? c1 : (k == 1 ? c2 : c0);
    float wr = along(dx, dy, pr.x - cx, pr.y - cy);
    float wl = along(dx, dy, pl.x - cx, pl.y - cy);
    uint32_t wexit, unused, clad;
    fetch.words(m, t0, (uint32_t)k, (uint32_t)k, wexit, unused, clad);
    const bool start_end = (p0.x == px && p0.y == py) || (p1.x == px && p1.y == py) || (p2.x == px && p2.y == py) ||
                           d2 == 0.0f;
    float sxy = start_end ? 1.0f : edge_sigma(cl, wl, cr, wr, inv_d2);
    uint32_t t = t0;
    float s = 0.0f;          // sigma reached
    float clad_s = 0.0f;     // sum of dsigma over cladding triangles
    ok = true;
    for (uint32_t step = 0; step < kMaxSteps; ++step) {
        const float* bcol = m.beta_col + (size_t)t * (uint32_t)layers;
        // this column's beta at every level's current layer: independent loads, one round trip
        float b[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int lvl = (kReflect && (cp[g] & 1)) ? (layers - 1 - r[g]) : r[g];
            b[g] = (g < nlev && !clad) ? bcol[lvl] : 0.0f;
        }
        const float snext = fmaxf(fminf(sxy, 1.0f), s);
        const bool at_end = snext >= 1.0f;
        const float sig0 = s;
        const float sig1 = at_end ? 1.0f : snext;
        clad_s = fmaf((float)clad, sig1 - sig0, clad_s);
        steps += (unsigned long long)nlev;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (g < nlev) {
                float a = sig0;
                float bg = b[g];
                while (snp[g] < sig1) {
                    // the level's ray enters the next layer inside this triangle
                    const float sp = fmaxf(snp[g], a);
                    bsum[g] = fmaf(bg, sp - a, bsum[g]);
                    a = sp;
                    steps++;
                    const float dz = ze - ((float)(lv0 + g) + 0.5f);
                    const bool up = dz > 0.0f;
                    const int rr = r[g] + (up ? 1 : -1);
                    if (kReflect && (rr == layers || rr < 0)) {
                        if ((uint32_t)abs(cp[g]) < nrefl) {
                            const float Dg = sqrtf(d2 + dz * h * dz * h);
                            const float vz = dz * h / Dg;
                            const uint32_t pl_ = up ? ((cp[g] & 1) ? 0u : 1u) : ((cp[g] & 1) ? 1u : 0u);
                            out_rf[g] *= surface_factor(m, t, pl_, fmaxf(0.0f, 1.0f - vz * vz));
                        }
                        r[g] = up ? 0 : layers - 1;
                        cp[g] += up ? 1 : -1;
                    } else {
                        r[g] = rr;
                    }
                    const int lvl = (kReflect && (cp[g] & 1)) ? (layers - 1 - r[g]) : r[g];
                    bg = clad ? 0.0f : bcol[lvl];
                    const int plane = cp[g] * layers + r[g] + (up ? 1 : 0);
                    snp[g] = fmaf((float)plane - ze, fast_rcp(dz), 1.0f);
                }
                bsum[g] = fmaf(bg, sig1 - a, bsum[g]);
            }
        }
        s = snext;
        bool done = at_end;
        if (!done && Fetch::boundary(wexit)) {
            // leaving the mesh sideways: only rounding short of the end (a
            // sample point on the rim) counts as arrival
            if ((1.0f - s) * Dxy <= tol) done = true;
            else { ok = false; return; }
        }
        if (done) {
            const float stot = fminf(s, 1.0f);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const float dz = ze - ((float)(lv0 + g) + 0.5f);
                D[g] = sqrtf(d2 + dz * h * dz * h);
                out_bl[g] = bsum[g] * D[g];
                out_cl[g] = clad_s * D[g];
                out_gl[g] = (stot - clad_s) * D[g];
                if (kReflect) {
                    // an end point on a reflecting plane takes its last factors here
                    const float vz = dz * h / D[g];
                    const float sin2 = fmaxf(0.0f, 1.0f - vz * vz);
                    for (uint32_t n = (uint32_t)abs(cp[g]); n < nrefl; ++n) {
                        const uint32_t pl_ = (n & 1u) ? 1u - first : first;
                        out_rf[g] *= surface_factor(m, t, pl_, sin2);
                    }
                }
            }
            return;
        }
        // enter the neighbour through its edge f: L = vertex f, R = vertex f+1, N = vertex f+2
        const uint32_t tn = Fetch::next(wexit), f = Fetch::edge(wexit);
        const uint32_t e1 = f == 2u ? 0u : f + 1u, v = f == 0u ? 2u : f - 1u;
        const float2 pn = fetch.vertex(m, tn, v);
        uint32_t wr_n, wl_n;
        fetch.words(m, tn, e1, v, wr_n, wl_n, clad);
        const float X = pn.x - cx, Y = pn.y - cy;
        const float c = side_of(dx, dy, X, Y), a = along(dx, dy, X, Y);
        const bool left = c >= 0.0f;
        cl = left ? c : cl;
        wl = left ? a : wl;
        cr = left ? cr : c;
        wr = left ? wr : a;
        wexit = left ? wr_n : wl_n;
        const bool reach = (pn.x == px) & (pn.y == py);
        sxy = reach ? 1.0f : edge_sigma(cl, wl, cr, wr, inv_d2);
        t = tn;
    }
    ok = false;
}

// Importance values of one column item: prisms (t0, lv0 .. lv0 + nlev - 1)
// of reflection slice k for the sample at (px, py) on level `plane`; out[g]
// receives the entry of level lv0 + g (the transfer factor gain * surface
// factors when `transfer` is set).  Returns the prism steps walked.
template <class Fetch>
HASE_HD inline unsigned long long column_entries(const MeshDev& m, const Fetch& fetch, const float* c1, const float* c2,
                                                 uint32_t n_lambda, float inv_lambda, uint32_t t0, int lv0, int nlev,
                                                 uint32_t k, float px, float py, int plane, bool transfer,
                                                 float out[kLevelGroup])
{
    const uint32_t nr = (k + 1u) >> 1, fp = k & 1u;
    const int end_plane = image_plane(plane, (int)m.layers, nr, fp);
    float bl[kLevelGroup], gl[kLevelGroup], cl[kLevelGroup], rf[kLevelGroup], len[kLevelGroup];
    bool ok;
    unsigned long long steps = 0;
    if (nr == 0)
        column_importance<kGroupDirect, false>(m, fetch, t0, lv0, nlev, px, py, end_plane, 0u, fp, bl, gl, cl, rf,
                                               len, steps, ok);
    else
        column_importance<kGroupMirror, true>(m, fetch, t0, lv0, nlev, px, py, end_plane, nr, fp, bl, gl, cl, rf,
                                              len, steps, ok);
#pragma unroll
    for (int g = 0; g < kLevelGroup; ++g) {
        float value = 0.0f;
        if (g < nlev) {
            const uint32_t j = t0 + (uint32_t)(lv0 + g) * m.T;
            const float bv = m.bv[j];
            if (ok && (bv > 0.0f || transfer)) {
                WalkOut w;
                w.beta_len = bl[g]; w.gain_len = gl[g]; w.clad_len = cl[g];
                const float gain = spectral_gain(c1, c2, n_lambda, inv_lambda, m.clad_alpha, w);
                value = transfer ? gain * rf[g] : bv * gain * rf[g] / (4.0f * kPi * len[g] * len[g]);
                if (!(value >= 0.0f) || std::isinf(value)) value = 0.0f;
            }
        }
        out[g] = value;
    }
    return steps;
}

}  // namespace hase
