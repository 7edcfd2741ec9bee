// Host build of the device prism walk (haseongpu_amd/csrc/prism_walk.h) and
// the importance map's column walks (importance_walk.h) for CPU tests: the same source, compiled for the host, walked over meshes the
// test packs into the device record layouts (48-byte global records, and the
// 32-byte compact or 16-byte indexed records staged in LDS; lds_kind 0, 1
// or 2).  Test infrastructure only.
#include <algorithm>
#include <vector>

#include "../../haseongpu_amd/csrc/importance_walk.h"
#include "../../haseongpu_amd/csrc/prism_walk.h"

using namespace hase;

extern "C" int walk_rays_host(uint32_t T, uint32_t layers, float h, const uint32_t* rec, const uint32_t* lds_rec,
                              int lds_kind, const float* beta, const float* refl, const float* nidx, uint32_t n,
                              const float* start, const uint32_t* prism, const float* end, const uint32_t* nrefl,
                              const uint32_t* first, float* out, uint32_t* steps, int32_t* status)
{
    MeshDev m{};
    m.rec = reinterpret_cast<const uint4*>(rec);
    m.beta = beta;
    m.refl = refl;
    m.T = T;
    m.layers = layers;
    m.h = h;
    if (nidx) {
        m.has_nidx = 1;
        m.nin_b = nidx[0]; m.nout_b = nidx[1]; m.nin_t = nidx[2]; m.nout_t = nidx[3];
    }
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t t = prism[i] % T, l = prism[i] / T;
        WalkOut w;
        const uint4* staged = reinterpret_cast<const uint4*>(lds_rec);
        if (lds_kind == (int)kLdsCompact)
            w = walk_ray(m, start[3 * i], start[3 * i + 1], start[3 * i + 2], t, l, end[3 * i], end[3 * i + 1],
                         end[3 * i + 2], nrefl[i], first[i], staged_fetch<kLdsCompact>(staged, T));
        else if (lds_kind == (int)kLdsIndexed)
            w = walk_ray(m, start[3 * i], start[3 * i + 1], start[3 * i + 2], t, l, end[3 * i], end[3 * i + 1],
                         end[3 * i + 2], nrefl[i], first[i], staged_fetch<kLdsIndexed>(staged, T));
        else
            w = walk_ray(m, start[3 * i], start[3 * i + 1], start[3 * i + 2], t, l, end[3 * i], end[3 * i + 1],
                         end[3 * i + 2], nrefl[i], first[i]);
        out[5 * i + 0] = w.beta_len;
        out[5 * i + 1] = w.gain_len;
        out[5 * i + 2] = w.clad_len;
        out[5 * i + 3] = w.refl;
        out[5 * i + 4] = w.length;
        steps[i] = w.steps;
        status[i] = w.status;
    }
    return 0;
}

// Walked importance map of one sample (every reflection slice and prism, in
// the caller's prism order) by the importance kernel's column walks.
extern "C" int importance_host(uint32_t T, uint32_t layers, float h, const uint32_t* rec, const uint32_t* lds_rec,
                               int lds_kind, const float* beta, const float* bv, const float* refl, const float* nidx, float clad_alpha,
                               const float* c1, const float* c2, uint32_t n_lambda, float px, float py, int plane,
                               uint32_t R, int transfer, float* out, unsigned long long* steps)
{
    MeshDev m{};
    m.rec = reinterpret_cast<const uint4*>(rec);
    m.beta = beta;
    m.bv = bv;
    m.refl = refl;
    m.T = T;
    m.layers = layers;
    m.h = h;
    m.clad_alpha = clad_alpha;
    if (nidx) {
        m.has_nidx = 1;
        m.nin_b = nidx[0]; m.nout_b = nidx[1]; m.nin_t = nidx[2]; m.nout_t = nidx[3];
    }
    // column-major copy of beta, as the library keeps it for the column walks
    std::vector<float> col((size_t)T * layers);
    for (uint32_t l = 0; l < layers; ++l)
        for (uint32_t t = 0; t < T; ++t) col[(size_t)t * layers + l] = beta[t + (size_t)l * T];
    m.beta_col = col.data();
    const uint32_t P = T * layers;
    const ColumnItems ci(T, layers, R);
    *steps = 0;
    for (uint32_t it = 0; it < ci.per_sample(); ++it) {
        uint32_t t0, k;
        int lv0, nlev;
        ci.decode(it, t0, k, lv0, nlev);
        float v[kLevelGroup];
        const uint4* staged = reinterpret_cast<const uint4*>(lds_rec);
        if (lds_kind == (int)kLdsCompact)
            *steps += column_entries(m, staged_fetch<kLdsCompact>(staged, T), c1, c2, n_lambda, 1.0f / (float)n_lambda,
                                     t0, lv0, nlev, k, px, py, plane, transfer != 0, v);
        else if (lds_kind == (int)kLdsIndexed)
            *steps += column_entries(m, staged_fetch<kLdsIndexed>(staged, T), c1, c2, n_lambda, 1.0f / (float)n_lambda,
                                     t0, lv0, nlev, k, px, py, plane, transfer != 0, v);
        else
            *steps += column_entries(m, GlobalTriangles{}, c1, c2, n_lambda, 1.0f / (float)n_lambda, t0, lv0, nlev, k,
                                     px, py, plane, transfer != 0, v);
        for (int g = 0; g < nlev; ++g) out[(size_t)k * P + t0 + (size_t)(lv0 + g) * T] = v[g];
    }
    return 0;
}
