// tet_fold.hip - host-only: prism arrays from the Tet4 cells of a 2.2.0
// HostMesh (three tetrahedra per wedge), the arrays hase_mesh_t takes.
//
// The split being undone is utils/convert_vtk_topology.py:38-53: wedge
// w = t + l*T with lower corners in ascending global order (a, b, c) and the
// corners above them (d, e, f) becomes cells 3w .. 3w+2.  Which of the three
// cells carries the base or the top face is found from the vertex levels,
// not assumed.  Reflectivity and refractive-index lookups follow
// include/core/mesh.hpp:300-332.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hase_ase.h"
#include "common.h"

namespace hase {
namespace {

struct Tet4Lookup {
    const hase_tet4_host_mesh_t& m;

    int surface(uint32_t cell, uint32_t face) const
    {
        return m.cell_face_boundaries ? m.cell_face_boundaries[4u * cell + face] : 0;
    }
    float reflectivity(uint32_t cell, uint32_t face) const
    {
        const int id = surface(cell, face);
        if (id > 0 && (uint32_t)id < m.number_of_surfaces && m.surface_reflectivities)
            return m.surface_reflectivities[id];
        const uint32_t legacy = cell * 2u + face % 2u;
        return (m.reflectivities && legacy < m.number_of_reflectivities) ? m.reflectivities[legacy] : 0.0f;
    }
    float index(uint32_t cell, uint32_t face, bool inside) const
    {
        const int id = surface(cell, face);
        const float* table = inside ? m.surface_refractive_index_inside : m.surface_refractive_index_outside;
        if (id > 0 && (uint32_t)id < m.number_of_surfaces && table) return table[id];
        const uint32_t legacy = 2u * (face % 2u) + (inside ? 0u : 1u);
        return m.refractive_indices ? m.refractive_indices[legacy] : 1.0f;
    }
};

double tet_volume(const double* P, uint32_t Nm, const uint32_t* v)
{
    double d[3][3];
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) d[k][c] = P[v[k + 1] + c * Nm] - P[v[0] + c * Nm];
    const double det = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
                       d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
                       d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
    return std::fabs(det) / 6.0;
}

void fold(const hase_tet4_host_mesh_t& m, hase_prism_arrays_t& out)
{
    const uint32_t Np = m.number_of_points, L = m.number_of_levels, C = m.number_of_cells;
    const uint32_t Nm = m.number_of_mesh_points;
    if (!m.points || !m.cell_point_indices || !m.beta_volume)
        throw HaseError(HASE_E_INVALID, "points, cell_point_indices and beta_volume are required");
    if (L < 2 || Np < 3) throw HaseError(HASE_E_INVALID, "need at least 2 levels of at least 3 points");
    if ((uint64_t)Np * L != Nm)
        throw HaseError(HASE_E_INVALID, "number_of_mesh_points must be number_of_points * number_of_levels "
                                        "(sample points are the mesh points of an extruded slab)");
    if (C == 0 || C % (3u * (L - 1u)))
        throw HaseError(HASE_E_INVALID, "number_of_cells must be 3 cells per wedge times (levels - 1) layers");
    const uint32_t W = C / 3u, T = W / (L - 1u);
    if (!out.points || !out.triangle_point_indices || !out.triangle_neighbors || !out.forbidden_edge ||
        !out.beta_volume)
        throw HaseError(HASE_E_INVALID, "points, triangle_point_indices, triangle_neighbors, forbidden_edge and "
                                        "beta_volume outputs are required");
    const double* X = m.points;
    const double* Y = m.points + Nm;
    const double* Z = m.points + 2u * (size_t)Nm;
    // every level repeats the transverse points; the planes are equally spaced
    const double h = Z[Np] - Z[0];
    if (!(h > 0.0)) throw HaseError(HASE_E_INVALID, "level 1 must lie above level 0");
    for (uint32_t l = 0; l < L; ++l)
        for (uint32_t p = 0; p < Np; ++p) {
            const uint32_t i = p + l * Np;
            if (X[i] != X[p] || Y[i] != Y[p])
                throw HaseError(HASE_E_INVALID, "point " + std::to_string(i) + " is not point " +
                                                    std::to_string(p) + " moved up " + std::to_string(l) + " levels");
            const double want = Z[0] + l * h;
            if (std::fabs(Z[i] - want) > 1e-9 * (1.0 + std::fabs(want)))
                throw HaseError(HASE_E_INVALID, "level " + std::to_string(l) + " is not a plane at " +
                                                    std::to_string(l) + " * thickness");
        }
    std::vector<uint32_t> base(3u * (size_t)T);
    Tet4Lookup look{m};
    for (uint32_t w = 0; w < W; ++w) {
        const uint32_t t = w % T, l = w / T;
        const uint32_t* v = m.cell_point_indices + 12u * (size_t)w;
        uint32_t lower[3], upper[3], nl = 0, nu = 0;
        for (int k = 0; k < 12; ++k) {
            const uint32_t i = v[k];
            if (i >= Nm) throw HaseError(HASE_E_INVALID, "cell point index out of range");
            const uint32_t lv = i / Np;
            uint32_t* set = lv == l ? lower : lv == l + 1u ? upper : nullptr;
            uint32_t& n = lv == l ? nl : nu;
            if (!set)
                throw HaseError(HASE_E_INVALID, "cells " + std::to_string(3u * w) + ".." + std::to_string(3u * w + 2u) +
                                                    " are not wedge " + std::to_string(w) + " (triangle t + level l * T)");
            if (std::find(set, set + n, i % Np) == set + n) {
                if (n == 3) throw HaseError(HASE_E_INVALID, "a wedge's cells span more than six points");
                set[n++] = i % Np;
            }
        }
        if (nl != 3 || nu != 3) throw HaseError(HASE_E_INVALID, "a wedge's cells must span six points");
        std::sort(lower, lower + 3);
        std::sort(upper, upper + 3);
        if (!std::equal(lower, lower + 3, upper))
            throw HaseError(HASE_E_INVALID, "a wedge's top corners must lie above its base corners");
        if (l == 0) {
            std::copy(lower, lower + 3, &base[3u * (size_t)t]);
        } else if (!std::equal(lower, lower + 3, &base[3u * (size_t)t])) {
            throw HaseError(HASE_E_INVALID, "wedge " + std::to_string(w) + " is not above triangle " +
                                                std::to_string(t));
        }
        // prism beta: the cells' common value, else their volume-weighted mean
        const double b0 = m.beta_volume[3u * w];
        double bsum = 0.0, vsum = 0.0;
        bool same = true;
        for (uint32_t c = 3u * w; c < 3u * w + 3u; ++c) {
            same = same && m.beta_volume[c] == b0;
            const double vol = tet_volume(m.points, Nm, m.cell_point_indices + 4u * (size_t)c);
            bsum += m.beta_volume[c] * vol;
            vsum += vol;
        }
        out.beta_volume[w] = same ? b0 : bsum / vsum;
        if (m.cladding_cell_types) {
            const uint32_t type = m.cladding_cell_types[3u * (size_t)t];
            for (uint32_t c = 3u * w; c < 3u * w + 3u; ++c)
                if (m.cladding_cell_types[c] != type)
                    throw HaseError(HASE_E_INVALID, "triangle " + std::to_string(t) +
                                                        " has cells of different cladding types; the prism "
                                                        "mesh keeps one type per triangle");
        }
    }
    for (uint32_t p = 0; p < Np; ++p) {
        out.points[p] = X[p];
        out.points[Np + p] = Y[p];
    }
    // counter-clockwise, starting at the lowest point index
    for (uint32_t t = 0; t < T; ++t) {
        uint32_t a = base[3u * t], b = base[3u * t + 1u], c = base[3u * t + 2u];
        const double cross = (X[b] - X[a]) * (Y[c] - Y[a]) - (Y[b] - Y[a]) * (X[c] - X[a]);
        if (cross == 0.0) throw HaseError(HASE_E_INVALID, "triangle " + std::to_string(t) + " has no area");
        if (cross < 0.0) std::swap(b, c);
        out.triangle_point_indices[t] = a;
        out.triangle_point_indices[t + T] = b;
        out.triangle_point_indices[t + 2u * T] = c;
        if (out.cladding_cell_types && m.cladding_cell_types)
            out.cladding_cell_types[t] = m.cladding_cell_types[3u * (size_t)t];
        else if (out.cladding_cell_types)
            out.cladding_cell_types[t] = 0u;
    }
    // neighbours: edge k joins vertex k and k+1; match the sorted point pairs
    std::vector<std::pair<uint64_t, uint32_t>> edges(3u * (size_t)T);
    for (uint32_t t = 0; t < T; ++t)
        for (uint32_t k = 0; k < 3; ++k) {
            const uint64_t p = out.triangle_point_indices[t + k * T];
            const uint64_t q = out.triangle_point_indices[t + ((k + 1u) % 3u) * T];
            edges[3u * t + k] = {std::min(p, q) << 32 | std::max(p, q), 3u * t + k};
            out.triangle_neighbors[t + k * T] = -1;
            out.forbidden_edge[t + k * T] = -1;
        }
    std::sort(edges.begin(), edges.end());
    for (size_t i = 0; i + 1 < edges.size(); ++i) {
        if (edges[i].first != edges[i + 1].first) continue;
        if (i + 2 < edges.size() && edges[i + 2].first == edges[i].first)
            throw HaseError(HASE_E_INVALID, "an edge is shared by more than two triangles");
        const uint32_t e0 = edges[i].second, e1 = edges[i + 1].second;
        const uint32_t t0 = e0 / 3u, k0 = e0 % 3u, t1 = e1 / 3u, k1 = e1 % 3u;
        out.triangle_neighbors[t0 + k0 * T] = (int32_t)t1;
        out.forbidden_edge[t0 + k0 * T] = (int32_t)k1;
        out.triangle_neighbors[t1 + k1 * T] = (int32_t)t0;
        out.forbidden_edge[t1 + k1 * T] = (int32_t)k0;
        ++i;
    }
    // surface data of the two planes: the face of the wedge's cells whose three
    // corners all lie on the plane
    auto plane_face = [&](uint32_t w, uint32_t level, uint32_t& cell, uint32_t& face) {
        for (uint32_t c = 3u * w; c < 3u * w + 3u; ++c)
            for (uint32_t f = 0; f < 4u; ++f) {
                bool on = true;
                for (uint32_t k = 0; k < 4u; ++k)
                    if (k != f && m.cell_point_indices[4u * (size_t)c + k] / Np != level) on = false;
                if (on) {
                    cell = c;
                    face = f;
                    return;
                }
            }
        throw HaseError(HASE_E_INVALID, "wedge " + std::to_string(w) + " has no face on level " + std::to_string(level));
    };
    float nidx[4] = {0, 0, 0, 0};
    for (uint32_t t = 0; t < T; ++t) {
        uint32_t cb, fb, ct, ft;
        plane_face(t, 0u, cb, fb);
        plane_face(t + (L - 2u) * T, L - 1u, ct, ft);
        if (out.reflectivities) {
            out.reflectivities[2u * t] = look.reflectivity(cb, fb);
            out.reflectivities[2u * t + 1u] = look.reflectivity(ct, ft);
        }
        const float here[4] = {look.index(cb, fb, true), look.index(cb, fb, false), look.index(ct, ft, true),
                               look.index(ct, ft, false)};
        if (t == 0) std::copy(here, here + 4, nidx);
        else if (!std::equal(here, here + 4, nidx))
            throw HaseError(HASE_E_INVALID, "refractive indices differ between triangles 0 and " + std::to_string(t) +
                                                "; the prism mesh takes one pair per plane");
    }
    std::copy(nidx, nidx + 4, out.refractive_indices);
    out.thickness = h;
    out.number_of_triangles = T;
}

}  // namespace
}  // namespace hase

extern "C" int hase_fold_tet4_mesh(const hase_tet4_host_mesh_t* mesh, hase_prism_arrays_t* out)
{
    return hase::guarded([&] {
        if (!mesh || !out) throw hase::HaseError(HASE_E_INVALID, "mesh and out must not be NULL");
        hase::fold(*mesh, *out);
    });
}
