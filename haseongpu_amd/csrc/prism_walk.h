// prism_walk.h - FP32 prism walk of the ASE ray tracer, shared by the HIP
// kernels and host-side checks (the functions are __host__ __device__).
//
// The walk replaces the legacy propagateRay (retired: src/kernels/propagateRay.cpp
// is an empty translation unit) and follows the gain and cladding rules of
// include/kernels/forward/rayWalk.hpp:170-208 and the surface reflectance of
// include/kernels/reflection.hpp:21-47.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#define HASE_HD __host__ __device__

namespace hase {

constexpr uint32_t kKeyTag = 0x48415345u;
constexpr uint32_t kTagRay = 0u;
constexpr uint32_t kTagOffset = 1u;
constexpr uint32_t kMaxSteps = 10000u;
constexpr float kPi = 3.14159265358979f;

HASE_HD inline float bits_to_float(uint32_t u) { return __builtin_bit_cast(float, u); }

// Triangle records, 48 B per triangle t at byte 48 t:
//   {x0, y0, x1, y1}  {x2, y2, n0, n1}  {n2, v0 | cladding << 31, v1, v2}
// with vertex coordinates as FP32 bits (vertex k at byte 8 k), v_k the mesh
// point index of vertex k and n_k the neighbour word of edge k (vertex k to
// k+1): the neighbour << 2 | the shared edge's index inside the neighbour in
// bits 0-30 (all ones at the boundary) and this triangle's cladding flag in
// bit 31, so every neighbour word read also tells the walk what its current
// triangle is made of.
// Compact records, 32 B per triangle for staging the mesh in LDS:
//   {x0, y0, x1, y1}  {x2, y2, W & 0xFFFFFFFF, W >> 32}
// with W = f0 | f1 << 21 | f2 << 42 | cladding << 63 and f_k = neighbour << 2 |
// shared edge (21 bits: neighbour < 2^19) or 0x1FFFFF at the boundary.
// Indexed records, for meshes beyond kLdsTriangles: 16 B per triangle,
//   {v0 | v1 << 16, v2 | cladding << 31, W & 0xFFFFFFFF, W >> 32}
// with v_k the point index of vertex k (< 2^16) and W as above, followed by
// the vertex table, the FP32 (x, y) of every mesh point, two per 16 bytes.
// One more LDS read per new vertex than the compact records, but a Delaunay
// disk (about T / 2 points) fits up to ~8,100 triangles instead of 4800.
constexpr uint32_t kLdsTriangles = 4800;     // compact records: 150 KiB of LDS
constexpr uint32_t kLdsStageWords = 10176;   // 16-byte words of LDS a staged kernel holds (159 KiB)
constexpr uint32_t kLdsCompact = 1u, kLdsIndexed = 2u;
// Meshes that fit neither LDS format keep the compact records in global
// memory: the walk reads one vertex and one neighbour word per step from
// them (two loads, 32 B per triangle) instead of three loads from the
// 48-byte records.
constexpr uint32_t kCompactGlobal = 3u;

struct MeshDev {
    const uint4* rec;      // 3 records per triangle
    const uint4* lds_rec;  // records of kind lds_kind (staged in LDS, or read from global memory), or null
    uint32_t lds_kind;     // 0: none, kLdsCompact, kLdsIndexed, kCompactGlobal
    const uint4* cmp_rec;  // the compact records in global memory (meshes below 2^19 triangles), or null
    uint32_t lds_words;    // 16-byte words to stage
    const float* beta;     // [P] level-major: t + l * T, zero in cladding prisms
    const float* beta_col; // [P] column-major: t * layers + l, zero in cladding prisms
    const float* bv;       // [P] beta * prism volume, level-major
    const float* refl;     // [2T] or nullptr
    float nin_b, nout_b, nin_t, nout_t;
    int has_nidx;
    uint32_t T, layers, np;
    float h, clad_alpha;
    float steps_per_cm;    // estimated triangle crossings per cm of xy path (1.2 sqrt(T / mesh area))
};

struct WalkOut {
    float beta_len, gain_len, clad_len, refl, length;
    uint32_t steps;
    int status;
};

HASE_HD inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// (2k+1) * 2^-24 from the top 23 bits: exact in FP32, strictly inside (0,1)
HASE_HD inline float u32_to_unit(uint32_t x)
{
    return (float)(((x >> 9) << 1) | 1u) * (1.0f / 16777216.0f);
}

HASE_HD inline double systematic_offset(uint32_t seed, uint32_t sample, uint32_t launch)
{
    uint32_t r[4];
    philox4x32_10(0xFFFFFFFFu, sample, launch, kTagOffset, seed, kKeyTag, r);
    return ((double)(r[0] >> 5) * 67108864.0 + (double)(r[1] >> 6)) * (1.0 / 9007199254740992.0);
}

HASE_HD inline float surface_factor(const MeshDev& m, uint32_t t, uint32_t plane, float sin2)
{
    if (m.has_nidx) {
        float nin = plane ? m.nin_t : m.nin_b;
        float nout = plane ? m.nout_t : m.nout_b;
        if (nin > 0.0f && nout > 0.0f) {
            float ratio = nin / nout;
            if (ratio * ratio * sin2 > 1.0f) return 1.0f;
        }
    }
    if (!m.refl) return 0.0f;
    float r = m.refl[2u * t + plane];
    return r > 0.0f ? r : 0.0f;
}

constexpr uint32_t kNoPoint = 0xFFFFFFFFu;

// a * b for a, b < 2^24 (full-rate 24-bit multiply on the device); triangle
// and level counts are kept below 2^24
HASE_HD inline uint32_t mul24(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    return a * b;
#endif
}

// The three records of triangle t.  Offsets are formed in 32 bits from the
// table base so that the loads use the scalar-base, 32-bit-offset form.
HASE_HD inline void load_triangle(const MeshDev& m, uint32_t t, uint4& q0, uint4& q1, uint4& q2)
{
    const char* base = reinterpret_cast<const char*>(m.rec);
    const uint32_t off = mul24(t, 48u);
    q0 = *reinterpret_cast<const uint4*>(base + off);
    q1 = *reinterpret_cast<const uint4*>(base + off + 16u);
    q2 = *reinterpret_cast<const uint4*>(base + off + 32u);
}

HASE_HD inline void load_corners(const MeshDev& m, uint32_t t, uint4& q0, uint4& q1)
{
    const char* base = reinterpret_cast<const char*>(m.rec);
    const uint32_t off = mul24(t, 48u);
    q0 = *reinterpret_cast<const uint4*>(base + off);
    q1 = *reinterpret_cast<const uint4*>(base + off + 16u);
}

HASE_HD inline float load_f32(const float* table, uint32_t index)
{
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(table) + index * 4u);
}

HASE_HD inline uint32_t record_cladding(const uint4& q2) { return q2.y >> 31; }

// Neighbour word of the 48-byte records: next triangle, its entry edge,
// boundary test (bit 31 is the owning triangle's cladding flag).
HASE_HD inline uint32_t nw_next(uint32_t w) { return (w >> 2) & 0x1FFFFFFFu; }
HASE_HD inline uint32_t nw_edge(uint32_t w) { return w & 3u; }
HASE_HD inline bool nw_boundary(uint32_t w) { return (w & 0x7FFFFFFFu) == 0x7FFFFFFFu; }

// A fetch gives the walk one triangle in the layout of the 48-byte records,
// as three vectors (kept as separate values so that the walk's selects stay
// selects of values, not of addresses into a stack copy):
//   a = {x0, y0, x1, y1}, b = {x2, y2, nw0, nw1}, c = {nw2, v0 | cladding << 31, v1, v2}
// with nw_k = neighbour << 2 | shared edge index, or the fetch's kBoundary.

// From the 48-byte records in global memory; the end point is identified by
// its mesh point index pe.
struct GlobalTriangles {
    static constexpr uint32_t kBoundary = 0xFFFFFFFFu;   // neighbour word at the mesh boundary
    HASE_HD void operator()(const MeshDev& m, uint32_t t, uint4& a, uint4& b, uint4& c) const
    {
        load_triangle(m, t, a, b, c);
    }
    // side-test walk: vertex v of triangle t
    HASE_HD float2 vertex(const MeshDev& m, uint32_t t, uint32_t v) const
    {
        const char* base = reinterpret_cast<const char*>(m.rec);
        return *reinterpret_cast<const float2*>(base + mul24(t, 48u) + 8u * v);
    }
    // neighbour words of edges e1 and e2 of t, and t's cladding flag
    HASE_HD void words(const MeshDev& m, uint32_t t, uint32_t e1, uint32_t e2, uint32_t& w1, uint32_t& w2,
                       uint32_t& clad) const
    {
        const char* base = reinterpret_cast<const char*>(m.rec) + mul24(t, 48u) + 24u;
        w1 = *reinterpret_cast<const uint32_t*>(base + 4u * e1);
        w2 = *reinterpret_cast<const uint32_t*>(base + 4u * e2);
        clad = w1 >> 31;
    }
    HASE_HD static uint32_t next(uint32_t w) { return nw_next(w); }
    HASE_HD static uint32_t edge(uint32_t w) { return nw_edge(w); }
    HASE_HD static bool boundary(uint32_t w) { return nw_boundary(w); }
    // the end is a corner other than the one facing the exit edge
    HASE_HD static int touches(const uint4&, const uint4&, const uint4& c, bool c1, bool c2, uint32_t pe, float,
                               float)
    {
        const uint32_t v0 = c.y & 0x7FFFFFFFu, v1 = c.z, v2 = c.w;
        uint32_t opposite = c1 ? v0 : v2;
        opposite = c2 ? v1 : opposite;
        const int vertex = (int)(v0 == pe) | (int)(v1 == pe) | (int)(v2 == pe);
        return vertex & (int)(opposite != pe);
    }
};

// From the 32-byte compact records staged in LDS; the end point is
// identified by its stored FP32 coordinates (ex, ey), which equal the
// corners' of the same mesh point bit for bit.
struct LdsTriangles {
    static constexpr uint32_t kBoundary = 0x1FFFFFu;
    const uint4* rec;
    HASE_HD void operator()(const MeshDev&, uint32_t t, uint4& a, uint4& b, uint4& c) const
    {
        const uint4 q1 = rec[2u * t + 1u];
        a = rec[2u * t];
        b = make_uint4(q1.x, q1.y, q1.z & 0x1FFFFFu, ((q1.z >> 21) | (q1.w << 11)) & 0x1FFFFFu);
        c = make_uint4((q1.w >> 10) & 0x1FFFFFu, q1.w & 0x80000000u, 0u, 0u);
    }
    HASE_HD float2 vertex(const MeshDev&, uint32_t t, uint32_t v) const
    {
        return reinterpret_cast<const float2*>(rec)[4u * t + v];
    }
    HASE_HD void words(const MeshDev&, uint32_t t, uint32_t e1, uint32_t e2, uint32_t& w1, uint32_t& w2,
                       uint32_t& clad) const
    {
        const uint64_t W = reinterpret_cast<const uint64_t*>(rec)[4u * t + 3u];
        w1 = (uint32_t)(W >> (21u * e1)) & 0x1FFFFFu;
        w2 = (uint32_t)(W >> (21u * e2)) & 0x1FFFFFu;
        clad = (uint32_t)(W >> 63);
    }
    HASE_HD static uint32_t next(uint32_t w) { return w >> 2; }
    HASE_HD static uint32_t edge(uint32_t w) { return w & 3u; }
    HASE_HD static bool boundary(uint32_t w) { return w == 0x1FFFFFu; }
    HASE_HD static int touches(const uint4& a, const uint4& b, const uint4&, bool c1, bool c2, uint32_t, float ex,
                               float ey)
    {
        const float x0 = bits_to_float(a.x), y0 = bits_to_float(a.y);
        const float x1 = bits_to_float(a.z), y1 = bits_to_float(a.w);
        const float x2 = bits_to_float(b.x), y2 = bits_to_float(b.y);
        float ox = c1 ? x0 : x2, oy = c1 ? y0 : y2;
        ox = c2 ? x1 : ox;
        oy = c2 ? y1 : oy;
        const int vertex = ((int)(x0 == ex) & (int)(y0 == ey)) | ((int)(x1 == ex) & (int)(y1 == ey)) |
                           ((int)(x2 == ex) & (int)(y2 == ey));
        return vertex & ((int)(ox != ex) | (int)(oy != ey));
    }
};

// From the 16-byte indexed records and the vertex table staged in LDS
// (vtx = the table, right after the T records); words, decoders and the end
// test are those of the compact records.
struct LdsIndexedTriangles : LdsTriangles {
    const float2* vtx;
    HASE_HD uint32_t point(uint32_t t, uint32_t v) const
    {
        const uint2 q = reinterpret_cast<const uint2*>(rec)[2u * t];
        const uint32_t i01 = v == 1u ? q.x >> 16 : q.x & 0xFFFFu;
        return v == 2u ? (q.y & 0xFFFFu) : i01;
    }
    HASE_HD void operator()(const MeshDev&, uint32_t t, uint4& a, uint4& b, uint4& c) const
    {
        const uint4 q = rec[t];
        const float2 p0 = vtx[q.x & 0xFFFFu], p1 = vtx[q.x >> 16], p2 = vtx[q.y & 0xFFFFu];
        a = make_uint4(__builtin_bit_cast(uint32_t, p0.x), __builtin_bit_cast(uint32_t, p0.y),
                       __builtin_bit_cast(uint32_t, p1.x), __builtin_bit_cast(uint32_t, p1.y));
        b = make_uint4(__builtin_bit_cast(uint32_t, p2.x), __builtin_bit_cast(uint32_t, p2.y), q.z & 0x1FFFFFu,
                       ((q.z >> 21) | (q.w << 11)) & 0x1FFFFFu);
        c = make_uint4((q.w >> 10) & 0x1FFFFFu, q.w & 0x80000000u, 0u, 0u);
    }
    HASE_HD float2 vertex(const MeshDev&, uint32_t t, uint32_t v) const { return vtx[point(t, v)]; }
    HASE_HD void words(const MeshDev&, uint32_t t, uint32_t e1, uint32_t e2, uint32_t& w1, uint32_t& w2,
                       uint32_t& clad) const
    {
        const uint64_t W = reinterpret_cast<const uint64_t*>(rec)[2u * t + 1u];
        w1 = (uint32_t)(W >> (21u * e1)) & 0x1FFFFFu;
        w2 = (uint32_t)(W >> (21u * e2)) & 0x1FFFFFu;
        clad = (uint32_t)(W >> 63);
    }
};

// Fetch of records staged in LDS at s (kind kLdsCompact or kLdsIndexed: T
// records, then the indexed format's vertex table).
template <uint32_t kKind>
using StagedFetch = std::conditional_t<kKind == kLdsIndexed, LdsIndexedTriangles, LdsTriangles>;

template <uint32_t kKind>
HASE_HD inline StagedFetch<kKind> staged_fetch(const uint4* s, uint32_t T)
{
    if constexpr (kKind == kLdsIndexed)
        return LdsIndexedTriangles{{s}, reinterpret_cast<const float2*>(s + T)};
    else
        return LdsTriangles{s};
}

// Exit edge among the candidates s_k (edge k joins corners k and k+1): its
// parameter, the neighbour word behind it, and whether it ends at the end
// point.
struct ExitEdge {
    float s;
    uint32_t nw;
    int touches;
};

template <class Fetch>
HASE_HD inline ExitEdge exit_edge(const uint4& a, const uint4& b, const uint4& c, float s0, float s1, float s2,
                                  uint32_t pe, float ex, float ey)
{
    ExitEdge x{s0, b.z, 0};
    const bool c1 = s1 < x.s;
    x.s = c1 ? s1 : x.s;
    x.nw = c1 ? b.w : x.nw;
    const bool c2 = s2 < x.s;
    x.s = c2 ? s2 : x.s;
    x.nw = c2 ? c.x : x.nw;
    x.touches = Fetch::touches(a, b, c, c1, c2, pe, ex, ey);
    return x;
}

// 1/x with the hardware reciprocal (1 ulp) on the device.
HASE_HD inline float fast_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// Prism walk from a (inside triangle t, layer `layer` of copy 0) to the
// unfolded end point e, parametrised by sigma in [0, 1] along the straight
// line (xy and height are both linear in sigma; lengths are sigma * D).
// Heights are in layer units (zeta = z / h), so an end point on a level plane
// (every sample point and its mirror images) sits on an exact integer plane
// whose crossing is the end of the walk.
//
// Triangle steps use side tests against the line through a and e: with
// d = e - a in xy, a vertex p has side c(p) = d x (p - a) (left of the line
// when >= 0) and along-line coordinate w(p) = d . (p - a).  The walk carries
// the two ends of the edge it entered through, L (left) and R (right), with
// their c and w; in the next triangle only the third vertex N is new: left
// of the line, the line leaves through edge R-N, else through N-L, and it
// crosses that edge where c vanishes, at
//   sigma = (c_L w_R - c_R w_L) / ((c_L - c_R) |d|^2).
// One vertex and one side test per triangle instead of three edge
// intersections.  A triangle whose new vertex is the end point's (x, y) is
// the last one (the line reaches it at sigma = 1, exactly).  Level-plane
// crossings are steps of their own; reflections are the copy boundaries of
// the unfolded height axis.
//
// The walk is resumable: walk_begin sets up a WalkState and walk_advance
// takes up to `budget` prism steps, so a lane can hand a finished ray back
// and start another while its neighbours are still walking.
struct WalkState {
    float ax, ay, za;         // start (x, y in cm; height in layer units)
    float ex, ey;             // end (x, y)
    float dx, dy, inv_d2;     // e - a in xy and 1 / |e - a|^2
    float D, inv_dz, tol, sin2;
    float ze;                 // end height (layer units)
    int zdir;
    uint32_t nrefl, first;
    uint32_t t;               // current triangle
    uint32_t clad;            // its cladding flag
    float cl, wl, cr, wr;     // side and along-line values of the entry edge's ends L and R
    float sxy;                // sigma where the line leaves t (1: the end lies in t)
    uint32_t wexit;           // neighbour word across that edge
    int u, copy, r;           // unfolded layer, mirror copy, layer in copy
    float sig;                // sigma reached
    float beta;               // beta of the current prism
    uint32_t applied;         // surface factors applied
    float beta_sig, clad_sig, refl;   // sum beta * dsigma, sum dsigma over cladding, surface factors
    uint32_t steps;
    int status;               // 0 ok, 1 left the mesh, 2 step limit
    int done;
};

HASE_HD inline float side_of(float dx, float dy, float x, float y) { return dx * y - dy * x; }
HASE_HD inline float along(float dx, float dy, float x, float y) { return dx * x + dy * y; }

// sigma at which the line crosses the edge from L (c >= 0) to R (c < 0)
HASE_HD inline float edge_sigma(float cl, float wl, float cr, float wr, float inv_d2)
{
    return (cl * wr - cr * wl) * inv_d2 * fast_rcp(cl - cr);
}

template <bool kReflect>
HASE_HD inline int folded_level(int u, int copy, int r, int layers)
{
    return kReflect ? ((copy & 1) ? (layers - 1 - r) : r) : u;
}

// Starts a walk in triangle t (corners p0, p1, p2 as the fetch returns them).
template <class Fetch>
HASE_HD inline void walk_begin(const MeshDev& m, WalkState& w, float ax, float ay, float za, uint32_t t,
                               uint32_t layer, float2 p0, float2 p1, float2 p2, float ex, float ey, float ze,
                               uint32_t nrefl, uint32_t first, const Fetch& fetch)
{
    const float h = m.h;
    w.beta_sig = 0.0f; w.clad_sig = 0.0f; w.refl = 1.0f;
    w.steps = 0; w.status = 0; w.done = 0;
    w.ax = ax; w.ay = ay; w.za = za; w.ex = ex; w.ey = ey;
    w.nrefl = nrefl; w.first = first;
    w.t = t;
    w.u = (int)layer;
    w.copy = 0;
    w.r = (int)layer;
    w.sig = 0.0f;
    w.applied = 0;
    const float dx = ex - ax, dy = ey - ay, dzeta = ze - za, dz = dzeta * h;
    const float D = sqrtf(dx * dx + dy * dy + dz * dz);
    w.D = D;
    w.dx = dx;
    w.dy = dy;
    if (!(D > 0.0f)) {
        w.done = 1;
        return;
    }
    const float d2 = dx * dx + dy * dy;
    w.inv_d2 = d2 > 0.0f ? 1.0f / d2 : 0.0f;
    // sin^2 of the angle to the plane normal, as 1 - cos^2: never above 1,
    // so a ratio of refractive indices of 1 never reflects totally (d2 / D^2
    // can round above 1 for grazing rays)
    const float vz = dz / D;
    w.sin2 = fmaxf(0.0f, 1.0f - vz * vz);
    w.tol = 2e-5f * (1.0f + D);
    w.inv_dz = dzeta != 0.0f ? 1.0f / dzeta : 0.0f;
    w.zdir = dzeta > 0.0f ? 1 : (dzeta < 0.0f ? -1 : 0);
    w.ze = ze;
    // exit of the start triangle: the edge whose start vertex is right of the
    // line and whose end vertex is left of it (counter-clockwise order)
    const float c0 = side_of(dx, dy, p0.x - ax, p0.y - ay), c1 = side_of(dx, dy, p1.x - ax, p1.y - ay);
    const float c2 = side_of(dx, dy, p2.x - ax, p2.y - ay);
    const bool b0 = c0 >= 0.0f, b1 = c1 >= 0.0f, b2 = c2 >= 0.0f;
    int k = (!b0 && b1) ? 0 : (!b1 && b2) ? 1 : (!b2 && b0) ? 2 : -1;
    if (k < 0) {
        // every vertex on one side (the line grazes a vertex, or rounding):
        // leave past the vertex furthest to the right, or enter left of the
        // one furthest to the left
        if (b0) {
            k = (c0 <= c1 && c0 <= c2) ? 0 : (c1 <= c2 ? 1 : 2);
        } else {
            const int km = (c0 >= c1 && c0 >= c2) ? 0 : (c1 >= c2 ? 1 : 2);
            k = km == 0 ? 2 : km - 1;
        }
    }
    const float2 pr = k == 0 ? p0 : (k == 1 ? p1 : p2);
    const float2 pl = k == 0 ? p1 : (k == 1 ? p2 : p0);
    w.cr = k == 0 ? c0 : (k == 1 ? c1 : c2);
    w.cl = k == 0 ? c1 : (k == 1 ? c2 : c0);
    w.wr = along(dx, dy, pr.x - ax, pr.y - ay);
    w.wl = along(dx, dy, pl.x - ax, pl.y - ay);
    uint32_t wk, unused;
    fetch.words(m, t, (uint32_t)k, (uint32_t)k, wk, unused, w.clad);
    w.wexit = wk;
    const bool at_end = (p0.x == ex && p0.y == ey) || (p1.x == ex && p1.y == ey) || (p2.x == ex && p2.y == ey) ||
                        d2 == 0.0f;
    w.sxy = at_end ? 1.0f : edge_sigma(w.cl, w.wl, w.cr, w.wr, w.inv_d2);
    w.beta = load_f32(m.beta, t + mul24(layer, m.T));
}

// kReflect = false is the direct walk (start and end inside the slab), which
// needs neither the mirror folding nor surface factors.
template <bool kReflect, class Fetch = GlobalTriangles>
HASE_HD inline void walk_advance(const MeshDev& m, WalkState& w, uint32_t budget, const Fetch& fetch = Fetch())
{
    const int layers = (int)m.layers;
    for (uint32_t it = 0; it < budget && !w.done; ++it) {
        if (w.steps >= kMaxSteps) {
            w.status = 2;
            w.done = 1;
            return;
        }
        w.steps++;
        // next level plane along the line, measured back from the end: the
        // end's own plane (every sample point and mirror image lies on one) is
        // reached at sigma = 1 exactly, and a level ray (inv_dz = 0) meets no
        // plane before the end
        const int plane = w.u + (w.zdir > 0 ? 1 : 0);
        const float sz = fmaf((float)plane - w.ze, w.inv_dz, 1.0f);
        const bool zstep = sz <= w.sxy;
        float snext = fminf(fminf(w.sxy, sz), 1.0f);
        snext = fmaxf(snext, w.sig);
        const float seg = snext - w.sig;
        w.clad_sig = fmaf((float)w.clad, seg, w.clad_sig);
        // the walk's beta copy is zero in cladding prisms
        w.beta_sig = fmaf(w.beta, seg, w.beta_sig);
        w.sig = snext;
        if (w.sig >= 1.0f) {
            if (kReflect) {
                // an end point on a reflecting plane takes its last factor here
                while (w.applied < w.nrefl) {
                    const uint32_t pl = (w.applied & 1u) ? 1u - w.first : w.first;
                    w.refl *= surface_factor(m, w.t, pl, w.sin2);
                    w.applied++;
                }
            }
            w.done = 1;
            return;
        }
        if (zstep) {
            if (kReflect) {
                if (w.zdir > 0) {
                    if (++w.r == layers) {
                        if (w.applied < w.nrefl) {
                            w.refl *= surface_factor(m, w.t, (w.copy & 1) ? 0u : 1u, w.sin2);
                            w.applied++;
                        }
                        w.r = 0;
                        w.copy++;
                    }
                } else {
                    if (--w.r < 0) {
                        if (w.applied < w.nrefl) {
                            w.refl *= surface_factor(m, w.t, (w.copy & 1) ? 1u : 0u, w.sin2);
                            w.applied++;
                        }
                        w.r = layers - 1;
                        w.copy--;
                    }
                }
            }
            w.u += w.zdir;
        } else {
            if (Fetch::boundary(w.wexit)) {
                // leaving the mesh sideways: only rounding short of the end (a
                // sample point on the rim) counts as arrival
                w.done = 1;
                if ((1.0f - w.sig) * w.D > w.tol) w.status = 1;
                return;
            }
            // enter the neighbour through its edge f: L = vertex f, R = vertex f+1, N = vertex f+2
            const uint32_t t = Fetch::next(w.wexit), f = Fetch::edge(w.wexit);
            const uint32_t e1 = f == 2u ? 0u : f + 1u, v = f == 0u ? 2u : f - 1u;
            const float2 pn = fetch.vertex(m, t, v);
            uint32_t wr_n, wl_n;
            fetch.words(m, t, e1, v, wr_n, wl_n, w.clad);
            const float X = pn.x - w.ax, Y = pn.y - w.ay;
            const float c = side_of(w.dx, w.dy, X, Y), a = along(w.dx, w.dy, X, Y);
            const bool left = c >= 0.0f;
            // left of the line: leave through R-N (edge f+1), N becomes L; else through N-L (edge f+2)
            w.cl = left ? c : w.cl;
            w.wl = left ? a : w.wl;
            w.cr = left ? w.cr : c;
            w.wr = left ? w.wr : a;
            w.wexit = left ? wr_n : wl_n;
            const bool at_end = (pn.x == w.ex) & (pn.y == w.ey);
            w.sxy = at_end ? 1.0f : edge_sigma(w.cl, w.wl, w.cr, w.wr, w.inv_d2);
            w.t = t;
        }
        const int level = folded_level<kReflect>(w.u, w.copy, w.r, layers);
        w.beta = load_f32(m.beta, w.t + mul24((uint32_t)level, m.T));
    }
}

HASE_HD inline WalkOut walk_result(const WalkState& s)
{
    WalkOut w;
    w.beta_len = s.beta_sig * s.D;
    w.gain_len = (s.sig - s.clad_sig) * s.D;
    w.clad_len = s.clad_sig * s.D;
    w.refl = s.refl;
    w.length = s.D;
    w.steps = s.steps;
    w.status = s.status;
    return w;
}

// Whole walk of one ray from triangle t (corners as the fetch returns them).
template <bool kReflect, class Fetch = GlobalTriangles>
HASE_HD inline WalkOut walk_ray_t(const MeshDev& m, float ax, float ay, float za, uint32_t t, uint32_t layer,
                                  float2 p0, float2 p1, float2 p2, float ex, float ey, float ze, uint32_t nrefl,
                                  uint32_t first, const Fetch& fetch = Fetch())
{
    WalkState s;
    walk_begin(m, s, ax, ay, za, t, layer, p0, p1, p2, ex, ey, ze, nrefl, first, fetch);
    walk_advance<kReflect>(m, s, kMaxSteps + 1u, fetch);
    return walk_result(s);
}

template <class Fetch = GlobalTriangles>
HASE_HD inline WalkOut walk_ray(const MeshDev& m, float ax, float ay, float za, uint32_t t, uint32_t layer,
                                float2 p0, float2 p1, float2 p2, float ex, float ey, float ze, uint32_t nrefl,
                                uint32_t first, const Fetch& fetch = Fetch())
{
    return nrefl == 0 ? walk_ray_t<false>(m, ax, ay, za, t, layer, p0, p1, p2, ex, ey, ze, 0u, first, fetch)
                      : walk_ray_t<true>(m, ax, ay, za, t, layer, p0, p1, p2, ex, ey, ze, nrefl, first, fetch);
}

// Whole walk with the start triangle's corners read through the fetch.
template <class Fetch = GlobalTriangles>
HASE_HD inline WalkOut walk_ray(const MeshDev& m, float ax, float ay, float za, uint32_t t, uint32_t layer, float ex,
                                float ey, float ze, uint32_t nrefl, uint32_t first, const Fetch& fetch = Fetch())
{
    return walk_ray(m, ax, ay, za, t, layer, fetch.vertex(m, t, 0u), fetch.vertex(m, t, 1u), fetch.vertex(m, t, 2u),
                    ex, ey, ze, nrefl, first, fetch);
}

// Height, in layer units, of the mirror image of a sample point on plane
// `lv` after `nrefl` alternating reflections at z = 0 (plane 0) and
// z = layers * h (plane 1), the first at `first`.  Integer arithmetic, so
// the image lies exactly on a plane of the unfolded axis.
HASE_HD inline int image_plane(int lv, int layers, uint32_t nrefl, uint32_t first)
{
    int mr = (int)nrefl;
    if (nrefl == 0) return lv;
    if (first == 1u) return (nrefl & 1u) ? mr * layers + (layers - lv) : mr * layers + lv;
    return (nrefl & 1u) ? -(mr - 1) * layers - lv : -mr * layers + lv;
}

constexpr float kLog2e = 1.44269504088896341f;

// 2^x with the hardware exponential on the device.
HASE_HD inline float fast_exp2(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_exp2f(x);
#else
    return exp2f(x);
#endif
}

// Gain of a walked path averaged over the spectral samples.  The exponent is
// linear in the walk's length sums, so one walk serves every wavelength:
// E = nTot (se + sa) sum(beta l) - nTot sa sum(l) - alpha_clad sum(l_clad)
// (g = nTot (beta (se + sa) - sa), include/kernels/forward/rayWalk.hpp:170-182).
// The tables hold the coefficients in base 2, c1 = nTot (se + sa) log2(e) and
// c2 = nTot sa log2(e), so each wavelength costs two FMAs and one exp2.
#if defined(__HIP_DEVICE_COMPILE__)
// read-only for the whole launch and indexed uniformly across the wavefront:
// the constant address space lets them be read through the scalar cache
typedef const __attribute__((address_space(4))) float ConstF32;
#else
typedef const float ConstF32;
#endif

HASE_HD inline float spectral_gain(const float* c1, const float* c2, uint32_t n_lambda, float inv_lambda,
                                   float clad_alpha, const WalkOut& w)
{
    ConstF32* k1 = (ConstF32*)c1;
    ConstF32* k2 = (ConstF32*)c2;
    const float tail = clad_alpha * kLog2e * w.clad_len;
    float g = 0.0f;
#pragma unroll 4
    for (uint32_t lam = 0; lam < n_lambda; ++lam)
        g += fast_exp2(fmaf(k1[lam], w.beta_len, fmaf(-k2[lam], w.gain_len, -tail)));
    return g * inv_lambda;
}

}  // namespace hase
