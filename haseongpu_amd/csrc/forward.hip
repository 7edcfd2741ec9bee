// forward.hip - forward Tet4 ASE estimator: rays leave beta*V-stratified
// points of a tetrahedral gain medium in isotropic directions and score the
// track-length integral of their gain into the vertices of every cell they
// cross; each cell's flux is its mean lumped-vertex density.
//
// Follows the reference's present forward path: AccumulateForwardPhiAse /
// ForwardAseCellPolicy
// (include/kernels/forward/accumulation.hpp:300-405), the walk and its tie
// and recovery rules (policyRay.hpp:379-465, rayWalk.hpp:106-208,
// rayTransition.hpp:17-188), midpoint vertex weights (barycentric.hpp:69-99),
// stratified beta*V sources and wavelengths (volumeSampling.hpp:19-137),
// face planes and lumped vertex volumes (include/core/mesh.hpp:630-725,
// include/kernels/vertexAccumulation.hpp:26-39), FinalizeForwardVolumePhiAse
// (include/kernels/forwardPhiAseMapping.hpp:47-147), the adaptive loop
// (include/core/forwardPhiAseEvaluator.hpp:114-178, types.hpp:333-396) and
// the surface reservoir method for reflections (include/core/forwardSrm.hpp:
// 120-366, accumulation.hpp:407-834): rays leaving through a reflective face
// leave their reflected direction in that face's weighted reservoir (one per
// RSE batch), and passes restart rays from face centroids until the
// remaining weight is below the tolerance, stops changing or keeps growing.
//
// Layout in HBM (C cells, Np points): face planes 4 x double4 per cell
// (128 B, one cache line), a 64-B cell record (neighbour links: neighbour * 4
// + its local face, -1 at the boundary; vertex ids; beta; cladding flag),
// beta*V prefix f64[C], lumped vertex volumes f64[2 Np], scores f64[8 x 8 x 2 Np]
// (eight atomic copies x eight RSE batches x gain | cladding vertex copies),
// visits and dropped
// u32[C].  One thread walks one ray in FP64 with contraction off, so a ray
// takes exactly the cells the FP64 oracle takes.  Vertex scores are FP64
// hardware atomics, which bound the kernel; since consecutive cells of a walk
// share three vertices, a ray keeps the current cell's four vertex shares in
// registers and writes a share only when the walk leaves its vertex (about
// one atomic per segment instead of four).  Summation order, and exp/expm1 in
// their last bits, differ from the oracle's.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <memory>
#include <numeric>
#include <vector>

#include "common.h"
#include "prism_walk.h"

namespace hase {
namespace fwd {

constexpr uint32_t kBatches = 8;
// copies of the vertex scores: work-group b adds into copy b % kCopies, which
// spreads the FP64 atomics of concurrent work-groups over more addresses (the
// copies are folded into copy 0 before the scores are read)
constexpr uint32_t kCopies = 8;
constexpr uint32_t kKeyTag = 0x46574431u;
constexpr uint32_t kMaxRecovery = 32;
constexpr uint32_t kDirectionCalls = 32;
constexpr double kTol = 64.0 * DBL_EPSILON;   // rayWalk.hpp:17
constexpr double kProbe = 1.0e-12;            // rayTransition.hpp:18
constexpr int kTraceBlock = 256;

enum { kEntered = 0, kBoundary = 1, kFailed = 2 };

// Everything but the face planes that a walk step reads of its cell, in one
// 64-byte record: a step touches two cache lines (planes, record).
struct CellAux {
    int32_t link[4];   // neighbour * 4 + its local face, -1 at the boundary
    uint32_t verts[4]; // point ids of the 4 vertices
    double beta;
    uint32_t clad;
    int32_t bface[4];  // reflective boundary face index of each face, -1 if none
    uint32_t pad;
};
static_assert(sizeof(CellAux) == 64, "cell record is one half cache line");

struct TetDev {
    const double4* planes;    // 4 per cell
    const CellAux* aux;
    const double* bv_prefix;
    const double* px;
    const double* py;
    const double* pz;
    uint32_t n_cells, n_points;
    double n_tot, clad_alpha;
};

struct TraceArgs {
    TetDev m;
    const double* sigma_a;
    const double* sigma_e;
    uint32_t n_lambda;
    uint32_t seed, launch, rays;
    double shift[kBatches];
    uint32_t phase[kBatches];
    double* scores;
    uint32_t* visits;
    uint32_t* dropped;
    unsigned long long* segments;
};

struct Hit {
    int face;
    double length;
    uint32_t tied;
};

struct Cross {
    uint32_t cell;
    int forbidden;
    int status;
    int face;          // boundary face when status == kBoundary
};

// Reflective boundary face: centroid (where reflected rays restart), outward
// unit normal and surface data (include/kernels/forward/rayWalk.hpp:42-84,
// reflection.hpp:21-47).
struct BFace {
    double c[3];
    double n[3];
    double refl, n_in, n_out;
    uint32_t face;     // cell * 4 + local face
    uint32_t pad;
};

struct Slot {
    double d[3];
    double w;
    uint32_t sigma;
    uint32_t pad;
};

// One side (input or output) of the surface reservoirs of all eight batches.
struct Reservoir {
    uint32_t* count;        // [8][nb]
    double* face_weight;    // [8][nb]
    Slot* slots;            // [8][nb][S]
};

struct SrmArgs {
    const BFace* faces;
    uint32_t nb, slots, pass;
    Reservoir out;
};

__device__ __forceinline__ double plane_at(double4 pl, double x, double y, double z)
{
#pragma clang fp contract(off)
    return pl.x * x + pl.y * y + pl.z * z + pl.w;
}

__device__ __forceinline__ double plane_dir(double4 pl, double x, double y, double z)
{
#pragma clang fp contract(off)
    return pl.x * x + pl.y * y + pl.z * z;
}

__device__ __forceinline__ Hit next_face(const TetDev& m, uint32_t cell, const double p[3], const double u[3],
                                         int forbidden)
{
#pragma clang fp contract(off)
    Hit h{-1, DBL_MAX, 0u};
    double4 pl[4];
    double c[4], d[4], cand[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pl[k] = m.planes[4 * (size_t)cell + k];
        c[k] = plane_at(pl[k], p[0], p[1], p[2]);
        d[k] = plane_dir(pl[k], u[0], u[1], u[2]);
        cand[k] = 0.0;
        if (k != forbidden && d[k] < 0.0) {
            const double len = -c[k] / d[k];
            cand[k] = (len > 0.0 && len <= DBL_MAX) ? len : 0.0;
            if (cand[k] > 0.0 && cand[k] < h.length) {
                h.length = cand[k];
                h.face = k;
            }
        }
    }
    if (h.face < 0) return h;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k != forbidden && cand[k] > 0.0 && fabs(c[k] + h.length * d[k]) <= kTol) h.tied |= 1u << k;
    h.tied |= 1u << h.face;
    h.face = __builtin_ctz(h.tied);
    return h;
}

// Crossing at a point where several faces meet (rayTransition.hpp:74-160).
__device__ Cross recover(const TetDev& m, uint32_t cell, int exit_face, const double p[3], const double u[3])
{
#pragma clang fp contract(off)
    Cross r{cell, -1, kEntered, -1};
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t recent0 = cell, recent1 = kNone, recent2 = kNone;
    for (uint32_t step = 0; step < kMaxRecovery; ++step) {
        const int32_t link = m.aux[cell].link[exit_face];
        if (link < 0) {
            r.cell = cell;
            r.status = kBoundary;
            r.face = exit_face;
            return r;
        }
        cell = (uint32_t)link >> 2;
        r.cell = cell;
        r.forbidden = link & 3;
        recent2 = recent1;
        recent1 = recent0;
        recent0 = cell;
        int fallback = -1, preferred = -1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (preferred >= 0) continue;
            const double4 pl = m.planes[4 * (size_t)cell + k];
            double c = plane_at(pl, p[0], p[1], p[2]);
            if (fabs(c) <= kTol) c = 0.0;
            const double d = plane_dir(pl, u[0], u[1], u[2]);
            if (c + kProbe * d >= 0.0) continue;
            if (fallback < 0) fallback = k;
            const int32_t cl = m.aux[cell].link[k];
            if (cl < 0) continue;
            const uint32_t cn = (uint32_t)cl >> 2;
            if (cn != recent0 && cn != recent1 && cn != recent2) preferred = k;
        }
        exit_face = preferred >= 0 ? preferred : fallback;
        if (exit_face < 0) return r;
    }
    r.status = kFailed;
    return r;
}

__device__ __forceinline__ Cross cross(const TetDev& m, uint32_t cell, Hit h, const double p[3], const double u[3])
{
    if (h.tied & (h.tied - 1u)) return recover(m, cell, h.face, p, u);
    const int32_t link = m.aux[cell].link[h.face];
    if (link < 0) return Cross{cell, -1, kBoundary, h.face};
    return Cross{(uint32_t)link >> 2, link & 3, kEntered, -1};
}

// Start of ray r of a launch: contiguous RSE batch, stratified beta*V cell,
// folded-cube point (Rocchini & Cignoni 2000), Marsaglia (1972) direction,
// stratified wavelength.
__device__ __forceinline__ void ray_start(const TraceArgs& a, uint32_t r, uint32_t& batch, uint32_t& cell,
                                          double p[3], double u[3], uint32_t& lambda)
{
#pragma clang fp contract(off)
    const uint32_t base = a.rays / kBatches, rem = a.rays % kBatches;
    const uint32_t split = rem * (base + 1u);
    const uint32_t b = r < split ? r / (base + 1u) : rem + (r - split) / base;
    const uint32_t count = base + (b < rem ? 1u : 0u);
    const uint32_t first = b * base + (b < rem ? b : rem);
    const uint32_t i = r - first;
    batch = b;
    const uint32_t n = a.m.n_cells;
    const double total = a.m.bv_prefix[n - 1];
    const double target = ((double)i + a.shift[b]) / (double)count * total;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (target < a.m.bv_prefix[mid]) hi = mid;
        else lo = mid + 1u;
    }
    cell = lo < n ? lo : n - 1u;
    uint32_t o[4];
    philox4x32_10(r, a.launch, 0u, 0u, a.seed, kKeyTag, o);
    double s = (double)u32_to_unit(o[0]), t = (double)u32_to_unit(o[1]), v = (double)u32_to_unit(o[2]);
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
    const double w0 = 1.0 - s - t - v;
    const uint32_t* cv = a.m.aux[cell].verts;
    p[0] = w0 * a.m.px[cv[0]] + s * a.m.px[cv[1]] + t * a.m.px[cv[2]] + v * a.m.px[cv[3]];
    p[1] = w0 * a.m.py[cv[0]] + s * a.m.py[cv[1]] + t * a.m.py[cv[2]] + v * a.m.py[cv[3]];
    p[2] = w0 * a.m.pz[cv[0]] + s * a.m.pz[cv[1]] + t * a.m.pz[cv[2]] + v * a.m.pz[cv[3]];
    u[0] = 0.0;
    u[1] = 0.0;
    u[2] = 1.0;
    bool done = false;
    for (uint32_t call = 0; call < kDirectionCalls && !done; ++call) {
        philox4x32_10(r, a.launch, call + 1u, 0u, a.seed, kKeyTag, o);
#pragma unroll
        for (int pair = 0; pair < 2; ++pair) {
            if (done) continue;
            const double x1 = 2.0 * (double)u32_to_unit(o[2 * pair]) - 1.0;
            const double x2 = 2.0 * (double)u32_to_unit(o[2 * pair + 1]) - 1.0;
            const double q = x1 * x1 + x2 * x2;
            if (q >= 1.0) continue;
            const double rr = sqrt(1.0 - q);
            u[0] = 2.0 * x1 * rr;
            u[1] = 2.0 * x2 * rr;
            u[2] = 1.0 - 2.0 * q;
            done = true;
        }
    }
    lambda = a.n_lambda ? (uint32_t)(((uint64_t)i * a.n_lambda / count + a.phase[b]) % a.n_lambda) : 0u;
}

struct Exit {
    uint32_t cell;
    int face;          // boundary face the walk left through, -1 if it did not
    double weight;     // source weight times accumulated gain there
};

// Walks one ray from `p` in cell `cell` (entered through `forbidden`),
// scoring weight * gain * track integral into the vertices of every cell.
__device__ void walk(const TraceArgs& a, uint32_t cell, double p[3], const double u[3], int forbidden,
                     double weight, uint32_t batch, double sa, double se, uint32_t& segments, Exit& ex)
{
#pragma clang fp contract(off)
    const TetDev& m = a.m;
    double* scores = a.scores + ((size_t)(blockIdx.x % kCopies) * kBatches + batch) * 2u * m.n_points;
    double gain = 1.0;
    bool dropped = true;
    ex.face = -1;
    ex.weight = 0.0;
    // vertex shares of the last scored cell, not yet added to `scores`
    uint32_t pend_v[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    double pend_s[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t step = 0; step < kMaxSteps; ++step) {
        const Hit h = next_face(m, cell, p, u, forbidden);
        if (h.face < 0) {
            // on a face with no exit ahead: leave through a face the ray
            // already lies on (rayTransition.hpp:41-72, policyRay.hpp:394-425)
            int face = -1;
            bool near = false;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                near |= fabs(plane_at(m.planes[4 * (size_t)cell + k], p[0], p[1], p[2])) <= kTol;
            if (near) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (face >= 0 || k == forbidden) continue;
                    const double4 pl = m.planes[4 * (size_t)cell + k];
                    if (fabs(plane_at(pl, p[0], p[1], p[2])) <= kTol && plane_dir(pl, u[0], u[1], u[2]) < 0.0)
                        face = k;
                }
            }
            if (face < 0) break;
            const Cross t = recover(m, cell, face, p, u);
            cell = t.cell;
            if (t.status == kFailed) break;
            if (t.status == kBoundary) {
                dropped = false;
                ex.face = t.face;
                ex.weight = weight * gain;
                break;
            }
            forbidden = t.forbidden;
            continue;
        }
        // score the segment (accumulation.hpp:310-345)
        const CellAux& ax = m.aux[cell];
        const bool clad = ax.clad != 0;
        const double g = clad ? -m.clad_alpha : m.n_tot * (ax.beta * (se + sa) - sa);
        const double gl = g * h.length;
        // one expm1 gives the track integral and the gain factor
        const double em1 = expm1(gl);
        const double contribution = weight * gain * (gl == 0.0 ? h.length : em1 / g);
        if (isfinite(contribution)) {
            atomicAdd(&a.visits[cell], 1u);
            ++segments;
            const double half = 0.5 * h.length;
            const double mx = p[0] + u[0] * half, my = p[1] + u[1] * half, mz = p[2] + u[2] * half;
            double w[4], sum = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                w[k] = fmax(0.0, plane_at(m.planes[4 * (size_t)cell + k], mx, my, mz));
                sum += w[k];
            }
            const uint32_t off = clad ? m.n_points : 0u;
            const uint32_t vid[4] = {ax.verts[0] + off, ax.verts[1] + off, ax.verts[2] + off, ax.verts[3] + off};
            // consecutive cells share a face: carry the pending shares of the
            // vertices this cell keeps, write out the ones it drops
            double carried[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                bool kept = false;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (pend_v[k] == vid[j]) {
                        carried[j] = pend_s[k];
                        kept = true;
                    }
                if (!kept && pend_s[k] != 0.0) unsafeAtomicAdd(&scores[pend_v[k]], pend_s[k]);
            }
            const bool flat = sum <= DBL_EPSILON;
            const double inv = flat ? 0.0 : 1.0 / sum;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                pend_v[k] = vid[k];
                pend_s[k] = carried[k] + contribution * (flat ? 0.25 : w[k] * inv);
            }
        } else {
            atomicAdd(&a.dropped[cell], 1u);
        }
        gain *= 1.0 + em1;
        p[0] = p[0] + u[0] * h.length;
        p[1] = p[1] + u[1] * h.length;
        p[2] = p[2] + u[2] * h.length;
        const Cross t = cross(m, cell, h, p, u);
        cell = t.cell;
        if (t.status == kFailed) break;
        if (t.status == kBoundary) {
            dropped = false;
            ex.face = t.face;
            ex.weight = weight * gain;
            break;
        }
        forbidden = t.forbidden;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (pend_s[k] != 0.0) unsafeAtomicAdd(&scores[pend_v[k]], pend_s[k]);
    if (dropped) atomicAdd(&a.dropped[cell], 1u);
    ex.cell = cell;
}

__device__ __forceinline__ void unit_vector(double v[3])
{
#pragma clang fp contract(off)
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

// depositReflection (accumulation.hpp:409-455): weighted reservoir sampling
// of the reflected direction at reflective boundary face f of `batch`.
__device__ void deposit(const SrmArgs& s, uint32_t batch, uint32_t f, const double dir[3], double w, uint32_t sigma,
                        uint32_t seed, uint32_t launch, uint32_t ray)
{
#pragma clang fp contract(off)
    if (s.slots == 0 || !(w > 0.0) || !isfinite(w)) return;
    const BFace bf = s.faces[f];
    double reflectance = fmax(0.0, bf.refl);
    if (bf.n_in > 0.0 && bf.n_out > 0.0) {
        double d[3] = {dir[0], dir[1], dir[2]};
        unit_vector(d);
        const double cosi = fmin(1.0, fabs(d[0] * bf.n[0] + d[1] * bf.n[1] + d[2] * bf.n[2]));
        const double sin2 = fmax(0.0, 1.0 - cosi * cosi);
        const double ratio = bf.n_in / bf.n_out;
        if (ratio * ratio * sin2 > 1.0) reflectance = 1.0;
    }
    const double rw = w * reflectance;
    if (!(rw > 0.0) || !isfinite(rw)) return;
    const size_t fi = (size_t)batch * s.nb + f;
    const uint32_t seen = atomicAdd(&s.out.count[fi], 1u);
    const double prior = atomicAdd(&s.out.face_weight[fi], rw);
    uint32_t slot = seen;
    if (slot >= s.slots) {
        uint32_t o[4];
        philox4x32_10(ray, launch, 0x10000u + s.pass, 1u, seed, kKeyTag, o);
        const double keep = fmin(1.0, (double)s.slots * rw / (prior + rw));
        if ((double)u32_to_unit(o[0]) >= keep) return;
        slot = (uint32_t)((double)u32_to_unit(o[1]) * (double)s.slots);
        if (slot >= s.slots) slot = s.slots - 1;
    }
    const double dn = 2.0 * (dir[0] * bf.n[0] + dir[1] * bf.n[1] + dir[2] * bf.n[2]);
    double out[3] = {dir[0] - bf.n[0] * dn, dir[1] - bf.n[1] * dn, dir[2] - bf.n[2] * dn};
    unit_vector(out);
    Slot& sl = s.out.slots[fi * s.slots + slot];
    sl.d[0] = out[0];
    sl.d[1] = out[1];
    sl.d[2] = out[2];
    sl.w = rw;
    sl.sigma = sigma;
}

// Primary rays (AccumulateForwardPhiAse / AccumulateForwardPhiAseReservoir):
// with kReflect a ray leaving through a reflective face deposits there.
template <bool kReflect>
__global__ void __launch_bounds__(kTraceBlock) hase_forward_trace_kernel(TraceArgs a, SrmArgs s)
{
#pragma clang fp contract(off)
    const uint32_t r = blockIdx.x * (uint32_t)blockDim.x + threadIdx.x;
    if (r >= a.rays) return;
    uint32_t batch, cell, lambda, segments = 0;
    double p[3], u[3];
    ray_start(a, r, batch, cell, p, u, lambda);
    Exit ex;
    walk(a, cell, p, u, -1, 1.0, batch, a.sigma_a[lambda], a.sigma_e[lambda], segments, ex);
    if (kReflect && ex.face >= 0) {
        const int32_t f = a.m.aux[ex.cell].bface[ex.face];
        if (f >= 0) deposit(s, batch, (uint32_t)f, u, ex.weight, lambda, a.seed, a.launch, r);
    }
    atomicAdd(a.segments, (unsigned long long)segments);
}

struct PassArgs {
    Reservoir in;
    const double* cdf;          // [8][nb] normalised inclusive prefix of the input face weights
    double source[kBatches];    // weight of each reflected ray of a batch, 0 when the batch has stopped
    double shift[kBatches];     // stratification offset of the face choice
};

// One reflected pass (AccumulateReflectedForwardPhiAse,
// accumulation.hpp:679-832): a face by reservoir weight (stratified when
// every face can get a ray), a slot by slot weight, a walk from the face
// centroid away from the face, deposits into the output reservoir.
__global__ void __launch_bounds__(kTraceBlock) hase_forward_reflect_kernel(TraceArgs a, SrmArgs s, PassArgs q)
{
#pragma clang fp contract(off)
    const uint32_t r = blockIdx.x * (uint32_t)blockDim.x + threadIdx.x;
    if (r >= a.rays || s.nb == 0) return;
    const uint32_t base = a.rays / kBatches, rem = a.rays % kBatches;
    const uint32_t split = rem * (base + 1u);
    const uint32_t b = r < split ? r / (base + 1u) : rem + (r - split) / base;
    const uint32_t count = base + (b < rem ? 1u : 0u);
    const uint32_t i = r - (b * base + (b < rem ? b : rem));
    if (!(q.source[b] > 0.0)) return;
    uint32_t o[4];
    philox4x32_10(r, a.launch, 0x10000u + s.pass, 2u, a.seed, kKeyTag, o);
    const double target =
        s.nb <= count ? ((double)i + q.shift[b]) / (double)count : (double)u32_to_unit(o[0]);
    const double* cdf = q.cdf + (size_t)b * s.nb;
    uint32_t lo = 0, hi = s.nb;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (cdf[mid] <= target) lo = mid + 1u;
        else hi = mid;
    }
    const uint32_t f = lo < s.nb ? lo : s.nb - 1u;
    const size_t fi = (size_t)b * s.nb + f;
    const uint32_t filled = min(q.in.count[fi], s.slots);
    if (filled == 0) return;
    const Slot* sl = q.in.slots + fi * s.slots;
    double slot_weight = 0.0;
    for (uint32_t k = 0; k < filled; ++k) slot_weight += sl[k].w;
    if (!(slot_weight > 0.0)) return;
    const double t2 = (double)u32_to_unit(o[1]) * slot_weight;
    double cum = 0.0;
    uint32_t pick = filled - 1u;
    for (uint32_t k = 0; k < filled; ++k) {
        cum += sl[k].w;
        if (cum >= t2) {
            pick = k;
            break;
        }
    }
    double dir[3] = {sl[pick].d[0], sl[pick].d[1], sl[pick].d[2]};
    unit_vector(dir);
    const uint32_t lam = sl[pick].sigma;
    const BFace bf = s.faces[f];
    double p[3] = {bf.c[0], bf.c[1], bf.c[2]};
    uint32_t segments = 0;
    Exit ex;
    walk(a, bf.face >> 2, p, dir, (int)(bf.face & 3u), q.source[b], b, a.sigma_a[lam], a.sigma_e[lam], segments,
         ex);
    if (ex.face >= 0) {
        const int32_t g = a.m.aux[ex.cell].bface[ex.face];
        if (g >= 0) deposit(s, b, (uint32_t)g, dir, ex.weight, lam, a.seed, a.launch, r);
    }
    atomicAdd(a.segments, (unsigned long long)segments);
}

// Per batch (one work-group each): inclusive prefix of the face weights of a
// reservoir side, normalised by its total; totals[b] = the total.
__global__ void __launch_bounds__(1024) hase_forward_face_cdf_kernel(const double* face_weight, uint32_t nb,
                                                                     double* cdf, double* totals)
{
    __shared__ double wave_sums[16];
    __shared__ double carry_s;
    const uint32_t b = blockIdx.x, t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    const double* w = face_weight + (size_t)b * nb;
    double* out = cdf + (size_t)b * nb;
    if (t == 0) carry_s = 0.0;
    __syncthreads();
    for (uint32_t start = 0; start < nb; start += 1024u) {
        const uint32_t i = start + t;
        double v = i < nb ? w[i] : 0.0;
        for (int d = 1; d < 64; d <<= 1) {
            const double x = __shfl_up(v, d, 64);
            if ((int)lane >= d) v += x;
        }
        if (lane == 63u) wave_sums[wv] = v;
        __syncthreads();
        double before = carry_s;
        for (uint32_t k = 0; k < wv; ++k) before += wave_sums[k];
        if (i < nb) out[i] = before + v;
        __syncthreads();
        if (t == 1023u) carry_s = before + v;
        __syncthreads();
    }
    const double total = carry_s;
    for (uint32_t i = t; i < nb; i += 1024u) out[i] = total > 0.0 ? out[i] / total : 0.0;
    if (t == 0) totals[b] = total;
}

__global__ void hase_forward_clear_reservoir_kernel(uint32_t* count, double* face_weight, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) {
        count[i] = 0u;
        face_weight[i] = 0.0;
    }
}

struct FinalizeArgs {
    TetDev m;
    const double* volume;
    const double* lumped;
    const double* scores;
    const uint32_t* dropped;
    uint32_t counts[kBatches];
    double rays;
    double total;
    double rate;
    double sigma_a, sigma_e;
    double threshold;
    uint32_t target;
    float* phi;
    double* se;
    double* rse;
    double* dndt;
    uint32_t* convergence;
    uint32_t* open_cells;
};

// FinalizeForwardVolumePhiAse (forwardPhiAseMapping.hpp:75-147) plus the
// convergence record and stop count (types.hpp:356-396).
__global__ void __launch_bounds__(256) hase_forward_finalize_kernel(FinalizeArgs a)
{
#pragma clang fp contract(off)
    const uint32_t c = blockIdx.x * (uint32_t)blockDim.x + threadIdx.x;
    if (c >= a.m.n_cells) return;
    const TetDev& m = a.m;
    const CellAux& ax = m.aux[c];
    const uint32_t off = ax.clad ? m.n_points : 0u;
    const uint32_t vid[4] = {ax.verts[0] + off, ax.verts[1] + off, ax.verts[2] + off, ax.verts[3] + off};
    const double vol = a.volume[c];
    const size_t nv = 2u * (size_t)m.n_points;
    double rel = DBL_MAX, abs_err = DBL_MAX, estimate = 0.0;
    if (a.rays > 0.0 && vol > 0.0) {
        double score_sum = 0.0, mean_sum = 0.0, mean_sq = 0.0;
        uint32_t active = 0;
        for (uint32_t b = 0; b < kBatches; ++b) {
            double dens = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double lv = a.lumped[vid[k]];
                dens += lv > 0.0 ? a.scores[b * nv + vid[k]] / lv : 0.0;
            }
            dens /= 4.0;
            const double bs = dens * vol;
            score_sum += bs;
            if (a.counts[b] == 0) continue;
            const double mean = bs / (double)a.counts[b];
            mean_sum += mean;
            mean_sq += mean * mean;
            ++active;
        }
        estimate = score_sum * a.total / (a.rays * vol);
        if (a.dropped[c] == 0 && active >= 2) {
            const double n = (double)active;
            const double mean = mean_sum / n;
            if (mean == 0.0) {
                rel = __builtin_nan("");
                abs_err = 0.0;
            } else {
                const double var = fmax(0.0, (mean_sq - mean_sum * mean_sum / n) / (n - 1.0));
                rel = sqrt(var / n) / fabs(mean);
                abs_err = rel * fabs(estimate) * a.rate;
            }
        }
    }
    const float f = (float)(estimate * a.rate);
    a.phi[c] = f;
    a.se[c] = abs_err;
    a.rse[c] = rel;
    a.dndt[c] = (ax.beta * (a.sigma_e + a.sigma_a) - a.sigma_a) * (double)f;
    const bool ok = a.dropped[c] == 0 && isfinite(rel) && rel * rel <= a.threshold;
    if (ok && a.convergence[c] == 0) a.convergence[c] = a.target;
    if (!ok) atomicAdd(a.open_cells, 1u);
}

// copy 0 += copies 1..kCopies-1, which are cleared (n = one copy's length)
__global__ void hase_forward_fold_kernel(double* scores, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = scores[i];
#pragma unroll
    for (uint32_t c = 1; c < kCopies; ++c) {
        s += scores[c * n + i];
        scores[c * n + i] = 0.0;
    }
    scores[i] = s;
}

__global__ void hase_forward_clear_kernel(double* scores, size_t n_scores, uint32_t* counters, size_t n_counters)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n_scores) scores[i] = 0.0;
    if (i < n_counters) counters[i] = 0u;
}

}  // namespace fwd
}  // namespace hase

struct hase_forward {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t n_cells = 0, n_points = 0, n_lambda = 0;
    uint32_t min_rays = 0, max_rays = 0, steps = 0, seed = 0;
    double threshold = 0.0, max_sigma_a = 0.0, max_sigma_e = 0.0, n_tot = 0.0, tau = 0.0, clad_alpha = 0.0;
    std::vector<uint32_t> verts;
    std::vector<uint8_t> clad;
    std::vector<double> volume;
    std::vector<double> beta;
    std::vector<double> prefix;
    std::vector<hase::fwd::CellAux> aux_host;
    hase::DevBuf<double4> planes;
    hase::DevBuf<hase::fwd::CellAux> aux;
    hase::DevBuf<double> prefix_dev, px, py, pz, volume_dev, lumped, sigma_a, sigma_e, scores;
    // counters: visits [C], dropped [C], convergence [C], open cells [1]
    hase::DevBuf<uint32_t> counters;
    hase::DevBuf<unsigned long long> segments;
    hase::DevBuf<float> phi;
    hase::DevBuf<double> se, rse, dndt;
    hase_forward_stats_t stats{};
    // surface reservoir method (reflections)
    bool reflect = false;                       // experiment.use_reflections
    uint32_t nb = 0, slots = 32, max_passes = 40, streak = 3;
    double tolerance = 1.0e-4;
    hase::DevBuf<hase::fwd::BFace> bfaces;
    hase::DevBuf<uint32_t> res_count[2];
    hase::DevBuf<double> res_weight[2];
    hase::DevBuf<hase::fwd::Slot> res_slots[2];
    hase::DevBuf<double> face_cdf, face_totals;
    uint32_t srm_status[hase::fwd::kBatches] = {0};
    uint32_t srm_passes[hase::fwd::kBatches] = {0};
    double srm_remaining[hase::fwd::kBatches] = {0.0};

    ~hase_forward()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
    hase::fwd::TetDev view() const
    {
        return hase::fwd::TetDev{planes.p, aux.p, prefix_dev.p, px.p, py.p, pz.p,
                                 n_cells, n_points, n_tot, clad_alpha};
    }
};

namespace hase {
namespace fwd {

using hase::HaseError;

static void validate(const hase_tet_mesh_t* mesh, const hase_experiment_t* exp, const hase_compute_t* comp)
{
    if (!mesh || !exp || !comp) throw HaseError(HASE_E_INVALID, "mesh, experiment and compute must not be NULL");
    if (!mesh->points || !mesh->cells || !mesh->beta_volume)
        throw HaseError(HASE_E_INVALID, "points, cells and beta_volume are required");
    if (mesh->number_of_cells == 0 || mesh->number_of_points < 4)
        throw HaseError(HASE_E_INVALID, "a Tet4 mesh needs at least one cell and four points");
    if (mesh->number_of_cells >= (1u << 29)) throw HaseError(HASE_E_INVALID, "number_of_cells must be < 2^29");
    if (mesh->number_of_points >= (1u << 30)) throw HaseError(HASE_E_INVALID, "number_of_points must be < 2^30");
    if (!(mesh->n_tot > 0.0) || !(mesh->crystal_t_fluo > 0.0) || !std::isfinite(mesh->n_tot) ||
        !std::isfinite(mesh->crystal_t_fluo))
        throw HaseError(HASE_E_INVALID, "n_tot and crystal_t_fluo must be positive");
    if (!(mesh->cladding_absorption >= 0.0) || !std::isfinite(mesh->cladding_absorption))
        throw HaseError(HASE_E_INVALID, "cladding_absorption must be finite and >= 0");
    const size_t np = mesh->number_of_points;
    for (size_t i = 0; i < 3 * np; ++i)
        if (!std::isfinite(mesh->points[i])) throw HaseError(HASE_E_INVALID, "points must be finite");
    for (size_t i = 0; i < 4 * (size_t)mesh->number_of_cells; ++i)
        if (mesh->cells[i] >= mesh->number_of_points)
            throw HaseError(HASE_E_INVALID, "cell " + std::to_string(i / 4) + " names a point out of range");
    for (size_t c = 0; c < mesh->number_of_cells; ++c)
        if (!std::isfinite(mesh->beta_volume[c])) throw HaseError(HASE_E_INVALID, "beta_volume must be finite");
    if (exp->min_rays_per_sample == 0) throw HaseError(HASE_E_INVALID, "min rays must be > 0");
    if (exp->max_rays_per_sample < exp->min_rays_per_sample)
        throw HaseError(HASE_E_INVALID, "max rays must be >= min rays");
    if (!(exp->mse_threshold >= 0.0)) throw HaseError(HASE_E_INVALID, "mse_threshold must be >= 0");
    if (exp->number_of_wavelengths == 0 || !exp->sigma_a || !exp->sigma_e)
        throw HaseError(HASE_E_INVALID, "at least one (sigma_a, sigma_e) pair is required");
    for (uint32_t w = 0; w < exp->number_of_wavelengths; ++w)
        if (!std::isfinite(exp->sigma_a[w]) || !std::isfinite(exp->sigma_e[w]) || exp->sigma_a[w] < 0.0 ||
            exp->sigma_e[w] < 0.0)
            throw HaseError(HASE_E_INVALID, "cross sections must be finite and >= 0");
}

// Face planes, neighbours, volumes and lumped vertex volumes on the host,
// in the operation order of the oracle so the walks match bit for bit.
static void prepare(hase_forward* f, const hase_tet_mesh_t* mesh)
{
    const uint32_t n = mesh->number_of_cells, np = mesh->number_of_points;
    const double* px = mesh->points;
    const double* py = mesh->points + np;
    const double* pz = mesh->points + 2 * (size_t)np;
    std::vector<double4> planes(4 * (size_t)n);
    std::vector<int32_t> link(4 * (size_t)n, -1);
    struct Key {
        uint32_t v[3];
        uint32_t slot;
    };
    std::vector<Key> keys(4 * (size_t)n);
    std::vector<double> lumped(2 * (size_t)np, 0.0);
    f->verts.assign(mesh->cells, mesh->cells + 4 * (size_t)n);
    f->clad.assign(n, 0);
    f->volume.assign(n, 0.0);
    for (uint32_t c = 0; c < n; ++c) {
        const uint32_t* cv = mesh->cells + 4 * (size_t)c;
        f->clad[c] = (mesh->cladding_cell_types && mesh->cladding_cell_types[c] == mesh->cladding_number) ? 1 : 0;
        for (uint32_t k = 0; k < 4; ++k) {
            uint32_t fv[3], j = 0;
            for (uint32_t q = 0; q < 4; ++q)
                if (q != k) fv[j++] = cv[q];
            std::sort(fv, fv + 3);
            const uint32_t o = cv[k];
            const double ax = px[fv[0]], ay = py[fv[0]], az = pz[fv[0]];
            const double bx = px[fv[1]] - ax, by = py[fv[1]] - ay, bz = pz[fv[1]] - az;
            const double cx = px[fv[2]] - ax, cy = py[fv[2]] - ay, cz = pz[fv[2]] - az;
            const double nx = by * cz - bz * cy, ny = bz * cx - bx * cz, nz = bx * cy - by * cx;
            const double den = nx * (px[o] - ax) + ny * (py[o] - ay) + nz * (pz[o] - az);
            if (std::fabs(den) <= DBL_EPSILON)
                throw HaseError(HASE_E_INVALID, "cell " + std::to_string(c) + " is degenerate");
            const double s = 1.0 / den;
            double4 pl;
            pl.x = nx * s;
            pl.y = ny * s;
            pl.z = nz * s;
            pl.w = -(pl.x * ax + pl.y * ay + pl.z * az);
            planes[4 * (size_t)c + k] = pl;
            keys[4 * (size_t)c + k] = Key{{fv[0], fv[1], fv[2]}, 4 * c + k};
        }
        const double ax = px[cv[0]], ay = py[cv[0]], az = pz[cv[0]];
        const double bx = px[cv[1]] - ax, by = py[cv[1]] - ay, bz = pz[cv[1]] - az;
        const double cx = px[cv[2]] - ax, cy = py[cv[2]] - ay, cz = pz[cv[2]] - az;
        const double dx = px[cv[3]] - ax, dy = py[cv[3]] - ay, dz = pz[cv[3]] - az;
        const double vol =
            std::fabs((by * cz - bz * cy) * dx + (bz * cx - bx * cz) * dy + (bx * cy - by * cx) * dz) / 6.0;
        f->volume[c] = vol;
        const uint32_t off = f->clad[c] ? np : 0u;
        for (uint32_t k = 0; k < 4; ++k) lumped[off + cv[k]] += vol / 4.0;
    }
    std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
        for (int i = 0; i < 3; ++i)
            if (a.v[i] != b.v[i]) return a.v[i] < b.v[i];
        return a.slot < b.slot;
    });
    auto same = [](const Key& a, const Key& b) { return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2]; };
    for (size_t i = 0; i + 1 < keys.size(); ++i) {
        if (!same(keys[i], keys[i + 1])) continue;
        if (i + 2 < keys.size() && same(keys[i], keys[i + 2]))
            throw HaseError(HASE_E_INVALID, "a face is shared by more than two cells (cell " +
                                                std::to_string(keys[i].slot / 4) + ")");
        const uint32_t a = keys[i].slot, b = keys[i + 1].slot;
        link[a] = (int32_t)((b / 4) * 4 + b % 4);
        link[b] = (int32_t)((a / 4) * 4 + a % 4);
        ++i;
    }
    f->planes.upload(planes.data(), planes.size(), f->stream);
    f->aux_host.assign(n, CellAux{});
    for (uint32_t c = 0; c < n; ++c) {
        CellAux& ax = f->aux_host[c];
        for (uint32_t k = 0; k < 4; ++k) {
            ax.link[k] = link[4 * (size_t)c + k];
            ax.verts[k] = f->verts[4 * (size_t)c + k];
        }
        ax.clad = f->clad[c];
        for (uint32_t k = 0; k < 4; ++k) ax.bface[k] = -1;
    }
    f->px.upload(px, np, f->stream);
    f->py.upload(py, np, f->stream);
    f->pz.upload(pz, np, f->stream);
    f->lumped.upload(lumped.data(), lumped.size(), f->stream);
    f->volume_dev.upload(f->volume.data(), n, f->stream);
    HIP_CHECK(hipStreamSynchronize(f->stream));
}

static void set_beta(hase_forward* f, const double* beta)
{
    f->beta.assign(beta, beta + f->n_cells);
    f->prefix.resize(f->n_cells);
    double run = 0.0;
    for (uint32_t c = 0; c < f->n_cells; ++c) {
        run += f->beta[c] * f->volume[c];
        f->prefix[c] = run;
    }
    for (uint32_t c = 0; c < f->n_cells; ++c) f->aux_host[c].beta = f->beta[c];
    f->aux.upload(f->aux_host.data(), f->n_cells, f->stream);
    f->prefix_dev.upload(f->prefix.data(), f->n_cells, f->stream);
    HIP_CHECK(hipStreamSynchronize(f->stream));
}

static std::unique_ptr<hase_forward> create(const hase_tet_mesh_t* mesh, const hase_experiment_t* exp,
                                            const hase_compute_t* comp)
{
    validate(mesh, exp, comp);
    auto f = std::make_unique<hase_forward>();
    f->device = comp->device;
    HIP_CHECK(hipSetDevice(f->device));
    HIP_CHECK(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
    f->n_cells = mesh->number_of_cells;
    f->n_points = mesh->number_of_points;
    f->n_lambda = exp->number_of_wavelengths;
    f->min_rays = exp->min_rays_per_sample;
    f->max_rays = exp->max_rays_per_sample;
    f->steps = comp->adaptive_steps;
    f->seed = comp->rng_seed;
    f->threshold = exp->mse_threshold;
    f->max_sigma_a = exp->max_sigma_a;
    f->max_sigma_e = exp->max_sigma_e;
    f->n_tot = mesh->n_tot;
    f->tau = mesh->crystal_t_fluo;
    f->clad_alpha = mesh->cladding_absorption;
    f->reflect = exp->use_reflections != 0;
    prepare(f.get(), mesh);
    set_beta(f.get(), mesh->beta_volume);
    f->sigma_a.upload(exp->sigma_a, f->n_lambda, f->stream);
    f->sigma_e.upload(exp->sigma_e, f->n_lambda, f->stream);
    f->scores.alloc((size_t)kCopies * kBatches * 2u * f->n_points);
    f->counters.alloc(3 * (size_t)f->n_cells + 1);
    f->segments.alloc(1);
    f->phi.alloc(f->n_cells);
    f->se.alloc(f->n_cells);
    f->rse.alloc(f->n_cells);
    f->dndt.alloc(f->n_cells);
    HIP_CHECK(hipStreamSynchronize(f->stream));
    return f;
}

static void fold(hase_forward* f)
{
    const size_t n = (size_t)kBatches * 2u * f->n_points;
    hipLaunchKernelGGL(hase_forward_fold_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, f->stream,
                       f->scores.p, n);
    HIP_CHECK(hipGetLastError());
}

static void clear(hase_forward* f)
{
    const size_t ns = (size_t)kCopies * kBatches * 2u * f->n_points, nc = 3 * (size_t)f->n_cells + 1;
    const size_t n = std::max(ns, nc);
    hipLaunchKernelGGL(hase_forward_clear_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, f->stream,
                       f->scores.p, ns, f->counters.p, nc);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemsetAsync(f->segments.p, 0, sizeof(unsigned long long), f->stream));
}

static TraceArgs trace_args(hase_forward* f, uint32_t seed, uint32_t launch, uint32_t rays)
{
    TraceArgs a{};
    a.m = f->view();
    a.sigma_a = f->sigma_a.p;
    a.sigma_e = f->sigma_e.p;
    a.n_lambda = f->n_lambda;
    a.seed = seed;
    a.launch = launch;
    a.rays = rays;
    for (uint32_t b = 0; b < kBatches; ++b) {
        uint32_t o[4];
        philox4x32_10(0xFFFFFFFFu, b, launch, 2u, seed, kKeyTag, o);
        a.shift[b] = (double)u32_to_unit(o[0]);
        a.phase[b] = o[1] % f->n_lambda;
    }
    a.scores = f->scores.p;
    a.visits = f->counters.p;
    a.dropped = f->counters.p + f->n_cells;
    a.segments = f->segments.p;
    return a;
}

// One launch of `rays` rays; returns its device time in ms.
static SrmArgs srm_args(hase_forward* f, int out_side, uint32_t pass)
{
    SrmArgs s{};
    s.faces = f->bfaces.p;
    s.nb = f->nb;
    s.slots = f->slots;
    s.pass = pass;
    s.out = Reservoir{f->res_count[out_side].p, f->res_weight[out_side].p, f->res_slots[out_side].p};
    return s;
}

static void clear_side(hase_forward* f, int side)
{
    const size_t n = (size_t)kBatches * f->nb;
    hipLaunchKernelGGL(hase_forward_clear_reservoir_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       f->stream, f->res_count[side].p, f->res_weight[side].p, n);
    HIP_CHECK(hipGetLastError());
}

// Normalised face prefix of a reservoir side per batch; returns the totals.
static void face_totals(hase_forward* f, int side, double totals[kBatches])
{
    hipLaunchKernelGGL(hase_forward_face_cdf_kernel, dim3(kBatches), dim3(1024), 0, f->stream,
                       f->res_weight[side].p, f->nb, f->face_cdf.p, f->face_totals.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(totals, f->face_totals.p, kBatches * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipStreamSynchronize(f->stream));
}

// One launch of `rays` rays (with reflected passes when reflections are on);
// returns the device time of its ray kernels in ms.
static double launch_trace(hase_forward* f, uint32_t seed, uint32_t launch, uint32_t rays)
{
    if (rays == 0 || !(f->prefix.back() > 0.0)) return 0.0;
    const TraceArgs a = trace_args(f, seed, launch, rays);
    const bool srm = f->reflect && f->nb > 0 && f->slots > 0;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    const unsigned blocks = (unsigned)(((uint64_t)rays + kTraceBlock - 1) / kTraceBlock);
    double ms_total = 0.0;
    auto timed = [&](auto&& enqueue) {
        HIP_CHECK(hipEventRecord(e0, f->stream));
        enqueue();
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipEventRecord(e1, f->stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        ms_total += ms;
    };
    if (!srm) {
        timed([&] {
            hipLaunchKernelGGL(hase_forward_trace_kernel<false>, dim3(blocks), dim3(kTraceBlock), 0, f->stream, a,
                               SrmArgs{});
        });
    } else {
        // forwardSrm.hpp:120-366, per RSE batch: primary rays deposit into
        // side 0, then passes alternate sides until every batch has stopped
        clear_side(f, 0);
        timed([&] {
            hipLaunchKernelGGL(hase_forward_trace_kernel<true>, dim3(blocks), dim3(kTraceBlock), 0, f->stream, a,
                               srm_args(f, 0, 0));
        });
        double initial[kBatches], previous[kBatches], current[kBatches];
        uint32_t grow[kBatches] = {0};
        bool active[kBatches];
        face_totals(f, 0, initial);
        for (uint32_t b = 0; b < kBatches; ++b) {
            previous[b] = initial[b];
            active[b] = initial[b] > 0.0;
            f->srm_status[b] = active[b] ? HASE_SRM_MAX_PASSES : HASE_SRM_CONVERGED;
            f->srm_passes[b] = 0;
            f->srm_remaining[b] = active[b] ? 1.0 : 0.0;
        }
        int in = 0;
        for (uint32_t pass = 1; pass <= f->max_passes; ++pass) {
            bool any = false;
            for (uint32_t b = 0; b < kBatches; ++b) any |= active[b];
            if (!any) break;
            const int out = 1 - in;
            clear_side(f, out);
            PassArgs q{};
            q.in = Reservoir{f->res_count[in].p, f->res_weight[in].p, f->res_slots[in].p};
            q.cdf = f->face_cdf.p;
            for (uint32_t b = 0; b < kBatches; ++b) {
                const uint32_t count = rays / kBatches + (b < rays % kBatches ? 1u : 0u);
                q.source[b] = active[b] && count ? previous[b] / (double)count : 0.0;
                uint32_t o[4];
                philox4x32_10(0xFFFFFFFEu, b, launch, pass, seed, kKeyTag, o);
                q.shift[b] = (double)u32_to_unit(o[0]);
            }
            timed([&] {
                hipLaunchKernelGGL(hase_forward_reflect_kernel, dim3(blocks), dim3(kTraceBlock), 0, f->stream, a,
                                   srm_args(f, out, pass), q);
            });
            face_totals(f, out, current);
            // stop rules (forwardSrm.hpp:322-361)
            for (uint32_t b = 0; b < kBatches; ++b) {
                if (!active[b]) continue;
                const double w = current[b];
                f->srm_passes[b] = pass;
                f->srm_remaining[b] = w / initial[b];
                if (w > previous[b]) {
                    if (++grow[b] >= f->streak) {
                        f->srm_status[b] = HASE_SRM_DIVERGED;
                        active[b] = false;
                        continue;
                    }
                } else {
                    grow[b] = 0;
                    if (std::fabs(w - previous[b]) / std::max(w, 1.0e-30) < f->tolerance) {
                        f->srm_status[b] = HASE_SRM_STABLE;
                        active[b] = false;
                        continue;
                    }
                }
                if (f->srm_remaining[b] < f->tolerance) {
                    f->srm_status[b] = HASE_SRM_CONVERGED;
                    active[b] = false;
                    continue;
                }
                previous[b] = w;
            }
            in = out;
        }
        for (uint32_t b = 0; b < kBatches; ++b) {
            f->stats.srm_passes = std::max(f->stats.srm_passes, f->srm_passes[b]);
            f->stats.srm_remaining = std::max(f->stats.srm_remaining, f->srm_remaining[b]);
            f->stats.srm_status = std::max(f->stats.srm_status, (uint32_t)f->srm_status[b]);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms_total;
}

// Reflective boundary faces and their reservoirs (hase_forward_set_reflections).
static void set_reflections(hase_forward* f, const hase_forward_reflections_t* r)
{
    if (!r) throw HaseError(HASE_E_INVALID, "reflections must not be NULL");
    if (!r->face_boundaries) throw HaseError(HASE_E_INVALID, "face_boundaries must not be NULL");
    if (r->number_of_surfaces > 0 && !r->reflectivity)
        throw HaseError(HASE_E_INVALID, "reflectivity is required for the surfaces");
    for (uint32_t i = 0; i < r->number_of_surfaces; ++i) {
        if (r->reflectivity && (!std::isfinite(r->reflectivity[i]) || r->reflectivity[i] < 0.0f ||
                                r->reflectivity[i] > 1.0f))
            throw HaseError(HASE_E_INVALID, "surface reflectivities must lie in [0, 1]");
        if ((r->n_inside && !(r->n_inside[i] >= 0.0f)) || (r->n_outside && !(r->n_outside[i] >= 0.0f)))
            throw HaseError(HASE_E_INVALID, "refractive indices must be >= 0");
    }
    if (!(r->tolerance > 0.0) || r->max_passes == 0 || r->divergence_streak == 0)
        throw HaseError(HASE_E_INVALID, "tolerance must be > 0, max_passes and divergence_streak >= 1");
    const uint32_t n = f->n_cells;
    const uint32_t* cells = f->verts.data();
    std::vector<double> hx(f->n_points), hy(f->n_points), hz(f->n_points);
    HIP_CHECK(hipMemcpy(hx.data(), f->px.p, f->n_points * sizeof(double), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(hy.data(), f->py.p, f->n_points * sizeof(double), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(hz.data(), f->pz.p, f->n_points * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<BFace> faces;
    for (uint32_t c = 0; c < n; ++c) {
        CellAux& ax = f->aux_host[c];
        for (uint32_t k = 0; k < 4; ++k) {
            ax.bface[k] = -1;
            const int32_t id = r->face_boundaries[4 * (size_t)c + k];
            if (ax.link[k] >= 0 || id <= 0 || (uint32_t)id >= r->number_of_surfaces) continue;
            const uint32_t* cv = cells + 4 * (size_t)c;
            uint32_t fv[3], j = 0;
            for (uint32_t q = 0; q < 4; ++q)
                if (q != k) fv[j++] = cv[q];
            std::sort(fv, fv + 3);
            BFace bf{};
            // face centroid and outward normal (rayWalk.hpp:42-84)
            bf.c[0] = (hx[fv[0]] + hx[fv[1]] + hx[fv[2]]) * (1.0 / 3.0);
            bf.c[1] = (hy[fv[0]] + hy[fv[1]] + hy[fv[2]]) * (1.0 / 3.0);
            bf.c[2] = (hz[fv[0]] + hz[fv[1]] + hz[fv[2]]) * (1.0 / 3.0);
            const double ctr[3] = {(hx[cv[0]] + hx[cv[1]] + hx[cv[2]] + hx[cv[3]]) * 0.25,
                                   (hy[cv[0]] + hy[cv[1]] + hy[cv[2]] + hy[cv[3]]) * 0.25,
                                   (hz[cv[0]] + hz[cv[1]] + hz[cv[2]] + hz[cv[3]]) * 0.25};
            const double bx = hx[fv[1]] - hx[fv[0]], by = hy[fv[1]] - hy[fv[0]], bz = hz[fv[1]] - hz[fv[0]];
            const double cx = hx[fv[2]] - hx[fv[0]], cy = hy[fv[2]] - hy[fv[0]], cz = hz[fv[2]] - hz[fv[0]];
            double nr[3] = {by * cz - bz * cy, bz * cx - bx * cz, bx * cy - by * cx};
            const double len = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
            const double inv = len <= DBL_EPSILON ? 0.0 : 1.0 / len;
            for (int d = 0; d < 3; ++d) nr[d] = nr[d] * inv;
            if (nr[0] * (ctr[0] - bf.c[0]) + nr[1] * (ctr[1] - bf.c[1]) + nr[2] * (ctr[2] - bf.c[2]) > 0.0)
                for (int d = 0; d < 3; ++d) nr[d] = -nr[d];
            for (int d = 0; d < 3; ++d) bf.n[d] = nr[d];
            bf.refl = r->reflectivity ? (double)r->reflectivity[id] : 0.0;
            bf.n_in = r->n_inside ? (double)r->n_inside[id] : 0.0;
            bf.n_out = r->n_outside ? (double)r->n_outside[id] : 0.0;
            bf.face = 4 * c + k;
            ax.bface[k] = (int32_t)faces.size();
            faces.push_back(bf);
        }
    }
    f->nb = (uint32_t)faces.size();
    f->slots = r->reservoir_slots;
    f->max_passes = r->max_passes;
    f->streak = r->divergence_streak;
    f->tolerance = r->tolerance;
    if (f->nb && (size_t)kBatches * f->nb * f->slots > ((size_t)1 << 31))
        throw HaseError(HASE_E_INVALID, "reservoir too large: reduce reservoir_slots");
    f->aux.upload(f->aux_host.data(), n, f->stream);
    if (f->nb) {
        f->bfaces.upload(faces.data(), faces.size(), f->stream);
        for (int side = 0; side < 2; ++side) {
            f->res_count[side].alloc((size_t)kBatches * f->nb);
            f->res_weight[side].alloc((size_t)kBatches * f->nb);
            f->res_slots[side].alloc(std::max<size_t>(1, (size_t)kBatches * f->nb * f->slots));
        }
        f->face_cdf.alloc((size_t)kBatches * f->nb);
        f->face_totals.alloc(kBatches);
    }
    HIP_CHECK(hipStreamSynchronize(f->stream));
}

static uint32_t adaptive_target(uint32_t min_rays, uint32_t max_rays, uint32_t steps, uint32_t k)
{
    if (max_rays <= min_rays || steps == 0) return min_rays;
    if (k >= steps) return max_rays;
    const double growth = std::pow((double)max_rays / (double)min_rays, 1.0 / (double)steps);
    double r = std::ceil((double)min_rays * std::pow(growth, (double)k));
    r = std::min(std::max(r, (double)min_rays), (double)max_rays);
    return (uint32_t)r;
}

static void evaluate(hase_forward* f, hase_forward_result_t* out)
{
    if (!out || !out->phi_ase) throw HaseError(HASE_E_INVALID, "result.phi_ase must not be NULL");
    HIP_CHECK(hipSetDevice(f->device));
    const auto t0 = std::chrono::steady_clock::now();
    clear(f);
    f->stats = hase_forward_stats_t{};
    uint32_t counts[kBatches] = {0};
    uint64_t done = 0;
    uint32_t launches = 0, open = 0;
    for (uint32_t k = 0;; ++k) {
        const uint32_t target = adaptive_target(f->min_rays, f->max_rays, f->steps, k);
        const uint32_t rays = (uint32_t)(target - done);
        for (uint32_t b = 0; b < kBatches; ++b) counts[b] += rays / kBatches + (b < rays % kBatches ? 1u : 0u);
        f->stats.trace_ms += launch_trace(f, f->seed, launches, rays);
        done = target;
        ++launches;
        fold(f);
        FinalizeArgs fa{};
        fa.m = f->view();
        fa.volume = f->volume_dev.p;
        fa.lumped = f->lumped.p;
        fa.scores = f->scores.p;
        fa.dropped = f->counters.p + f->n_cells;
        for (uint32_t b = 0; b < kBatches; ++b) fa.counts[b] = counts[b];
        fa.rays = (double)done;
        fa.total = f->prefix.back();
        fa.rate = f->n_tot / f->tau;
        fa.sigma_a = f->max_sigma_a;
        fa.sigma_e = f->max_sigma_e;
        fa.threshold = f->threshold;
        fa.target = target;
        fa.phi = f->phi.p;
        fa.se = f->se.p;
        fa.rse = f->rse.p;
        fa.dndt = f->dndt.p;
        fa.convergence = f->counters.p + 2 * (size_t)f->n_cells;
        fa.open_cells = f->counters.p + 3 * (size_t)f->n_cells;
        HIP_CHECK(hipMemsetAsync(fa.open_cells, 0, sizeof(uint32_t), f->stream));
        hipLaunchKernelGGL(hase_forward_finalize_kernel, dim3((f->n_cells + 255) / 256), dim3(256), 0, f->stream,
                           fa);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(&open, fa.open_cells, sizeof(uint32_t), hipMemcpyDeviceToHost, f->stream));
        HIP_CHECK(hipStreamSynchronize(f->stream));
        if (open == 0 || target >= f->max_rays || f->steps == 0 || f->max_rays <= f->min_rays) break;
    }
    const size_t n = f->n_cells;
    HIP_CHECK(hipMemcpyAsync(out->phi_ase, f->phi.p, n * sizeof(float), hipMemcpyDeviceToHost, f->stream));
    if (out->standard_error)
        HIP_CHECK(hipMemcpyAsync(out->standard_error, f->se.p, n * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    if (out->relative_standard_error)
        HIP_CHECK(hipMemcpyAsync(out->relative_standard_error, f->rse.p, n * sizeof(double), hipMemcpyDeviceToHost,
                                 f->stream));
    if (out->dndt_ase)
        HIP_CHECK(hipMemcpyAsync(out->dndt_ase, f->dndt.p, n * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    if (out->ray_visits)
        HIP_CHECK(hipMemcpyAsync(out->ray_visits, f->counters.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 f->stream));
    if (out->dropped_rays)
        HIP_CHECK(hipMemcpyAsync(out->dropped_rays, f->counters.p + n, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 f->stream));
    if (out->convergence_rays)
        HIP_CHECK(hipMemcpyAsync(out->convergence_rays, f->counters.p + 2 * n, n * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, f->stream));
    unsigned long long seg = 0;
    std::vector<uint32_t> dropped(n);
    HIP_CHECK(hipMemcpyAsync(&seg, f->segments.p, sizeof(seg), hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipMemcpyAsync(dropped.data(), f->counters.p + n, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             f->stream));
    HIP_CHECK(hipStreamSynchronize(f->stream));
    f->stats.rays = done;
    f->stats.segments = seg;
    f->stats.dropped_rays = std::accumulate(dropped.begin(), dropped.end(), (uint64_t)0);
    f->stats.launches = launches;
    f->stats.open_cells = open;
    f->stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace fwd
}  // namespace hase

extern "C" {

int hase_forward_create(const hase_tet_mesh_t* mesh, const hase_experiment_t* experiment,
                        const hase_compute_t* compute, hase_forward_t** out)
{
    return hase::guarded([&] {
        if (!out) throw hase::HaseError(HASE_E_INVALID, "out must not be NULL");
        *out = nullptr;
        *out = hase::fwd::create(mesh, experiment, compute).release();
    });
}

int hase_forward_update_beta(hase_forward_t* f, const double* beta_volume)
{
    return hase::guarded([&] {
        if (!f || !beta_volume) throw hase::HaseError(HASE_E_INVALID, "forward and beta_volume must not be NULL");
        for (uint32_t c = 0; c < f->n_cells; ++c)
            if (!std::isfinite(beta_volume[c])) throw hase::HaseError(HASE_E_INVALID, "beta_volume must be finite");
        HIP_CHECK(hipSetDevice(f->device));
        hase::fwd::set_beta(f, beta_volume);
    });
}

int hase_forward_evaluate(hase_forward_t* f, hase_forward_result_t* result)
{
    return hase::guarded([&] {
        if (!f) throw hase::HaseError(HASE_E_INVALID, "forward must not be NULL");
        hase::fwd::evaluate(f, result);
    });
}

int hase_forward_trace(hase_forward_t* f, uint32_t seed, uint32_t launch, uint32_t rays, double* scores,
                       uint32_t* ray_visits, uint32_t* dropped_rays, uint64_t* segments)
{
    return hase::guarded([&] {
        if (!f || !scores || !ray_visits || !dropped_rays)
            throw hase::HaseError(HASE_E_INVALID, "forward, scores, ray_visits and dropped_rays must not be NULL");
        HIP_CHECK(hipSetDevice(f->device));
        hase::fwd::clear(f);
        f->stats = hase_forward_stats_t{};
        f->stats.trace_ms = hase::fwd::launch_trace(f, seed, launch, rays);
        hase::fwd::fold(f);
        const size_t n = f->n_cells;
        HIP_CHECK(hipMemcpyAsync(scores, f->scores.p, (size_t)hase::fwd::kBatches * 2u * f->n_points * sizeof(double),
                                 hipMemcpyDeviceToHost, f->stream));
        HIP_CHECK(hipMemcpyAsync(ray_visits, f->counters.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, f->stream));
        HIP_CHECK(hipMemcpyAsync(dropped_rays, f->counters.p + n, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 f->stream));
        unsigned long long seg = 0;
        HIP_CHECK(hipMemcpyAsync(&seg, f->segments.p, sizeof(seg), hipMemcpyDeviceToHost, f->stream));
        HIP_CHECK(hipStreamSynchronize(f->stream));
        if (segments) *segments = seg;
    });
}

int hase_forward_set_reflections(hase_forward_t* f, const hase_forward_reflections_t* reflections)
{
    return hase::guarded([&] {
        if (!f) throw hase::HaseError(HASE_E_INVALID, "forward must not be NULL");
        HIP_CHECK(hipSetDevice(f->device));
        hase::fwd::set_reflections(f, reflections);
    });
}

int hase_forward_stats(const hase_forward_t* f, hase_forward_stats_t* stats)
{
    return hase::guarded([&] {
        if (!f || !stats) throw hase::HaseError(HASE_E_INVALID, "forward and stats must not be NULL");
        *stats = f->stats;
    });
}

int hase_forward_destroy(hase_forward_t* f)
{
    return hase::guarded([&] {
        if (!f) return;
        (void)hipSetDevice(f->device);
        (void)hipStreamSynchronize(f->stream);
        delete f;
    });
}

int hase_calc_forward_phi_ase(const hase_tet_mesh_t* mesh, const hase_experiment_t* experiment,
                              const hase_compute_t* compute, hase_forward_result_t* result)
{
    return hase::guarded([&] {
        auto f = hase::fwd::create(mesh, experiment, compute);
        hase::fwd::evaluate(f.get(), result);
    });
}

}  // extern "C"
