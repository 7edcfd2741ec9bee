// hase_ase.hip - MI355X (gfx950) importance-sampled ASE ray tracer behind the
// C ABI of include/hase_ase.h.
//
// Device pipeline per batch of sample points:
//   1. hase_importance_kernel (one work-group per sample): walks from every
//      prism centroid (and its mirror slices) to the sample point, forms the
//      defensive-mixture proposal and writes its FP64 inclusive prefix.  This
//      fuses the legacy importanceSamplingPropagation and
//      importanceSamplingDistribution passes (retired; shells at
//      include/kernels/importanceSampling.hpp:10-18).
//   2. hase_trace_kernel (work-groups of one sample's ray chunk): maps each ray
//      to its (slice, prism) by systematic sampling on the device prefix (the
//      legacy host loop mapRaysToPrisms, shell include/kernels/mapRaysToTets.hpp),
//      draws the start point from a counter-based Philox stream, walks the
//      prisms to the sample point in FP32 (legacy propagateRay) and reduces
//      the estimator's sum and sum of squares in FP64 with wavefront shuffles.
//   3. hase_finalize_kernel: pools the chunk partials in a fixed order, forms
//      phiAse / mse / dndtAse in FP64 and compacts the samples that still need
//      rays (adaptive escalation, include/core/types.hpp:333-354).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hase_ase.h"
#include "common.h"
#include "context.h"
#include "prism_walk.h"
#include "importance_walk.h"
#include "proposal.h"
#include "device_common.h"

namespace {

using namespace hase;

// ------------------------------------------------------------------ device

// Defensive-mixture proposal and its FP64 inclusive prefix for one sample
// (one work-group): q_e = (1 - eps) imp_e / I + eps * bv_j / (R * Bsum), with
// I the work-group sum of the thread partials `local`.  value(e) returns
// imp_e and must be readable by every thread (the entries of other threads
// are read after a barrier).  red holds kBlock/64 + 2 doubles of LDS.
constexpr int kScanPer = 8;

// (slice k, prism j = t + l*T) of a proposal entry e = k*P + j.
struct EntryIndex {
    uint32_t k, j, t, l;
    __device__ __forceinline__ void init(uint32_t e, uint32_t P, uint32_t T)
    {
        k = e / P;
        j = e - k * P;
        l = j / T;
        t = j - l * T;
    }
};

// Entry e = k * P + l * T + t of loops that visit e, e + step, e + 2 step, ...:
// decoded once, then advanced without an integer division per entry.
struct EntryCursor {
    EntryIndex x;
    uint32_t T, layers;
    __device__ __forceinline__ void start(uint32_t e, uint32_t P, uint32_t T_)
    {
        T = T_;
        layers = P / T_;
        x.init(e, P, T_);
    }
    __device__ __forceinline__ void advance(uint32_t step)
    {
        uint32_t t = x.t + step, l = x.l, k = x.k;
        while (t >= T) {
            t -= T;
            if (++l == layers) {
                l = 0;
                ++k;
            }
        }
        x.t = t;
        x.l = l;
        x.k = k;
        x.j = t + l * T;
    }
};

// LDS for one pass of mixture_prefix: kBlock * kScanPer doubles, one pad
// slot per kScanPer so a thread's consecutive run is bank-conflict free.
constexpr int kScanLds = kBlock * kScanPer + kBlock * kScanPer / kScanPer;
__device__ __forceinline__ int scan_slot(int i) { return i + i / kScanPer; }

// Defensive-mixture proposal and its FP64 inclusive prefix for one sample
// (one work-group): q_e = (1 - eps) imp_e / I + eps * bv_j / (R * Bsum), with
// I the work-group sum of the thread partials `local`.  value(x, e) returns
// imp_e.  Each pass covers kBlock * kScanPer entries: they are computed and
// stored in coalesced order (entry base + i*kBlock + tid), while the scan
// runs over kScanPer consecutive entries per thread, transposed through LDS.
// red holds kBlock/64 + 2 doubles, buf kScanLds doubles of LDS.
template <class F>
__device__ __forceinline__ void mixture_prefix(double local, uint32_t n, uint32_t P, uint32_t T, const float* bv,
                                               double eps, double Bsum, uint32_t R, F value, double* pre,
                                               double* Qout, double* red, double* buf, uint32_t* bucket)
{
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double* carry_s = red + kBlock / 64;
    double* total_s = red + kBlock / 64 + 1;
    // buckets beyond the total (rounding) keep the last entry
    if (bucket)
        for (uint32_t k = (uint32_t)tid; k < kPrefixBuckets; k += (uint32_t)kBlock) bucket[k] = n - 1u;
    local = wave_sum(local);
    if (lane == 0) red[wid] = local;
    __syncthreads();
    if (tid == 0) {
        double I = 0.0;
        for (int i = 0; i < kBlock / 64; ++i) I += red[i];
        *total_s = I;
        *carry_s = 0.0;
    }
    __syncthreads();
    const double I = *total_s;
    const bool use_imp = I > 0.0 && isfinite(I);
    const double wI = use_imp ? (1.0 - eps) / I : 0.0;
    const double wB = (use_imp ? eps : 1.0) / (Bsum * (double)R);
    // entries base + i * kBlock + tid: one cursor stepping by kBlock
    EntryCursor cur;
    cur.start((uint32_t)tid, P, T);
    for (uint32_t base = 0; base < n; base += (uint32_t)kBlock * kScanPer) {
#pragma unroll
        for (int i = 0; i < kScanPer; ++i) {
            const uint32_t e = base + (uint32_t)(i * kBlock + tid);
            double q = 0.0;
            if (e < n) q = wI * (double)value(cur.x, e) + wB * (double)bv[cur.x.j];
            buf[scan_slot(i * kBlock + tid)] = q;
            cur.advance((uint32_t)kBlock);
        }
        __syncthreads();
        double v[kScanPer];
        double run = 0.0;
#pragma unroll
        for (int i = 0; i < kScanPer; ++i) {
            run += buf[scan_slot(tid * kScanPer + i)];
            v[i] = run;
        }
        const double inc = wave_inclusive_scan(run, lane);
        double before = __shfl_up(inc, 1, 64);
        if (lane == 0) before = 0.0;
        if (lane == 63) red[wid] = inc;
        __syncthreads();
        double off = *carry_s;
        for (int i = 0; i < wid; ++i) off += red[i];
        const double excl = off + before;
#pragma unroll
        for (int i = 0; i < kScanPer; ++i) buf[scan_slot(tid * kScanPer + i)] = excl + v[i];
        if (bucket) {
            // entry e takes the buckets k with prefix(e - 1) <= k / B < prefix(e),
            // i.e. k in [ceil(prefix(e - 1) B), ceil(prefix(e) B)): B is a power of
            // two, so the products are exact and one ceil per entry suffices
            const double B = (double)kPrefixBuckets;
            uint32_t k0 = (uint32_t)fmin(ceil(excl * B), B);
#pragma unroll
            for (int i = 0; i < kScanPer; ++i) {
                const uint32_t e = base + (uint32_t)(tid * kScanPer + i);
                const uint32_t k1 = (uint32_t)fmin(ceil((excl + v[i]) * B), B);
                if (e < n)
                    for (uint32_t k = k0; k < k1; ++k) bucket[k] = e;
                k0 = k1;
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kScanPer; ++i) {
            const uint32_t e = base + (uint32_t)(i * kBlock + tid);
            if (e < n) pre[e] = buf[scan_slot(i * kBlock + tid)];
        }
        if (tid == kBlock - 1) *carry_s = off + inc;
        __syncthreads();
    }
    if (tid == 0) *Qout = *carry_s;
}

struct ImportanceParams {
    MeshDev m;
    const float2* pts;
    uint32_t sample_base;     // first sample of the batch
    uint32_t slots;           // samples in this batch
    uint32_t R;               // reflection slices
    uint32_t n_lambda;
    const float* c1;          // [n_lambda] nTot (sigma_e + sigma_a) log2(e)
    const float* c2;          // [n_lambda] nTot sigma_a log2(e)
    float inv_lambda;
    double eps;               // defensive fraction
    double Bsum;              // sum of beta * V
    float* imp;               // [slots * R * P] scratch
    double* prefix;           // [slots * R * P]
    double* Q;                // [slots]
    uint32_t* buckets = nullptr; // [slots * kPrefixBuckets] or null
    unsigned long long* counters;
    const uint32_t* active;   // block -> slot (nullptr: block index is the slot)
    const uint32_t* sample_of;// slot -> sample (nullptr: sample_base + slot)
    int transfer;             // write the transfer factor gain * reflectance instead of the importance
    uint32_t parts;           // work-groups per sample; > 1: importance only, hase_prefix_kernel follows
};

// kKind: the records the walks read from global memory: 0 the 48-byte
// records, kCompactGlobal the 32-byte compact records.
template <uint32_t kKind>
__global__ void __launch_bounds__(kBlock) hase_importance_kernel(ImportanceParams p)
{
    const uint32_t row = blockIdx.x / p.parts, part = blockIdx.x % p.parts;
    const uint32_t slot = p.active ? p.active[row] : row;
    if (slot >= p.slots) return;
    const MeshDev& m = p.m;
    const uint32_t T = m.T, P = T * m.layers;
    const size_t n = (size_t)p.R * P;
    const uint32_t sample = p.sample_of ? p.sample_of[slot] : p.sample_base + slot;
    SampleGeom sp = sample_point(p.pts, m.np, sample);
    float* imp = p.imp + (size_t)slot * n;
    double* pre = p.prefix + (size_t)slot * n;
    const int tid = threadIdx.x, lane = tid & 63;
    __shared__ double red[kBlock / 64 + 2];
    __shared__ double buf[kScanLds];

    double local = 0.0;
    unsigned long long walks = 0, steps = 0;
    const ColumnItems ci(T, m.layers, p.R);
    const uint32_t items = ci.per_sample();
    // this work-group's share of the sample's column walks
    const uint32_t it0 = (uint32_t)((uint64_t)items * part / p.parts);
    const uint32_t it1 = (uint32_t)((uint64_t)items * (part + 1u) / p.parts);
    for (uint32_t it = it0 + tid; it < it1; it += kBlock) {
        uint32_t t0, k;
        int lv0, nlev;
        ci.decode(it, t0, k, lv0, nlev);
        float v[kLevelGroup];
        if constexpr (kKind != 0u)
            steps += column_entries(m, staged_fetch<kKind>(m.lds_rec, T), p.c1, p.c2, p.n_lambda, p.inv_lambda, t0, lv0,
                                    nlev, k, sp.px, sp.py, sp.plane, p.transfer != 0, v);
        else
            steps += column_entries(m, GlobalTriangles{}, p.c1, p.c2, p.n_lambda, p.inv_lambda, t0, lv0, nlev, k, sp.px,
                                    sp.py, sp.plane, p.transfer != 0, v);
        walks += (unsigned long long)nlev;
#pragma unroll
        for (int g = 0; g < kLevelGroup; ++g) {
            if (g < nlev) {
                imp[(size_t)k * P + t0 + (uint32_t)(lv0 + g) * T] = v[g];
                local += (double)v[g];
            }
        }
    }
    unsigned long long ws = wave_sum_u64(walks), ss = wave_sum_u64(steps);
    if (lane == 0 && ws) {
        atomicAdd(&p.counters[2], ws);
        atomicAdd(&p.counters[3], ss);
    }
    if (p.transfer || p.parts > 1u) return;
    mixture_prefix(local, (uint32_t)n, P, T, m.bv, p.eps, p.Bsum, p.R,
                   [&](const EntryIndex&, uint32_t e) { return imp[e]; }, pre, p.Q + slot, red, buf,
                   p.buckets ? p.buckets + (size_t)slot * kPrefixBuckets : nullptr);
}

// Walked importance of whole batches for meshes of up to kLdsTriangles
// triangles: one persistent 1024-thread work-group per CU copies the compact
// triangle records into LDS once, then its wavefronts take 256 consecutive
// column items of the batch at a time, round-robin over the grid (a batch's
// items are numbered slot-major, so a wavefront's items are neighbouring
// triangle columns of one sample and slice), and walk them in four rounds
// of 64 sorted by estimated work.  Only the
// importance values are written; hase_prefix_kernel forms each sample's
// mixture prefix afterwards (not for transfer maps).
constexpr int kStagedBlock = 1024;

template <uint32_t kKind>
__global__ void __launch_bounds__(kStagedBlock) hase_importance_staged_kernel(ImportanceParams p)
{
    __shared__ uint4 s_tri[kLdsStageWords];
    const MeshDev& m = p.m;
    const uint32_t T = m.T, P = T * m.layers;
    for (uint32_t i = threadIdx.x; i < m.lds_words; i += blockDim.x) s_tri[i] = m.lds_rec[i];
    __syncthreads();
    const auto fetch = staged_fetch<kKind>(s_tri, T);
    const size_t n = (size_t)p.R * P;
    const ColumnItems ci(T, m.layers, p.R);
    const uint32_t per_slot = ci.per_sample();
    const uint64_t total = (uint64_t)p.slots * per_slot;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves_per_block = blockDim.x / 64u;
    const uint64_t wave = (uint64_t)blockIdx.x * waves_per_block + threadIdx.x / 64u;
    const uint64_t stride = (uint64_t)gridDim.x * waves_per_block * 256u;
    unsigned long long steps = 0;
    // Items 256 at a time per wavefront, sorted by estimated work (the
    // column's triangle crossings, from its centroid's xy distance to the
    // sample at 1.2 sqrt(T / mesh area) per cm, times its levels, plus their
    // plane crossings) and taken as four rounds of 64: lanes of a round walk
    // columns of similar length, whose plane crossings fall on similar steps.
    for (uint64_t base = wave * 256u; base < total; base += stride) {
        // slot and item of the group's first item (wavefront-uniform); item
        // base + li is li further on, in the same slot or the next ones
        const uint64_t slot0 = base / per_slot;
        const uint32_t it0 = (uint32_t)(base - slot0 * per_slot);
        auto locate = [&](uint32_t li, uint32_t& slot, uint32_t& it) {
            slot = (uint32_t)slot0;
            it = it0 + li;
            while (it >= per_slot) {
                it -= per_slot;
                ++slot;
            }
            return slot < p.slots;
        };
        uint32_t key[4];
#pragma unroll
        for (uint32_t r = 0; r < 4u; ++r) {
            const uint32_t li = r * 64u + lane;
            key[r] = 0xFFFFFF00u | li;   // past the end: sorted last, skipped
            uint32_t slot, it;
            if (locate(li, slot, it)) {
                const uint32_t sample = p.sample_of ? p.sample_of[slot] : p.sample_base + slot;
                const SampleGeom sp = sample_point(p.pts, m.np, sample);
                uint32_t t0, k;
                int lv0, nlev;
                ci.decode(it, t0, k, lv0, nlev);
                const float2 a = fetch.vertex(m, t0, 0u), b = fetch.vertex(m, t0, 1u), c = fetch.vertex(m, t0, 2u);
                const float dx = (a.x + b.x + c.x) * (1.0f / 3.0f) - sp.px, dy = (a.y + b.y + c.y) * (1.0f / 3.0f) - sp.py;
                const float ze = (float)image_plane(sp.plane, (int)m.layers, (k + 1u) >> 1, k & 1u);
                const float est = (float)nlev * (sqrtf(dx * dx + dy * dy) * m.steps_per_cm +
                                                 fabsf((float)lv0 + 0.5f * (float)nlev - ze));
                key[r] = ((uint32_t)fminf(est * 4.0f, 16777214.0f) << 8) | li;
            }
        }
        wave_sort256(key, (int)lane);
        uint32_t items4 = 0u;   // the four group indices, 8 bits each
#pragma unroll
        for (uint32_t r = 0; r < 4u; ++r) items4 |= (key[r] & 0xFFu) << (8u * r);
#pragma unroll 1
        for (uint32_t r = 0; r < 4u; ++r, items4 >>= 8) {
            uint32_t slot, it;
            if (!locate(items4 & 0xFFu, slot, it)) continue;
            const uint32_t sample = p.sample_of ? p.sample_of[slot] : p.sample_base + slot;
            const SampleGeom sp = sample_point(p.pts, m.np, sample);
            uint32_t t0, k;
            int lv0, nlev;
            ci.decode(it, t0, k, lv0, nlev);
            float v[kLevelGroup];
            steps += column_entries(m, fetch, p.c1, p.c2, p.n_lambda, p.inv_lambda, t0, lv0, nlev, k, sp.px, sp.py,
                                    sp.plane, p.transfer != 0, v);
            float* imp = p.imp + (size_t)slot * n + (size_t)k * P + t0;
#pragma unroll
            for (int g = 0; g < kLevelGroup; ++g)
                if (g < nlev) imp[(size_t)(lv0 + g) * T] = v[g];
        }
    }
    // every entry of every slot is one level walk
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&p.counters[2], (unsigned long long)p.slots * n);
    const unsigned long long ss = wave_sum_u64(steps);
    if (lane == 0 && ss) atomicAdd(&p.counters[3], ss);
}

// beta * V proposal shared by every sample (one work-group): the mixture with
// no importance term, q_e = bv_j / (R * Bsum), its FP64 prefix, total and
// buckets in slot 0 of the prefix tables.
__global__ void __launch_bounds__(kBlock) hase_beta_volume_prefix_kernel(ImportanceParams p)
{
    const MeshDev& m = p.m;
    const uint32_t T = m.T, P = T * m.layers;
    __shared__ double red[kBlock / 64 + 2];
    __shared__ double buf[kScanLds];
    mixture_prefix(0.0, p.R * P, P, T, m.bv, p.eps, p.Bsum, p.R, [&](const EntryIndex&, uint32_t) { return 0.0f; },
                   p.prefix, p.Q, red, buf, p.buckets);
}

// Prefix of importance maps written by split work-groups (one per sample here).
__global__ void __launch_bounds__(kBlock) hase_prefix_kernel(ImportanceParams p)
{
    const uint32_t slot = p.active ? p.active[blockIdx.x] : blockIdx.x;
    if (slot >= p.slots) return;
    const MeshDev& m = p.m;
    const uint32_t T = m.T, P = T * m.layers;
    const uint32_t n = p.R * P;
    const float* imp = p.imp + (size_t)slot * n;
    __shared__ double red[kBlock / 64 + 2];
    __shared__ double buf[kScanLds];
    double local = 0.0;
    for (uint32_t e = threadIdx.x; e < n; e += kBlock) local += (double)imp[e];
    mixture_prefix(local, n, P, T, m.bv, p.eps, p.Bsum, p.R, [&](const EntryIndex&, uint32_t e) { return imp[e]; },
                   p.prefix + (size_t)slot * n, p.Q + slot, red, buf,
                   p.buckets ? p.buckets + (size_t)slot * kPrefixBuckets : nullptr);
}

// Learned proposal for one sample per work-group: the network's log transfer
// factor y of the prism's cluster (one value per reflection slice and
// cluster) times the prism's own beta * V / (4 pi d^2), d the distance from
// the prism centroid to the sample's image for that slice; then the same
// defensive mixture and prefix as the walked importance map.
struct LearnedParams {
    MeshDev m;
    const float2* pts;
    const float2* centroid;   // [T] device order
    const float* logits;      // [slots * R * C]
    const uint32_t* cl_tri;   // [T] cluster column of each device triangle
    const uint32_t* cl_lev;   // [layers] first cluster of each layer's cluster row
    uint32_t slots, sample_base, R, C, TG, ct, cl;
    double eps, Bsum;
    double* prefix;
    double* Q;
    uint32_t* buckets = nullptr;   // [slots * kPrefixBuckets] or null
};

// An entry's slice and layer change only every T entries of a thread's
// ascending run, so the slice's image plane, the layer's row of cluster
// logits and the height difference are kept per thread and recomputed on a
// change; per entry remain the cluster column, the logit, the centroid and
// beta * V.  The reciprocal is the hardware one (1 ulp).
struct LearnedRow {
    uint32_t k = ~0u, l = ~0u;
    float zi = 0.0f, dz = 0.0f;
    const float* row = nullptr;
};

__device__ __forceinline__ float learned_importance(const LearnedParams& p, const SampleGeom& sp, const float* lg,
                                                    const EntryIndex& x, LearnedRow& r)
{
    const MeshDev& m = p.m;
    if (x.k != r.k) {
        r.k = x.k;
        r.zi = (float)image_plane(sp.plane, (int)m.layers, (x.k + 1u) >> 1, x.k & 1u);
        r.l = ~0u;
    }
    if (x.l != r.l) {
        r.l = x.l;
        r.row = lg + (size_t)x.k * p.C + p.cl_lev[x.l];
        r.dz = ((float)x.l + 0.5f - r.zi) * m.h;
    }
    const float y = fminf(r.row[p.cl_tri[x.t]], 80.0f);
    const float2 cxy = p.centroid[x.t];
    const float dx = cxy.x - sp.px, dy = cxy.y - sp.py;
    const float d2 = fmaf(dx, dx, fmaf(dy, dy, r.dz * r.dz));
    const float v = m.bv[x.j] * __expf(y) * __builtin_amdgcn_rcpf(4.0f * kPi * d2);
    return (v >= 0.0f && !isinf(v)) ? v : 0.0f;
}

__global__ void __launch_bounds__(kBlock) hase_learned_importance_kernel(LearnedParams p)
{
    const uint32_t slot = blockIdx.x;
    if (slot >= p.slots) return;
    const MeshDev& m = p.m;
    const uint32_t T = m.T, P = T * m.layers;
    const uint32_t n = p.R * P;
    const SampleGeom sp = sample_point(p.pts, m.np, p.sample_base + slot);
    const float* lg = p.logits + (size_t)slot * p.R * p.C;
    LearnedRow row;
    auto value = [&](const EntryIndex& x, uint32_t) { return learned_importance(p, sp, lg, x, row); };
    __shared__ double red[kBlock / 64 + 2];
    __shared__ double buf[kScanLds];
    double local = 0.0;
    EntryCursor cur;
    cur.start(threadIdx.x, P, T);
    for (uint32_t e = threadIdx.x; e < n; e += kBlock) {
        local += (double)value(cur.x, e);
        cur.advance((uint32_t)kBlock);
    }
    row = LearnedRow{};
    mixture_prefix(local, n, P, T, m.bv, p.eps, p.Bsum, p.R, value, p.prefix + (size_t)slot * n, p.Q + slot, red,
                   buf, p.buckets ? p.buckets + (size_t)slot * kPrefixBuckets : nullptr);
}

struct TraceParams {
    MeshDev m;
    const float2* pts;
    uint32_t sample_base;
    const uint32_t* sample_of;// slot -> sample (nullptr: sample_base + slot)
    uint32_t R;
    uint32_t n_lambda;
    const float* c1;          // [n_lambda] nTot (sigma_e + sigma_a) log2(e)
    const float* c2;          // [n_lambda] nTot sigma_a log2(e)
    float inv_lambda;
    uint32_t seed, launch, nrays;
    uint32_t chunks;          // ray chunks per active sample
    uint32_t rays_per_thread;
    const uint32_t* active;   // [active count] batch slots
    const double* prefix;
    const double* Q;
    const uint32_t* buckets;  // [slots * kPrefixBuckets]
    double2* partial;         // [slots * max_chunks]
    uint32_t* partial_drop;   // [slots * max_chunks]
    uint32_t max_chunks;
    unsigned long long* counters;
    uint32_t shared;          // 1: every slot follows the proposal of slot 0 (beta * V sampling)
};

// blockIdx -> logical block so that consecutive logical blocks (the chunks of
// one sample) run on one XCD and share its L2 (blocks are dealt to the 8 XCDs
// round-robin).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    const uint32_t xcds = 8u;
    uint32_t q = nb / xcds, rem = nb % xcds, x = b % xcds, k = b / xcds;
    return x < rem ? x * (q + 1u) + k : rem * (q + 1u) + (x - rem) * q + k;
}

__global__ void __launch_bounds__(kBlock) hase_trace_kernel(TraceParams p)
{
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t a = lb / p.chunks, chunk = lb % p.chunks;
    const uint32_t slot = p.active[a];
    const MeshDev& m = p.m;
    const uint32_t T = m.T, P = T * m.layers;
    const size_t n = (size_t)p.R * P;
    const uint32_t sample = p.sample_of ? p.sample_of[slot] : p.sample_base + slot;
    const SampleGeom sp = sample_point(p.pts, m.np, sample);
    const uint32_t own = p.shared ? 0u : slot;
    const double* pre = p.prefix + (size_t)own * n;
    const double Q = p.Q[own];
    const double U = systematic_offset(p.seed, sample, p.launch);
    const double scale = Q / (double)p.nrays;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t rays_per_block = kBlock * p.rays_per_thread;

    double s1 = 0.0, s2 = 0.0;
    unsigned long long steps = 0, rays = 0;
    uint32_t drops = 0;
    for (uint32_t i = 0; i < p.rays_per_thread; ++i) {
        uint32_t ray = chunk * rays_per_block + i * kBlock + tid;
        if (ray >= p.nrays) break;
        double target = ((double)ray + U) * scale;
        // first entry whose prefix exceeds the target
        const uint32_t* bk = p.buckets + (size_t)own * kPrefixBuckets;
        const double kt = target * (double)kPrefixBuckets;
        const uint32_t kb = kt >= (double)(kPrefixBuckets - 1u) ? kPrefixBuckets - 1u : (uint32_t)kt;
        size_t lo = bk[kb], hi = (kb + 1u < kPrefixBuckets ? (size_t)bk[kb + 1u] : n - 1) + 1;
        while (lo < hi) {
            size_t mid = (lo + hi) >> 1;
            if (pre[mid] > target) hi = mid; else lo = mid + 1;
        }
        size_t e = lo < n ? lo : n - 1;
        double q = pre[e] - (e ? pre[e - 1] : 0.0);
        uint32_t k = (uint32_t)(e / P), j = (uint32_t)(e % P);
        uint32_t t = j % T, l = j / T;
        uint32_t rnd[4];
        philox4x32_10(ray, sample, p.launch, kTagRay, p.seed, kKeyTag, rnd);
        float u1 = u32_to_unit(rnd[0]), u2 = u32_to_unit(rnd[1]), u3 = u32_to_unit(rnd[2]);
        if (u1 + u2 > 1.0f) { u1 = 1.0f - u1; u2 = 1.0f - u2; }
        const GlobalTriangles fetch;
        const float2 p0 = fetch.vertex(m, t, 0u), p1 = fetch.vertex(m, t, 1u), p2 = fetch.vertex(m, t, 2u);
        float ax = p0.x + u1 * (p1.x - p0.x) + u2 * (p2.x - p0.x);
        float ay = p0.y + u1 * (p1.y - p0.y) + u2 * (p2.y - p0.y);
        float az = (float)l + u3;   // layer units
        uint32_t nr = (k + 1u) >> 1, fp = k & 1u;
        float ez = (float)image_plane(sp.plane, (int)m.layers, nr, fp);
        WalkOut w = walk_ray(m, ax, ay, az, t, l, p0, p1, p2, sp.px, sp.py, ez, nr, fp, fetch);
        steps += w.steps;
        rays++;
        if (w.status != 0 || !(w.length > 0.0f)) {
            drops++;
            continue;
        }
        float g = spectral_gain(p.c1, p.c2, p.n_lambda, p.inv_lambda, m.clad_alpha, w);
        float f = g * w.refl * m.bv[j] / (4.0f * kPi * w.length * w.length);
        double X = (double)f * (Q / q);
        if (!isfinite(X)) {
            drops++;
            continue;
        }
        s1 += X;
        s2 += X * X;
    }
    __shared__ double r1[kBlock / 64], r2[kBlock / 64];
    __shared__ uint32_t rd[kBlock / 64];
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    unsigned long long ws = wave_sum_u64(steps), wr = wave_sum_u64(rays);
    unsigned long long wd = wave_sum_u64((unsigned long long)drops);
    if (lane == 0) {
        r1[wid] = s1;
        r2[wid] = s2;
        rd[wid] = (uint32_t)wd;
        if (wr) {
            atomicAdd(&p.counters[0], wr);
            atomicAdd(&p.counters[1], ws);
        }
    }
    __syncthreads();
    if (tid == 0) {
        double a1 = 0.0, a2 = 0.0;
        uint32_t ad = 0;
        for (int i = 0; i < kBlock / 64; ++i) { a1 += r1[i]; a2 += r2[i]; ad += rd[i]; }
        size_t at = (size_t)slot * p.max_chunks + chunk;
        p.partial[at] = make_double2(a1, a2);
        p.partial_drop[at] = ad;
    }
}

struct FinalizeParams {
    uint32_t slots, chunks, max_chunks, nrays;
    const double2* partial;
    const uint32_t* partial_drop;
    double* s1;
    double* s2;
    uint32_t* N;
    uint32_t* dropped;
    int32_t* active_flag;
    uint32_t* active_out;
    uint32_t* active_count;
    double mse_threshold;
    double phi_scale;           // nTot / tau
    float* phi;
    double* mse;
    double* se;
};

__global__ void hase_finalize_kernel(FinalizeParams p)
{
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= p.slots || !p.active_flag[slot]) return;
    double a1 = p.s1[slot], a2 = p.s2[slot];
    uint32_t drop = p.dropped[slot];
    for (uint32_t c = 0; c < p.chunks; ++c) {
        double2 v = p.partial[(size_t)slot * p.max_chunks + c];
        a1 += v.x;
        a2 += v.y;
        drop += p.partial_drop[(size_t)slot * p.max_chunks + c];
    }
    uint32_t N = p.N[slot] + p.nrays;
    p.s1[slot] = a1;
    p.s2[slot] = a2;
    p.N[slot] = N;
    p.dropped[slot] = drop;
    // relative standard error, src/core/calcForwardPhiAse.cpp:81-98; mse is its square
    double nn = (double)N, mse;
    if (N < 2u || !isfinite(a1) || !isfinite(a2)) mse = 1.79769313486231570e308;
    else if (a1 == 0.0) mse = __longlong_as_double(0x7ff8000000000000LL);
    else {
        double rv = (nn * a2 / (a1 * a1) - 1.0) / nn;
        mse = rv > 0.0 ? rv : 0.0;
    }
    double phi = p.phi_scale * a1 / nn;
    p.phi[slot] = (float)phi;
    p.mse[slot] = mse;
    p.se[slot] = (mse >= 0.0 && mse < 1e300) ? sqrt(mse) * fabs(phi) : mse;
    bool converged = drop == 0u && isfinite(mse) && mse <= p.mse_threshold;
    if (converged) {
        p.active_flag[slot] = 0;
    } else {
        uint32_t at = atomicAdd(p.active_count, 1u);
        p.active_out[at] = slot;
    }
}

}  // namespace

namespace hase {

// Single-ray check in the caller's units: start / end (x, y, z in cm) and the
// caller's prism index.  Heights go to layer units in FP64 (an end within
// 1e-5 layers of a plane is put on it), the prism to device triangle order,
// then the device walk runs and its sums are written per ray.
__global__ void hase_single_rays_kernel(MeshDev m, double h, const uint32_t* user2dev, uint32_t n,
                                        const float* start, const uint32_t* prism, const float* end,
                                        const uint32_t* nrefl, const uint32_t* first, float* bl, float* gl, float* cl,
                                        float* refl, uint32_t* steps, int32_t* status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = prism[i];
    const uint32_t t = user2dev[j % m.T], l = j / m.T;
    const float za = (float)((double)start[3 * i + 2] / h);
    const double ze = (double)end[3 * i + 2] / h, zr = rint(ze);
    const float zef = (float)(fabs(ze - zr) < 1e-5 ? zr : ze);
    const WalkOut w = walk_ray(m, start[3 * i], start[3 * i + 1], za, t, l, end[3 * i], end[3 * i + 1], zef,
                               nrefl[i], first[i]);
    bl[i] = w.beta_len;
    gl[i] = w.gain_len;
    cl[i] = w.clad_len;
    refl[i] = w.refl;
    steps[i] = w.steps;
    status[i] = w.status;
}

}  // namespace hase

namespace {

using namespace hase;

// Flux of the resident slots into the flux of every sample.
__global__ void hase_phi_scatter_kernel(const float* __restrict__ phi, const uint32_t* __restrict__ sample_of,
                                        uint32_t base, uint32_t slots, float* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < slots) out[sample_of ? sample_of[i] : base + i] = phi[i];
}

// beta [P] (FP64, device prism order) -> FP32 beta, beta * V, column-major
// beta, and per-block sums of beta * V (fixed order: deterministic).
__global__ void __launch_bounds__(256) hase_refresh_beta_kernel(const double* __restrict__ beta,
                                                                const double* __restrict__ area,
                                                                const uint32_t* __restrict__ clad, double h,
                                                                uint32_t T, uint32_t layers, float* __restrict__ b32,
                                                                float* __restrict__ bv, float* __restrict__ bcol,
                                                                double* __restrict__ block_sums)
{
    __shared__ double red[256];
    const uint32_t P = T * layers;
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    double w = 0.0;
    if (j < P) {
        const double v = beta[j];
        const uint32_t t = j % T, l = j / T;
        w = v > 0.0 ? v * area[t] * h : 0.0;
        const float walk = clad[t] ? 0.0f : (float)v;   // the walks' copy: no beta in cladding prisms
        b32[j] = walk;
        bv[j] = (float)w;
        bcol[(size_t)t * layers + l] = walk;
    }
    red[threadIdx.x] = w;
    __syncthreads();
    for (uint32_t k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------- host

}  // namespace

namespace {

void validate(const hase_mesh_t* mesh, const hase_experiment_t* e, const hase_compute_t* c)
{
    if (!mesh || !e || !c) throw HaseError(HASE_E_INVALID, "mesh, experiment and compute must not be NULL");
    if (mesh->number_of_points < 3 || mesh->number_of_triangles < 1)
        throw HaseError(HASE_E_INVALID, "mesh needs at least 3 points and 1 triangle");
    if (mesh->number_of_levels < 2) throw HaseError(HASE_E_INVALID, "number_of_levels must be at least 2");
    if (!(mesh->thickness > 0.0) || !std::isfinite(mesh->thickness))
        throw HaseError(HASE_E_INVALID, "thickness must be positive");
    if (!mesh->points || !mesh->triangle_point_indices || !mesh->triangle_neighbors || !mesh->forbidden_edge ||
        !mesh->beta_volume)
        throw HaseError(HASE_E_INVALID, "mesh arrays points, triangle_point_indices, triangle_neighbors, "
                                        "forbidden_edge and beta_volume are required");
    if (!(mesh->n_tot > 0.0) || !(mesh->crystal_t_fluo > 0.0))
        throw HaseError(HASE_E_INVALID, "n_tot and crystal_t_fluo must be positive");
    if (e->min_rays_per_sample == 0) throw HaseError(HASE_E_INVALID, "min_rays_per_sample must be positive");
    if (e->max_rays_per_sample < e->min_rays_per_sample)
        throw HaseError(HASE_E_INVALID, "max_rays_per_sample must be >= min_rays_per_sample");
    if (e->number_of_wavelengths == 0 || !e->sigma_a || !e->sigma_e)
        throw HaseError(HASE_E_INVALID, "at least one wavelength (sigma_a, sigma_e) is required");
    if (!(e->mse_threshold >= 0.0)) throw HaseError(HASE_E_INVALID, "mse_threshold must be >= 0");
    if (!(e->defensive_fraction >= 0.0 && e->defensive_fraction < 1.0))
        throw HaseError(HASE_E_INVALID, "defensive_fraction must lie in [0, 1)");
    uint64_t S = (uint64_t)mesh->number_of_points * mesh->number_of_levels;
    if (S > 0xFFFFFFFFull) throw HaseError(HASE_E_INVALID, "too many sample points");
    if (c->max_sample_range < c->min_sample_range || c->max_sample_range >= S)
        throw HaseError(HASE_E_INVALID, "sample range must satisfy min <= max < number_of_points * number_of_levels");
    if (c->max_repetitions == 0) throw HaseError(HASE_E_INVALID, "max_repetitions must be >= 1");
    uint64_t P = (uint64_t)mesh->number_of_triangles * (mesh->number_of_levels - 1);
    // 24-bit triangle and level indices, 31-bit point indices
    if (P * 3u > 0xFFFFFFFFull || mesh->number_of_triangles >= (1u << 24) || mesh->number_of_levels >= (1u << 24) ||
        mesh->number_of_points >= 0x80000000u)
        throw HaseError(HASE_E_INVALID, "mesh too large");
}

void set_params(hase_context* ctx, const hase_experiment_t* e, const hase_compute_t* c)
{
    ctx->exp = *e;
    ctx->sigma_a.assign(e->sigma_a, e->sigma_a + e->number_of_wavelengths);
    ctx->sigma_e.assign(e->sigma_e, e->sigma_e + e->number_of_wavelengths);
    ctx->exp.sigma_a = ctx->sigma_a.data();
    ctx->exp.sigma_e = ctx->sigma_e.data();
    ctx->comp = *c;
    std::vector<float> c1(e->number_of_wavelengths), c2(e->number_of_wavelengths);
    for (uint32_t i = 0; i < e->number_of_wavelengths; ++i) {
        // base-2 coefficients of the spectral gain (its exponent in units of log2)
        const double log2e = 1.4426950408889634074;
        c1[i] = (float)(ctx->n_tot * (e->sigma_e[i] + e->sigma_a[i]) * log2e);
        c2[i] = (float)(ctx->n_tot * e->sigma_a[i] * log2e);
    }
    ctx->c1.upload(c1.data(), c1.size(), ctx->stream);
    ctx->c2.upload(c2.data(), c2.size(), ctx->stream);
    ctx->log_sigma = (e->max_sigma_a > 0.0 && e->max_sigma_e > 0.0)
                         ? (float)std::log(e->max_sigma_e / e->max_sigma_a) : 0.0f;
}

void upload_beta(hase_context* ctx, const double* beta_volume, const double* beta_cells)
{
    const uint32_t T = ctx->T, P = ctx->P;
    ctx->beta_host.assign(beta_volume, beta_volume + P);
    std::vector<float> b(P), bv(P);
    double B = 0.0;
    for (uint32_t j = 0; j < P; ++j) {
        // device prism j = t_dev + l * T reads the caller's prism dev2user[t_dev] + l * T
        double v = beta_volume[ctx->dev2user[j % T] + (j / T) * (size_t)T];
        if (!std::isfinite(v)) throw HaseError(HASE_E_INVALID, "beta_volume must be finite");
        // the walks' copy: cladding prisms absorb with -alpha and take no beta
        b[j] = ctx->clad_flag[j % T] ? 0.0f : (float)v;
        double w = v > 0.0 ? v * ctx->tri_area[j % T] * ctx->h : 0.0;
        bv[j] = (float)w;
        B += w;
    }
    ctx->Bsum = B;
    ctx->beta.upload(b.data(), P, ctx->stream);
    ctx->bv.upload(bv.data(), P, ctx->stream);
    std::vector<float> bc(P);
    const uint32_t layers = ctx->L - 1;
    for (uint32_t j = 0; j < P; ++j) bc[(size_t)(j % T) * layers + j / T] = b[j];
    ctx->beta_col.upload(bc.data(), P, ctx->stream);
    if (beta_cells) {
        ctx->beta_cells.assign(beta_cells, beta_cells + ctx->S);
        ctx->bcells.upload(ctx->beta_cells.data(), ctx->S, ctx->stream);
    } else if (ctx->beta_cells.empty()) {
        ctx->beta_cells.assign(ctx->S, 0.0);
        ctx->bcells.upload(ctx->beta_cells.data(), ctx->S, ctx->stream);
    }
    if (ctx->proposal_kind == HASE_PROPOSAL_LEARNED) build_clusters(ctx);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

// Index of (x, y) on a 2^16 x 2^16 Hilbert curve.
uint64_t hilbert_index(uint32_t x, uint32_t y)
{
    const uint32_t n = 1u << 16;
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

// Device triangle order: triangles sorted by the Hilbert index of their
// centroid (quantised to 16 bits per axis over the centroids' bounding box),
// ties by caller index.  Lanes that take consecutive device triangles then
// walk from neighbouring prisms, which keeps their records in the same cache
// lines and their step counts alike.  The proposal prefix runs in this order.
std::vector<uint32_t> triangle_order(const hase_mesh_t* mesh)
{
    const uint32_t T = mesh->number_of_triangles, np = mesh->number_of_points;
    const double* px = mesh->points;
    const double* py = mesh->points + np;
    std::vector<double> cx(T), cy(T);
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (uint32_t t = 0; t < T; ++t) {
        uint32_t a = mesh->triangle_point_indices[t], b = mesh->triangle_point_indices[t + (size_t)T],
                 c = mesh->triangle_point_indices[t + 2 * (size_t)T];
        cx[t] = (px[a] + px[b] + px[c]) / 3.0;
        cy[t] = (py[a] + py[b] + py[c]) / 3.0;
        x0 = std::min(x0, cx[t]); x1 = std::max(x1, cx[t]);
        y0 = std::min(y0, cy[t]); y1 = std::max(y1, cy[t]);
    }
    auto quant = [](double v, double lo, double hi) -> uint32_t {
        if (!(hi > lo)) return 0u;
        double q = std::floor((v - lo) / (hi - lo) * 65535.0);
        return (uint32_t)std::min(65535.0, std::max(0.0, q));
    };
    std::vector<uint64_t> key(T);
    for (uint32_t t = 0; t < T; ++t) key[t] = hilbert_index(quant(cx[t], x0, x1), quant(cy[t], y0, y1));
    std::vector<uint32_t> order(T);
    for (uint32_t t = 0; t < T; ++t) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    return order;
}

std::unique_ptr<hase_context> create_context(const hase_mesh_t* mesh, const hase_experiment_t* e,
                                             const hase_compute_t* c)
{
    validate(mesh, e, c);
    int count = 0;
    HIP_CHECK(hipGetDeviceCount(&count));
    if (count <= 0) throw HaseError(HASE_E_DEVICE, "no HIP device is visible");
    if (c->device < 0 || c->device >= count) throw HaseError(HASE_E_INVALID, "device ordinal out of range");
    HIP_CHECK(hipSetDevice(c->device));
    auto ctx = std::make_unique<hase_context>();
    ctx->device = c->device;
    if (g_arena) {
        // allocation-free evaluation: the null stream, rounds launched one by
        // one (instantiating a graph allocates), no timing events
        ctx->stream = nullptr;
        ctx->tuning.use_graph = 0u;
    } else {
        HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    }
    const uint32_t np = mesh->number_of_points, T = mesh->number_of_triangles, L = mesh->number_of_levels;
    ctx->np = np; ctx->T = T; ctx->L = L; ctx->P = T * (L - 1); ctx->S = np * L;
    ctx->h = mesh->thickness;
    ctx->n_tot = mesh->n_tot;
    ctx->tau = mesh->crystal_t_fluo;
    ctx->clad_alpha = mesh->cladding_absorption;
    const double* px = mesh->points;
    const double* py = mesh->points + np;
    for (size_t i = 0; i < 3u * (size_t)T; ++i)
        if (mesh->triangle_point_indices[i] >= np)
            throw HaseError(HASE_E_INVALID, "triangle point index out of range");
    ctx->dev2user = triangle_order(mesh);
    ctx->user2dev.assign(T, 0u);
    for (uint32_t t = 0; t < T; ++t) ctx->user2dev[ctx->dev2user[t]] = t;
    // triangle records in device order: {x0,y0,x1,y1} {x2,y2,n0,n1} {n2,clad,0,0};
    // n_k = device neighbour << 2 | shared edge index inside the neighbour
    std::vector<uint4> rec(3u * (size_t)T);
    ctx->tri_area.resize(T);
    ctx->clad_flag.resize(T);
    ctx->centroid_host.resize(T);
    ctx->tri_vertex.resize(3u * (size_t)T);
    for (uint32_t td = 0; td < T; ++td) {
        const uint32_t t = ctx->dev2user[td];
        uint32_t v[3];
        uint32_t nw[3];
        const uint32_t clad = mesh->cladding_cell_types && mesh->cladding_cell_types[t] == mesh->cladding_number ? 1u : 0u;
        for (int k = 0; k < 3; ++k) {
            v[k] = mesh->triangle_point_indices[t + (size_t)k * T];
            ctx->tri_vertex[3u * td + k] = v[k];
            int32_t nb = mesh->triangle_neighbors[t + (size_t)k * T];
            int32_t fe = mesh->forbidden_edge[t + (size_t)k * T];
            if (nb >= 0) {
                if ((uint32_t)nb >= T || fe < 0 || fe > 2)
                    throw HaseError(HASE_E_INVALID, "triangle_neighbors / forbidden_edge out of range");
                nw[k] = (ctx->user2dev[(uint32_t)nb] << 2) | (uint32_t)fe;
            } else {
                nw[k] = 0x7FFFFFFFu;
            }
            nw[k] |= clad << 31;   // the owning triangle's cladding flag rides in every neighbour word
        }
        double cr = (px[v[1]] - px[v[0]]) * (py[v[2]] - py[v[0]]) - (py[v[1]] - py[v[0]]) * (px[v[2]] - px[v[0]]);
        if (!(cr > 0.0))
            throw HaseError(HASE_E_INVALID, "triangles must be non-degenerate and counter-clockwise");
        ctx->tri_area[td] = 0.5 * cr;
        ctx->clad_flag[td] = clad;
        auto fb = [](double x) { float f = (float)x; uint32_t u; std::memcpy(&u, &f, 4); return u; };
        rec[3u * td] = make_uint4(fb(px[v[0]]), fb(py[v[0]]), fb(px[v[1]]), fb(py[v[1]]));
        rec[3u * td + 1u] = make_uint4(fb(px[v[2]]), fb(py[v[2]]), nw[0], nw[1]);
        rec[3u * td + 2u] = make_uint4(nw[2], v[0] | (clad << 31), v[1], v[2]);
        const float x0 = (float)px[v[0]], x1 = (float)px[v[1]], x2 = (float)px[v[2]];
        const float y0 = (float)py[v[0]], y1 = (float)py[v[1]], y2 = (float)py[v[2]];
        ctx->centroid_host[td] = make_float2((x0 + x1 + x2) * (1.0f / 3.0f), (y0 + y1 + y2) * (1.0f / 3.0f));
    }
    {
        double area = 0.0;
        for (uint32_t td = 0; td < T; ++td) area += ctx->tri_area[td];
        ctx->steps_per_cm = 1.2 * std::sqrt((double)T / area);
    }
    ctx->rec.upload(rec.data(), rec.size(), ctx->stream);
    ctx->user2dev_dev.upload(ctx->user2dev.data(), T, ctx->stream);
    const bool indexed_fits = T < (1u << 19) && np <= 0x10000u && (size_t)T + (np + 1u) / 2u <= hase::kLdsStageWords;
    // compact records (corners + 21-bit neighbour fields + cladding bit) of
    // every mesh below 2^19 triangles: staged in LDS up to kLdsTriangles,
    // else read from global memory by the walks that do not stage
    std::vector<uint4> crec;
    if (T < (1u << 19)) {
        crec.resize(2u * (size_t)T);
        for (uint32_t td = 0; td < T; ++td) {
            const uint4 a = rec[3u * td], b = rec[3u * td + 1u], c = rec[3u * td + 2u];
            const uint32_t nwk[3] = {b.z, b.w, c.x};
            uint64_t W = 0;
            for (int k = 0; k < 3; ++k) {
                const uint64_t f = nw_boundary(nwk[k]) ? 0x1FFFFFull
                                                       : (uint64_t)(nw_next(nwk[k]) << 2 | nw_edge(nwk[k]));
                W |= f << (21 * k);
            }
            W |= (uint64_t)(c.y >> 31) << 63;
            crec[2u * td] = a;
            crec[2u * td + 1u] = make_uint4(b.x, b.y, (uint32_t)W, (uint32_t)(W >> 32));
        }
    }
    if (T <= hase::kLdsTriangles || (!indexed_fits && !crec.empty())) {
        ctx->lds_rec.upload(crec.data(), crec.size(), ctx->stream);
        ctx->lds_kind = T <= hase::kLdsTriangles ? hase::kLdsCompact : hase::kCompactGlobal;
    } else if (indexed_fits) {
        // indexed records (point indices + the compact records' neighbour
        // word) followed by the vertex table; the compact records stay in
        // global memory beside them
        std::vector<uint4> lrec((size_t)T + (np + 1u) / 2u, make_uint4(0u, 0u, 0u, 0u));
        for (uint32_t td = 0; td < T; ++td) {
            const uint4 c = rec[3u * td + 2u], q = crec[2u * td + 1u];
            const uint32_t clad = c.y >> 31;
            lrec[td] = make_uint4((c.y & 0x7FFFFFFFu) | c.z << 16, c.w | clad << 31, q.z, q.w);
        }
        float2* vtx = reinterpret_cast<float2*>(lrec.data() + T);
        for (uint32_t i = 0; i < np; ++i) vtx[i] = make_float2((float)px[i], (float)py[i]);
        ctx->lds_rec.upload(lrec.data(), lrec.size(), ctx->stream);
        ctx->lds_kind = hase::kLdsIndexed;
        ctx->cmp_rec.upload(crec.data(), crec.size(), ctx->stream);
    }
    if (ctx->lds_staged()) {
        // staged walks run one 1024-thread work-group per CU
        ctx->tuning.stage_lds = 1u;
        ctx->tuning.block_size = 1024u;
        ctx->tuning.rays_per_item = 512u;
    } else if (ctx->lds_kind == hase::kCompactGlobal) {
        ctx->tuning.stage_lds = 1u;
    }
    ctx->centroid.upload(ctx->centroid_host.data(), T, ctx->stream);
    std::vector<float2>& pts = ctx->pts_host;
    pts.resize(np);
    for (uint32_t i = 0; i < np; ++i) pts[i] = make_float2((float)px[i], (float)py[i]);
    ctx->pts.upload(pts.data(), np, ctx->stream);
    if (mesh->reflectivities) {
        ctx->has_refl = true;
        std::vector<float> refl(2u * (size_t)T);
        for (uint32_t td = 0; td < T; ++td) {
            refl[2u * td] = mesh->reflectivities[2u * (size_t)ctx->dev2user[td]];
            refl[2u * td + 1u] = mesh->reflectivities[2u * (size_t)ctx->dev2user[td] + 1u];
        }
        ctx->refl.upload(refl.data(), refl.size(), ctx->stream);
    }
    if (mesh->refractive_indices) {
        ctx->has_nidx = true;
        for (int i = 0; i < 4; ++i) ctx->nidx[i] = mesh->refractive_indices[i];
    }
    ctx->counters.alloc(4);
    set_params(ctx.get(), e, c);
    upload_beta(ctx.get(), mesh->beta_volume, mesh->beta_cells);
    return ctx;
}

}  // namespace

namespace hase {

MeshDev mesh_view(const hase_context* ctx)
{
    MeshDev m;
    m.rec = ctx->rec.p;
    m.lds_rec = ctx->lds_rec.p;
    m.lds_kind = ctx->lds_rec.p ? ctx->lds_kind : 0u;
    m.cmp_rec = ctx->lds_kind == kLdsIndexed ? ctx->cmp_rec.p : (ctx->lds_rec.p ? ctx->lds_rec.p : nullptr);
    m.lds_words = (uint32_t)ctx->lds_rec.n;
    m.beta = ctx->beta.p;
    m.beta_col = ctx->beta_col.p;
    m.bv = ctx->bv.p;
    m.refl = ctx->has_refl ? ctx->refl.p : nullptr;
    m.nin_b = ctx->nidx[0]; m.nout_b = ctx->nidx[1]; m.nin_t = ctx->nidx[2]; m.nout_t = ctx->nidx[3];
    m.has_nidx = ctx->has_nidx ? 1 : 0;
    m.T = ctx->T;
    m.layers = ctx->L - 1;
    m.np = ctx->np;
    m.h = (float)ctx->h;
    m.clad_alpha = (float)ctx->clad_alpha;
    m.steps_per_cm = (float)ctx->steps_per_cm;
    return m;
}

uint32_t slices(const hase_context* ctx)
{
    uint32_t K = ctx->exp.use_reflections ? (ctx->exp.max_reflections ? ctx->exp.max_reflections : 1u) : 0u;
    return 1u + 2u * K;
}

}  // namespace hase

namespace {

struct BatchTimers {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> imp, trace, fin, inf, prop;
    double queue_ms = 0.0;
};

// Device buffers for batches of up to `want` resident samples; returns the
// batch size.  Unless the caller fixed sample_batch, the importance prefix and
// scratch of a batch stay within `budget` bytes; budget 0 takes the context's
// own, fixed at its first batch: 6 GiB, or for a mesh whose entries need more
// than that for 2048 samples (~22 MB per sample at 600k prisms, where a few
// hundred samples per batch leave most CUs idle in the per-sample proposal
// kernels) up to those 2048 samples' bytes, within a quarter of the device
// memory then free and 48 GiB.  Every context keeps its buffers, so a
// greedier default would starve the contexts created after it.
uint32_t reserve_batch(hase_context* ctx, uint32_t want, size_t budget = 0, bool proposal_only = false)
{
    const hase_compute_t& c = ctx->comp;
    const hase_experiment_t& e = ctx->exp;
    const uint32_t R = slices(ctx);
    const size_t n_ent = (size_t)R * ctx->P;
    if (n_ent >= ((size_t)1 << 32))
        throw HaseError(HASE_E_INVALID, "reflection slices times prisms must stay below 2^32");
    const bool learned = ctx->proposal_kind == HASE_PROPOSAL_LEARNED;
    const bool shared = ctx->proposal_kind == HASE_PROPOSAL_BETA_VOLUME;
    const uint32_t rpt = c.rays_per_thread ? c.rays_per_thread : 8u;
    const uint32_t rays_per_block = kBlock * rpt;
    const uint32_t max_chunks = (e.max_rays_per_sample + rays_per_block - 1) / rays_per_block;
    const bool queue = ctx->tuning.mode == HASE_TRACE_QUEUE;
    // per-slot partial sums: queue items in the work queue, chunks otherwise
    const size_t chunks = queue ? queue_items_per_slot(ctx, ctx->tuning) : max_chunks;
    uint32_t batch = c.sample_batch;
    if (g_arena) {
        // allocation-free call: as many slots as the scratch left after the
        // mesh holds (64 KiB kept for the per-call tables and alignment)
        const size_t tables_b = n_ent * 12u + kPrefixBuckets * 4u;
        const size_t per_slot = (shared ? 0u : tables_b) + chunks * 20u + 256u +
                                (learned ? (size_t)R * ctx->C * 4u : 0u);
        size_t left = g_arena->remaining();
        const size_t keep = ((size_t)64 << 10) + (shared ? tables_b : 0u);
        left = left > keep ? left - keep : 0u;
        const size_t fit = std::min<size_t>(left / per_slot, 8192);
        if (fit == 0)
            throw HaseError(HASE_E_MEMORY, "the no-allocation scratch cannot hold one sample point's buffers: " +
                                               std::to_string(per_slot + keep) + " bytes needed, " +
                                               std::to_string(g_arena->remaining()) + " left after the mesh");
        batch = batch ? std::min<uint32_t>(batch, (uint32_t)fit) : (uint32_t)fit;
    }
    if (batch == 0) {
        const uint32_t per_item = ctx->tuning.mode == HASE_TRACE_QUEUE
                                      ? (uint32_t)queue_items_per_slot(ctx, ctx->tuning)
                                      : max_chunks;
        size_t per_sample = (shared ? 0u : n_ent * 12u + kPrefixBuckets * 4u) + (size_t)per_item * 12u + 96u +
                            (learned ? (size_t)R * ctx->C * 4u : 0u);
        if (budget == 0) {
            if (ctx->batch_budget == 0) {
                const size_t floor_b = (size_t)6 << 30, want_b = per_sample * 2048u;
                ctx->batch_budget = floor_b;
                if (want_b > floor_b) {
                    size_t free_b = 0, total_b = 0;
                    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
                    ctx->batch_budget = std::max(floor_b, std::min({want_b, free_b / 4u, (size_t)48 << 30}));
                }
            }
            budget = ctx->batch_budget;
        }
        batch = (uint32_t)std::max<size_t>(1, std::min<size_t>(budget / per_sample, 8192));
    }
    batch = std::min(batch, want);
    // the shared beta * V proposal keeps one slot of prefix tables
    const size_t tables = shared ? 1u : batch;
    if (!shared) ctx->imp.alloc(tables * n_ent);
    ctx->prefix.alloc(tables * n_ent);
    ctx->buckets.alloc(tables * kPrefixBuckets);
    if (learned) ctx->logits.alloc((size_t)batch * R * ctx->C);
    ctx->Q.alloc(batch);
    ctx->slot_sample.alloc(batch);
    ctx->counters.alloc(4);
    if (proposal_only) {
        queue_targets(ctx);
        return batch;
    }
    ctx->s1.alloc(batch);
    ctx->s2.alloc(batch);
    ctx->N.alloc(batch);
    ctx->dropped.alloc(batch);
    ctx->active_flag.alloc(batch);
    ctx->active_list.alloc(batch);
    ctx->active_next.alloc(batch);
    ctx->active_count.alloc(1);
    ctx->slot_sample.alloc(batch);
    ctx->partial.alloc((size_t)batch * chunks);
    ctx->partial_drop.alloc((size_t)batch * chunks);
    ctx->phi.alloc(batch);
    ctx->mse.alloc(batch);
    ctx->se.alloc(batch);
    ctx->counters.alloc(4);
    if (ctx->tuning.mode == HASE_TRACE_QUEUE) queue_reserve(ctx, batch);
    return batch;
}

// Work-groups per sample for the walked importance map: enough to give a
// small batch two work-groups per CU.
uint32_t importance_parts(const hase_context* ctx, uint32_t slots)
{
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const uint32_t want = 2u * (uint32_t)std::max(1, cus);
    if (slots >= want) return 1u;
    return std::min<uint32_t>(64u, (want + slots - 1u) / slots);
}

// Walked importance of `slots` samples (and, unless ip.transfer is set, their
// mixture prefixes): the LDS-staged persistent kernel when the mesh's compact
// records exist, else one work-group per sample (split into ip.parts for
// small batches).  ip.parts is set here.
void launch_importance(hase_context* ctx, ImportanceParams& ip, uint32_t slots)
{
    if (ctx->lds_rec.p && ctx->lds_staged()) {
        int cus = 0;
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        ip.parts = 1u;
        if (ctx->lds_kind == kLdsIndexed)
            hipLaunchKernelGGL(hase_importance_staged_kernel<kLdsIndexed>, dim3((uint32_t)std::max(1, cus)),
                               dim3(kStagedBlock), 0, ctx->stream, ip);
        else
            hipLaunchKernelGGL(hase_importance_staged_kernel<kLdsCompact>, dim3((uint32_t)std::max(1, cus)),
                               dim3(kStagedBlock), 0, ctx->stream, ip);
        HIP_CHECK(hipGetLastError());
        if (!ip.transfer) {
            hipLaunchKernelGGL(hase_prefix_kernel, dim3(slots), dim3(kBlock), 0, ctx->stream, ip);
            HIP_CHECK(hipGetLastError());
        }
        return;
    }
    // a small batch would leave most of the chip idle: split each sample's walks
    ip.parts = importance_parts(ctx, slots);
    if (ctx->lds_rec.p && ctx->lds_kind == kCompactGlobal)
        hipLaunchKernelGGL(hase_importance_kernel<kCompactGlobal>, dim3(slots * ip.parts), dim3(kBlock), 0,
                           ctx->stream, ip);
    else
        hipLaunchKernelGGL(hase_importance_kernel<0u>, dim3(slots * ip.parts), dim3(kBlock), 0, ctx->stream, ip);
    HIP_CHECK(hipGetLastError());
    if (ip.parts > 1u && !ip.transfer) {
        hipLaunchKernelGGL(hase_prefix_kernel, dim3(slots), dim3(kBlock), 0, ctx->stream, ip);
        HIP_CHECK(hipGetLastError());
    }
}

// The beta * V proposal every sample shares, into slot 0 of the prefix
// tables (once per evaluation: beta may have changed since the last).
void build_shared_prefix(hase_context* ctx, BatchTimers& tm)
{
    const hase_experiment_t& e = ctx->exp;
    ImportanceParams ip{};
    ip.m = mesh_view(ctx);
    ip.R = slices(ctx);
    ip.eps = e.defensive_fraction;
    ip.Bsum = ctx->Bsum;
    ip.prefix = ctx->prefix.p;
    ip.Q = ctx->Q.p;
    ip.buckets = ctx->buckets.p;
    hipEvent_t i0 = record(ctx);
    hipLaunchKernelGGL(hase_beta_volume_prefix_kernel, dim3(1), dim3(kBlock), 0, ctx->stream, ip);
    HIP_CHECK(hipGetLastError());
    tm.imp.emplace_back(i0, record(ctx));
    ctx->stats.importance_launches++;
}

// Proposal prefix and Q of slots [0, slots): the walked importance map or the
// learned proposal (which needs contiguous samples b0 + slot); nothing for
// the beta * V proposal, built once per evaluation by build_shared_prefix.
void build_proposal(hase_context* ctx, uint32_t b0, const uint32_t* sample_of, uint32_t slots, bool learned,
                    BatchTimers& tm)
{
    if (ctx->proposal_kind == HASE_PROPOSAL_BETA_VOLUME && !learned) return;
    const hase_experiment_t& e = ctx->exp;
    const uint32_t R = slices(ctx);
    const MeshDev m = mesh_view(ctx);
    if (learned) {
        hipEvent_t a0 = record(ctx);
        launch_proposal_inference(ctx, b0, slots, ctx->logits.p);
        hipEvent_t a1 = record(ctx);
        LearnedParams lp;
        lp.m = m;
        lp.pts = ctx->pts.p;
        lp.centroid = ctx->centroid.p;
        lp.logits = ctx->logits.p;
        lp.cl_tri = ctx->cl_tri.p;
        lp.cl_lev = ctx->cl_lev.p;
        lp.slots = slots;
        lp.sample_base = b0;
        lp.R = R;
        lp.C = ctx->C;
        lp.TG = ctx->TG;
        lp.ct = ctx->cluster_t;
        lp.cl = ctx->cluster_l;
        lp.eps = e.defensive_fraction;
        lp.Bsum = ctx->Bsum;
        lp.prefix = ctx->prefix.p;
        lp.Q = ctx->Q.p;
        lp.buckets = ctx->buckets.p;
        hipLaunchKernelGGL(hase_learned_importance_kernel, dim3(slots), dim3(kBlock), 0, ctx->stream, lp);
        HIP_CHECK(hipGetLastError());
        tm.inf.emplace_back(a0, a1);
        tm.prop.emplace_back(a1, record(ctx));
    } else {
        ImportanceParams ip;
        ip.m = m;
        ip.pts = ctx->pts.p;
        ip.sample_base = b0;
        ip.slots = slots;
        ip.R = R;
        ip.n_lambda = e.number_of_wavelengths;
        ip.c1 = ctx->c1.p;
        ip.c2 = ctx->c2.p;
        ip.inv_lambda = 1.0f / (float)e.number_of_wavelengths;
        ip.eps = e.defensive_fraction;
        ip.Bsum = ctx->Bsum;
        ip.imp = ctx->imp.p;
        ip.prefix = ctx->prefix.p;
        ip.Q = ctx->Q.p;
        ip.buckets = ctx->buckets.p;
        ip.counters = ctx->counters.p;
        ip.active = nullptr;
        ip.sample_of = sample_of;
        ip.transfer = 0;
        hipEvent_t i0 = record(ctx);
        launch_importance(ctx, ip, slots);
        tm.imp.emplace_back(i0, record(ctx));
        ctx->stats.importance_launches++;
    }

}


void run_batch(hase_context* ctx, hase_result_t* res, uint32_t b0, const uint32_t* samples, uint32_t slots,
               bool learned, uint32_t launch0, std::vector<uint32_t>* deferred, BatchTimers& tm)
{
    const hase_compute_t& c = ctx->comp;
    const hase_experiment_t& e = ctx->exp;
    const uint32_t R = slices(ctx);
    const uint32_t rpt = c.rays_per_thread ? c.rays_per_thread : 8u;
    const uint32_t rays_per_block = kBlock * rpt;
    const uint32_t max_chunks = (e.max_rays_per_sample + rays_per_block - 1) / rays_per_block;
    const MeshDev m = mesh_view(ctx);
    const Schedule sched{e.min_rays_per_sample, e.max_rays_per_sample, c.adaptive_steps, c.max_repetitions};
    (void)R;
    std::vector<uint32_t> iota(slots);
    for (uint32_t i = 0; i < slots; ++i) iota[i] = i;
    std::vector<int32_t> ones(slots, 1);
    const uint32_t* sample_of = nullptr;
    if (samples) {
        HIP_CHECK(hipMemcpyAsync(ctx->slot_sample.p, samples, slots * sizeof(uint32_t), hipMemcpyHostToDevice,
                                 ctx->stream));
        sample_of = ctx->slot_sample.p;
    }
    HIP_CHECK(hipMemsetAsync(ctx->s1.p, 0, slots * sizeof(double), ctx->stream));
    HIP_CHECK(hipMemsetAsync(ctx->s2.p, 0, slots * sizeof(double), ctx->stream));
    HIP_CHECK(hipMemsetAsync(ctx->N.p, 0, slots * sizeof(uint32_t), ctx->stream));
    HIP_CHECK(hipMemsetAsync(ctx->dropped.p, 0, slots * sizeof(uint32_t), ctx->stream));
    HIP_CHECK(hipMemcpyAsync(ctx->active_flag.p, ones.data(), slots * sizeof(int32_t), hipMemcpyHostToDevice,
                             ctx->stream));
    HIP_CHECK(hipMemcpyAsync(ctx->active_list.p, iota.data(), slots * sizeof(uint32_t), hipMemcpyHostToDevice,
                             ctx->stream));

    build_proposal(ctx, b0, sample_of, slots, learned, tm);
    std::vector<uint32_t> flags;
    if (ctx->tuning.mode == HASE_TRACE_QUEUE) {
        tm.queue_ms += queue_rounds(ctx, b0, samples != nullptr, slots, launch0, learned && deferred);
        if (learned && deferred) {
            flags.resize(slots);
            HIP_CHECK(hipMemcpy(flags.data(), ctx->q_flags.p, slots * sizeof(uint32_t), hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < slots; ++i)
                if (flags[i] & 2u) {
                    deferred->push_back(samples ? samples[i] : b0 + i);
                    ctx->pstats.fallback_samples++;
                }
        }
    }
    uint32_t active = ctx->tuning.mode == HASE_TRACE_QUEUE ? 0u : slots;
    uint32_t N = 0, k = 0, extra = 0, launch = launch0;
    while (active > 0) {
        const uint32_t target = sched.target(k);
        uint32_t nrays;
        if (target > N) {
            nrays = target - N;
        } else if (N >= sched.max_rays || target >= sched.max_rays) {
            if (extra + 1u >= sched.reps) break;
            extra++;
            nrays = sched.max_rays;
        } else if (k >= sched.steps) {
            break;   // flat schedule (no increases): nothing more to add
        } else {
            k++;
            continue;
        }
        const uint32_t chunks = (nrays + rays_per_block - 1) / rays_per_block;
        TraceParams tp;
        tp.m = m;
        tp.pts = ctx->pts.p;
        tp.sample_base = b0;
        tp.sample_of = sample_of;
        tp.R = R;
        tp.n_lambda = e.number_of_wavelengths;
        tp.c1 = ctx->c1.p;
        tp.c2 = ctx->c2.p;
        tp.inv_lambda = 1.0f / (float)e.number_of_wavelengths;
        tp.seed = c.rng_seed;
        tp.launch = launch;
        tp.nrays = nrays;
        tp.chunks = chunks;
        tp.rays_per_thread = rpt;
        tp.active = ctx->active_list.p;
        tp.prefix = ctx->prefix.p;
        tp.Q = ctx->Q.p;
        tp.buckets = ctx->buckets.p;
        tp.partial = ctx->partial.p;
        tp.partial_drop = ctx->partial_drop.p;
        tp.max_chunks = max_chunks;
        tp.counters = ctx->counters.p;
        tp.shared = ctx->proposal_kind == HASE_PROPOSAL_BETA_VOLUME ? 1u : 0u;
        hipEvent_t t0 = record(ctx);
        hipLaunchKernelGGL(hase_trace_kernel, dim3(active * chunks), dim3(kBlock), 0, ctx->stream, tp);
        HIP_CHECK(hipGetLastError());
        tm.trace.emplace_back(t0, record(ctx));
        ctx->stats.trace_launches++;

        HIP_CHECK(hipMemsetAsync(ctx->active_count.p, 0, sizeof(uint32_t), ctx->stream));
        FinalizeParams fp;
        fp.slots = slots;
        fp.chunks = chunks;
        fp.max_chunks = max_chunks;
        fp.nrays = nrays;
        fp.partial = ctx->partial.p;
        fp.partial_drop = ctx->partial_drop.p;
        fp.s1 = ctx->s1.p;
        fp.s2 = ctx->s2.p;
        fp.N = ctx->N.p;
        fp.dropped = ctx->dropped.p;
        fp.active_flag = ctx->active_flag.p;
        fp.active_out = ctx->active_next.p;
        fp.active_count = ctx->active_count.p;
        fp.mse_threshold = e.mse_threshold;
        fp.phi_scale = ctx->n_tot / ctx->tau;
        fp.phi = ctx->phi.p;
        fp.mse = ctx->mse.p;
        fp.se = ctx->se.p;
        hipEvent_t f0 = record(ctx);
        hipLaunchKernelGGL(hase_finalize_kernel, dim3((slots + 255) / 256), dim3(256), 0, ctx->stream, fp);
        HIP_CHECK(hipGetLastError());
        tm.fin.emplace_back(f0, record(ctx));
        HIP_CHECK(hipMemcpyAsync(&active, ctx->active_count.p, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        std::swap(ctx->active_list.p, ctx->active_next.p);
        N += nrays;
        launch++;
        if (target < sched.max_rays) k++;
        if (learned && deferred && launch == launch0 + 1u && active > 0u) {
            // samples the learned proposal did not settle with the first budget: hand them back
            std::vector<uint32_t> miss(active);
            HIP_CHECK(hipMemcpy(miss.data(), ctx->active_list.p, active * sizeof(uint32_t), hipMemcpyDeviceToHost));
            for (uint32_t slot : miss) deferred->push_back(samples ? samples[slot] : b0 + slot);
            ctx->pstats.fallback_samples += active;
            break;
        }
    }
    if (ctx->phi_keep) {
        hipLaunchKernelGGL(hase_phi_scatter_kernel, dim3((slots + 255) / 256), dim3(256), 0, ctx->stream,
                           (const float*)ctx->phi.p, sample_of, b0, slots, ctx->phi_all.p);
        HIP_CHECK(hipGetLastError());
    }
    std::vector<float> h_phi(slots);
    std::vector<double> h_mse(slots), h_se(slots);
    std::vector<uint32_t> h_N(slots), h_drop(slots);
    HIP_CHECK(hipMemcpyAsync(h_phi.data(), ctx->phi.p, slots * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(h_mse.data(), ctx->mse.p, slots * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(h_se.data(), ctx->se.p, slots * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(h_N.data(), ctx->N.p, slots * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(h_drop.data(), ctx->dropped.p, slots * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (uint32_t i = 0; i < slots; ++i) {
        const uint32_t s = samples ? samples[i] : b0 + i;
        const double phi = (double)h_phi[i];
        res->phi_ase[s] = h_phi[i];
        res->mse[s] = h_mse[i];
        res->total_rays[s] = h_N[i];
        res->dndt_ase[s] = (ctx->beta_cells[s] * (e.max_sigma_e + e.max_sigma_a) - e.max_sigma_a) * phi;
        if (res->standard_error) res->standard_error[s] = h_se[i];
        if (res->dropped_rays) res->dropped_rays[s] = h_drop[i];
    }
}

void evaluate(hase_context* ctx, hase_result_t* res)
{
    if (!res || !res->phi_ase || !res->mse || !res->total_rays || !res->dndt_ase)
        throw HaseError(HASE_E_INVALID, "result arrays phi_ase, mse, total_rays and dndt_ase are required");
    HIP_CHECK(hipSetDevice(ctx->device));
    const hase_compute_t& c = ctx->comp;
    const uint32_t S = ctx->S;
    for (uint32_t s = 0; s < S; ++s) {
        res->phi_ase[s] = 0.0f;
        res->mse[s] = 0.0;
        res->total_rays[s] = 0;
        res->dndt_ase[s] = 0.0;
        if (res->standard_error) res->standard_error[s] = 0.0;
        if (res->dropped_rays) res->dropped_rays[s] = 0;
    }
    for (auto ev : ctx->events) (void)hipEventDestroy(ev);
    ctx->events.clear();
    ctx->stats = hase_stats_t{};
    ctx->pstats = hase_proposal_stats_t{};
    ctx->pstats.clusters = ctx->C;
    if (ctx->phi_keep) {
        ctx->phi_all.alloc(S);
        HIP_CHECK(hipMemsetAsync(ctx->phi_all.p, 0, (size_t)S * sizeof(float), ctx->stream));
    }
    if (!(ctx->Bsum > 0.0)) return;   // no emitting volume: the estimate is zero
    const bool learned = ctx->proposal_kind == HASE_PROPOSAL_LEARNED;

    const uint32_t first = c.min_sample_range, last = c.max_sample_range;
    const uint32_t batch = reserve_batch(ctx, last - first + 1u);
    HIP_CHECK(hipMemsetAsync(ctx->counters.p, 0, 4 * sizeof(unsigned long long), ctx->stream));
    ctx->q_generation++;

    BatchTimers tm;
    hipEvent_t ev_begin = record(ctx);
    if (ctx->proposal_kind == HASE_PROPOSAL_BETA_VOLUME) build_shared_prefix(ctx, tm);
    std::vector<uint32_t> deferred;
    std::vector<uint32_t>* defer = (learned && ctx->proposal_fallback) ? &deferred : nullptr;
    for (uint32_t b0 = first; b0 <= last; b0 += batch) {
        const uint32_t slots = std::min(batch, last - b0 + 1u);
        run_batch(ctx, res, b0, nullptr, slots, learned, 0u, defer, tm);
        if (last - b0 < batch) break;
    }
    // learned-proposal misses restart on the walked map, batched across the whole range
    for (size_t d0 = 0; d0 < deferred.size(); d0 += batch) {
        const uint32_t slots = (uint32_t)std::min<size_t>(batch, deferred.size() - d0);
        run_batch(ctx, res, 0u, deferred.data() + d0, slots, false, 1u, nullptr, tm);
    }
    hipEvent_t ev_end = record(ctx);
    if (ev_end)
        HIP_CHECK(hipEventSynchronize(ev_end));
    else
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (auto& pr : tm.imp) ctx->stats.importance_ms += elapsed(pr.first, pr.second);
    for (auto& pr : tm.trace) ctx->stats.trace_ms += elapsed(pr.first, pr.second);
    for (auto& pr : tm.fin) ctx->stats.finalize_ms += elapsed(pr.first, pr.second);
    ctx->stats.trace_ms += tm.queue_ms;
    for (auto& pr : tm.inf) ctx->pstats.inference_ms += elapsed(pr.first, pr.second);
    for (auto& pr : tm.prop) ctx->pstats.proposal_ms += elapsed(pr.first, pr.second);
    ctx->stats.total_ms = elapsed(ev_begin, ev_end);
    unsigned long long cnt[4];
    HIP_CHECK(hipMemcpy(cnt, ctx->counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
    ctx->stats.rays = cnt[0];
    ctx->stats.ray_steps = cnt[1];
    ctx->stats.importance_walks = cnt[2];
    ctx->stats.importance_steps = cnt[3];
    uint64_t dropped = 0;
    for (uint32_t s = first; s <= last; ++s)
        if (res->dropped_rays) dropped += res->dropped_rays[s];
    ctx->stats.dropped_rays = dropped;
}


// ------------------------------------------------------------------ batches
}  // namespace

namespace hase {

void evaluate_context(hase_context* ctx, hase_result_t* res) { evaluate(ctx, res); }

void validate_arguments(const hase_mesh_t* mesh, const hase_experiment_t* experiment, const hase_compute_t* compute)
{
    validate(mesh, experiment, compute);
}

void calc_phi_ase(const hase_mesh_t* mesh, const hase_experiment_t* experiment, const hase_compute_t* compute,
                  hase_result_t* result)
{
    auto ctx = create_context(mesh, experiment, compute);
    evaluate(ctx.get(), result);
}

void refresh_beta_device(hase_context* ctx, const double* beta, double* block_sums)
{
    HIP_CHECK(hipSetDevice(ctx->device));
    const uint32_t T = ctx->T, layers = ctx->L - 1, P = ctx->P;
    static_assert(sizeof(double) == 8, "FP64");
    if (ctx->tri_area_dev.n < T) ctx->tri_area_dev.upload(ctx->tri_area.data(), T, ctx->stream);
    if (ctx->clad_dev.n < T) ctx->clad_dev.upload(ctx->clad_flag.data(), T, ctx->stream);
    const uint32_t blocks = (P + 255u) / 256u;
    hipLaunchKernelGGL(hase_refresh_beta_kernel, dim3(blocks), dim3(256), 0, ctx->stream, beta,
                       (const double*)ctx->tri_area_dev.p, (const uint32_t*)ctx->clad_dev.p, ctx->h, T, layers,
                       ctx->beta.p, ctx->bv.p, ctx->beta_col.p, block_sums);
    HIP_CHECK(hipGetLastError());
    std::vector<double> sums(blocks);
    HIP_CHECK(hipMemcpyAsync(sums.data(), block_sums, blocks * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (ctx->proposal_kind == HASE_PROPOSAL_LEARNED) {
        // the cluster representatives are built on the host from beta
        std::vector<double> b(P);
        HIP_CHECK(hipMemcpyAsync(b.data(), beta, (size_t)P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (uint32_t j = 0; j < P; ++j) ctx->beta_host[ctx->dev2user[j % T] + (size_t)(j / T) * T] = b[j];
        build_clusters(ctx);
    }
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double B = 0.0;
    for (double v : sums) B += v;
    ctx->Bsum = B;
    ctx->q_generation++;
}

}  // namespace hase

struct hase_batch {
    std::vector<hase_context*> ctx;
    int device = 0;
    hipStream_t stream = nullptr;
    hase::DevBuf<hase::ProblemDev> problems;
    hase::DevBuf<uint32_t> slot_problem, flags, k, extra, launch, nrays, item_rays, target, list, item_off, scalars, N,
        dropped;
    hase::DevBuf<double> s1, s2, mse, se;
    hase::DevBuf<float> phi;
    hase::DevBuf<double2> partial;
    hase::DevBuf<uint32_t> partial_drop;
    hase::DevBuf<unsigned long long> counters;
    hase::GraphCache graph;
    uint64_t generation = 1;
    hase_stats_t stats{};
    ~hase_batch()
    {
        graph.reset();
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace {

// Samples from which a problem of a batch whose records fit in LDS runs on
// its own queue (one resident batch of sample points).
constexpr uint32_t kSoloSamples = 8192;

void batch_evaluate(hase_batch* b, hase_result_t* const* results)
{
    if (!results) throw HaseError(HASE_E_INVALID, "results must not be NULL");
    const size_t n = b->ctx.size();
    for (size_t p = 0; p < n; ++p) {
        hase_result_t* r = results[p];
        if (!r || !r->phi_ase || !r->mse || !r->total_rays || !r->dndt_ase)
            throw HaseError(HASE_E_INVALID, "every result needs phi_ase, mse, total_rays and dndt_ase");
    }
    HIP_CHECK(hipSetDevice(b->device));
    hase_tuning_t tu = b->ctx[0]->tuning;
    tu.mode = HASE_TRACE_QUEUE;
    check_tuning(tu);
    b->stats = hase_stats_t{};
    b->generation++;
    // A problem whose triangle records fit in LDS and whose samples fill a
    // resident batch of their own (kSoloSamples) runs on its own queue, as
    // when evaluated alone: its trace kernel then reads the records from LDS,
    // which the shared queue (items of several problems in one work-group)
    // cannot.  The shared queue takes the rest, the problems too small to
    // fill the chip by themselves, where one queue for all pays.
    std::vector<char> solo(n, 0);
    hase_stats_t solo_stats{};
    for (size_t p = 0; p < n; ++p) {
        hase_context* ctx = b->ctx[p];
        const hase_tuning_t& ct = ctx->tuning;
        const uint32_t count = ctx->comp.max_sample_range - ctx->comp.min_sample_range + 1u;
        solo[p] = ctx->lds_rec.p != nullptr && ctx->lds_staged() && ct.mode == HASE_TRACE_QUEUE && ct.stage_lds &&
                  ct.refill_lanes == 0u &&
                  count >= kSoloSamples;
        if (!solo[p]) continue;
        evaluate(ctx, results[p]);
        const hase_stats_t& st = ctx->stats;
        solo_stats.importance_ms += st.importance_ms + ctx->pstats.inference_ms + ctx->pstats.proposal_ms;
        solo_stats.trace_ms += st.trace_ms;
        solo_stats.finalize_ms += st.finalize_ms;
        solo_stats.rays += st.rays;
        solo_stats.ray_steps += st.ray_steps;
        solo_stats.importance_walks += st.importance_walks;
        solo_stats.importance_steps += st.importance_steps;
        solo_stats.dropped_rays += st.dropped_rays;
        solo_stats.trace_launches += st.trace_launches;
        solo_stats.importance_launches += st.importance_launches;
    }
    // Batch memory per problem: what it would take evaluated alone (all its
    // samples, at most 8192 within 6 GiB), scaled down evenly when the
    // problems together would need more than half the free device memory
    // (each at least one sample's bytes); a context that already holds
    // larger buffers keeps using them.  Small shares leave the largest
    // problems many small waves after the others have finished.
    std::vector<size_t> budget(n, 0);
    {
        size_t free_b = 0, total_b = 0;
        HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
        std::vector<size_t> unit(n), held(n);
        size_t want_sum = 0;
        for (size_t p = 0; p < n; ++p) {
            const hase_context* ctx = b->ctx[p];
            if (solo[p]) continue;
            const size_t count = std::min<size_t>(8192, ctx->comp.max_sample_range - ctx->comp.min_sample_range + 1u);
            const bool shared = ctx->proposal_kind == HASE_PROPOSAL_BETA_VOLUME;
            unit[p] = (shared ? 0u : (size_t)slices(ctx) * ctx->P * 12u + kPrefixBuckets * 4u) +
                      (size_t)queue_items_per_slot(ctx, tu) * 12u + 96u +
                      (ctx->proposal_kind == HASE_PROPOSAL_LEARNED ? (size_t)slices(ctx) * ctx->C * 4u : 0u);
            held[p] = ctx->imp.n * sizeof(float) + ctx->prefix.n * sizeof(double);
            budget[p] = std::min(unit[p] * count, (size_t)6 << 30);
            want_sum += budget[p] > held[p] ? budget[p] - held[p] : 0u;
        }
        const double scale = want_sum > free_b / 2u ? (double)(free_b / 2u) / (double)want_sum : 1.0;
        for (size_t p = 0; p < n; ++p) {
            const size_t extra = budget[p] > held[p] ? budget[p] - held[p] : 0u;
            budget[p] = std::max({held[p] + (size_t)(scale * (double)extra), unit[p], held[p]});
        }
    }
    std::vector<uint32_t> batch(n, 0), next(n), last(n);
    std::vector<std::vector<uint32_t>> deferred(n);
    std::vector<BatchTimers> tm(n);
    for (size_t p = 0; p < n; ++p) {
        hase_context* ctx = b->ctx[p];
        if (solo[p]) {   // evaluated above
            next[p] = ctx->comp.max_sample_range + 1u;
            last[p] = ctx->comp.max_sample_range;
            continue;
        }
        hase_result_t* r = results[p];
        for (uint32_t s = 0; s < ctx->S; ++s) {
            r->phi_ase[s] = 0.0f;
            r->mse[s] = 0.0;
            r->total_rays[s] = 0;
            r->dndt_ase[s] = 0.0;
            if (r->standard_error) r->standard_error[s] = 0.0;
            if (r->dropped_rays) r->dropped_rays[s] = 0;
        }
        for (auto ev : ctx->events) (void)hipEventDestroy(ev);
        ctx->events.clear();
        ctx->stats = hase_stats_t{};
        ctx->pstats = hase_proposal_stats_t{};
        ctx->pstats.clusters = ctx->C;
        next[p] = ctx->comp.min_sample_range;
        last[p] = ctx->comp.max_sample_range;
        if (!(ctx->Bsum > 0.0)) {   // no emitting volume: zeros
            next[p] = last[p] + 1u;
            continue;
        }
        batch[p] = reserve_batch(ctx, last[p] - next[p] + 1u, budget[p], true);
        HIP_CHECK(hipMemsetAsync(ctx->counters.p, 0, 4 * sizeof(unsigned long long), ctx->stream));
        if (ctx->proposal_kind == HASE_PROPOSAL_BETA_VOLUME) build_shared_prefix(ctx, tm[p]);
    }
    b->counters.alloc(4);
    HIP_CHECK(hipMemsetAsync(b->counters.p, 0, 4 * sizeof(unsigned long long), b->stream));
    double queue_ms = 0.0;
    for (;;) {
        // one wave: the next resident batch of every problem with samples left
        std::vector<uint32_t> part, slots_of, slot0_of;
        uint32_t total = 0;
        size_t items = 0;
        for (size_t p = 0; p < n; ++p) {
            if (next[p] > last[p]) continue;
            const uint32_t sl = std::min(batch[p], last[p] - next[p] + 1u);
            part.push_back((uint32_t)p);
            slots_of.push_back(sl);
            slot0_of.push_back(total);
            total += sl;
            items += (size_t)sl * queue_items_per_slot(b->ctx[p], tu);
        }
        if (part.empty()) break;
        for (size_t w = 0; w < part.size(); ++w) {
            hase_context* ctx = b->ctx[part[w]];
            build_proposal(ctx, next[part[w]], nullptr, slots_of[w], ctx->proposal_kind == HASE_PROPOSAL_LEARNED,
                           tm[part[w]]);
        }
        for (size_t w = 0; w < part.size(); ++w) HIP_CHECK(hipStreamSynchronize(b->ctx[part[w]]->stream));
        for (auto* buf : {&b->slot_problem, &b->flags, &b->k, &b->extra, &b->launch, &b->nrays, &b->item_rays,
                          &b->target, &b->list, &b->N, &b->dropped})
            buf->alloc(total);
        b->item_off.alloc((size_t)total + 1u);
        b->scalars.alloc(kQueueScalars);
        b->s1.alloc(total);
        b->s2.alloc(total);
        b->mse.alloc(total);
        b->se.alloc(total);
        b->phi.alloc(total);
        b->partial.alloc(items);
        b->partial_drop.alloc(items);
        b->problems.alloc(part.size());
        std::vector<ProblemDev> table(part.size());
        std::vector<uint32_t> owner(total);
        uint32_t rounds = 0;
        for (size_t w = 0; w < part.size(); ++w) {
            hase_context* ctx = b->ctx[part[w]];
            const bool learned = ctx->proposal_kind == HASE_PROPOSAL_LEARNED;
            table[w] = queue_problem(ctx, slot0_of[w], next[part[w]], false, 0u, learned && ctx->proposal_fallback);
            for (uint32_t i = 0; i < slots_of[w]; ++i) owner[slot0_of[w] + i] = (uint32_t)w;
            const Schedule sc{ctx->exp.min_rays_per_sample, ctx->exp.max_rays_per_sample, ctx->comp.adaptive_steps,
                              ctx->comp.max_repetitions};
            rounds = std::max(rounds, queue_max_rounds(sc));
        }
        HIP_CHECK(hipMemcpyAsync(b->problems.p, table.data(), table.size() * sizeof(ProblemDev),
                                 hipMemcpyHostToDevice, b->stream));
        HIP_CHECK(hipMemcpyAsync(b->slot_problem.p, owner.data(), total * sizeof(uint32_t), hipMemcpyHostToDevice,
                                 b->stream));
        QueueState qs;
        std::memset(&qs, 0, sizeof(qs));
        qs.problems = b->problems.p;
        qs.slot_problem = b->slot_problem.p;
        qs.flags = b->flags.p;
        qs.k = b->k.p;
        qs.extra = b->extra.p;
        qs.launch = b->launch.p;
        qs.nrays = b->nrays.p;
        qs.item_rays = b->item_rays.p;
        qs.target = b->target.p;
        qs.list = b->list.p;
        qs.item_off = b->item_off.p;
        qs.scalars = b->scalars.p;
        qs.s1 = b->s1.p;
        qs.s2 = b->s2.p;
        qs.N = b->N.p;
        qs.dropped = b->dropped.p;
        qs.phi = b->phi.p;
        qs.mse = b->mse.p;
        qs.se = b->se.p;
        qs.partial = b->partial.p;
        qs.partial_drop = b->partial_drop.p;
        qs.counters = b->counters.p;
        qs.rays_per_item = tu.rays_per_item;
        qs.queue_chunk = tu.queue_chunk;
        qs.refill = tu.refill_lanes;
        qs.budget = tu.step_budget;
        b->generation++;   // buffers may have moved: capture again
        bool table_compact = true;
        for (size_t w = 0; w < part.size(); ++w) table_compact = table_compact && table[w].m.cmp_rec != nullptr;
        queue_ms += queue_run(b->stream, b->device, tu, qs, total, rounds, b->graph, b->generation, nullptr,
                              table_compact);
        b->stats.trace_launches += rounds;
        std::vector<float> h_phi(total);
        std::vector<double> h_mse(total), h_se(total);
        std::vector<uint32_t> h_N(total), h_drop(total), h_flags(total);
        HIP_CHECK(hipMemcpyAsync(h_phi.data(), b->phi.p, total * sizeof(float), hipMemcpyDeviceToHost, b->stream));
        HIP_CHECK(hipMemcpyAsync(h_mse.data(), b->mse.p, total * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        HIP_CHECK(hipMemcpyAsync(h_se.data(), b->se.p, total * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        HIP_CHECK(hipMemcpyAsync(h_N.data(), b->N.p, total * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
        HIP_CHECK(hipMemcpyAsync(h_drop.data(), b->dropped.p, total * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 b->stream));
        HIP_CHECK(hipMemcpyAsync(h_flags.data(), b->flags.p, total * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 b->stream));
        HIP_CHECK(hipStreamSynchronize(b->stream));
        for (size_t w = 0; w < part.size(); ++w) {
            const uint32_t p = part[w];
            hase_context* ctx = b->ctx[p];
            const hase_experiment_t& e = ctx->exp;
            hase_result_t* r = results[p];
            for (uint32_t i = 0; i < slots_of[w]; ++i) {
                const uint32_t g = slot0_of[w] + i, s = next[p] + i;
                r->phi_ase[s] = h_phi[g];
                r->mse[s] = h_mse[g];
                r->total_rays[s] = h_N[g];
                r->dndt_ase[s] = (ctx->beta_cells[s] * (e.max_sigma_e + e.max_sigma_a) - e.max_sigma_a) * (double)h_phi[g];
                if (r->standard_error) r->standard_error[s] = h_se[g];
                if (r->dropped_rays) r->dropped_rays[s] = h_drop[g];
                if (h_flags[g] & kQueueDeferred) {
                    deferred[p].push_back(s);
                    ctx->pstats.fallback_samples++;
                }
            }
            next[p] += slots_of[w];
        }
    }
    // learned-proposal misses restart on each problem's walked map
    for (size_t p = 0; p < n; ++p) {
        if (deferred[p].empty()) continue;
        hase_context* ctx = b->ctx[p];
        const uint32_t bs = reserve_batch(ctx, (uint32_t)deferred[p].size(), budget[p]);
        for (size_t d0 = 0; d0 < deferred[p].size(); d0 += bs) {
            const uint32_t sl = (uint32_t)std::min<size_t>(bs, deferred[p].size() - d0);
            run_batch(ctx, results[p], 0u, deferred[p].data() + d0, sl, false, 1u, nullptr, tm[p]);
        }
    }
    unsigned long long cnt[4] = {0, 0, 0, 0};
    HIP_CHECK(hipMemcpy(cnt, b->counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
    b->stats.rays = cnt[0];
    b->stats.ray_steps = cnt[1];
    b->stats.trace_ms = queue_ms;
    for (size_t p = 0; p < n; ++p) {
        hase_context* ctx = b->ctx[p];
        if (solo[p]) continue;
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (auto& pr : tm[p].imp) b->stats.importance_ms += elapsed(pr.first, pr.second);
        for (auto& pr : tm[p].trace) b->stats.trace_ms += elapsed(pr.first, pr.second);
        for (auto& pr : tm[p].inf) ctx->pstats.inference_ms += elapsed(pr.first, pr.second);
        for (auto& pr : tm[p].prop) ctx->pstats.proposal_ms += elapsed(pr.first, pr.second);
        b->stats.importance_ms += ctx->pstats.inference_ms + ctx->pstats.proposal_ms;
        b->stats.trace_ms += tm[p].queue_ms;
        unsigned long long c2[4];
        HIP_CHECK(hipMemcpy(c2, ctx->counters.p, sizeof(c2), hipMemcpyDeviceToHost));
        b->stats.rays += c2[0];
        b->stats.ray_steps += c2[1];
        b->stats.importance_walks += c2[2];
        b->stats.importance_steps += c2[3];
        b->stats.importance_launches += ctx->stats.importance_launches;
        uint64_t dr = 0;
        if (results[p]->dropped_rays)
            for (uint32_t s = ctx->comp.min_sample_range; s <= ctx->comp.max_sample_range; ++s) dr += results[p]->dropped_rays[s];
        b->stats.dropped_rays += dr;
    }
    b->stats.importance_ms += solo_stats.importance_ms;
    b->stats.trace_ms += solo_stats.trace_ms;
    b->stats.finalize_ms += solo_stats.finalize_ms;
    b->stats.rays += solo_stats.rays;
    b->stats.ray_steps += solo_stats.ray_steps;
    b->stats.importance_walks += solo_stats.importance_walks;
    b->stats.importance_steps += solo_stats.importance_steps;
    b->stats.dropped_rays += solo_stats.dropped_rays;
    b->stats.trace_launches += solo_stats.trace_launches;
    b->stats.importance_launches += solo_stats.importance_launches;
}

// Autotuner of the persistent work queue: coordinate descent over the
// kernel's launch shape on one adaptive batch of the first probe samples,
// each setting timed (best of two) on the same proposal.
void autotune(hase_context* ctx, uint32_t probe, hase_tuning_t* best)
{
    HIP_CHECK(hipSetDevice(ctx->device));
    hase_tuning_t cur = ctx->tuning;
    cur.mode = HASE_TRACE_QUEUE;
    cur.probe_ms = 0.0;
    if (!(ctx->Bsum > 0.0)) {
        ctx->tuning = cur;
        *best = cur;
        return;
    }
    static const uint32_t blocks[] = {128u, 256u, 512u, 1024u};
    static const uint32_t items[] = {256u, 512u, 1024u, 2048u};
    static const uint32_t chunks[] = {1u, 2u, 4u};
    static const uint32_t refills[] = {0u, 8u, 16u, 32u};
    static const uint32_t budgets[] = {4u, 8u, 16u};
    const uint32_t first = ctx->comp.min_sample_range, last = ctx->comp.max_sample_range;
    const uint32_t want = std::min(probe ? probe : 512u, last - first + 1u);
    const bool learned = ctx->proposal_kind == HASE_PROPOSAL_LEARNED;
    ctx->tuning = cur;
    ctx->tuning.rays_per_item = items[0];   // largest item count any candidate needs
    const uint32_t slots = reserve_batch(ctx, want);
    HIP_CHECK(hipMemsetAsync(ctx->counters.p, 0, 4 * sizeof(unsigned long long), ctx->stream));
    BatchTimers tm;
    if (ctx->proposal_kind == HASE_PROPOSAL_BETA_VOLUME) build_shared_prefix(ctx, tm);
    build_proposal(ctx, first, nullptr, slots, learned, tm);
    auto time_it = [&](const hase_tuning_t& t) {
        ctx->tuning = t;
        queue_reserve(ctx, slots);
        ctx->q_generation++;
        float ms = std::numeric_limits<float>::infinity();
        for (int rep = 0; rep < 2; ++rep) ms = std::min(ms, queue_rounds(ctx, first, false, slots, 0u, false));
        return (double)ms;
    };
    double best_ms = time_it(cur);
    for (int pass = 0; pass < 2; ++pass) {
        auto sweep = [&](const uint32_t* vals, size_t n, uint32_t hase_tuning_t::*field) {
            for (size_t i = 0; i < n; ++i) {
                if (cur.*field == vals[i]) continue;
                hase_tuning_t t = cur;
                t.*field = vals[i];
                const double ms = time_it(t);
                if (ms < 0.98 * best_ms) {
                    best_ms = ms;
                    cur = t;
                }
            }
        };
        if (ctx->lds_rec.p) {
            // one work-group per CU holds a staged mesh: try it with the largest
            // block (compact records in global memory keep the block size)
            hase_tuning_t t = cur;
            t.stage_lds = cur.stage_lds ? 0u : 1u;
            t.block_size = t.stage_lds && ctx->lds_staged() ? 1024u : cur.block_size;
            const double ms = time_it(t);
            if (ms < 0.98 * best_ms) {
                best_ms = ms;
                cur = t;
            }
        }
        sweep(blocks, 4, &hase_tuning_t::block_size);
        sweep(items, 4, &hase_tuning_t::rays_per_item);
        sweep(chunks, 3, &hase_tuning_t::queue_chunk);
        sweep(refills, 4, &hase_tuning_t::refill_lanes);
        sweep(budgets, 3, &hase_tuning_t::step_budget);
    }
    cur.probe_ms = best_ms;
    ctx->tuning = cur;
    ctx->q_generation++;
    *best = cur;
}

}  // namespace

namespace hase {
thread_local std::string g_last_error;
thread_local Arena* g_arena = nullptr;

hipEvent_t record(hase_context* ctx)
{
    if (g_arena) return nullptr;   // no events are created in allocation-free calls
    hipEvent_t ev;
    HIP_CHECK(hipEventCreate(&ev));
    HIP_CHECK(hipEventRecord(ev, ctx->stream));
    ctx->events.push_back(ev);
    return ev;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.0f;
    if (!a || !b) return ms;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms;
}

// One batch of resident samples: proposal (walked map or learned), then
// adaptive trace launches until every slot converges or reaches its budget.
// Slot i holds sample samples[i] (or b0 + i when samples is null).  With the
// learned proposal and `deferred`, samples that miss the threshold after the
// first launch are handed back (they restart on the walked map later, from
// launch index 1, so their random streams stay distinct).
}  // namespace hase

extern "C" {

int hase_calc_phi_ase(const hase_mesh_t* mesh, const hase_experiment_t* experiment, const hase_compute_t* compute,
                      hase_result_t* result)
{
    return guarded([&] { calc_phi_ase(mesh, experiment, compute, result); });
}

int hase_context_create(const hase_mesh_t* mesh, const hase_experiment_t* experiment, const hase_compute_t* compute,
                        hase_context_t** out_context)
{
    return guarded([&] {
        if (!out_context) throw HaseError(HASE_E_INVALID, "out_context must not be NULL");
        *out_context = nullptr;
        *out_context = create_context(mesh, experiment, compute).release();
    });
}

int hase_context_update_beta(hase_context_t* context, const double* beta_volume, const double* beta_cells)
{
    return guarded([&] {
        if (!context || !beta_volume) throw HaseError(HASE_E_INVALID, "context and beta_volume must not be NULL");
        HIP_CHECK(hipSetDevice(context->device));
        upload_beta(context, beta_volume, beta_cells);
    });
}

int hase_context_evaluate(hase_context_t* context, hase_result_t* result)
{
    return guarded([&] {
        if (!context) throw HaseError(HASE_E_INVALID, "context must not be NULL");
        evaluate(context, result);
    });
}

int hase_context_stats(const hase_context_t* context, hase_stats_t* stats)
{
    return guarded([&] {
        if (!context || !stats) throw HaseError(HASE_E_INVALID, "context and stats must not be NULL");
        *stats = context->stats;
    });
}

void hase_context_destroy(hase_context_t* context)
{
    if (context) {
        (void)hipSetDevice(context->device);
        (void)hipStreamSynchronize(context->stream);
        delete context;
    }
}

int hase_context_trace_rays(hase_context_t* context, uint32_t n, const float* start, const uint32_t* start_prism,
                            const float* end, const uint32_t* reflections, const uint32_t* first_plane,
                            float* beta_length, float* gain_length, float* clad_length, float* reflectance,
                            uint32_t* steps, int32_t* status)
{
    return guarded([&] {
        if (!context) throw HaseError(HASE_E_INVALID, "context must not be NULL");
        if (n == 0) return;
        if (!start || !start_prism || !end || !reflections || !first_plane || !beta_length || !gain_length ||
            !clad_length || !reflectance || !steps || !status)
            throw HaseError(HASE_E_INVALID, "all ray arrays are required");
        for (uint32_t i = 0; i < n; ++i)
            if (start_prism[i] >= context->P) throw HaseError(HASE_E_INVALID, "start_prism out of range");
        HIP_CHECK(hipSetDevice(context->device));
        hipStream_t st = context->stream;
        DevBuf<float> d_start, d_end, d_bl, d_gl, d_cl, d_rf;
        DevBuf<uint32_t> d_prism, d_nr, d_fp, d_steps;
        DevBuf<int32_t> d_status;
        d_start.upload(start, 3u * (size_t)n, st);
        d_end.upload(end, 3u * (size_t)n, st);
        d_prism.upload(start_prism, n, st);
        d_nr.upload(reflections, n, st);
        d_fp.upload(first_plane, n, st);
        d_bl.alloc(n); d_gl.alloc(n); d_cl.alloc(n); d_rf.alloc(n); d_steps.alloc(n); d_status.alloc(n);
        hipLaunchKernelGGL(hase_single_rays_kernel, dim3((n + 127) / 128), dim3(128), 0, st, mesh_view(context),
                           context->h, context->user2dev_dev.p, n, d_start.p, d_prism.p, d_end.p, d_nr.p, d_fp.p,
                           d_bl.p, d_gl.p, d_cl.p, d_rf.p, d_steps.p, d_status.p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(beta_length, d_bl.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(gain_length, d_gl.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(clad_length, d_cl.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(reflectance, d_rf.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(steps, d_steps.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(status, d_status.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    });
}

int hase_context_importance(hase_context_t* context, uint32_t sample, float* importance, uint32_t capacity)
{
    return guarded([&] {
        if (!context || !importance) throw HaseError(HASE_E_INVALID, "context and importance must not be NULL");
        if (sample >= context->S) throw HaseError(HASE_E_INVALID, "sample out of range");
        const uint32_t R = slices(context);
        const size_t n_ent = (size_t)R * context->P;
        if (capacity < n_ent) throw HaseError(HASE_E_INVALID, "importance capacity must be at least R * P");
        if (!(context->Bsum > 0.0)) {
            std::fill(importance, importance + n_ent, 0.0f);
            return;
        }
        HIP_CHECK(hipSetDevice(context->device));
        context->imp.alloc(std::max(context->imp.n, n_ent));
        context->prefix.alloc(std::max(context->prefix.n, n_ent));
        context->Q.alloc(std::max<size_t>(context->Q.n, 1));
        const hase_experiment_t& e = context->exp;
        ImportanceParams ip;
        ip.m = mesh_view(context);
        ip.pts = context->pts.p;
        ip.sample_base = sample;
        ip.slots = 1;
        ip.R = R;
        ip.n_lambda = e.number_of_wavelengths;
        ip.c1 = context->c1.p;
        ip.c2 = context->c2.p;
        ip.inv_lambda = 1.0f / (float)e.number_of_wavelengths;
        ip.eps = e.defensive_fraction;
        ip.Bsum = context->Bsum;
        ip.imp = context->imp.p;
        ip.prefix = context->prefix.p;
        ip.Q = context->Q.p;
        ip.counters = context->counters.p;
        ip.active = nullptr;
        ip.sample_of = nullptr;
        ip.transfer = 0;
        launch_importance(context, ip, 1u);
        std::vector<float> dev(n_ent);
        HIP_CHECK(hipMemcpyAsync(dev.data(), context->imp.p, n_ent * sizeof(float), hipMemcpyDeviceToHost,
                                 context->stream));
        HIP_CHECK(hipStreamSynchronize(context->stream));
        const uint32_t T = context->T, P = context->P;
        for (size_t e = 0; e < n_ent; ++e) {
            size_t k = e / P, j = e % P;
            importance[k * P + context->dev2user[j % T] + (j / T) * T] = dev[e];
        }
    });
}

int hase_context_set_proposal(hase_context_t* context, const hase_proposal_t* proposal)
{
    return guarded([&] {
        if (!context || !proposal) throw HaseError(HASE_E_INVALID, "context and proposal must not be NULL");
        if (proposal->kind == HASE_PROPOSAL_PLAIN || proposal->kind == HASE_PROPOSAL_BETA_VOLUME) {
            context->proposal_kind = proposal->kind;
            return;
        }
        if (proposal->kind != HASE_PROPOSAL_LEARNED) throw HaseError(HASE_E_INVALID, "unknown proposal kind");
        if (!proposal->params || !proposal->input_shift || !proposal->input_scale)
            throw HaseError(HASE_E_INVALID, "a learned proposal needs params, input_shift and input_scale");
        if (proposal->cluster_triangles == 0 || proposal->cluster_levels == 0)
            throw HaseError(HASE_E_INVALID, "cluster_triangles and cluster_levels must be at least 1");
        for (size_t i = 0; i < kMlpParams; ++i)
            if (!std::isfinite(proposal->params[i])) throw HaseError(HASE_E_INVALID, "params must be finite");
        for (int i = 0; i < kMlpIn; ++i)
            if (!std::isfinite(proposal->input_shift[i]) || !std::isfinite(proposal->input_scale[i]))
                throw HaseError(HASE_E_INVALID, "input_shift and input_scale must be finite");
        HIP_CHECK(hipSetDevice(context->device));
        std::vector<uint4> frags((size_t)kFrags * 64);
        std::vector<float> vec(kMlpVec);
        pack_mlp_bf16(proposal->params, frags.data(), vec.data());
        context->mlp_frags.upload(frags.data(), frags.size(), context->stream);
        context->mlp_vec.upload(vec.data(), vec.size(), context->stream);
        for (int i = 0; i < kMlpIn; ++i) {
            context->in_shift[i] = proposal->input_shift[i];
            context->in_scale[i] = proposal->input_scale[i];
        }
        context->cluster_t = std::min(proposal->cluster_triangles, context->T);
        context->cluster_l = std::min(proposal->cluster_levels, context->L - 1);
        context->proposal_fallback = proposal->fallback != 0;
        build_clusters(context);
        context->proposal_kind = HASE_PROPOSAL_LEARNED;
        HIP_CHECK(hipStreamSynchronize(context->stream));
    });
}

int hase_context_transfer_map(hase_context_t* context, uint32_t first, uint32_t count, float* transfer,
                              size_t capacity)
{
    return guarded([&] {
        if (!context || !transfer) throw HaseError(HASE_E_INVALID, "context and transfer must not be NULL");
        if (count == 0 || first >= context->S || count > context->S - first)
            throw HaseError(HASE_E_INVALID, "sample range out of bounds");
        const uint32_t R = slices(context), T = context->T, P = context->P;
        const size_t n_ent = (size_t)R * P;
        if (capacity < (size_t)count * n_ent) throw HaseError(HASE_E_INVALID, "transfer capacity must be count * R * P");
        HIP_CHECK(hipSetDevice(context->device));
        const uint32_t batch = (uint32_t)std::max<size_t>(1, std::min<size_t>(count, ((size_t)1 << 28) / n_ent));
        context->imp.alloc(std::max(context->imp.n, (size_t)batch * n_ent));
        context->counters.alloc(4);
        const hase_experiment_t& e = context->exp;
        std::vector<float> dev((size_t)batch * n_ent);
        for (uint32_t b0 = 0; b0 < count; b0 += batch) {
            const uint32_t slots = std::min(batch, count - b0);
            ImportanceParams ip;
            ip.m = mesh_view(context);
            ip.pts = context->pts.p;
            ip.sample_base = first + b0;
            ip.slots = slots;
            ip.R = R;
            ip.n_lambda = e.number_of_wavelengths;
            ip.c1 = context->c1.p;
            ip.c2 = context->c2.p;
            ip.inv_lambda = 1.0f / (float)e.number_of_wavelengths;
            ip.eps = e.defensive_fraction;
            ip.Bsum = context->Bsum;
            ip.imp = context->imp.p;
            ip.prefix = nullptr;
            ip.Q = nullptr;
            ip.counters = context->counters.p;
            ip.active = nullptr;
            ip.sample_of = nullptr;
            ip.transfer = 1;
            launch_importance(context, ip, slots);
            HIP_CHECK(hipMemcpyAsync(dev.data(), context->imp.p, (size_t)slots * n_ent * sizeof(float),
                                     hipMemcpyDeviceToHost, context->stream));
            HIP_CHECK(hipStreamSynchronize(context->stream));
            for (uint32_t i = 0; i < slots; ++i) {
                const float* src = dev.data() + (size_t)i * n_ent;
                float* dst = transfer + (size_t)(b0 + i) * n_ent;
                for (size_t x = 0; x < n_ent; ++x) {
                    const size_t k = x / P, j = x % P;
                    dst[k * P + context->dev2user[j % T] + (j / T) * (size_t)T] = src[x];
                }
            }
        }
    });
}

int hase_context_proposal_logits(hase_context_t* context, uint32_t first, uint32_t count, float* logits,
                                 size_t capacity, float* clusters, uint32_t* n_clusters)
{
    return guarded([&] {
        if (!context || !logits) throw HaseError(HASE_E_INVALID, "context and logits must not be NULL");
        if (context->proposal_kind != HASE_PROPOSAL_LEARNED)
            throw HaseError(HASE_E_INVALID, "no learned proposal is set on this context");
        if (count == 0 || first >= context->S || count > context->S - first)
            throw HaseError(HASE_E_INVALID, "sample range out of bounds");
        const uint32_t R = slices(context), C = context->C;
        const size_t per = (size_t)R * C;
        if (capacity < (size_t)count * per) throw HaseError(HASE_E_INVALID, "logits capacity must be count * R * C");
        HIP_CHECK(hipSetDevice(context->device));
        const uint32_t batch = (uint32_t)std::max<size_t>(1, std::min<size_t>(count, ((size_t)1 << 26) / per));
        context->logits.alloc(std::max(context->logits.n, (size_t)batch * per));
        for (uint32_t b0 = 0; b0 < count; b0 += batch) {
            const uint32_t slots = std::min(batch, count - b0);
            launch_proposal_inference(context, first + b0, slots, context->logits.p);
            HIP_CHECK(hipMemcpyAsync(logits + (size_t)b0 * per, context->logits.p, (size_t)slots * per * sizeof(float),
                                     hipMemcpyDeviceToHost, context->stream));
        }
        HIP_CHECK(hipStreamSynchronize(context->stream));
        if (clusters)
            for (uint32_t c = 0; c < C; ++c) {
                clusters[4 * c] = context->clusters_host[c].x;
                clusters[4 * c + 1] = context->clusters_host[c].y;
                clusters[4 * c + 2] = context->clusters_host[c].z;
                clusters[4 * c + 3] = context->clusters_host[c].w;
            }
        if (n_clusters) *n_clusters = C;
    });
}

int hase_context_proposal_stats(const hase_context_t* context, hase_proposal_stats_t* stats)
{
    return guarded([&] {
        if (!context || !stats) throw HaseError(HASE_E_INVALID, "context and stats must not be NULL");
        *stats = context->pstats;
    });
}

int hase_context_set_tuning(hase_context_t* context, const hase_tuning_t* tuning)
{
    return guarded([&] {
        if (!context || !tuning) throw HaseError(HASE_E_INVALID, "context and tuning must not be NULL");
        check_tuning(*tuning);
        context->tuning = *tuning;
        context->q_generation++;
    });
}

int hase_context_get_tuning(const hase_context_t* context, hase_tuning_t* tuning)
{
    return guarded([&] {
        if (!context || !tuning) throw HaseError(HASE_E_INVALID, "context and tuning must not be NULL");
        *tuning = context->tuning;
    });
}

int hase_context_autotune(hase_context_t* context, uint32_t probe_samples, hase_tuning_t* best)
{
    return guarded([&] {
        if (!context || !best) throw HaseError(HASE_E_INVALID, "context and best must not be NULL");
        autotune(context, probe_samples, best);
    });
}

int hase_batch_create(uint32_t n_problems, hase_context_t* const* contexts, hase_batch_t** out_batch)
{
    return guarded([&] {
        if (!out_batch) throw HaseError(HASE_E_INVALID, "out_batch must not be NULL");
        *out_batch = nullptr;
        if (n_problems == 0 || !contexts) throw HaseError(HASE_E_INVALID, "a batch needs at least one context");
        auto b = std::make_unique<hase_batch>();
        for (uint32_t i = 0; i < n_problems; ++i) {
            if (!contexts[i]) throw HaseError(HASE_E_INVALID, "contexts must not be NULL");
            if (contexts[i]->device != contexts[0]->device)
                throw HaseError(HASE_E_INVALID, "every context of a batch must be on one device");
            for (uint32_t j = 0; j < i; ++j)
                if (contexts[j] == contexts[i]) throw HaseError(HASE_E_INVALID, "a context appears twice in the batch");
            b->ctx.push_back(contexts[i]);
        }
        b->device = contexts[0]->device;
        HIP_CHECK(hipSetDevice(b->device));
        HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        *out_batch = b.release();
    });
}

int hase_batch_evaluate(hase_batch_t* batch, hase_result_t* const* results)
{
    return guarded([&] {
        if (!batch) throw HaseError(HASE_E_INVALID, "batch must not be NULL");
        batch_evaluate(batch, results);
    });
}

int hase_batch_stats(const hase_batch_t* batch, hase_stats_t* stats)
{
    return guarded([&] {
        if (!batch || !stats) throw HaseError(HASE_E_INVALID, "batch and stats must not be NULL");
        *stats = batch->stats;
    });
}

void hase_batch_destroy(hase_batch_t* batch)
{
    if (batch) {
        (void)hipSetDevice(batch->device);
        delete batch;
    }
}

const char* hase_last_error(void) { return hase::g_last_error.c_str(); }

int hase_abi_version(void) { return HASE_ABI_VERSION; }

const char* hase_build_info(void)
{
    return "hase_ase gfx950 wave64 fp32-geometry fp64-statistics";
}

}  // extern "C"
