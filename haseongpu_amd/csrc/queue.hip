// queue.hip - adaptive rounds of a batch of samples as one persistent
// work-queue kernel per round, with the per-sample schedule on the device.
//
// A round is three kernels:
//   plan      (one work-group): for every slot still active, the rays of this
//             round from its own schedule state (N, k, extra;
//             include/core/types.hpp:333-354 with the pooled-launch rule), the
//             list of slots with work and the exclusive prefix of their queue
//             items (rays_per_item rays each);
//   trace     (persistent grid): each wavefront takes queue_chunk items at a
//             time from a device cursor; its lanes walk the item's rays and,
//             whenever refill_lanes of them are idle, start the next rays of
//             the item, so a long ray no longer holds 63 finished lanes; the
//             wavefront's FP64 sums are written per item (no atomics);
//   finalize  (one lane per listed slot): pools its items in order, forms
//             phiAse / mse in FP64 and decides convergence.
// All round inputs live in device memory, so the rounds of a batch replay one
// captured hipGraph back to back; rounds with nothing left to do are empty.
// Slots of several problems (meshes with their own proposal, spectra and
// schedule) share one queue: per-problem constants come from a table.
// Ray r of a slot's launch uses the same Philox counters and systematic
// offset as the one-launch-per-round kernel, so both paths trace identical
// rays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/hase_ase.h"
#include "common.h"
#include "context.h"
#include "device_common.h"
#include "prism_walk.h"
#include "queue.h"

namespace hase {
namespace {

constexpr int kPlanBlock = 1024;
constexpr int kQueueMaxBlock = 1024;
enum { kNActive = 0, kHead = 1, kSlots = 2 };

__global__ void __launch_bounds__(kPlanBlock) hase_queue_plan_kernel(QueueState q)
{
    __shared__ uint32_t part_items[kPlanBlock], part_slots[kPlanBlock];
    const uint32_t slots = q.scalars[kSlots];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (slots + kPlanBlock - 1) / kPlanBlock;
    const uint32_t s0 = min(slots, tid * per), s1 = min(slots, s0 + per);
    uint32_t my_items = 0, my_slots = 0;
    for (uint32_t s = s0; s < s1; ++s) {
        uint32_t nr = 0;
        if (q.flags[s] & kQueueActive) {
            const ProblemDev& pd = q.problems[q.slot_problem[s]];
            uint32_t kk = q.k[s], ex = q.extra[s], tg = 0;
            const uint32_t n = q.N[s];
            for (;;) {
                tg = pd.targets[min(kk, pd.n_targets - 1u)];
                if (tg > n) { nr = tg - n; break; }
                if (n >= pd.max_rays || tg >= pd.max_rays) {
                    if (ex + 1u >= pd.reps) break;
                    ex++;
                    nr = pd.max_rays;
                    break;
                }
                if (kk + 1u >= pd.n_targets) break;   // flat schedule: nothing more to add
                kk++;
            }
            q.k[s] = kk;
            q.extra[s] = ex;
            q.target[s] = tg;
            if (nr == 0u) q.flags[s] &= ~kQueueActive;
        }
        q.nrays[s] = nr;
        // item length: the tuned one, or longer so the round fits kQueueItemsPerSlot items
        uint32_t ir = max(q.rays_per_item, (nr + kQueueItemsPerSlot - 1u) / kQueueItemsPerSlot);
        ir = (ir + 63u) & ~63u;
        q.item_rays[s] = ir;
        if (nr) {
            my_items += (nr + ir - 1u) / ir;
            my_slots++;
        }
    }
    part_items[tid] = my_items;
    part_slots[tid] = my_slots;
    __syncthreads();
    // inclusive Hillis-Steele scans of both counts
    for (uint32_t off = 1; off < kPlanBlock; off <<= 1) {
        const uint32_t a = tid >= off ? part_items[tid - off] : 0u;
        const uint32_t b = tid >= off ? part_slots[tid - off] : 0u;
        __syncthreads();
        part_items[tid] += a;
        part_slots[tid] += b;
        __syncthreads();
    }
    uint32_t item = part_items[tid] - my_items, at = part_slots[tid] - my_slots;
    for (uint32_t s = s0; s < s1; ++s) {
        const uint32_t nr = q.nrays[s];
        if (!nr) continue;
        q.list[at] = s;
        q.item_off[at] = item;
        item += (nr + q.item_rays[s] - 1u) / q.item_rays[s];
        at++;
    }
    if (tid == kPlanBlock - 1) {
        q.item_off[part_slots[tid]] = part_items[tid];
        q.scalars[kNActive] = part_slots[tid];
        q.scalars[kHead] = 0u;
    }
}

// Per-lane ray of a queue item: walk state plus what its weight needs.
struct LaneRay {
    WalkState w;
    double q;        // proposal probability of the drawn (slice, prism)
    uint32_t j;      // prism (device order)
};

struct ItemCtx {
    SampleGeom sp;
    const double* pre;
    const uint32_t* bk;       // prefix buckets of the item's sample
    size_t n;
    double Q, U, scale;
    uint32_t sample, launch, r0, seed;
};

// Start point, start triangle's corners, end image and proposal probability
// of ray `ray` of an item.
struct RayStart {
    float ax, ay, az, ez;
    float2 p0, p1, p2;
    uint32_t t, l, nr, fp;
};

// Proposal entry of ray `ray` of an item: the first entry whose prefix
// exceeds the ray's systematic target, searched between the target's buckets.
__device__ __forceinline__ size_t ray_entry(const ItemCtx& c, uint32_t ray)
{
    const double target = ((double)ray + c.U) * c.scale;
    const double kt = target * (double)kPrefixBuckets;
    const uint32_t kb = kt >= (double)(kPrefixBuckets - 1u) ? kPrefixBuckets - 1u : (uint32_t)kt;
    size_t lo = c.bk[kb], hi = (kb + 1u < kPrefixBuckets ? (size_t)c.bk[kb + 1u] : c.n - 1) + 1;
    while (lo < hi) {
        const size_t mid = (lo + hi) >> 1;
        if (c.pre[mid] > target) hi = mid; else lo = mid + 1;
    }
    return lo < c.n ? lo : c.n - 1;
}

// Estimated prism steps of a ray from entry e to the item's sample: triangle
// crossings over the xy distance from the start triangle's first corner,
// plus the level planes between the start layer's middle and the end image.
// Only used to order the rays of a wavefront (longer walks together).
template <class Fetch>
__device__ __forceinline__ float ray_steps_estimate(const MeshDev& m, const ItemCtx& c, size_t e, const Fetch& fetch)
{
    const uint32_t T = m.T, P = T * m.layers;
    const uint32_t k = (uint32_t)(e / P), j = (uint32_t)(e % P);
    const uint32_t t = j % T, l = j / T;
    const float2 p0 = fetch.vertex(m, t, 0u);
    const float ez = (float)image_plane(c.sp.plane, (int)m.layers, (k + 1u) >> 1, k & 1u);
    const float dx = p0.x - c.sp.px, dy = p0.y - c.sp.py;
    return sqrtf(dx * dx + dy * dy) * m.steps_per_cm + fabsf((float)l + 0.5f - ez);
}

template <class Fetch>
__device__ __forceinline__ RayStart draw_ray_at(const MeshDev& m, const ItemCtx& c, uint32_t ray, size_t e, double& q,
                                                uint32_t& jout, const Fetch& fetch)
{
    const uint32_t T = m.T, P = T * m.layers;
    q = c.pre[e] - (e ? c.pre[e - 1] : 0.0);
    const uint32_t k = (uint32_t)(e / P), j = (uint32_t)(e % P);
    const uint32_t t = j % T, l = j / T;
    jout = j;
    uint32_t rnd[4];
    philox4x32_10(ray, c.sample, c.launch, kTagRay, c.seed, kKeyTag, rnd);
    float u1 = u32_to_unit(rnd[0]), u2 = u32_to_unit(rnd[1]), u3 = u32_to_unit(rnd[2]);
    if (u1 + u2 > 1.0f) { u1 = 1.0f - u1; u2 = 1.0f - u2; }
    RayStart r;
    r.p0 = fetch.vertex(m, t, 0u);
    r.p1 = fetch.vertex(m, t, 1u);
    r.p2 = fetch.vertex(m, t, 2u);
    r.ax = r.p0.x + u1 * (r.p1.x - r.p0.x) + u2 * (r.p2.x - r.p0.x);
    r.ay = r.p0.y + u1 * (r.p1.y - r.p0.y) + u2 * (r.p2.y - r.p0.y);
    r.az = (float)l + u3;   // layer units
    r.t = t;
    r.l = l;
    r.nr = (k + 1u) >> 1;
    r.fp = k & 1u;
    r.ez = (float)image_plane(c.sp.plane, (int)m.layers, r.nr, r.fp);
    return r;
}

template <class Fetch>
__device__ __forceinline__ RayStart draw_ray(const MeshDev& m, const ItemCtx& c, uint32_t ray, double& q, uint32_t& jout,
                                             const Fetch& fetch)
{
    return draw_ray_at(m, c, ray, ray_entry(c, ray), q, jout, fetch);
}

__device__ __forceinline__ void start_ray(const MeshDev& m, const ItemCtx& c, uint32_t ray, LaneRay& lr)
{
    const GlobalTriangles fetch;
    const RayStart r = draw_ray(m, c, ray, lr.q, lr.j, fetch);
    walk_begin(m, lr.w, r.ax, r.ay, r.az, r.t, r.l, r.p0, r.p1, r.p2, c.sp.px, c.sp.py, r.ez, r.nr, r.fp, fetch);
}

//
// kTable: per-problem constants from the device table (batches of several
// problems); otherwise from the kernel argument `one` (a single problem), so
// they sit in scalar registers.  kLds (single problem whose records fit in
// LDS: kLdsCompact or kLdsIndexed, 0 for none): the work-group first copies
// the staged triangle records into LDS and the walks read them there; beta
// stays in global memory.  kCompactGlobal: the walks read the compact
// records of the item's problem from global memory (single problems too
// large for LDS, and the problem table of a batch).
template <bool kRefill, bool kTable, uint32_t kLds>
__global__ void __launch_bounds__(kQueueMaxBlock) hase_queue_trace_kernel(QueueState q, ProblemDev one)
{
    const uint32_t n_active = q.scalars[kNActive];
    if (n_active == 0u) return;
    const uint4* lds_rec = nullptr;
    if constexpr (kLds == kLdsCompact || kLds == kLdsIndexed) {
        __shared__ uint4 s_tri[kLdsStageWords];
        const uint32_t n = one.m.lds_words;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s_tri[i] = one.m.lds_rec[i];
        __syncthreads();
        lds_rec = s_tri;
    }
    const uint32_t total = q.item_off[n_active];
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long steps = 0, rays = 0;
    for (;;) {
        uint32_t first = 0;
        if (lane == 0) first = atomicAdd(q.scalars + kHead, q.queue_chunk);
        // wavefront-uniform from here on: item, slot and problem live in scalar registers
        first = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(first, 0, 64));
        if (first >= total) break;
        const uint32_t last = min(total, first + q.queue_chunk);
        for (uint32_t item = first; item < last; ++item) {
            // listed slot of the item: last a with item_off[a] <= item
            uint32_t lo = 0, hi = n_active;
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (q.item_off[mid] <= item) lo = mid; else hi = mid;
            }
            const uint32_t slot = q.list[lo];
            const ProblemDev& pd = kTable ? q.problems[q.slot_problem[slot]] : one;
            const MeshDev m = pd.m;
            const uint32_t local = slot - pd.slot0;
            const uint32_t nrays = q.nrays[slot];
            const uint32_t ir = q.item_rays[slot];
            ItemCtx c;
            c.r0 = (item - q.item_off[lo]) * ir;
            const uint32_t cnt = min(ir, nrays - c.r0);
            c.sample = pd.sample_of ? pd.sample_of[local] : pd.sample_base + local;
            c.launch = q.launch[slot];
            c.seed = pd.seed;
            c.sp = sample_point(pd.pts, m.np, c.sample);
            c.n = pd.n;
            const uint32_t own = pd.shared ? 0u : local;   // slot whose proposal the rays follow
            c.pre = pd.prefix + (size_t)own * pd.n;
            c.bk = pd.buckets + (size_t)own * kPrefixBuckets;
            c.Q = pd.Q[own];
            c.U = systematic_offset(pd.seed, c.sample, c.launch);
            c.scale = c.Q / (double)nrays;

            double s1 = 0.0, s2 = 0.0;
            uint32_t drops = 0;
            auto finish = [&](const WalkOut& w, double qprob, uint32_t j) {
                steps += w.steps;
                rays++;
                bool ok = w.status == 0 && w.length > 0.0f;
                if (ok) {
                    const float g = spectral_gain(pd.c1, pd.c2, pd.n_lambda, pd.inv_lambda, m.clad_alpha, w);
                    const float f = g * w.refl * m.bv[j] / (4.0f * kPi * w.length * w.length);
                    const double X = (double)f * (c.Q / qprob);
                    if (isfinite(X)) {
                        s1 += X;
                        s2 += X * X;
                    } else {
                        ok = false;
                    }
                }
                drops += ok ? 0u : 1u;
            };
            if (!kRefill) {
                // Rays in groups of 256: each lane finds the entries of four,
                // the wavefront sorts the group by estimated walk length and
                // walks it as four rounds of 64 rays of similar length (a
                // round lasts as long as its longest walk).
                auto run_group = [&](uint32_t g0, const auto& fetch) {
                    uint32_t key[4];
                    uint32_t ent[4];
#pragma unroll
                    for (uint32_t r = 0; r < 4u; ++r) {
                        const uint32_t li = r * 64u + (uint32_t)lane;
                        key[r] = 0xFFFFFF00u | li;   // past the item's end: sorted last, skipped
                        ent[r] = 0u;
                        if (g0 + li < cnt) {
                            const size_t e = ray_entry(c, c.r0 + g0 + li);
                            ent[r] = (uint32_t)e;
                            const float est = fminf(ray_steps_estimate(m, c, e, fetch) * 16.0f, 16777214.0f);
                            key[r] = ((uint32_t)est << 8) | li;
                        }
                    }
                    wave_sort256(key, lane);
                    // round r: the ray of sorted key r of this lane, whose
                    // entry lane li % 64 found in its register li / 64
                    uint32_t mine[4];
                    uint32_t rays4 = 0u;   // the four group indices, 8 bits each
#pragma unroll
                    for (uint32_t r = 0; r < 4u; ++r) {
                        const int src = (int)(key[r] & 63u);
                        const uint32_t e0 = (uint32_t)__shfl((int)ent[0], src, 64);
                        const uint32_t e1 = (uint32_t)__shfl((int)ent[1], src, 64);
                        const uint32_t e2 = (uint32_t)__shfl((int)ent[2], src, 64);
                        const uint32_t e3 = (uint32_t)__shfl((int)ent[3], src, 64);
                        const uint32_t sel = (key[r] >> 6) & 3u;
                        mine[r] = sel == 0u ? e0 : (sel == 1u ? e1 : (sel == 2u ? e2 : e3));
                        rays4 |= (key[r] & 0xFFu) << (8u * r);
                    }
                    // one copy of the walk: each round takes slot 0, then the
                    // slots move down (static register indices, no scratch)
#pragma unroll 1
                    for (uint32_t r = 0; r < 4u; ++r) {
                        const uint32_t li = rays4 & 0xFFu;
                        if (g0 + li < cnt) {
                            double qprob;
                            uint32_t j;
                            const RayStart s = draw_ray_at(m, c, c.r0 + g0 + li, (size_t)mine[0], qprob, j, fetch);
                            const WalkOut w = walk_ray(m, s.ax, s.ay, s.az, s.t, s.l, s.p0, s.p1, s.p2, c.sp.px,
                                                       c.sp.py, s.ez, s.nr, s.fp, fetch);
                            finish(w, qprob, j);
                        }
                        rays4 >>= 8;
                        mine[0] = mine[1]; mine[1] = mine[2]; mine[2] = mine[3];
                    }
                };
                for (uint32_t g0 = 0; g0 < cnt; g0 += 256u) {
                    if constexpr (kLds != 0u) {
                        // staged records, or the item's problem's compact records in global memory
                        run_group(g0, staged_fetch<kLds>(kLds == kCompactGlobal ? m.cmp_rec : lds_rec, m.T));
                    } else {
                        run_group(g0, GlobalTriangles{});
                    }
                }
            } else {
                LaneRay lr;
                bool have = (uint32_t)lane < cnt;
                if (have) start_ray(m, c, c.r0 + (uint32_t)lane, lr);
                uint32_t cursor = 64u;   // next ray of the item not yet started (wavefront-uniform)
                for (;;) {
                    if (have && !lr.w.done) {
                        if (lr.w.nrefl) walk_advance<true>(m, lr.w, q.budget);
                        else walk_advance<false>(m, lr.w, q.budget);
                    }
                    if (have && lr.w.done) {
                        finish(walk_result(lr.w), lr.q, lr.j);
                        have = false;
                    }
                    const unsigned long long idle = __ballot(!have);
                    const uint32_t nidle = (uint32_t)__popcll(idle);
                    if (cursor >= cnt) {
                        if (nidle == 64u) break;
                    } else if (nidle >= q.refill || nidle == 64u) {
                        if (!have) {
                            const uint32_t at = cursor + (uint32_t)__popcll(idle & below);
                            if (at < cnt) {
                                start_ray(m, c, c.r0 + at, lr);
                                have = true;
                            }
                        }
                        cursor += nidle;
                    }
                }
            }
            s1 = wave_sum(s1);
            s2 = wave_sum(s2);
            const unsigned long long d = wave_sum_u64((unsigned long long)drops);
            if (lane == 0) {
                q.partial[item] = make_double2(s1, s2);
                q.partial_drop[item] = (uint32_t)d;
            }
        }
    }
    const unsigned long long ws = wave_sum_u64(steps), wr = wave_sum_u64(rays);
    if (lane == 0 && wr) {
        atomicAdd(&q.counters[0], wr);
        atomicAdd(&q.counters[1], ws);
    }
}

__global__ void hase_queue_finalize_kernel(QueueState q)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= q.scalars[kNActive]) return;
    const uint32_t slot = q.list[a];
    const ProblemDev& pd = q.problems[q.slot_problem[slot]];
    double a1 = q.s1[slot], a2 = q.s2[slot];
    uint32_t drop = q.dropped[slot];
    for (uint32_t it = q.item_off[a]; it < q.item_off[a + 1]; ++it) {
        const double2 v = q.partial[it];
        a1 += v.x;
        a2 += v.y;
        drop += q.partial_drop[it];
    }
    const uint32_t N = q.N[slot] + q.nrays[slot];
    q.s1[slot] = a1;
    q.s2[slot] = a2;
    q.N[slot] = N;
    q.dropped[slot] = drop;
    const uint32_t launch = q.launch[slot] + 1u;
    q.launch[slot] = launch;
    if (q.target[slot] < pd.max_rays) q.k[slot]++;
    // relative standard error, src/core/calcForwardPhiAse.cpp:81-98; mse is its square
    const double nn = (double)N;
    double mse;
    if (N < 2u || !isfinite(a1) || !isfinite(a2)) mse = 1.79769313486231570e308;
    else if (a1 == 0.0) mse = __longlong_as_double(0x7ff8000000000000LL);
    else {
        const double rv = (nn * a2 / (a1 * a1) - 1.0) / nn;
        mse = rv > 0.0 ? rv : 0.0;
    }
    const double phi = pd.phi_scale * a1 / nn;
    q.phi[slot] = (float)phi;
    q.mse[slot] = mse;
    q.se[slot] = (mse >= 0.0 && mse < 1e300) ? sqrt(mse) * fabs(phi) : mse;
    const bool converged = drop == 0u && isfinite(mse) && mse <= pd.mse_threshold;
    if (converged) {
        q.flags[slot] = 0u;
    } else if (pd.defer && launch == pd.launch0 + 1u) {
        q.flags[slot] = kQueueDeferred;
    }
}

__global__ void hase_queue_reset_kernel(QueueState q, uint32_t slots)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < slots) {
        q.flags[s] = kQueueActive;
        q.k[s] = 0u;
        q.extra[s] = 0u;
        q.launch[s] = q.problems[q.slot_problem[s]].launch0;
        q.s1[s] = 0.0;
        q.s2[s] = 0.0;
        q.N[s] = 0u;
        q.dropped[s] = 0u;
    }
    if (s == 0) {
        q.scalars[kNActive] = 0u;
        q.scalars[kHead] = 0u;
        q.scalars[kSlots] = slots;
    }
}

}  // namespace

void check_tuning(const hase_tuning_t& t)
{
    if (t.mode != HASE_TRACE_QUEUE && t.mode != HASE_TRACE_ROUNDS) throw HaseError(HASE_E_INVALID, "unknown trace mode");
    if (t.block_size < 64 || t.block_size > (uint32_t)kQueueMaxBlock || t.block_size % 64)
        throw HaseError(HASE_E_INVALID, "block_size must be a multiple of 64 in [64, 1024]");
    if (t.stage_lds > 1) throw HaseError(HASE_E_INVALID, "stage_lds must be 0 or 1");
    if (t.rays_per_item < 64 || t.rays_per_item % 64 || t.rays_per_item > (1u << 20))
        throw HaseError(HASE_E_INVALID, "rays_per_item must be a multiple of 64 up to 2^20");
    if (t.queue_chunk < 1 || t.queue_chunk > 1024) throw HaseError(HASE_E_INVALID, "queue_chunk must be in [1, 1024]");
    if (t.refill_lanes > 64) throw HaseError(HASE_E_INVALID, "refill_lanes must be in [0, 64]");
    if (t.step_budget < 1 || t.step_budget > kMaxSteps + 1u)
        throw HaseError(HASE_E_INVALID, "step_budget must be in [1, 10001]");
}

uint32_t queue_max_rounds(const Schedule& s)
{
    // launches of a sample that never converges, from the same rules as the plan kernel
    uint32_t N = 0, k = 0, extra = 0, rounds = 0;
    for (;;) {
        const uint32_t tg = s.target(k);
        uint32_t nr = 0;
        if (tg > N) nr = tg - N;
        else if (N >= s.max_rays || tg >= s.max_rays) {
            if (extra + 1u >= s.reps) break;
            extra++;
            nr = s.max_rays;
        } else if (k >= s.steps || s.max_rays <= s.min_rays) break;
        else { k++; continue; }
        N += nr;
        rounds++;
        if (tg < s.max_rays) k++;
        if (rounds > 4096u) throw HaseError(HASE_E_INTERNAL, "adaptive schedule does not terminate");
    }
    return rounds;
}

size_t queue_items_per_slot(const hase_context* ctx, const hase_tuning_t& tu)
{
    const size_t need = (ctx->exp.max_rays_per_sample + tu.rays_per_item - 1) / tu.rays_per_item;
    return std::min<size_t>(need, kQueueItemsPerSlot);
}

void queue_targets(hase_context* ctx)
{
    const Schedule sc{ctx->exp.min_rays_per_sample, ctx->exp.max_rays_per_sample, ctx->comp.adaptive_steps,
                      ctx->comp.max_repetitions};
    const bool flat = sc.max_rays <= sc.min_rays || sc.steps == 0;
    std::vector<uint32_t> targets(flat ? 1u : sc.steps + 1u);
    for (uint32_t k = 0; k < targets.size(); ++k) targets[k] = sc.target(k);
    ctx->q_targets.upload(targets.data(), targets.size(), ctx->stream);
}

void queue_reserve(hase_context* ctx, uint32_t batch)
{
    ctx->q_flags.alloc(batch);
    ctx->q_k.alloc(batch);
    ctx->q_extra.alloc(batch);
    ctx->q_launch.alloc(batch);
    ctx->q_nrays.alloc(batch);
    ctx->q_item_rays.alloc(batch);
    ctx->q_target.alloc(batch);
    ctx->q_list.alloc(batch);
    ctx->q_item_off.alloc((size_t)batch + 1u);
    ctx->q_scalars.alloc(kQueueScalars);
    ctx->q_problems.alloc(1);
    if (ctx->q_slot_problem.n < batch) {
        ctx->q_slot_problem.alloc(batch);
        HIP_CHECK(hipMemsetAsync(ctx->q_slot_problem.p, 0, batch * sizeof(uint32_t), ctx->stream));
    }
    const size_t items = (size_t)batch * queue_items_per_slot(ctx, ctx->tuning);
    ctx->partial.alloc(std::max(ctx->partial.n, items));
    ctx->partial_drop.alloc(std::max(ctx->partial_drop.n, items));
    queue_targets(ctx);
}

ProblemDev queue_problem(const hase_context* ctx, uint32_t slot0, uint32_t base, bool mapped, uint32_t launch0,
                         bool defer)
{
    ProblemDev pd;
    std::memset(&pd, 0, sizeof(pd));
    const hase_experiment_t& e = ctx->exp;
    pd.m = mesh_view(ctx);
    pd.pts = ctx->pts.p;
    pd.c1 = ctx->c1.p;
    pd.c2 = ctx->c2.p;
    pd.prefix = ctx->prefix.p;
    pd.Q = ctx->Q.p;
    pd.buckets = ctx->buckets.p;
    pd.sample_of = mapped ? ctx->slot_sample.p : nullptr;
    pd.targets = ctx->q_targets.p;
    pd.n = (size_t)slices(ctx) * ctx->P;
    pd.mse_threshold = e.mse_threshold;
    pd.phi_scale = ctx->n_tot / ctx->tau;
    pd.n_lambda = e.number_of_wavelengths;
    pd.R = slices(ctx);
    pd.seed = ctx->comp.rng_seed;
    pd.slot0 = slot0;
    pd.sample_base = base;
    pd.n_targets = (uint32_t)ctx->q_targets.n;
    pd.max_rays = e.max_rays_per_sample;
    pd.reps = ctx->comp.max_repetitions;
    pd.launch0 = launch0;
    pd.defer = defer ? 1u : 0u;
    pd.shared = ctx->proposal_kind == HASE_PROPOSAL_BETA_VOLUME ? 1u : 0u;
    pd.inv_lambda = 1.0f / (float)e.number_of_wavelengths;
    return pd;
}

float queue_run(hipStream_t stream, int device, const hase_tuning_t& tu, const QueueState& qs, uint32_t slots,
                uint32_t rounds, GraphCache& gc, uint64_t generation, const ProblemDev* single, bool table_compact)
{
    int cus = 0, per = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const bool refill = tu.refill_lanes != 0u;
    const bool table = single == nullptr;
    // LDS staging: single problem, no lane refill, mesh small enough
    const uint32_t lds = tu.stage_lds && !table && !refill && single->m.lds_rec != nullptr ? single->m.lds_kind : 0u;
    auto kernel = lds == kLdsIndexed ? hase_queue_trace_kernel<false, false, kLdsIndexed>
                : lds == kLdsCompact ? hase_queue_trace_kernel<false, false, kLdsCompact>
                : lds == kCompactGlobal ? hase_queue_trace_kernel<false, false, kCompactGlobal>
                : table && table_compact && tu.stage_lds && !refill ? hase_queue_trace_kernel<false, true, kCompactGlobal>
                : refill ? (table ? hase_queue_trace_kernel<true, true, 0u> : hase_queue_trace_kernel<true, false, 0u>)
                         : (table ? hase_queue_trace_kernel<false, true, 0u> : hase_queue_trace_kernel<false, false, 0u>);
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, (int)tu.block_size, 0));
    ProblemDev one;
    std::memset(&one, 0, sizeof(one));
    if (single) one = *single;
    const int grid = std::max(1, cus) * std::max(1, per);
    const uint32_t fin_blocks = std::max(1u, (slots + 255u) / 256u);
    // allocation-free calls (an arena installed) create no events and report no time
    const bool timed = g_arena == nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timed) {
        HIP_CHECK(hipEventCreate(&t0));
        HIP_CHECK(hipEventCreate(&t1));
    }
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } guard{t0, t1};
    hipLaunchKernelGGL(hase_queue_reset_kernel, dim3((slots + 255u) / 256u), dim3(256), 0, stream, qs, slots);
    HIP_CHECK(hipGetLastError());
    auto launch_round = [&]() {
        hipLaunchKernelGGL(hase_queue_plan_kernel, dim3(1), dim3(kPlanBlock), 0, stream, qs);
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(tu.block_size), 0, stream, qs, one);
        hipLaunchKernelGGL(hase_queue_finalize_kernel, dim3(fin_blocks), dim3(256), 0, stream, qs);
    };
    if (timed) HIP_CHECK(hipEventRecord(t0, stream));
    if (tu.use_graph && timed) {
        // kernel arguments are fixed within a generation, up to the slot count and the
        // single problem's constants (its sample base changes from batch to batch)
        uint64_t key = 1469598103934665603ull;
        auto mix = [&key](const void* data, size_t n) {
            const unsigned char* b = (const unsigned char*)data;
            for (size_t i = 0; i < n; ++i) {
                key ^= b[i];
                key *= 1099511628211ull;
            }
        };
        mix(&generation, sizeof(generation));
        mix(&slots, sizeof(slots));
        mix(&one, sizeof(one));
        if (!gc.exec || gc.key != key) {
            gc.reset();
            HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            launch_round();
            hipGraph_t g = nullptr;
            HIP_CHECK(hipStreamEndCapture(stream, &g));
            gc.graph = g;
            HIP_CHECK(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0));
            gc.key = key;
        }
        for (uint32_t r = 0; r < rounds; ++r) HIP_CHECK(hipGraphLaunch(gc.exec, stream));
    } else {
        for (uint32_t r = 0; r < rounds; ++r) {
            launch_round();
            HIP_CHECK(hipGetLastError());
        }
    }
    float ms = 0.0f;
    if (!timed) {
        HIP_CHECK(hipStreamSynchronize(stream));
        return ms;
    }
    HIP_CHECK(hipEventRecord(t1, stream));
    HIP_CHECK(hipEventSynchronize(t1));
    HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
    return ms;
}

QueueState queue_state(hase_context* ctx)
{
    QueueState qs;
    std::memset(&qs, 0, sizeof(qs));
    qs.problems = ctx->q_problems.p;
    qs.slot_problem = ctx->q_slot_problem.p;
    qs.flags = ctx->q_flags.p;
    qs.k = ctx->q_k.p;
    qs.extra = ctx->q_extra.p;
    qs.launch = ctx->q_launch.p;
    qs.nrays = ctx->q_nrays.p;
    qs.item_rays = ctx->q_item_rays.p;
    qs.target = ctx->q_target.p;
    qs.list = ctx->q_list.p;
    qs.item_off = ctx->q_item_off.p;
    qs.scalars = ctx->q_scalars.p;
    qs.s1 = ctx->s1.p;
    qs.s2 = ctx->s2.p;
    qs.N = ctx->N.p;
    qs.dropped = ctx->dropped.p;
    qs.phi = ctx->phi.p;
    qs.mse = ctx->mse.p;
    qs.se = ctx->se.p;
    qs.partial = ctx->partial.p;
    qs.partial_drop = ctx->partial_drop.p;
    qs.counters = ctx->counters.p;
    qs.rays_per_item = ctx->tuning.rays_per_item;
    qs.queue_chunk = ctx->tuning.queue_chunk;
    qs.refill = ctx->tuning.refill_lanes;
    qs.budget = ctx->tuning.step_budget;
    return qs;
}

// Adaptive rounds of one resident batch of a single context whose proposal
// prefix and Q are in place: slots [0, slots) hold samples base + slot, or
// ctx->slot_sample when mapped.  Leaves phi / mse / se / N / dropped /
// q_flags per slot on the device; returns the device time of the rounds.
float queue_rounds(hase_context* ctx, uint32_t base, bool mapped, uint32_t slots, uint32_t launch0, bool defer)
{
    const ProblemDev pd = queue_problem(ctx, 0u, base, mapped, launch0, defer);
    HIP_CHECK(hipMemcpyAsync(ctx->q_problems.p, &pd, sizeof(pd), hipMemcpyHostToDevice, ctx->stream));
    const Schedule sc{ctx->exp.min_rays_per_sample, ctx->exp.max_rays_per_sample, ctx->comp.adaptive_steps,
                      ctx->comp.max_repetitions};
    const float ms = queue_run(ctx->stream, ctx->device, ctx->tuning, queue_state(ctx), slots, queue_max_rounds(sc),
                               ctx->q_graph, ctx->q_generation, &pd);
    ctx->stats.trace_launches += queue_max_rounds(sc);
    return ms;
}

}  // namespace hase
