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

namespace hase {
namespace {

constexpr uint32_t kActive = 1u, kDeferred = 2u;
constexpr int kPlanBlock = 1024;
constexpr int kQueueMaxBlock = 1024;

// device scalars: [0] listed slots, [1] queue cursor, [2] slots, [3] sample
// base, [4] slot->sample map in use, [5] first launch, [6] defer misses
enum { kNActive = 0, kHead = 1, kSlots = 2, kBase = 3, kMapped = 4, kLaunch0 = 5, kDefer = 6, kScalars = 8 };

struct PlanParams {
    uint32_t* flags;
    uint32_t* k;
    uint32_t* extra;
    const uint32_t* N;
    uint32_t* nrays;
    uint32_t* target;
    uint32_t* list;
    uint32_t* item_off;
    uint32_t* scalars;
    const uint32_t* targets;   // schedule target of increase k, k < n_targets
    uint32_t n_targets, max_rays, reps, rays_per_item;
};

__global__ void __launch_bounds__(kPlanBlock) hase_queue_plan_kernel(PlanParams p)
{
    __shared__ uint32_t part_items[kPlanBlock], part_slots[kPlanBlock];
    const uint32_t slots = p.scalars[kSlots];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (slots + kPlanBlock - 1) / kPlanBlock;
    const uint32_t s0 = min(slots, tid * per), s1 = min(slots, s0 + per);
    uint32_t my_items = 0, my_slots = 0;
    for (uint32_t s = s0; s < s1; ++s) {
        uint32_t nr = 0;
        if (p.flags[s] & kActive) {
            uint32_t kk = p.k[s], ex = p.extra[s], tg = 0;
            const uint32_t n = p.N[s];
            for (;;) {
                tg = p.targets[min(kk, p.n_targets - 1u)];
                if (tg > n) { nr = tg - n; break; }
                if (n >= p.max_rays || tg >= p.max_rays) {
                    if (ex + 1u >= p.reps) break;
                    ex++;
                    nr = p.max_rays;
                    break;
                }
                if (kk + 1u >= p.n_targets) break;   // flat schedule: nothing more to add
                kk++;
            }
            p.k[s] = kk;
            p.extra[s] = ex;
            p.target[s] = tg;
            if (nr == 0u) p.flags[s] &= ~kActive;
        }
        p.nrays[s] = nr;
        if (nr) {
            my_items += (nr + p.rays_per_item - 1u) / p.rays_per_item;
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
        const uint32_t nr = p.nrays[s];
        if (!nr) continue;
        p.list[at] = s;
        p.item_off[at] = item;
        item += (nr + p.rays_per_item - 1u) / p.rays_per_item;
        at++;
    }
    if (tid == kPlanBlock - 1) {
        p.item_off[part_slots[tid]] = part_items[tid];
        p.scalars[kNActive] = part_slots[tid];
        p.scalars[kHead] = 0u;
    }
}

struct QueueParams {
    MeshDev m;
    const float2* pts;
    const uint32_t* sample_of;
    const uint32_t* scalars;
    uint32_t R, n_lambda;
    const float* c1;
    const float* c2;
    float inv_lambda;
    uint32_t seed;
    const uint32_t* list;
    const uint32_t* item_off;
    const uint32_t* nrays;
    const uint32_t* launch;
    uint32_t* head;
    uint32_t rays_per_item, queue_chunk, refill, budget;
    const double* prefix;
    const double* Q;
    double2* partial;
    uint32_t* partial_drop;
    unsigned long long* counters;
};

// Per-lane ray of a queue item: walk state plus what its weight needs.
struct LaneRay {
    WalkState w;
    double q;        // proposal probability of the drawn (slice, prism)
    uint32_t j;      // prism (device order)
};

struct ItemCtx {
    SampleGeom sp;
    const double* pre;
    size_t n;
    double Q, U, scale;
    uint32_t sample, launch, r0;
};

__device__ __forceinline__ void start_ray(const QueueParams& p, const ItemCtx& c, uint32_t ray, LaneRay& lr)
{
    const MeshDev& m = p.m;
    const uint32_t T = m.T, P = T * m.layers;
    const double target = ((double)ray + c.U) * c.scale;
    // first entry whose prefix exceeds the target
    size_t lo = 0, hi = c.n;
    while (lo < hi) {
        const size_t mid = (lo + hi) >> 1;
        if (c.pre[mid] > target) hi = mid; else lo = mid + 1;
    }
    const size_t e = lo < c.n ? lo : c.n - 1;
    lr.q = c.pre[e] - (e ? c.pre[e - 1] : 0.0);
    const uint32_t k = (uint32_t)(e / P), j = (uint32_t)(e % P);
    const uint32_t t = j % T, l = j / T;
    lr.j = j;
    uint32_t rnd[4];
    philox4x32_10(ray, c.sample, c.launch, kTagRay, p.seed, kKeyTag, rnd);
    float u1 = u32_to_unit(rnd[0]), u2 = u32_to_unit(rnd[1]), u3 = u32_to_unit(rnd[2]);
    if (u1 + u2 > 1.0f) { u1 = 1.0f - u1; u2 = 1.0f - u2; }
    const uint4 q0 = m.rec[3u * t], q1 = m.rec[3u * t + 1u];
    const float x0 = bits_to_float(q0.x), y0 = bits_to_float(q0.y);
    const float x1 = bits_to_float(q0.z), y1 = bits_to_float(q0.w);
    const float x2 = bits_to_float(q1.x), y2 = bits_to_float(q1.y);
    const float ax = x0 + u1 * (x1 - x0) + u2 * (x2 - x0);
    const float ay = y0 + u1 * (y1 - y0) + u2 * (y2 - y0);
    const float az = (float)l + u3;   // layer units
    const uint32_t nr = (k + 1u) >> 1, fp = k & 1u;
    const float ez = (float)image_plane(c.sp.plane, (int)m.layers, nr, fp);
    walk_begin(lr.w, m.h, ax, ay, az, t, l, c.sp.px, c.sp.py, ez, nr, fp);
}

__global__ void __launch_bounds__(kQueueMaxBlock) hase_queue_trace_kernel(QueueParams p)
{
    const uint32_t n_active = p.scalars[kNActive];
    if (n_active == 0u) return;
    const uint32_t total = p.item_off[n_active];
    const uint32_t base_sample = p.scalars[kBase];
    const bool mapped = p.scalars[kMapped] != 0u;
    const int lane = threadIdx.x & 63;
    const MeshDev& m = p.m;
    const size_t n = (size_t)p.R * m.T * m.layers;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long steps = 0, rays = 0;
    for (;;) {
        uint32_t first = 0;
        if (lane == 0) first = atomicAdd(p.head, p.queue_chunk);
        first = __shfl(first, 0, 64);
        if (first >= total) break;
        const uint32_t last = min(total, first + p.queue_chunk);
        for (uint32_t item = first; item < last; ++item) {
            // listed slot of the item: last a with item_off[a] <= item
            uint32_t lo = 0, hi = n_active;
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (p.item_off[mid] <= item) lo = mid; else hi = mid;
            }
            const uint32_t slot = p.list[lo];
            const uint32_t nrays = p.nrays[slot];
            ItemCtx c;
            c.r0 = (item - p.item_off[lo]) * p.rays_per_item;
            const uint32_t cnt = min(p.rays_per_item, nrays - c.r0);
            c.sample = mapped ? p.sample_of[slot] : base_sample + slot;
            c.launch = p.launch[slot];
            c.sp = sample_point(p.pts, m.np, c.sample);
            c.pre = p.prefix + (size_t)slot * n;
            c.n = n;
            c.Q = p.Q[slot];
            c.U = systematic_offset(p.seed, c.sample, c.launch);
            c.scale = c.Q / (double)nrays;

            double s1 = 0.0, s2 = 0.0;
            uint32_t drops = 0;
            LaneRay lr;
            bool have = (uint32_t)lane < cnt;
            if (have) start_ray(p, c, c.r0 + (uint32_t)lane, lr);
            uint32_t cursor = 64u;   // next ray of the item not yet started (wavefront-uniform)
            for (;;) {
                if (have && !lr.w.done) {
                    if (lr.w.nrefl) walk_advance<true>(m, lr.w, p.budget);
                    else walk_advance<false>(m, lr.w, p.budget);
                }
                if (have && lr.w.done) {
                    const WalkOut w = walk_result(lr.w);
                    steps += w.steps;
                    rays++;
                    bool ok = w.status == 0 && w.length > 0.0f;
                    if (ok) {
                        const float g = spectral_gain(p.c1, p.c2, p.n_lambda, p.inv_lambda, m.clad_alpha, w);
                        const float f = g * w.refl * m.bv[lr.j] / (4.0f * kPi * w.length * w.length);
                        const double X = (double)f * (c.Q / lr.q);
                        if (isfinite(X)) {
                            s1 += X;
                            s2 += X * X;
                        } else {
                            ok = false;
                        }
                    }
                    drops += ok ? 0u : 1u;
                    have = false;
                }
                const unsigned long long idle = __ballot(!have);
                const uint32_t nidle = (uint32_t)__popcll(idle);
                if (cursor >= cnt) {
                    if (nidle == 64u) break;
                } else if (nidle >= p.refill || nidle == 64u) {
                    if (!have) {
                        const uint32_t at = cursor + (uint32_t)__popcll(idle & below);
                        if (at < cnt) {
                            start_ray(p, c, c.r0 + at, lr);
                            have = true;
                        }
                    }
                    cursor += nidle;
                }
            }
            s1 = wave_sum(s1);
            s2 = wave_sum(s2);
            const unsigned long long d = wave_sum_u64((unsigned long long)drops);
            if (lane == 0) {
                p.partial[item] = make_double2(s1, s2);
                p.partial_drop[item] = (uint32_t)d;
            }
        }
    }
    const unsigned long long ws = wave_sum_u64(steps), wr = wave_sum_u64(rays);
    if (lane == 0 && wr) {
        atomicAdd(&p.counters[0], wr);
        atomicAdd(&p.counters[1], ws);
    }
}

struct QueueFinalizeParams {
    const uint32_t* scalars;
    const uint32_t* list;
    const uint32_t* item_off;
    const uint32_t* nrays;
    const uint32_t* target;
    uint32_t* flags;
    uint32_t* k;
    uint32_t* launch;
    const double2* partial;
    const uint32_t* partial_drop;
    double* s1;
    double* s2;
    uint32_t* N;
    uint32_t* dropped;
    uint32_t max_rays;
    double mse_threshold, phi_scale;
    float* phi;
    double* mse;
    double* se;
};

__global__ void hase_queue_finalize_kernel(QueueFinalizeParams p)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= p.scalars[kNActive]) return;
    const uint32_t slot = p.list[a];
    double a1 = p.s1[slot], a2 = p.s2[slot];
    uint32_t drop = p.dropped[slot];
    for (uint32_t it = p.item_off[a]; it < p.item_off[a + 1]; ++it) {
        const double2 v = p.partial[it];
        a1 += v.x;
        a2 += v.y;
        drop += p.partial_drop[it];
    }
    const uint32_t N = p.N[slot] + p.nrays[slot];
    p.s1[slot] = a1;
    p.s2[slot] = a2;
    p.N[slot] = N;
    p.dropped[slot] = drop;
    const uint32_t launch = p.launch[slot] + 1u;
    p.launch[slot] = launch;
    if (p.target[slot] < p.max_rays) p.k[slot]++;
    // relative standard error, src/core/calcForwardPhiAse.cpp:81-98; mse is its square
    const double nn = (double)N;
    double mse;
    if (N < 2u || !isfinite(a1) || !isfinite(a2)) mse = 1.79769313486231570e308;
    else if (a1 == 0.0) mse = __longlong_as_double(0x7ff8000000000000LL);
    else {
        const double rv = (nn * a2 / (a1 * a1) - 1.0) / nn;
        mse = rv > 0.0 ? rv : 0.0;
    }
    const double phi = p.phi_scale * a1 / nn;
    p.phi[slot] = (float)phi;
    p.mse[slot] = mse;
    p.se[slot] = (mse >= 0.0 && mse < 1e300) ? sqrt(mse) * fabs(phi) : mse;
    const bool converged = drop == 0u && isfinite(mse) && mse <= p.mse_threshold;
    if (converged) {
        p.flags[slot] = 0u;
    } else if (p.scalars[kDefer] && launch == p.scalars[kLaunch0] + 1u) {
        p.flags[slot] = kDeferred;
    }
}

__global__ void hase_queue_reset_kernel(uint32_t slots, uint32_t base, uint32_t mapped, uint32_t launch0,
                                        uint32_t defer, uint32_t* flags, uint32_t* k, uint32_t* extra,
                                        uint32_t* launch, double* s1, double* s2, uint32_t* N, uint32_t* dropped,
                                        uint32_t* scalars)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < slots) {
        flags[s] = kActive;
        k[s] = 0u;
        extra[s] = 0u;
        launch[s] = launch0;
        s1[s] = 0.0;
        s2[s] = 0.0;
        N[s] = 0u;
        dropped[s] = 0u;
    }
    if (s == 0) {
        scalars[kNActive] = 0u;
        scalars[kHead] = 0u;
        scalars[kSlots] = slots;
        scalars[kBase] = base;
        scalars[kMapped] = mapped;
        scalars[kLaunch0] = launch0;
        scalars[kDefer] = defer;
    }
}

int queue_grid(const hase_context* ctx)
{
    int cus = 0, per = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, hase_queue_trace_kernel,
                                                           (int)ctx->tuning.block_size, 0));
    return std::max(1, cus) * std::max(1, per);
}

}  // namespace

void check_tuning(const hase_tuning_t& t)
{
    if (t.mode != HASE_TRACE_QUEUE && t.mode != HASE_TRACE_ROUNDS) throw HaseError(HASE_E_INVALID, "unknown trace mode");
    if (t.block_size < 64 || t.block_size > (uint32_t)kQueueMaxBlock || t.block_size % 64)
        throw HaseError(HASE_E_INVALID, "block_size must be a multiple of 64 in [64, 1024]");
    if (t.rays_per_item < 64 || t.rays_per_item % 64 || t.rays_per_item > (1u << 20))
        throw HaseError(HASE_E_INVALID, "rays_per_item must be a multiple of 64 up to 2^20");
    if (t.queue_chunk < 1 || t.queue_chunk > 1024) throw HaseError(HASE_E_INVALID, "queue_chunk must be in [1, 1024]");
    if (t.refill_lanes < 1 || t.refill_lanes > 64) throw HaseError(HASE_E_INVALID, "refill_lanes must be in [1, 64]");
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

void queue_reserve(hase_context* ctx, uint32_t batch)
{
    ctx->q_flags.alloc(batch);
    ctx->q_k.alloc(batch);
    ctx->q_extra.alloc(batch);
    ctx->q_launch.alloc(batch);
    ctx->q_nrays.alloc(batch);
    ctx->q_target.alloc(batch);
    ctx->q_list.alloc(batch);
    ctx->q_item_off.alloc((size_t)batch + 1u);
    ctx->q_scalars.alloc(kScalars);
    const uint32_t per_slot = (ctx->exp.max_rays_per_sample + ctx->tuning.rays_per_item - 1) / ctx->tuning.rays_per_item;
    const size_t items = (size_t)batch * per_slot;
    ctx->partial.alloc(std::max(ctx->partial.n, items));
    ctx->partial_drop.alloc(std::max(ctx->partial_drop.n, items));
    ctx->q_items = items;
    const Schedule sc{ctx->exp.min_rays_per_sample, ctx->exp.max_rays_per_sample, ctx->comp.adaptive_steps,
                      ctx->comp.max_repetitions};
    const bool flat = sc.max_rays <= sc.min_rays || sc.steps == 0;
    std::vector<uint32_t> targets(flat ? 1u : sc.steps + 1u);
    for (uint32_t k = 0; k < targets.size(); ++k) targets[k] = sc.target(k);
    ctx->q_targets.upload(targets.data(), targets.size(), ctx->stream);
}

// Adaptive rounds of one resident batch whose proposal prefix and Q are in
// place: slots [0, slots) hold samples base + slot, or ctx->slot_sample when
// mapped.  Leaves phi / mse / se / N / dropped / q_flags per slot on the
// device; returns the device time of the rounds.
float queue_rounds(hase_context* ctx, uint32_t base, bool mapped, uint32_t slots, uint32_t launch0, bool defer)
{
    const hase_tuning_t& tu = ctx->tuning;
    const hase_experiment_t& e = ctx->exp;
    const uint32_t n_targets = (uint32_t)ctx->q_targets.n;
    hipLaunchKernelGGL(hase_queue_reset_kernel, dim3((slots + 255u) / 256u), dim3(256), 0, ctx->stream, slots, base,
                       mapped ? 1u : 0u, launch0, defer ? 1u : 0u, ctx->q_flags.p, ctx->q_k.p, ctx->q_extra.p,
                       ctx->q_launch.p, ctx->s1.p, ctx->s2.p, ctx->N.p, ctx->dropped.p, ctx->q_scalars.p);
    HIP_CHECK(hipGetLastError());

    PlanParams pp;
    pp.flags = ctx->q_flags.p;
    pp.k = ctx->q_k.p;
    pp.extra = ctx->q_extra.p;
    pp.N = ctx->N.p;
    pp.nrays = ctx->q_nrays.p;
    pp.target = ctx->q_target.p;
    pp.list = ctx->q_list.p;
    pp.item_off = ctx->q_item_off.p;
    pp.scalars = ctx->q_scalars.p;
    pp.targets = ctx->q_targets.p;
    pp.n_targets = n_targets;
    pp.max_rays = e.max_rays_per_sample;
    pp.reps = ctx->comp.max_repetitions;
    pp.rays_per_item = tu.rays_per_item;

    QueueParams qp;
    qp.m = mesh_view(ctx);
    qp.pts = ctx->pts.p;
    qp.sample_of = ctx->slot_sample.p;
    qp.scalars = ctx->q_scalars.p;
    qp.R = slices(ctx);
    qp.n_lambda = e.number_of_wavelengths;
    qp.c1 = ctx->c1.p;
    qp.c2 = ctx->c2.p;
    qp.inv_lambda = 1.0f / (float)e.number_of_wavelengths;
    qp.seed = ctx->comp.rng_seed;
    qp.list = ctx->q_list.p;
    qp.item_off = ctx->q_item_off.p;
    qp.nrays = ctx->q_nrays.p;
    qp.launch = ctx->q_launch.p;
    qp.head = ctx->q_scalars.p + kHead;
    qp.rays_per_item = tu.rays_per_item;
    qp.queue_chunk = tu.queue_chunk;
    qp.refill = tu.refill_lanes;
    qp.budget = tu.step_budget;
    qp.prefix = ctx->prefix.p;
    qp.Q = ctx->Q.p;
    qp.partial = ctx->partial.p;
    qp.partial_drop = ctx->partial_drop.p;
    qp.counters = ctx->counters.p;

    QueueFinalizeParams fp;
    fp.scalars = ctx->q_scalars.p;
    fp.list = ctx->q_list.p;
    fp.item_off = ctx->q_item_off.p;
    fp.nrays = ctx->q_nrays.p;
    fp.target = ctx->q_target.p;
    fp.flags = ctx->q_flags.p;
    fp.k = ctx->q_k.p;
    fp.launch = ctx->q_launch.p;
    fp.partial = ctx->partial.p;
    fp.partial_drop = ctx->partial_drop.p;
    fp.s1 = ctx->s1.p;
    fp.s2 = ctx->s2.p;
    fp.N = ctx->N.p;
    fp.dropped = ctx->dropped.p;
    fp.max_rays = e.max_rays_per_sample;
    fp.mse_threshold = e.mse_threshold;
    fp.phi_scale = ctx->n_tot / ctx->tau;
    fp.phi = ctx->phi.p;
    fp.mse = ctx->mse.p;
    fp.se = ctx->se.p;

    const int grid = queue_grid(ctx);
    const uint32_t fin_blocks = (uint32_t)((ctx->q_list.n + 255) / 256);
    auto launch_round = [&]() {
        hipLaunchKernelGGL(hase_queue_plan_kernel, dim3(1), dim3(kPlanBlock), 0, ctx->stream, pp);
        hipLaunchKernelGGL(hase_queue_trace_kernel, dim3(grid), dim3(tu.block_size), 0, ctx->stream, qp);
        hipLaunchKernelGGL(hase_queue_finalize_kernel, dim3(std::max(1u, fin_blocks)), dim3(256), 0, ctx->stream, fp);
    };
    const Schedule sc{e.min_rays_per_sample, e.max_rays_per_sample, ctx->comp.adaptive_steps,
                      ctx->comp.max_repetitions};
    const uint32_t rounds = queue_max_rounds(sc);

    hipEvent_t t0 = record(ctx);
    if (tu.use_graph) {
        // kernel arguments are fixed within a generation (one evaluation or probe)
        if (!ctx->q_exec || ctx->q_key != ctx->q_generation) {
            if (ctx->q_exec) (void)hipGraphExecDestroy(ctx->q_exec);
            if (ctx->q_graph) (void)hipGraphDestroy(ctx->q_graph);
            ctx->q_exec = nullptr;
            ctx->q_graph = nullptr;
            HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            launch_round();
            hipGraph_t g = nullptr;
            const hipError_t cap = hipStreamEndCapture(ctx->stream, &g);
            HIP_CHECK(cap);
            ctx->q_graph = g;
            HIP_CHECK(hipGraphInstantiate(&ctx->q_exec, ctx->q_graph, nullptr, nullptr, 0));
            ctx->q_key = ctx->q_generation;
        }
        for (uint32_t r = 0; r < rounds; ++r) HIP_CHECK(hipGraphLaunch(ctx->q_exec, ctx->stream));
    } else {
        for (uint32_t r = 0; r < rounds; ++r) {
            launch_round();
            HIP_CHECK(hipGetLastError());
        }
    }
    hipEvent_t t1 = record(ctx);
    HIP_CHECK(hipEventSynchronize(t1));
    ctx->stats.trace_launches += rounds;
    return elapsed(t0, t1);
}

}  // namespace hase
