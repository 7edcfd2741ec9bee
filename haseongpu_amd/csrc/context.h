// context.h - the persistent device context behind hase_context_t.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/hase_ase.h"
#include "common.h"
#include "prism_walk.h"
#include "queue.h"

struct hase_context {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t np = 0, T = 0, L = 0, P = 0, S = 0;
    double h = 0.0;
    double n_tot = 0.0, tau = 0.0;
    double clad_alpha = 0.0;
    bool has_refl = false, has_nidx = false;
    float nidx[4] = {1, 1, 1, 1};
    std::vector<double> tri_area;      // [T] device order
    double steps_per_cm = 0.0;         // 1.2 sqrt(T / mesh area): triangle crossings per cm, estimated
    std::vector<uint32_t> clad_flag;   // [T] device order
    std::vector<uint32_t> dev2user;    // device triangle -> caller triangle
    std::vector<uint32_t> user2dev;    // caller triangle -> device triangle
    std::vector<double> beta_host;     // [P]
    std::vector<double> beta_cells;    // [S]
    std::vector<uint32_t> tri_vertex;  // [3T] vertex k of device triangle t at 3t + k
    std::vector<float2> pts_host;      // [Np] point coordinates as stored on the device
    double Bsum = 0.0;
    hase_experiment_t exp{};
    std::vector<double> sigma_a, sigma_e;
    hase_compute_t comp{};
    hase::DevBuf<uint4> rec, lds_rec;
    hase::DevBuf<uint4> cmp_rec;       // compact records beside indexed ones (kind kLdsIndexed)
    uint32_t lds_kind = 0;             // format of lds_rec: 0 (none), kLdsCompact, kLdsIndexed or kCompactGlobal
    bool lds_staged() const { return lds_kind == hase::kLdsCompact || lds_kind == hase::kLdsIndexed; }
    hase::DevBuf<float2> pts;
    hase::DevBuf<float> beta, beta_col, bv, refl, c1, c2;
    hase::DevBuf<float> imp;
    hase::DevBuf<double> prefix, Q, s1, s2, bcells;
    hase::DevBuf<uint32_t> buckets;    // [batch * kPrefixBuckets] prefix buckets per resident sample
    hase::DevBuf<double2> partial;
    hase::DevBuf<uint32_t> partial_drop, N, dropped, active_list, active_next, active_count, slot_sample;
    hase::DevBuf<int32_t> active_flag;
    hase::DevBuf<float> phi;
    hase::DevBuf<double> mse, se;
    hase::DevBuf<unsigned long long> counters;
    hase_stats_t stats{};
    std::vector<hipEvent_t> events;

    // learned proposal
    int proposal_kind = HASE_PROPOSAL_PLAIN;
    bool proposal_fallback = true;
    uint32_t cluster_t = 1, cluster_l = 1;   // triangles / layers per cluster
    uint32_t TG = 0, LG = 0, C = 0;          // triangle groups, layer groups, clusters per slice
    float in_shift[HASE_MLP_INPUTS] = {}, in_scale[HASE_MLP_INPUTS] = {};
    float log_sigma = 0.0f;                  // log(max_sigma_e / max_sigma_a)
    std::vector<float2> centroid_host;       // [T] triangle centroids, device order
    hase::DevBuf<float2> centroid;
    hase::DevBuf<uint4> mlp_frags;           // BF16 operand fragments
    hase::DevBuf<float> mlp_vec;             // FP32 biases and output weights
    std::vector<float4> clusters_host;       // [C] (x, y, level, beta) representatives
    hase::DevBuf<float4> clusters;
    hase::DevBuf<uint32_t> cl_tri, cl_lev;          // cluster column per triangle, cluster row start per layer
    hase::DevBuf<float> logits;              // [slots * R * C]
    hase_proposal_stats_t pstats{};

    // persistent work queue
    hase_tuning_t tuning{HASE_TRACE_QUEUE, 512u, 256u, 1u, 0u, 16u, 1u, 0u, 0.0};
    hase::DevBuf<uint32_t> q_flags, q_k, q_extra, q_launch, q_nrays, q_item_rays, q_target;   // [slots]
    hase::DevBuf<uint32_t> q_list, q_item_off;                                  // [slots], [slots + 1]
    hase::DevBuf<uint32_t> q_scalars;                                           // n_active, head, slots
    hase::DevBuf<uint32_t> q_targets;                                           // schedule targets by step
    hase::DevBuf<hase::ProblemDev> q_problems;                                  // this context as a queue problem
    hase::DevBuf<uint32_t> q_slot_problem;                                      // zeros
    hase::GraphCache q_graph;
    uint64_t q_generation = 1;                                                  // bumped when launch inputs may change
    size_t batch_budget = 0;                                                    // bytes of batch buffers, fixed at first use

    // device copy of the flux of every sample, kept when phi_keep is set
    // (the time-stepped loop reads it without a host round trip)
    bool phi_keep = false;
    hase::DevBuf<float> phi_all;                                                // [S]
    hase::DevBuf<double> tri_area_dev;                                          // [T] device order
    hase::DevBuf<uint32_t> clad_dev;                                            // [T] device order
    hase::DevBuf<uint32_t> user2dev_dev;                                        // [T] caller -> device triangle

    ~hase_context()
    {
        q_graph.reset();
        for (auto e : events) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};


namespace hase {

// Adaptive ray schedule (include/core/types.hpp:333-354): cumulative target
// of increase k, min_rays * (max_rays / min_rays)^(k / steps) rounded up.
struct Schedule {
    uint32_t min_rays, max_rays, steps, reps;
    uint32_t target(uint32_t k) const
    {
        if (max_rays <= min_rays || steps == 0) return min_rays;
        if (k >= steps) return max_rays;
        const double growth = std::pow((double)max_rays / (double)min_rays, 1.0 / (double)steps);
        double t = std::ceil((double)min_rays * std::pow(growth, (double)k));
        t = std::min(std::max(t, (double)min_rays), (double)max_rays);
        return (uint32_t)t;
    }
};

MeshDev mesh_view(const hase_context* ctx);
uint32_t slices(const hase_context* ctx);

// events recorded on the context stream (destroyed at the next evaluation)
hipEvent_t record(hase_context* ctx);
float elapsed(hipEvent_t a, hipEvent_t b);

// evaluation of the context's sample range into host results (hase_ase.hip)
void evaluate_context(hase_context* ctx, hase_result_t* res);
// one-shot evaluation: a context built for the call, evaluated and destroyed
// (its device buffers come from the calling thread's arena when one is
// installed)
void calc_phi_ase(const hase_mesh_t* mesh, const hase_experiment_t* experiment, const hase_compute_t* compute,
                  hase_result_t* result);
// host-side precondition checks of a one-shot evaluation (no HIP call)
void validate_arguments(const hase_mesh_t* mesh, const hase_experiment_t* experiment, const hase_compute_t* compute);
// Replaces the walk's copies of beta (FP32 beta, beta * V, column-major beta,
// the sum of beta * V) by those of `beta` [P], FP64 in device prism order,
// already on the device; block_sums holds ceil(P / 256) doubles.
void refresh_beta_device(hase_context* ctx, const double* beta, double* block_sums);

// single-ray check in the caller's units (hase_ase.hip): heights in cm, caller
// prism order; one thread per ray
__global__ void hase_single_rays_kernel(MeshDev m, double h, const uint32_t* user2dev, uint32_t n,
                                        const float* start, const uint32_t* prism, const float* end,
                                        const uint32_t* nrefl, const uint32_t* first, float* bl, float* gl, float* cl,
                                        float* refl, uint32_t* steps, int32_t* status);

// persistent work queue (queue.hip)
void check_tuning(const hase_tuning_t& t);
uint32_t queue_max_rounds(const Schedule& s);
void queue_targets(hase_context* ctx);
void queue_reserve(hase_context* ctx, uint32_t batch);
QueueState queue_state(hase_context* ctx);
float queue_rounds(hase_context* ctx, uint32_t base, bool mapped, uint32_t slots, uint32_t launch0, bool defer);
}  // namespace hase
