// queue.h - device tables of the persistent work-queue ray scheduler.
//
// One queue round serves any number of problems (meshes with their own
// proposal, spectra and schedule): every resident sample is a global slot g
// that belongs to problem slot_problem[g]; per-problem constants come from a
// ProblemDev table in device memory, per-slot state from flat arrays.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/hase_ase.h"
#include "prism_walk.h"

struct hase_context;

namespace hase {

struct ProblemDev {
    MeshDev m;
    const float2* pts;
    const float* c1;           // [n_lambda] nTot (sigma_e + sigma_a) log2(e)
    const float* c2;           // [n_lambda] nTot sigma_a log2(e)
    const double* prefix;      // [local slots * n] proposal prefix
    const double* Q;           // [local slots] proposal totals
    const uint32_t* buckets;   // [local slots * kPrefixBuckets] prefix buckets
    const uint32_t* sample_of; // local slot -> sample, or null: sample_base + local slot
    const uint32_t* targets;   // cumulative ray target of schedule increase k, k < n_targets
    size_t n;                  // R * P entries per slot
    double mse_threshold, phi_scale;
    uint32_t n_lambda, R, seed, slot0, sample_base, n_targets, max_rays, reps, launch0, defer;
    uint32_t shared;           // 1: every slot reads the proposal of slot 0 (beta * V sampling)
    float inv_lambda;
};

struct QueueState {
    const ProblemDev* problems;
    const uint32_t* slot_problem;   // [slots]
    uint32_t* flags;
    uint32_t* k;
    uint32_t* extra;
    uint32_t* launch;
    uint32_t* nrays;
    uint32_t* item_rays;            // rays per queue item of each slot this round
    uint32_t* target;
    uint32_t* list;                 // [slots] listed slots of the round
    uint32_t* item_off;             // [slots + 1] first queue item of each listed slot
    uint32_t* scalars;              // [0] listed slots, [1] queue cursor, [2] slots
    double* s1;
    double* s2;
    uint32_t* N;
    uint32_t* dropped;
    float* phi;
    double* mse;
    double* se;
    double2* partial;               // [items] per-item FP64 sums
    uint32_t* partial_drop;
    unsigned long long* counters;   // [0] rays, [1] prism steps
    uint32_t rays_per_item, queue_chunk, refill, budget;
};

constexpr uint32_t kQueueActive = 1u, kQueueDeferred = 2u;
constexpr int kQueueScalars = 4;
// Queue items of one slot in one round at most: larger rounds get longer
// items, so the per-item partials stay at 16 KB per slot.
constexpr uint32_t kQueueItemsPerSlot = 1024u;

struct GraphCache {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t key = 0;
    void reset()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key = 0;
    }
};

// Problem constants of a context whose resident batch starts at sample
// `base` (or follows ctx->slot_sample when mapped).
ProblemDev queue_problem(const hase_context* ctx, uint32_t slot0, uint32_t base, bool mapped, uint32_t launch0,
                         bool defer);

// Items a slot of this context can need in one round.
size_t queue_items_per_slot(const hase_context* ctx, const hase_tuning_t& tu);

// Runs `rounds` adaptive rounds over `slots` global slots on `stream`
// (the state is reset first); returns the device time in ms.  `single`
// (a batch of one problem) passes the problem as a kernel argument instead
// of the device table; `table_compact`: every problem of the table has its
// compact records in global memory (the walks then read those).
float queue_run(hipStream_t stream, int device, const hase_tuning_t& tu, const QueueState& qs, uint32_t slots,
                uint32_t rounds, GraphCache& gc, uint64_t generation, const ProblemDev* single = nullptr,
                bool table_compact = false);

}  // namespace hase
