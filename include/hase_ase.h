/*
 * hase_ase.h - C ABI of the MI355X importance-sampled ASE ray tracer.
 *
 * One entry point computes, for every sample point of an extruded
 * triangle-prism gain medium, the amplified-spontaneous-emission flux
 * phiAse, its convergence statistic mse, the number of rays spent and the
 * population derivative dndtAse.  It is the C form of the legacy
 * calcPhiAse entry whose slot is now an empty header
 * (include/core/calcPhiAse.hpp) and whose present C++ shape is
 * `int startSimulation<bool>(ExperimentParameters&, ComputeParameters&,
 * Result&, HostMesh&)` (include/core/simulation.hpp:42-47, returns 0 on
 * success at :332).
 *
 * Conventions (all arrays are borrowed for the duration of a call):
 *   T  = number_of_triangles, L = number_of_levels, P = T * (L - 1) prisms,
 *   Np = number_of_points,    S = Np * L sample points, W = wavelengths.
 *   Prism j = t + l * T (level-major).  Sample s = p + l * Np, located at
 *   (x_p, y_p, l * thickness).  Per-triangle edge arrays are column-major:
 *   entry (t, k) is stored at t + k * T (pyInclude/geometry/core.py:769-773).
 *   Edge k of triangle t runs from vertex k to vertex (k + 1) % 3.
 *
 * Errors: no C++ exception crosses this boundary.  Every function returns
 * HASE_OK (0) or a negative HASE_E_* code and records a message that
 * hase_last_error() returns (thread-local).  Precondition checks follow
 * src/openpmd/OpenPmdParser.cpp:1408-1426 (minRays > 0, maxRays >= minRays,
 * threshold >= 0).
 *
 * Threading: one GPU, one HIP stream per context.  Calls on different
 * contexts may run concurrently; calls on one context must be serialised by
 * the caller.
 */
#ifndef HASE_ASE_H
#define HASE_ASE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HASE_ABI_VERSION 2

enum hase_status {
    HASE_OK = 0,
    HASE_E_INVALID = -1,   /* argument or precondition violated */
    HASE_E_DEVICE = -2,    /* HIP runtime error (no device, launch failure) */
    HASE_E_MEMORY = -3,    /* device or host allocation failed */
    HASE_E_INTERNAL = -4   /* unexpected internal failure */
};

/* Extruded triangle-prism mesh.  Mirrors the legacy prism fields that
 * HostMesh still carries (include/core/mesh.hpp:532-538, 301-332) and the
 * arrays built by pyInclude/geometry/core.py:671-774. */
typedef struct hase_mesh {
    uint32_t number_of_points;      /* Np: transverse points per level */
    uint32_t number_of_triangles;   /* T */
    uint32_t number_of_levels;      /* L >= 2 z planes */
    double thickness;               /* distance between adjacent planes, cm */
    const double* points;           /* [2*Np]: x[0..Np) then y[0..Np) (core.py:1022) */
    const uint32_t* triangle_point_indices; /* [3T] column-major, counter-clockwise */
    const int32_t* triangle_neighbors;      /* [3T] neighbour across edge k, -1 = boundary */
    const int32_t* forbidden_edge;          /* [3T] index of the shared edge inside that neighbour */
    const double* beta_volume;      /* [P] excited-state fraction per prism */
    const double* beta_cells;       /* [S] excited-state fraction per sample point (dndtAse) */
    const uint32_t* cladding_cell_types; /* [T] or NULL */
    uint32_t cladding_number;       /* triangles whose type equals this absorb */
    double cladding_absorption;     /* cm^-1 */
    double n_tot;                   /* active-ion density, cm^-3 */
    double crystal_t_fluo;          /* fluorescence lifetime tau, s */
    const float* reflectivities;    /* [2T] or NULL: index t*2 + {0 bottom, 1 top}
                                       (legacy index cell*2 + face%2, mesh.hpp:308) */
    const float* refractive_indices;/* [4] or NULL: {bottom inside, bottom outside,
                                       top inside, top outside} (mesh.hpp:312-332) */
} hase_mesh_t;

/* Mirrors ExperimentParameters (include/core/types.hpp:248-331) with the
 * legacy mseThreshold in place of relativeStandardErrorThreshold. */
typedef struct hase_experiment {
    uint32_t min_rays_per_sample;
    uint32_t max_rays_per_sample;
    uint32_t number_of_wavelengths; /* W >= 1 */
    const double* sigma_a;          /* [W] absorption cross sections, cm^2 */
    const double* sigma_e;          /* [W] emission cross sections, cm^2 */
    double max_sigma_a;             /* pair used by the importance map and dndtAse */
    double max_sigma_e;
    double mse_threshold;           /* stop when mse (the squared relative standard error) <= this */
    int32_t use_reflections;        /* top/bottom reflections by mirrored images */
    uint32_t max_reflections;       /* reflection order cap when use_reflections (0 -> 1) */
    double defensive_fraction;      /* share of the proposal drawn from beta*V, in [0, 1) */
} hase_experiment_t;

/* Mirrors ComputeParameters (include/core/types.hpp:79-158). */
typedef struct hase_compute {
    uint32_t max_repetitions;       /* >= 1: pooled launches at the maximum ray count */
    uint32_t adaptive_steps;        /* geometric increases from min to max rays */
    uint32_t min_sample_range;      /* inclusive first sample computed */
    uint32_t max_sample_range;      /* inclusive last sample computed */
    uint32_t rng_seed;
    int32_t device;                 /* HIP device ordinal */
    uint32_t rays_per_thread;       /* tuning: rays a lane takes per work item (0 = auto) */
    uint32_t block_size;            /* tuning: threads per ray work-group (0 = auto) */
    uint32_t sample_batch;          /* tuning: samples resident at once (0 = auto: up to 8192 in 6 GiB, or up
                                       to 2048 for meshes too large for that, within a quarter of the free
                                       device memory and 48 GiB) */
} hase_compute_t;

/* Per-sample outputs, caller-owned, each with S entries.  Samples outside
 * [min_sample_range, max_sample_range] are written as zero.  Mirrors
 * Result (include/core/types.hpp:201-246): mse takes the slot of
 * relativeStandardError (mse = relative standard error squared),
 * standard_error = sqrt(mse) * |phiAse|. */
typedef struct hase_result {
    float* phi_ase;
    double* mse;
    uint32_t* total_rays;
    double* dndt_ase;
    double* standard_error;         /* optional, may be NULL */
    uint32_t* dropped_rays;         /* optional, may be NULL */
} hase_result_t;

/* Counters and device timings of the last evaluation of a context. */
typedef struct hase_stats {
    double importance_ms;           /* summed device time of the importance-map kernel */
    double trace_ms;                /* summed device time of the ray-trace kernel */
    double finalize_ms;             /* summed device time of the statistics kernel */
    double total_ms;                /* device time from first to last launch */
    uint64_t rays;                  /* rays traced (all samples, all launches) */
    uint64_t ray_steps;             /* prism steps walked by those rays */
    uint64_t importance_walks;      /* centroid walks of the importance map */
    uint64_t importance_steps;      /* prism steps of those walks */
    uint64_t dropped_rays;
    uint32_t trace_launches;        /* launches of the ray-trace kernel */
    uint32_t importance_launches;
} hase_stats_t;

typedef struct hase_context hase_context_t;

/* Whole evaluation in one call: upload, importance map, adaptive rays,
 * results.  Replaces calcPhiAse / startSimulation
 * (include/core/simulation.hpp:42-333). */
int hase_calc_phi_ase(const hase_mesh_t* mesh,
                      const hase_experiment_t* experiment,
                      const hase_compute_t* compute,
                      hase_result_t* result);

/* Bytes of the fixed device scratch behind hase_calc_phi_ase_noalloc. */
#define HASE_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)

/* hase_calc_phi_ase without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * that forbids hipMalloc, tight serving loops).  Same arguments, results and
 * random streams as hase_calc_phi_ase; every device buffer is carved out of
 * one static device array of HASE_NOALLOC_SCRATCH_BYTES per GPU, held from
 * the first call on that GPU until the process ends.  The call creates no
 * stream, event or graph (it runs on the null stream and returns when the
 * results are on the host).  A problem whose mesh plus one sample point's
 * proposal tables (12 B per prism and reflection slice, ~28 KB more) does
 * not fit returns HASE_E_MEMORY with the sizes in hase_last_error(); larger
 * problems are evaluated in as many sample batches as the scratch takes.
 * Calls on one GPU share its array and are serialised by a lock within the
 * process; the array keeps the last call's contents, and each call
 * initialises what it reads.  Another process's calls have their own copy. */
int hase_calc_phi_ase_noalloc(const hase_mesh_t* mesh,
                              const hase_experiment_t* experiment,
                              const hase_compute_t* compute,
                              hase_result_t* result);

/* Persistent context: keeps the mesh resident on the device between
 * evaluations, like ForwardPhiAseContext
 * (include/core/forwardPhiAseEvaluator.hpp:180-204). */
int hase_context_create(const hase_mesh_t* mesh,
                        const hase_experiment_t* experiment,
                        const hase_compute_t* compute,
                        hase_context_t** out_context);

/* Replace beta per prism and per sample point without re-uploading the
 * geometry (refreshDynamicMeshes, forwardPhiAseEvaluator.hpp:407-433).
 * beta_cells may be NULL to keep the current values. */
int hase_context_update_beta(hase_context_t* context,
                             const double* beta_volume,
                             const double* beta_cells);

/* Run one evaluation with the context's parameters. */
int hase_context_evaluate(hase_context_t* context, hase_result_t* result);

/* Counters and device timings of the last evaluation. */
int hase_context_stats(const hase_context_t* context, hase_stats_t* stats);

void hase_context_destroy(hase_context_t* context);

/* Single-ray check: walk n rays from explicit start points to explicit
 * sample points and return, per ray, the FP32 gain exponent parts and the
 * reflectivity product.  start/end are [3n] (x, y, z interleaved; end may be
 * an unfolded mirror image when reflections > 0), start_prism [n],
 * reflections [n] and first_plane [n] (0 bottom, 1 top).  Outputs are [n]:
 * beta_length = sum(beta*l) over gain prisms, gain_length = sum(l) over
 * gain prisms, clad_length = sum(l) over cladding prisms, reflectance =
 * product of surface factors, steps = prisms visited, status = 0 ok.
 * Plays the role of the legacy propagateRay (retired,
 * src/kernels/propagateRay.cpp:9-10) seen from the host, like the
 * reference's one-thread TraverseOneRay test kernel
 * (tests/forwardRayTraversal.cpp:98-152). */
int hase_context_trace_rays(hase_context_t* context,
                            uint32_t n,
                            const float* start,
                            const uint32_t* start_prism,
                            const float* end,
                            const uint32_t* reflections,
                            const uint32_t* first_plane,
                            float* beta_length,
                            float* gain_length,
                            float* clad_length,
                            float* reflectance,
                            uint32_t* steps,
                            int32_t* status);

/* Importance map of one sample: the unnormalised importance of every
 * (reflection slice, prism) pair, [R*P] with R = 1 + 2*K, K = reflection
 * order cap.  Replaces importanceSamplingPropagation (retired, shell at
 * include/kernels/importanceSampling.hpp:10-18). */
int hase_context_importance(hase_context_t* context,
                            uint32_t sample,
                            float* importance,
                            uint32_t capacity);

/* ------------------------------------------------------------------------
 * Learned proposal.  The reference has no counterpart (its importance map is
 * the centroid-propagation pass, retired shell
 * include/kernels/importanceSampling.hpp:10-18); these entry points extend
 * the map above.
 *
 * Network: HASE_MLP_INPUTS inputs -> 4 hidden layers of HASE_MLP_HIDDEN
 * SiLU units -> 1 output, the natural log of the transfer factor (spectral
 * gain times reflectance) from a prism centroid to a sample point.  The
 * importance of prism j for sample s is then beta_j V_j exp(y) / (4 pi d^2).
 * Inputs, in order: sample (or mirror image) x, y, z; prism centroid x, y, z
 * (cm); level; beta; centroid-to-sample distance (cm); log(max_sigma_e /
 * max_sigma_a).  Parameters are FP32, row-major W[out][in] per layer, in the
 * order W1 b1 W2 b2 W3 b3 W4 b4 w5 b5 (hase_mlp_param_count() floats).
 * Every hase_mlp_* call takes inputs already normalised by the caller.
 * ---------------------------------------------------------------------- */
#define HASE_MLP_INPUTS 10
#define HASE_MLP_HIDDEN 128

#define HASE_PROPOSAL_PLAIN 0
#define HASE_PROPOSAL_LEARNED 1
/* Rays drawn in proportion to beta * V (each reflection slice 1/R of it),
 * one table shared by every sample point and no per-sample map: the
 * estimator without importance information, the way the present method
 * chooses its ray sources (include/kernels/forwardPhiAseMapping.hpp:21-49
 * with the scan of include/core/calcForwardPhiAse.hpp:347-378).  Only the
 * kind field of hase_proposal_t is read. */
#define HASE_PROPOSAL_BETA_VOLUME 2

typedef struct hase_train {
    uint32_t epochs;
    uint32_t batch;                 /* rows per Adam step */
    float learning_rate;
    float beta1, beta2, epsilon;    /* Adam moments */
    float weight_decay;             /* decoupled (AdamW) */
    uint32_t seed;                  /* row shuffling */
} hase_train_t;

typedef struct hase_proposal {
    int32_t kind;                   /* HASE_PROPOSAL_PLAIN, _LEARNED or _BETA_VOLUME */
    const float* params;            /* [hase_mlp_param_count()] FP32 weights (kept as BF16 on the device) */
    const float* input_shift;       /* [HASE_MLP_INPUTS] input normalisation: (x - shift) * scale */
    const float* input_scale;       /* [HASE_MLP_INPUTS] */
    uint32_t cluster_triangles;     /* triangles per evaluated cluster (device order), >= 1 */
    uint32_t cluster_levels;        /* prism layers per evaluated cluster, >= 1 */
    int32_t fallback;               /* 1: samples that miss mse_threshold after the first
                                       launch restart on the plain importance map */
} hase_proposal_t;

/* Counters of the learned proposal during the last evaluation. */
typedef struct hase_proposal_stats {
    double inference_ms;            /* BF16 network kernel */
    double proposal_ms;             /* per-prism proposal and prefix from the network output */
    uint64_t evaluations;           /* network rows evaluated */
    uint32_t fallback_samples;      /* samples moved to the plain map */
    uint32_t clusters;              /* clusters per reflection slice */
} hase_proposal_stats_t;

uint32_t hase_mlp_param_count(void);

/* FP32 forward pass on the device: outputs[i] = net(inputs[i*HASE_MLP_INPUTS ..]). */
int hase_mlp_forward(const float* params, uint32_t n, const float* inputs, float* outputs, int32_t device);

/* Weighted squared error L = sum w_i (y_i - t_i)^2 / sum w_i (weights NULL:
 * all 1) and its gradient with respect to params, by the device backward
 * kernels. */
int hase_mlp_loss_gradient(const float* params, uint32_t n, const float* inputs, const float* targets,
                           const float* weights, double* loss, float* gradient, int32_t device);

/* Adam(W) training on the device; params are updated in place and the mean
 * batch loss of every epoch is written to epoch_loss[epochs] (may be NULL). */
int hase_mlp_train(float* params, uint32_t n, const float* inputs, const float* targets, const float* weights,
                   const hase_train_t* config, float* epoch_loss, int32_t device);

/* BF16-MFMA inference (FP32 accumulate) on explicit rows, the arithmetic
 * the learned proposal uses. */
int hase_mlp_forward_bf16(const float* params, uint32_t n, const float* inputs, float* outputs, int32_t device);

/* Select the proposal of later evaluations (kind PLAIN restores the map). */
int hase_context_set_proposal(hase_context_t* context, const hase_proposal_t* proposal);

/* Transfer factor (spectral gain times reflectance) of the centroid walk of
 * every (reflection slice, prism) pair for samples [first, first+count):
 * transfer[(s*R + k)*P + j], caller prism order; the network's training
 * target is its logarithm.  Non-walkable pairs are written as 0. */
int hase_context_transfer_map(hase_context_t* context, uint32_t first, uint32_t count, float* transfer,
                              size_t capacity);

/* Network outputs of the learned proposal for samples [first, first+count):
 * logits[(s*R + k)*C + c], C clusters per slice; clusters (may be NULL)
 * receives C rows of (x, y, level, beta) of each cluster's representative
 * prism (area-weighted centroid, mean level, volume-weighted beta). */
int hase_context_proposal_logits(hase_context_t* context, uint32_t first, uint32_t count, float* logits,
                                 size_t capacity, float* clusters, uint32_t* n_clusters);

int hase_context_proposal_stats(const hase_context_t* context, hase_proposal_stats_t* stats);

/* ------------------------------------------------------------------------
 * Ray scheduling.  No counterpart in the reference, whose adaptive loop is a
 * host loop of launches (include/core/forwardPhiAseEvaluator.hpp:114-178);
 * here the default runs each adaptive round of a batch of samples as one
 * persistent work-queue kernel: every wavefront takes queue items (a run of
 * rays_per_item rays of one sample) from a device cursor, its lanes walk
 * those rays (optionally refilling with new ones when refill_lanes of them
 * are idle),
 * and the per-sample schedule lives on the device, so the rounds of a batch
 * replay as one captured hipGraph with a single host synchronisation.
 * ---------------------------------------------------------------------- */
#define HASE_TRACE_QUEUE 0
#define HASE_TRACE_ROUNDS 1

typedef struct hase_tuning {
    uint32_t mode;            /* HASE_TRACE_QUEUE (default) or HASE_TRACE_ROUNDS (one launch per round) */
    uint32_t block_size;      /* queue kernel threads per work-group: 64..1024, multiple of 64 */
    uint32_t rays_per_item;   /* rays per queue item, multiple of 64 */
    uint32_t queue_chunk;     /* items a wavefront takes per queue access, >= 1 */
    uint32_t refill_lanes;    /* idle lanes that trigger a refill, 1..64; 0: each lane walks its rays to the end */
    uint32_t step_budget;     /* prism steps between refill checks, >= 1 */
    uint32_t use_graph;       /* 1: replay the rounds of a batch as a captured hipGraph */
    uint32_t stage_lds;       /* 1: walk compact triangle records, staged in LDS when they fit (32-byte records
                                 up to 4800 triangles, 16-byte indexed ones up to ~8,100), else read from
                                 global memory; 0: the 48-byte global records (refill always uses these) */
    double probe_ms;          /* hase_context_autotune: device time of the chosen setting's probe */
} hase_tuning_t;

int hase_context_set_tuning(hase_context_t* context, const hase_tuning_t* tuning);
int hase_context_get_tuning(const hase_context_t* context, hase_tuning_t* tuning);

/* Autotuner: times one adaptive round of the first probe_samples samples of
 * the context's range (0: 512) under candidate settings (coordinate descent
 * over block_size, rays_per_item, queue_chunk, refill_lanes, step_budget),
 * keeps the fastest on the context and returns it. */
int hase_context_autotune(hase_context_t* context, uint32_t probe_samples, hase_tuning_t* best);

/* Batch of independent problems: the resident samples of every context run
 * their adaptive rounds together in one persistent work-queue kernel per
 * round (uneven budgets, different meshes, spectra and schedules), so small
 * problems no longer leave the GPU idle.  Each context keeps its proposal
 * (walked or learned) and its outputs; ray scheduling follows the first
 * context's tuning.  A context whose triangle records fit in LDS (up to
 * ~8,100 triangles), whose own tuning stages them there, and which has at
 * least 8192 sample points is evaluated on its own queue first, as
 * hase_context_evaluate would (its samples fill the chip alone and its
 * trace kernel reads the records from LDS); the shared queue takes the
 * others.  Contexts must be on one device, stay owned by the
 * caller and outlive the batch.  No counterpart in the reference (one
 * problem per calcPhiASE run, include/core/simulation.hpp:42-333). */
typedef struct hase_batch hase_batch_t;

int hase_batch_create(uint32_t n_problems, hase_context_t* const* contexts, hase_batch_t** out_batch);

/* results[i] receives context i's outputs, as hase_context_evaluate would. */
int hase_batch_evaluate(hase_batch_t* batch, hase_result_t* const* results);

/* Summed counters and device times of the last batch evaluation. */
int hase_batch_stats(const hase_batch_t* batch, hase_stats_t* stats);

void hase_batch_destroy(hase_batch_t* batch);

/* ------------------------------------------------------------------------
 * Time-stepped beta loop: the reference's CompiledSimulationRunner
 * (include/core/timeSteppedSimulation.hpp:132-615) on the prism mesh.  The
 * state is beta per prism, FP64, resident on the device next to the
 * context; each stage of an integrator refreshes the context's walk copies
 * of beta and its importance map from that state (no host round trip of
 * beta), evaluates the ASE flux at the sample points, averages it over the
 * six vertices of each prism (the legacy wedge mean of
 * tests/data/laserPumpCladding/fixed_legacy_no_reflection_pump3_wedge_reference/generate_reference.py:138-141)
 * and composes the derivative of include/kernels/derivativeComposition.hpp:27-60:
 *   d beta/dt = pump - (beta (sigma_e + sigma_a) - sigma_a) phi - beta / tau,
 * with sigma_e = max_sigma_e, sigma_a = sigma_a at the peak of sigma_e
 * (timeSteppedSimulation.hpp:45-54) and tau = crystal_t_fluo.  Cladding
 * prisms keep their beta.  The pump is local and linear in beta,
 *   pump_j = pump_rate_j - pump_saturation_j * beta_j   [1/s],
 * the form of a pump of fixed intensity on a two-level absorber; it stands
 * in for the reference's ray-traced pump (include/kernels/generalPump.hpp),
 * which is not part of this library.  Every ASE evaluation uses the
 * context's settings and seed, as the reference's runner re-evaluates with
 * one ExperimentParameters.
 * ---------------------------------------------------------------------- */
#define HASE_INTEGRATOR_EXPLICIT_EULER 0               /* "explicit-euler" */
#define HASE_INTEGRATOR_HEUN 1                         /* "heun" */
#define HASE_INTEGRATOR_MIDPOINT 2                     /* "midpoint" */
#define HASE_INTEGRATOR_RUNGE_KUTTA_4 3                /* "runge-kutta-4" */
#define HASE_INTEGRATOR_FROZEN_PHI_ASE_RUNGE_KUTTA_4 4 /* "frozen-phi-ase-runge-kutta-4": one ASE flux per step */
#define HASE_INTEGRATOR_IMPLICIT_EULER 5               /* "implicit-euler": fixed-point iterations */
#define HASE_INTEGRATOR_EXPONENTIAL_EULER 6            /* "exponential-euler": exact fluorescence decay */

/* Run control (include/core/simulationRunControl.hpp:76-81, 115-127). */
typedef struct hase_run_control {
    double time_step;               /* seconds, > 0 */
    uint32_t integrator;            /* HASE_INTEGRATOR_* */
    uint32_t implicit_iterations;   /* implicit Euler iterations (0 counts as 1) */
    uint32_t first_simulation_step; /* index of the first step (activity windows count from it) */
    uint32_t ase_steps;             /* ASE is evaluated in steps < ase_steps */
    uint32_t pump_steps;            /* the pump acts in steps < pump_steps */
    uint32_t pre_pump;              /* 1: no ASE in step 0 */
} hase_run_control_t;

/* State after the last completed step (include/core/simulationSnapshot.hpp:18-26);
 * every pointer may be NULL, each non-NULL one receives its array.  Prism
 * arrays [P] are in the caller's prism order (j = t + l * T); sample arrays
 * [S] come from the last ASE evaluation (zeros when ASE was off). */
typedef struct hase_time_state {
    uint32_t step;                  /* steps completed since creation */
    uint32_t ase_evaluations;       /* ASE evaluations since creation */
    double time;                    /* step * time_step, seconds */
    double* beta_volume;            /* [P] */
    double* phi_ase;                /* [P] six-vertex mean of the sample flux */
    double* dndt_ase;               /* [P] (beta (sigma_e + sigma_a) - sigma_a) phi of the last stage */
    double* dndt_pump;              /* [P] pump term of the last stage */
    float* phi_samples;             /* [S] */
    double* mse;                    /* [S] */
    uint32_t* total_rays;           /* [S] */
    double* standard_error;         /* [S] */
} hase_time_state_t;

typedef struct hase_simulation hase_simulation_t;

/* Starts a loop on `context` (which must outlive it, cover all S samples and
 * is not used by the caller while the loop exists) from beta_volume [P]
 * (NULL: the context's current beta).  pump_rate [P] is required,
 * pump_saturation [P] may be NULL (0).  Mirrors the construction and
 * validation of timeSteppedSimulation.hpp:56-129, 140-175. */
int hase_simulation_create(hase_context_t* context, const hase_run_control_t* run, const double* beta_volume,
                           const double* pump_rate, const double* pump_saturation, hase_simulation_t** out_sim);

/* Replaces the pump terms from the next stage on. */
int hase_simulation_set_pump(hase_simulation_t* sim, const double* pump_rate, const double* pump_saturation);

/* Overwrites the beta state (the beta_volume control field of the
 * synchronized-debug mode, timeSteppedSimulation.hpp:228-236). */
int hase_simulation_set_beta(hase_simulation_t* sim, const double* beta_volume);

/* Advances n_steps time steps (timeSteppedSimulation.hpp:177-241). */
int hase_simulation_step(hase_simulation_t* sim, uint32_t n_steps);

/* Copies the state after the last completed step. */
int hase_simulation_state(const hase_simulation_t* sim, hase_time_state_t* state);

/* Totals of the loop so far (each pointer may be NULL): wall time of the ASE
 * evaluations (beta refresh, proposal, rays), device time of the integrator
 * kernels, rays traced. */
int hase_simulation_times(const hase_simulation_t* sim, double* ase_ms, double* integrator_ms, uint64_t* rays);

void hase_simulation_destroy(hase_simulation_t* sim);

/* ---------------------------------------------------------------------
 * Forward Tet4 estimator: the reference's present method
 * (include/core/calcForwardPhiAse.hpp:60-300, ForwardPhiAseContext).  Rays start at beta*V-stratified points of a
 * tetrahedral gain medium, travel in isotropic directions until they leave
 * it, and score the track-length integral of their gain into the vertices of
 * every traversed cell; a cell's phiAse is the mean lumped-vertex density of
 * its four vertices (gain and cladding cells keep separate vertex copies),
 * its relative standard error the spread of eight batch means
 * (include/kernels/forwardPhiAseMapping.hpp:47-147).
 *
 * Of hase_experiment_t it reads min/max_rays_per_sample (rays of the WHOLE
 * evaluation, grown geometrically over compute.adaptive_steps as in
 * include/core/types.hpp:333-354), the wavelength tables (each ray takes a
 * stratified wavelength), max_sigma_a/e (dndtAse) and mse_threshold (stop
 * when every cell has relative_standard_error^2 <= it) and use_reflections
 * (surface reservoir passes over the surfaces of
 * hase_forward_set_reflections).  Of hase_compute_t: adaptive_steps,
 * rng_seed, device. */
typedef struct hase_tet_mesh {
    uint32_t number_of_points;
    uint32_t number_of_cells;
    const double* points;           /* [3*Np] SoA: x then y then z (mesh.hpp:182-187) */
    const uint32_t* cells;          /* [4C] point indices, cell-major (cellPointIndices) */
    const double* beta_volume;      /* [C] excited-state fraction per cell */
    const uint32_t* cladding_cell_types; /* [C] or NULL (claddingCellTypes) */
    uint32_t cladding_number;       /* cells whose type equals this absorb */
    uint32_t reserved;
    double cladding_absorption;     /* cm^-1 */
    double n_tot;                   /* cm^-3 */
    double crystal_t_fluo;          /* s */
} hase_tet_mesh_t;

/* Per-cell outputs, caller-owned, C entries each; all but phi_ase may be NULL.
 * Mirrors the forward Result (phiAse, standardError, relativeStandardError,
 * dndtAse; include/core/types.hpp:201-246) plus the walk counters
 * (cellRayVisits, cellDroppedRays; accumulation.hpp:110-117). */
typedef struct hase_forward_result {
    float* phi_ase;
    double* standard_error;
    double* relative_standard_error;  /* DBL_MAX where undetermined, NaN where the flux is zero */
    double* dndt_ase;
    uint32_t* ray_visits;             /* scored segments in the cell */
    uint32_t* dropped_rays;           /* rays whose walk failed in the cell, non-finite scores */
    uint32_t* convergence_rays;       /* ray target at which the cell first met the threshold, 0 = never */
} hase_forward_result_t;

/* Surface-reservoir pass outcome (SrmStatus, include/core/types.hpp:174-181). */
enum hase_srm_status {
    HASE_SRM_DISABLED = 0,
    HASE_SRM_CONVERGED = 1,    /* remaining reflected weight < tolerance */
    HASE_SRM_STABLE = 2,       /* remaining weight changed by < tolerance */
    HASE_SRM_MAX_PASSES = 3,   /* pass limit reached */
    HASE_SRM_DIVERGED = 4      /* weight grew divergence_streak passes in a row */
};

typedef struct hase_forward_stats {
    double trace_ms;                /* summed device time of the ray kernels */
    double total_ms;                /* wall time of the evaluation */
    uint64_t rays;
    uint64_t segments;              /* scored cell segments (primary and reflected rays) */
    uint64_t dropped_rays;
    uint32_t launches;              /* adaptive launches */
    uint32_t open_cells;            /* cells short of the threshold at the end */
    uint32_t srm_status;            /* worst hase_srm_status over batches and launches */
    uint32_t srm_passes;            /* most reflected passes of a batch */
    double srm_remaining;           /* largest remaining weight fraction */
} hase_forward_stats_t;

/* Reflective surfaces of the Tet4 mesh for the surface reservoir method
 * (include/core/forwardSrm.hpp, the mesh's cellFaceBoundaries and surface
 * tables, include/core/mesh.hpp:300-332; controls of ExperimentParameters,
 * types.hpp:328-330).  A boundary face (no neighbour) whose surface id is in
 * (0, number_of_surfaces) reflects with that surface's reflectivity, or
 * totally when (n_inside / n_outside)^2 sin^2(theta) > 1; other boundary faces
 * absorb.  Used when the experiment's use_reflections is set. */
typedef struct hase_forward_reflections {
    const int32_t* face_boundaries; /* [4C] surface id of face k (opposite vertex k) of each cell */
    uint32_t number_of_surfaces;    /* entries of the tables below (id 0 unused) */
    uint32_t reservoir_slots;       /* surfaceReservoirSize, 32 */
    const float* reflectivity;      /* [number_of_surfaces] */
    const float* n_inside;          /* [number_of_surfaces] or NULL (no total reflection) */
    const float* n_outside;         /* [number_of_surfaces] or NULL */
    uint32_t max_passes;            /* reflectionMaxIterations, 40 */
    uint32_t divergence_streak;     /* 3 */
    double tolerance;               /* reflectionTolerance, 1e-4 */
} hase_forward_reflections_t;

typedef struct hase_forward hase_forward_t;

/* Prepares a Tet4 mesh (face planes, neighbours by shared faces, volumes,
 * lumped vertex volumes) on the device.  Degenerate cells and faces shared by
 * more than two cells are rejected. */
int hase_forward_create(const hase_tet_mesh_t* mesh, const hase_experiment_t* experiment,
                        const hase_compute_t* compute, hase_forward_t** out);

/* New beta per cell (the beta*V source prefix is rebuilt). */
int hase_forward_update_beta(hase_forward_t* forward, const double* beta_volume);

/* Adaptive evaluation into `result`. */
int hase_forward_evaluate(hase_forward_t* forward, hase_forward_result_t* result);

/* One launch of `rays` rays with the given seed and launch index, without
 * finalisation: scores [8 * 2 * Np] (batch-major; gain vertex copies then
 * cladding copies), visits and dropped [C], scored segments. */
int hase_forward_trace(hase_forward_t* forward, uint32_t seed, uint32_t launch, uint32_t rays, double* scores,
                       uint32_t* ray_visits, uint32_t* dropped_rays, uint64_t* segments);

/* Sets the reflective surfaces (the whole table is replaced). */
int hase_forward_set_reflections(hase_forward_t* forward, const hase_forward_reflections_t* reflections);

int hase_forward_stats(const hase_forward_t* forward, hase_forward_stats_t* stats);
int hase_forward_destroy(hase_forward_t* forward);

/* One-shot create + evaluate + destroy (calcForwardPhiAse). */
int hase_calc_forward_phi_ase(const hase_tet_mesh_t* mesh, const hase_experiment_t* experiment,
                              const hase_compute_t* compute, hase_forward_result_t* result);

/* ------------------------------------------------------------------------
 * Prism arrays from a 2.2.0 HostMesh.  In 2.2.0 the HostMesh of an extruded
 * slab holds Tet4 cells, three per wedge (utils/convert_vtk_topology.py:38-53
 * splits wedge (a,b,c | d,e,f), base vertices in ascending global order, into
 * [a,b,c,d], [b,e,c,d], [c,e,f,d]; wedge w = t + l*T becomes cells 3w..3w+2),
 * and its legacy alias fields trianglePointIndices / triangleNeighbors /
 * forbiddenEdge are copies of the Tet4 arrays cellPointIndices /
 * cellNeighborCells / cellFaceBoundaries (include/core/mesh.hpp:605-625), not
 * prism arrays.  hase_fold_tet4_mesh folds the cells back into the prism
 * arrays hase_mesh_t takes.  Host-only: it needs no device.
 *
 * Each triangle comes out counter-clockwise and starting at its lowest point
 * index (the ascending base order of the split, swapped when clockwise);
 * neighbours and forbidden edges are derived from those triangles.  Prism
 * beta is the tetrahedra's common value, or their volume-weighted mean when
 * they differ; a triangle's cladding type must be the same in all its cells.
 * Bottom / top reflectivity of triangle t is that of the base face of its
 * lowest cells / the top face of its highest cells, looked up as
 * getSurfaceReflectivity does (include/core/mesh.hpp:300-310: the face's
 * surface id in surface_reflectivities when valid, else the legacy entry
 * cell*2 + face%2); refractive indices likewise per plane (mesh.hpp:312-332),
 * which must agree over all triangles. */
typedef struct hase_tet4_host_mesh {
    uint32_t number_of_mesh_points;      /* numberOfMeshPoints = Np * L */
    uint32_t number_of_cells;            /* numberOfCells = 3 * T * (L - 1) */
    uint32_t number_of_points;           /* numberOfPoints: points per level, Np */
    uint32_t number_of_levels;           /* numberOfLevels, L >= 2 */
    const double* points;                /* [3 * Nm] SoA x, y, z; point p + l*Np is p on level l */
    const uint32_t* cell_point_indices;  /* [4 * C] cell-major */
    const double* beta_volume;           /* [C] */
    const uint32_t* cladding_cell_types; /* [C] or NULL */
    const int32_t* cell_face_boundaries; /* [4 * C] or NULL: surface id of face k (opposite vertex k) */
    uint32_t number_of_surfaces;         /* entries of the three surface tables below */
    const float* surface_reflectivities; /* [number_of_surfaces] or NULL */
    const float* surface_refractive_index_inside;  /* [number_of_surfaces] or NULL */
    const float* surface_refractive_index_outside; /* [number_of_surfaces] or NULL */
    uint32_t number_of_reflectivities;   /* entries of the legacy reflectivities */
    const float* reflectivities;         /* legacy per-cell table or NULL */
    const float* refractive_indices;     /* legacy [4] or NULL */
} hase_tet4_host_mesh_t;

/* Caller-owned outputs, T = C / (3 (L - 1)) triangles, P = T (L - 1). */
typedef struct hase_prism_arrays {
    double* points;                      /* [2 Np]: x then y */
    uint32_t* triangle_point_indices;    /* [3T] column-major */
    int32_t* triangle_neighbors;         /* [3T] column-major, -1 at the boundary */
    int32_t* forbidden_edge;             /* [3T] column-major */
    double* beta_volume;                 /* [P] level-major */
    uint32_t* cladding_cell_types;       /* [T] or NULL */
    float* reflectivities;               /* [2T] or NULL: t*2 + {0 bottom, 1 top} */
    float refractive_indices[4];         /* {bottom in, bottom out, top in, top out} */
    double thickness;                    /* distance of levels 0 and 1 */
    uint32_t number_of_triangles;        /* T */
} hase_prism_arrays_t;

/* Replaces the legacy alias fields of HostMesh (include/core/mesh.hpp:532-538,
 * 605-625) with real prism arrays. */
int hase_fold_tet4_mesh(const hase_tet4_host_mesh_t* mesh, hase_prism_arrays_t* out);

/* Most recent error message of this thread ("" if none). */
const char* hase_last_error(void);

/* ABI version (HASE_ABI_VERSION) and a build description string. */
int hase_abi_version(void);
const char* hase_build_info(void);

#ifdef __cplusplus
}
#endif

#endif /* HASE_ASE_H */
