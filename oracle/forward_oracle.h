/*
 * forward_oracle.h - FP64 CPU restatement of the reference's forward Tet4
 * ASE estimator (TEST INFRASTRUCTURE ONLY: the checker of the HIP forward
 * path, never the thing measured or shipped).
 *
 * Rays start at beta*V-stratified points of the tetrahedral gain medium,
 * travel in isotropic directions until they leave the mesh, and score the
 * track-length integral of their accumulated gain into the four vertices of
 * every traversed cell (barycentric weights of the segment midpoint, gain and
 * cladding kept apart); a cell's flux is the mean lumped-vertex density over
 * its vertices, its relative standard error the spread of eight batch means.
 *
 * Followed:
 *   include/kernels/forward/accumulation.hpp:300-405 (AccumulateForwardPhiAse,
 *     ForwardAseCellPolicy), rayWalk.hpp:106-208 (nextFaceIntersection, gain,
 *     track-length integral), rayTransition.hpp:17-188 (ties, recovery),
 *     policyRay.hpp:379-465 (walk), barycentric.hpp:69-99 (midpoint vertex
 *     weights), volumeSampling.hpp:19-32, 80-137 (stratified beta*V source,
 *     stratified wavelength), include/core/mesh.hpp:630-700 (face planes,
 *     beta*V prefix), include/kernels/vertexAccumulation.hpp:26-39 (lumped
 *     material vertex volumes), include/kernels/forwardPhiAseMapping.hpp:
 *     47-147 (FinalizeForwardVolumePhiAse), include/core/types.hpp:333-396
 *     (adaptive targets, stop rule).
 * Chosen here (no reference counterpart, documented in the library):
 *   the eight RSE batches are contiguous ray ranges of a launch; the random
 *   streams are Philox4x32-10 counters (ray, launch, slot); the start point
 *   folds three uniforms into the tetrahedron (Rocchini & Cignoni, "Generating
 *   random points in a tetrahedron", JGT 5(4), 2000) and the direction is
 *   Marsaglia's (1972) rejection method, so neither needs a transcendental
 *   function and the HIP path draws bit-identical rays; the track-length
 *   integral uses expm1.
 */
#ifndef HASE_FORWARD_ORACLE_H
#define HASE_FORWARD_ORACLE_H

#include <stdint.h>

#define ORF_BATCHES 8u
#define ORF_MAX_STEPS 10000u
#define ORF_MAX_RECOVERY 32u
#define ORF_DIRECTION_CALLS 32u

typedef struct {
    uint32_t n_points, n_cells;
    const double* px;
    const double* py;
    const double* pz;
    const uint32_t* cells;       /* 4 point indices per cell, cell-major */
    const uint32_t* cell_types;  /* NULL = all gain */
    uint32_t cladding_number;
    double cladding_absorption;
    const double* beta;          /* per cell */
    double n_tot, tau;
    /* derived by orf_prepare, caller-allocated */
    double* planes;              /* 16 per cell: face k -> (gx, gy, gz, d), lambda_k(x) = g.x + d */
    int32_t* nbr;                /* 4 per cell: cell across face k, -1 = boundary */
    int32_t* nbr_face;           /* 4 per cell: that cell's local face index */
    double* volume;              /* per cell */
    double* bv_prefix;           /* per cell: running sum of beta * volume */
    double* lumped;              /* 2 * n_points: (gain | cladding) lumped vertex volumes */
} orf_mesh;

typedef struct {
    int32_t face;        /* exit face, -1 = none */
    double length;
    uint32_t tied_mask;
} orf_hit;

enum { ORF_ENTERED = 0, ORF_BOUNDARY = 1, ORF_FAILED = 2 };

typedef struct {
    uint32_t cell;
    int32_t forbidden;
    int32_t boundary_face;
    int32_t status;
} orf_transition;

typedef struct {
    uint32_t cell;
    double p[3];
    double u[3];
    uint32_t lambda;
    uint32_t batch;
    uint32_t batch_index;
    uint32_t batch_count;
} orf_ray;

typedef struct {
    uint32_t cell;
    int32_t face;      /* boundary face the walk left through, -1 if it did not */
    double weight;     /* source weight times accumulated gain there */
} orf_exit;

int orf_prepare(orf_mesh* m);
void orf_set_beta(orf_mesh* m, const double* beta);

orf_hit orf_next_face(const orf_mesh* m, uint32_t cell, const double p[3], const double u[3], int32_t forbidden);
orf_transition orf_recover(const orf_mesh* m, uint32_t cell, int32_t exit_face, const double p[3],
                           const double u[3]);
orf_transition orf_cross(const orf_mesh* m, uint32_t cell, orf_hit hit, const double p[3], const double u[3]);

void orf_batch_range(uint32_t rays, uint32_t batch, uint32_t* first, uint32_t* count);
void orf_batch_controls(uint32_t seed, uint32_t launch, uint32_t batch, uint32_t n_lambda, double* shift,
                        uint32_t* phase);
void orf_point_in_tet(const double u3[3], double bary[4]);
int orf_direction(uint32_t seed, uint32_t launch, uint32_t ray, double u[3]);
void orf_ray_start(const orf_mesh* m, uint32_t seed, uint32_t launch, uint32_t rays, uint32_t ray,
                   uint32_t n_lambda, orf_ray* out);

/* One ray: returns 0 (left the mesh), 1 (dropped).  scores: [2 * n_points]
 * of this ray's batch; segments counts scored segments. */
int orf_walk(const orf_mesh* m, double sigma_a, double sigma_e, uint32_t cell, const double p0[3],
             const double u[3], double* scores, uint32_t* visits, uint32_t* dropped, uint64_t* segments);

/* Walk from a point with a forbidden entry face and a source weight; on
 * leaving the mesh `out` names the boundary face and the incident weight. */
int orf_walk_from(const orf_mesh* m, double sigma_a, double sigma_e, uint32_t cell, const double p0[3],
                  const double u[3], int32_t forbidden, double weight, double* scores, uint32_t* visits,
                  uint32_t* dropped, uint64_t* segments, orf_exit* out);

/* One launch of `rays` rays: accumulates into scores [8 * 2 * n_points],
 * visits and dropped [n_cells]; batch_counts [8] gain each batch's rays. */
void orf_trace(const orf_mesh* m, const double* sigma_a, const double* sigma_e, uint32_t n_lambda, uint32_t seed,
               uint32_t launch, uint32_t rays, double* scores, uint32_t* visits, uint32_t* dropped,
               uint32_t* batch_counts, uint64_t* segments, int threads);

void orf_finalize(const orf_mesh* m, const double* scores, const uint32_t* batch_counts, uint64_t rays,
                  const uint32_t* dropped, double sigma_a, double sigma_e, float* phi, double* se, double* rse,
                  double* dndt);

/* Adaptive evaluation (rays per target from min to max over `steps`
 * geometric increases; stop when every cell has rse^2 <= threshold). */
int orf_evaluate(const orf_mesh* m, const double* sigma_a, const double* sigma_e, uint32_t n_lambda,
                 double max_sigma_a, double max_sigma_e, uint32_t min_rays, uint32_t max_rays, uint32_t steps,
                 double threshold, uint32_t seed, float* phi, double* se, double* rse, double* dndt,
                 uint32_t* visits, uint32_t* dropped, uint32_t* convergence_rays, uint64_t* total_rays,
                 int threads);

/* ---------------------------------------------------------- reflections
 * Surface reservoir method (include/core/forwardSrm.hpp:120-366,
 * accumulation.hpp:407-834): rays that leave through a reflective boundary
 * face leave their reflected direction and weight in that face's weighted
 * reservoir (per RSE batch); each later pass draws faces by reservoir weight
 * and slots by slot weight and restarts from the face centroid, until the
 * remaining weight falls below the tolerance, stops changing or grows
 * `divergence_streak` passes in a row. */
enum { ORF_SRM_DISABLED = 0, ORF_SRM_CONVERGED = 1, ORF_SRM_STABLE = 2, ORF_SRM_MAX_PASSES = 3,
       ORF_SRM_DIVERGED = 4 };

typedef struct {
    uint32_t n_faces;          /* reflective boundary faces (surface id > 0, no neighbour) */
    uint32_t* face;            /* cell * 4 + local face */
    int32_t* face_index;       /* 4 per cell: reflective face index or -1 */
    double* centroid;          /* 3 per face */
    double* normal;            /* 3 per face, outward unit */
    double* reflectivity;      /* per face */
    double* n_in;              /* per face, 0 = no total-internal-reflection test */
    double* n_out;
} orf_surfaces;

typedef struct {
    uint32_t slots;            /* reservoir slots per face */
    uint32_t max_passes;
    uint32_t divergence_streak;
    double tolerance;
    /* results of the last launch, per batch */
    int32_t status[ORF_BATCHES];
    uint32_t passes[ORF_BATCHES];
    double remaining[ORF_BATCHES];
    /* order in which rays deposit into the reservoirs: ray k of a pass is
     * processed at position (k * order_stride) mod count (0 or 1: ray order).
     * The device deposits in atomic arrival order; a stride coprime to the
     * ray count reproduces the effect of another order on the CPU. */
    uint32_t order_stride;
} orf_srm;

int orf_surfaces_build(const orf_mesh* m, const int32_t* face_surface, uint32_t n_surfaces,
                       const float* reflectivity, const float* n_in, const float* n_out, orf_surfaces* out);
void orf_surfaces_free(orf_surfaces* s);
double orf_reflectance(const orf_surfaces* s, uint32_t f, const double dir[3]);

/* One launch with reflections (single-threaded, rays in order): primary rays,
 * then reflected passes per batch. */
void orf_trace_srm(const orf_mesh* m, const orf_surfaces* sf, orf_srm* srm, const double* sigma_a,
                   const double* sigma_e, uint32_t n_lambda, uint32_t seed, uint32_t launch, uint32_t rays,
                   double* scores, uint32_t* visits, uint32_t* dropped, uint32_t* batch_counts, uint64_t* segments);

int orf_evaluate_srm(const orf_mesh* m, const orf_surfaces* sf, orf_srm* srm, const double* sigma_a,
                     const double* sigma_e, uint32_t n_lambda, double max_sigma_a, double max_sigma_e,
                     uint32_t min_rays, uint32_t max_rays, uint32_t steps, double threshold, uint32_t seed,
                     float* phi, double* se, double* rse, double* dndt, uint32_t* visits, uint32_t* dropped,
                     uint32_t* convergence_rays, uint64_t* total_rays);

#endif
