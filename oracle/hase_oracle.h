/*
 * hase_oracle.h - CPU restatement (FP64, OpenMP) of the importance-sampled
 * ASE ray tracer on the extruded triangle-prism mesh.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library:
 * only tests, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  The product never calls it and has no CPU fallback.
 *
 * The legacy backward estimator is retired in the reference (CHANGELOG.md:5-11,
 * 29-33; include/core/calcPhiAse.hpp is empty), so this restatement follows the
 * pieces that survive:
 *   - gain coefficient and cladding: include/kernels/forward/rayWalk.hpp:170-208
 *   - physical scaling phi = (nTot/tau) * flux: docs/source/theoryAndModel.rst:139-163
 *   - surface reflectance with total internal reflection:
 *     include/kernels/reflection.hpp:21-47, lookups include/core/mesh.hpp:301-332
 *   - prism arrays: pyInclude/geometry/core.py:671-774
 *   - adaptive ray schedule: include/core/types.hpp:333-354
 *   - relative standard error: src/core/calcForwardPhiAse.cpp:81-98
 *   - Philox4x32-10: Salmon et al., "Parallel random numbers: as easy as
 *     1, 2, 3", SC'11 (the reference uses alpaka3's Philox4x32x10,
 *     include/kernels/forward/accumulation.hpp:389).
 */
#ifndef HASE_ORACLE_H
#define HASE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t np, nt, nl;        /* points, triangles, levels */
    double h;                   /* thickness between planes */
    const double* px;           /* [np] */
    const double* py;           /* [np] */
    const uint32_t* tri;        /* [3nt] column-major, counter-clockwise */
    const int32_t* nbr;         /* [3nt] neighbour across edge k (-1 boundary) */
    const int32_t* fe;          /* [3nt] shared edge index inside the neighbour */
    const double* beta;         /* [nt*(nl-1)] level-major */
    const uint32_t* clad;       /* [nt] or NULL */
    uint32_t clad_number;
    double clad_alpha;          /* cm^-1 */
    const double* refl;         /* [2nt] or NULL: t*2 + {0 bottom, 1 top} */
    const double* nidx;         /* [4] or NULL */
} orc_mesh;

typedef struct {
    double beta_len;            /* sum beta_j * l_j over gain prisms */
    double gain_len;            /* sum l_j over gain prisms */
    double clad_len;            /* sum l_j over cladding prisms */
    double refl;                /* product of surface factors */
    double length;              /* total (unfolded) path length */
    uint32_t steps;             /* prisms visited */
    int32_t status;             /* 0 ok, 1 left the mesh, 2 step limit, 3 bad input */
} orc_walk;

typedef struct {
    uint32_t min_rays, max_rays;
    uint32_t n_lambda;
    const double* sigma_a;      /* [n_lambda] */
    const double* sigma_e;      /* [n_lambda] */
    double max_sigma_a, max_sigma_e;
    double mse_threshold;
    int32_t use_reflections;
    uint32_t max_reflections;
    double defensive;
    uint32_t max_repetitions, adaptive_steps;
    uint32_t sample_min, sample_max;   /* inclusive */
    uint32_t seed;
    double n_tot, tau;
    const double* beta_cells;   /* [np*nl] */
    int32_t threads;            /* OpenMP threads (<=0: runtime default) */
    const double* importance;   /* NULL: walked map; else an external proposal importance,
                                   [(s - sample_min) * R*P + k*P + j], caller prism order */
    int32_t fallback;           /* with importance: samples that miss the threshold after the
                                   first launch restart on the walked map */
    const uint32_t* sample_list;/* NULL: the range [sample_min, sample_max]; else these samples
                                   (each within the range), spread over the threads */
    uint32_t n_list;
} orc_params;

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u32_to_unit(uint32_t x);

void orc_image_point(const orc_mesh* m, const double p[3], uint32_t reflections,
                     uint32_t first_plane, double out[3]);
int orc_walk_ray(const orc_mesh* m, const double a[3], uint32_t prism, const double e[3],
                 uint32_t reflections, uint32_t first_plane, orc_walk* w);

uint32_t orc_adaptive_target(uint32_t min_rays, uint32_t max_rays, uint32_t steps, uint32_t k);
double orc_relative_standard_error(double s1, double s2, double n);

int orc_importance(const orc_mesh* m, const orc_params* prm, uint32_t sample, double* imp);
void orc_triangle_order(const orc_mesh* m, uint32_t* order);
int orc_proposal_prefix(const orc_mesh* m, const orc_params* prm, const double* imp,
                        const uint32_t* order, double* prefix);
void orc_ray_start(const orc_mesh* m, uint32_t seed, uint32_t sample, uint32_t launch,
                   uint32_t ray, uint32_t prism, double out[3]);
double orc_systematic_offset(uint32_t seed, uint32_t sample, uint32_t launch);

int orc_calc_phi_ase(const orc_mesh* m, const orc_params* prm, double* phi, double* mse,
                     uint32_t* total_rays, double* dndt, uint32_t* dropped, double* s1_out,
                     double* s2_out);

#ifdef __cplusplus
}
#endif

#endif
