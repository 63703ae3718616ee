/*
 * oracle.h -- CPU restatement of the reference's per-step liquid solver. TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker of the HIP solver. Only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it; the product library never links or calls it.
 *
 * Each function follows the reference kernel it restates (file:line in oracle.c). Arithmetic is FP32 where the
 * reference is FP32, evaluated in the written order with FMA contraction off (the library is built with
 * -ffp-contract=off), so it is bit-comparable with the HIP kernels that use the same order. The pressure solve
 * keeps FP32 vectors with FP64 dot products and alpha/beta, and a separate FP64 solve serves as the accuracy judge.
 */
#ifndef SLOSH_ORACLE_H_
#define SLOSH_ORACLE_H_

#include "../include/slosh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct osim osim;

osim *osim_create(const slosh_settings *s);
void osim_destroy(osim *o);
int osim_emit(osim *o, const slosh_emit_request *r);
int osim_set_colliders(osim *o, const slosh_collider *c, uint32_t n);
int osim_set_manipulators(osim *o, const slosh_manipulator *m, uint32_t n);
int osim_write_particles(osim *o, const slosh_particle *p, uint32_t n);
int osim_write_pressure(osim *o, const float *dense);
/* Writes the density-projection potential q (SLOSH_FLAG_DENSITY_PROJECTION), kept by a step with SOLVER_NONE. */
int osim_write_density_q(osim *o, const float *dense);
/* One step (stop_after = SLOSH_STAGE_*). */
int osim_step_until(osim *o, float dt, const float g[3], int stop_after);
uint32_t osim_num_particles(const osim *o);
int osim_read_particles(const osim *o, slosh_particle *dst, uint32_t cap);
/* Dense x-fastest copy of a field (same field ids and plane layout as slosh_read_grid). */
int osim_read_grid(const osim *o, int field, void *dst, uint64_t bytes);
int osim_get_results(const osim *o, slosh_results *out);

/* FP64 solve of the current pressure system (operator of the last step's classification) to rel_tol with an FP64
 * multigrid-preconditioned CG started from zero. Writes the dense solution (0 outside fluid cells), returns the
 * iteration count or -1. */
int osim_solve_fp64(const osim *o, double rel_tol, int max_iter, double *dense_out, double *rel_res_out);
/* ||b - A p||_2 / ||b||_2 of a dense pressure field p (NULL: the solver's own FP64 pressure) against the current
 * system, in FP64. */
double osim_true_residual(const osim *o, const float *dense_p);
/* Same for an FP64 dense field. */
double osim_true_residual_d(const osim *o, const double *dense_p);

/* The last step's collider coupling sums, 16 int32 per collider in 1/1024 units (the layout of
 * slosh_get_collider_feedback before decoding); copies min(count, cap) colliders, returns the count. */
int osim_coupling_raw(const osim *o, int32_t *dst, uint32_t cap);

/* Threads of the parallel (OpenMP) loops; results are the same for any count. */
void oracle_set_threads(int n);
int oracle_threads(void);

/* Building blocks exposed for known-answer tests. */
uint32_t oracle_pcg_hash(uint32_t v);
float oracle_hash_to_unit(uint32_t h);
float oracle_collider_sdf(const slosh_collider *c, const float p[3]);
float oracle_ghost_theta(float rho_fluid, float rho_air);
/* Unit normal of a collider's surface by central differences of its SDF with step eps (SloshCommon.ush:197-205). */
void oracle_collider_normal(const slosh_collider *c, const float p[3], float eps, float n[3]);
/* Velocity of a collider's material point p: v + w x (p - centre) (SloshCommon.ush:232-235). */
void oracle_collider_velocity(const slosh_collider *c, const float p[3], float v[3]);
/* The collider response G2P applies to one particle (position x, velocity nv, updated in place). */
void oracle_collider_response(const slosh_collider *cols, uint32_t ncol, float radius, float friction,
                              float cell_size, float x[3], float nv[3]);

#ifdef __cplusplus
}
#endif
#endif
