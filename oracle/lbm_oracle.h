/*
 * lbm_oracle.h - FP64 CPU restatement of the reference time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the CPU baseline of the benchmark; nothing in the product
 * (lbm_mutiwing_amd/) links, loads or calls it.
 *
 * The functions mirror include/lbm_hip.h one for one (orc_* for lbm_*) and
 * take the same lbm_config, so a test drives both with identical inputs.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include "../include/lbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

int orc_create(const lbm_config *cfg, orc_t **out);
int orc_destroy(orc_t *o);
int orc_reset(orc_t *o);
int orc_set_motion_table(orc_t *o, const double *rows, int32_t nrows,
                         int32_t ncols, double *end_time);
int orc_set_stream_segment(orc_t *o, const double *start_row,
                           const double *target_row, int32_t ncols);
int orc_build_stream_target_row(int32_t wing_count, const double *start_row,
                                double target_time, const double *delta_row,
                                double *target_row);
int orc_advance(orc_t *o, int32_t n_last, double *force_rows,
                double *motion_rows, double *row_sum, double *last_row,
                double *err);
int orc_get_state(orc_t *o, int32_t *n_next, int32_t *movenum, int32_t *flag2,
                  int32_t *row_len, int32_t *adjoint_n);
int orc_get_fields(orc_t *o, double *rho, double *u, double *v, double *pre);
int orc_get_populations(orc_t *o, double *f, double *f0);
int orc_set_populations(orc_t *o, const double *f, const double *f0);
/* Sum of f over every node and population, compensated (lbm_get_mass). */
int orc_get_mass(orc_t *o, double *mass);
int orc_get_markers(orc_t *o, double *xb1, double *yb1, double *xb2,
                    double *yb2);
int orc_get_ibm(orc_t *o, int32_t capacity, int32_t *adjoint_ij,
                int32_t *wing_id, double *wing_force, double *wing_moment);
int orc_set_body_rel_pos(orc_t *o, int32_t wing, double pos);
int orc_get_body_rel(orc_t *o, double *rel);
int orc_set_body_rel(orc_t *o, const double *rel);
int orc_set_dynamics(orc_t *o, const lbm_dynamics *d);
int orc_set_step(orc_t *o, int32_t n_next, int32_t movenum);
int orc_last_error(const orc_t *o, char *buf, size_t n);

/* Test hooks exposing single reference routines. */
void orc_mrt_matrices(double m[81], double minv[81]);  /* matm, nizhen */
void orc_weights(double w[9]);                          /* D2Q9model   */
double orc_delta_kernel(double disr);                   /* smooth=1 branch */
/* The rigid-body update of one step on wing 1's 11 body_rel fields. */
void orc_body_step(const lbm_dynamics *d, double scale, double uu, double fx, double mom, double *st);
/* IBM system of the last step: A (nseg x nseg, column-major), rhs (nseg x 2)
 * and solution comj (nseg x 2); nseg = wing_count*(points_per_wing-1). */
int orc_get_ibm_system(orc_t *o, double *a, double *rhs, double *sol);
int orc_collide_node(orc_t *o, const double *f, double rt, double ut, double vt,
                     double gfx, double gfy, double *out);
int orc_bc_node(orc_t *o, double rb, double ub, double vb, double rn, double un,
                double vn, const double *f0n, double *out);
int orc_get_ibm_nodes(orc_t *o, int32_t capacity, double *us, double *vs, double *u,
                      double *v, double *rho);
int orc_get_rates(orc_t *o, double *s);
/* LODI update of one far-field node (right = 0: the left boundary): rho, u,
 * v at the boundary node and the next two inward -> its updated rho, u, v. */
int orc_lodi_node(const orc_t *o, int32_t right, const double *r, const double *u, const double *v, double *out);
/* Threads used by the lattice loops (OpenMP); 1 when built without it. */
int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
