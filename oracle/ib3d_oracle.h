/*
 * ib3d_oracle.h -- CPU restatement of the IB3D per-timestep hot path.
 *
 * TEST INFRASTRUCTURE.  This library is the correctness oracle: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and only as the checker.  The product path (the HIP library) never
 * links or calls it.
 *
 * Every routine restates one reference subroutine expression by expression
 * (same operand order, no contraction, IEEE double) so results are bitwise
 * comparable with the reference built from its own sources.  Arrays use the
 * reference's Fortran layout: column-major, 1-based, padded fields of
 * (nx+2)(ny+2)(nz+2) doubles with i fastest.
 */
#ifndef IB3D_ORACLE_H
#define IB3D_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

orc_state *orc_create(int64_t nx, int64_t ny, int64_t nz);
void orc_free(orc_state *s);
const char *orc_last_error(orc_state *s);

/* scalars of readInput (subReadInput.f90:22-24) */
int orc_set_params(orc_state *s, double re, double dt_order, double eps1, int64_t inor);
/* node coordinates read from gridx/y/z.dat: nx+1, ny+1, nz+1 values */
int orc_set_grid(orc_state *s, const double *gx, const double *gy, const double *gz);
/* surface mesh: nodes (n,3) row-major, connectivity (m,3) 1-based row-major */
int orc_set_mesh(orc_state *s, int64_t n_nodes, const double *nodes, int64_t n_elems, const int64_t *conn);

int orc_shift_nodes_initial(orc_state *s, double xs, double ys, double zs);
int orc_compute_surface_norm(orc_state *s);
int orc_tagging(orc_state *s, const int64_t window[6]);
int orc_cell_count(orc_state *s);
/* moving body: prescribed displacement, selective retagging, wall motion */
int orc_shift_nodes_moving(orc_state *s, double xs, double ys, double zs, double dx, double dy, double dz);
int orc_selective_retagging(orc_state *s);
int orc_set_wall_motion(orc_state *s, const double vel[3], const double acc[3]);
/* per-body force coefficients of the last stress call (4 per body) */
int orc_set_bodies(orc_state *s, const int32_t *body, int32_t nbody);
int orc_stress_bodies(orc_state *s, double *out);
int orc_compute_norm_distance(orc_state *s);
int orc_initial_conditions(orc_state *s);
int orc_coefficient_matrix(orc_state *s);
int orc_non_uni_coeff(orc_state *s);

int orc_ns_momentum(orc_state *s);
int orc_velocity_bc(orc_state *s);
int orc_solid_cell_bc(orc_state *s);
int orc_velocity_forcing(orc_state *s);
int orc_pressure_forcing(orc_state *s);
int orc_compute_div(orc_state *s, double *divmax);
int orc_redblack_sor(orc_state *s, double epsi, int64_t *isum, double *derr2);
int orc_correct_pressure(orc_state *s);
int orc_correct_velocity(orc_state *s);
int orc_poisson_solver(orc_state *s, int64_t *niter, double *divmax, double *derr_std);
/* first != 0: stressCal1 (finds and caches stencil cells), else stressCal2.
 * out[4] = {Cd_visc, Cd_press, Cl_visc, Cl_press} */
int orc_stress(orc_state *s, int first, double out[4]);

/* one full time step in the order of main.f90:152-174 (stress included) */
int orc_step(orc_state *s, int first, int64_t *niter, double *divmax, double *derr_std,
             double coef[4]);

/* ---- north-star Poisson modes (no reference counterpart) ---------------- */
/* FP64 L2 norms over fluid cells of r = b/dt - A pc and of b/dt */
int orc_residual_l2(orc_state *s, double *rnorm, double *bnorm);
/* RB-SOR with a chosen omega from the current pc, checking ||r||/||b/dt||
 * every `check` sweeps; stops at tol or max_sweeps. */
int orc_sor_l2(orc_state *s, double omega, double tol, int64_t max_sweeps, int64_t check,
               int64_t *sweeps, double *relres);

/* raw access to a named array (see the table in ib3d_oracle.c):
 * *type = 8 for double, 64 for int64 */
int orc_array(orc_state *s, const char *name, void **ptr, int64_t *count, int *type);
/* counts: ib, fluid, solid, red, black, ibElems, ibNodes */
int orc_counts(orc_state *s, int64_t out[7]);

#ifdef __cplusplus
}
#endif
#endif
