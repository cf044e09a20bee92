/*
 * LULESH oracle: a serial / OpenMP CPU restatement of the reference HPX
 * Lagrange leapfrog path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this library: it is used by tests/, by __graft_entry__.smoke() as the
 * checker, and by bench.py's cpu_baseline leg (as the timed CPU port).
 *
 * Every function follows the reference arithmetic order of
 * lulesh.cc / lulesh-init.cc (file:line cited at each definition).  The
 * library is compiled with -ffp-contract=off so that the expression shapes
 * of the reference map one-to-one onto IEEE operations.
 */
#ifndef LULESH_ORACLE_H
#define LULESH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* lulesh.h:30 */
enum { ORA_VOLUME_ERROR = -1, ORA_QSTOP_ERROR = -2 };

typedef struct ora_domain ora_domain;

/* Domain::Domain, lulesh-init.cc:10-183 (single rank, tp = 1). */
ora_domain *ora_create(int32_t nx, int32_t nr, int32_t balance, int32_t cost);
void ora_destroy(ora_domain *d);

/* Field access by reference accessor name ("x", "e", "nodelist", ...).
 * Returns a pointer to the live array and its length; NULL if unknown. */
double *ora_field_f64(ora_domain *d, const char *name, int64_t *len);
int32_t *ora_field_i32(ora_domain *d, const char *name, int64_t *len);

/* Scalars: time, deltatime, dtcourant, dthydro, stoptime, dtmax, ... */
double ora_get_scalar(ora_domain *d, const char *name);
void ora_set_scalar(ora_domain *d, const char *name, double value);
int32_t ora_cycle(ora_domain *d);
void ora_set_cycle(ora_domain *d, int32_t cycle);
int32_t ora_num_reg(ora_domain *d);
int32_t ora_reg_elem_size(ora_domain *d, int32_t r);
int32_t *ora_reg_elem_list(ora_domain *d, int32_t r);
int32_t ora_reg_rep(ora_domain *d, int32_t r);

/* Ensemble member: scale the deposit of element 0 by `factor` and recompute
 * the initial time step as the constructor does (lulesh-init.cc:172-181). */
void ora_scale_energy(ora_domain *d, double factor);

/* Threads used by the OpenMP loops (1 = serial).  Results do not depend on it. */
void ora_set_threads(int32_t nthreads);
int32_t ora_max_threads(void);

/* TimeIncrement, lulesh.cc:182-229 */
void ora_time_increment(ora_domain *d);

/* Phases of LagrangeLeapFrogWithTasks (lulesh.cc:1977-2240).  Each returns
 * 0 or ORA_VOLUME_ERROR where the reference calls exit(VolumeError). */
/* Phase 1: stress + hourglass corner forces (lulesh.cc:2017-2038).  If the
 * corner pointers are non-NULL they receive the six 8*numElem arrays in the
 * reference layout [8*elem + corner]: sx,sy,sz,hx,hy,hz. */
int ora_calc_corner_forces(ora_domain *d, double *sx, double *sy, double *sz,
                           double *hx, double *hy, double *hz);
/* Phases 1-3: forces, corner->node gather, acceleration, BCs, velocity and
 * position (lulesh.cc:2017-2110). */
int ora_lagrange_nodal(ora_domain *d);
/* Phase 4: kinematics, volume update, monotonic-Q gradients (2112-2136). */
int ora_lagrange_elements(ora_domain *d);
/* Phase 5: per-region monotonic Q, EOS x rep, sound speed (2138-2202). */
int ora_apply_material_properties(ora_domain *d);
/* Phase 6: Courant / hydro constraints (2204-2239). */
void ora_calc_time_constraints(ora_domain *d);
/* The whole cycle (all six phases). */
int ora_lagrange_leapfrog(ora_domain *d);

/* The reference main loop (lulesh.cc:2374-2386): while time < stoptime and
 * cycle < its: TimeIncrement + LagrangeLeapFrog.  Returns 0 or an error. */
int ora_run(ora_domain *d, int32_t its);

/* VerifyAndWriteFinalOutput numbers (lulesh-util.cc:60-115):
 * out[0] = e(0), out[1] = MaxAbsDiff, out[2] = TotalAbsDiff, out[3] = MaxRelDiff */
void ora_report(ora_domain *d, double out[4]);

/* Element-level helpers exported for closed-form unit tests. */
double ora_calc_elem_volume(const double x[8], const double y[8], const double z[8]);
void ora_calc_elem_shape_function_derivatives(const double x[8], const double y[8],
                                              const double z[8], double b[24], double *volume);
void ora_calc_elem_node_normals(double pf[24], const double x[8], const double y[8],
                                const double z[8]);
void ora_calc_elem_volume_derivative(double dv[24], const double x[8], const double y[8],
                                     const double z[8]);
double ora_calc_elem_characteristic_length(const double x[8], const double y[8],
                                           const double z[8], double volume);
void ora_calc_elem_velocity_gradient(const double xv[8], const double yv[8],
                                     const double zv[8], const double b[24], double detJ,
                                     double d[6]);
/* Hourglass force of one element (lulesh.cc:958-1071): hg[24] = (hgfx,hgfy,hgfz). */
void ora_calc_elem_fb_hourglass(const double x[8], const double y[8], const double z[8],
                                const double xd[8], const double yd[8], const double zd[8],
                                double determ, double ss, double mass, double hgcoef,
                                double hg[24]);
/* EOS of one element (lulesh.cc:1692-1924 restricted to one element).
 * in:  e, delv, p, q, qq, ql, vnewc;  out: e_new, p_new, q_new, ss. */
void ora_eval_eos_elem(const double in[7], double out[4]);

#ifdef __cplusplus
}
#endif
#endif
