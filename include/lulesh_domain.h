/*
 * lulesh_domain.h -- C-ABI of the host side: Domain construction, the host
 * TimeIncrement and the final report.
 *
 * The C++ class behind the opaque handle keeps the reference Domain's
 * accessor names (lulesh.h:130-606); these entry points let a non-C++ caller
 * (the Python test and bench harness, or any FFI) build and inspect it.
 */
#ifndef LULESH_DOMAIN_H
#define LULESH_DOMAIN_H

#include <stdint.h>

#include "lulesh_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lulesh_domain lulesh_domain;

/* new Domain(numRanks=1, col=row=plane=0, nx, tp=1, nr, balance, cost):
 * replaces lulesh.cc:2364 -> lulesh-init.cc:10-183.  NULL on bad input. */
lulesh_domain *lulesh_domain_create(int32_t nx, int32_t nr, int32_t balance, int32_t cost);
void lulesh_domain_destroy(lulesh_domain *d);

/* Fill a plain-pointer view of the Domain's live arrays (for lulesh_hip_create). */
int lulesh_domain_get_view(lulesh_domain *d, lulesh_domain_view *out);
/* Copy the time state of a view back into the Domain (time, deltatime,
 * dtcourant, dthydro, cycle). */
int lulesh_domain_set_time_state(lulesh_domain *d, const lulesh_time_state *ts);
int lulesh_domain_get_time_state(lulesh_domain *d, lulesh_time_state *ts);

/* TimeIncrement, lulesh.cc:182-229, on the host Domain. */
void lulesh_time_increment(lulesh_domain *d);

/* LagrangeLeapFrog(Domain&): the cycle body of lulesh.cc:2378 run through the
 * device context (lulesh_hip_leapfrog with the Domain's deltatime); stores
 * dtcourant / dthydro into the Domain. */
int lulesh_lagrange_leapfrog(lulesh_domain *d, lulesh_ctx *ctx);

/* VerifyAndWriteFinalOutput numbers, lulesh-util.cc:60-97:
 * out = {e(0), MaxAbsDiff, TotalAbsDiff, MaxRelDiff}. */
void lulesh_verify_report(lulesh_domain *d, double out[4]);

/* Region repetition count of the EOS (lulesh.cc:2159-2169). */
int32_t lulesh_region_rep(int32_t reg, int32_t numReg, int32_t cost);

/* Scale the energy deposited in element 0 by `factor` and recompute the
 * initial time step from it, as the constructor does for the unscaled
 * deposit (lulesh-init.cc:172-181).  For ensembles of Sedov problems. */
int lulesh_domain_scale_energy(lulesh_domain *d, double factor);

/* The visualization dump of DumpToVisit (lulesh-viz.cc:27-197, called for
 * -v at lulesh.cc:2397-2400) from the Domain's current host arrays, as VTK
 * legacy unstructured-grid files (VisIt reads them; the reference's Silo
 * library is not used): mesh, regions, e/p/v/q per element and
 * speed/xd/yd/zd per node, time and cycle.  Writes
 * <dir>/lulesh_plot_c<cycle>.vtk, or for pieces > 1 (-f) that many slabs of
 * element planes, lulesh_plot_c<cycle>.<k>.vtk, listed in
 * lulesh_plot_c<cycle>.visit.  0 on success, -3 bad arguments, -4 a file
 * could not be written. */
int lulesh_domain_write_viz(lulesh_domain *d, const char *dir, int32_t pieces);

/* Parameters of a seeded ensemble (no counterpart in the reference): problem
 * sizes uniform in [smin, smax], region counts from {11, 16, 21}, energy
 * factors uniform in [0.5, 2), all drawn from one std::mt19937_64 stream
 * (the reference's libc rand() stream is left untouched). */
int lulesh_batch_params(uint64_t seed, int32_t nprob, int32_t smin, int32_t smax,
                        int32_t *sizes, int32_t *regions, double *factors);

#ifdef __cplusplus
}
#endif
#endif
