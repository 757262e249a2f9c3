/* ewald-3fts.h -- iterative FTS mobility solve (drop-in for the
 * reference's ewald-3fts.h, fluid-rest frame). */
#ifndef LIBSTOKES_HIP_EWALD_3FTS_H
#define LIBSTOKES_HIP_EWALD_3FTS_H

#include "stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces ewald-3fts.c:263-305: given f[3np], t[3np], e[5np] return
 * u[3np], o[3np], s[5np] by restarted GMRES over the GPU operator
 * (unknowns (U, O, S), A x = (U, O, 0) - M (0, 0, S), b = M (F, T, 0) -
 * (0, 0, E), ewald-3fts.c:145-253).  iter: NULL (use sys->it) or sys->it;
 * max, restart and eps of the parameters are used (defaults 2000, 30,
 * 1e-12: relative residual). */
void solve_mob_3fts_0 (struct stokes *sys, struct iter *iter, const double *f,
                       const double *t, const double *e, double *u, double *o,
                       double *s);

#ifdef __cplusplus
}
#endif
#endif
