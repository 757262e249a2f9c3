/* ewald-3f.h -- iterative F resistance, mobility and mixed problems
 * (drop-in for the reference's ewald-3f.h without the lubrication
 * variants).  The resistance and mixed problems are solved by restarted
 * GMRES over the GPU operator with the parameters kept by stokes_set_iter
 * (sys->it: restart, max, eps on |b - A x| / |b|; defaults 30, 2000,
 * 1e-12); the mobility problem is one product.  A wrong sys->version is
 * reset to F with a message, as the reference does. */
#ifndef LIBSTOKES_HIP_EWALD_3F_H
#define LIBSTOKES_HIP_EWALD_3F_H

#include "stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces ewald-3f.c:43-71: natural resistance problem in the labo frame,
 * u[3np] (shifted by the imposed flow first) in, f[3np] out */
void solve_res_3f (struct stokes *sys, const double *u, double *f);

/* replaces ewald-3f.c:81-98: the same in the fluid-rest frame, M f = u */
void solve_res_3f_0 (struct stokes *sys, const double *u, double *f);

/* replaces ewald-3f.c:109-130: natural mobility problem, u = M f shifted
 * to the labo frame; f[3np] in, u[3np] out */
void solve_mob_3f (struct stokes *sys, const double *f, double *u);

/* replaces ewald-3f.c:268-321: mobility problem with fixed particles
 * (nm..np-1 of sys): f[3nm], uf[3nf] (labo frame) in; u[3nm] (labo frame),
 * ff[3nf] out.  Unknowns (U_m, F_f), A x = [U_m, 0_f] - M [0_m, F_f],
 * b = M [F_m, 0_f] - [0_m, U_f] (ewald-3f.c:144-258).  As the reference
 * does (ewald-3f.c:301), uf is shifted with the positions of particles
 * 0..nf-1.  nm == np is the mobility problem */
void solve_mix_3f (struct stokes *sys, const double *f, const double *uf,
                   double *u, double *ff);

#ifdef __cplusplus
}
#endif
#endif
