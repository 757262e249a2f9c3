/* ewald-3f.h -- iterative F resistance, mobility and mixed problems,
 * without and with lubrication (drop-in for the reference's ewald-3f.h).  The resistance and mixed problems are solved by restarted
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

/* with lubrication (the correction L of calc_lub_3f, lub.h).  Replace
 * ewald-3f.c:373-406 (solve_res_lub_3f_0: M f = (I + M L) u, GMRES on M)
 * and 334-362 (solve_res_lub_3f: u shifted to the fluid-rest frame first) */
void solve_res_lub_3f_0 (struct stokes *sys, const double *u, double *f);
void solve_res_lub_3f (struct stokes *sys, const double *u, double *f);

/* replaces ewald-3f.c:410-436: y = (I + M L) x, n = 3 np, user_data the
 * struct stokes */
void atimes_mob_lub_3f (int n, const double *x, double *y, void *user_data);

/* replaces ewald-3f.c:446-481: (I + M L) u = M f by GMRES, u shifted to
 * the labo frame */
void solve_mob_lub_3f (struct stokes *sys, const double *f, double *u);

/* replaces ewald-3f.c:646-705: fixed particles nm..np-1; unknowns (U_m,
 * F_f) of A x = (I + M L) [U_m, 0_f] - M [0_m, F_f] = M [F_m, 0_f] -
 * (I + M L) [0_m, U_f] (ewald-3f.c:495-644); uf shifted with the positions
 * of particles 0..nf-1 as the reference does (ewald-3f.c:685).  nm == np
 * is solve_mob_lub_3f */
void solve_mix_lub_3f (struct stokes *sys, const double *f, const double *uf,
                       double *u, double *ff);

#ifdef __cplusplus
}
#endif
#endif
