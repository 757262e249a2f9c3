/* ewald-3ft.h -- iterative FT resistance, mobility and mixed problems
 * (drop-in for the reference's ewald-3ft.h, with the lubrication
 * variants).  The resistance and mixed problems are solved by restarted
 * GMRES over the GPU operator with the parameters kept by stokes_set_iter
 * (sys->it: restart, max, eps on |b - A x| / |b|; defaults 30, 2000,
 * 1e-12); the mobility problem is one product.  Arrays are 3 per particle
 * (u, o, f, t separately), as in the reference.  A wrong sys->version is
 * reset to FT with a message, as the reference does. */
#ifndef LIBSTOKES_HIP_EWALD_3FT_H
#define LIBSTOKES_HIP_EWALD_3FT_H

#include "stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces ewald-3ft.c:46-79: natural resistance problem in the labo
 * frame, u, o[3np] (shifted by the imposed flow first) in, f, t[3np] out */
void solve_res_3ft (struct stokes *sys, const double *u, const double *o,
                    double *f, double *t);

/* replaces ewald-3ft.c:91-122: the same in the fluid-rest frame */
void solve_res_3ft_0 (struct stokes *sys, const double *u, const double *o,
                      double *f, double *t);

/* replaces ewald-3ft.c:135-169: natural mobility problem, (u, o) = M (f, t)
 * shifted to the labo frame */
void solve_mob_3ft (struct stokes *sys, const double *f, const double *t,
                    double *u, double *o);

/* replaces ewald-3ft.c:323-382: mobility problem with fixed particles
 * (nm..np-1 of sys): f, t[3nm], uf, of[3nf] (labo frame) in; u, o[3nm]
 * (labo frame), ff, tf[3nf] out.  Unknowns (U, O)_m, (F, T)_f, as
 * solve_mix_3f with six values per particle (ewald-3ft.c:185-308).  As the
 * reference does (ewald-3ft.c:359-360), uf and of are shifted with the
 * positions of particles 0..nf-1.  nm == np is the mobility problem */
void solve_mix_3ft (struct stokes *sys, const double *f, const double *t,
                    const double *uf, const double *of, double *u, double *o,
                    double *ff, double *tf);

/* with lubrication (calc_lub_3ft, lub.h), six values per particle as the F
 * forms: replace ewald-3ft.c:441-485 (solve_res_lub_3ft_0), 395-428
 * (solve_res_lub_3ft), 527-569 (solve_mob_lub_3ft: (I + M L) (u, o) = M
 * (f, t), velocities shifted to the labo frame) and 747-805
 * (solve_mix_lub_3ft: fixed particles nm..np-1, uf and of shifted with the
 * positions of particles 0..nf-1 as ewald-3ft.c:782-783 do) */
void solve_res_lub_3ft_0 (struct stokes *sys, const double *u, const double *o,
                          double *f, double *t);
void solve_res_lub_3ft (struct stokes *sys, const double *u, const double *o,
                        double *f, double *t);
void solve_mob_lub_3ft (struct stokes *sys, const double *f, const double *t,
                        double *u, double *o);
void solve_mix_lub_3ft (struct stokes *sys, const double *f, const double *t,
                        const double *uf, const double *of, double *u, double *o,
                        double *ff, double *tf);

#ifdef __cplusplus
}
#endif
#endif
