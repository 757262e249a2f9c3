/* ewald-3fts.h -- iterative FTS resistance, mobility and mixed problems
 * (drop-in for the reference's ewald-3fts.h, with the lubrication
 * variants), solved by restarted GMRES over the GPU operator.  The
 * iteration parameters are those kept by stokes_set_iter (sys->it):
 * restart, max and eps (relative residual |b - A x| / |b|); defaults 30,
 * 2000, 1e-12 when none were set. */
#ifndef LIBSTOKES_HIP_EWALD_3FTS_H
#define LIBSTOKES_HIP_EWALD_3FTS_H

#include "stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces ewald-3fts.c:50-80: natural resistance problem in the
 * fluid-rest frame, M (f, t, s) = (u, o, e); u[3np], o[3np], e[5np] in,
 * f[3np], t[3np], s[5np] out */
void solve_res_3fts_0 (struct stokes *sys, const double *u, const double *o,
                       const double *e, double *f, double *t, double *s);

/* replaces ewald-3fts.c:94-143: the same in the labo frame (u, o, e shifted
 * by the imposed flow sys->Ui, Oi, Ei first) */
void solve_res_3fts (struct stokes *sys, const double *u, const double *o,
                     const double *e, double *f, double *t, double *s);

/* replaces ewald-3fts.c:263-305: natural mobility problem in the fluid-rest
 * frame; f[3np], t[3np], e[5np] in, u[3np], o[3np], s[5np] out (unknowns
 * (U, O, S), A x = (U, O, 0) - M (0, 0, S), b = M (F, T, 0) - (0, 0, E),
 * ewald-3fts.c:145-253).  iter: NULL (use sys->it) or a parameter set */
void solve_mob_3fts_0 (struct stokes *sys, struct iter *iter, const double *f,
                       const double *t, const double *e, double *u, double *o,
                       double *s);

/* replaces ewald-3fts.c:315-352: the same in the labo frame (e shifted
 * before, u and o shifted back after) */
void solve_mob_3fts (struct stokes *sys, const double *f, const double *t,
                     const double *e, double *u, double *o, double *s);

/* replaces ewald-3fts.c:531-593: mobility problem with fixed particles
 * (particles nm..np-1 of sys), fluid-rest frame: f, t[3nm], e[5nm], uf,
 * of[3nf], ef[5nf] in; u, o[3nm], s[5nm], ff, tf[3nf], sf[5nf] out
 * (ewald-3fts.c:363-525); nm == np is the mobility problem */
void solve_mix_3fts_0 (struct stokes *sys, struct iter *iter, const double *f,
                       const double *t, const double *e, const double *uf,
                       const double *of, const double *ef, double *u,
                       double *o, double *s, double *ff, double *tf,
                       double *sf);

/* replaces ewald-3fts.c:603-667: the same in the labo frame.  As the
 * reference does (ewald-3fts.c:652), the fixed particles' velocities are
 * shifted with the positions of particles 0..nf-1 */
void solve_mix_3fts (struct stokes *sys, const double *f, const double *t,
                     const double *e, const double *uf, const double *of,
                     const double *ef, double *u, double *o, double *s,
                     double *ff, double *tf, double *sf);

/* natural resistance / mobility problems with lubrication (ewald-3fts.c:
 * 676-719, 733-770; 923-951, 966-1000): _0 in the fluid-rest frame, the
 * others shift the known velocities / strains to it and the velocities
 * back; GMRES on the device over M and the lubrication correction */
void solve_res_lub_3fts_0 (struct stokes *sys, const double *u,
                           const double *o, const double *e, double *f,
                           double *t, double *s);
void solve_res_lub_3fts (struct stokes *sys, const double *u, const double *o,
                         const double *e, double *f, double *t, double *s);
void solve_mob_lub_3fts_0 (struct stokes *sys, const double *f,
                           const double *t, const double *e, double *u,
                           double *o, double *s);
void solve_mob_lub_3fts (struct stokes *sys, const double *f, const double *t,
                         const double *e, double *u, double *o, double *s);

/* replaces ewald-3fts.c:1192-1234 and 1255-1306: the mixed problem with
 * lubrication, fixed particles nm..np-1.  Unknowns (U, O, S)_mobile,
 * (F, T, S)_fixed by GMRES on A x = (I + M L) [(U, O, 0)_m, 0_f] -
 * M [(0, 0, S)_m, (F, T, S)_f] with b = M [(F, T, 0)_m, 0_f] - (I + M L)
 * [(0, 0, E)_m, (U, O, E)_f] (ewald-3fts.c:1015-1190).  _0: fluid-rest
 * frame; the other shifts uf, of, ef with the positions of particles
 * 0..nf-1 and e with those of 0..nm-1, as the reference does
 * (ewald-3fts.c:1288-1291), and u, o back.  nm == np is the mobility
 * problem with lubrication */
void solve_mix_lub_3fts_0 (struct stokes *sys, const double *f, const double *t,
                           const double *e, const double *uf, const double *of,
                           const double *ef, double *u, double *o, double *s,
                           double *ff, double *tf, double *sf);
void solve_mix_lub_3fts (struct stokes *sys, const double *f, const double *t,
                         const double *e, const double *uf, const double *of,
                         const double *ef, double *u, double *o, double *s,
                         double *ff, double *tf, double *sf);

/* imposed-flow shifts of particles 0..n-1: replace f.c:569-610
 * (shift_labo_to_rest_U), f.c:621-660 (shift_rest_to_labo_U), ft.c:901-920
 * (shift_labo_to_rest_O), ft.c:935-955 (shift_rest_to_labo_O) and
 * fts.c:2303-2320 (shift_labo_to_rest_E) */
void shift_labo_to_rest_U (struct stokes *sys, int n, const double *u,
                           double *u0);
void shift_rest_to_labo_U (struct stokes *sys, int n, double *u);
void shift_labo_to_rest_O (struct stokes *sys, int n, const double *o,
                           double *o0);
void shift_rest_to_labo_O (struct stokes *sys, int n, double *o);
void shift_labo_to_rest_E (struct stokes *sys, int n, const double *e,
                           double *e0);

#ifdef __cplusplus
}
#endif
#endif
