/* ewald-3fts-matrix.h -- dense FP64 mobility solves of the FTS version
 * (drop-in for the reference's ewald-3fts-matrix.h, fluid-rest frame). */
#ifndef LIBSTOKES_HIP_EWALD_3FTS_MATRIX_H
#define LIBSTOKES_HIP_EWALD_3FTS_MATRIX_H

#include "stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces ewald-3fts-matrix.c:570-652: given f[3np], t[3np], e[5np]
 * (fluid-rest frame) return u[3np], o[3np], s[5np]; dense matrix and solve
 * on the GPU in FP64 */
void solve_mob_3fts_matrix_0 (struct stokes *sys, const double *f,
                              const double *t, const double *e, double *u,
                              double *o, double *s);
/* replaces ewald-3fts-matrix.c:1192-1290: particles [0, nm) mobile with
 * f, t, e [nm*3, nm*3, nm*5]; particles [nm, np) fixed with uf, of, ef
 * [nf*3, nf*3, nf*5]; returns u, o, s of the mobile and ff, tf, sf of the
 * fixed particles */
void solve_mix_3fts_matrix_0 (struct stokes *sys, const double *f,
                              const double *t, const double *e,
                              const double *uf, const double *of,
                              const double *ef, double *u, double *o,
                              double *s, double *ff, double *tf, double *sf);

/* the same problems with lubrication by dense FP64 solves on the GPU
 * (ewald-3fts-matrix.c:264-330, 332-380; 731-838, 840-880) */
void solve_res_lub_3fts_matrix_0 (struct stokes *sys, const double *u,
                                  const double *o, const double *e, double *f,
                                  double *t, double *s);
void solve_res_lub_3fts_matrix (struct stokes *sys, const double *u,
                                const double *o, const double *e, double *f,
                                double *t, double *s);
void solve_mob_lub_3fts_matrix_0 (struct stokes *sys, const double *f,
                                  const double *t, const double *e, double *u,
                                  double *o, double *s);
void solve_mob_lub_3fts_matrix (struct stokes *sys, const double *f,
                                const double *t, const double *e, double *u,
                                double *o, double *s);

#ifdef __cplusplus
}
#endif
#endif
