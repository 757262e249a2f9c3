/* lub.h -- the lubrication correction (drop-in for the reference's lub.h,
 * two-body-res.h, minv-poly.h and the lubrication routines of f.h, ft.h,
 * fts.h).
 *
 * The correction of a near pair is the exact two-body resistance minus the
 * inverse of its two-body far-field (RPY) mobility, applied as the pair's
 * resistance blocks to (U, O, E).  Pairs: r > 0 and r < lubmax (equal
 * spheres) or s = 2 r / (a1 + a2) < lubmax (unequal), over the periodic
 * image cells -imax..imax; r (or s) below sqrt (lubmin2) is clamped.  FP64,
 * computed on the GPU.
 */
#ifndef LIBSTOKES_HIP_LUB_H
#define LIBSTOKES_HIP_LUB_H

#include "stokes.h"
#include "twobody.h"

#ifdef __cplusplus
extern "C" {
#endif

/* lub.c:42-73, 82-127: 0 when the pair takes the correction, 1 otherwise */
int cond_lub (const double *x1, const double *x2, double lubmax2);
int cond_lub_poly (const double *x1, const double *x2, double a1, double a2,
                   double lubmax2);

/* lub.c:164-368 (F), 370-574 (FT), 576-768 (FTS): y = L x for all
 * particles, x the velocities (U; U, O; U, O, E) and y the forces (F; F, T;
 * F, T, S), np * 3 / 6 / 11 values, host pointers.  Reads sys->pos, sys->a
 * (NULL: equal spheres of radius 1), sys->lubmin2, sys->lubmax,
 * sys->twobody_nmax and sys->twobody_lub (unequal spheres), the box and the
 * shear shift.  On failure prints and exits, as the reference's
 * allocations do. */
void calc_lub_3f (struct stokes *sys, const double *u, double *f);
void calc_lub_3ft (struct stokes *sys, const double *uo, double *ft);
void calc_lub_3fts (struct stokes *sys, const double *uoe, double *fts);

/* two-body-res.c:31: equal spheres, res[22] in the dimensional form at
 * radius 1 (the lubrication form at lambda = 1 to order 50) */
void scalar_two_body_res (double s, double *res);
/* f.c:183, ft.c:414, fts.c:1490: the inverse of the two-body far-field
 * mobility of equal spheres, (11, 12) pairs (4, 10 or 22 values) */
void scalar_minv_f (double s, double *scalar_f);
void scalar_minv_ft (double s, double *scalar_ft);
void scalar_minv_fts (double s, double *scalar_fts);
/* minv-poly.c:39-560: the same for radii aa, ab, (11, 12, 21, 22) per
 * function (8, 20 or 44 values), dimensional */
void scalars_minv_f_poly (double r, double aa, double ab, double *scalar_f);
void scalars_minv_ft_poly (double r, double aa, double ab, double *scalar_ft);
void scalars_minv_fts_poly (double r, double aa, double ab,
                            double *scalar_fts);
/* minv-poly.c:565-660: dimensional to SD form */
void scalars_res_poly_scale_SD (int version, double a1, double a2,
                                double *scalar);
/* minv-poly.c:669-757, 759-857: the correction's scalars, lub[22] (11, 12)
 * or lub[44] (11, 12, 21, 22), dimensional */
void scalars_lub_poly (int version, double r, double a1, double a2,
                       struct twobody_f *f12, int n, int flag_lub,
                       double *lub);
void scalars_lub_poly_full (int version, double r, double a1, double a2,
                            struct twobody_f *f12, struct twobody_f *f21,
                            int n, int flag_lub, double *lub);

/* lub-matrix.c:76-223 (F), 225-371 (FT), 373-507 (FTS): the dense
 * extracted-form matrix of L, (np D)^2 row-major, host pointer, as the
 * reference's (pair blocks with matrix_fts_ij's signs); built on the GPU */
void make_matrix_lub_3f (struct stokes *sys, double *mat);
void make_matrix_lub_3ft (struct stokes *sys, double *mat);
void make_matrix_lub_3fts (struct stokes *sys, double *mat);
/* the same for sys->version into device memory.  No counterpart in the
 * reference (host pointers only). */
int stokes_hip_make_matrix_lub_dev (struct stokes *sys, double *mat_dev);

/* Lubricated FTS solves with device pointers (11 values per particle, the
 * reference's particle-wise packing).  res: (U, O, E) -> (F, T, S) =
 * (M^-1 + L) (U, O, E) by GMRES on M with b = (I + M L) (U, O, E)
 * (solve_res_lub_3fts_0, ewald-3fts.c:676-719); mob: (F, T, E) -> (U, O,
 * S) by GMRES on A (u, o, s) = -M (0, 0, s) + (I + M L) (u, o, 0)
 * (solve_mob_lub_3fts_0, 923-951); stokes_set_iter's parameters.
 * matrix: kind 0 / 1 the same problems by a dense LU solve of the assembled
 * FP64 matrices (solve_res_lub_3fts_matrix_0, ewald-3fts-matrix.c:264-330;
 * solve_mob_lub_3fts_matrix_0, 731-838).  No counterpart in the reference
 * (host pointers only). */
int stokes_hip_solve_res_lub_3fts_dev (struct stokes *sys, const double *uoe_dev,
                                       double *fts_dev, int *iters,
                                       double *relres);
int stokes_hip_solve_mob_lub_3fts_dev (struct stokes *sys, const double *fte_dev,
                                       double *uos_dev, int *iters,
                                       double *relres);
int stokes_hip_solve_lub_3fts_matrix_dev (struct stokes *sys, int kind,
                                          const double *b_dev, double *x_dev);

/* the lubricated solves of sys->version (D = 3, 6, 11 values per particle)
 * with device pointers, fluid-rest frame.  res: velocities of every
 * particle -> forces, M f = (I + M L) v by GMRES (solve_res_lub_3f_0,
 * ewald-3f.c:373-406, and its FT / FTS forms); mix: known [(F, T, E)_m,
 * (U, O, E)_f] -> unknowns [(U, O, S)_m, (F, T, S)_f], fixed particles
 * sys->nm..np-1 (solve_mix_lub_3f, ewald-3f.c:646-705; solve_mix_lub_3ft,
 * ewald-3ft.c:747-805; solve_mix_lub_3fts_0, ewald-3fts.c:1192-1234; the
 * mobility problem when nm == np).  No counterpart in the reference (host
 * pointers only). */
int stokes_hip_solve_res_lub_dev (struct stokes *sys, const double *v_dev,
                                  double *f_dev, int *iters, double *relres);
int stokes_hip_solve_mix_lub_dev (struct stokes *sys, const double *b_dev,
                                  double *x_dev, int *iters, double *relres);

/* y = L x with device pointers (the operator of calc_lub_3f / 3ft / 3fts
 * for sys->version); 0 or a STOKES_HIP_E* code.  No counterpart in the
 * reference (its correction runs on the host). */
int stokes_hip_lub_dev (struct stokes *sys, const double *x_dev,
                        double *y_dev);

/* the coefficient polynomials of the eleven families for orders <= nmax:
 * tab[f][k (N + 2) + q] is the coefficient of lambda^q in f_k, N >= nmax
 * the order they were computed to (held until the process ends).  No
 * counterpart in the reference (it links them as generated code). */
int stokes_hip_twobody_coef (int nmax, const double **tab, int *N);

#ifdef __cplusplus
}
#endif

#endif
