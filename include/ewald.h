/* ewald.h -- the mobility operator of the MI355X libstokes (drop-in).
 *
 * Same symbols and prototypes as the reference's ewald.h; every operator
 * runs on the GPU (hand-written HIP kernels for gfx950).  The void entry
 * points keep the reference's behaviour on failure (message on stderr and
 * exit(1), as memory-check.h:23-26); stokes_hip.h has int-returning
 * variants.
 */
#ifndef LIBSTOKES_HIP_EWALD_H
#define LIBSTOKES_HIP_EWALD_H

#include "stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces ewald.c:48-186: 11 real-space Ewald scalars of (xi, r) */
void scalars_ewald_real (int version, double xi, double r,
                         double *xa, double *ya, double *yb,
                         double *xc, double *yc, double *xg, double *yg,
                         double *yh, double *xm, double *ym, double *zm);
/* replaces ewald.c:203-359 (dimensional form) */
void scalars_ewald_real_poly (int version, double xi, double r,
                              double aa, double ab,
                              double *xa, double *ya, double *yb,
                              double *xc, double *yc, double *xg, double *yg,
                              double *yh, double *xm, double *ym, double *zm);
/* replaces ewald.c:382-420 */
void scalars_mob_poly_scale_SD_ (int version, double a1,
                                 double *xa, double *ya, double *yb,
                                 double *xc, double *yc, double *xg,
                                 double *yg, double *yh, double *xm,
                                 double *ym, double *zm);

/* replaces ewald.c:432-447: periodic -> Ewald, else open-BC RPY.
 * The atimes callback of libiter, chebyshev.h and the BD code. */
void atimes_3all (int n, const double *x, double *y, void *user_data);
/* replaces ewald.c:459-472 */
void make_matrix_mob_3all (struct stokes *sys, double *mat);
/* replaces ewald.c:484-505 */
void atimes_3all_matrix (int n, const double *x, double *y, void *user_data);
/* replaces ewald.c:519-1081 */
void atimes_ewald_3all (int n, const double *x, double *y, void *user_data);
/* replaces ewald.c:1092-1677 */
void make_matrix_mob_ewald_3all (struct stokes *sys, double *mat);

#ifdef __cplusplus
}
#endif
#endif
