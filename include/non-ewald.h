/* non-ewald.h -- open-boundary RPY mobility (drop-in for non-ewald.h). */
#ifndef LIBSTOKES_HIP_NON_EWALD_H
#define LIBSTOKES_HIP_NON_EWALD_H

#include "stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces non-ewald.c:49-90 */
void scalars_nonewald (int version, double r, double *scalar);
/* replaces non-ewald.c:330-493 */
void atimes_nonewald_3all (int n, const double *x, double *y, void *user_data);
/* replaces non-ewald.c:494- (make_matrix_mob_nonewald_3all) */
void make_matrix_mob_nonewald_3all (struct stokes *sys, double *mat);

#ifdef __cplusplus
}
#endif
#endif
