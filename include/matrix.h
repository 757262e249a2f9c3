/* matrix.h -- extracted-form dense products (drop-in for matrix.h). */
#ifndef LIBSTOKES_HIP_MATRIX_H
#define LIBSTOKES_HIP_MATRIX_H

#ifdef __cplusplus
extern "C" {
#endif

/* replaces matrix.c:365-397: y[n1] = mat[n1,n2] . x[n2] (on the GPU) */
void dot_prod_matrix (const double *mat, int n1, int n2, const double *x,
                      double *y);
/* replaces matrix.c:447-478: y = m . T(x), T the S transform of the
 * extracted form (on the GPU) */
void multiply_extmat_with_extvec_3fts (int np, const double *m,
                                       const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
