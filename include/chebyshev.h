/* chebyshev.h -- Chebyshev operator functions (drop-in for chebyshev.h).
 * The recursion calls the given atimes; with atimes_3all it runs the GPU
 * operator once per order.  stokes_hip.h has the device-resident block
 * recursion. */
#ifndef LIBSTOKES_HIP_CHEBYSHEV_H
#define LIBSTOKES_HIP_CHEBYSHEV_H

#ifdef __cplusplus
extern "C" {
#endif

/* replaces chebyshev.c:33-61 */
void chebyshev_coef (int ncheb, double (*func) (double), double x0,
                     double x1, double *a);
/* replaces chebyshev.c:76-107 */
double chebyshev_eval (int ncheb, const double *a, double x0, double x1,
                       double x);
/* replaces chebyshev.c:163-228 */
void chebyshev_eval_atimes (int ncheb, const double *a, int n,
                            const double *y, double *z,
                            double eign0, double eign1,
                            void (*atimes) (int, const double *, double *,
                                            void *),
                            void *user_data);
/* replaces chebyshev.c:241-269 */
double chebyshev_error_minvsqrt (int n, const double *y, const double *z,
                                 void (*atimes) (int, const double *,
                                                 double *, void *),
                                 void *user_data);
/* replaces chebyshev.c:282-310 */
double chebyshev_error_Rsqrt (int n, const double *y, const double *z,
                              void (*atimes) (int, const double *, double *,
                                              void *),
                              void *user_data);

#ifdef __cplusplus
}
#endif
#endif
