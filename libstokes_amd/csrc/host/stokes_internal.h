/* stokes_internal.h -- host-side helpers shared by the C files of the
 * library (not installed). */
#ifndef LIBSTOKES_HIP_INTERNAL_H
#define LIBSTOKES_HIP_INTERNAL_H

#include <stdio.h>
#include "stokes.h"
#include "../shk.h"

/* iterative-solver parameters kept by stokes_set_iter (the fields libiter's
 * struct iter carries: solver name, maximum iterations, restart, tolerance,
 * debug flag and stream); the GMRES of iterative.c reads max, restart and
 * eps */
struct iter {
  char solver[32];
  int max, restart;
  double eps;
  int debug;
  FILE *out;
};

/* per-struct-stokes device context */
typedef struct stokes_ctx {
  struct stokes *sys;
  shk_dev *dev;
  int prec;                 /* SHK_FP64 / SHK_MIXED */
  int fix_recip_g;          /* reciprocal G term with the xi-independent sign */
  int timing;
  int valid;                /* parameters uploaded and current */
  unsigned long uploads;    /* count of full parameter uploads */
  shk_params prm_user;      /* parameters of sys itself at the last upload */
  double split_tr;          /* > 0: compute at the splitting of this tratio */
  double split_used_tr;     /* splitting of the uploaded tables (0: sys->xi) */
  shk_params prm;           /* last uploaded parameters */
  double shear_shift;       /* shift the k table and images were built for */
  int shear_mode;
  int version;
  const double *a_ptr;      /* radii pointer seen at upload */
  const double *pos_dev;    /* device positions given by the caller */
  /* host-side staging for the legacy (host pointer) API */
  double *x_dev, *y_dev;
  size_t xy_bytes;
  double *mat_dev;
  size_t mat_bytes;
  /* device-resident Chebyshev work space */
  double *cheb_dev;
  size_t cheb_bytes;
} stokes_ctx;

/* find or create the context of sys; NULL and a message on failure */
stokes_ctx *stokes_hip_ctx (struct stokes *sys);
/* the context of sys if one exists (none is created) */
stokes_ctx *stokes_hip_ctx_find (struct stokes *sys);
/* the device parameters p and host lattice tables t of sys computed at the
 * splitting of split_tr (0 or a non-periodic / sheared system: sys's own
 * xi), the reciprocal G term with the reference's sign unless fix_g;
 * release t with stokes_hip_tables_free */
int stokes_hip_tables (const struct stokes *sys, double split_tr, int fix_g,
                       shk_params *p, shk_tables *t);
void stokes_hip_tables_free (shk_tables *t);
/* upload the parameters / positions of sys for the next operator call */
int stokes_hip_prepare (stokes_ctx *c, int positions_on_device);
void stokes_hip_release_context (struct stokes *sys);
void stokes_hip_invalidate (struct stokes *sys);
/* context not tied to a system (free-standing dense products) */
shk_dev *stokes_hip_global_dev (void);
/* sh = s with the Ewald splitting of tratio (stokes_set_xi at
 * xi_by_tratio, same ewald_eps) and its own image table; release the table
 * with stokes_hip_shadow_free */
void stokes_hip_shadow_split (const struct stokes *s, double tratio,
                              struct stokes *sh);
void stokes_hip_shadow_free (struct stokes *sh);
/* restarted GMRES over a device operator op (ud, x, y): y = A x, with the
 * parameters of stokes_set_iter; returns an SHK_E* code */
int stokes_hip_gmres (stokes_ctx *c, size_t n,
                      int (*op) (void *, const double *, double *), void *ud,
                      const double *b_dev, double *x_dev, int *iters,
                      double *relres);
/* the lubrication parameters of sys (tables of the coefficients or the
 * equal-sphere f_k(1), 11 x 51 values into f1) */
int stokes_hip_lub_params (struct stokes *sys, shk_lub *lp, const double **tab,
                           double *f1);

/* record an error message for stokes_hip_last_error() */
void stokes_hip_set_error (const char *fmt, ...);
int stokes_dim (int version);
/* Clenshaw recursion of chebyshev.c:163-228 on the device operator of dev
 * (positions already uploaded): z = sum_j a_j T_j(M') y for nrhs columns of
 * length len / nrhs; work holds 4 len doubles.  Launches only (no
 * allocation, no wait), so it may be captured into a graph.  Returns an
 * SHK_E* code. */
int stokes_hip_cheb_launch (shk_dev *dev, int prec, int ncheb, const double *a,
                            int nrhs, size_t len, const double *y, double *z,
                            double eig0, double eig1, double *work);

/* restarted GMRES for the FTS problems on the device operator of c
 * (parameters and positions already uploaded): kind 0 resistance (b = (U,
 * O, E) -> x = (F, T, S)), kind 1 mixed with nm mobile particles (nm = np:
 * the mobility problem, b = (F, T, E) -> x = (U, O, S)); ws holds
 * stokes_hip_fts_ws (np, m) doubles of device memory.  Returns an SHK_E*
 * code. */
size_t stokes_hip_fts_ws (int np, int m);
/* the same for D values per particle, the first qs of a mobile particle
 * its known forces (F: 3, 3; FT: 6, 6; FTS: 11, 6) */
size_t stokes_hip_mix_ws (int np, int D, int m);
int stokes_hip_mix_solve (stokes_ctx *c, int D, int qs, int kind, int nm,
                          const double *b_dev, double *x_dev, int m, int maxit,
                          double eps, double *ws, int *iters, double *relres);
/* GMRES parameters of stokes_set_iter (restart, max, eps; defaults 30,
 * 2000, 1e-12) */
void stokes_hip_iter_params (const struct stokes *sys, int *m, int *maxit,
                             double *eps);
int stokes_hip_fts_solve (stokes_ctx *c, int kind, int nm, const double *b_dev,
                          double *x_dev, int m, int maxit, double eps,
                          double *ws, int *iters, double *relres);
/* Lanczos extreme eigenvalues of a symmetric operator of dimension n given
 * as op (ud, x, y) on host vectors (returns a STOKES_HIP_E* code) */
typedef int (*stokes_hip_host_op) (void *ud, const double *x, double *y);
int stokes_hip_lanczos_op (int n, int niter, unsigned long seed, double margin,
                           stokes_hip_host_op op, void *ud, double *lmin,
                           double *lmax);

/* MT19937 + polar method, the reference's Brownian random stream */
typedef struct stokes_hip_mt {
  unsigned long mt[624];
  int idx;
  int has_saved;
  double saved;
} stokes_hip_mt;
void stokes_hip_mt_seed (stokes_hip_mt *r, unsigned long seed);
double stokes_hip_mt_gaussian (stokes_hip_mt *r);

/* stokes_hip_solve_mix_lub_dev at the positions the device context holds
 * now (a Brownian step's stage sets them), without the set-up of the
 * public entry point (lub.c) */
int stokes_hip_solve_mix_lub_here (struct stokes *sys, const double *b_dev,
                                   double *x_dev, int *iters, double *relres);

#endif
