/* iterative.c -- restarted GMRES on device vectors and the FTS natural
 * mobility problem solved with it (the libiter route of
 * solve_mob_3fts_0, ewald-3fts.c:263-305, with the operator of
 * atimes_mob_3fts, 200-253, and the right-hand side of calc_b_mob_3fts,
 * 145-188).
 *
 * Unknowns x = (U, O, S) per particle; A x = (U, O, 0) - M (0, 0, S);
 * b = M (F, T, 0) - (0, 0, E).  GMRES(m): Arnoldi with classical
 * Gram-Schmidt applied twice (basis products on the GPU, the small
 * Hessenberg problem and Givens rotations on the host), restarts until the
 * relative residual |b - A x| / |b| falls below eps or max iterations.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ewald-3fts.h"
#include "stokes_hip.h"
#include "stokes_internal.h"

#define GMRES_MAX_RESTART 64

typedef int (*dev_op) (void *ud, const double *x, double *y);

/* |v|^2 of a device vector (one value back to the host) */
static int
dev_dot (shk_dev *d, const double *u, const double *v, size_t n, double *tmp,
         double *out)
{
  int rc = shk_dev_vtw (d, u, 1, n, v, tmp);
  if (!rc) rc = shk_dev_d2h (d, out, tmp, sizeof (double));
  return rc;
}

static int
gmres_dev (shk_dev *d, size_t n, dev_op op, void *ud, const double *b,
           double *x, int m, int maxit, double tol, double *V, double *w,
           double *hd, int *iters, double *relres)
{
  double H[GMRES_MAX_RESTART + 1][GMRES_MAX_RESTART];
  double cs[GMRES_MAX_RESTART], sn[GMRES_MAX_RESTART];
  double g[GMRES_MAX_RESTART + 1], y[GMRES_MAX_RESTART], h[GMRES_MAX_RESTART + 1];
  double bn2 = 0.0, res = 1.0;
  int total = 0, rc;
  if ((rc = dev_dot (d, b, b, n, hd, &bn2))) return rc;
  double bnorm = sqrt (bn2);
  if ((rc = shk_dev_scal (d, x, NULL, 0.0, n))) return rc;
  if (bnorm == 0.0)
    {
      *iters = 0;
      *relres = 0.0;
      return 0;
    }
  while (total < maxit)
    {
      /* r = b - A x into w (x = 0 on the first pass: r = b) */
      if (total == 0)
        rc = shk_dev_scal (d, w, b, 1.0, n);
      else
        {
          rc = op (ud, x, w);
          if (!rc) rc = shk_dev_axpby (d, w, b, 1.0, -1.0, n);
        }
      if (rc) return rc;
      double beta2;
      if ((rc = dev_dot (d, w, w, n, hd, &beta2))) return rc;
      double beta = sqrt (beta2);
      res = beta / bnorm;
      if (res < tol) break;
      if ((rc = shk_dev_scal (d, V, w, 1.0 / beta, n))) return rc;
      for (int k = 0; k <= m; k++) g[k] = 0.0;
      g[0] = beta;
      int jend = 0;
      for (int j = 0; j < m && total < maxit; j++)
        {
          double *vj = V + (size_t) j * n;
          if ((rc = op (ud, vj, w))) return rc;
          for (int k = 0; k <= j + 1; k++) H[k][j] = 0.0;
          for (int pass = 0; pass < 2; pass++)
            {
              rc = shk_dev_vtw (d, V, j + 1, n, w, hd);
              if (!rc) rc = shk_dev_d2h (d, h, hd, sizeof (double) * (j + 1));
              if (!rc) rc = shk_dev_vh (d, V, hd, j + 1, n, -1.0, w);
              if (rc) return rc;
              for (int k = 0; k <= j; k++) H[k][j] += h[k];
            }
          double hn2;
          if ((rc = dev_dot (d, w, w, n, hd, &hn2))) return rc;
          double hn = sqrt (hn2);
          H[j + 1][j] = hn;
          /* previous rotations, then a new one zeroing H[j+1][j] */
          for (int k = 0; k < j; k++)
            {
              double t = cs[k] * H[k][j] + sn[k] * H[k + 1][j];
              H[k + 1][j] = -sn[k] * H[k][j] + cs[k] * H[k + 1][j];
              H[k][j] = t;
            }
          double den = hypot (H[j][j], H[j + 1][j]);
          cs[j] = den > 0.0 ? H[j][j] / den : 1.0;
          sn[j] = den > 0.0 ? H[j + 1][j] / den : 0.0;
          H[j][j] = den;
          H[j + 1][j] = 0.0;
          g[j + 1] = -sn[j] * g[j];
          g[j] = cs[j] * g[j];
          total++;
          jend = j + 1;
          res = fabs (g[j + 1]) / bnorm;
          if (res < tol || hn == 0.0) break;
          if ((rc = shk_dev_scal (d, V + (size_t) (j + 1) * n, w, 1.0 / hn, n)))
            return rc;
        }
      /* y = H^-1 g (upper triangular), x += V y */
      for (int k = jend - 1; k >= 0; k--)
        {
          double t = g[k];
          for (int l = k + 1; l < jend; l++) t -= H[k][l] * y[l];
          y[k] = H[k][k] != 0.0 ? t / H[k][k] : 0.0;
        }
      rc = shk_dev_h2d (d, hd, y, sizeof (double) * jend);
      if (!rc) rc = shk_dev_vh (d, V, hd, jend, n, 1.0, x);
      if (rc) return rc;
      if (res < tol) break;
    }
  /* true residual of the result */
  if ((rc = op (ud, x, w))) return rc;
  if ((rc = shk_dev_axpby (d, w, b, 1.0, -1.0, n))) return rc;
  double r2;
  if ((rc = dev_dot (d, w, w, n, hd, &r2))) return rc;
  *iters = total;
  *relres = sqrt (r2) / bnorm;
  return 0;
}

typedef struct {
  shk_dev *d;
  int np, prec;
  double *t1, *t2;
} mob_op;

/* A x = (U, O, 0) - M (0, 0, S)  (atimes_mob_3fts, ewald-3fts.c:200-253) */
static int
mob_apply (void *ud, const double *x, double *y)
{
  mob_op *m = ud;
  int rc = shk_dev_fts_split (m->d, x, m->np, NULL, m->t1);
  if (!rc) rc = shk_dev_atimes (m->d, m->t1, m->t2, 1, m->prec);
  if (!rc) rc = shk_dev_fts_combine (m->d, x, m->t2, m->np, 0, y);
  return rc;
}

static void
iter_params (const struct stokes *sys, const struct iter *it, int *m,
             int *maxit, double *eps)
{
  *m = 30;
  *maxit = 2000;
  *eps = 1e-12;
  if (it == NULL) it = sys->it;
  if (it != NULL)
    {
      if (it->restart > 0) *m = it->restart;
      if (it->max > 0) *maxit = it->max;
      if (it->eps > 0.0) *eps = it->eps;
    }
  if (*m > GMRES_MAX_RESTART) *m = GMRES_MAX_RESTART;
}

int
stokes_hip_solve_mob_3fts_dev (struct stokes *sys, const double *fte_dev,
                               double *uos_dev, int *iters, double *relres)
{
  if (sys == NULL || fte_dev == NULL || uos_dev == NULL)
    {
      stokes_hip_set_error ("solve_mob_3fts: NULL argument");
      return STOKES_HIP_EARG;
    }
  if (sys->version != 2)
    {
      stokes_hip_set_error ("solve_mob_3fts: FTS version only (version %d)",
                            sys->version);
      return STOKES_HIP_EARG;
    }
  stokes_ctx *c = stokes_hip_ctx (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  int m, maxit;
  double eps;
  iter_params (sys, NULL, &m, &maxit, &eps);
  size_t n = (size_t) sys->np * 11;
  double *buf = NULL;
  size_t nd = n * (size_t) (m + 1) + 4 * n + (size_t) (m + 2);
  rc = shk_dev_malloc (c->dev, (void **) &buf, sizeof (double) * nd);
  if (rc)
    {
      stokes_hip_set_error ("solve_mob_3fts: %s", shk_last_error ());
      return STOKES_HIP_ENOMEM;
    }
  double *V = buf, *w = V + n * (size_t) (m + 1), *b = w + n, *t1 = b + n,
    *t2 = t1 + n, *hd = t2 + n;
  mob_op op = {c->dev, sys->np, c->prec, t1, t2};
  /* b = M (F, T, 0) - (0, 0, E)  (calc_b_mob_3fts, ewald-3fts.c:145-188) */
  rc = shk_dev_fts_split (c->dev, fte_dev, sys->np, t1, NULL);
  if (!rc) rc = shk_dev_atimes (c->dev, t1, t2, 1, c->prec);
  if (!rc) rc = shk_dev_fts_combine (c->dev, fte_dev, t2, sys->np, 1, b);
  int it = 0;
  double rr = 0.0;
  if (!rc)
    rc = gmres_dev (c->dev, n, mob_apply, &op, b, uos_dev, m, maxit, eps, V, w,
                    hd, &it, &rr);
  shk_dev_free (c->dev, buf);
  if (rc)
    {
      stokes_hip_set_error ("solve_mob_3fts: %s", shk_last_error ());
      return rc;
    }
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return 0;
}

void
solve_mob_3fts_0 (struct stokes *sys, struct iter *iter, const double *f,
                  const double *t, const double *e, double *u, double *o,
                  double *s)
{
  if (sys->version != 2)
    {
      fprintf (stderr, "libstokes solve_mob_3fts_0 :"
               " the version is wrong. reset to FTS\n");
      sys->version = 2;
    }
  struct iter *keep = sys->it;
  if (iter != NULL) sys->it = iter;
  int np = sys->np;
  size_t n = (size_t) np * 11;
  double *h = malloc (sizeof (double) * n);
  double *bd = NULL, *xd = NULL;
  if (h == NULL)
    {
      fprintf (stderr, "allocation failed in solve_mob_3fts_0\n");
      exit (1);
    }
  for (int i = 0; i < np; i++)
    {
      memcpy (h + 11 * i, f + 3 * i, sizeof (double) * 3);
      memcpy (h + 11 * i + 3, t + 3 * i, sizeof (double) * 3);
      memcpy (h + 11 * i + 6, e + 5 * i, sizeof (double) * 5);
    }
  int rc = stokes_hip_malloc ((void **) &bd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_malloc ((void **) &xd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_memcpy_h2d (bd, h, sizeof (double) * n);
  if (!rc) rc = stokes_hip_solve_mob_3fts_dev (sys, bd, xd, NULL, NULL);
  if (!rc) rc = stokes_hip_memcpy_d2h (h, xd, sizeof (double) * n);
  if (bd) stokes_hip_free (bd);
  if (xd) stokes_hip_free (xd);
  sys->it = keep;
  if (rc)
    {
      fprintf (stderr, "libstokes-hip: solve_mob_3fts_0 failed (%d): %s\n",
               rc, stokes_hip_last_error ());
      exit (1);
    }
  for (int i = 0; i < np; i++)
    {
      memcpy (u + 3 * i, h + 11 * i, sizeof (double) * 3);
      memcpy (o + 3 * i, h + 11 * i + 3, sizeof (double) * 3);
      memcpy (s + 5 * i, h + 11 * i + 6, sizeof (double) * 5);
    }
  free (h);
}
