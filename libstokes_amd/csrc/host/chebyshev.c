/* chebyshev.c -- Chebyshev operator functions and spectral bounds (host C,
 * driving the GPU operator).
 *
 * chebyshev_coef / chebyshev_eval restate chebyshev.c:33-107 (scalar
 * set-up work).  chebyshev_eval_atimes keeps the reference's callback
 * contract (chebyshev.c:163-228); when the callback is this library's own
 * operator the whole Clenshaw recursion runs on the GPU with the vectors
 * resident in HBM (stokes_hip_cheb_eval_dev), otherwise the recursion calls
 * the user's atimes on host vectors as the reference does.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "chebyshev.h"
#include "ewald.h"
#include "non-ewald.h"
#include "stokes_hip.h"
#include "stokes_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

void
chebyshev_coef (int ncheb, double (*func) (double), double x0, double x1,
                double *a)
{
  /* Chebyshev-Gauss nodes on [x0, x1]; a_j = 2/n sum_k f(x_k) T_j(t_k) */
  double half = 0.5 * (x1 - x0), mid = 0.5 * (x1 + x0);
  for (int j = 0; j < ncheb; j++)
    {
      double acc = 0.0;
      for (int k = 1; k <= ncheb; k++)
        {
          double th = M_PI * ((double) k - 0.5) / (double) ncheb;
          acc += func (cos (th) * half + mid) * cos ((double) j * th);
        }
      a[j] = 2.0 / (double) ncheb * acc;
    }
  a[0] *= 0.5;
}

double
chebyshev_eval (int ncheb, const double *a, double x0, double x1, double x)
{
  double t = (x - 0.5 * (x1 + x0)) / (0.5 * (x1 - x0));
  double b1 = 0.0, b2 = 0.0;
  for (int j = ncheb - 1; j >= 1; j--)
    {
      double b = 2.0 * t * b1 - b2 + a[j];
      b2 = b1;
      b1 = b;
    }
  return t * b1 - b2 + a[0];
}

static int
is_own_operator (void (*atimes) (int, const double *, double *, void *))
{
  return atimes == atimes_3all || atimes == atimes_ewald_3all
    || atimes == atimes_nonewald_3all;
}

/* ------------------------------------------------------------------ */
/* device-resident Clenshaw recursion                                   */
/* ------------------------------------------------------------------ */

int
stokes_hip_cheb_eval_dev (struct stokes *sys, int ncheb, const double *a,
                          int nrhs, const double *y, double *z, double eig0,
                          double eig1)
{
  if (ncheb < 1 || nrhs < 1 || !(eig1 > eig0))
    {
      stokes_hip_set_error ("chebyshev: invalid ncheb/nrhs/bounds");
      return STOKES_HIP_EARG;
    }
  stokes_ctx *c = stokes_hip_ctx (sys);
  if (!c) return STOKES_HIP_ENODEV;
  size_t n = (size_t) sys->np * stokes_dim (sys->version);
  size_t len = n * (size_t) nrhs;
  size_t need = sizeof (double) * 4 * len;
  if (c->cheb_bytes < need)
    {
      if (c->cheb_dev) shk_dev_free (c->dev, c->cheb_dev);
      c->cheb_dev = NULL;
      c->cheb_bytes = 0;
      if (shk_dev_malloc (c->dev, (void **) &c->cheb_dev, need))
        {
          stokes_hip_set_error ("chebyshev: %s", shk_last_error ());
          return STOKES_HIP_ENOMEM;
        }
      c->cheb_bytes = need;
    }
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  rc = stokes_hip_cheb_launch (c->dev, c->prec, ncheb, a, nrhs, len, y, z,
                               eig0, eig1, c->cheb_dev);
  if (rc)
    {
      stokes_hip_set_error ("chebyshev: %s", shk_last_error ());
      return STOKES_HIP_EHIP;
    }
  return 0;
}

int
stokes_hip_cheb_launch (shk_dev *dev, int prec, int ncheb, const double *a,
                        int nrhs, size_t len, const double *y, double *z,
                        double eig0, double eig1, double *work)
{
  double *x = work, *b0 = x + len, *b1 = b0 + len, *b2 = b1 + len;
  double da = 2.0 / (eig1 - eig0);
  double db = -(eig1 + eig0) / (eig1 - eig0);
  int rc;
  /* b_{n} = b_{n+1} = 0; the first operator product M.0 = 0 is skipped,
   * so b_{n-1} = a_{n-1} y */
  if (ncheb == 1)
    return shk_dev_scal (dev, z, y, a[0], len);
  rc = shk_dev_scal (dev, b0, NULL, 0.0, len);
  if (!rc) rc = shk_dev_scal (dev, b1, y, a[ncheb - 1], len);
  if (rc) return rc;
  for (int j = ncheb - 2; j >= 1; j--)
    {
      rc = shk_dev_atimes (dev, b1, x, nrhs, prec);
      if (rc) return rc;
      rc = shk_dev_axpby4 (dev, b2, x, b1, b0, y, 2.0 * da, 2.0 * db, a[j], len);
      if (rc) return rc;
      double *t = b0; b0 = b1; b1 = b2; b2 = t;
    }
  rc = shk_dev_atimes (dev, b1, x, nrhs, prec);
  if (!rc) rc = shk_dev_axpby4 (dev, z, x, b1, b0, y, da, db, a[0], len);
  return rc;
}

void
chebyshev_eval_atimes (int ncheb, const double *a, int n, const double *y,
                       double *z, double eign0, double eign1,
                       void (*atimes) (int, const double *, double *, void *),
                       void *user_data)
{
  if (is_own_operator (atimes))
    {
      struct stokes *sys = (struct stokes *) user_data;
      stokes_ctx *c = stokes_hip_ctx (sys);
      int rc = c ? 0 : STOKES_HIP_ENODEV;
      double *yd = NULL, *zd = NULL;
      size_t bytes = sizeof (double) * (size_t) n;
      if (!rc && n != sys->np * stokes_dim (sys->version)) rc = STOKES_HIP_EARG;
      if (!rc) rc = stokes_hip_malloc ((void **) &yd, bytes);
      if (!rc) rc = stokes_hip_malloc ((void **) &zd, bytes);
      if (!rc) rc = stokes_hip_memcpy_h2d (yd, y, bytes);
      if (!rc) rc = stokes_hip_cheb_eval_dev (sys, ncheb, a, 1, yd, zd, eign0,
                                              eign1);
      if (!rc) rc = stokes_hip_sync (sys);
      if (!rc) rc = stokes_hip_memcpy_d2h (z, zd, bytes);
      if (yd) stokes_hip_free (yd);
      if (zd) stokes_hip_free (zd);
      if (rc)
        {
          fprintf (stderr, "libstokes-hip: chebyshev_eval_atimes failed (%d):"
                   " %s\n", rc, stokes_hip_last_error ());
          exit (1);
        }
      return;
    }
  /* generic operator: the reference's recursion on host vectors */
  double da = 2.0 / (eign1 - eign0);
  double db = -(eign1 + eign0) / (eign1 - eign0);
  double *w = calloc ((size_t) 4 * n, sizeof (double));
  if (w == NULL)
    {
      fprintf (stderr, "allocation failed in chebyshev_eval_atimes\n");
      exit (1);
    }
  double *x = w, *b0 = w + n, *b1 = w + 2 * n, *b2 = w + 3 * n;
  for (int j = ncheb - 1; j >= 1; j--)
    {
      atimes (n, b1, x, user_data);
      for (int k = 0; k < n; k++)
        b2[k] = 2.0 * da * x[k] + 2.0 * db * b1[k] - b0[k] + a[j] * y[k];
      double *t = b0; b0 = b1; b1 = b2; b2 = t;
    }
  atimes (n, b1, x, user_data);
  for (int k = 0; k < n; k++)
    z[k] = da * x[k] + db * b1[k] - b0[k] + a[0] * y[k];
  free (w);
}

double
chebyshev_error_minvsqrt (int n, const double *y, const double *z,
                          void (*atimes) (int, const double *, double *, void *),
                          void *user_data)
{
  double *x = malloc (sizeof (double) * (size_t) n);
  if (x == NULL) { fprintf (stderr, "allocation failed\n"); exit (1); }
  atimes (n, z, x, user_data);
  double yy = 0.0, zz = 0.0;
  for (int k = 0; k < n; k++) { yy += y[k] * y[k]; zz += z[k] * x[k]; }
  free (x);
  if (zz <= 0.0)
    {
      fprintf (stderr, "negative norm! %e\n", zz);
      exit (1);
    }
  return fabs (zz - yy) / zz;
}

double
chebyshev_error_Rsqrt (int n, const double *y, const double *z,
                       void (*atimes) (int, const double *, double *, void *),
                       void *user_data)
{
  double *x = malloc (sizeof (double) * (size_t) n);
  if (x == NULL) { fprintf (stderr, "allocation failed\n"); exit (1); }
  atimes (n, y, x, user_data);
  double yy = 0.0, zz = 0.0;
  for (int k = 0; k < n; k++) { yy += y[k] * x[k]; zz += z[k] * z[k]; }
  free (x);
  if (yy <= 0.0)
    {
      fprintf (stderr, "negative norm! %e\n", yy);
      exit (1);
    }
  return fabs (yy - zz) / yy;
}

/* ------------------------------------------------------------------ */
/* Lanczos bounds                                                       */
/* ------------------------------------------------------------------ */

/* number of eigenvalues of the symmetric tridiagonal (alpha, beta) below x
 * (Sturm sequence count) */
static int
sturm_count (int m, const double *al, const double *be, double x)
{
  int cnt = 0;
  double q = al[0] - x;
  if (q < 0.0) cnt++;
  for (int i = 1; i < m; i++)
    {
      double d = (q != 0.0) ? q : 1e-300;
      q = al[i] - x - be[i - 1] * be[i - 1] / d;
      if (q < 0.0) cnt++;
    }
  return cnt;
}

/* k-th smallest eigenvalue (0-based) by bisection in Gershgorin bounds */
static double
tridiag_eig (int m, const double *al, const double *be, int kth)
{
  double lo = al[0], hi = al[0];
  for (int i = 0; i < m; i++)
    {
      double r = (i > 0 ? fabs (be[i - 1]) : 0.0)
        + (i < m - 1 ? fabs (be[i]) : 0.0);
      lo = fmin (lo, al[i] - r);
      hi = fmax (hi, al[i] + r);
    }
  for (int it = 0; it < 200; it++)
    {
      double mid = 0.5 * (lo + hi);
      if (sturm_count (m, al, be, mid) > kth) hi = mid;
      else lo = mid;
      if (hi - lo <= 1e-15 * fmax (fabs (lo), fabs (hi))) break;
    }
  return 0.5 * (lo + hi);
}

static int
sys_op (void *ud, const double *x, double *y)
{
  struct stokes *sys = ud;
  return stokes_hip_atimes (sys, sys->np * stokes_dim (sys->version), x, y);
}

int
stokes_hip_lanczos_bounds (struct stokes *sys, int niter, unsigned long seed,
                           double margin, double *lmin, double *lmax)
{
  return stokes_hip_lanczos_op (sys->np * stokes_dim (sys->version), niter,
                                seed, margin, sys_op, sys, lmin, lmax);
}

int
stokes_hip_lanczos_op (int n, int niter, unsigned long seed, double margin,
                       stokes_hip_host_op op, void *ud, double *lmin,
                       double *lmax)
{
  if (niter < 2) niter = 2;
  if (niter > n) niter = n;
  double *V = malloc (sizeof (double) * (size_t) n * (size_t) (niter + 1));
  double *w = malloc (sizeof (double) * (size_t) n);
  double *al = malloc (sizeof (double) * (size_t) niter);
  double *be = malloc (sizeof (double) * (size_t) niter);
  if (!V || !w || !al || !be)
    {
      free (V); free (w); free (al); free (be);
      stokes_hip_set_error ("lanczos: host allocation failed");
      return STOKES_HIP_ENOMEM;
    }
  /* deterministic start vector (splitmix64 of the seed) */
  unsigned long long st = seed * 0x9E3779B97F4A7C15ULL + 1;
  double nrm = 0.0;
  for (int k = 0; k < n; k++)
    {
      st += 0x9E3779B97F4A7C15ULL;
      unsigned long long zz = st;
      zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ULL;
      zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBULL;
      zz ^= zz >> 31;
      V[k] = (double) (zz >> 11) * (1.0 / 9007199254740992.0) - 0.5;
      nrm += V[k] * V[k];
    }
  nrm = sqrt (nrm);
  for (int k = 0; k < n; k++) V[k] /= nrm;
  int m = 0, rc = 0;
  for (int j = 0; j < niter; j++)
    {
      double *vj = V + (size_t) j * n;
      rc = op (ud, vj, w);
      if (rc) break;
      double a = 0.0;
      for (int k = 0; k < n; k++) a += w[k] * vj[k];
      al[j] = a;
      m = j + 1;
      /* full reorthogonalisation (twice) against all previous vectors */
      for (int pass = 0; pass < 2; pass++)
        for (int i = 0; i <= j; i++)
          {
            double *vi = V + (size_t) i * n, d = 0.0;
            for (int k = 0; k < n; k++) d += w[k] * vi[k];
            for (int k = 0; k < n; k++) w[k] -= d * vi[k];
          }
      double b = 0.0;
      for (int k = 0; k < n; k++) b += w[k] * w[k];
      b = sqrt (b);
      if (j == niter - 1 || b < 1e-13 * fabs (a)) break;
      be[j] = b;
      double *vn = V + (size_t) (j + 1) * n;
      for (int k = 0; k < n; k++) vn[k] = w[k] / b;
    }
  if (!rc)
    {
      double e0 = tridiag_eig (m, al, be, 0);
      double e1 = tridiag_eig (m, al, be, m - 1);
      double span = e1 - e0;
      *lmin = e0 - margin * span;
      if (*lmin <= 0.0) *lmin = e0 * (1.0 - margin);
      *lmax = e1 + margin * span;
    }
  free (V); free (w); free (al); free (be);
  return rc;
}
