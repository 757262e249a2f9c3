/* iterative.c -- restarted GMRES on device vectors and the FTS problems
 * solved with it, the libiter route of ewald-3fts.c:
 *   resistance  M (F, T, S) = (U, O, E)                     (:50-143)
 *   mobility    unknowns (U, O, S) from (F, T, E)           (:146-360)
 *   mixed       mobile particles as the mobility problem, fixed ones
 *               with unknowns (F, T, S) from (U, O, E)      (:363-670)
 * For the mixed problem (the mobility problem is its nm = np case) the
 * unknowns are x = [(U, O, S)_mobile, (F, T, S)_fixed],
 *   A x = [(U, O, 0)_m, 0_f] - M [(0, 0, S)_m, (F, T, S)_f]   (:440-525)
 *   b   = M [(F, T, 0)_m, 0_f] - [(0, 0, E)_m, (U, O, E)_f]   (:363-438)
 * GMRES(m): Arnoldi with classical Gram-Schmidt applied twice (basis
 * products on the GPU, the small Hessenberg problem and Givens rotations on
 * the host), restarts until the relative residual |b - A x| / |b| falls
 * below eps or max iterations.  The labo-frame wrappers shift by the
 * imposed flow (Ui, Oi, Ei) as f.c:569-660, ft.c:901-960, fts.c:2303-2330.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ewald-3f.h"
#include "ewald-3ft.h"
#include "ewald-3fts.h"
#include "ewald.h"
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
  int np, nm, prec;
  int D, qs;                    /* values per particle, known forces of a
                                   mobile particle (11, 6 for FTS) */
  double *t1, *t2;
} mix_op;

/* A x = [(U, O, 0)_m, 0_f] - M [(0, 0, S)_m, (F, T, S)_f]
 * (atimes_mix_3fts, ewald-3fts.c:440-525; atimes_mob_3fts for nm = np);
 * F / FT: A x = [U_m, 0_f] - M [0_m, F_f] (atimes_mix_3f, ewald-3f.c) */
static int
mix_apply (void *ud, const double *x, double *y)
{
  mix_op *m = ud;
  int rc = shk_dev_mix_split (m->d, x, m->np, m->nm, m->D, m->qs, NULL, m->t1);
  if (!rc) rc = shk_dev_atimes (m->d, m->t1, m->t2, 1, m->prec);
  if (!rc) rc = shk_dev_mix_combine (m->d, x, m->t2, m->np, m->nm, m->D, m->qs, 0, y);
  return rc;
}

/* A x = M x (the resistance problem, ewald-3fts.c:50-80) */
static int
res_apply (void *ud, const double *x, double *y)
{
  mix_op *m = ud;
  return shk_dev_atimes (m->d, x, y, 1, m->prec);
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

size_t
stokes_hip_fts_ws (int np, int m)
{
  const size_t n = (size_t) np * 11;
  if (m > GMRES_MAX_RESTART) m = GMRES_MAX_RESTART;
  return n * (size_t) (m + 1) + 4 * n + (size_t) (m + 2);
}

int
stokes_hip_fts_solve (stokes_ctx *c, int kind, int nm, const double *b_dev,
                      double *x_dev, int m, int maxit, double eps, double *ws,
                      int *iters, double *relres)
{
  return stokes_hip_mix_solve (c, 11, 6, kind, nm, b_dev, x_dev, m, maxit, eps,
                               ws, iters, relres);
}

size_t
stokes_hip_mix_ws (int np, int D, int m)
{
  const size_t n = (size_t) np * D;
  if (m > GMRES_MAX_RESTART) m = GMRES_MAX_RESTART;
  return n * (size_t) (m + 1) + 4 * n + (size_t) (m + 2);
}

int
stokes_hip_mix_solve (stokes_ctx *c, int D, int qs, int kind, int nm,
                      const double *b_dev, double *x_dev, int m, int maxit,
                      double eps, double *ws, int *iters, double *relres)
{
  const int np = c->sys->np;
  const size_t n = (size_t) np * D;
  if (m > GMRES_MAX_RESTART) m = GMRES_MAX_RESTART;
  double *V = ws, *w = V + n * (size_t) (m + 1), *b = w + n, *t1 = b + n,
    *t2 = t1 + n, *hd = t2 + n;
  mix_op op = {c->dev, np, nm, c->prec, D, qs, t1, t2};
  int rc;
  if (kind == 0)
    rc = shk_dev_scal (c->dev, b, b_dev, 1.0, n);
  else
    {
      /* b = M [(F, T, 0)_m, 0_f] - [(0, 0, E)_m, (U, O, E)_f]
       * (calc_b_mix_3fts, ewald-3fts.c:363-438; calc_b_mix_3f: b = M
       * [F_m, 0_f] - [0_m, U_f]) */
      rc = shk_dev_mix_split (c->dev, b_dev, np, nm, D, qs, t1, NULL);
      if (!rc) rc = shk_dev_atimes (c->dev, t1, t2, 1, c->prec);
      if (!rc) rc = shk_dev_mix_combine (c->dev, b_dev, t2, np, nm, D, qs, 1, b);
    }
  int it = 0;
  double rr = 0.0;
  if (!rc)
    rc = gmres_dev (c->dev, n, kind == 0 ? res_apply : mix_apply, &op, b,
                    x_dev, m, maxit, eps, V, w, hd, &it, &rr);
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return rc;
}

int
stokes_hip_gmres (stokes_ctx *c, size_t n, int (*op) (void *, const double *, double *),
                  void *ud, const double *b_dev, double *x_dev, int *iters,
                  double *relres)
{
  int m, maxit;
  double eps;
  iter_params (c->sys, NULL, &m, &maxit, &eps);
  double *ws = NULL;
  const size_t words = n * (size_t) (m + 1) + n + (size_t) (m + 2);
  int rc = shk_dev_malloc (c->dev, (void **) &ws, sizeof (double) * words);
  if (rc) return rc;
  int it = 0;
  double rr = 0.0;
  rc = gmres_dev (c->dev, n, op, ud, b_dev, x_dev, m, maxit, eps, ws,
                  ws + n * (size_t) (m + 1), ws + n * (size_t) (m + 2), &it, &rr);
  shk_dev_free (c->dev, ws);
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return rc;
}

void
stokes_hip_iter_params (const struct stokes *sys, int *m, int *maxit,
                        double *eps)
{
  iter_params (sys, NULL, m, maxit, eps);
}

/* values per particle of a version (F 3, FT 6, FTS 11) and the known
 * forces of a mobile particle in the mixed problem (3, 6, 6) */
static int
ver_dim (int version)
{
  return version == 0 ? 3 : (version == 1 ? 6 : 11);
}

static int
ver_qs (int version)
{
  return version == 0 ? 3 : 6;
}

/* kind 0: resistance, b_dev = (U, O, E) per particle -> x = (F, T, S);
 * kind 1: mixed with nm mobile particles, b_dev = [(F, T, E)_m,
 * (U, O, E)_f] -> x = [(U, O, S)_m, (F, T, S)_f]; the F and FT versions
 * the same with (U), (F) or (U, O), (F, T) per particle.  version: the
 * version the caller requires of sys */
static int
solve_dev (struct stokes *sys, int version, int kind, int nm,
           const double *b_dev, double *x_dev, int *iters, double *relres,
           const char *what)
{
  if (sys == NULL || b_dev == NULL || x_dev == NULL)
    {
      stokes_hip_set_error ("%s: NULL argument", what);
      return STOKES_HIP_EARG;
    }
  if (sys->version != version)
    {
      stokes_hip_set_error ("%s: version %d required (version %d)", what,
                            version, sys->version);
      return STOKES_HIP_EARG;
    }
  if (nm < 0 || nm > sys->np)
    {
      stokes_hip_set_error ("%s: nm = %d outside [0, %d]", what, nm, sys->np);
      return STOKES_HIP_EARG;
    }
  stokes_ctx *c = stokes_hip_ctx (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  int m, maxit;
  double eps;
  iter_params (sys, NULL, &m, &maxit, &eps);
  const int D = ver_dim (version);
  double *buf = NULL;
  rc = shk_dev_malloc (c->dev, (void **) &buf,
                       sizeof (double) * stokes_hip_mix_ws (sys->np, D, m));
  if (rc)
    {
      stokes_hip_set_error ("%s: %s", what, shk_last_error ());
      return STOKES_HIP_ENOMEM;
    }
  int it = 0;
  double rr = 0.0;
  rc = stokes_hip_mix_solve (c, D, ver_qs (version), kind, nm, b_dev, x_dev,
                             m, maxit, eps, buf, &it, &rr);
  shk_dev_free (c->dev, buf);
  if (rc)
    {
      stokes_hip_set_error ("%s: %s", what, shk_last_error ());
      return rc;
    }
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return 0;
}

int
stokes_hip_solve_mob_3fts_dev (struct stokes *sys, const double *fte_dev,
                               double *uos_dev, int *iters, double *relres)
{
  return solve_dev (sys, 2, 1, sys ? sys->np : 0, fte_dev, uos_dev, iters,
                        relres, "solve_mob_3fts");
}

int
stokes_hip_solve_mix_3fts_dev (struct stokes *sys, const double *b_dev,
                               double *x_dev, int *iters, double *relres)
{
  return solve_dev (sys, 2, 1, sys ? sys->nm : 0, b_dev, x_dev, iters,
                        relres, "solve_mix_3fts");
}

int
stokes_hip_solve_res_3fts_dev (struct stokes *sys, const double *uoe_dev,
                               double *fts_dev, int *iters, double *relres)
{
  return solve_dev (sys, 2, 0, sys ? sys->np : 0, uoe_dev, fts_dev, iters,
                        relres, "solve_res_3fts");
}

/* ---- reference entry points (host arrays; void, exit(1) on failure) ---- */

/* as the reference's solve_* do: a wrong version is reset with a message
 * (ewald-3f.c:50-55, ewald-3ft.c:53-58, ewald-3fts.c:54-59) */
static void
check_version (struct stokes *sys, int version, const char *name)
{
  static const char *vn[] = {"F", "FT", "FTS"};
  if (sys->version != version)
    {
      fprintf (stderr, "libstokes %s : the version is wrong. reset to %s\n",
               name, vn[version]);
      sys->version = version;
    }
}

static void
check_fts (struct stokes *sys, const char *name)
{
  check_version (sys, 2, name);
}

/* pack per-particle blocks of 3, 3, 5 from three arrays into h[11 n] */
static void
pack_335 (int n, double *h, const double *a, const double *b, const double *c)
{
  for (int i = 0; i < n; i++)
    {
      memcpy (h + 11 * i, a + 3 * i, sizeof (double) * 3);
      memcpy (h + 11 * i + 3, b + 3 * i, sizeof (double) * 3);
      memcpy (h + 11 * i + 6, c + 5 * i, sizeof (double) * 5);
    }
}

static void
unpack_335 (int n, const double *h, double *a, double *b, double *c)
{
  for (int i = 0; i < n; i++)
    {
      memcpy (a + 3 * i, h + 11 * i, sizeof (double) * 3);
      memcpy (b + 3 * i, h + 11 * i + 3, sizeof (double) * 3);
      memcpy (c + 5 * i, h + 11 * i + 6, sizeof (double) * 5);
    }
}

/* solve on host vectors h[D np] of sys->version (in: b, out: x) with
 * sys->it replaced by iter when not NULL */
static void
solve_host (struct stokes *sys, struct iter *iter, int kind, int nm,
            double *h, const char *name)
{
  struct iter *keep = sys->it;
  if (iter != NULL) sys->it = iter;
  size_t n = (size_t) sys->np * ver_dim (sys->version);
  double *bd = NULL, *xd = NULL;
  int rc = stokes_hip_malloc ((void **) &bd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_malloc ((void **) &xd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_memcpy_h2d (bd, h, sizeof (double) * n);
  if (!rc) rc = solve_dev (sys, sys->version, kind, nm, bd, xd, NULL, NULL,
                           name);
  if (!rc) rc = stokes_hip_memcpy_d2h (h, xd, sizeof (double) * n);
  if (bd) stokes_hip_free (bd);
  if (xd) stokes_hip_free (xd);
  sys->it = keep;
  if (rc)
    {
      fprintf (stderr, "libstokes-hip: %s failed (%d): %s\n", name, rc,
               stokes_hip_last_error ());
      exit (1);
    }
}

static double *
alloc_h (size_t n, const char *name)
{
  double *h = malloc (sizeof (double) * n);
  if (h == NULL)
    {
      fprintf (stderr, "allocation failed in %s\n", name);
      exit (1);
    }
  return h;
}

void
solve_res_3fts_0 (struct stokes *sys, const double *u, const double *o,
                  const double *e, double *f, double *t, double *s)
{
  check_fts (sys, "solve_res_3fts_0");
  double *h = alloc_h ((size_t) sys->np * 11, "solve_res_3fts_0");
  pack_335 (sys->np, h, u, o, e);
  solve_host (sys, NULL, 0, sys->np, h, "solve_res_3fts_0");
  unpack_335 (sys->np, h, f, t, s);
  free (h);
}

void
solve_mob_3fts_0 (struct stokes *sys, struct iter *iter, const double *f,
                  const double *t, const double *e, double *u, double *o,
                  double *s)
{
  check_fts (sys, "solve_mob_3fts_0");
  double *h = alloc_h ((size_t) sys->np * 11, "solve_mob_3fts_0");
  pack_335 (sys->np, h, f, t, e);
  solve_host (sys, iter, 1, sys->np, h, "solve_mob_3fts_0");
  unpack_335 (sys->np, h, u, o, s);
  free (h);
}

void
solve_mix_3fts_0 (struct stokes *sys, struct iter *iter, const double *f,
                  const double *t, const double *e, const double *uf,
                  const double *of, const double *ef, double *u, double *o,
                  double *s, double *ff, double *tf, double *sf)
{
  check_fts (sys, "solve_mix_3fts_0");
  int np = sys->np, nm = sys->nm, nf = np - nm;
  if (nm == np)
    {
      solve_mob_3fts_0 (sys, iter, f, t, e, u, o, s);
      return;
    }
  double *h = alloc_h ((size_t) np * 11, "solve_mix_3fts_0");
  pack_335 (nm, h, f, t, e);
  pack_335 (nf, h + 11 * (size_t) nm, uf, of, ef);
  solve_host (sys, iter, 1, nm, h, "solve_mix_3fts_0");
  unpack_335 (nm, h, u, o, s);
  unpack_335 (nf, h + 11 * (size_t) nm, ff, tf, sf);
  free (h);
}

/* ---- F and FT versions (ewald-3f.c, ewald-3ft.c) ---- */

/* h[6 n] particle-wise (a_i, b_i) from two arrays of 3 per particle and
 * back (set_ft_by_FT / set_FT_by_ft) */
static void
pack_33 (int n, double *h, const double *a, const double *b)
{
  for (int i = 0; i < n; i++)
    {
      memcpy (h + 6 * i, a + 3 * i, sizeof (double) * 3);
      memcpy (h + 6 * i + 3, b + 3 * i, sizeof (double) * 3);
    }
}

static void
unpack_33 (int n, const double *h, double *a, double *b)
{
  for (int i = 0; i < n; i++)
    {
      memcpy (a + 3 * i, h + 6 * i, sizeof (double) * 3);
      memcpy (b + 3 * i, h + 6 * i + 3, sizeof (double) * 3);
    }
}

void
solve_res_3f_0 (struct stokes *sys, const double *u, double *f)
{
  check_version (sys, 0, "solve_res_3f_0");
  const size_t n = (size_t) sys->np * 3;
  double *h = alloc_h (n, "solve_res_3f_0");
  memcpy (h, u, sizeof (double) * n);
  solve_host (sys, NULL, 0, sys->np, h, "solve_res_3f_0");
  memcpy (f, h, sizeof (double) * n);
  free (h);
}

void
solve_res_3f (struct stokes *sys, const double *u, double *f)
{
  check_version (sys, 0, "solve_res_3f");
  double *u0 = alloc_h ((size_t) sys->np * 3, "solve_res_3f");
  shift_labo_to_rest_U (sys, sys->np, u, u0);
  solve_res_3f_0 (sys, u0, f);
  free (u0);
}

void
solve_mob_3f (struct stokes *sys, const double *f, double *u)
{
  check_version (sys, 0, "solve_mob_3f");
  atimes_3all (sys->np * 3, f, u, sys);
  shift_rest_to_labo_U (sys, sys->np, u);
}

void
solve_mix_3f (struct stokes *sys, const double *f, const double *uf,
              double *u, double *ff)
{
  check_version (sys, 0, "solve_mix_3f");
  const int np = sys->np, nm = sys->nm, nf = np - nm;
  if (nm == np)
    {
      solve_mob_3f (sys, f, u);
      return;
    }
  double *h = alloc_h ((size_t) np * 3, "solve_mix_3f");
  memcpy (h, f, sizeof (double) * 3 * (size_t) nm);
  /* the shift takes the positions of particles 0..nf-1, as the reference
   * does (ewald-3f.c:301) */
  shift_labo_to_rest_U (sys, nf, uf, h + 3 * (size_t) nm);
  solve_host (sys, NULL, 1, nm, h, "solve_mix_3f");
  memcpy (u, h, sizeof (double) * 3 * (size_t) nm);
  memcpy (ff, h + 3 * (size_t) nm, sizeof (double) * 3 * (size_t) nf);
  free (h);
  shift_rest_to_labo_U (sys, nm, u);
}

void
solve_res_3ft_0 (struct stokes *sys, const double *u, const double *o,
                 double *f, double *t)
{
  check_version (sys, 1, "solve_res_3ft_0");
  double *h = alloc_h ((size_t) sys->np * 6, "solve_res_3ft_0");
  pack_33 (sys->np, h, u, o);
  solve_host (sys, NULL, 0, sys->np, h, "solve_res_3ft_0");
  unpack_33 (sys->np, h, f, t);
  free (h);
}

void
solve_res_3ft (struct stokes *sys, const double *u, const double *o,
               double *f, double *t)
{
  check_version (sys, 1, "solve_res_3ft");
  const size_t n3 = (size_t) sys->np * 3;
  double *u0 = alloc_h (2 * n3, "solve_res_3ft"), *o0 = u0 + n3;
  shift_labo_to_rest_U (sys, sys->np, u, u0);
  shift_labo_to_rest_O (sys, sys->np, o, o0);
  solve_res_3ft_0 (sys, u0, o0, f, t);
  free (u0);
}

void
solve_mob_3ft (struct stokes *sys, const double *f, const double *t,
               double *u, double *o)
{
  check_version (sys, 1, "solve_mob_3ft");
  const size_t n6 = (size_t) sys->np * 6;
  double *h = alloc_h (2 * n6, "solve_mob_3ft"), *y = h + n6;
  pack_33 (sys->np, h, f, t);
  atimes_3all (sys->np * 6, h, y, sys);
  unpack_33 (sys->np, y, u, o);
  free (h);
  shift_rest_to_labo_U (sys, sys->np, u);
  shift_rest_to_labo_O (sys, sys->np, o);
}

void
solve_mix_3ft (struct stokes *sys, const double *f, const double *t,
               const double *uf, const double *of, double *u, double *o,
               double *ff, double *tf)
{
  check_version (sys, 1, "solve_mix_3ft");
  const int np = sys->np, nm = sys->nm, nf = np - nm;
  if (nm == np)
    {
      solve_mob_3ft (sys, f, t, u, o);
      return;
    }
  double *h = alloc_h ((size_t) np * 6 + 6 * (size_t) nf, "solve_mix_3ft");
  double *uf0 = h + (size_t) np * 6, *of0 = uf0 + 3 * (size_t) nf;
  /* the shifts take the positions of particles 0..nf-1, as the reference
   * does (ewald-3ft.c:359-360) */
  shift_labo_to_rest_U (sys, nf, uf, uf0);
  shift_labo_to_rest_O (sys, nf, of, of0);
  pack_33 (nm, h, f, t);
  pack_33 (nf, h + 6 * (size_t) nm, uf0, of0);
  solve_host (sys, NULL, 1, nm, h, "solve_mix_3ft");
  unpack_33 (nm, h, u, o);
  unpack_33 (nf, h + 6 * (size_t) nm, ff, tf);
  free (h);
  shift_rest_to_labo_U (sys, nm, u);
  shift_rest_to_labo_O (sys, nm, o);
}

/* ---- imposed flow: labo frame <-> fluid-rest frame ---- */

/* imposed velocity Ui + Oi x pos_i + Ei . pos_i at particle i (Ei in the
 * extracted form xx xy xz yz yy, zz = -xx - yy) */
static void
imposed_u (const struct stokes *sys, int i, double *v)
{
  const double *E = sys->Ei, Ezz = -sys->Ei[0] - sys->Ei[4];
  const double *x = sys->pos + 3 * (size_t) i;
  v[0] = sys->Ui[0] + sys->Oi[1] * x[2] - sys->Oi[2] * x[1]
    + E[0] * x[0] + E[1] * x[1] + E[2] * x[2];
  v[1] = sys->Ui[1] + sys->Oi[2] * x[0] - sys->Oi[0] * x[2]
    + E[1] * x[0] + E[4] * x[1] + E[3] * x[2];
  v[2] = sys->Ui[2] + sys->Oi[0] * x[1] - sys->Oi[1] * x[0]
    + E[2] * x[0] + E[3] * x[1] + Ezz * x[2];
}

/* u0 = u - imposed velocity, particles 0..n-1 (f.c:569-610) */
void
shift_labo_to_rest_U (struct stokes *sys, int n, const double *u, double *u0)
{
  for (int i = 0; i < n; i++)
    {
      double v[3];
      imposed_u (sys, i, v);
      for (int c = 0; c < 3; c++) u0[3 * i + c] = u[3 * i + c] - v[c];
    }
}

/* u += imposed velocity, particles 0..n-1 (f.c:621-660) */
void
shift_rest_to_labo_U (struct stokes *sys, int n, double *u)
{
  for (int i = 0; i < n; i++)
    {
      double v[3];
      imposed_u (sys, i, v);
      for (int c = 0; c < 3; c++) u[3 * i + c] += v[c];
    }
}

/* o0 = o - Oi (ft.c:901-920) */
void
shift_labo_to_rest_O (struct stokes *sys, int n, const double *o, double *o0)
{
  for (int i = 0; i < 3 * n; i++) o0[i] = o[i] - sys->Oi[i % 3];
}

/* o += Oi (ft.c:935-955) */
void
shift_rest_to_labo_O (struct stokes *sys, int n, double *o)
{
  for (int i = 0; i < 3 * n; i++) o[i] += sys->Oi[i % 3];
}

/* e0 = e - Ei (fts.c:2303-2320) */
void
shift_labo_to_rest_E (struct stokes *sys, int n, const double *e, double *e0)
{
  for (int i = 0; i < 5 * n; i++) e0[i] = e[i] - sys->Ei[i % 5];
}

void
solve_res_3fts (struct stokes *sys, const double *u, const double *o,
                const double *e, double *f, double *t, double *s)
{
  check_fts (sys, "solve_res_3fts");
  int np = sys->np;
  double *u0 = alloc_h ((size_t) 11 * np, "solve_res_3fts");
  double *o0 = u0 + 3 * (size_t) np, *e0 = o0 + 3 * (size_t) np;
  shift_labo_to_rest_U (sys, np, u, u0);
  shift_labo_to_rest_O (sys, np, o, o0);
  shift_labo_to_rest_E (sys, np, e, e0);
  solve_res_3fts_0 (sys, u0, o0, e0, f, t, s);
  free (u0);
}

void
solve_mob_3fts (struct stokes *sys, const double *f, const double *t,
                const double *e, double *u, double *o, double *s)
{
  check_fts (sys, "solve_mob_3fts");
  int np = sys->np;
  double *e0 = alloc_h ((size_t) 5 * np, "solve_mob_3fts");
  shift_labo_to_rest_E (sys, np, e, e0);
  solve_mob_3fts_0 (sys, sys->it, f, t, e0, u, o, s);
  free (e0);
  shift_rest_to_labo_U (sys, np, u);
  shift_rest_to_labo_O (sys, np, o);
}

void
solve_mix_3fts (struct stokes *sys, const double *f, const double *t,
                const double *e, const double *uf, const double *of,
                const double *ef, double *u, double *o, double *s,
                double *ff, double *tf, double *sf)
{
  check_fts (sys, "solve_mix_3fts");
  int np = sys->np, nm = sys->nm, nf = np - nm;
  if (nm == np)
    {
      solve_mob_3fts (sys, f, t, e, u, o, s);
      return;
    }
  double *buf = alloc_h ((size_t) 11 * nf + 5 * (size_t) nm, "solve_mix_3fts");
  double *uf0 = buf, *of0 = uf0 + 3 * (size_t) nf, *ef0 = of0 + 3 * (size_t) nf;
  double *e0 = ef0 + 5 * (size_t) nf;
  /* as ewald-3fts.c:652: the fixed particles' velocities are shifted with
   * the positions of particles 0..nf-1, not nm..np-1 (kept for parity) */
  shift_labo_to_rest_U (sys, nf, uf, uf0);
  shift_labo_to_rest_O (sys, nf, of, of0);
  shift_labo_to_rest_E (sys, nf, ef, ef0);
  shift_labo_to_rest_E (sys, nm, e, e0);
  solve_mix_3fts_0 (sys, sys->it, f, t, e0, uf0, of0, ef0, u, o, s, ff, tf, sf);
  free (buf);
  shift_rest_to_labo_U (sys, nm, u);
  shift_rest_to_labo_O (sys, nm, o);
}
