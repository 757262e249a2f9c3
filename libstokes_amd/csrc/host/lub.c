/* lub.c -- the lubrication correction of the reference's lub.c (calc_lub_3f,
 * calc_lub_3ft, calc_lub_3fts) on the GPU.
 *
 * The pair scalars (two-body resistance minus the far-field inverse) are
 * evaluated per near pair in FP64 by the device kernels (lub_kernels.h);
 * the coefficient polynomials of unequal spheres come from twobody.c, the
 * equal-sphere series from f_k(1).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "stokes_hip.h"
#include "lub.h"
#include "ewald.h"
#include "ewald-3f.h"
#include "ewald-3ft.h"
#include "ewald-3fts.h"
#include "ewald-3fts-matrix.h"
#include "stokes_internal.h"

int
stokes_hip_lub_params (struct stokes *sys, shk_lub *lp, const double **tab, double *f1)
{
  memset (lp, 0, sizeof (*lp));
  lp->lubmin2 = sys->lubmin2;
  lp->lubmax = sys->lubmax;
  lp->nmax = sys->twobody_nmax;
  lp->flag_lub = sys->twobody_lub;
  lp->amax = 1.0;
  if (sys->a != NULL)
    for (int i = 0; i < sys->np; i++)
      if (i == 0 || sys->a[i] > lp->amax) lp->amax = sys->a[i];
  if (sys->a != NULL)
    {
      if (lp->nmax < 2)
        {
          stokes_hip_set_error ("lubrication: twobody_nmax %d < 2", lp->nmax);
          return STOKES_HIP_EARG;
        }
      int rc = stokes_hip_twobody_coef (lp->nmax, tab, &lp->N);
      if (rc) return rc;
      lp->coef = tab;
    }
  else
    {
      /* the equal-sphere series of two-body-res.c: f_k(1), k <= 50 */
      void (*fam[11]) (int, double, double *) = {
        twobody_XA, twobody_YA, twobody_YB, twobody_XC, twobody_YC, twobody_XG,
        twobody_YG, twobody_YH, twobody_XM, twobody_YM, twobody_ZM};
      for (int k = 0; k < 11; k++) fam[k] (50, 1.0, f1 + 51 * k);
      lp->f1 = f1;
    }
  return 0;
}

static int
lub_dev (struct stokes *sys, int version, const double *x_dev, double *y_dev)
{
  stokes_ctx *c = stokes_hip_ctx (sys);
  if (!c) return STOKES_HIP_ENODEV;
  if (sys->slip != NULL)
    {
      stokes_hip_set_error ("lubrication: slip particles are not supported");
      return STOKES_HIP_EUNSUP;
    }
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  shk_lub lp;
  const double *tab[11];
  double f1[11 * 51];
  rc = stokes_hip_lub_params (sys, &lp, tab, f1);
  if (rc) return rc;
  rc = shk_dev_lub (c->dev, version, &lp, x_dev, y_dev);
  if (rc)
    {
      stokes_hip_set_error ("lubrication: %s", shk_last_error ());
      return rc;
    }
  return 0;
}

int
stokes_hip_lub_dev (struct stokes *sys, const double *x_dev, double *y_dev)
{
  if (!sys || sys->version < 0 || sys->version > 2)
    {
      stokes_hip_set_error ("lub_dev: invalid system");
      return STOKES_HIP_EARG;
    }
  return lub_dev (sys, sys->version, x_dev, y_dev);
}

/* host vectors: staged through device buffers of this call */
static void
lub_host (struct stokes *sys, int version, const double *x, double *y, const char *name)
{
  const size_t n = (size_t) sys->np * (version == 0 ? 3 : (version == 1 ? 6 : 11));
  void *xd = NULL, *yd = NULL;
  int rc = stokes_hip_malloc (&xd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_malloc (&yd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_memcpy_h2d (xd, x, sizeof (double) * n);
  if (!rc) rc = lub_dev (sys, version, xd, yd);
  if (!rc) rc = stokes_hip_memcpy_d2h (y, yd, sizeof (double) * n);
  if (xd) stokes_hip_free (xd);
  if (yd) stokes_hip_free (yd);
  if (rc)
    {
      fprintf (stderr, "%s: %s\n", name, stokes_hip_last_error ());
      exit (1);
    }
}

void calc_lub_3f (struct stokes *sys, const double *u, double *f) { lub_host (sys, 0, u, f, "calc_lub_3f"); }
void calc_lub_3ft (struct stokes *sys, const double *uo, double *ft) { lub_host (sys, 1, uo, ft, "calc_lub_3ft"); }
void calc_lub_3fts (struct stokes *sys, const double *uoe, double *fts) { lub_host (sys, 2, uoe, fts, "calc_lub_3fts"); }

/* ------------------------------------------------------------------ */
/* dense lubrication matrix                                             */
/* ------------------------------------------------------------------ */

static int
lub_matrix_dev (struct stokes *sys, int version, double *mat_dev)
{
  stokes_ctx *c = stokes_hip_ctx (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  shk_lub lp;
  const double *tab[11];
  double f1[11 * 51];
  if ((rc = stokes_hip_lub_params (sys, &lp, tab, f1))) return rc;
  rc = shk_dev_lub_matrix (c->dev, version, &lp, mat_dev);
  if (rc) stokes_hip_set_error ("lubrication matrix: %s", shk_last_error ());
  return rc;
}

int
stokes_hip_make_matrix_lub_dev (struct stokes *sys, double *mat_dev)
{
  if (!sys || sys->version < 0 || sys->version > 2)
    {
      stokes_hip_set_error ("make_matrix_lub_dev: invalid system");
      return STOKES_HIP_EARG;
    }
  return lub_matrix_dev (sys, sys->version, mat_dev);
}

static void
lub_matrix_host (struct stokes *sys, int version, double *mat, const char *name)
{
  const size_t n = (size_t) sys->np * (version == 0 ? 3 : (version == 1 ? 6 : 11));
  void *md = NULL;
  int rc = stokes_hip_malloc (&md, sizeof (double) * n * n);
  if (!rc) rc = lub_matrix_dev (sys, version, md);
  if (!rc) rc = stokes_hip_memcpy_d2h (mat, md, sizeof (double) * n * n);
  if (md) stokes_hip_free (md);
  if (rc)
    {
      fprintf (stderr, "%s: %s\n", name, stokes_hip_last_error ());
      exit (1);
    }
}

void make_matrix_lub_3f (struct stokes *sys, double *mat) { lub_matrix_host (sys, 0, mat, "make_matrix_lub_3f"); }
void make_matrix_lub_3ft (struct stokes *sys, double *mat) { lub_matrix_host (sys, 1, mat, "make_matrix_lub_3ft"); }
void make_matrix_lub_3fts (struct stokes *sys, double *mat) { lub_matrix_host (sys, 2, mat, "make_matrix_lub_3fts"); }

/* ------------------------------------------------------------------ */
/* lubricated solves                                                    */
/* ------------------------------------------------------------------ */

typedef struct {
  stokes_ctx *c;
  shk_lub lp;
  int np, nm;                 /* particles nm..np-1 are fixed */
  int version, D, qs;         /* D values per particle, the first qs of a
                                 mobile particle's are its known forces */
  double *t1, *t2, *t3;
} lub_op;

/* A x = M x (the resistance problem's operator) */
static int
op_mob (void *ud, const double *x, double *y)
{
  lub_op *o = ud;
  return shk_dev_atimes (o->c->dev, x, y, 1, o->c->prec);
}

/* A x = (I + M L) lo(x) - M hi(x), with lo(x) = [(u, o, 0)_mobile,
 * 0_fixed] and hi(x) = [(0, 0, s)_mobile, (f, t, s)_fixed] for the
 * unknowns x = [(u, o, s)_mobile, (f, t, s)_fixed] (F: u / f, FT: u, o /
 * f, t; atimes_mob_lub_3f, ewald-3f.c:410-436; atimes_mix_lub_3f, 566-644;
 * atimes_mix_lub_3ft, ewald-3ft.c:657-746; atimes_mob_lub_3fts,
 * ewald-3fts.c:835-907; atimes_mix_lub_3fts, 1088-1190) */
static int
op_mix_lub (void *ud, const double *x, double *y)
{
  lub_op *o = ud;
  shk_dev *d = o->c->dev;
  const size_t n = (size_t) o->np * o->D;
  int rc = shk_dev_mix_split (d, x, o->np, o->nm, o->D, o->qs, o->t1, o->t2);
  if (!rc) rc = shk_dev_atimes (d, o->t2, y, 1, o->c->prec);
  if (!rc) rc = shk_dev_axpby (d, y, o->t1, 1.0, -1.0, n);
  if (!rc) rc = shk_dev_lub (d, o->version, &o->lp, o->t1, o->t3);
  if (!rc) rc = shk_dev_atimes (d, o->t3, o->t2, 1, o->c->prec);
  if (!rc) rc = shk_dev_axpby (d, y, o->t2, 1.0, 1.0, n);
  return rc;
}

/* sys->version, nm mobile particles.  kind 0: velocities (U, O, E) of
 * every particle -> (F, T, S) = M^-1 (I + M L) (U, O, E), GMRES on M with
 * b = (I + M L) (U, O, E) (solve_res_lub_3f_0, ewald-3f.c:373-406;
 * solve_res_lub_3ft_0, ewald-3ft.c:441-485; solve_res_lub_3fts_0,
 * ewald-3fts.c:676-719); kind 1: known [(F, T, E)_mobile, (U, O,
 * E)_fixed] -> unknowns [(U, O, S)_mobile, (F, T, S)_fixed], GMRES on
 * op_mix_lub with b = M lo(known) - (I + M L) hi(known) (calc_b_mix_lub_3f,
 * ewald-3f.c:495-564; calc_b_mix_lub_3ft, ewald-3ft.c:585-655;
 * calc_b_mob_lub_3fts, ewald-3fts.c:776-833; calc_b_mix_lub_3fts,
 * 1015-1076), nm = np being the mobility problem.  matrix (FTS, nm = np):
 * dense solve instead of GMRES.  All in the fluid-rest frame. */
static int
lub_solve_dev_pos (struct stokes *sys, int version, int kind, int nm, int matrix,
                   const double *b_dev, double *x_dev, int *iters, double *relres,
                   const char *what, int keep_pos);

static int
lub_solve_dev (struct stokes *sys, int version, int kind, int nm, int matrix,
               const double *b_dev, double *x_dev, int *iters, double *relres,
               const char *what)
{
  return lub_solve_dev_pos (sys, version, kind, nm, matrix, b_dev, x_dev, iters, relres,
                            what, 0);
}

/* keep_pos: the positions the device context holds now (set by the
 * caller, e.g. a Brownian step's stage), no set-up or upload */
static int
lub_solve_dev_pos (struct stokes *sys, int version, int kind, int nm, int matrix,
                   const double *b_dev, double *x_dev, int *iters, double *relres,
                   const char *what, int keep_pos)
{
  static const char *vn[3] = {"F", "FT", "FTS"};
  if (!sys || sys->version != version || !b_dev || !x_dev)
    {
      stokes_hip_set_error ("%s: an %s system and device vectors are required", what,
                            vn[version]);
      return STOKES_HIP_EARG;
    }
  if (nm < 0 || nm > sys->np || (matrix && (version != 2 || nm != sys->np)))
    {
      stokes_hip_set_error ("%s: %d mobile particles of %d", what, nm, sys->np);
      return STOKES_HIP_EARG;
    }
  stokes_ctx *c = stokes_hip_ctx (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = keep_pos ? 0 : stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  lub_op o;
  memset (&o, 0, sizeof (o));
  o.c = c;
  o.np = sys->np;
  o.nm = nm;
  o.version = version;
  o.D = version == 0 ? 3 : (version == 1 ? 6 : 11);
  o.qs = version == 2 ? 6 : o.D;
  const double *tab[11];
  double f1[11 * 51];
  if ((rc = stokes_hip_lub_params (sys, &o.lp, tab, f1))) return rc;
  shk_dev *d = c->dev;
  const size_t n = (size_t) sys->np * o.D;
  if (matrix)
    {
      double *M = NULL, *L = NULL;
      rc = shk_dev_malloc (d, (void **) &M, sizeof (double) * n * n);
      if (!rc) rc = shk_dev_malloc (d, (void **) &L, sizeof (double) * n * n);
      if (!rc) rc = shk_dev_matrix (d, M);
      if (!rc) rc = shk_dev_lub_matrix (d, 2, &o.lp, L);
      if (!rc) rc = shk_dev_solve_lub (d, M, L, sys->np, kind, b_dev, x_dev);
      if (M) shk_dev_free (d, M);
      if (L) shk_dev_free (d, L);
      if (rc) stokes_hip_set_error ("%s: %s", what, shk_last_error ());
      return rc;
    }
  double *ws = NULL;
  rc = shk_dev_malloc (d, (void **) &ws, sizeof (double) * 4 * n);
  if (rc)
    {
      stokes_hip_set_error ("%s: %s", what, shk_last_error ());
      return STOKES_HIP_ENOMEM;
    }
  double *b = ws;
  o.t1 = ws + n;
  o.t2 = ws + 2 * n;
  o.t3 = ws + 3 * n;
  if (kind == 0)
    {
      rc = shk_dev_lub (d, version, &o.lp, b_dev, o.t1);
      if (!rc) rc = shk_dev_atimes (d, o.t1, b, 1, c->prec);
      if (!rc) rc = shk_dev_axpby (d, b, b_dev, 1.0, 1.0, n);
      if (!rc) rc = stokes_hip_gmres (c, n, op_mob, &o, b, x_dev, iters, relres);
    }
  else
    {
      rc = shk_dev_mix_split (d, b_dev, sys->np, nm, o.D, o.qs, o.t1, o.t2);
      if (!rc) rc = shk_dev_atimes (d, o.t1, b, 1, c->prec);
      if (!rc) rc = shk_dev_axpby (d, b, o.t2, -1.0, 1.0, n);
      if (!rc) rc = shk_dev_lub (d, version, &o.lp, o.t2, o.t3);
      if (!rc) rc = shk_dev_atimes (d, o.t3, o.t1, 1, c->prec);
      if (!rc) rc = shk_dev_axpby (d, b, o.t1, -1.0, 1.0, n);
      if (!rc) rc = stokes_hip_gmres (c, n, op_mix_lub, &o, b, x_dev, iters, relres);
    }
  shk_dev_free (d, ws);
  if (rc) stokes_hip_set_error ("%s: %s", what, shk_last_error ());
  return rc;
}

int
stokes_hip_solve_res_lub_3fts_dev (struct stokes *sys, const double *uoe_dev,
                                   double *fts_dev, int *iters, double *relres)
{
  return lub_solve_dev (sys, 2, 0, sys ? sys->np : 0, 0, uoe_dev, fts_dev, iters, relres,
                        "solve_res_lub_3fts");
}

int
stokes_hip_solve_mob_lub_3fts_dev (struct stokes *sys, const double *fte_dev,
                                   double *uos_dev, int *iters, double *relres)
{
  return lub_solve_dev (sys, 2, 1, sys ? sys->np : 0, 0, fte_dev, uos_dev, iters, relres,
                        "solve_mob_lub_3fts");
}

int
stokes_hip_solve_lub_3fts_matrix_dev (struct stokes *sys, int kind,
                                      const double *b_dev, double *x_dev)
{
  if (kind != 0 && kind != 1)
    {
      stokes_hip_set_error ("solve_lub_3fts_matrix: kind %d", kind);
      return STOKES_HIP_EARG;
    }
  return lub_solve_dev (sys, 2, kind, sys ? sys->np : 0, 1, b_dev, x_dev, NULL, NULL,
                        "solve_lub_3fts_matrix");
}

int
stokes_hip_solve_res_lub_dev (struct stokes *sys, const double *v_dev, double *f_dev,
                              int *iters, double *relres)
{
  if (!sys || sys->version < 0 || sys->version > 2)
    {
      stokes_hip_set_error ("solve_res_lub: no system");
      return STOKES_HIP_EARG;
    }
  return lub_solve_dev (sys, sys->version, 0, sys->np, 0, v_dev, f_dev, iters, relres,
                        "solve_res_lub");
}

int
stokes_hip_solve_mix_lub_here (struct stokes *sys, const double *b_dev, double *x_dev,
                               int *iters, double *relres)
{
  return lub_solve_dev_pos (sys, sys->version, 1, sys->nm, 0, b_dev, x_dev, iters, relres,
                            "solve_mix_lub", 1);
}

int
stokes_hip_solve_mix_lub_dev (struct stokes *sys, const double *b_dev, double *x_dev,
                              int *iters, double *relres)
{
  if (!sys || sys->version < 0 || sys->version > 2)
    {
      stokes_hip_set_error ("solve_mix_lub: no system");
      return STOKES_HIP_EARG;
    }
  return lub_solve_dev (sys, sys->version, 1, sys->nm, 0, b_dev, x_dev, iters, relres,
                        "solve_mix_lub");
}

/* ---- the reference's host-array entry points ---- */

static void
pack11 (int np, double *h, const double *a, const double *b, const double *c)
{
  for (int i = 0; i < np; i++)
    {
      memcpy (h + 11 * (size_t) i, a + 3 * (size_t) i, sizeof (double) * 3);
      memcpy (h + 11 * (size_t) i + 3, b + 3 * (size_t) i, sizeof (double) * 3);
      memcpy (h + 11 * (size_t) i + 6, c + 5 * (size_t) i, sizeof (double) * 5);
    }
}

static void
unpack11 (int np, const double *h, double *a, double *b, double *c)
{
  for (int i = 0; i < np; i++)
    {
      memcpy (a + 3 * (size_t) i, h + 11 * (size_t) i, sizeof (double) * 3);
      memcpy (b + 3 * (size_t) i, h + 11 * (size_t) i + 3, sizeof (double) * 3);
      memcpy (c + 5 * (size_t) i, h + 11 * (size_t) i + 6, sizeof (double) * 5);
    }
}

/* the reference resets a wrong version with a message (ewald-3fts.c:680-685) */
static void
lub_host_solve (struct stokes *sys, int kind, int matrix, const double *a,
                const double *b, const double *c, double *x1, double *x2,
                double *x3, const char *name)
{
  if (sys->version != 2)
    {
      fprintf (stderr, "libstokes %s : the version is wrong. reset to FTS\n", name);
      sys->version = 2;
    }
  const size_t n = (size_t) sys->np * 11;
  double *h = malloc (sizeof (double) * n);
  void *bd = NULL, *xd = NULL;
  if (!h)
    {
      fprintf (stderr, "%s: allocation failed\n", name);
      exit (1);
    }
  pack11 (sys->np, h, a, b, c);
  int rc = stokes_hip_malloc (&bd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_malloc (&xd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_memcpy_h2d (bd, h, sizeof (double) * n);
  if (!rc) rc = lub_solve_dev (sys, 2, kind, sys->np, matrix, bd, xd, NULL, NULL, name);
  if (!rc) rc = stokes_hip_memcpy_d2h (h, xd, sizeof (double) * n);
  if (bd) stokes_hip_free (bd);
  if (xd) stokes_hip_free (xd);
  if (rc)
    {
      fprintf (stderr, "%s: %s\n", name, stokes_hip_last_error ());
      exit (1);
    }
  unpack11 (sys->np, h, x1, x2, x3);
  free (h);
}

void
solve_res_lub_3fts_0 (struct stokes *sys, const double *u, const double *o,
                      const double *e, double *f, double *t, double *s)
{
  lub_host_solve (sys, 0, 0, u, o, e, f, t, s, "solve_res_lub_3fts_0");
}

void
solve_mob_lub_3fts_0 (struct stokes *sys, const double *f, const double *t,
                      const double *e, double *u, double *o, double *s)
{
  lub_host_solve (sys, 1, 0, f, t, e, u, o, s, "solve_mob_lub_3fts_0");
}

void
solve_res_lub_3fts_matrix_0 (struct stokes *sys, const double *u,
                             const double *o, const double *e, double *f,
                             double *t, double *s)
{
  lub_host_solve (sys, 0, 1, u, o, e, f, t, s, "solve_res_lub_3fts_matrix_0");
}

void
solve_mob_lub_3fts_matrix_0 (struct stokes *sys, const double *f,
                             const double *t, const double *e, double *u,
                             double *o, double *s)
{
  lub_host_solve (sys, 1, 1, f, t, e, u, o, s, "solve_mob_lub_3fts_matrix_0");
}

static double *
lub_alloc (size_t n)
{
  double *p = malloc (sizeof (double) * (n ? n : 1));
  if (!p)
    {
      fprintf (stderr, "lubricated solve: allocation failed\n");
      exit (1);
    }
  return p;
}

/* labo frame: the known velocities / strains go to the fluid-rest frame
 * first, the velocities come back (ewald-3fts.c:733-770, 966-1000) */
static void
res_labo (struct stokes *sys, int matrix, const double *u, const double *o,
          const double *e, double *f, double *t, double *s)
{
  const int np = sys->np;
  double *u0 = lub_alloc (3 * (size_t) np), *o0 = lub_alloc (3 * (size_t) np);
  double *e0 = lub_alloc (5 * (size_t) np);
  shift_labo_to_rest_U (sys, np, u, u0);
  shift_labo_to_rest_O (sys, np, o, o0);
  shift_labo_to_rest_E (sys, np, e, e0);
  if (matrix) solve_res_lub_3fts_matrix_0 (sys, u0, o0, e0, f, t, s);
  else solve_res_lub_3fts_0 (sys, u0, o0, e0, f, t, s);
  free (u0); free (o0); free (e0);
}

static void
mob_labo (struct stokes *sys, int matrix, const double *f, const double *t,
          const double *e, double *u, double *o, double *s)
{
  const int np = sys->np;
  double *e0 = lub_alloc (5 * (size_t) np);
  shift_labo_to_rest_E (sys, np, e, e0);
  if (matrix) solve_mob_lub_3fts_matrix_0 (sys, f, t, e0, u, o, s);
  else solve_mob_lub_3fts_0 (sys, f, t, e0, u, o, s);
  free (e0);
  shift_rest_to_labo_U (sys, np, u);
  shift_rest_to_labo_O (sys, np, o);
}

void
solve_res_lub_3fts (struct stokes *sys, const double *u, const double *o,
                    const double *e, double *f, double *t, double *s)
{
  res_labo (sys, 0, u, o, e, f, t, s);
}

void
solve_mob_lub_3fts (struct stokes *sys, const double *f, const double *t,
                    const double *e, double *u, double *o, double *s)
{
  mob_labo (sys, 0, f, t, e, u, o, s);
}

void
solve_res_lub_3fts_matrix (struct stokes *sys, const double *u,
                           const double *o, const double *e, double *f,
                           double *t, double *s)
{
  res_labo (sys, 1, u, o, e, f, t, s);
}

void
solve_mob_lub_3fts_matrix (struct stokes *sys, const double *f,
                           const double *t, const double *e, double *u,
                           double *o, double *s)
{
  mob_labo (sys, 1, f, t, e, u, o, s);
}

/* ---- lubricated F and FT solves, the FTS mixed problem (host arrays) ---- */

/* h[D np] in (known values), out (solution); exits with a message on a
 * failure, as the reference's solvers have no error return */
static void
lub_host_v (struct stokes *sys, int version, int kind, int nm, double *h,
            const char *name)
{
  static const char *vn[3] = {"F", "FT", "FTS"};
  if (sys->version != version)
    {
      fprintf (stderr, "libstokes %s : the version is wrong. reset to %s\n", name,
               vn[version]);
      sys->version = version;
    }
  const size_t n = (size_t) sys->np * (version == 0 ? 3 : (version == 1 ? 6 : 11));
  void *bd = NULL, *xd = NULL;
  int rc = stokes_hip_malloc (&bd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_malloc (&xd, sizeof (double) * n);
  if (!rc) rc = stokes_hip_memcpy_h2d (bd, h, sizeof (double) * n);
  if (!rc) rc = lub_solve_dev (sys, version, kind, nm, 0, bd, xd, NULL, NULL, name);
  if (!rc) rc = stokes_hip_memcpy_d2h (h, xd, sizeof (double) * n);
  if (bd) stokes_hip_free (bd);
  if (xd) stokes_hip_free (xd);
  if (rc)
    {
      fprintf (stderr, "%s: %s\n", name, stokes_hip_last_error ());
      exit (1);
    }
}

/* (a_i, b_i) per particle of 6 from two arrays of 3 and back */
static void
pack_ft (int n, double *h, const double *a, const double *b)
{
  for (int i = 0; i < n; i++)
    {
      memcpy (h + 6 * (size_t) i, a + 3 * (size_t) i, sizeof (double) * 3);
      memcpy (h + 6 * (size_t) i + 3, b + 3 * (size_t) i, sizeof (double) * 3);
    }
}

static void
unpack_ft (int n, const double *h, double *a, double *b)
{
  for (int i = 0; i < n; i++)
    {
      memcpy (a + 3 * (size_t) i, h + 6 * (size_t) i, sizeof (double) * 3);
      memcpy (b + 3 * (size_t) i, h + 6 * (size_t) i + 3, sizeof (double) * 3);
    }
}

void
solve_res_lub_3f_0 (struct stokes *sys, const double *u, double *f)
{
  const size_t n = (size_t) sys->np * 3;
  double *h = lub_alloc (n);
  memcpy (h, u, sizeof (double) * n);
  lub_host_v (sys, 0, 0, sys->np, h, "solve_res_lub_3f_0");
  memcpy (f, h, sizeof (double) * n);
  free (h);
}

void
solve_res_lub_3f (struct stokes *sys, const double *u, double *f)
{
  double *u0 = lub_alloc (3 * (size_t) sys->np);
  shift_labo_to_rest_U (sys, sys->np, u, u0);
  solve_res_lub_3f_0 (sys, u0, f);
  free (u0);
}

/* y = (I + M L) x (ewald-3f.c:410-436) */
void
atimes_mob_lub_3f (int n, const double *x, double *y, void *user_data)
{
  struct stokes *sys = user_data;
  if (sys->version != 0)
    {
      fprintf (stderr, "libstokes: version is wrong. reset to F.\n");
      sys->version = 0;
    }
  double *l = lub_alloc ((size_t) n), *ml = lub_alloc ((size_t) n);
  calc_lub_3f (sys, x, l);
  atimes_3all (n, l, ml, sys);
  for (int i = 0; i < n; i++) y[i] = x[i] + ml[i];
  free (l);
  free (ml);
}

void
solve_mob_lub_3f (struct stokes *sys, const double *f, double *u)
{
  const size_t n = (size_t) sys->np * 3;
  double *h = lub_alloc (n);
  memcpy (h, f, sizeof (double) * n);
  lub_host_v (sys, 0, 1, sys->np, h, "solve_mob_lub_3f");
  memcpy (u, h, sizeof (double) * n);
  free (h);
  shift_rest_to_labo_U (sys, sys->np, u);
}

void
solve_mix_lub_3f (struct stokes *sys, const double *f, const double *uf, double *u,
                  double *ff)
{
  const int np = sys->np, nm = sys->nm, nf = np - nm;
  if (nm == np)
    {
      solve_mob_lub_3f (sys, f, u);
      return;
    }
  double *h = lub_alloc ((size_t) np * 3);
  memcpy (h, f, sizeof (double) * 3 * (size_t) nm);
  /* the shift takes the positions of particles 0..nf-1, as the reference
   * does (ewald-3f.c:685) */
  shift_labo_to_rest_U (sys, nf, uf, h + 3 * (size_t) nm);
  lub_host_v (sys, 0, 1, nm, h, "solve_mix_lub_3f");
  memcpy (u, h, sizeof (double) * 3 * (size_t) nm);
  memcpy (ff, h + 3 * (size_t) nm, sizeof (double) * 3 * (size_t) nf);
  free (h);
  shift_rest_to_labo_U (sys, nm, u);
}

void
solve_res_lub_3ft_0 (struct stokes *sys, const double *u, const double *o, double *f,
                     double *t)
{
  double *h = lub_alloc ((size_t) sys->np * 6);
  pack_ft (sys->np, h, u, o);
  lub_host_v (sys, 1, 0, sys->np, h, "solve_res_lub_3ft_0");
  unpack_ft (sys->np, h, f, t);
  free (h);
}

void
solve_res_lub_3ft (struct stokes *sys, const double *u, const double *o, double *f,
                   double *t)
{
  const size_t n3 = (size_t) sys->np * 3;
  double *u0 = lub_alloc (2 * n3), *o0 = u0 + n3;
  shift_labo_to_rest_U (sys, sys->np, u, u0);
  shift_labo_to_rest_O (sys, sys->np, o, o0);
  solve_res_lub_3ft_0 (sys, u0, o0, f, t);
  free (u0);
}

void
solve_mob_lub_3ft (struct stokes *sys, const double *f, const double *t, double *u,
                   double *o)
{
  double *h = lub_alloc ((size_t) sys->np * 6);
  pack_ft (sys->np, h, f, t);
  lub_host_v (sys, 1, 1, sys->np, h, "solve_mob_lub_3ft");
  unpack_ft (sys->np, h, u, o);
  free (h);
  shift_rest_to_labo_U (sys, sys->np, u);
  shift_rest_to_labo_O (sys, sys->np, o);
}

void
solve_mix_lub_3ft (struct stokes *sys, const double *f, const double *t,
                   const double *uf, const double *of, double *u, double *o, double *ff,
                   double *tf)
{
  const int np = sys->np, nm = sys->nm, nf = np - nm;
  if (nm == np)
    {
      solve_mob_lub_3ft (sys, f, t, u, o);
      return;
    }
  const size_t nf3 = (size_t) nf * 3;
  double *h = lub_alloc ((size_t) np * 6), *w = lub_alloc (2 * nf3 + 1);
  pack_ft (nm, h, f, t);
  /* positions of particles 0..nf-1 for the shift, as the reference's
   * (ewald-3ft.c:782-783) */
  shift_labo_to_rest_U (sys, nf, uf, w);
  shift_labo_to_rest_O (sys, nf, of, w + nf3);
  pack_ft (nf, h + 6 * (size_t) nm, w, w + nf3);
  lub_host_v (sys, 1, 1, nm, h, "solve_mix_lub_3ft");
  unpack_ft (nm, h, u, o);
  unpack_ft (nf, h + 6 * (size_t) nm, ff, tf);
  free (h);
  free (w);
  shift_rest_to_labo_U (sys, nm, u);
  shift_rest_to_labo_O (sys, nm, o);
}

void
solve_mix_lub_3fts_0 (struct stokes *sys, const double *f, const double *t,
                      const double *e, const double *uf, const double *of,
                      const double *ef, double *u, double *o, double *s, double *ff,
                      double *tf, double *sf)
{
  const int np = sys->np, nm = sys->nm, nf = np - nm;
  if (nm == np)
    {
      solve_mob_lub_3fts_0 (sys, f, t, e, u, o, s);
      return;
    }
  double *h = lub_alloc ((size_t) np * 11);
  pack11 (nm, h, f, t, e);
  pack11 (nf, h + 11 * (size_t) nm, uf, of, ef);
  lub_host_v (sys, 2, 1, nm, h, "solve_mix_lub_3fts_0");
  unpack11 (nm, h, u, o, s);
  unpack11 (nf, h + 11 * (size_t) nm, ff, tf, sf);
  free (h);
}

void
solve_mix_lub_3fts (struct stokes *sys, const double *f, const double *t, const double *e,
                    const double *uf, const double *of, const double *ef, double *u,
                    double *o, double *s, double *ff, double *tf, double *sf)
{
  const int np = sys->np, nm = sys->nm, nf = np - nm;
  if (nm == np)
    {
      solve_mob_lub_3fts (sys, f, t, e, u, o, s);
      return;
    }
  const size_t nf3 = (size_t) nf * 3;
  double *w = lub_alloc (2 * nf3 + 5 * (size_t) nf + 5 * (size_t) nm);
  double *uf0 = w, *of0 = w + nf3, *ef0 = w + 2 * nf3, *e0 = ef0 + 5 * (size_t) nf;
  /* positions of particles 0..nf-1 (and 0..nm-1 for e), as the
   * reference's shifts (ewald-3fts.c:1288-1291) */
  shift_labo_to_rest_U (sys, nf, uf, uf0);
  shift_labo_to_rest_O (sys, nf, of, of0);
  shift_labo_to_rest_E (sys, nf, ef, ef0);
  shift_labo_to_rest_E (sys, nm, e, e0);
  solve_mix_lub_3fts_0 (sys, f, t, e0, uf0, of0, ef0, u, o, s, ff, tf, sf);
  free (w);
  shift_rest_to_labo_U (sys, nm, u);
  shift_rest_to_labo_O (sys, nm, o);
}
