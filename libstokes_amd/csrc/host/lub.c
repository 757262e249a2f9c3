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
#include "stokes_internal.h"

static int
lub_params (struct stokes *sys, shk_lub *lp, const double **tab, double *f1)
{
  memset (lp, 0, sizeof (*lp));
  lp->lubmin2 = sys->lubmin2;
  lp->lubmax = sys->lubmax;
  lp->nmax = sys->twobody_nmax;
  lp->flag_lub = sys->twobody_lub;
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
  rc = lub_params (sys, &lp, tab, f1);
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
