/* brownian.c -- the mid-point Brownian-dynamics step, F version, on the
 * GPU (restates BD_evolve_mid, brownian.c:2352-2560, with BD_calc_FB's
 * atimes route 1545-1874, BD_calc_forces 2273-2311, evolve_Euler_3all
 * 1889-1930 and check_overlap 299-375).
 *
 * One attempt of the step is a captured graph (see bd_record): draw y,
 * z = M^{-1/2} y by Chebyshev, the accuracy dots, f = F + fact z, u = M(x) f,
 * x_mid = x + dt/2 u, overlap count, u = M(x_mid) f, x_new = x + dt u,
 * overlap count, advance the draw counter, copy the status to pinned host
 * memory.  fact and dt are read by the kernels from device memory, so the
 * graph is replayed unchanged for every dt.  The host then takes the
 * reference's decisions on the status (refit, halve dt, redraw, accept).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "chebyshev.h"
#include "stokes_hip.h"
#include "stokes_internal.h"

#define BD_MAX_ATTEMPTS 64
#define BD_LANCZOS_ITER 160
/* Lanczos Ritz values approach the extreme eigenvalues from inside, the
 * lower one slowly; a Chebyshev fit of 1/sqrt(x) grows fast below its
 * interval, so the lower bound is taken well below the Ritz value, more so
 * at each refit (level = refits so far, at most 3) */
#define BD_LOW_MARGIN(level) (0.1 * ((level) + 1))
#define BD_HIGH_MARGIN(level) (0.02 * ((level) + 1))

struct stokes_hip_bd {
  struct stokes *sys;
  stokes_ctx *c;
  const double *F;
  double peclet, eps, dt_lim, rmin_ol;
  int ncheb;
  unsigned long long seed;
  int fact_fix;
  int real_cache;              /* dense real-space matrix at x (32 products) */
  int have_bounds;
  int fits;                    /* Lanczos fits so far */
  double eig[2];
  double *acoef;               /* host [ncheb] */
  size_t n;                    /* 3 np */
  double *buf;                 /* device: y z w f u xm xn work[4n] */
  double *y, *z, *w, *f, *u, *xm, *xn, *work;
  void *small;                 /* device: counter, scalars, status */
  unsigned long long *counter;
  shk_bd_scalars *sc_dev;
  shk_bd_status *st_dev;
  void *pinned;                /* host: scalars, status */
  shk_bd_scalars *sc_host;
  shk_bd_status *st_host;
  double sc_fact, sc_dt;       /* values held by sc_dev */
  /* captured step */
  void *graph;
  stokes_ctx *graph_c;
  double *graph_x;
  unsigned long graph_gen, graph_uploads;
  int graph_prec, graph_ol, graph_rc;
};

static double
bd_inv_sqrt (double x)
{
  return 1.0 / sqrt (x);
}

/* brownian.c:2177-2187 */
static double
bd_fact (double peclet, double dt)
{
  return peclet > 0.0 ? sqrt (2.0 / (peclet * dt)) : 0.0;
}

static int
bd_fail (int rc, const char *where)
{
  if (rc == SHK_ENOMEM || rc == STOKES_HIP_ENOMEM)
    {
      stokes_hip_set_error ("%s: %s", where, shk_last_error ());
      return STOKES_HIP_ENOMEM;
    }
  stokes_hip_set_error ("%s: %s", where, shk_last_error ());
  return STOKES_HIP_EHIP;
}

void
stokes_hip_bd_destroy (stokes_hip_bd *bd)
{
  if (!bd) return;
  /* the system may already be gone: no context lookup; hipFree waits for
   * the device */
  shk_dev_graph_destroy (bd->graph);
  if (bd->buf) shk_dev_free (NULL, bd->buf);
  if (bd->small) shk_dev_free (NULL, bd->small);
  shk_host_free (bd->pinned);
  free (bd->acoef);
  free (bd);
}

int
stokes_hip_bd_create (struct stokes *sys, const double *F_dev, double peclet,
                      int n_cheb, double eps, double dt_lim, double rmin_ol,
                      unsigned long long seed, stokes_hip_bd **out)
{
  if (!out || !sys)
    {
      stokes_hip_set_error ("bd_create: NULL argument");
      return STOKES_HIP_EARG;
    }
  *out = NULL;
  if (sys->version != 0)
    {
      stokes_hip_set_error ("bd_create: only the F version (version 0) is"
                            " supported, got %d", sys->version);
      return STOKES_HIP_EUNSUP;
    }
  if (sys->a != NULL || sys->shear_mode != 0 || sys->nm != sys->np)
    {
      stokes_hip_set_error ("bd_create: needs equal radii, no shear and every"
                            " particle mobile");
      return STOKES_HIP_EUNSUP;
    }
  if (n_cheb < 2 || !(eps > 0.0) || dt_lim < 0.0 || sys->np <= 0)
    {
      stokes_hip_set_error ("bd_create: invalid n_cheb/eps/dt_lim/np");
      return STOKES_HIP_EARG;
    }
  stokes_ctx *c = stokes_hip_ctx (sys);
  if (!c) return STOKES_HIP_ENODEV;
  stokes_hip_bd *bd = calloc (1, sizeof (*bd));
  if (!bd)
    {
      stokes_hip_set_error ("bd_create: host allocation failed");
      return STOKES_HIP_ENOMEM;
    }
  bd->sys = sys;
  bd->c = c;
  bd->F = F_dev;
  bd->peclet = peclet;
  bd->eps = eps;
  bd->dt_lim = dt_lim;
  bd->rmin_ol = rmin_ol;
  bd->ncheb = n_cheb;
  bd->seed = seed;
  bd->n = (size_t) sys->np * 3;
  /* the dense real-space matrix ((3N)^2 values) when it stays below 4 GiB
   * in FP64 */
  bd->real_cache = sys->periodic && (double) bd->n * (double) bd->n * 8.0
    <= 4.0 * 1024.0 * 1024.0 * 1024.0;
  bd->sc_fact = bd->sc_dt = -1.0;
  bd->acoef = calloc ((size_t) n_cheb, sizeof (double));
  const size_t n = bd->n;
  int rc = bd->acoef ? 0 : SHK_ENOMEM;
  if (!rc) rc = shk_dev_malloc (c->dev, (void **) &bd->buf, sizeof (double) * 11 * n);
  if (!rc) rc = shk_dev_malloc (c->dev, &bd->small, 256);
  if (!rc) rc = shk_host_alloc (&bd->pinned, 256);
  if (rc)
    {
      stokes_hip_bd_destroy (bd);
      stokes_hip_set_error ("bd_create: allocation failed: %s", shk_last_error ());
      return STOKES_HIP_ENOMEM;
    }
  bd->y = bd->buf;
  bd->z = bd->y + n;
  bd->w = bd->z + n;
  bd->f = bd->w + n;
  bd->u = bd->f + n;
  bd->xm = bd->u + n;
  bd->xn = bd->xm + n;
  bd->work = bd->xn + n;
  bd->counter = (unsigned long long *) bd->small;
  bd->sc_dev = (shk_bd_scalars *) ((char *) bd->small + 64);
  bd->st_dev = (shk_bd_status *) ((char *) bd->small + 128);
  bd->sc_host = (shk_bd_scalars *) bd->pinned;
  bd->st_host = (shk_bd_status *) ((char *) bd->pinned + 128);
  memset (bd->pinned, 0, 256);
  /* draw counter = 0, work vectors = 0 */
  unsigned long long zero = 0;
  rc = shk_dev_h2d (c->dev, bd->counter, &zero, sizeof (zero));
  if (!rc) rc = shk_dev_scal (c->dev, bd->buf, NULL, 0.0, 11 * n);
  if (!rc) rc = shk_dev_sync (c->dev);
  if (rc)
    {
      stokes_hip_bd_destroy (bd);
      return bd_fail (rc, "bd_create");
    }
  *out = bd;
  return 0;
}

int
stokes_hip_bd_set_bounds (stokes_hip_bd *bd, double lmin, double lmax)
{
  if (!bd || !(lmin > 0.0) || !(lmax > lmin))
    {
      stokes_hip_set_error ("bd_set_bounds: need 0 < lmin < lmax");
      return STOKES_HIP_EARG;
    }
  bd->eig[0] = lmin;
  bd->eig[1] = lmax;
  chebyshev_coef (bd->ncheb, bd_inv_sqrt, lmin, lmax, bd->acoef);
  bd->have_bounds = 1;
  shk_dev_graph_destroy (bd->graph);   /* coefficients are baked in */
  bd->graph = NULL;
  return 0;
}

int
stokes_hip_bd_set_fact_fix (stokes_hip_bd *bd, int enable)
{
  if (!bd)
    {
      stokes_hip_set_error ("bd_set_fact_fix: NULL handle");
      return STOKES_HIP_EARG;
    }
  bd->fact_fix = enable != 0;
  return 0;
}

int
stokes_hip_bd_set_real_cache (stokes_hip_bd *bd, int enable)
{
  if (!bd)
    {
      stokes_hip_set_error ("bd_set_real_cache: NULL handle");
      return STOKES_HIP_EARG;
    }
  if (enable && !bd->sys->periodic)
    {
      stokes_hip_set_error ("bd_set_real_cache: periodic systems only");
      return STOKES_HIP_EARG;
    }
  bd->real_cache = enable != 0;
  return 0;
}

int
stokes_hip_bd_last_noise (stokes_hip_bd *bd, double *y_dev, double *z_dev)
{
  if (!bd)
    {
      stokes_hip_set_error ("bd_last_noise: NULL handle");
      return STOKES_HIP_EARG;
    }
  stokes_ctx *c = stokes_hip_ctx (bd->sys);
  if (!c) return STOKES_HIP_ENODEV;
  size_t bytes = sizeof (double) * bd->n;
  int rc = 0;
  if (y_dev) rc = shk_dev_d2d (c->dev, y_dev, bd->y, bytes);
  if (!rc && z_dev) rc = shk_dev_d2d (c->dev, z_dev, bd->z, bytes);
  if (!rc) rc = shk_dev_sync (c->dev);
  return rc ? bd_fail (rc, "bd_last_noise") : 0;
}

/* the kernels of one attempt, in stream order (captured) */
static int
bd_record (stokes_hip_bd *bd, shk_dev *d, int prec, int check_ol, double *x)
{
  const struct stokes *s = bd->sys;
  const size_t n = bd->n;
  int rc = shk_dev_bd_status_clear (d, bd->st_dev);
  if (!rc) rc = shk_dev_gauss (d, bd->seed, bd->counter, n, bd->y);
  /* BD_calc_FB at x: z = M^{-1/2} y and its accuracy (chebyshev.c:241-269) */
  if (!rc) rc = shk_dev_set_pos (d, x, 1);
  /* the 32 products at x share one dense real-space matrix */
  if (!rc && bd->real_cache) rc = shk_dev_real_cache (d, 1, prec);
  if (!rc) rc = stokes_hip_cheb_launch (d, prec, bd->ncheb, bd->acoef, 1, n,
                                        bd->y, bd->z, bd->eig[0], bd->eig[1],
                                        bd->work);
  if (!rc) rc = shk_dev_atimes (d, bd->z, bd->w, 1, prec);
  if (!rc) rc = shk_dev_bd_dots (d, bd->y, bd->z, bd->w, n, bd->st_dev);
  /* step 1: drift at x, half step */
  if (!rc) rc = shk_dev_bd_force (d, bd->f, bd->F, bd->z, bd->sc_dev, n);
  if (!rc) rc = shk_dev_atimes (d, bd->f, bd->u, 1, prec);
  if (!rc) rc = shk_dev_bd_euler (d, bd->xm, x, bd->u, bd->sc_dev, 0.5, n);
  if (!rc && check_ol)
    rc = shk_dev_bd_overlap (d, bd->xm, NULL, s->np, s->periodic, s->lx, s->ly,
                             s->lz, 0, 0.0, bd->rmin_ol, &bd->st_dev->overlap[0]);
  /* step 2: drift at x_mid with the same forces, full step from x */
  if (!rc && bd->real_cache) rc = shk_dev_real_cache (d, 0, prec);
  if (!rc) rc = shk_dev_set_pos (d, bd->xm, 1);
  if (!rc) rc = shk_dev_atimes (d, bd->f, bd->u, 1, prec);
  if (!rc) rc = shk_dev_bd_euler (d, bd->xn, x, bd->u, bd->sc_dev, 1.0, n);
  if (!rc && check_ol)
    rc = shk_dev_bd_overlap (d, bd->xn, NULL, s->np, s->periodic, s->lx, s->ly,
                             s->lz, 0, 0.0, bd->rmin_ol, &bd->st_dev->overlap[1]);
  if (!rc) rc = shk_dev_counter_add (d, bd->counter, 1);
  if (!rc) rc = shk_dev_d2h_async (d, bd->st_host, bd->st_dev, sizeof (shk_bd_status));
  return rc;
}

static int
bd_capture (stokes_hip_bd *bd, stokes_ctx *c, double *x)
{
  shk_dev *d = c->dev;
  const int check_ol = bd->sys->rmin == 0.0;
  shk_dev_graph_destroy (bd->graph);
  bd->graph = NULL;
  /* every work buffer the operator uses must exist before the capture:
   * one eager product allocates them */
  shk_dev_timing (d, 0);
  int rc = shk_dev_set_pos (d, x, 1);
  if (!rc) rc = shk_dev_atimes (d, bd->z, bd->w, 1, c->prec);
  if (!rc && bd->real_cache)
    {
      rc = shk_dev_real_cache (d, 1, c->prec);
      if (!rc) rc = shk_dev_atimes (d, bd->z, bd->w, 1, c->prec);
      int rc2 = shk_dev_real_cache (d, 0, c->prec);
      if (!rc) rc = rc2;
    }
  if (!rc) rc = shk_dev_sync (d);
  if (rc)
    {
      shk_dev_timing (d, c->timing);
      return bd_fail (rc, "bd_step (warm-up)");
    }
  rc = shk_dev_capture_begin (d);
  if (rc)
    {
      shk_dev_timing (d, c->timing);
      return bd_fail (rc, "bd_step (capture)");
    }
  rc = bd_record (bd, d, c->prec, check_ol, x);
  if (rc)
    {
      shk_dev_capture_abort (d);
      shk_dev_timing (d, c->timing);
      return bd_fail (rc, "bd_step (capture)");
    }
  rc = shk_dev_capture_end (d, &bd->graph);
  shk_dev_timing (d, c->timing);
  if (rc) return bd_fail (rc, "bd_step (instantiate)");
  bd->graph_c = c;
  bd->graph_x = x;
  bd->graph_gen = shk_dev_generation (d);
  bd->graph_uploads = c->uploads;
  bd->graph_prec = c->prec;
  bd->graph_ol = check_ol;
  bd->graph_rc = bd->real_cache;
  return 0;
}

static int
bd_fit_bounds (stokes_hip_bd *bd, stokes_ctx *c, double *x)
{
  double l0, l1;
  int level = bd->fits < 3 ? bd->fits : 3;
  c->pos_dev = x;
  int rc = 0;
  if (bd->real_cache)
    {
      /* every Lanczos product is at x: one dense real-space matrix */
      rc = stokes_hip_prepare (c, 1);
      if (!rc) rc = shk_dev_real_cache (c->dev, 1, c->prec);
      if (rc) return bd_fail (rc, "bd_fit_bounds");
    }
  rc = stokes_hip_lanczos_bounds (bd->sys, BD_LANCZOS_ITER * (level + 1),
                                  1 + bd->seed + bd->fits, 0.0, &l0, &l1);
  if (bd->real_cache) shk_dev_real_cache (c->dev, 0, c->prec);
  if (rc) return rc;
  bd->fits++;
  return stokes_hip_bd_set_bounds (bd, l0 * (1.0 - BD_LOW_MARGIN (level)),
                                   l1 * (1.0 + BD_HIGH_MARGIN (level)));
}

int
stokes_hip_bd_step (stokes_hip_bd *bd, double *x_dev, double dt,
                    stokes_hip_bd_info *info)
{
  if (!bd || !x_dev || !(dt > 0.0))
    {
      stokes_hip_set_error ("bd_step: need a handle, positions and dt > 0");
      return STOKES_HIP_EARG;
    }
  stokes_ctx *c = stokes_hip_ctx (bd->sys);
  if (!c) return STOKES_HIP_ENODEV;
  shk_dev *d = c->dev;
  const double *pos_dev_saved = c->pos_dev;
  stokes_hip_bd_info inf;
  memset (&inf, 0, sizeof (inf));
  int rc = 0;
  c->pos_dev = x_dev;
  /* parameters of sys current on the device (positions from x_dev) */
  rc = stokes_hip_prepare (c, 1);
  if (rc) goto out;
  if (!bd->have_bounds)
    {
      rc = bd_fit_bounds (bd, c, x_dev);
      if (rc) goto out;
    }

  double dt_local = dt;
  double fact = bd_fact (bd->peclet, dt_local);
  int fresh = 1;           /* new Gaussian draw for this attempt */
  int refit_done = 0;      /* BD_calc_FB refits at most once per draw */
  for (;;)
    {
      if (inf.attempts >= BD_MAX_ATTEMPTS)
        {
          stokes_hip_set_error ("bd_step: no acceptable step after %d attempts",
                                inf.attempts);
          rc = STOKES_HIP_EARG;
          goto out;
        }
      if (!bd->graph || bd->graph_c != c || bd->graph_x != x_dev
          || bd->graph_gen != shk_dev_generation (d)
          || bd->graph_uploads != c->uploads || bd->graph_prec != c->prec
          || bd->graph_ol != (bd->sys->rmin == 0.0)
          || bd->graph_rc != bd->real_cache)
        {
          rc = bd_capture (bd, c, x_dev);
          if (rc) goto out;
        }
      if (fact != bd->sc_fact || dt_local != bd->sc_dt)
        {
          bd->sc_host->fact = fact;
          bd->sc_host->dt = dt_local;
          rc = shk_dev_h2d_async (d, bd->sc_dev, bd->sc_host, sizeof (shk_bd_scalars));
          if (rc) { rc = bd_fail (rc, "bd_step"); goto out; }
          bd->sc_fact = fact;
          bd->sc_dt = dt_local;
        }
      if (!fresh)
        {
          /* replay the previous draw */
          rc = shk_dev_counter_add (d, bd->counter, (unsigned long long) -1);
          if (rc) { rc = bd_fail (rc, "bd_step"); goto out; }
        }
      rc = shk_dev_graph_launch (d, bd->graph);
      if (!rc) rc = shk_dev_sync (d);
      if (rc) { rc = bd_fail (rc, "bd_step (graph)"); goto out; }
      inf.attempts++;
      const shk_bd_status st = *bd->st_host;
      inf.overlaps[0] = st.overlap[0];
      inf.overlaps[1] = st.overlap[1];
      if (!(st.zmz > 0.0))
        {
          stokes_hip_set_error ("bd_step: z.Mz = %g is not positive", st.zmz);
          rc = STOKES_HIP_EARG;
          goto out;
        }
      inf.err_cheb = fabs (st.zmz - st.yy) / st.zmz;
      if (inf.err_cheb > bd->eps && !refit_done)
        {
          /* brownian.c:1707-1727: new bounds at x, same draw */
          rc = bd_fit_bounds (bd, c, x_dev);
          if (rc) goto out;
          inf.refits++;
          refit_done = 1;
          fresh = 0;
          continue;
        }
      int stage = st.overlap[0] > 0 ? 1 : (st.overlap[1] > 0 ? 2 : 0);
      if (stage == 0)
        break;
      /* brownian.c:2435-2459 (stage 1), 2501-2525 (stage 2) */
      dt_local *= 0.5;
      inf.halvings++;
      if (dt_local < bd->dt_lim)
        {
          dt_local = dt;
          if (bd->fact_fix) fact = bd_fact (bd->peclet, dt_local);
          inf.redraws++;
          fresh = 1;
          refit_done = 0;
        }
      else
        {
          fact = bd_fact (bd->peclet, dt_local);
          fresh = 0;
        }
    }
  /* accepted: x <- x_new */
  rc = shk_dev_d2d (d, x_dev, bd->xn, sizeof (double) * bd->n);
  if (rc) { rc = bd_fail (rc, "bd_step"); goto out; }
  inf.dt = dt_local;
out:
  inf.eig[0] = bd->eig[0];
  inf.eig[1] = bd->eig[1];
  c->pos_dev = pos_dev_saved;
  if (info) *info = inf;
  return rc;
}
