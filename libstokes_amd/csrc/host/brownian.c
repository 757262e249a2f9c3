/* brownian.c -- the mid-point Brownian-dynamics step on the GPU, F, FT and
 * FTS versions (restates BD_evolve_mid, brownian.c:2352-2560, with
 * BD_calc_FB's atimes route 1545-1874, BD_atimes_mob_FU 494-739,
 * BD_calc_forces 2273-2311, evolve_Euler_3all 1889-1930 and check_overlap
 * 299-375).
 *
 * One attempt of the step (see bd_record): draw y, z = M^{-1/2} y by
 * Chebyshev, the accuracy dots, f = F + fact z, u = M(x) f, x_mid = x + dt/2
 * u, overlap count, u = M(x_mid) f, x_new = x + dt u and the quaternions
 * q_new = q + dt dQ/dt(q, O) for FT / FTS (the mid-point scheme;
 * BD_evolve_BB03, 2579-2920, and BD_evolve_BM97, 2970-3190, change the
 * intermediate step and the velocity combination), overlap count, advance the draw
 * counter, copy the status to pinned host memory.  M is the mobility of
 * the forces (F) or of the forces and torques (FT, FTS: particle-wise (F,
 * T) -> (U, O)); for FTS every product is the mobility problem with E = 0
 * solved by GMRES (BD_atimes_mob_FU, brownian.c:597-634).  For F and FT an
 * attempt is one captured graph; fact and dt are read by the kernels from
 * device memory, so the graph is replayed unchanged for every dt.  For FTS
 * the GMRES iterations are steered from the host and the attempt runs
 * eagerly.  The host then takes the reference's decisions on the status
 * (refit, halve dt, redraw, accept).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "chebyshev.h"
#include "lub.h"
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
  const double *F, *T;         /* device [3 np] or NULL */
  int ver;                     /* sys->version */
  int np, nm;                  /* particles, mobile particles (0..nm-1) */
  double *posfull;             /* device [3 np]: mobile from x, fixed from
                                  sys->pos (nm < np only) */
  double *fpad, *upad;         /* device [D np]: zero-padded F / FT products */
  double *sv, *sqsv;           /* device [n]: the Global scaling of the noise
                                  operator (1/a, 1/a^3) and its square root;
                                  NULL for equal radii */
  double *a_dev;               /* device [np] radii, NULL for equal radii */
  /* the reference's random stream (MT19937 + polar, host side): draws
   * written to pinned memory before an attempt, copied by the attempt */
  stokes_hip_mt *mt;
  double *y_host;
  double peclet, eps, dt_lim, rmin_ol;
  int ncheb;
  unsigned long long seed;
  int fact_fix;
  int scheme;                  /* 0 mid-point, 1 BB03, 2 BM97 */
  double bb_n;                 /* BB03 step parameter */
  int real_cache;              /* dense real-space matrix at x (32 products) */
  int lub;                     /* flag_lub: the drift with lubrication */
  int have_bounds;
  int fits;                    /* Lanczos fits so far */
  double eig[2];
  double *acoef;               /* host [ncheb] */
  size_t n;                    /* operator dimension: 3 np (F), 6 np */
  double *buf;                 /* device: y z w f u work[4n] xm xn qn */
  double *y, *z, *w, *f, *u, *xm, *xn, *qn, *work, *u0, *f2;
  const double *pcur;          /* device: positions of the products now */
  int shear;                   /* sys->shear_mode of the step */
  double t0, s0;               /* BD_set_shear_shift_ref (brownian.c:274) */
  double sh_cur;               /* the cell shift the system holds */
  double sh_after;             /* the shift the last attempt's second stage
                                  set */
  shk_flow flow;               /* the system's imposed flow (Ui, Oi, Ei) */
  int has_flow;
  shk_flow graph_flow;         /* the flow the graph was captured with */
  double *xs;                  /* device [3 nm]: the intermediate
                                  configuration a failed attempt leaves in
                                  the system (the retry's first stage) */
  /* FTS: GMRES of the noise products (brownian.c:597-600: restart 20,
   * max 2000, eps 1e-6) and work space */
  int fts_m, fts_max;
  double fts_eps;
  double *fts_buf, *fts_b, *fts_x, *fts_ws;
  long fts_iters;
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
  double *graph_x, *graph_q;
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
  if (bd->fts_buf) shk_dev_free (NULL, bd->fts_buf);
  shk_host_free (bd->y_host);
  free (bd->mt);
  if (bd->small) shk_dev_free (NULL, bd->small);
  shk_host_free (bd->pinned);
  free (bd->acoef);
  free (bd);
}

/* no imposed flow: Ui, Oi, Ei all zero */
static int
bd_fluid_at_rest (const struct stokes *sys)
{
  for (int k = 0; k < 3; k++)
    if (sys->Ui[k] != 0.0 || sys->Oi[k] != 0.0) return 0;
  for (int k = 0; k < 5; k++)
    if (sys->Ei[k] != 0.0) return 0;
  return 1;
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
  if (sys->version < 0 || sys->version > 2)
    {
      stokes_hip_set_error ("bd_create: version %d", sys->version);
      return STOKES_HIP_EARG;
    }
  if ((sys->shear_mode != 0 && !sys->periodic) || sys->nm < 1
      || sys->nm > sys->np)
    {
      stokes_hip_set_error ("bd_create: needs 1 <= nm <= np (and a periodic"
                            " system for shear)");
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
  bd->ver = sys->version;
  bd->np = sys->np;
  bd->nm = sys->nm;
  bd->n = (size_t) sys->nm * (sys->version == 0 ? 3 : 6);
  bd->fts_m = 20;
  bd->fts_max = 2000;
  bd->fts_eps = 1e-6;
  /* the dense real-space matrix ((D N)^2 values) when it stays below 4 GiB
   * in FP64 */
  const double dn = (double) sys->np * stokes_dim (sys->version);
  bd->real_cache = sys->periodic && dn * dn * 8.0
    <= 4.0 * 1024.0 * 1024.0 * 1024.0;
  bd->sc_fact = bd->sc_dt = -1.0;
  bd->acoef = calloc ((size_t) n_cheb, sizeof (double));
  const size_t n = bd->n, np = (size_t) sys->np;
  const size_t dfull = (size_t) sys->np * (sys->version == 0 ? 3 : 6);
  const size_t nbuf = 11 * n + 10 * np + 3 * np + 2 * dfull
    + (sys->a ? 2 * n + np : 0) + 3 * np;
  int rc = bd->acoef ? 0 : SHK_ENOMEM;
  if (!rc) rc = shk_dev_malloc (c->dev, (void **) &bd->buf, sizeof (double) * nbuf);
  /* GMRES work space: the FTS products, and the F / FT drift with fixed
   * particles (a mixed problem) */
  const int dmix = sys->version == 2 ? 11 : (sys->version == 0 ? 3 : 6);
  if (!rc && (bd->ver == 2 || sys->nm < sys->np))
    rc = shk_dev_malloc (c->dev, (void **) &bd->fts_buf, sizeof (double)
                         * (2 * dmix * np + stokes_hip_mix_ws (sys->np, dmix, 64)));
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
  bd->work = bd->u + n;
  bd->xm = bd->work + 4 * n;
  bd->xn = bd->xm + 3 * np;
  bd->qn = bd->xn + 3 * np;
  bd->u0 = bd->qn + 4 * np;
  bd->f2 = bd->u0 + n;
  bd->posfull = bd->f2 + n;
  bd->fpad = bd->posfull + 3 * np;
  bd->upad = bd->fpad + dfull;
  if (sys->a)
    {
      bd->sv = bd->upad + dfull;
      bd->sqsv = bd->sv + n;
      bd->a_dev = bd->sqsv + n;
    }
  bd->xs = bd->buf + nbuf - 3 * np;
  bd->bb_n = 1.0;
  if (bd->fts_buf)
    {
      bd->fts_b = bd->fts_buf;
      bd->fts_x = bd->fts_b + dmix * np;
      bd->fts_ws = bd->fts_x + dmix * np;
    }
  bd->counter = (unsigned long long *) bd->small;
  bd->sc_dev = (shk_bd_scalars *) ((char *) bd->small + 64);
  bd->st_dev = (shk_bd_status *) ((char *) bd->small + 128);
  bd->sc_host = (shk_bd_scalars *) bd->pinned;
  bd->st_host = (shk_bd_status *) ((char *) bd->pinned + 128);
  memset (bd->pinned, 0, 256);
  /* draw counter = 0, work vectors = 0 */
  unsigned long long zero = 0;
  rc = shk_dev_h2d (c->dev, bd->counter, &zero, sizeof (zero));
  if (!rc) rc = shk_dev_scal (c->dev, bd->buf, NULL, 0.0, nbuf);
  if (!rc && bd->fts_buf)
    rc = shk_dev_scal (c->dev, bd->fts_buf, NULL, 0.0, 2 * dmix * np);
  if (!rc && sys->a)
    {
      /* BD_scale_P_to_G_uo (brownian.c:430-480): U / a, O / a^3 */
      const int dd = sys->version == 0 ? 3 : 6;
      double *h = malloc (sizeof (double) * (2 * n + np));
      if (!h) rc = SHK_ENOMEM;
      else
        {
          for (int i = 0; i < sys->nm; i++)
            for (int q = 0; q < dd; q++)
              {
                const double a = sys->a[i];
                const double v = q < 3 ? 1.0 / a : 1.0 / (a * a * a);
                h[dd * i + q] = v;
                h[n + dd * i + q] = sqrt (v);
              }
          for (size_t i = 0; i < np; i++) h[2 * n + i] = sys->a[i];
          rc = shk_dev_h2d (c->dev, bd->sv, h, sizeof (double) * (2 * n + np));
          free (h);
        }
    }
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
stokes_hip_bd_set_rng_mt19937 (stokes_hip_bd *bd, unsigned long seed)
{
  if (!bd)
    {
      stokes_hip_set_error ("bd_set_rng_mt19937: NULL handle");
      return STOKES_HIP_EARG;
    }
  if (!bd->mt)
    {
      bd->mt = malloc (sizeof (stokes_hip_mt));
      void *h = NULL;
      if (!bd->mt || shk_host_alloc (&h, sizeof (double) * bd->n))
        {
          free (bd->mt);
          bd->mt = NULL;
          stokes_hip_set_error ("bd_set_rng_mt19937: allocation failed");
          return STOKES_HIP_ENOMEM;
        }
      bd->y_host = h;
    }
  stokes_hip_mt_seed (bd->mt, seed);
  shk_dev_graph_destroy (bd->graph);   /* the draw source is baked in */
  bd->graph = NULL;
  return 0;
}

int
stokes_hip_bd_set_scheme (stokes_hip_bd *bd, int scheme, double bb_n)
{
  if (!bd || scheme < 0 || scheme > 2 || (scheme == 1 && !(bb_n > 0.0)))
    {
      stokes_hip_set_error ("bd_set_scheme: scheme 0 (mid-point), 1 (BB03,"
                            " bb_n > 0) or 2 (BM97)");
      return STOKES_HIP_EARG;
    }
  bd->scheme = scheme;
  if (scheme == 1) bd->bb_n = bb_n;
  shk_dev_graph_destroy (bd->graph);   /* the scheme is baked in */
  bd->graph = NULL;
  return 0;
}

int
stokes_hip_bd_set_lub (stokes_hip_bd *bd, int enable)
{
  if (!bd)
    {
      stokes_hip_set_error ("bd_set_lub: no integrator");
      return STOKES_HIP_EARG;
    }
  if (enable && !bd->fts_buf)
    {
      /* the known values and the solution of the lubricated mixed problem */
      const int dmix = bd->ver == 2 ? 11 : (bd->ver == 0 ? 3 : 6);
      const size_t len = 2 * (size_t) dmix * bd->np;
      int rc = shk_dev_malloc (bd->c->dev, (void **) &bd->fts_buf, sizeof (double) * len);
      if (!rc) rc = shk_dev_scal (bd->c->dev, bd->fts_buf, NULL, 0.0, len);
      if (rc)
        {
          stokes_hip_set_error ("bd_set_lub: allocation failed: %s", shk_last_error ());
          return STOKES_HIP_ENOMEM;
        }
      bd->fts_b = bd->fts_buf;
      bd->fts_x = bd->fts_b + (size_t) dmix * bd->np;
    }
  bd->lub = enable != 0;
  shk_dev_graph_destroy (bd->graph);
  bd->graph = NULL;
  return 0;
}

int
stokes_hip_bd_set_torque (stokes_hip_bd *bd, const double *T_dev)
{
  if (!bd || bd->ver == 0)
    {
      stokes_hip_set_error ("bd_set_torque: needs an FT or FTS integrator");
      return STOKES_HIP_EARG;
    }
  bd->T = T_dev;
  shk_dev_graph_destroy (bd->graph);   /* the pointer is baked in */
  bd->graph = NULL;
  return 0;
}

int
stokes_hip_bd_set_fts_iter (stokes_hip_bd *bd, int restart, int max,
                            double eps)
{
  if (!bd || bd->ver != 2 || restart < 1 || restart > 64 || max < 1
      || !(eps > 0.0))
    {
      stokes_hip_set_error ("bd_set_fts_iter: needs an FTS integrator, "
                            "1 <= restart <= 64, max >= 1, eps > 0");
      return STOKES_HIP_EARG;
    }
  bd->fts_m = restart;
  bd->fts_max = max;
  bd->fts_eps = eps;
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

/* the mobility of the integrator, y = M x on device vectors of bd->n:
 * F and FT the operator itself; FTS the mobility problem (F, T, E = 0) ->
 * (U, O) by GMRES (BD_atimes_mob_FU, brownian.c:597-634), with the noise
 * parameters (noise != 0) or those of stokes_set_iter (the step's solves,
 * solve_mix_3all) */
static int bd_mob (stokes_hip_bd *bd, shk_dev *d, int prec, const double *x,
                   double *y, int noise);

/* scale != 0 applies the Global scaling of BD_atimes_mob_FU (unequal radii,
 * brownian.c:736): the noise operator; the drift products do not carry it */
static int
bd_op (stokes_hip_bd *bd, shk_dev *d, int prec, const double *x, double *y,
       int noise, int scale)
{
  int rc = bd_mob (bd, d, prec, x, y, noise);
  if (!rc && scale && bd->sv) rc = shk_dev_dmul (d, y, bd->sv, bd->n);
  return rc;
}

static int
bd_mob (stokes_hip_bd *bd, shk_dev *d, int prec, const double *x, double *y,
        int noise)
{
  /* the drift is solve_mix_3all in the labo frame: with an imposed flow
   * the known velocities and strains go to the fluid-rest frame first and
   * the velocities come back with the flow at the positions of the product
   * (shift_rest_to_labo_U / _O); the noise operator stays in the rest
   * frame */
  const int flow = !noise && bd->has_flow;
  const int nf = bd->np - bd->nm;
  if (!noise && bd->lub)
    {
      /* the drift with lubrication (flag_lub: solve_mix_3all ->
       * solve_mix_lub_3f / 3ft / 3fts, ode.c:41-160): the known values
       * [(F, T, E = 0)_mobile, (U, O, E)_fixed] in the fluid-rest frame
       * (the fixed particles at rest; with an imposed flow the shifts of
       * the drift without lubrication), the lubricated mixed problem by
       * GMRES with stokes_set_iter's parameters, the mobile velocities
       * back in the labo frame */
      const int D = bd->ver == 0 ? 3 : (bd->ver == 1 ? 6 : 11);
      const int dv = bd->ver == 2 ? 6 : D;
      int rc = 0, it = 0;
      double rr = 0.0;
      if (bd->ver == 2) rc = shk_dev_ft_fts (d, x, bd->nm, 0, bd->fts_b);
      else rc = shk_dev_scal (d, bd->fts_b, x, 1.0, bd->n);
      if (!rc && nf > 0)
        rc = shk_dev_scal (d, bd->fts_b + (size_t) D * bd->nm, NULL, 0.0, (size_t) D * nf);
      if (!rc && flow && bd->ver == 2)
        rc = shk_dev_bd_imposed (d, bd->fts_b, bd->pcur, bd->nm, 11, 6, 11, &bd->flow, -1.0);
      if (!rc && flow && nf > 0)
        rc = shk_dev_bd_imposed (d, bd->fts_b + (size_t) D * bd->nm, bd->pcur, nf, D, 0, D,
                                 &bd->flow, -1.0);
      if (!rc) rc = stokes_hip_solve_mix_lub_here (bd->sys, bd->fts_b, bd->fts_x, &it, &rr);
      bd->fts_iters += it;
      if (!rc)
        {
          if (bd->ver == 2) rc = shk_dev_ft_fts (d, bd->fts_x, bd->nm, 1, y);
          else rc = shk_dev_scal (d, y, bd->fts_x, 1.0, bd->n);
        }
      if (!rc && flow)
        rc = shk_dev_bd_imposed (d, y, bd->pcur, bd->nm, dv, 0, dv, &bd->flow, 1.0);
      return rc;
    }
  if (bd->ver != 2)
    {
      const int D = bd->ver == 0 ? 3 : 6;
      if (bd->nm == bd->np)
        {
          int rc = shk_dev_atimes (d, x, y, 1, prec);
          if (!rc && flow)
            rc = shk_dev_bd_imposed (d, y, bd->pcur, bd->nm, D, 0, D, &bd->flow, 1.0);
          return rc;
        }
      if (noise)
        {
          /* the noise operator with fixed particles: zero forces there, the
           * mobile rows kept (brownian.c:521-547, 562-588); the tail of
           * fpad stays zero */
          int rc = shk_dev_scal (d, bd->fpad, x, 1.0, bd->n);
          if (!rc) rc = shk_dev_atimes (d, bd->fpad, bd->upad, 1, prec);
          if (!rc) rc = shk_dev_scal (d, y, bd->upad, 1.0, bd->n);
          return rc;
        }
      /* the drift (solve_mix_3all -> solve_mix_3f / 3ft, ewald-3f.c:269-330):
       * the fixed particles at rest (their velocities 0, forces unknown),
       * a mixed problem by GMRES with stokes_set_iter's parameters; the
       * tail of fts_b holds the fixed particles' rest-frame velocities:
       * zero, or minus the imposed flow at the positions of particles
       * 0..nf-1, as the shift of ewald-3f.c:301 takes them */
      int m, maxit, it = 0;
      double eps, rr = 0.0;
      stokes_hip_iter_params (bd->sys, &m, &maxit, &eps);
      int rc = shk_dev_scal (d, bd->fts_b, x, 1.0, bd->n);
      if (!rc) rc = shk_dev_scal (d, bd->fts_b + bd->n, NULL, 0.0, (size_t) D * nf);
      if (!rc && flow)
        rc = shk_dev_bd_imposed (d, bd->fts_b + bd->n, bd->pcur, nf, D, 0, D, &bd->flow, -1.0);
      if (!rc) rc = stokes_hip_mix_solve (bd->c, D, D, 1, bd->nm, bd->fts_b,
                                          bd->fts_x, m, maxit, eps, bd->fts_ws,
                                          &it, &rr);
      bd->fts_iters += it;
      if (!rc) rc = shk_dev_scal (d, y, bd->fts_x, 1.0, bd->n);
      if (!rc && flow)
        rc = shk_dev_bd_imposed (d, y, bd->pcur, bd->nm, D, 0, D, &bd->flow, 1.0);
      return rc;
    }
  /* FTS: the mixed problem with the fixed particles at rest (U = O = E = 0
   * in the rest frame; with an imposed flow the drift's known values are
   * E = -Ei for the mobile particles and -(u_inf, Oi, Ei) for the fixed
   * ones, ewald-3fts.c:603-667); nm = np is the mobility problem */
  int m = bd->fts_m, maxit = bd->fts_max, it = 0;
  double eps = bd->fts_eps, rr = 0.0;
  if (!noise) stokes_hip_iter_params (bd->sys, &m, &maxit, &eps);
  int rc = shk_dev_ft_fts (d, x, bd->nm, 0, bd->fts_b);
  if (!rc && nf > 0)
    rc = shk_dev_scal (d, bd->fts_b + (size_t) 11 * bd->nm, NULL, 0.0, (size_t) 11 * nf);
  if (!rc && flow)
    rc = shk_dev_bd_imposed (d, bd->fts_b, bd->pcur, bd->nm, 11, 6, 11, &bd->flow, -1.0);
  if (!rc && flow)
    rc = shk_dev_bd_imposed (d, bd->fts_b + (size_t) 11 * bd->nm, bd->pcur, nf, 11, 0, 11,
                             &bd->flow, -1.0);
  if (!rc) rc = stokes_hip_fts_solve (bd->c, 1, bd->nm, bd->fts_b, bd->fts_x,
                                      m, maxit, eps, bd->fts_ws, &it, &rr);
  bd->fts_iters += it;
  if (!rc) rc = shk_dev_ft_fts (d, bd->fts_x, bd->nm, 1, y);
  if (!rc && flow)
    rc = shk_dev_bd_imposed (d, y, bd->pcur, bd->nm, 6, 0, 6, &bd->flow, 1.0);
  return rc;
}

/* positions of all particles for the operator: x [3 nm] itself when every
 * particle is mobile, else x copied into the mobile part of posfull (the
 * fixed part uploaded by bd_step_q) */
static const double *
bd_pos (stokes_hip_bd *bd, shk_dev *d, const double *x, int *rc)
{
  if (bd->nm == bd->np) return x;
  *rc = shk_dev_scal (d, bd->posfull, x, 1.0, 3 * (size_t) bd->nm);
  return bd->posfull;
}

/* Clenshaw recursion of chebyshev.c:163-228 over bd_op: z = sum_j a_j
 * T_j(M') y (the same launches as stokes_hip_cheb_launch for F and FT) */
static int
bd_cheb (stokes_hip_bd *bd, shk_dev *d, int prec, const double *y, double *z)
{
  const size_t len = bd->n;
  const double *a = bd->acoef;
  const double e0 = bd->eig[0], e1 = bd->eig[1];
  double *x = bd->work, *b0 = x + len, *b1 = b0 + len, *b2 = b1 + len;
  const double da = 2.0 / (e1 - e0), db = -(e1 + e0) / (e1 - e0);
  const int ncheb = bd->ncheb;
  int rc = shk_dev_scal (d, b0, NULL, 0.0, len);
  if (!rc) rc = shk_dev_scal (d, b1, y, a[ncheb - 1], len);
  for (int j = ncheb - 2; j >= 1 && !rc; j--)
    {
      rc = bd_op (bd, d, prec, b1, x, 1, 1);
      if (!rc) rc = shk_dev_axpby4 (d, b2, x, b1, b0, y, 2.0 * da, 2.0 * db,
                                    a[j], len);
      double *t = b0; b0 = b1; b1 = b2; b2 = t;
    }
  if (!rc) rc = bd_op (bd, d, prec, b1, x, 1, 1);
  if (!rc) rc = shk_dev_axpby4 (d, z, x, b1, b0, y, da, db, a[0], len);
  return rc;
}

/* the kernels of one attempt, in stream order (captured for F and FT) */
/* sheared systems: the cell shift of the products that follow (the
 * system's parameters uploaded again, an eager attempt only) */
static int
bd_shift (stokes_hip_bd *bd, double shift)
{
  bd->sys->shear_shift = shift;
  return stokes_hip_prepare (bd->c, 1);
}

static int
bd_record (stokes_hip_bd *bd, shk_dev *d, int prec, int check_ol, double *x,
           double *q, const double *p1, int noise, const double *sh)
{
  /* sh (sheared systems): the cell shifts of the noise and first stage,
   * of the second stage, and of BB03's deterministic product (which the
   * reference takes at t and at x: BD_add_FP (BD, x, f) moves the system
   * back there, brownian.c:2209-2212, 2800-2816) */
  const int shear = bd->shear && sh;
  const struct stokes *s = bd->sys;
  const size_t n = bd->n;
  const size_t n3 = (size_t) bd->nm * 3;
  /* p1: the positions the system holds for the noise and the first stage
   * (x, or the intermediate configuration of a failed attempt); noise = 0
   * keeps z of the previous attempt (REDO_scale) */
  const int cache = bd->real_cache && noise;
  int rc = shk_dev_bd_status_clear (d, bd->st_dev);
  if (!rc && noise)
    rc = bd->mt ? shk_dev_h2d_async (d, bd->y, bd->y_host, sizeof (double) * n)
                : shk_dev_gauss (d, bd->seed, bd->counter, n, bd->y);
  /* BD_calc_FB: z = M^{-1/2} y and its accuracy (chebyshev.c:241-269) */
  if (!rc && shear) rc = bd_shift (bd, sh[0]);
  const double *px = rc ? NULL : bd_pos (bd, d, p1, &rc);
  if (!rc) rc = shk_dev_set_pos (d, px, 1);
  bd->pcur = px;
  /* the products at p1 share one dense real-space matrix */
  if (!rc && cache) rc = shk_dev_real_cache (d, 1, prec);
  if (!rc && noise) rc = bd_cheb (bd, d, prec, bd->y, bd->z);
  if (!rc && noise) rc = bd_op (bd, d, prec, bd->z, bd->w, 1, 1);
  if (!rc && noise) rc = shk_dev_bd_dots (d, bd->y, bd->z, bd->w, n, bd->st_dev);
  /* stage 1: drift at x (BB03: the Brownian force alone), Euler step to
   * the intermediate configuration: dt / 2 (mid-point), dt / n (BB03), dt
   * (BM97); the intermediate orientation is not formed, no force here
   * depends on it */
  const int sch = bd->scheme;
  const double h1 = sch == 0 ? 0.5 : (sch == 1 ? 1.0 / bd->bb_n : 1.0);
  const double *Fx = sch == 1 ? NULL : bd->F, *Tx = sch == 1 ? NULL : bd->T;
  if (!rc)
    rc = bd->ver == 0
      ? shk_dev_bd_force (d, bd->f, Fx, bd->z, bd->sc_dev, n)
      : shk_dev_bd_force_ft (d, bd->f, Fx, Tx, bd->z, bd->sc_dev, bd->nm);
  if (!rc) rc = bd_op (bd, d, prec, bd->f, bd->u0, 0, 0);
  if (!rc)
    rc = bd->ver == 0
      ? shk_dev_bd_euler (d, bd->xm, x, bd->u0, bd->sc_dev, h1, n3)
      : shk_dev_bd_euler_ft (d, bd->xm, x, NULL, NULL, bd->u0, bd->sc_dev, h1,
                             bd->nm);
  /* the mobile particles only (check_overlap, brownian.c:305-310) */
  if (!rc && check_ol)
    rc = shk_dev_bd_overlap (d, bd->xm, bd->a_dev, bd->nm, s->periodic, s->lx, s->ly,
                             s->lz, shear ? s->shear_mode : 0, shear ? sh[0] : 0.0,
                             bd->rmin_ol, &bd->st_dev->overlap[0]);
  /* stage 2: drift at the intermediate configuration with the same forces;
   * mid-point: u = M(x1) f; BM97: u = (u0 + M(x1) f) / 2; BB03: u =
   * M(x) F + u0 + n/2 (M(x1) fact z - u0), its deterministic product back
   * at x (brownian.c:2800-2816); then the full step from x */
  if (!rc && cache) rc = shk_dev_real_cache (d, 0, prec);
  if (!rc && shear) rc = bd_shift (bd, sh[1]);
  const double *pm = rc ? NULL : bd_pos (bd, d, bd->xm, &rc);
  if (!rc) rc = shk_dev_set_pos (d, pm, 1);
  bd->pcur = pm;
  if (!rc) rc = bd_op (bd, d, prec, bd->f, bd->u, 0, 0);
  if (!rc && sch == 2) rc = shk_dev_axpby (d, bd->u, bd->u0, 0.5, 0.5, n);
  if (!rc && sch == 1)
    {
      rc = shk_dev_axpby (d, bd->u0, bd->u, 0.5 * bd->bb_n, 1.0 - 0.5 * bd->bb_n, n);
      if (!rc)
        rc = bd->ver == 0
          ? shk_dev_bd_force (d, bd->f2, bd->F, NULL, bd->sc_dev, n)
          : shk_dev_bd_force_ft (d, bd->f2, bd->F, bd->T, NULL, bd->sc_dev, bd->nm);
      if (!rc && shear) rc = bd_shift (bd, sh[2]);
      const double *p0 = rc ? NULL : bd_pos (bd, d, x, &rc);
      if (!rc) rc = shk_dev_set_pos (d, p0, 1);
      bd->pcur = p0;
      if (!rc) rc = bd_op (bd, d, prec, bd->f2, bd->u, 0, 0);
      if (!rc) rc = shk_dev_axpby (d, bd->u, bd->u0, 1.0, 1.0, n);
    }
  if (!rc)
    rc = bd->ver == 0
      ? shk_dev_bd_euler (d, bd->xn, x, bd->u, bd->sc_dev, 1.0, n3)
      : shk_dev_bd_euler_ft (d, bd->xn, x, q ? bd->qn : NULL, q, bd->u,
                             bd->sc_dev, 1.0, bd->nm);
  if (!rc && check_ol)
    rc = shk_dev_bd_overlap (d, bd->xn, bd->a_dev, bd->nm, s->periodic, s->lx, s->ly,
                             s->lz, shear ? s->shear_mode : 0,
                             shear ? sh[sch == 1 ? 2 : 1] : 0.0, bd->rmin_ol,
                             &bd->st_dev->overlap[1]);
  if (!rc && noise) rc = shk_dev_counter_add (d, bd->counter, 1);
  if (!rc) rc = shk_dev_d2h_async (d, bd->st_host, bd->st_dev, sizeof (shk_bd_status));
  return rc;
}

static int
bd_capture (stokes_hip_bd *bd, stokes_ctx *c, double *x, double *q)
{
  shk_dev *d = c->dev;
  const int check_ol = bd->sys->rmin == 0.0;
  shk_dev_graph_destroy (bd->graph);
  bd->graph = NULL;
  /* every work buffer the operator uses must exist before the capture:
   * one eager product allocates them */
  shk_dev_timing (d, 0);
  int rc = 0;
  const double *px = bd_pos (bd, d, x, &rc);
  if (!rc) rc = shk_dev_set_pos (d, px, 1);
  bd->pcur = px;
  if (!rc) rc = bd_op (bd, d, c->prec, bd->z, bd->w, 1, 1);
  if (!rc && bd->real_cache)
    {
      rc = shk_dev_real_cache (d, 1, c->prec);
      if (!rc) rc = bd_op (bd, d, c->prec, bd->z, bd->w, 1, 1);
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
  rc = bd_record (bd, d, c->prec, check_ol, x, q, x, 1, NULL);
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
  bd->graph_q = q;
  bd->graph_gen = shk_dev_generation (d);
  bd->graph_uploads = c->uploads;
  bd->graph_prec = c->prec;
  bd->graph_ol = check_ol;
  bd->graph_rc = bd->real_cache;
  bd->graph_flow = bd->flow;
  return 0;
}

/* host-vector form of bd_op (noise parameters) for the Lanczos bounds */
static int
bd_host_op (void *ud, const double *x, double *y)
{
  stokes_hip_bd *bd = ud;
  shk_dev *d = bd->c->dev;
  const size_t bytes = sizeof (double) * bd->n;
  double *xd = bd->work, *yd = bd->work + bd->n;
  /* unequal radii: the noise operator S M (S the Global scaling) has the
   * spectrum of the symmetric S^1/2 M S^1/2, which Lanczos takes */
  int rc = shk_dev_h2d (d, xd, x, bytes);
  if (!rc && bd->sqsv) rc = shk_dev_dmul (d, xd, bd->sqsv, bd->n);
  if (!rc) rc = bd_op (bd, d, bd->c->prec, xd, yd, 1, 0);
  if (!rc && bd->sqsv) rc = shk_dev_dmul (d, yd, bd->sqsv, bd->n);
  if (!rc) rc = shk_dev_d2h (d, y, yd, bytes);
  return rc ? bd_fail (rc, "bd_fit_bounds") : 0;
}

static int
bd_fit_bounds (stokes_hip_bd *bd, stokes_ctx *c, double *x)
{
  double l0, l1;
  int level = bd->fits < 3 ? bd->fits : 3;
  /* the products are at x (an attempt leaves x_mid on the device) */
  int rc = 0;
  c->pos_dev = bd_pos (bd, c->dev, x, &rc);
  bd->pcur = c->pos_dev;
  if (!rc) rc = stokes_hip_prepare (c, 1);
  if (rc) return rc;
  if (bd->real_cache)
    {
      /* every Lanczos product is at x: one dense real-space matrix */
      rc = shk_dev_real_cache (c->dev, 1, c->prec);
      if (rc) return bd_fail (rc, "bd_fit_bounds");
    }
  rc = stokes_hip_lanczos_op ((int) bd->n, BD_LANCZOS_ITER * (level + 1),
                              1 + bd->seed + bd->fits, 0.0, bd_host_op, bd,
                              &l0, &l1);
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
  return stokes_hip_bd_step_q (bd, x_dev, NULL, dt, info);
}

static int bd_step (stokes_hip_bd *bd, double *x_dev, double *q_dev,
                    int have_t, double t, double dt, stokes_hip_bd_info *info);

int
stokes_hip_bd_step_q (stokes_hip_bd *bd, double *x_dev, double *q_dev,
                      double dt, stokes_hip_bd_info *info)
{
  return bd_step (bd, x_dev, q_dev, 0, 0.0, dt, info);
}

int
stokes_hip_bd_step_t (stokes_hip_bd *bd, double *x_dev, double *q_dev,
                      double t, double dt, stokes_hip_bd_info *info)
{
  return bd_step (bd, x_dev, q_dev, 1, t, dt, info);
}

int
stokes_hip_bd_set_shear_ref (stokes_hip_bd *bd, double t0, double s0)
{
  if (!bd)
    {
      stokes_hip_set_error ("bd_set_shear_ref: NULL handle");
      return STOKES_HIP_EARG;
    }
  bd->t0 = t0;
  bd->s0 = s0;
  return 0;
}

static int
bd_step (stokes_hip_bd *bd, double *x_dev, double *q_dev, int have_t,
         double t, double dt, stokes_hip_bd_info *info)
{
  if (!bd || !x_dev || !(dt > 0.0))
    {
      stokes_hip_set_error ("bd_step: need a handle, positions and dt > 0");
      return STOKES_HIP_EARG;
    }
  if (bd->ver == 0) q_dev = NULL;      /* evolve_Euler_3all ignores q for F */
  bd->shear = bd->sys->shear_mode != 0;
  if (bd->shear && !have_t)
    {
      stokes_hip_set_error ("bd_step: a sheared system needs the time of the"
                            " step (stokes_hip_bd_step_t)");
      return STOKES_HIP_EARG;
    }
  if (bd->shear)
    bd->sys->shear_shift = stokes_get_shear_shift (bd->sys, t, bd->t0, bd->s0);
  /* the imposed flow of the system now (solve_mix_3all's labo frame) */
  for (int k = 0; k < 3; k++)
    {
      bd->flow.U[k] = bd->sys->Ui[k];
      bd->flow.O[k] = bd->sys->Oi[k];
    }
  for (int k = 0; k < 5; k++) bd->flow.E[k] = bd->sys->Ei[k];
  bd->has_flow = !bd_fluid_at_rest (bd->sys);
  stokes_ctx *c = stokes_hip_ctx (bd->sys);
  if (!c) return STOKES_HIP_ENODEV;
  shk_dev *d = c->dev;
  const double *pos_dev_saved = c->pos_dev;
  stokes_hip_bd_info inf;
  memset (&inf, 0, sizeof (inf));
  int rc = 0;
  if (bd->nm < bd->np)
    {
      /* the fixed particles where the system holds them now */
      rc = shk_dev_h2d (d, bd->posfull + 3 * (size_t) bd->nm,
                        bd->sys->pos + 3 * bd->nm,
                        sizeof (double) * 3 * (size_t) (bd->np - bd->nm));
      if (rc) { rc = bd_fail (rc, "bd_step"); goto out; }
    }
  c->pos_dev = bd_pos (bd, d, x_dev, &rc);
  if (rc) { rc = bd_fail (rc, "bd_step"); goto out; }
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
  /* an attempt that steers GMRES from the host runs eagerly (FTS, and the
   * F / FT drift with fixed particles); the others replay a graph */
  const int eager = bd->ver == 2 || bd->nm < bd->np || bd->shear || bd->lub;
  /* the cell shift at t (stokes_set_shear_shift at REDO, brownian.c:2393) */
  const double sh_t = bd->shear
    ? stokes_get_shear_shift (bd->sys, t, bd->t0, bd->s0) : 0.0;
  bd->sh_cur = sh_t;
  int fresh = 1;           /* new Gaussian draw for this attempt */
  int refit_done = 0;      /* BD_calc_FB refits at most once per draw */
  /* stale: the system holds the intermediate configuration of a failed
   * attempt (bd->xs), as the reference's REDO_scale leaves it after an
   * overlap in the second stage of the mid-point and BM97 schemes
   * (brownian.c:2523, 3162; BB03 has moved back to x for its deterministic
   * product, 2808, so its retries start from x); such an attempt runs
   * eagerly.  keep: the attempt keeps the noise of the
   * previous one (REDO_scale); the graph's replays at x recompute it from
   * the same draw, which gives the same values */
  int stale = 0, keep = 0;
  for (;;)
    {
      const int run_eager = eager || stale;
      if (inf.attempts >= BD_MAX_ATTEMPTS)
        {
          stokes_hip_set_error ("bd_step: no acceptable step after %d attempts",
                                inf.attempts);
          rc = STOKES_HIP_EARG;
          goto out;
        }
      if (!run_eager
          && (!bd->graph || bd->graph_c != c || bd->graph_x != x_dev
              || bd->graph_q != q_dev
              || bd->graph_gen != shk_dev_generation (d)
              || bd->graph_uploads != c->uploads || bd->graph_prec != c->prec
              || bd->graph_ol != (bd->sys->rmin == 0.0)
              || bd->graph_rc != bd->real_cache
              || memcmp (&bd->graph_flow, &bd->flow, sizeof (shk_flow)) != 0))
        {
          rc = bd_capture (bd, c, x_dev, q_dev);
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
      if (keep)
        ;
      else if (!fresh)
        {
          /* replay the previous draw (the host stream keeps it in y_host) */
          rc = shk_dev_counter_add (d, bd->counter, (unsigned long long) -1);
          if (rc) { rc = bd_fail (rc, "bd_step"); goto out; }
        }
      else if (bd->mt)
        /* BD_calc_FB's n draws in order (brownian.c:1604-1609) */
        for (size_t k = 0; k < bd->n; k++)
          bd->y_host[k] = stokes_hip_mt_gaussian (bd->mt);
      if (run_eager)
        {
          /* the shifts of this attempt: where the system is, at t + h
           * (h = dt / 2, dt / n, dt), and at t */
          const double h = bd->scheme == 0 ? 0.5 * dt_local
            : (bd->scheme == 1 ? dt_local / bd->bb_n : dt_local);
          const double sh[3] = {
            bd->sh_cur,
            bd->shear ? stokes_get_shear_shift (bd->sys, t + h, bd->t0, bd->s0) : 0.0,
            sh_t};
          rc = bd_record (bd, d, c->prec, bd->sys->rmin == 0.0, x_dev, q_dev,
                          stale ? bd->xs : x_dev, !keep, sh);
          if (!rc && bd->shear)
            bd->sh_after = sh[bd->scheme == 1 ? 2 : 1];
          if (bd->real_cache && rc)
            shk_dev_real_cache (d, 0, c->prec);
        }
      else
        rc = shk_dev_graph_launch (d, bd->graph);
      if (!rc) rc = shk_dev_sync (d);
      if (rc) { rc = bd_fail (rc, "bd_step (graph)"); goto out; }
      inf.attempts++;
      const shk_bd_status st = *bd->st_host;
      inf.overlaps[0] = st.overlap[0];
      inf.overlaps[1] = st.overlap[1];
      if (!keep && !(st.zmz > 0.0))
        {
          stokes_hip_set_error ("bd_step: z.Mz = %g is not positive", st.zmz);
          rc = STOKES_HIP_EARG;
          goto out;
        }
      if (!keep) inf.err_cheb = fabs (st.zmz - st.yy) / st.zmz;
      if (!keep && inf.err_cheb > bd->eps && !refit_done)
        {
          /* brownian.c:1707-1727: new bounds where the noise is taken,
           * same draw */
          if (bd->shear) bd->sys->shear_shift = bd->sh_cur;
          rc = bd_fit_bounds (bd, c, stale ? bd->xs : x_dev);
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
          /* REDO_FB (stage 1: the positions stay) or REDO (stage 2: x) */
          dt_local = dt;
          if (bd->fact_fix) fact = bd_fact (bd->peclet, dt_local);
          inf.redraws++;
          fresh = 1;
          refit_done = 0;
          keep = 0;
          if (stage == 2)
            {
              stale = 0;
              bd->sh_cur = sh_t;
            }
        }
      else
        {
          /* REDO_scale: the noise kept; after the second stage of the
           * mid-point and BM97 schemes the system holds this attempt's
           * intermediate configuration (BB03's is back at x) */
          fact = bd_fact (bd->peclet, dt_local);
          fresh = 0;
          if (stage == 2 && bd->scheme != 1)
            {
              rc = shk_dev_d2d (d, bd->xs, bd->xm, sizeof (double) * 3 * bd->nm);
              if (rc) { rc = bd_fail (rc, "bd_step"); goto out; }
              stale = 1;
              /* and the shift the second stage left */
              bd->sh_cur = bd->sh_after;
            }
          keep = stale;
        }
    }
  /* accepted: x <- x_new */
  rc = shk_dev_d2d (d, x_dev, bd->xn, sizeof (double) * 3 * bd->nm);
  if (!rc && q_dev)
    rc = shk_dev_d2d (d, q_dev, bd->qn, sizeof (double) * 4 * bd->nm);
  if (rc) { rc = bd_fail (rc, "bd_step"); goto out; }
  inf.dt = dt_local;
out:
  inf.eig[0] = bd->eig[0];
  inf.eig[1] = bd->eig[1];
  c->pos_dev = pos_dev_saved;
  if (info) *info = inf;
  return rc;
}
