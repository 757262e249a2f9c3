/* ensemble.c -- the mobility product (and the FP64 lubrication correction)
 * of many independent periodic systems -- boxes of uneven size, radii and
 * box length -- as one graph launch of batched kernels.
 *
 * Every phase of the operator is one kernel over the work items of all
 * boxes (ens_dev.hip); the graph holds those kernels, the real-space phase
 * beside the reciprocal phases on a forked stream.  Box b reads its
 * positions, x and y at offsets off[b] (particles) of packed device arrays:
 * positions 3 off[b], vectors dim off[b].  Box parameters (box size, Ewald
 * splitting, radii, lubrication parameters) are taken when the boxes are
 * uploaded -- stokes_hip_ensemble_refresh() re-reads them; positions are
 * re-read on every launch.
 *
 * The autotuner chooses each box's Ewald splitting and the tile shape of
 * the batched kernels: it times the batch at each candidate splitting (all
 * boxes alike), fits the time of the real-space phase to the candidate
 * tests and pair images of the boxes and the time of the reciprocal phases
 * to their (k, particle) and phase-table work, gives each box the
 * splitting the fitted costs make cheapest, keeps that assignment if the
 * whole launch is faster than the best common splitting, and then sweeps
 * the tile shape (64-particle chunks per real-space wave, particles per
 * structure-factor split, k per back-projection split) one parameter at a
 * time on the whole launch.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "stokes_hip.h"
#include "stokes_internal.h"

struct stokes_hip_ensemble {
  int nbox;
  struct stokes **sys;
  size_t *off;                  /* particle offsets, nbox + 1 */
  double *tr;                   /* per box: splitting computed at (0: own xi) */
  int *fix_g;                   /* per box: xi-independent G sign */
  int prec;
  int jw, jper, kper;           /* tile shape */
  shk_ens *dev;
  int uploaded, lub_set;
  void *graph;
  const double *g_pos, *g_x;
  double *g_y, *g_ylub;
  unsigned long g_gen;
  int g_cap;
};

void
stokes_hip_ensemble_destroy (stokes_hip_ensemble *e)
{
  if (!e) return;
  shk_dev_graph_destroy (e->graph);
  shk_ens_destroy (e->dev);
  free (e->sys);
  free (e->off);
  free (e->tr);
  free (e->fix_g);
  free (e);
}

int
stokes_hip_ensemble_create (struct stokes *const *sys, int nbox,
                            stokes_hip_ensemble **out)
{
  if (!out || !sys || nbox <= 0)
    {
      stokes_hip_set_error ("ensemble: need nbox > 0 systems");
      return STOKES_HIP_EARG;
    }
  *out = NULL;
  stokes_hip_ensemble *e = calloc (1, sizeof (*e));
  if (!e) return STOKES_HIP_ENOMEM;
  e->nbox = nbox;
  e->sys = malloc (sizeof (struct stokes *) * nbox);
  e->off = malloc (sizeof (size_t) * (nbox + 1));
  e->tr = calloc (nbox, sizeof (double));
  e->fix_g = calloc (nbox, sizeof (int));
  if (!e->sys || !e->off || !e->tr || !e->fix_g)
    {
      stokes_hip_ensemble_destroy (e);
      return STOKES_HIP_ENOMEM;
    }
  e->prec = -1;
  e->jw = 2;
  e->jper = 256;
  e->kper = 64;
  e->off[0] = 0;
  for (int b = 0; b < nbox; b++)
    {
      struct stokes *s = sys[b];
      if (!s || s->np <= 0 || s->pos == NULL)
        {
          stokes_hip_ensemble_destroy (e);
          stokes_hip_set_error ("ensemble: box %d has no particles", b);
          return STOKES_HIP_EARG;
        }
      if (!s->periodic || s->version != sys[0]->version
          || (s->a == NULL) != (sys[0]->a == NULL) || s->slip != NULL)
        {
          stokes_hip_ensemble_destroy (e);
          stokes_hip_set_error ("ensemble: box %d must be periodic, without slip, with the"
                                " version and kind of radii of box 0", b);
          return STOKES_HIP_EUNSUP;
        }
      e->sys[b] = s;
      e->off[b + 1] = e->off[b] + (size_t) s->np;
      /* a box's own context settings carry over: splitting, G sign,
       * precision (that of box 0 for all) */
      stokes_ctx *c = stokes_hip_ctx_find (s);
      if (c)
        {
          e->tr[b] = c->split_tr;
          e->fix_g[b] = c->fix_recip_g;
          if (b == 0) e->prec = c->prec;
        }
    }
  if (e->prec < 0) e->prec = SHK_FP64;
  e->dev = shk_ens_create ();
  if (!e->dev)
    {
      stokes_hip_set_error ("ensemble: %s", shk_last_error ());
      stokes_hip_ensemble_destroy (e);
      return STOKES_HIP_ENODEV;
    }
  *out = e;
  return 0;
}

static void
drop_graph (stokes_hip_ensemble *e)
{
  shk_dev_graph_destroy (e->graph);
  e->graph = NULL;
}

/* the boxes' tables at the current splittings and tile shape */
static int
upload (stokes_hip_ensemble *e)
{
  shk_ens_box *bx = calloc (e->nbox, sizeof (shk_ens_box));
  if (!bx) return STOKES_HIP_ENOMEM;
  int rc = 0, b;
  for (b = 0; b < e->nbox && !rc; b++)
    {
      struct stokes *s = e->sys[b];
      rc = stokes_hip_tables (s, e->tr[b], e->fix_g[b], &bx[b].p, &bx[b].t);
      if (rc) break;
      bx[b].a = s->a;
      bx[b].jw = e->jw;
      bx[b].sj = (s->np + e->jper - 1) / e->jper;
      bx[b].sk = bx[b].p.nk > 0 ? (bx[b].p.nk + e->kper - 1) / e->kper : 1;
    }
  if (!rc)
    {
      rc = shk_ens_set (e->dev, e->nbox, bx, e->prec);
      if (rc) stokes_hip_set_error ("ensemble: %s", shk_last_error ());
    }
  for (int k = 0; k < b && k < e->nbox; k++) stokes_hip_tables_free (&bx[k].t);
  free (bx);
  drop_graph (e);
  e->uploaded = rc == 0;
  e->lub_set = 0;
  return rc;
}

static int
set_lub (stokes_hip_ensemble *e)
{
  shk_lub *lp = calloc (e->nbox, sizeof (shk_lub));
  const double **tab = calloc ((size_t) 11 * e->nbox, sizeof (double *));
  double *f1 = calloc ((size_t) 11 * 51, sizeof (double));
  int rc = (!lp || !tab || !f1) ? STOKES_HIP_ENOMEM : 0;
  for (int b = 0; b < e->nbox && !rc; b++)
    {
      /* the f_k(1) table is the same for every box: fill it once */
      rc = stokes_hip_lub_params (e->sys[b], &lp[b], tab + 11 * b, f1);
    }
  if (!rc)
    {
      rc = shk_ens_set_lub (e->dev, lp);
      if (rc) stokes_hip_set_error ("ensemble lubrication: %s", shk_last_error ());
    }
  free (lp);
  free (tab);
  free (f1);
  drop_graph (e);
  e->lub_set = rc == 0;
  return rc;
}

int
stokes_hip_ensemble_refresh (stokes_hip_ensemble *e)
{
  if (!e)
    {
      stokes_hip_set_error ("ensemble: NULL handle");
      return STOKES_HIP_EARG;
    }
  drop_graph (e);
  e->uploaded = 0;
  e->lub_set = 0;
  return 0;
}

int
stokes_hip_ensemble_set_precision (stokes_hip_ensemble *e, int mode)
{
  if (!e || (mode != STOKES_HIP_FP64 && mode != STOKES_HIP_MIXED))
    {
      stokes_hip_set_error ("ensemble: invalid precision");
      return STOKES_HIP_EARG;
    }
  e->prec = mode == STOKES_HIP_MIXED ? SHK_MIXED : SHK_FP64;
  return stokes_hip_ensemble_refresh (e);
}

int
stokes_hip_ensemble_set_split (stokes_hip_ensemble *e, int box, double tratio)
{
  if (!e || box < -1 || box >= e->nbox || tratio < 0.0)
    {
      stokes_hip_set_error ("ensemble: invalid box or splitting");
      return STOKES_HIP_EARG;
    }
  for (int b = 0; b < e->nbox; b++)
    if (box < 0 || b == box) e->tr[b] = tratio;
  return stokes_hip_ensemble_refresh (e);
}

int
stokes_hip_ensemble_set_tiles (stokes_hip_ensemble *e, int jw, int jper, int kper)
{
  if (!e || jw < 1 || jper < 16 || kper < 16)
    {
      stokes_hip_set_error ("ensemble: tiles need jw >= 1, jper >= 16, kper >= 16");
      return STOKES_HIP_EARG;
    }
  e->jw = jw;
  e->jper = jper;
  e->kper = kper;
  return stokes_hip_ensemble_refresh (e);
}

int
stokes_hip_ensemble_get_tiles (const stokes_hip_ensemble *e, int *jw, int *jper,
                               int *kper)
{
  if (!e) return STOKES_HIP_EARG;
  if (jw) *jw = e->jw;
  if (jper) *jper = e->jper;
  if (kper) *kper = e->kper;
  return 0;
}

static int
ens_run (stokes_hip_ensemble *e, const double *pos, const double *x, double *y,
         double *ylub)
{
  if (!e || !pos || !x || (!y && !ylub))
    {
      stokes_hip_set_error ("ensemble: NULL argument");
      return STOKES_HIP_EARG;
    }
  int rc = 0;
  if (!e->uploaded && (rc = upload (e))) return rc;
  if (ylub && !e->lub_set && (rc = set_lub (e))) return rc;
  for (int attempt = 0; attempt < 2; attempt++)
    {
      if (ylub && attempt == 0 && !e->graph)
        {
          /* size the pair list at these positions before the capture */
          if ((rc = shk_ens_lub_prepare (e->dev, pos, NULL)))
            {
              stokes_hip_set_error ("ensemble lubrication: %s", shk_last_error ());
              return rc;
            }
        }
      int stale = e->graph == NULL || e->g_pos != pos || e->g_x != x || e->g_y != y
        || e->g_ylub != ylub || e->g_gen != shk_ens_generation (e->dev)
        || e->g_cap != shk_ens_lub_capacity (e->dev);
      if (stale)
        {
          drop_graph (e);
          void *s = shk_ens_stream (e->dev);
          rc = shk_stream_capture_begin (s);
          if (!rc)
            {
              rc = shk_ens_launch (e->dev, pos, x, y, ylub, 0);
              if (rc) shk_stream_capture_abort (s);
              else rc = shk_stream_capture_end (s, &e->graph);
            }
          if (rc)
            {
              stokes_hip_set_error ("ensemble capture: %s", shk_last_error ());
              return STOKES_HIP_EHIP;
            }
          e->g_pos = pos; e->g_x = x; e->g_y = y; e->g_ylub = ylub;
          e->g_gen = shk_ens_generation (e->dev);
          e->g_cap = shk_ens_lub_capacity (e->dev);
        }
      void *s = shk_ens_stream (e->dev);
      rc = shk_graph_launch_on (e->graph, s);
      if (!rc) rc = shk_stream_sync (s);
      if (rc)
        {
          stokes_hip_set_error ("ensemble launch: %s", shk_last_error ());
          return STOKES_HIP_EHIP;
        }
      if (!ylub || shk_ens_lub_last_total (e->dev) <= shk_ens_lub_capacity (e->dev))
        return 0;
      /* more near pairs than the list holds: grow it and run again */
      if ((rc = shk_ens_lub_prepare (e->dev, pos, NULL)))
        {
          stokes_hip_set_error ("ensemble lubrication: %s", shk_last_error ());
          return rc;
        }
    }
  stokes_hip_set_error ("ensemble lubrication: pair list could not be sized");
  return STOKES_HIP_EHIP;
}

int
stokes_hip_ensemble_atimes (stokes_hip_ensemble *e, const double *pos_dev,
                            const double *x_dev, double *y_dev)
{
  if (!y_dev)
    {
      stokes_hip_set_error ("ensemble: NULL argument");
      return STOKES_HIP_EARG;
    }
  return ens_run (e, pos_dev, x_dev, y_dev, NULL);
}

int
stokes_hip_ensemble_lub (stokes_hip_ensemble *e, const double *pos_dev,
                         const double *x_dev, double *ylub_dev)
{
  if (!ylub_dev)
    {
      stokes_hip_set_error ("ensemble: NULL argument");
      return STOKES_HIP_EARG;
    }
  return ens_run (e, pos_dev, x_dev, NULL, ylub_dev);
}

int
stokes_hip_ensemble_atimes_lub (stokes_hip_ensemble *e, const double *pos_dev,
                                const double *x_dev, double *y_dev, double *ylub_dev)
{
  if (!y_dev || !ylub_dev)
    {
      stokes_hip_set_error ("ensemble: NULL argument");
      return STOKES_HIP_EARG;
    }
  return ens_run (e, pos_dev, x_dev, y_dev, ylub_dev);
}

int
stokes_hip_ensemble_phase_ms (stokes_hip_ensemble *e, const double *pos_dev,
                              const double *x_dev, double *y_dev, double *ylub_dev,
                              double *ms)
{
  if (!e || !pos_dev || !x_dev || (!y_dev && !ylub_dev) || !ms)
    {
      stokes_hip_set_error ("ensemble: NULL argument");
      return STOKES_HIP_EARG;
    }
  int rc = 0;
  if (!e->uploaded && (rc = upload (e))) return rc;
  if (ylub_dev)
    {
      if (!e->lub_set && (rc = set_lub (e))) return rc;
      if ((rc = shk_ens_lub_prepare (e->dev, pos_dev, NULL)))
        {
          stokes_hip_set_error ("ensemble lubrication: %s", shk_last_error ());
          return rc;
        }
    }
  rc = shk_ens_launch (e->dev, pos_dev, x_dev, y_dev, ylub_dev, 1);
  if (!rc) rc = shk_ens_phase_ms (e->dev, ms);
  if (rc)
    {
      stokes_hip_set_error ("ensemble: %s", shk_last_error ());
      return STOKES_HIP_EHIP;
    }
  return 0;
}

size_t
stokes_hip_ensemble_offset (const stokes_hip_ensemble *e, int box)
{
  if (!e || box < 0 || box > e->nbox) return 0;
  return e->off[box];
}

/* ------------------------------------------------------------------ */
/* autotuner                                                            */
/* ------------------------------------------------------------------ */

static double
now_s (void)
{
  struct timespec ts;
  clock_gettime (CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* best of reps graph launches (after one untimed launch), seconds */
static int
time_launch (stokes_hip_ensemble *e, const double *pos, const double *x, double *y,
             int reps, double *t_out)
{
  int rc = ens_run (e, pos, x, y, NULL);
  if (rc) return rc;
  double best = 1e30;
  for (int r = 0; r < reps; r++)
    {
      double t0 = now_s ();
      if ((rc = ens_run (e, pos, x, y, NULL))) return rc;
      double t = now_s () - t0;
      if (t < best) best = t;
    }
  *t_out = best;
  return 0;
}

/* least squares of t[u] ~ c0 f0[u] + c1 f1[u] over u < n with c >= 0 */
static void
fit2 (int n, const double *f0, const double *f1, const double *t, double *c)
{
  double a00 = 0, a01 = 0, a11 = 0, b0 = 0, b1 = 0;
  for (int u = 0; u < n; u++)
    {
      a00 += f0[u] * f0[u]; a01 += f0[u] * f1[u]; a11 += f1[u] * f1[u];
      b0 += f0[u] * t[u]; b1 += f1[u] * t[u];
    }
  double det = a00 * a11 - a01 * a01;
  c[0] = c[1] = 0.0;
  if (det > 1e-12 * a00 * a11 && a00 > 0 && a11 > 0)
    {
      c[0] = (b0 * a11 - b1 * a01) / det;
      c[1] = (a00 * b1 - a01 * b0) / det;
    }
  if (!(c[0] >= 0.0 && c[1] >= 0.0))
    {
      /* one feature alone: the one that fits better */
      double c0 = a00 > 0 ? b0 / a00 : 0.0, c1 = a11 > 0 ? b1 / a11 : 0.0;
      double r0 = 0, r1 = 0;
      for (int u = 0; u < n; u++)
        {
          r0 += (t[u] - c0 * f0[u]) * (t[u] - c0 * f0[u]);
          r1 += (t[u] - c1 * f1[u]) * (t[u] - c1 * f1[u]);
        }
      c[0] = r0 <= r1 ? (c0 > 0 ? c0 : 0.0) : 0.0;
      c[1] = r0 <= r1 ? 0.0 : (c1 > 0 ? c1 : 0.0);
    }
}

int
stokes_hip_ensemble_autotune (stokes_hip_ensemble *e, const double *trs,
                              int ntr, int reps, double *best_tr)
{
  static const double def_trs[] = {10.0, 30.0, 100.0, 300.0, 1000.0, 3000.0};
  if (!e)
    {
      stokes_hip_set_error ("ensemble: NULL handle");
      return STOKES_HIP_EARG;
    }
  if (!trs || ntr <= 0)
    {
      trs = def_trs;
      ntr = (int) (sizeof (def_trs) / sizeof (def_trs[0]));
    }
  if (reps < 1) reps = 1;
  const int nb = e->nbox, D = stokes_dim (e->sys[0]->version);
  const size_t ntot = e->off[nb];
  /* packed positions from the systems, a random x */
  double *hpos = malloc (sizeof (double) * 3 * ntot);
  double *hx = malloc (sizeof (double) * D * ntot);
  double *feat = calloc ((size_t) nb * ntr * 4, sizeof (double));
  double *R = calloc (ntr, sizeof (double)), *K = calloc (ntr, sizeof (double));
  double *T = calloc (ntr, sizeof (double));
  double *F = calloc ((size_t) ntr * 4, sizeof (double));
  double *keep_tr = malloc (sizeof (double) * nb);
  void *dpos = NULL, *dx = NULL, *dy = NULL;
  int rc = (!hpos || !hx || !feat || !R || !K || !T || !F || !keep_tr) ? STOKES_HIP_ENOMEM : 0;
  if (!rc)
    {
      unsigned long st = 12345;
      for (int b = 0; b < nb; b++)
        memcpy (hpos + 3 * e->off[b], e->sys[b]->pos, sizeof (double) * 3 * e->sys[b]->np);
      for (size_t k = 0; k < D * ntot; k++)
        {
          st = st * 6364136223846793005UL + 1442695040888963407UL;
          hx[k] = (double) (st >> 11) / 9007199254740992.0 - 0.5;
        }
      rc = stokes_hip_malloc (&dpos, sizeof (double) * 3 * ntot);
      if (!rc) rc = stokes_hip_malloc (&dx, sizeof (double) * D * ntot);
      if (!rc) rc = stokes_hip_malloc (&dy, sizeof (double) * D * ntot);
      if (!rc) rc = stokes_hip_memcpy_h2d (dpos, hpos, sizeof (double) * 3 * ntot);
      if (!rc) rc = stokes_hip_memcpy_h2d (dx, hx, sizeof (double) * D * ntot);
    }
  memcpy (keep_tr, e->tr, sizeof (double) * nb);
  /* 1. every candidate splitting for all boxes: phase times and work */
  int ubest = 0;
  for (int u = 0; u < ntr && !rc; u++)
    {
      for (int b = 0; b < nb; b++) e->tr[b] = trs[u];
      if ((rc = upload (e))) break;
      double ms[4], best[3] = {1e30, 1e30, 1e30};
      for (int r = 0; r <= reps && !rc; r++)
        {
          rc = stokes_hip_ensemble_phase_ms (e, dpos, dx, dy, NULL, ms);
          if (r == 0) continue;                      /* untimed first launch */
          for (int k = 0; k < 3; k++) if (ms[k] < best[k]) best[k] = ms[k];
        }
      if (rc) break;
      R[u] = best[0];
      K[u] = best[1];
      if ((rc = time_launch (e, dpos, dx, dy, reps, &T[u]))) break;
      if (T[u] < T[ubest]) ubest = u;
      for (int b = 0; b < nb; b++)
        {
          int info[8];
          shk_ens_box_info (e->dev, b, info);
          const double np = info[0], nj = (info[0] + 63) / 64;
          const struct stokes *s = e->sys[b];
          {
            struct stokes sh;
            stokes_hip_shadow_split (s, trs[u], &sh);
            double rmax = sqrt (sh.rmax2);
            double vol = s->lx * s->ly * s->lz;
            double *f = feat + ((size_t) b * ntr + u) * 4;
            f[0] = np * 64.0 * nj * info[1];                       /* candidate tests */
            f[1] = np * np * 4.18879020478639 * rmax * rmax * rmax / vol; /* pair images */
            f[2] = np * info[3];                                   /* (k, particle) */
            f[3] = np * info[4];                                   /* phase-table entries */
            stokes_hip_shadow_free (&sh);
          }
        }
    }
  if (!rc)
    {
      /* 2. fitted costs of the two real-space and the two reciprocal
       * features; each box takes its cheapest candidate */
      for (int u = 0; u < ntr; u++)
        for (int q = 0; q < 4; q++)
          {
            double sum = 0.0;
            for (int b = 0; b < nb; b++) sum += feat[((size_t) b * ntr + u) * 4 + q];
            F[(size_t) q * ntr + u] = sum;
          }
      double cr[2], ck[2];
      fit2 (ntr, F, F + ntr, R, cr);
      fit2 (ntr, F + 2 * ntr, F + 3 * ntr, K, ck);
      for (int b = 0; b < nb; b++)
        {
          int ub = 0;
          double cb = 1e300;
          for (int u = 0; u < ntr; u++)
            {
              const double *f = feat + ((size_t) b * ntr + u) * 4;
              double c = cr[0] * f[0] + cr[1] * f[1] + ck[0] * f[2] + ck[1] * f[3];
              if (c < cb) { cb = c; ub = u; }
            }
          e->tr[b] = trs[ub];
        }
      double tmix = 1e30;
      rc = upload (e);
      if (!rc) rc = time_launch (e, dpos, dx, dy, reps, &tmix);
      if (!rc && tmix > T[ubest])
        {
          for (int b = 0; b < nb; b++) e->tr[b] = trs[ubest];
          rc = upload (e);
        }
    }
  /* 3. tile shape, one parameter at a time on the whole launch */
  if (!rc)
    {
      static const int jws[] = {1, 2, 4}, jpers[] = {128, 256, 512, 1024},
        kpers[] = {32, 64, 128, 256};
      double tbest;
      rc = time_launch (e, dpos, dx, dy, reps, &tbest);
      for (int pass = 0; pass < 3 && !rc; pass++)
        {
          const int *vals = pass == 0 ? jws : (pass == 1 ? jpers : kpers);
          const int nv = pass == 0 ? 3 : 4;
          int *field = pass == 0 ? &e->jw : (pass == 1 ? &e->jper : &e->kper);
          int keep = *field;
          for (int v = 0; v < nv && !rc; v++)
            {
              if (vals[v] == keep) continue;
              int prev = *field;
              *field = vals[v];
              double t = 1e30;
              if (!(rc = upload (e))) rc = time_launch (e, dpos, dx, dy, reps, &t);
              if (!rc && t < tbest) { tbest = t; keep = vals[v]; }
              else *field = prev;
            }
          *field = keep;
          if (!rc) rc = upload (e);
        }
    }
  if (rc) memcpy (e->tr, keep_tr, sizeof (double) * nb);
  else if (best_tr) memcpy (best_tr, e->tr, sizeof (double) * nb);
  if (dpos) stokes_hip_free (dpos);
  if (dx) stokes_hip_free (dx);
  if (dy) stokes_hip_free (dy);
  free (hpos); free (hx); free (feat); free (R); free (K); free (T); free (F);
  free (keep_tr);
  drop_graph (e);
  return rc;
}
