/* ensemble.c -- the mobility product of many independent systems (boxes of
 * uneven size and radii) as one graph launch.
 *
 * Each box keeps its own device context (tables, work buffers); the graph
 * runs every box's operator on a pool of streams forked from and joined
 * back to one capture stream, so small boxes overlap on the GPU.  Box b
 * reads its positions, x and y at offsets off[b] (particles) of packed
 * device arrays: positions 3 off[b], vectors dim off[b].  Box parameters
 * (box size, Ewald splitting, radii, precision) are taken when the graph is
 * captured -- stokes_hip_ensemble_refresh() recaptures after changes;
 * positions are re-read on every launch.
 */
#include <stdlib.h>
#include <string.h>
#include "stokes_hip.h"
#include "stokes_internal.h"

#define ENS_STREAMS 8

struct stokes_hip_ensemble {
  int nbox;
  struct stokes **sys;
  size_t *off;                  /* particle offsets, nbox + 1 */
  void *main;                   /* capture / launch stream */
  void *pool[ENS_STREAMS];
  void *fork, *join[ENS_STREAMS];
  void *graph;
  const double *g_pos, *g_x;
  double *g_y;
  unsigned long *gen;           /* per box: buffer generation at capture */
};

void
stokes_hip_ensemble_destroy (stokes_hip_ensemble *e)
{
  if (!e) return;
  if (e->main) shk_stream_sync (e->main);
  shk_dev_graph_destroy (e->graph);
  for (int k = 0; k < ENS_STREAMS; k++)
    {
      shk_stream_destroy (e->pool[k]);
      shk_event_destroy (e->join[k]);
    }
  shk_event_destroy (e->fork);
  shk_stream_destroy (e->main);
  free (e->sys);
  free (e->off);
  free (e->gen);
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
  e->gen = calloc (nbox, sizeof (unsigned long));
  if (!e->sys || !e->off || !e->gen)
    {
      stokes_hip_ensemble_destroy (e);
      return STOKES_HIP_ENOMEM;
    }
  e->off[0] = 0;
  for (int b = 0; b < nbox; b++)
    {
      if (!sys[b] || sys[b]->np <= 0 || sys[b]->pos == NULL)
        {
          stokes_hip_ensemble_destroy (e);
          stokes_hip_set_error ("ensemble: box %d has no particles", b);
          return STOKES_HIP_EARG;
        }
      e->sys[b] = sys[b];
      e->off[b + 1] = e->off[b] + (size_t) sys[b]->np;
    }
  int rc = shk_stream_create (&e->main);
  if (!rc) rc = shk_event_create (&e->fork);
  for (int k = 0; k < ENS_STREAMS && !rc; k++)
    {
      rc = shk_stream_create (&e->pool[k]);
      if (!rc) rc = shk_event_create (&e->join[k]);
    }
  if (rc)
    {
      stokes_hip_set_error ("ensemble: %s", shk_last_error ());
      stokes_hip_ensemble_destroy (e);
      return STOKES_HIP_EHIP;
    }
  *out = e;
  return 0;
}

int
stokes_hip_ensemble_refresh (stokes_hip_ensemble *e)
{
  if (!e)
    {
      stokes_hip_set_error ("ensemble: NULL handle");
      return STOKES_HIP_EARG;
    }
  shk_dev_graph_destroy (e->graph);
  e->graph = NULL;
  return 0;
}

int
stokes_hip_ensemble_autotune (stokes_hip_ensemble *e, const double *trs,
                              int ntr, int reps, double *best_tr)
{
  if (!e)
    {
      stokes_hip_set_error ("ensemble: NULL handle");
      return STOKES_HIP_EARG;
    }
  for (int b = 0; b < e->nbox; b++)
    {
      double bt = 0.0;
      int rc = stokes_hip_autotune (e->sys[b], 1, trs, ntr, reps, &bt, NULL);
      if (rc) return rc;
      if (best_tr) best_tr[b] = bt;
    }
  return stokes_hip_ensemble_refresh (e);
}

static int
ens_capture (stokes_hip_ensemble *e, const double *pos, const double *x,
             double *y)
{
  int rc = 0;
  shk_dev_graph_destroy (e->graph);
  e->graph = NULL;
  /* parameters up to date and every work buffer allocated (one eager
   * product per box) before the capture */
  for (int b = 0; b < e->nbox && !rc; b++)
    {
      struct stokes *s = e->sys[b];
      stokes_ctx *c = stokes_hip_ctx (s);
      if (!c) return STOKES_HIP_ENODEV;
      int dim = stokes_dim (s->version);
      const double *keep = c->pos_dev;
      c->pos_dev = pos + 3 * e->off[b];
      rc = stokes_hip_prepare (c, 1);
      c->pos_dev = keep;
      if (rc) return rc;
      rc = shk_dev_atimes (c->dev, x + dim * e->off[b], y + dim * e->off[b], 1,
                           c->prec);
      if (!rc) rc = shk_dev_sync (c->dev);
      if (rc)
        {
          stokes_hip_set_error ("ensemble box %d: %s", b, shk_last_error ());
          return rc;
        }
    }
  rc = shk_stream_capture_begin (e->main);
  if (!rc) rc = shk_event_record (e->fork, e->main);
  for (int k = 0; k < ENS_STREAMS && !rc; k++)
    rc = shk_stream_wait (e->pool[k], e->fork);
  for (int b = 0; b < e->nbox && !rc; b++)
    {
      struct stokes *s = e->sys[b];
      stokes_ctx *c = stokes_hip_ctx (s);
      int dim = stokes_dim (s->version);
      void *old = shk_dev_swap_stream (c->dev, e->pool[b % ENS_STREAMS]);
      shk_dev_timing (c->dev, 0);     /* a capture cannot wait on events */
      rc = shk_dev_set_pos (c->dev, pos + 3 * e->off[b], 1);
      if (!rc)
        rc = shk_dev_atimes (c->dev, x + dim * e->off[b], y + dim * e->off[b], 1,
                             c->prec);
      shk_dev_timing (c->dev, c->timing);
      shk_dev_swap_stream (c->dev, old);
      e->gen[b] = shk_dev_generation (c->dev);
    }
  for (int k = 0; k < ENS_STREAMS && !rc; k++)
    {
      rc = shk_event_record (e->join[k], e->pool[k]);
      if (!rc) rc = shk_stream_wait (e->main, e->join[k]);
    }
  if (rc)
    {
      shk_stream_capture_abort (e->main);
      stokes_hip_set_error ("ensemble capture: %s", shk_last_error ());
      return STOKES_HIP_EHIP;
    }
  rc = shk_stream_capture_end (e->main, &e->graph);
  if (rc)
    {
      stokes_hip_set_error ("ensemble capture: %s", shk_last_error ());
      return STOKES_HIP_EHIP;
    }
  e->g_pos = pos;
  e->g_x = x;
  e->g_y = y;
  return 0;
}

int
stokes_hip_ensemble_atimes (stokes_hip_ensemble *e, const double *pos_dev,
                            const double *x_dev, double *y_dev)
{
  if (!e || !pos_dev || !x_dev || !y_dev)
    {
      stokes_hip_set_error ("ensemble: NULL argument");
      return STOKES_HIP_EARG;
    }
  int stale = e->graph == NULL || e->g_pos != pos_dev || e->g_x != x_dev
    || e->g_y != y_dev;
  for (int b = 0; b < e->nbox && !stale; b++)
    {
      stokes_ctx *c = stokes_hip_ctx (e->sys[b]);
      if (!c) return STOKES_HIP_ENODEV;
      if (shk_dev_generation (c->dev) != e->gen[b]) stale = 1;
    }
  int rc = 0;
  if (stale)
    {
      rc = ens_capture (e, pos_dev, x_dev, y_dev);
      if (rc) return rc;
    }
  rc = shk_graph_launch_on (e->graph, e->main);
  if (!rc) rc = shk_stream_sync (e->main);
  if (rc)
    {
      stokes_hip_set_error ("ensemble launch: %s", shk_last_error ());
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
