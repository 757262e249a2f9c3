/* solve.c -- dense FP64 mobility solves of the FTS version (host side):
 * solve_mob_3fts_matrix_0 (ewald-3fts-matrix.c:570-652) and
 * solve_mix_3fts_matrix_0 (ewald-3fts-matrix.c:1192-1290), plus the
 * device-pointer entry point stokes_hip_solve_mix_dev.
 *
 * The reference builds the dense extracted-form matrix, turns it into the
 * true (F, T, S) -> (U, O, E) map, splits it into the known-force and
 * known-velocity blocks, inverts the latter with LAPACK (solve_linear,
 * matrix.c:266-312) and applies the merged matrix.  Here the matrix is
 * built and kept on the GPU and the known-velocity block is LU-factored and
 * solved there (solve.hip); the packing of the vectors per particle
 * (set_fts_by_FTS / set_FTS_by_fts, fts.c:1422-1489) is the reference's.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ewald-3fts-matrix.h"
#include "stokes_hip.h"
#include "stokes_internal.h"

int
stokes_hip_solve_mix_dev (struct stokes *sys, const double *b_dev,
                          double *x_dev)
{
  stokes_ctx *c = stokes_hip_ctx (sys);
  if (!c) return STOKES_HIP_ENODEV;
  if (sys->nm < 0 || sys->nm > sys->np)
    {
      stokes_hip_set_error ("solve: nm = %d outside [0, np = %d]", sys->nm,
                            sys->np);
      return STOKES_HIP_EARG;
    }
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  size_t n = (size_t) sys->np * stokes_dim (sys->version);
  double *mat = NULL;
  rc = shk_dev_malloc (c->dev, (void **) &mat, sizeof (double) * n * n);
  if (rc)
    {
      stokes_hip_set_error ("solve: %s", shk_last_error ());
      return STOKES_HIP_ENOMEM;
    }
  rc = shk_dev_matrix (c->dev, mat);
  if (!rc)
    rc = shk_dev_solve_mixed (c->dev, mat, sys->np, sys->nm, sys->version,
                              b_dev, x_dev);
  shk_dev_free (c->dev, mat);
  if (rc)
    {
      stokes_hip_set_error ("solve: %s", shk_last_error ());
      return rc;
    }
  return 0;
}

/* fts.c:1458-1489 */
static void
pack_fts (int n, double *fts, const double *f, const double *t,
          const double *s)
{
  for (int i = 0; i < n; i++)
    {
      memcpy (fts + 11 * i, f + 3 * i, sizeof (double) * 3);
      memcpy (fts + 11 * i + 3, t + 3 * i, sizeof (double) * 3);
      memcpy (fts + 11 * i + 6, s + 5 * i, sizeof (double) * 5);
    }
}

/* fts.c:1422-1456 */
static void
unpack_fts (int n, double *f, double *t, double *s, const double *fts)
{
  for (int i = 0; i < n; i++)
    {
      memcpy (f + 3 * i, fts + 11 * i, sizeof (double) * 3);
      memcpy (t + 3 * i, fts + 11 * i + 3, sizeof (double) * 3);
      memcpy (s + 5 * i, fts + 11 * i + 6, sizeof (double) * 5);
    }
}

static void
solve_host (struct stokes *sys, const double *b, double *x, const char *who)
{
  size_t bytes = sizeof (double) * 11 * (size_t) sys->np;
  double *bd = NULL, *xd = NULL;
  int rc = stokes_hip_malloc ((void **) &bd, bytes);
  if (!rc) rc = stokes_hip_malloc ((void **) &xd, bytes);
  if (!rc) rc = stokes_hip_memcpy_h2d (bd, b, bytes);
  if (!rc) rc = stokes_hip_solve_mix_dev (sys, bd, xd);
  if (!rc) rc = stokes_hip_memcpy_d2h (x, xd, bytes);
  if (bd) stokes_hip_free (bd);
  if (xd) stokes_hip_free (xd);
  if (rc)
    {
      fprintf (stderr, "libstokes-hip: %s failed (%d): %s\n", who, rc,
               stokes_hip_last_error ());
      exit (1);
    }
}

void
solve_mob_3fts_matrix_0 (struct stokes *sys, const double *f,
                         const double *t, const double *e, double *u,
                         double *o, double *s)
{
  if (sys->version != 2)
    {
      fprintf (stderr, "libstokes solve_mob_3fts_matrix_0 :"
               " the version is wrong. reset to FTS\n");
      sys->version = 2;
    }
  int np = sys->np, nm = sys->nm;
  double *b = malloc (sizeof (double) * 11 * (size_t) np);
  double *x = malloc (sizeof (double) * 11 * (size_t) np);
  if (!b || !x)
    {
      fprintf (stderr, "allocation failed in solve_mob_3fts_matrix_0\n");
      exit (1);
    }
  pack_fts (np, b, f, t, e);
  /* every particle mobile for this problem */
  sys->nm = np;
  solve_host (sys, b, x, "solve_mob_3fts_matrix_0");
  sys->nm = nm;
  unpack_fts (np, u, o, s, x);
  free (b);
  free (x);
}

void
solve_mix_3fts_matrix_0 (struct stokes *sys, const double *f,
                         const double *t, const double *e, const double *uf,
                         const double *of, const double *ef, double *u,
                         double *o, double *s, double *ff, double *tf,
                         double *sf)
{
  if (sys->version != 2)
    {
      fprintf (stderr, "libstokes solve_mix_3fts_matrix_0 :"
               " the version is wrong. reset to FTS\n");
      sys->version = 2;
    }
  int np = sys->np, nm = sys->nm;
  if (np == nm)
    {
      solve_mob_3fts_matrix_0 (sys, f, t, e, u, o, s);
      return;
    }
  int nf = np - nm;
  double *b = malloc (sizeof (double) * 11 * (size_t) np);
  double *x = malloc (sizeof (double) * 11 * (size_t) np);
  if (!b || !x)
    {
      fprintf (stderr, "allocation failed in solve_mix_3fts_matrix_0\n");
      exit (1);
    }
  pack_fts (nm, b, f, t, e);
  pack_fts (nf, b + 11 * (size_t) nm, uf, of, ef);
  solve_host (sys, b, x, "solve_mix_3fts_matrix_0");
  unpack_fts (nm, u, o, s, x);
  unpack_fts (nf, ff, tf, sf, x + 11 * (size_t) nm);
  free (b);
  free (x);
}
