/* stokes.c -- struct stokes set-up for the MI355X libstokes (host, C).
 *
 * The parameter selection (xi, r_max, k_max, image table) restates the
 * reference's stokes.c:654-776 and stokes.c:173-233 step for step: the
 * set of lattice images and reciprocal vectors decides which terms enter
 * the Ewald sums, so it must be chosen by the same rule to give the same
 * operator.  The reference's unused reciprocal table (stokes.c:235-307) is
 * not built here (nk stays 0); the device operator enumerates k itself.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "stokes.h"
#include "stokes_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static void *
xmalloc (size_t bytes, const char *who)
{
  void *p = malloc (bytes ? bytes : 1);
  if (p == NULL)
    {
      fprintf (stderr, "allocation failed in %s\n", who);
      exit (1);
    }
  return p;
}

static void
table_free (struct stokes *s)
{
  free (s->rlx); free (s->rly); free (s->rlz);
  free (s->rmx); free (s->rmy); free (s->rmz);
  s->rlx = s->rly = s->rlz = NULL;
  s->rmx = s->rmy = s->rmz = NULL;
  s->nr = 0;
  s->nk = 0;
  s->flag_table = 0;
}

struct stokes *
stokes_init (void)
{
  struct stokes *s = (struct stokes *) calloc (1, sizeof (struct stokes));
  if (s == NULL)
    {
      fprintf (stderr, "allocation failed in stokes_init\n");
      exit (1);
    }
  /* defaults of stokes.c:316-419 (everything else zero) */
  s->twobody_nmax = 100;
  s->twobody_lub = 1;
  return s;
}

void
stokes_free (struct stokes *s)
{
  if (s == NULL) return;
  stokes_hip_release_context (s);
  free (s->it);
  free (s->pos);
  free (s->a);
  table_free (s);
  free (s);
}

void
stokes_set_np (struct stokes *s, int np, int nm)
{
  s->np = np;
  s->nm = nm;
  free (s->pos);
  s->pos = (double *) xmalloc (sizeof (double) * 3 * (size_t) np,
                               "stokes_set_np");
}

void
stokes_set_Ui (struct stokes *s, double uix, double uiy, double uiz)
{
  s->Ui[0] = uix; s->Ui[1] = uiy; s->Ui[2] = uiz;
}

void
stokes_set_Oi (struct stokes *s, double oix, double oiy, double oiz)
{
  s->Oi[0] = oix; s->Oi[1] = oiy; s->Oi[2] = oiz;
}

void
stokes_set_Ei (struct stokes *s, double eixx, double eixy, double eixz,
               double eiyz, double eiyy)
{
  s->Ei[0] = eixx; s->Ei[1] = eixy; s->Ei[2] = eixz;
  s->Ei[3] = eiyz; s->Ei[4] = eiyy;
}

void
stokes_set_shear (struct stokes *s, int shear_mode, double shear_rate)
{
  s->shear_mode = shear_mode;
  s->shear_rate = shear_rate;
  memset (s->Ui, 0, sizeof (s->Ui));
  memset (s->Oi, 0, sizeof (s->Oi));
  memset (s->Ei, 0, sizeof (s->Ei));
  if (shear_mode == 1)
    {
      s->Oi[2] = -0.5 * shear_rate;
      s->Ei[1] = +0.5 * shear_rate;
    }
  else if (shear_mode == 2)
    {
      s->Oi[1] = +0.5 * shear_rate;
      s->Ei[2] = +0.5 * shear_rate;
    }
  else if (shear_mode != 0)
    fprintf (stderr, "stokes_set_shear(): invalid shear_mode %d\n",
             shear_mode);
}

double
stokes_get_shear_shift (struct stokes *s, double t, double t0, double s0)
{
  if (s->periodic == 0 || s->shear_mode == 0) return 0.0;
  double h;
  if (s->shear_mode == 1) h = s->ly;
  else if (s->shear_mode == 2) h = s->lz;
  else
    {
      fprintf (stderr, "stokes_set_shear(): invalid shear_mode %d\n",
               s->shear_mode);
      return 0.0;
    }
  /* bring the shift into [-lx/2, lx/2) */
  double shift = s0 + h * s->shear_rate * (t - t0);
  while (shift < -0.5 * s->lx) shift += s->lx;
  while (shift >= 0.5 * s->lx) shift -= s->lx;
  return shift;
}

void
stokes_set_shear_shift (struct stokes *s, double t, double t0, double s0)
{
  s->shear_shift = stokes_get_shear_shift (s, t, t0, s0);
}

void
stokes_set_l (struct stokes *s, double lx, double ly, double lz)
{
  s->pivol = M_PI / lx / ly / lz;
  s->lx = lx; s->ly = ly; s->lz = lz;
  /* 27-cell integer table: 0 is the primary cell, 1-8 the xy plane,
   * 9-17 the z = -1 layer, 18-26 the z = +1 layer (stokes.c:615-650) */
  static const int cx[9] = {0, -1, 0, 1, 0, -1, -1, 1, 1};
  static const int cy[9] = {0, 0, -1, 0, 1, -1, 1, -1, 1};
  for (int layer = 0; layer < 3; layer++)
    for (int c = 0; c < 9; c++)
      {
        int i = layer * 9 + c;
        s->ilx[i] = cx[c];
        s->ily[i] = cy[c];
        s->ilz[i] = layer == 0 ? 0 : (layer == 1 ? -1 : 1);
      }
}

/* tail estimates with the 1.15 safety factor (stokes.c:654-700) */
static double
tail_real (double r, double xi)
{
  double xr = xi * r / 1.15, xr2 = xr * xr, r2 = r * r;
  return erfc (xr) * (0.75 + 1.5 / r2) / r
    + exp (-xr2) * xi / sqrt (M_PI)
      * (4.5 + 3.0 * xr2 + (3.0 + xr2 * (14.0 + xr2 * (4.0 + xr2))) / r2);
}

static double
tail_recip (double k, double xi)
{
  double kx = k / xi / 1.15, kx2 = kx * kx, k2 = k * k;
  return 6.0 * M_PI * exp (-kx2 / 4.0) * (1.0 + k2 / 3.0) / k2
    * (1.0 + kx2 * (1.0 / 4.0 + kx2 / 8.0));
}

static void
table_push (struct stokes *s, int *cap, double x, double y, double z,
            int m1, int m2, int m3)
{
  if (s->nr == *cap)
    {
      *cap = *cap ? 2 * *cap : 256;
      s->rlx = realloc (s->rlx, sizeof (double) * *cap);
      s->rly = realloc (s->rly, sizeof (double) * *cap);
      s->rlz = realloc (s->rlz, sizeof (double) * *cap);
      s->rmx = realloc (s->rmx, sizeof (int) * *cap);
      s->rmy = realloc (s->rmy, sizeof (int) * *cap);
      s->rmz = realloc (s->rmz, sizeof (int) * *cap);
      if (!s->rlx || !s->rly || !s->rlz || !s->rmx || !s->rmy || !s->rmz)
        {
          fprintf (stderr, "allocation failed in stokes_set_xi\n");
          exit (1);
        }
    }
  int i = s->nr++;
  s->rlx[i] = x; s->rly[i] = y; s->rlz[i] = z;
  s->rmx[i] = m1; s->rmy[i] = m2; s->rmz[i] = m3;
}

void
stokes_set_xi (struct stokes *s, double xi, double ewald_eps)
{
  s->ewald_eps = ewald_eps;
  s->xi = xi;
  s->xi2 = xi * xi;
  s->xia2 = s->xi2;
  s->xiaspi = xi / sqrt (M_PI);

  s->self_a = 1.0  - s->xiaspi * (6.0 - 40.0 / 3.0 * s->xia2);
  s->self_c = 0.75 - s->xiaspi * s->xia2 * 10.0;
  s->self_m = 0.9  - s->xiaspi * s->xia2 * (12.0 - 30.24 * s->xia2);

  /* shrink r_max and k_max by 0.99 while the estimated tail (one shell
   * plus six neighbours one lattice step out) stays below ewald_eps */
  const double shrink = 0.99;
  double lmin = fmin (s->lx, fmin (s->ly, s->lz));
  double lmax = fmax (s->lx, fmax (s->ly, s->lz));
  double rmax = 10.0 * sqrt (-log (ewald_eps)) / xi;
  while (tail_real (rmax * shrink, xi)
         + 6.0 * tail_real (rmax * shrink + lmin, xi) < ewald_eps)
    rmax *= shrink;
  s->rmax2 = rmax * rmax;
  s->rmaxx = (int) (rmax / s->lx) + 1;
  s->rmaxy = (int) (rmax / s->ly) + 1;
  s->rmaxz = (int) (rmax / s->lz) + 1;

  double dk = 2.0 * M_PI / lmax;
  double kmax = 10.0 * sqrt (-log (ewald_eps)) * 2.0 * xi;
  while (tail_recip (kmax * shrink, xi)
         + 6.0 * tail_recip (kmax * shrink + dk, xi) < ewald_eps)
    kmax *= shrink;
  s->kmax = kmax;
  s->kmaxx = (int) (kmax * s->lx / 2.0 / M_PI);
  s->kmaxy = (int) (kmax * s->ly / 2.0 / M_PI);
  s->kmaxz = (int) (kmax * s->lz / 2.0 / M_PI);

  /* real-space image cells: keep a cell when the corner of the primary
   * cell nearest to it lies within r_max + |box diagonal| */
  table_free (s);
  int cap = 0;
  double diag = sqrt (s->lx * s->lx + s->ly * s->ly + s->lz * s->lz);
  for (int m1 = -s->rmaxx; m1 <= s->rmaxx; m1++)
    {
      double x = s->lx * (double) m1;
      double xx = x - (m1 > 0 ? 0.5 * s->lx : (m1 < 0 ? -0.5 * s->lx : 0.0));
      for (int m2 = -s->rmaxy; m2 <= s->rmaxy; m2++)
        {
          double y = s->ly * (double) m2;
          double yy = y - (m2 > 0 ? 0.5 * s->ly : (m2 < 0 ? -0.5 * s->ly : 0.0));
          for (int m3 = -s->rmaxz; m3 <= s->rmaxz; m3++)
            {
              double z = s->lz * (double) m3;
              double zz = z - (m3 > 0 ? 0.5 * s->lz
                               : (m3 < 0 ? -0.5 * s->lz : 0.0));
              if (sqrt (xx * xx + yy * yy + zz * zz) <= rmax + diag)
                table_push (s, &cap, x, y, z, m1, m2, m3);
            }
        }
    }
  s->flag_table = 1;
  stokes_hip_invalidate (s);
}

void
stokes_hip_shadow_split (const struct stokes *s, double tratio,
                         struct stokes *sh)
{
  *sh = *s;
  sh->rlx = sh->rly = sh->rlz = NULL;
  sh->rmx = sh->rmy = sh->rmz = NULL;
  sh->nr = 0;
  stokes_set_xi (sh, xi_by_tratio (sh, tratio), s->ewald_eps);
}

void
stokes_hip_shadow_free (struct stokes *sh)
{
  table_free (sh);
}

double
xi_by_tratio (struct stokes *s, double tratio)
{
  return pow (tratio, 1.0 / 6.0) * sqrt (M_PI)
    / pow (s->lx * s->ly * s->lz, 1.0 / 3.0);
}

void
stokes_set_iter (struct stokes *s, const char *solver, int max, int restart,
                 double eps, int debug, FILE *out)
{
  free (s->it);
  s->it = calloc (1, sizeof (struct iter));
  if (s->it == NULL) return;
  if (solver) snprintf (s->it->solver, sizeof (s->it->solver), "%s", solver);
  s->it->max = max;
  s->it->restart = restart;
  s->it->eps = eps;
  s->it->debug = debug;
  s->it->out = out;
}

void
stokes_set_pos (struct stokes *s, const double *pos)
{
  memcpy (s->pos, pos, sizeof (double) * 3 * (size_t) s->np);
}

void
stokes_set_pos_mobile (struct stokes *s, const double *pos)
{
  memcpy (s->pos, pos, sizeof (double) * 3 * (size_t) s->nm);
}

void
stokes_set_pos_fixed (struct stokes *s, const double *pos)
{
  memcpy (s->pos + 3 * (size_t) s->nm, pos,
          sizeof (double) * 3 * (size_t) (s->np - s->nm));
}

void
stokes_set_radius (struct stokes *s, const double *a)
{
  free (s->a);
  s->a = (double *) xmalloc (sizeof (double) * (size_t) s->np,
                             "stokes_set_radius");
  memcpy (s->a, a, sizeof (double) * (size_t) s->np);
}

void
stokes_unset_radius (struct stokes *s)
{
  free (s->a);
  s->a = NULL;
}
