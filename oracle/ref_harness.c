/* ref_harness.c -- glue that drives the compiled reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) together
 * with the reference's own ewald.c, f.c, ft.c, fts.c, non-ewald.c, bench.c,
 * chebyshev.c and KIrand.c, compiled where they lie under the reference tree.
 * The reference's stokes.c needs the absent libiter header, so the harness
 * fills struct stokes itself from an orc_sys prepared by the oracle's
 * restatement of stokes_set_l / stokes_set_xi (oracle.c).  No reference
 * source is copied; only its headers are included for the struct layout.
 */
#include <stdlib.h>
#include <string.h>
#include "stokes.h"      /* reference: struct stokes */
#include "ewald.h"       /* reference: atimes_ewald_3all, ... */
#include "non-ewald.h"
#include "chebyshev.h"
#include "KIrand.h"
#include "f.h"
#include "ft.h"
#include "fts.h"
#include "two-body-res.h"
#include "twobody.h"
#include "lub.h"
#include "oracle.h"

struct stokes *
ref_sys_new (const orc_sys *o)
{
  struct stokes *s = calloc (1, sizeof (struct stokes));
  s->version = o->version;
  s->np = s->nm = o->np;
  s->pos = malloc (sizeof (double) * 3 * o->np);
  memcpy (s->pos, o->pos, sizeof (double) * 3 * o->np);
  if (o->a)
    {
      s->a = malloc (sizeof (double) * o->np);
      memcpy (s->a, o->a, sizeof (double) * o->np);
    }
  s->rmin = o->rmin;
  s->periodic = o->periodic;
  s->shear_mode = o->shear_mode;
  s->shear_shift = o->shear_shift;
  s->ewald_eps = o->ewald_eps;
  s->rmax2 = o->rmax2;
  s->rmaxx = o->rmaxx; s->rmaxy = o->rmaxy; s->rmaxz = o->rmaxz;
  s->kmax = o->kmax;
  s->kmaxx = o->kmaxx; s->kmaxy = o->kmaxy; s->kmaxz = o->kmaxz;
  s->xi = o->xi; s->xi2 = o->xi2; s->xiaspi = o->xiaspi; s->xia2 = o->xia2;
  s->pivol = o->pivol;
  s->lx = o->lx; s->ly = o->ly; s->lz = o->lz;
  s->self_a = o->self_a; s->self_c = o->self_c; s->self_m = o->self_m;
  s->nr = o->nr;
  if (o->nr > 0)
    {
      size_t nd = sizeof (double) * o->nr, ni = sizeof (int) * o->nr;
      s->rlx = malloc (nd); memcpy (s->rlx, o->rlx, nd);
      s->rly = malloc (nd); memcpy (s->rly, o->rly, nd);
      s->rlz = malloc (nd); memcpy (s->rlz, o->rlz, nd);
      s->rmx = malloc (ni); memcpy (s->rmx, o->rmx, ni);
      s->rmy = malloc (ni); memcpy (s->rmy, o->rmy, ni);
      s->rmz = malloc (ni); memcpy (s->rmz, o->rmz, ni);
    }
  return s;
}

void
ref_sys_free (struct stokes *s)
{
  free (s->pos); free (s->a);
  free (s->rlx); free (s->rly); free (s->rlz);
  free (s->rmx); free (s->rmy); free (s->rmz);
  free (s);
}

void
ref_atimes (struct stokes *s, int n, const double *x, double *y)
{
  atimes_3all (n, x, y, s);
}

void
ref_matrix (struct stokes *s, double *mat)
{
  if (s->periodic) make_matrix_mob_ewald_3all (s, mat);
  else make_matrix_mob_nonewald_3all (s, mat);
}

double
ref_cpu_split (struct stokes *s, int which)
{
  return which == 1 ? s->cpu1 : (which == 2 ? s->cpu2 : s->cpu3);
}

static double ref_inv_sqrt (double x) { return 1.0 / __builtin_sqrt (x); }
static double ref_sqrt (double x) { return __builtin_sqrt (x); }
static double ref_cos (double x) { return __builtin_cos (x); }
static double ref_sin (double x) { return __builtin_sin (x); }

void
ref_cheb_coef (int ncheb, int func_id, double x0, double x1, double *a)
{
  double (*f) (double) = func_id == 0 ? ref_inv_sqrt
    : func_id == 1 ? ref_sqrt : func_id == 2 ? ref_cos : ref_sin;
  chebyshev_coef (ncheb, f, x0, x1, a);
}

double
ref_cheb_eval (int ncheb, const double *a, double x0, double x1, double x)
{
  return chebyshev_eval (ncheb, a, x0, x1, x);
}

typedef struct { int n; const double *m; } ref_dense;

static void
ref_dense_atimes (int n, const double *x, double *y, void *ud)
{
  const ref_dense *d = ud;
  for (int i = 0; i < n; i++)
    {
      double acc = 0.0;
      for (int j = 0; j < n; j++) acc += d->m[(size_t) i * n + j] * x[j];
      y[i] = acc;
    }
}

void
ref_cheb_eval_dense (int ncheb, const double *a, int n, const double *m,
                     const double *y, double *z, double e0, double e1)
{
  ref_dense d = {n, m};
  chebyshev_eval_atimes (ncheb, a, n, y, z, e0, e1, ref_dense_atimes, &d);
}

double
ref_cheb_error_minvsqrt_dense (int n, const double *m, const double *y,
                               const double *z)
{
  ref_dense d = {n, m};
  return chebyshev_error_minvsqrt (n, y, z, ref_dense_atimes, &d);
}

double
ref_cheb_error_Rsqrt_dense (int n, const double *m, const double *y,
                            const double *z)
{
  ref_dense d = {n, m};
  return chebyshev_error_Rsqrt (n, y, z, ref_dense_atimes, &d);
}

void
ref_cheb_eval_sys (int ncheb, const double *a, struct stokes *s,
                   const double *y, double *z, double e0, double e1)
{
  int n = s->np * (s->version == 0 ? 3 : (s->version == 1 ? 6 : 11));
  chebyshev_eval_atimes (ncheb, a, n, y, z, e0, e1, atimes_3all, s);
}

/* KIrand */
void *
ref_rng_new (unsigned long seed)
{
  struct KIrand *r = KIrand_init ();
  KIrand_init_genrand (r, seed);
  return r;
}

void ref_rng_free (void *r) { KIrand_free (r); }
unsigned long ref_rng_int32 (void *r) { return KIrand_genrand_int32 (r); }
double ref_rng_gaussian (void *r) { return KIrand_Gaussian (r); }

/* ---- lubrication ------------------------------------------------------ */

/* the reference's two-body tables twobody_XA ... twobody_ZM (XA.c ... ZM.c),
 * family k in that order */
void
ref_twobody_table (int k, int n, double l, double *f)
{
  void (*t[11]) (int, double, double *) = {
    twobody_XA, twobody_YA, twobody_YB, twobody_XC, twobody_YC, twobody_XG,
    twobody_YG, twobody_YH, twobody_XM, twobody_YM, twobody_ZM};
  t[k] (n, l, f);
}

/* twobody_scalars_res (twobody.c:2786-2814) with the table of l = ab / aa */
void
ref_twobody_res (int version, double r, double aa, double ab, int n,
                 int flag_lub, int flag_scale, double *res)
{
  struct twobody_f *f = twobody_f_init (n, ab / aa);
  twobody_scalars_res (version, r, aa, ab, f, n, flag_lub, flag_scale, res);
  twobody_f_free (f);
}

void ref_two_body_res (double s, double *res) { scalar_two_body_res (s, res); }
void ref_minv_fts (double s, double *res) { scalar_minv_fts (s, res); }

/* matrix_fts_self_atimes / matrix_fts_atimes (fts.c): y += B(e; sc) x with
 * sc the 11 scalars xa ya yb xc yc xg yg yh xm ym zm */
void
ref_fts_block (int self, const double *x, double *y, const double *e,
               const double *sc)
{
  if (self)
    matrix_fts_self_atimes (x, y, e[0], e[1], e[2], sc[0], sc[1], sc[2],
                            sc[3], sc[4], sc[5], sc[6], sc[7], sc[8], sc[9],
                            sc[10]);
  else
    matrix_fts_atimes (x, y, e[0], e[1], e[2], sc[0], sc[1], sc[2], sc[3],
                       sc[4], sc[5], sc[6], sc[7], sc[8], sc[9], sc[10]);
}

int
ref_cond_lub (const double *x1, const double *x2, double a1, double a2,
              double lubmax2, int poly)
{
  return poly ? cond_lub_poly (x1, x2, a1, a2, lubmax2)
              : cond_lub (x1, x2, lubmax2);
}

/* calc_lub_3f / _3ft / _3fts (lub.c:164-768) for equal spheres: the
 * reference's pair loop (j >= i, image cells -imax..imax, cond_lub) over
 * its compiled calc_lub_f_2b (f.c:234), calc_lub_ft_2b (ft.c:494) or
 * calc_lub_fts_2b (fts.c:1758-1869); lub.c itself also reaches the
 * unequal-sphere path, which needs LAPACK (minv-poly.c), so the loop is
 * driven here */
void
ref_lub_3fts_mono (struct stokes *sys, double lubmin2, double lubmax,
                   const double *uoe, double *fts)
{
  int np = sys->np;
  const int d = sys->version == 0 ? 3 : (sys->version == 1 ? 6 : 11);
  void (*pair) (struct stokes *, const double *, const double *,
                const double *, const double *, double *, double *)
    = sys->version == 0 ? calc_lub_f_2b
      : (sys->version == 1 ? calc_lub_ft_2b : calc_lub_fts_2b);
  sys->lubmin2 = lubmin2;
  sys->lubmax = lubmax;
  for (int i = 0; i < np * d; i++) fts[i] = 0.0;
  double lubmax2 = lubmax * lubmax;
  int ix0 = 0, iy0 = 0, iz0 = 0;
  if (sys->periodic)
    {
      double lm = lubmax;
      if (lm <= 0.0)
        {
          lm = sys->lx;
          if (sys->ly > lm) lm = sys->ly;
          if (sys->lz > lm) lm = sys->lz;
        }
      ix0 = (int) (lm / sys->lx) + 1;
      iy0 = (int) (lm / sys->ly) + 1;
      iz0 = (int) (lm / sys->lz) + 1;
    }
  for (int i = 0; i < np; i++)
    for (int j = i; j < np; j++)
      for (int ix = -ix0; ix <= ix0; ix++)
        for (int iy = -iy0; iy <= iy0; iy++)
          for (int iz = -iz0; iz <= iz0; iz++)
            {
              double p[3] = {sys->pos[3 * j] + sys->lx * ix,
                             sys->pos[3 * j + 1] + sys->ly * iy,
                             sys->pos[3 * j + 2] + sys->lz * iz};
              if (sys->shear_mode == 1) p[0] += iy * sys->shear_shift;
              else if (sys->shear_mode == 2) p[0] += iz * sys->shear_shift;
              if (cond_lub (sys->pos + 3 * i, p, lubmax2) == 0)
                pair (sys, uoe + d * i, uoe + d * j, sys->pos + 3 * i, p,
                      fts + d * i, fts + d * j);
            }
}

/* make_matrix_lub_3f / 3ft / 3fts (lub-matrix.c:76-507) for equal spheres:
 * the same pair loop over matrix_lub_{f,ft,fts}_2b (f.c:308, ft.c:579,
 * fts.c:1871); lub-matrix.c itself also names the unequal-sphere pair
 * routines, which need LAPACK (minv-poly.c), so the loop is driven here */
void
ref_lub_matrix_mono (struct stokes *sys, double lubmin2, double lubmax,
                     double *mat)
{
  int np = sys->np;
  const int d = sys->version == 0 ? 3 : (sys->version == 1 ? 6 : 11);
  const int n = np * d;
  void (*pair) (struct stokes *, int, int, const double *, const double *,
                int, double *)
    = sys->version == 0 ? matrix_lub_f_2b
      : (sys->version == 1 ? matrix_lub_ft_2b : matrix_lub_fts_2b);
  sys->lubmin2 = lubmin2;
  sys->lubmax = lubmax;
  for (long k = 0; k < (long) n * n; k++) mat[k] = 0.0;
  double lubmax2 = lubmax * lubmax;
  int ix0 = 0, iy0 = 0, iz0 = 0;
  if (sys->periodic)
    {
      double lm = lubmax;
      if (lm <= 0.0)
        {
          lm = sys->lx;
          if (sys->ly > lm) lm = sys->ly;
          if (sys->lz > lm) lm = sys->lz;
        }
      ix0 = (int) (lm / sys->lx) + 1;
      iy0 = (int) (lm / sys->ly) + 1;
      iz0 = (int) (lm / sys->lz) + 1;
    }
  for (int i = 0; i < np; i++)
    for (int j = i; j < np; j++)
      for (int ix = -ix0; ix <= ix0; ix++)
        for (int iy = -iy0; iy <= iy0; iy++)
          for (int iz = -iz0; iz <= iz0; iz++)
            {
              double p[3] = {sys->pos[3 * j] + sys->lx * ix,
                             sys->pos[3 * j + 1] + sys->ly * iy,
                             sys->pos[3 * j + 2] + sys->lz * iz};
              if (sys->shear_mode == 1) p[0] += iy * sys->shear_shift;
              else if (sys->shear_mode == 2) p[0] += iz * sys->shear_shift;
              if (cond_lub (sys->pos + 3 * i, p, lubmax2) == 0)
                pair (sys, i, j, sys->pos + 3 * i, p, n, mat);
            }
}
