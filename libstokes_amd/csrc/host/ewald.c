/* ewald.c -- the mobility operator entry points (host side, C).
 *
 * atimes_3all and friends keep the reference's prototypes (ewald.h) and
 * hand the work to the HIP layer: positions, image table and reciprocal
 * vectors go to the device, the operator runs there, and the result comes
 * back.  One device context is kept per struct stokes; the parameters are
 * re-uploaded when they change and the positions on every call, because
 * callers move sys->pos in place between calls (brownian.c:2392, 2466).
 */
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "stokes.h"
#include "ewald.h"
#include "non-ewald.h"
#include "matrix.h"
#include "stokes_hip.h"
#include "stokes_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* errors                                                               */
/* ------------------------------------------------------------------ */

static __thread char g_msg[512];

void
stokes_hip_set_error (const char *fmt, ...)
{
  va_list ap;
  va_start (ap, fmt);
  vsnprintf (g_msg, sizeof (g_msg), fmt, ap);
  va_end (ap);
}

const char *
stokes_hip_last_error (void)
{
  return g_msg;
}

static int
dev_fail (int rc, const char *where)
{
  stokes_hip_set_error ("%s: %s", where, shk_last_error ());
  return rc;
}

/* the reference's void entry points abort on failure (memory-check.h) */
static void
die_on (int rc, const char *where)
{
  if (rc == 0) return;
  fprintf (stderr, "libstokes-hip: %s failed (%d): %s\n", where, rc,
           stokes_hip_last_error ());
  exit (1);
}

int
stokes_dim (int version)
{
  return version == 0 ? 3 : (version == 1 ? 6 : 11);
}

/* ------------------------------------------------------------------ */
/* contexts                                                             */
/* ------------------------------------------------------------------ */

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static stokes_ctx **g_ctx = NULL;
static int g_nctx = 0, g_cap = 0;
static shk_dev *g_free_dev = NULL;

stokes_ctx *
stokes_hip_ctx (struct stokes *sys)
{
  stokes_ctx *c = NULL;
  pthread_mutex_lock (&g_lock);
  for (int i = 0; i < g_nctx; i++)
    if (g_ctx[i]->sys == sys) { c = g_ctx[i]; break; }
  if (c == NULL)
    {
      shk_dev *d = shk_dev_create ();
      if (d == NULL)
        stokes_hip_set_error ("no GPU context: %s", shk_last_error ());
      else
        {
          c = (stokes_ctx *) calloc (1, sizeof (stokes_ctx));
          c->sys = sys;
          c->dev = d;
          c->prec = SHK_FP64;
          if (g_nctx == g_cap)
            {
              g_cap = g_cap ? 2 * g_cap : 16;
              g_ctx = realloc (g_ctx, sizeof (stokes_ctx *) * g_cap);
            }
          g_ctx[g_nctx++] = c;
        }
    }
  pthread_mutex_unlock (&g_lock);
  return c;
}

/* the context of sys if one exists (none is created) */
stokes_ctx *
stokes_hip_ctx_find (struct stokes *sys)
{
  stokes_ctx *c = NULL;
  pthread_mutex_lock (&g_lock);
  for (int i = 0; i < g_nctx; i++)
    if (g_ctx[i]->sys == sys) { c = g_ctx[i]; break; }
  pthread_mutex_unlock (&g_lock);
  return c;
}

static void
ctx_free (stokes_ctx *c)
{
  if (c->x_dev) shk_dev_free (c->dev, c->x_dev);
  if (c->y_dev) shk_dev_free (c->dev, c->y_dev);
  if (c->mat_dev) shk_dev_free (c->dev, c->mat_dev);
  if (c->cheb_dev) shk_dev_free (c->dev, c->cheb_dev);
  shk_dev_destroy (c->dev);
  free (c);
}

void
stokes_hip_release_context (struct stokes *sys)
{
  pthread_mutex_lock (&g_lock);
  for (int i = 0; i < g_nctx; i++)
    if (g_ctx[i]->sys == sys)
      {
        stokes_ctx *c = g_ctx[i];
        g_ctx[i] = g_ctx[--g_nctx];
        pthread_mutex_unlock (&g_lock);
        ctx_free (c);
        return;
      }
  pthread_mutex_unlock (&g_lock);
}

void
stokes_hip_invalidate (struct stokes *sys)
{
  pthread_mutex_lock (&g_lock);
  for (int i = 0; i < g_nctx; i++)
    if (g_ctx[i]->sys == sys) g_ctx[i]->valid = 0;
  pthread_mutex_unlock (&g_lock);
}

shk_dev *
stokes_hip_global_dev (void)
{
  pthread_mutex_lock (&g_lock);
  if (g_free_dev == NULL)
    {
      g_free_dev = shk_dev_create ();
      if (g_free_dev == NULL)
        stokes_hip_set_error ("no GPU context: %s", shk_last_error ());
    }
  pthread_mutex_unlock (&g_lock);
  return g_free_dev;
}

/* ------------------------------------------------------------------ */
/* parameter upload                                                     */
/* ------------------------------------------------------------------ */

/* Reciprocal vectors of ewald.c:923-958 restricted to the half space
 * m1 > 0 | (m1 == 0 & m2 > 0) | (m1 == m2 == 0 & m3 > 0); every k is paired
 * with -k (identical contribution: cos is even, and the odd blocks B, G
 * come with the odd factor sin), so the weights carry a factor 2. */
static int
build_kvectors (const struct stokes *s, double gsign, double **kv_out,
                double **ks_out, int **km_out)
{
  int cap = 64, nk = 0;
  double *kv = malloc (sizeof (double) * 4 * cap);
  double *ks = malloc (sizeof (double) * 8 * cap);
  int *km = malloc (sizeof (int) * 4 * cap);
  double k0x = 2.0 * M_PI / s->lx;
  double k0y = 2.0 * M_PI / s->ly;
  double k0z = 2.0 * M_PI / s->lz;
  double sl = s->shear_shift / s->lx;
  for (int m1 = 0; m1 <= s->kmaxx; m1++)
    for (int m2 = -s->kmaxy; m2 <= s->kmaxy; m2++)
      for (int m3 = -s->kmaxz; m3 <= s->kmaxz; m3++)
        {
          if (m1 == 0 && (m2 < 0 || (m2 == 0 && m3 <= 0))) continue;
          double k1 = k0x * (double) m1;
          double k2 = k0y * (double) m2;
          double k3 = k0z * (double) m3;
          if (s->shear_mode == 1)
            k2 = k0y * ((double) m2 - sl * (double) m1);
          else if (s->shear_mode == 2)
            k3 = k0z * ((double) m3 - sl * (double) m1);
          double kk = k1 * k1 + k2 * k2 + k3 * k3;
          double k = sqrt (kk);
          if (s->kmax != 0.0 && k > s->kmax) continue;
          double k4z = kk / 4.0 / s->xi2;
          double kexp = s->pivol * (1.0 + k4z * (1.0 + 2.0 * k4z)) / kk
            * exp (-k4z);
          if (nk == cap)
            {
              cap *= 2;
              kv = realloc (kv, sizeof (double) * 4 * cap);
              ks = realloc (ks, sizeof (double) * 8 * cap);
              km = realloc (km, sizeof (int) * 4 * cap);
            }
          double *v = kv + 4 * nk, *w = ks + 8 * nk;
          km[4 * nk] = m1; km[4 * nk + 1] = m2; km[4 * nk + 2] = m3;
          /* first k of a lattice row (m1, m2 fixed, m3 consecutive) */
          km[4 * nk + 3] = !(nk > 0 && km[4 * nk - 4] == m1 && km[4 * nk - 3] == m2
                             && km[4 * nk - 2] == m3 - 1);
          v[0] = k1; v[1] = k2; v[2] = k3; v[3] = k;
          w[0] = 2.0 * 6.0 * (1.0 - kk / 3.0) * kexp;              /* ya */
          w[1] = 2.0 * 3.0 * k * kexp;                             /* yb */
          w[2] = 2.0 * 3.0 / 2.0 * kk * kexp;                      /* yc */
          /* yg; unequal spheres: the signed kexp of the G terms */
          w[3] = s->a ? gsign * kexp
            : gsign * 2.0 * 3.0 * (1.0 - 4.0 / 15.0 * kk) * k * kexp;
          w[4] = 2.0 * 3.0 / 2.0 * kk * kexp;                      /* yh */
          w[5] = 2.0 * 3.0 * (1.0 - kk / 5.0) * kk * kexp;         /* ym */
          w[6] = kk;
          w[7] = kexp;
          nk++;
        }
  *kv_out = kv;
  *ks_out = ks;
  if (km_out) *km_out = km;
  else free (km);
  return nk;
}

/* Reference-sign G term at the splitting of sys, added to reciprocal
 * tables built at another splitting with the xi-independent sign.
 *
 * With either sign s_G of the reciprocal G / G-tilde weight, the FTS sum is
 * y(xi; s_G) = y(xi; -1) + (1 + s_G) K_G(xi), K_G the reciprocal G part
 * with weight +yg.  y(xi; -1) does not depend on xi (to the Ewald
 * tolerance), so the reference's y(xi_1; +1) = y(xi_t; -1) + 2 K_G(xi_1):
 * twice the G weight at xi_1 (yg for equal spheres, kexp for unequal ones,
 * whose G scalars are formed in the kernels) goes onto the G weight of each
 * k of the xi_1 sphere -- the same lattice vector in the xi_t tables, or a
 * new G-only entry.  Returns the new count (the tables are reallocated), or
 * -1. */
static int
add_g_correction (const struct stokes *s, double **kv_io, double **ks_io,
                  int **km_io, int nk_t)
{
  int *km_t = *km_io;
  double *kv1 = NULL, *ks1 = NULL;
  int *km1 = NULL;
  int nk1 = build_kvectors (s, 1.0, &kv1, &ks1, &km1);
  int KX = 0, KY = 0, KZ = 0;
  for (int k = 0; k < nk_t; k++)
    {
      KX = km_t[4 * k] > KX ? km_t[4 * k] : KX;
      KY = abs (km_t[4 * k + 1]) > KY ? abs (km_t[4 * k + 1]) : KY;
      KZ = abs (km_t[4 * k + 2]) > KZ ? abs (km_t[4 * k + 2]) : KZ;
    }
  long ny = 2 * KY + 1, nz = 2 * KZ + 1;
  long ncell = (KX + 1) * ny * nz;
  int *slot = malloc (sizeof (int) * (ncell > 0 ? ncell : 1));
  double *kv = realloc (*kv_io, sizeof (double) * 4 * (nk_t + nk1 + 1));
  double *ks = realloc (*ks_io, sizeof (double) * 8 * (nk_t + nk1 + 1));
  int *km = realloc (km_t, sizeof (int) * 4 * (nk_t + nk1 + 1));
  if (kv) *kv_io = kv;
  if (ks) *ks_io = ks;
  if (km) *km_io = km;
  if (!slot || !kv || !ks || !km || !kv1 || !ks1 || !km1)
    {
      free (slot); free (kv1); free (ks1); free (km1);
      return -1;
    }
  km_t = km;
  for (long c = 0; c < ncell; c++) slot[c] = -1;
  for (int k = 0; k < nk_t; k++)
    slot[(km_t[4 * k] * ny + km_t[4 * k + 1] + KY) * nz + km_t[4 * k + 2] + KZ] = k;
  int nk = nk_t;
  for (int k = 0; k < nk1; k++)
    {
      int m1 = km1[4 * k], m2 = km1[4 * k + 1], m3 = km1[4 * k + 2];
      int t = -1;
      if (m1 <= KX && abs (m2) <= KY && abs (m3) <= KZ)
        t = slot[(m1 * ny + m2 + KY) * nz + m3 + KZ];
      if (t < 0)
        {
          t = nk++;
          memcpy (kv + 4 * t, kv1 + 4 * k, sizeof (double) * 4);
          for (int q = 0; q < 8; q++) ks[8 * t + q] = 0.0;
          ks[8 * t + 6] = ks1[8 * k + 6];   /* kk; kexp 0: G terms only */
          memcpy (km + 4 * t, km1 + 4 * k, sizeof (int) * 3);
          km[4 * t + 3] = !(t > 0 && km[4 * t - 4] == km[4 * t]
                            && km[4 * t - 3] == km[4 * t + 1]
                            && km[4 * t - 2] == km[4 * t + 2] - 1);
        }
      ks[8 * t + 3] += 2.0 * ks1[8 * k + 3];
    }
  free (slot); free (kv1); free (ks1); free (km1);
  return nk;
}

/* Lattice vectors that can hold an image inside r_max once x_j - x_i is
 * wrapped into the primary cell (|components| <= half the box): those with
 * |v| <= r_max + R, R the half diagonal, sorted by |v| (ties by index);
 * *ncore counts the leading ones with |v| <= r_max - R, inside r_max for
 * every pair.  Returns the count, or -1. */
typedef struct { double v[4]; int m[4]; } cand_t;

static int
cand_cmp (const void *a, const void *b)
{
  const cand_t *x = a, *y = b;
  if (x->v[3] < y->v[3]) return -1;
  if (x->v[3] > y->v[3]) return 1;
  for (int k = 0; k < 3; k++)
    if (x->m[k] != y->m[k]) return x->m[k] < y->m[k] ? -1 : 1;
  return 0;
}

static int
build_candidates (const struct stokes *s, double s2x, double s3x,
                  double **cand_out, int **candm_out, int *ncore)
{
  double rmax = sqrt (s->rmax2);
  double rh = 0.5 * sqrt (s->lx * s->lx + s->ly * s->ly + s->lz * s->lz);
  double R = rmax + rh;
  int M3 = (int) ceil (R / s->lz), M2 = (int) ceil (R / s->ly);
  int cap = 256, n = 0;
  cand_t *c = malloc (sizeof (cand_t) * cap);
  if (!c) return -1;
  for (int m3 = -M3; m3 <= M3; m3++)
    for (int m2 = -M2; m2 <= M2; m2++)
      {
        double off = m2 * s2x + m3 * s3x;
        int lo = (int) floor ((-R - off) / s->lx), hi = (int) ceil ((R - off) / s->lx);
        for (int m1 = lo; m1 <= hi; m1++)
          {
            double x = m1 * s->lx + off, y = m2 * s->ly, z = m3 * s->lz;
            double r = sqrt (x * x + y * y + z * z);
            if (r > R) continue;
            if (n == cap)
              {
                cap *= 2;
                cand_t *cc = realloc (c, sizeof (cand_t) * cap);
                if (!cc) { free (c); return -1; }
                c = cc;
              }
            c[n].v[0] = x; c[n].v[1] = y; c[n].v[2] = z; c[n].v[3] = r;
            c[n].m[0] = m1; c[n].m[1] = m2; c[n].m[2] = m3; c[n].m[3] = 0;
            n++;
          }
      }
  qsort (c, n, sizeof (cand_t), cand_cmp);
  double *cv = malloc (sizeof (double) * 4 * (n > 0 ? n : 1));
  int *cm = malloc (sizeof (int) * 4 * (n > 0 ? n : 1));
  if (!cv || !cm) { free (c); free (cv); free (cm); return -1; }
  int nc = 0;
  for (int k = 0; k < n; k++)
    {
      memcpy (cv + 4 * k, c[k].v, sizeof (double) * 4);
      memcpy (cm + 4 * k, c[k].m, sizeof (int) * 4);
      if (c[k].v[3] <= rmax - rh) nc = k + 1;
    }
  free (c);
  *cand_out = cv;
  *candm_out = cm;
  *ncore = nc;
  return n;
}

/* bitmap of the reference image table (stokes.c:173-233) over
 * [-rmaxx, rmaxx] x [-rmaxy, rmaxy] x [-rmaxz, rmaxz], z fastest */
static int
build_table_bits (const struct stokes *s, unsigned **bits_out, int *tx, int *ty,
                  int *tz)
{
  *tx = s->rmaxx; *ty = s->rmaxy; *tz = s->rmaxz;
  long nb = (long) (2 * *tx + 1) * (2 * *ty + 1) * (2 * *tz + 1);
  int words = (int) ((nb + 31) / 32);
  unsigned *b = calloc (words > 0 ? words : 1, sizeof (unsigned));
  if (!b) return -1;
  for (int m = 0; m < s->nr; m++)
    {
      int m1 = s->rmx[m], m2 = s->rmy[m], m3 = s->rmz[m];
      if (m1 < -*tx || m1 > *tx || m2 < -*ty || m2 > *ty || m3 < -*tz || m3 > *tz)
        continue;
      long idx = ((long) (m1 + *tx) * (2 * *ty + 1) + (m2 + *ty)) * (2 * *tz + 1)
        + (m3 + *tz);
      b[idx >> 5] |= 1u << (idx & 31);
    }
  *bits_out = b;
  return words;
}

static int
same_params (const shk_params *a, const shk_params *b)
{
  return memcmp (a, b, sizeof (shk_params)) == 0;
}

/* the parameters of sys as the device path takes them (pu), and whether
 * the lattice tables have to be rebuilt and uploaded (the return value) */
static int
prepare_params (stokes_ctx *c, shk_params *pup)
{
  struct stokes *s = c->sys;
  shk_params pu;
  memset (&pu, 0, sizeof (pu));
  pu.version = s->version;
  pu.np = s->np;
  pu.periodic = s->periodic ? 1 : 0;
  pu.poly = s->a != NULL;
  pu.lx = s->lx; pu.ly = s->ly; pu.lz = s->lz; pu.pivol = s->pivol;
  pu.xi = s->xi; pu.xi2 = s->xi2; pu.xiaspi = s->xiaspi; pu.xia2 = s->xia2;
  pu.rmax2 = s->rmax2; pu.rmin = s->rmin;
  pu.self_a = s->self_a; pu.self_c = s->self_c; pu.self_m = s->self_m;
  pu.nr = pu.periodic ? s->nr : 0;
  pu.gsign = c->fix_recip_g ? -1.0 : 1.0;
  const int split = c->split_tr > 0.0 && pu.periodic && s->shear_mode == 0;
  int full = !c->valid || c->shear_mode != s->shear_mode
    || c->shear_shift != s->shear_shift || pu.poly
    || c->a_ptr != s->a || (split ? c->split_tr : 0.0) != c->split_used_tr;
  /* nk is only known after enumeration; compare the rest */
  pu.nk = c->prm_user.nk;
  if (!full && !same_params (&pu, &c->prm_user)) full = 1;
  *pup = pu;
  return full;
}

/* the device parameters p and host lattice tables t of s (user parameters
 * pu from prepare_params) computed at the splitting of split_tr (0: s's own
 * xi); gcorr: restore the reference-sign G term of that other splitting.
 * Release t with stokes_hip_tables_free. */
static int
build_tables (const struct stokes *s, const shk_params *pu, double split_tr,
              int gcorr, shk_params *p_out, shk_tables *t_out)
{
  double *img = NULL, *kv = NULL, *ks = NULL, *cand = NULL;
  int *candm = NULL, *km = NULL;
  unsigned *tbits = NULL;
  shk_tables t;
  memset (&t, 0, sizeof (t));
  struct stokes sh;
  const struct stokes *src = s;
  shk_params p = *pu;
  const int split = split_tr > 0.0;
  if (split)
    {
      stokes_hip_shadow_split (s, split_tr, &sh);
      src = &sh;
      p.xi = sh.xi; p.xi2 = sh.xi2; p.xiaspi = sh.xiaspi; p.xia2 = sh.xia2;
      p.rmax2 = sh.rmax2;
      p.self_a = sh.self_a; p.self_c = sh.self_c; p.self_m = sh.self_m;
      p.nr = sh.nr;
      if (gcorr) p.gsign = -1.0;
    }
  if (p.periodic)
    {
      if (p.nr <= 0 || src->rlx == NULL)
        {
          if (split) stokes_hip_shadow_free (&sh);
          stokes_hip_set_error ("periodic system without image table"
                                " (call stokes_set_xi)");
          return STOKES_HIP_EARG;
        }
      img = malloc (sizeof (double) * 4 * p.nr);
      for (int m = 0; img && m < p.nr; m++)
        {
          double sx = 0.0;
          if (s->shear_mode == 1) sx = (double) src->rmy[m] * s->shear_shift;
          else if (s->shear_mode == 2) sx = (double) src->rmz[m] * s->shear_shift;
          img[4 * m + 0] = src->rlx[m] + sx;
          img[4 * m + 1] = src->rly[m];
          img[4 * m + 2] = src->rlz[m];
          img[4 * m + 3] = 0.0;
        }
      p.nk = build_kvectors (src, p.gsign, &kv, &ks, &km);
      if (split && gcorr && p.nk >= 0)
        p.nk = add_g_correction (s, &kv, &ks, &km, p.nk);
      t.km = km;
      t.kbasis_s1 = s->shear_mode == 1 ? s->shear_shift / s->lx : 0.0;
      t.kbasis_s2 = s->shear_mode == 2 ? s->shear_shift / s->lx : 0.0;
      for (int k = 0; k < p.nk && km; k++)
        {
          if (km[4 * k] > t.kmax1) t.kmax1 = km[4 * k];
          if (abs (km[4 * k + 1]) > t.kmax2) t.kmax2 = abs (km[4 * k + 1]);
          if (abs (km[4 * k + 2]) > t.kmax3) t.kmax3 = abs (km[4 * k + 2]);
        }
      t.img = img;
      t.kv = kv;
      t.ks = ks;
      t.s2x = s->shear_mode == 1 ? s->shear_shift : 0.0;
      t.s3x = s->shear_mode == 2 ? s->shear_shift : 0.0;
      t.ncand = build_candidates (src, t.s2x, t.s3x, &cand, &candm, &t.ncore);
      t.cand = cand;
      t.candm = candm;
      t.tbits_words = build_table_bits (src, &tbits, &t.tx, &t.ty, &t.tz);
      t.tbits = tbits;
      if (!img || t.ncand < 0 || t.tbits_words < 0 || p.nk < 0)
        {
          stokes_hip_tables_free (&t);
          if (split) stokes_hip_shadow_free (&sh);
          stokes_hip_set_error ("lattice tables: allocation failed");
          return STOKES_HIP_ENOMEM;
        }
    }
  else
    p.nk = 0;
  if (split) stokes_hip_shadow_free (&sh);
  *p_out = p;
  *t_out = t;
  return 0;
}

void
stokes_hip_tables_free (shk_tables *t)
{
  free ((void *) t->img); free ((void *) t->kv); free ((void *) t->ks);
  free ((void *) t->km); free ((void *) t->cand); free ((void *) t->candm);
  free ((void *) t->tbits);
  memset (t, 0, sizeof (*t));
}

int
stokes_hip_tables (const struct stokes *s, double split_tr, int fix_g,
                   shk_params *p, shk_tables *t)
{
  stokes_ctx c;
  memset (&c, 0, sizeof (c));
  c.sys = (struct stokes *) s;
  c.fix_recip_g = fix_g;
  shk_params pu;
  prepare_params (&c, &pu);
  const int split = split_tr > 0.0 && pu.periodic && s->shear_mode == 0;
  return build_tables (s, &pu, split ? split_tr : 0.0, s->version == 2 && !fix_g, p, t);
}

int
stokes_hip_prepare (stokes_ctx *c, int positions_on_device)
{
  struct stokes *s = c->sys;
  if (s->np <= 0 || s->pos == NULL)
    {
      stokes_hip_set_error ("system has no particles");
      return STOKES_HIP_EARG;
    }
  if (s->version < 0 || s->version > 2)
    {
      stokes_hip_set_error ("invalid version %d", s->version);
      return STOKES_HIP_EARG;
    }
  if (s->slip != NULL)
    {
      stokes_hip_set_error ("slip particles are not supported");
      return STOKES_HIP_EUNSUP;
    }
  shk_params pu;
  const int full = prepare_params (c, &pu);
  /* another splitting computes the same sum (to the Ewald tolerance) except
   * for the reference-sign G term of FTS, which add_g_correction restores */
  const int gcorr = s->version == 2 && !c->fix_recip_g;
  const int split = c->split_tr > 0.0 && pu.periodic && s->shear_mode == 0;
  if (full)
    {
      shk_params p;
      shk_tables t;
      int rc = build_tables (s, &pu, split ? c->split_tr : 0.0, gcorr, &p, &t);
      if (rc) return rc;
      rc = shk_dev_set_system (c->dev, &p, s->pos, s->a, p.periodic ? &t : NULL);
      stokes_hip_tables_free (&t);
      if (rc) return dev_fail (rc, "upload");
      c->prm = p;
      pu.nk = p.nk;
      c->prm_user = pu;
      c->split_used_tr = split ? c->split_tr : 0.0;
      c->shear_mode = s->shear_mode;
      c->shear_shift = s->shear_shift;
      c->a_ptr = s->a;
      c->valid = 1;
      c->uploads++;
      if (positions_on_device && c->pos_dev)
        {
          rc = shk_dev_set_pos (c->dev, c->pos_dev, 1);
          if (rc) return dev_fail (rc, "positions");
        }
      return 0;
    }
  int rc;
  if (positions_on_device && c->pos_dev)
    rc = shk_dev_set_pos (c->dev, c->pos_dev, 1);
  else
    rc = shk_dev_set_pos (c->dev, s->pos, 0);
  if (rc) return dev_fail (rc, "positions");
  return 0;
}

static int
ensure_xy (stokes_ctx *c, size_t bytes)
{
  if (c->xy_bytes >= bytes) return 0;
  if (c->x_dev) shk_dev_free (c->dev, c->x_dev);
  if (c->y_dev) shk_dev_free (c->dev, c->y_dev);
  c->x_dev = c->y_dev = NULL;
  c->xy_bytes = 0;
  int rc = shk_dev_malloc (c->dev, (void **) &c->x_dev, bytes);
  if (!rc) rc = shk_dev_malloc (c->dev, (void **) &c->y_dev, bytes);
  if (rc) return dev_fail (rc, "staging buffers");
  c->xy_bytes = bytes;
  return 0;
}

static void
record_times (stokes_ctx *c)
{
  if (!c->timing) return;
  c->sys->cpu2 = shk_dev_last_ms (c->dev, 0);
  c->sys->cpu3 = shk_dev_last_ms (c->dev, 1);
  c->sys->cpu1 = c->sys->cpu2 + c->sys->cpu3;
}

/* ------------------------------------------------------------------ */
/* new entry points                                                     */
/* ------------------------------------------------------------------ */

static stokes_ctx *
ctx_or_error (struct stokes *sys)
{
  if (sys == NULL)
    {
      stokes_hip_set_error ("NULL struct stokes");
      return NULL;
    }
  return stokes_hip_ctx (sys);
}

int
stokes_hip_set_recip_g_fix (struct stokes *sys, int enable)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  c->fix_recip_g = enable ? 1 : 0;
  c->valid = 0;
  return 0;
}

int
stokes_hip_set_real_cache (struct stokes *sys, int enable)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  if (enable && !sys->periodic)
    {
      stokes_hip_set_error ("set_real_cache: periodic systems only");
      return STOKES_HIP_EARG;
    }
  int rc = enable ? stokes_hip_prepare (c, c->pos_dev != NULL) : 0;
  if (rc) return rc;
  rc = shk_dev_real_cache (c->dev, enable != 0, c->prec);
  return rc ? dev_fail (rc, "set_real_cache") : 0;
}

int
stokes_hip_set_dense_blocks (struct stokes *sys, int enable)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = shk_dev_set_dense_blocks (c->dev, enable);
  return rc ? dev_fail (rc, "set_dense_blocks") : 0;
}

int
stokes_hip_set_overlap (struct stokes *sys, int enable)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = shk_dev_set_overlap (c->dev, enable);
  return rc ? dev_fail (rc, "set_overlap") : 0;
}

int
stokes_hip_set_split (struct stokes *sys, double tratio)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  if (tratio < 0.0 || tratio != tratio)
    {
      stokes_hip_set_error ("set_split: tratio must be >= 0");
      return STOKES_HIP_EARG;
    }
  c->split_tr = tratio;
  return 0;
}

static double
now_s (void)
{
  struct timespec ts;
  clock_gettime (CLOCK_MONOTONIC, &ts);
  return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

int
stokes_hip_autotune (struct stokes *sys, int nrhs, const double *trs, int ntr,
                     int reps, double *best_tr, double *best_ms)
{
  static const double dflt[] = {0.0, 10.0, 100.0, 300.0, 1000.0, 3000.0,
                                10000.0};
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  if (nrhs < 1 || reps < 1)
    {
      stokes_hip_set_error ("autotune: nrhs and reps must be >= 1");
      return STOKES_HIP_EARG;
    }
  if (trs == NULL || ntr <= 0)
    {
      trs = dflt;
      ntr = (int) (sizeof (dflt) / sizeof (dflt[0]));
    }
  size_t n = (size_t) sys->np * stokes_dim (sys->version) * (size_t) nrhs;
  double *x = NULL, *y = NULL;
  int rc = shk_dev_malloc (c->dev, (void **) &x, sizeof (double) * n);
  if (!rc) rc = shk_dev_malloc (c->dev, (void **) &y, sizeof (double) * n);
  if (!rc) rc = shk_dev_scal (c->dev, x, NULL, 0.0, n);
  if (rc)
    {
      if (x) shk_dev_free (c->dev, x);
      stokes_hip_set_error ("autotune: %s", shk_last_error ());
      return STOKES_HIP_ENOMEM;
    }
  const double keep = c->split_tr;
  double bt = keep, bms = -1.0;
  for (int k = 0; k < ntr && !rc; k++)
    {
      c->split_tr = trs[k];
      rc = stokes_hip_prepare (c, c->pos_dev != NULL);
      if (!rc) rc = shk_dev_atimes (c->dev, x, y, nrhs, c->prec);
      if (!rc) rc = shk_dev_sync (c->dev);
      double t0 = now_s ();
      for (int r = 0; r < reps && !rc; r++)
        rc = shk_dev_atimes (c->dev, x, y, nrhs, c->prec);
      if (!rc) rc = shk_dev_sync (c->dev);
      double ms = (now_s () - t0) * 1e3 / reps;
      if (!rc && (bms < 0.0 || ms < bms))
        {
          bms = ms;
          bt = trs[k];
        }
    }
  shk_dev_free (c->dev, x);
  shk_dev_free (c->dev, y);
  if (rc)
    {
      c->split_tr = keep;
      return dev_fail (rc, "autotune");
    }
  c->split_tr = bt;
  if (best_tr) *best_tr = bt;
  if (best_ms) *best_ms = bms;
  return 0;
}

int
stokes_hip_set_precision (struct stokes *sys, int mode)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  if (mode != STOKES_HIP_FP64 && mode != STOKES_HIP_MIXED)
    {
      stokes_hip_set_error ("invalid precision mode %d", mode);
      return STOKES_HIP_EARG;
    }
  c->prec = mode;
  return 0;
}

int
stokes_hip_set_timing (struct stokes *sys, int enable)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  c->timing = enable ? 1 : 0;
  shk_dev_timing (c->dev, c->timing);
  return 0;
}

int
stokes_hip_malloc (void **p, size_t bytes)
{
  shk_dev *d = stokes_hip_global_dev ();
  if (!d) return STOKES_HIP_ENODEV;
  int rc = shk_dev_malloc (d, p, bytes);
  return rc ? dev_fail (rc, "stokes_hip_malloc") : 0;
}

int
stokes_hip_free (void *p)
{
  shk_dev *d = stokes_hip_global_dev ();
  if (!d) return STOKES_HIP_ENODEV;
  int rc = shk_dev_free (d, p);
  return rc ? dev_fail (rc, "stokes_hip_free") : 0;
}

int
stokes_hip_memcpy_h2d (void *dst, const void *src, size_t bytes)
{
  shk_dev *d = stokes_hip_global_dev ();
  if (!d) return STOKES_HIP_ENODEV;
  /* the copy may overwrite memory that work queued on a system's stream
   * still reads: order it after all queued work */
  int rc = shk_device_sync ();
  if (!rc) rc = shk_dev_h2d (d, dst, src, bytes);
  if (!rc) rc = shk_dev_sync (d);
  return rc ? dev_fail (rc, "stokes_hip_memcpy_h2d") : 0;
}

int
stokes_hip_memcpy_d2h (void *dst, const void *src, size_t bytes)
{
  shk_dev *d = stokes_hip_global_dev ();
  if (!d) return STOKES_HIP_ENODEV;
  /* results may still be in flight on a system's stream */
  int rc = shk_device_sync ();
  if (!rc) rc = shk_dev_d2h (d, dst, src, bytes);
  return rc ? dev_fail (rc, "stokes_hip_memcpy_d2h") : 0;
}

int
stokes_hip_copy_calibrate (void *dst, const void *src, size_t bytes)
{
  shk_dev *d = stokes_hip_global_dev ();
  if (!d) return STOKES_HIP_ENODEV;
  if (bytes % 32)
    {
      stokes_hip_set_error ("copy size must be a multiple of 32 bytes");
      return STOKES_HIP_EARG;
    }
  int rc = shk_dev_copy_calibrate (d, dst, src, bytes);
  return rc ? dev_fail (rc, "stokes_hip_copy_calibrate") : 0;
}

int
stokes_hip_sync (struct stokes *sys)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = shk_dev_sync (c->dev);
  return rc ? dev_fail (rc, "stokes_hip_sync") : 0;
}

int
stokes_hip_set_pos_dev (struct stokes *sys, const double *pos_dev)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  c->pos_dev = pos_dev;
  return 0;
}

int
stokes_hip_use_host_pos (struct stokes *sys)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  c->pos_dev = NULL;
  return 0;
}

int
stokes_hip_atimes_dev (struct stokes *sys, int nrhs, const double *x_dev,
                       double *y_dev)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  rc = shk_dev_atimes (c->dev, x_dev, y_dev, nrhs, c->prec);
  if (rc) return dev_fail (rc, "atimes");
  record_times (c);
  return 0;
}

int
stokes_hip_atimes_dev_noalloc (struct stokes *sys, const double *x_dev,
                               double *y_dev)
{
  stokes_ctx *c = sys ? stokes_hip_ctx_find (sys) : NULL;
  if (!c || !c->valid)
    {
      stokes_hip_set_error ("atimes_dev_noalloc: no device set-up for this"
                            " system yet (one ordinary product, such as"
                            " stokes_hip_atimes_dev, sets it up)");
      return STOKES_HIP_EARG;
    }
  if (sys->a != NULL)
    {
      stokes_hip_set_error ("atimes_dev_noalloc: unequal radii re-upload the"
                            " system's tables on every product; equal"
                            " spheres only");
      return STOKES_HIP_EUNSUP;
    }
  shk_params pu;
  if (prepare_params (c, &pu))
    {
      stokes_hip_set_error ("atimes_dev_noalloc: the system's parameters"
                            " changed since its device set-up (one ordinary"
                            " product sets it up again)");
      return STOKES_HIP_EARG;
    }
  int rc = shk_dev_noalloc_begin (c->dev);
  if (rc) return dev_fail (rc, "atimes_dev_noalloc");
  /* positions only: the tables are current */
  rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (!rc)
    {
      int r2 = shk_dev_atimes (c->dev, x_dev, y_dev, 1, c->prec);
      if (r2) rc = dev_fail (r2, "atimes_dev_noalloc");
    }
  shk_dev_noalloc_end (c->dev);
  if (rc) return rc;
  record_times (c);
  return 0;
}

int
stokes_hip_atimes (struct stokes *sys, int n, const double *x, double *y)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int nn = sys->np * stokes_dim (sys->version);
  if (n != nn)
    {
      stokes_hip_set_error ("n = %d but np * %d = %d", n,
                            stokes_dim (sys->version), nn);
      return STOKES_HIP_EARG;
    }
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  size_t bytes = sizeof (double) * (size_t) n;
  if ((rc = ensure_xy (c, bytes))) return rc;
  rc = shk_dev_h2d (c->dev, c->x_dev, x, bytes);
  if (!rc) rc = shk_dev_atimes (c->dev, c->x_dev, c->y_dev, 1, c->prec);
  if (!rc) rc = shk_dev_d2h (c->dev, y, c->y_dev, bytes);
  if (rc) return dev_fail (rc, "atimes");
  record_times (c);
  return 0;
}

static int
ensure_mat (stokes_ctx *c, size_t bytes)
{
  if (c->mat_bytes >= bytes) return 0;
  if (c->mat_dev) shk_dev_free (c->dev, c->mat_dev);
  c->mat_dev = NULL;
  c->mat_bytes = 0;
  int rc = shk_dev_malloc (c->dev, (void **) &c->mat_dev, bytes);
  if (rc) return dev_fail (rc, "matrix buffer");
  c->mat_bytes = bytes;
  return 0;
}

int
stokes_hip_make_matrix_dev (struct stokes *sys, double *mat_dev)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  rc = shk_dev_matrix (c->dev, mat_dev);
  if (rc) return dev_fail (rc, "make_matrix");
  record_times (c);
  return 0;
}

int
stokes_hip_extmat_dev (struct stokes *sys, const double *mat_dev, int nrhs,
                       const double *x_dev, double *y_dev)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  rc = shk_dev_extmat (c->dev, mat_dev, x_dev, y_dev, nrhs);
  return rc ? dev_fail (rc, "extmat") : 0;
}

/* ------------------------------------------------------------------ */
/* reference API                                                        */
/* ------------------------------------------------------------------ */

void
atimes_ewald_3all (int n, const double *x, double *y, void *user_data)
{
  struct stokes *sys = (struct stokes *) user_data;
  if (!sys->periodic)
    {
      stokes_hip_set_error ("atimes_ewald_3all on a non-periodic system");
      die_on (STOKES_HIP_EARG, "atimes_ewald_3all");
    }
  die_on (stokes_hip_atimes (sys, n, x, y), "atimes_ewald_3all");
}

void
atimes_nonewald_3all (int n, const double *x, double *y, void *user_data)
{
  struct stokes *sys = (struct stokes *) user_data;
  if (sys->periodic)
    {
      stokes_hip_set_error ("atimes_nonewald_3all on a periodic system");
      die_on (STOKES_HIP_EARG, "atimes_nonewald_3all");
    }
  die_on (stokes_hip_atimes (sys, n, x, y), "atimes_nonewald_3all");
}

void
atimes_3all (int n, const double *x, double *y, void *user_data)
{
  die_on (stokes_hip_atimes ((struct stokes *) user_data, n, x, y),
          "atimes_3all");
}

int
stokes_hip_make_matrix (struct stokes *sys, double *mat)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  size_t n = (size_t) sys->np * stokes_dim (sys->version);
  int rc = ensure_mat (c, sizeof (double) * n * n);
  if (rc) return rc;
  if ((rc = stokes_hip_make_matrix_dev (sys, c->mat_dev))) return rc;
  rc = shk_dev_d2h (c->dev, mat, c->mat_dev, sizeof (double) * n * n);
  return rc ? dev_fail (rc, "make_matrix") : 0;
}

void
make_matrix_mob_3all (struct stokes *sys, double *mat)
{
  die_on (stokes_hip_make_matrix (sys, mat), "make_matrix_mob_3all");
}

void
make_matrix_mob_ewald_3all (struct stokes *sys, double *mat)
{
  if (!sys->periodic)
    {
      stokes_hip_set_error ("non-periodic system");
      die_on (STOKES_HIP_EARG, "make_matrix_mob_ewald_3all");
    }
  die_on (stokes_hip_make_matrix (sys, mat), "make_matrix_mob_ewald_3all");
}

void
make_matrix_mob_nonewald_3all (struct stokes *sys, double *mat)
{
  if (sys->periodic)
    {
      stokes_hip_set_error ("periodic system");
      die_on (STOKES_HIP_EARG, "make_matrix_mob_nonewald_3all");
    }
  die_on (stokes_hip_make_matrix (sys, mat), "make_matrix_mob_nonewald_3all");
}

/* ewald.c:484-505: build the matrix and multiply, all on the device */
static int
atimes_matrix_impl (struct stokes *sys, int n, const double *x, double *y)
{
  stokes_ctx *c = ctx_or_error (sys);
  if (!c) return STOKES_HIP_ENODEV;
  size_t nn = (size_t) sys->np * stokes_dim (sys->version);
  if ((size_t) n != nn)
    {
      stokes_hip_set_error ("n = %d does not match the system", n);
      return STOKES_HIP_EARG;
    }
  int rc = ensure_mat (c, sizeof (double) * nn * nn);
  if (!rc) rc = ensure_xy (c, sizeof (double) * nn);
  if (rc) return rc;
  if ((rc = stokes_hip_make_matrix_dev (sys, c->mat_dev))) return rc;
  rc = shk_dev_h2d (c->dev, c->x_dev, x, sizeof (double) * nn);
  if (!rc) rc = shk_dev_extmat (c->dev, c->mat_dev, c->x_dev, c->y_dev, 1);
  if (!rc) rc = shk_dev_d2h (c->dev, y, c->y_dev, sizeof (double) * nn);
  return rc ? dev_fail (rc, "atimes_3all_matrix") : 0;
}

void
atimes_3all_matrix (int n, const double *x, double *y, void *user_data)
{
  die_on (atimes_matrix_impl ((struct stokes *) user_data, n, x, y),
          "atimes_3all_matrix");
}

/* free-standing dense products on the library's global context */
static int
dense_product (int extracted, const double *mat, int n1, int n2,
               const double *x, double *y)
{
  shk_dev *d = stokes_hip_global_dev ();
  if (!d) return STOKES_HIP_ENODEV;
  size_t mb = sizeof (double) * (size_t) n1 * (size_t) n2;
  double *md = NULL, *xd = NULL, *yd = NULL;
  int rc = shk_dev_malloc (d, (void **) &md, mb);
  if (!rc) rc = shk_dev_malloc (d, (void **) &xd, sizeof (double) * n2);
  if (!rc) rc = shk_dev_malloc (d, (void **) &yd, sizeof (double) * n1);
  if (!rc) rc = shk_dev_h2d (d, md, mat, mb);
  if (!rc) rc = shk_dev_h2d (d, xd, x, sizeof (double) * n2);
  if (!rc) rc = shk_dev_gemv (d, md, n1, n2, xd, extracted, yd);
  if (!rc) rc = shk_dev_d2h (d, y, yd, sizeof (double) * n1);
  if (md) shk_dev_free (d, md);
  if (xd) shk_dev_free (d, xd);
  if (yd) shk_dev_free (d, yd);
  return rc ? dev_fail (rc, "dense product") : 0;
}

void
dot_prod_matrix (const double *mat, int n1, int n2, const double *x,
                 double *y)
{
  die_on (dense_product (0, mat, n1, n2, x, y), "dot_prod_matrix");
}

void
multiply_extmat_with_extvec_3fts (int np, const double *m, const double *x,
                                  double *y)
{
  die_on (dense_product (1, m, 11 * np, 11 * np, x, y),
          "multiply_extmat_with_extvec_3fts");
}
