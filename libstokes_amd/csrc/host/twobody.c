/* twobody.c -- two-body resistance coefficients f_k(lambda) and the
 * reference's twobody / lubrication-condition API (twobody.h, lub.h).
 *
 * The reference ships the coefficients as generated code (XA.c ... ZM.c:
 * polynomials in lambda to order ~150).  They are computed here instead,
 * once per process, from the published twin-multipole recurrences of
 * Jeffrey & Onishi (1984, J. Fluid Mech. 139, 261-290) for A, B, C, G, H
 * and Jeffrey (1992, Phys. Fluids A 4, 16-29) for M: coefficients P_npq,
 * V_npq, Q_npq of the expansion in the two radius ratios, azimuthal order m
 * = 0 (axial), 1 (transverse) or 2, started from the translation (P = V =
 * delta_{n1}), rotation (Q = delta_{n1}) or straining (P = V = delta_{n2})
 * harmonic.  f_k(lambda) = c 2^k sum_q X_{n (k - q) q} lambda^(q + o k%2)
 * with X, n, c and the parity offset o per family (table below).  The
 * recurrences run in long double; the stored coefficients are doubles.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "twobody.h"
#include "lub.h"
#include "stokes_hip.h"
#include "stokes_internal.h"

/* families in the order of struct twobody_f */
enum { TWOBODY_XA, TWOBODY_YA, TWOBODY_YB, TWOBODY_XC, TWOBODY_YC, TWOBODY_XG,
       TWOBODY_YG, TWOBODY_YH, TWOBODY_XM, TWOBODY_YM, TWOBODY_ZM };

static pthread_mutex_t tb_lock = PTHREAD_MUTEX_INITIALIZER;
static int tb_N = -1;
static double *tb_c[11];        /* [k (N + 2) + q] */

typedef struct {
  int N, m;
  long double *P, *V, *Q;       /* [(n (N + 1) + p) (N + 1) + q] */
} tb_run;

static size_t
tix (int N, int n, int p, int q)
{
  return ((size_t) n * (size_t) (N + 1) + (size_t) p) * (size_t) (N + 1) + (size_t) q;
}

/* one recurrence run: kind 0 starts P = V = 1 at harmonic n0, kind 1 Q = 1 */
static int
tb_run_exec (tb_run *r, int N, int m, int n0, int kind)
{
  const size_t sz = (size_t) (N + 1) * (N + 1) * (N + 1);
  r->N = N;
  r->m = m;
  r->P = calloc (sz, sizeof (long double));
  r->V = calloc (sz, sizeof (long double));
  r->Q = calloc (sz, sizeof (long double));
  /* per (n, s): binomial C(n + s, n + m) and the four coupling factors */
  const size_t ns = (size_t) (N + 1) * (N + 1);
  long double *fb = calloc (ns, sizeof (long double));
  long double *fA = calloc (ns, sizeof (long double));
  long double *fV = calloc (ns, sizeof (long double));
  long double *fQ = calloc (ns, sizeof (long double));   /* Q <- P in the Q equation */
  long double *fS = calloc (ns, sizeof (long double));   /* Q <- Q */
  if (!r->P || !r->V || !r->Q || !fb || !fA || !fV || !fQ || !fS)
    {
      free (r->P); free (r->V); free (r->Q);
      free (fb); free (fA); free (fV); free (fQ); free (fS);
      return -1;
    }
  const long double mm = (long double) m * (long double) m;
  for (int n = 1; n <= N; n++)
    for (int s = 1; s <= N; s++)
      {
        const long double ln = n, ls = s;
        /* C(n + s, n + m) by the product formula */
        long double b = 1.0L;
        const int top = n + s, k = n + m;
        if (k > top) b = 0.0L;
        else
          for (int i = 1; i <= top - k; i++) b = b * (long double) (k + i) / (long double) i;
        const size_t o = (size_t) n * (N + 1) + s;
        fb[o] = b;
        if (m == 0)
          {
            fA[o] = ln * (2 * ln + 1) * (2 * ln * ls - ln - ls + 2)
              / (2 * (ln + 1) * (2 * ls - 1) * (ln + ls));
            fV[o] = -ln * (4 * ln * ln - 1) / (2 * (ln + 1) * (2 * ls + 1));
            fQ[o] = 0.0L;
          }
        else
          {
            fA[o] = -(2 * ln + 1) * ((ln * ls + mm) * (2 * ln * ls - ls - ln + 2)
                                     - 3 * mm * (ln + ls))
              / (2 * (ln + 1) * ls * (ln + ls) * (2 * ls - 1));
            fV[o] = ln * (4 * ln * ln - 1) / (2 * (ln + 1) * (2 * ls + 1));
            fQ[o] = -3.0L / (2 * ln * ls * (ln + 1));
          }
        fS[o] = ls / (ln + 1);
      }
  if (kind == 0)
    {
      r->P[tix (N, n0, 0, 0)] = 1.0L;
      r->V[tix (N, n0, 0, 0)] = 1.0L;
    }
  else
    r->Q[tix (N, n0, 0, 0)] = 1.0L;
  const int lo = m > 1 ? m : 1;
  for (int t = 1; t <= N; t++)
    for (int p = 0; p <= t; p++)
      {
        const int q = t - p;
        /* terms read the index p - n + 1 or p - n: n <= p + 1 */
        for (int n = lo; n <= N && n <= p + 1; n++)
          {
            const long double ln = n;
            const long double fB = (m == 0 ? -1.0L : 1.0L) * ln * (2 * ln - 1) / (2 * (ln + 1));
            const long double fQP = -mm * 2 * (4 * ln * ln - 1) / (3 * (ln + 1));
            long double P = 0.0L, Vs = 0.0L, Q = 0.0L;
            for (int s = lo; s <= q; s++)
              {
                const size_t o = (size_t) n * (N + 1) + s;
                const long double b = fb[o];
                const int p1 = p - n + 1, p0 = p - n, pm = p - n - 1;
                long double acc = 0.0L;
                if (p1 >= 0)
                  {
                    acc += fA[o] * r->P[tix (N, s, q - s, p1)];
                    if (q - s - 2 >= 0) acc += fV[o] * r->V[tix (N, s, q - s - 2, p1)];
                    if (q - s - 1 >= 0) acc += fQP * r->Q[tix (N, s, q - s - 1, p1)];
                  }
                if (pm >= 0)
                  {
                    const long double pp = r->P[tix (N, s, q - s, pm)];
                    acc += fB * pp;
                    Vs += b * pp;
                  }
                P += b * acc;
                if (p0 >= 0)
                  {
                    long double qq = fQ[o] * r->P[tix (N, s, q - s, p0)];
                    if (q - s - 1 >= 0) qq += fS[o] * r->Q[tix (N, s, q - s - 1, p0)];
                    Q += b * qq;
                  }
              }
            const long double vc = (m == 0 ? -2.0L : 2.0L) * ln / ((ln + 1) * (2 * ln + 3));
            r->P[tix (N, n, p, q)] = P;
            r->V[tix (N, n, p, q)] = P + vc * Vs;
            r->Q[tix (N, n, p, q)] = Q;
          }
      }
  free (fb); free (fA); free (fV); free (fQ); free (fS);
  return 0;
}

static void
tb_run_free (tb_run *r)
{
  free (r->P); free (r->V); free (r->Q);
}

/* c 2^(k + e) sum_q X_{n (k - q) q} lambda^(q + (odd ? k%2 : 0)) into c[fam] */
static void
tb_collect (int fam, const tb_run *r, const long double *X, int n, long double c, int e,
            int odd)
{
  const int N = r->N;
  for (int k = 0; k <= N; k++)
    {
      const long double w = ldexpl (c, k + e);
      for (int q = 0; q <= k; q++)
        tb_c[fam][(size_t) k * (N + 2) + q + (odd ? (k & 1) : 0)]
          += (double) (w * X[tix (N, n, k - q, q)]);
    }
}

static int
tb_build (int N)
{
  for (int f = 0; f < 11; f++)
    {
      free (tb_c[f]);
      tb_c[f] = calloc ((size_t) (N + 1) * (N + 2), sizeof (double));
      if (!tb_c[f]) return -1;
    }
  /* (m, start harmonic, kind) and the families read from each run */
  tb_run r;
  if (tb_run_exec (&r, N, 0, 1, 0)) return -1;
  tb_collect (TWOBODY_XA, &r, r.P, 1, 1.0L, 0, 0);
  tb_collect (TWOBODY_XG, &r, r.P, 2, 0.75L, 0, 0);
  tb_run_free (&r);
  if (tb_run_exec (&r, N, 0, 1, 1)) return -1;
  tb_collect (TWOBODY_XC, &r, r.Q, 1, 1.0L, 0, 1);
  tb_run_free (&r);
  if (tb_run_exec (&r, N, 0, 2, 0)) return -1;
  tb_collect (TWOBODY_XM, &r, r.P, 2, 1.0L, 0, 1);
  tb_run_free (&r);
  if (tb_run_exec (&r, N, 1, 1, 0)) return -1;
  tb_collect (TWOBODY_YA, &r, r.P, 1, 1.0L, 0, 0);
  tb_collect (TWOBODY_YB, &r, r.Q, 1, 1.0L, 1, 0);
  tb_collect (TWOBODY_YG, &r, r.P, 2, 0.75L, 0, 0);
  tb_run_free (&r);
  if (tb_run_exec (&r, N, 1, 1, 1)) return -1;
  tb_collect (TWOBODY_YC, &r, r.Q, 1, 1.0L, 0, 1);
  tb_collect (TWOBODY_YH, &r, r.P, 2, -0.375L, 0, 1);
  tb_run_free (&r);
  if (tb_run_exec (&r, N, 1, 2, 0)) return -1;
  tb_collect (TWOBODY_YM, &r, r.P, 2, 1.0L, 0, 1);
  tb_run_free (&r);
  if (tb_run_exec (&r, N, 2, 2, 0)) return -1;
  tb_collect (TWOBODY_ZM, &r, r.P, 2, 1.0L, 0, 1);
  tb_run_free (&r);
  return 0;
}

int
stokes_hip_twobody_coef (int nmax, const double **tab, int *N)
{
  if (nmax < 0)
    {
      stokes_hip_set_error ("twobody: negative order %d", nmax);
      return STOKES_HIP_EARG;
    }
  pthread_mutex_lock (&tb_lock);
  int rc = 0;
  if (tb_N < nmax)
    {
      const int want = nmax > 100 ? nmax : 100;
      if (tb_build (want))
        {
          tb_N = -1;
          stokes_hip_set_error ("twobody: allocation failed");
          rc = STOKES_HIP_ENOMEM;
        }
      else
        tb_N = want;
    }
  if (!rc)
    {
      if (tab)
        for (int f = 0; f < 11; f++) tab[f] = tb_c[f];
      if (N) *N = tb_N;
    }
  pthread_mutex_unlock (&tb_lock);
  return rc;
}

/* f[0..n] of one family at lambda (Horner over the stored polynomials) */
static void
tb_eval (int fam, int n, double l, double *f)
{
  const double *tab[11];
  int N;
  if (stokes_hip_twobody_coef (n, tab, &N))
    {
      fprintf (stderr, "twobody: %s\n", stokes_hip_last_error ());
      exit (1);
    }
  for (int k = 0; k <= n; k++)
    {
      const double *c = tab[fam] + (size_t) k * (N + 2);
      double v = 0.0;
      for (int q = k + 1; q >= 0; q--) v = v * l + c[q];
      f[k] = v;
    }
}

void twobody_XA (int n, double l, double *f) { tb_eval (TWOBODY_XA, n, l, f); }
void twobody_YA (int n, double l, double *f) { tb_eval (TWOBODY_YA, n, l, f); }
void twobody_YB (int n, double l, double *f) { tb_eval (TWOBODY_YB, n, l, f); }
void twobody_XC (int n, double l, double *f) { tb_eval (TWOBODY_XC, n, l, f); }
void twobody_YC (int n, double l, double *f) { tb_eval (TWOBODY_YC, n, l, f); }
void twobody_XG (int n, double l, double *f) { tb_eval (TWOBODY_XG, n, l, f); }
void twobody_YG (int n, double l, double *f) { tb_eval (TWOBODY_YG, n, l, f); }
void twobody_YH (int n, double l, double *f) { tb_eval (TWOBODY_YH, n, l, f); }
void twobody_XM (int n, double l, double *f) { tb_eval (TWOBODY_XM, n, l, f); }
void twobody_YM (int n, double l, double *f) { tb_eval (TWOBODY_YM, n, l, f); }
void twobody_ZM (int n, double l, double *f) { tb_eval (TWOBODY_ZM, n, l, f); }

static void *
tb_alloc (size_t bytes)
{
  void *p = malloc (bytes);
  if (!p)
    {
      fprintf (stderr, "twobody: allocation failed\n");
      exit (1);
    }
  return p;
}

struct twobody_f *
twobody_f_init (int nmax, double lambda)
{
  struct twobody_f *f = tb_alloc (sizeof (*f));
  f->nmax = nmax;
  f->lambda = lambda;
  double **dst[11] = {&f->XA, &f->YA, &f->YB, &f->XC, &f->YC, &f->XG, &f->YG, &f->YH,
                      &f->XM, &f->YM, &f->ZM};
  for (int k = 0; k < 11; k++)
    {
      *dst[k] = tb_alloc (sizeof (double) * (size_t) (nmax + 1));
      tb_eval (k, nmax, lambda, *dst[k]);
    }
  return f;
}

void
twobody_f_free (struct twobody_f *f)
{
  if (!f) return;
  free (f->XA); free (f->YA); free (f->YB); free (f->XC); free (f->YC); free (f->XG);
  free (f->YG); free (f->YH); free (f->XM); free (f->YM); free (f->ZM);
  free (f);
}

struct twobody_f_list *
twobody_f_list_init (void)
{
  struct twobody_f_list *l = tb_alloc (sizeof (*l));
  l->n = 0;
  l->l = NULL;
  l->f = NULL;
  return l;
}

void
twobody_f_list_append (struct twobody_f_list *list, int nmax, double lambda)
{
  const int n = list->n + 1;
  double *ll = realloc (list->l, sizeof (double) * (size_t) n);
  struct twobody_f **ff = realloc (list->f, sizeof (struct twobody_f *) * (size_t) n);
  if (!ll || !ff)
    {
      fprintf (stderr, "twobody_f_list_append: allocation failed\n");
      exit (1);
    }
  list->l = ll;
  list->f = ff;
  list->l[n - 1] = lambda;
  list->f[n - 1] = twobody_f_init (nmax, lambda);
  list->n = n;
}

void
twobody_f_list_free (struct twobody_f_list *list)
{
  if (!list) return;
  for (int i = 0; i < list->n; i++) twobody_f_free (list->f[i]);
  free (list->f);
  free (list->l);
  free (list);
}

/* cond_lub (lub.c:42-73): 0 when 0 < r and (lubmax2 <= 0 or r^2 < lubmax2) */
int
cond_lub (const double *x1, const double *x2, double lubmax2)
{
  const double dx = x1[0] - x2[0], dy = x1[1] - x2[1], dz = x1[2] - x2[2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 == 0.0) return 1;
  if (lubmax2 <= 0.0) return 0;
  return r2 < lubmax2 ? 0 : 1;
}

/* cond_lub_poly (lub.c:82-127): the same on s^2 = 4 r^2 / (a1 + a2)^2 */
int
cond_lub_poly (const double *x1, const double *x2, double a1, double a2, double lubmax2)
{
  const double dx = x1[0] - x2[0], dy = x1[1] - x2[1], dz = x1[2] - x2[2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 == 0.0) return 1;
  if (lubmax2 <= 0.0) return 0;
  return r2 * 4.0 / (a1 + a2) / (a1 + a2) < lubmax2 ? 0 : 1;
}
