/* orc_lub.c -- FP64 CPU restatement of libstokes' lubrication correction.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP library; see oracle.h).
 *
 * Restates, per the reference file:line cited at each function:
 *  - the two-body coefficients f_k(lambda) of twobody_XA ... twobody_ZM
 *    (XA.c ... ZM.c: generated tables) from the published recurrences of
 *    Jeffrey & Onishi (1984, J. Fluid Mech. 139) and Jeffrey (1992, Phys.
 *    Fluids A 4), evaluated here in long double;
 *  - the far and lubrication forms of the 22 resistance scalars and their
 *    scalings (twobody.c:644-2814), and the equal-sphere series
 *    scalar_two_body_res (two-body-res.c:31);
 *  - the inverse of the two-body far-field (RPY) mobility, mode by mode
 *    (minv-poly.c:267-560, fts.c:1490);
 *  - the pair loop calc_lub_3fts (lub.c:576-768) with calc_lub_fts_2b
 *    (fts.c:1758-1869) and calc_lub_fts_2b_poly (fts.c:1991-2147).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ------------------------------------------------------------------ */
/* coefficient recurrences                                              */
/* ------------------------------------------------------------------ */

/* One run of the twin-multipole recurrences for azimuthal order m, the
 * coefficients P_npq, V_npq, Q_npq of the expansion in the two radius
 * ratios (p, q powers), started either from the force / strain harmonic n0
 * (P = V = delta_{n n0}: kind 0) or from the rotlet (Q = delta_{n 1}: kind
 * 1).  N is the largest p + q kept. */
typedef struct {
  int N;
  long double *P, *V, *Q;
} orc_rec;

static size_t
rix (int N, int n, int p, int q)
{
  return ((size_t) n * (N + 1) + (size_t) p) * (N + 1) + (size_t) q;
}

static long double
rget (const orc_rec *r, const long double *a, int n, int p, int q)
{
  if (n < 0 || n > r->N || p < 0 || q < 0 || p > r->N || q > r->N) return 0.0L;
  return a[rix (r->N, n, p, q)];
}

/* binomial coefficients C(n, k), n <= 2 N + 2, from Pascal's rule */
static long double *bt;
static int bt_n = -1;

static void
binom_init (int nmax)
{
  if (bt_n >= nmax) return;
  free (bt);
  bt = calloc ((size_t) (nmax + 1) * (nmax + 1), sizeof (long double));
  for (int n = 0; n <= nmax; n++)
    {
      bt[(size_t) n * (nmax + 1)] = 1.0L;
      for (int k = 1; k <= n; k++)
        bt[(size_t) n * (nmax + 1) + k] = bt[(size_t) (n - 1) * (nmax + 1) + k - 1]
          + (k <= n - 1 ? bt[(size_t) (n - 1) * (nmax + 1) + k] : 0.0L);
    }
  bt_n = nmax;
}

static long double
binom (int n, int k)
{
  if (k < 0 || k > n) return 0.0L;
  return bt[(size_t) n * (bt_n + 1) + k];
}

static void
rec_run (orc_rec *r, int N, int m, int n0, int kind)
{
  size_t sz = (size_t) (N + 1) * (N + 1) * (N + 1);
  r->N = N;
  r->P = calloc (sz, sizeof (long double));
  r->V = calloc (sz, sizeof (long double));
  r->Q = calloc (sz, sizeof (long double));
  if (kind == 0)
    {
      r->P[rix (N, n0, 0, 0)] = 1.0L;
      r->V[rix (N, n0, 0, 0)] = 1.0L;
    }
  else
    r->Q[rix (N, n0, 0, 0)] = 1.0L;
  const int nlo = m > 1 ? m : 1, slo = m > 1 ? m : 1;
  const long double mm = (long double) m * m;
  binom_init (2 * N + 4);
  /* every term of P, V and Q at (n, p, q) reads index p - n + 1 or p - n
   * of a coefficient: zero for n > p + 1 (only the start values live
   * there) */
  for (int t = 1; t <= N; t++)
    for (int p = 0; p <= t; p++)
      for (int n = nlo; n <= N && n <= p + 1; n++)
        {
          const int q = t - p;
          long double P = 0.0L, Vs = 0.0L, Q = 0.0L;
          const long double ln = n;
          for (int s = slo; s <= q; s++)
            {
              const long double ls = s;
              const long double b = binom (n + s, n + m);
              long double tA, tB, tV, tQ = 0.0L;
              if (m == 0)
                {
                  tA = ln * (2 * ln + 1) * (2 * ln * ls - ln - ls + 2)
                    / (2 * (ln + 1) * (2 * ls - 1) * (ln + ls));
                  tB = -ln * (2 * ln - 1) / (2 * (ln + 1));
                  tV = -ln * (4 * ln * ln - 1) / (2 * (ln + 1) * (2 * ls + 1));
                }
              else
                {
                  tA = -(2 * ln + 1) * ((ln * ls + mm) * (2 * ln * ls - ls - ln + 2)
                                        - 3 * mm * (ln + ls))
                    / (2 * (ln + 1) * ls * (ln + ls) * (2 * ls - 1));
                  tB = ln * (2 * ln - 1) / (2 * (ln + 1));
                  tV = ln * (4 * ln * ln - 1) / (2 * (ln + 1) * (2 * ls + 1));
                  tQ = -mm * 2 * (4 * ln * ln - 1) / (3 * (ln + 1));
                }
              P += b * (tA * rget (r, r->P, s, q - s, p - n + 1)
                        + tB * rget (r, r->P, s, q - s, p - n - 1)
                        + tV * rget (r, r->V, s, q - s - 2, p - n + 1)
                        + tQ * rget (r, r->Q, s, q - s - 1, p - n + 1));
              Vs += b * rget (r, r->P, s, q - s, p - n - 1);
              Q += b * (ls / (ln + 1) * rget (r, r->Q, s, q - s - 1, p - n)
                        - (m ? 3.0L / (2 * ln * ls * (ln + 1)) : 0.0L)
                          * rget (r, r->P, s, q - s, p - n));
            }
          const long double vc = (m == 0 ? -2.0L : 2.0L) * ln / ((ln + 1) * (2 * ln + 3));
          r->P[rix (N, n, p, q)] = P;
          r->V[rix (N, n, p, q)] = P + vc * Vs;
          r->Q[rix (N, n, p, q)] = Q;
        }
}

static void
rec_free (orc_rec *r)
{
  free (r->P); free (r->V); free (r->Q);
}

/* coefficient of lambda^q in f_k of family f: tb[f][k * (N + 2) + q] */
static int tb_N = -1;
static double *tb[11];

/* c * 2^(k + e2) * sum_q A(n, k - q, q) lambda^(q + shift k%2) */
static void
tb_store (int f, const orc_rec *r, const long double *A, int n, long double c,
          int e2, int shift)
{
  const int N = r->N;
  double *t = tb[f];
  for (int k = 0; k <= N; k++)
    {
      long double pw = ldexpl (c, k + e2);
      for (int q = 0; q <= k; q++)
        t[(size_t) k * (N + 2) + q + (shift ? k % 2 : 0)]
          += (double) (pw * rget (r, A, n, k - q, q));
    }
}

/* families in the order of struct twobody_f (twobody.h): XA YA YB XC YC
 * XG YG YH XM YM ZM */
int
orc_twobody_init (int N)
{
  if (tb_N >= N) return 0;
  for (int f = 0; f < 11; f++)
    {
      free (tb[f]);
      tb[f] = calloc ((size_t) (N + 1) * (N + 2), sizeof (double));
      if (!tb[f]) return -1;
    }
  orc_rec r;
  rec_run (&r, N, 0, 1, 0);                 /* axisymmetric translation */
  tb_store (0, &r, r.P, 1, 1.0L, 0, 0);     /* XA */
  tb_store (5, &r, r.P, 2, 0.75L, 0, 0);    /* XG */
  rec_free (&r);
  rec_run (&r, N, 0, 1, 1);                 /* axisymmetric rotation */
  tb_store (3, &r, r.Q, 1, 1.0L, 0, 1);     /* XC */
  rec_free (&r);
  rec_run (&r, N, 0, 2, 0);                 /* axisymmetric straining */
  tb_store (8, &r, r.P, 2, 1.0L, 0, 1);     /* XM */
  rec_free (&r);
  rec_run (&r, N, 1, 1, 0);                 /* transverse translation */
  tb_store (1, &r, r.P, 1, 1.0L, 0, 0);     /* YA */
  tb_store (2, &r, r.Q, 1, 1.0L, 1, 0);     /* YB */
  tb_store (6, &r, r.P, 2, 0.75L, 0, 0);    /* YG */
  rec_free (&r);
  rec_run (&r, N, 1, 1, 1);                 /* transverse rotation */
  tb_store (4, &r, r.Q, 1, 1.0L, 0, 1);     /* YC */
  tb_store (7, &r, r.P, 2, -0.375L, 0, 1);  /* YH */
  rec_free (&r);
  rec_run (&r, N, 1, 2, 0);                 /* transverse straining */
  tb_store (9, &r, r.P, 2, 1.0L, 0, 1);     /* YM */
  rec_free (&r);
  rec_run (&r, N, 2, 2, 0);                 /* m = 2 straining */
  tb_store (10, &r, r.P, 2, 1.0L, 0, 1);    /* ZM */
  rec_free (&r);
  tb_N = N;
  return 0;
}

/* f[0..n] of family fam at lambda (twobody_XA etc., XA.c ...) */
void
orc_twobody_table (int fam, int n, double l, double *f)
{
  if (orc_twobody_init (n > 100 ? n : 100)) return;
  for (int k = 0; k <= n; k++)
    {
      const double *c = tb[fam] + (size_t) k * (tb_N + 2);
      double v = 0.0;
      for (int q = k + 1; q >= 0; q--) v = v * l + c[q];
      f[k] = v;
    }
}

/* ------------------------------------------------------------------ */
/* far and lubrication forms                                            */
/* ------------------------------------------------------------------ */

/* twobody_far_with_f (twobody.c:689-998): the 1/s series, even terms for
 * one of (11, 12) and odd for the other, times the 12 prefactor */
static void
far_pair (const double *f, int n, double s1l, int odd11, double pre,
          double *v11, double *v12)
{
  double a = 0.0, b = 0.0, se = 1.0, so = s1l;
  const double s2 = s1l * s1l;
  for (int i = 0; i < n / 2; i++)
    {
      if (odd11) { a += f[2 * i + 1] / so; b += f[2 * i] / se; }
      else { a += f[2 * i] / se; b += f[2 * i + 1] / so; }
      se *= s2;
      so *= s2;
    }
  *v11 = a;
  *v12 = b * pre;
}

void
orc_twobody_far (int version, const double *f, int n, double l, double s,
                 double *out)
{
  if (n % 2) n--;
  const int N1 = n + 1;
  const double l1 = 1.0 + l, l12 = l1 * l1, l13 = l12 * l1, s1l = s * l1;
  /* family, odd for 11, 12 prefactor */
  static const int odd11[11] = {0, 0, 1, 0, 0, 1, 1, 0, 0, 0, 0};
  const double pre[11] = {-2.0 / l1, -2.0 / l1, -4.0 / l12, -8.0 / l13,
                          8.0 / l13, -4.0 / l12, -4.0 / l12, 8.0 / l13,
                          8.0 / l13, 8.0 / l13, -8.0 / l13};
  const int nf = version == 0 ? 2 : (version == 1 ? 5 : 11);
  for (int k = 0; k < nf; k++)
    far_pair (f + (size_t) k * N1, n, s1l, odd11[k], pre[k], out + 2 * k,
              out + 2 * k + 1);
}

/* the subtracted series term of the lubrication form: for m of the right
 * parity add f_m / (s(1+l))^m + a_m (2/s)^m to *acc */
typedef struct { double c1, c2, c3; int plus2; } lub_sub;

static double
lub_a (const lub_sub *u, int m)
{
  /* a_m = -c1 - 2 c2 / m + 4 c3 / d_m, d_m = m (m - 2) (m + 2 for the
   * G, H, M forms); at m = 2 of the m (m - 2) form the reference uses
   * -2 m in place of the vanishing product (twobody.c:1055-1062) */
  double d = u->plus2 ? (double) (m * (m + 2))
    : (m == 2 ? (double) (-2 * m) : (double) (m * (m - 2)));
  return -u->c1 - 2.0 * u->c2 / m + 4.0 * u->c3 / d;
}

/* sum over m = mlo .. mmax of one parity (odd: 1) */
static double
lub_series (const double *f, int mlo, int mmax, int odd, double s, double l,
            const lub_sub *u, double sign_f)
{
  const double s1l = s * (1.0 + l), s21 = 2.0 / s;
  double tm = 1.0, t2 = 1.0, acc = 0.0;
  for (int m = 1; m <= mmax; m++)
    {
      tm *= s1l;
      t2 *= s21;
      if (m < mlo || (m % 2) != odd) continue;
      acc += sign_f * (f[m] / tm + lub_a (u, m) * t2);
    }
  return acc;
}

/* twobody_lub_with_f (twobody.c:1977-2644): the near-contact forms of
 * Jeffrey & Onishi (1984) and Jeffrey (1992) plus the remaining series,
 * terms m = 1 .. mmax (the reference takes m < n, n even).  yc12_g5: the
 * factor of the g5 coefficient of YC12 (1 as twobody.c has it; 2 in the
 * equal-sphere series of two-body-res.c) */
void
orc_twobody_lub (int version, const double *f, int N1, int mmax, double l,
                 double s, double yc12_g5, double *out)
{
  const double l1 = 1.0 + l, l12 = l1 * l1, l13 = l12 * l1, l14 = l12 * l12;
  const double S1 = 1.0 - 4.0 / (s * s);
  const double ls1 = log (S1), ls2 = log ((s + 2.0) / (s - 2.0));
  const double *fXA = f, *fYA = f + N1, *fYB = f + 2 * N1, *fXC = f + 3 * N1;
  const double *fYC = f + 4 * N1, *fXG = f + 5 * N1, *fYG = f + 6 * N1;
  const double *fYH = f + 7 * N1, *fXM = f + 8 * N1, *fYM = f + 9 * N1;
  const double *fZM = f + 10 * N1;
  double g1, g2, g3, g4, g5, g6;
  lub_sub u;

  /* XA */
  g1 = 2.0 * l * l / l13;
  g2 = 0.2 * l * (1.0 + 7.0 * l + l * l) / l13;
  g3 = (1.0 + 18.0 * l - 29.0 * l * l + 18.0 * l * l * l + l * l * l * l) / (42.0 * l13);
  u = (lub_sub) {g1, g2, g3, 0};
  out[0] = g1 / S1 - g2 * ls1 - g3 * S1 * ls1 + fXA[0] - g1
    + lub_series (fXA, 2, mmax, 0, s, l, &u, 1.0);
  out[1] = (2.0 / s * g1 / S1 + g2 * ls2 + g3 * S1 * ls2 + 4.0 * g3 / s
            + lub_series (fXA, 1, mmax, 1, s, l, &u, 1.0)) * (-2.0 / l1);
  /* YA */
  g2 = 4.0 / 15.0 * l * (2.0 + l + 2.0 * l * l) / l13;
  g3 = 2.0 / 375.0 * (16.0 - 45.0 * l + 58.0 * l * l - 45.0 * l * l * l
                      + 16.0 * l * l * l * l) / l13;
  u = (lub_sub) {0.0, g2, g3, 0};
  out[2] = -g2 * ls1 - g3 * S1 * ls1 + fYA[0]
    + lub_series (fYA, 2, mmax, 0, s, l, &u, 1.0);
  out[3] = (g2 * ls2 + g3 * S1 * ls2 + 4.0 * g3 / s
            + lub_series (fYA, 1, mmax, 1, s, l, &u, 1.0)) * (-2.0 / l1);
  if (version == 0) return;
  /* YB */
  g2 = -0.2 * l * (4.0 + l) / l12;
  g3 = -0.004 * (32.0 - 33.0 * l + 83.0 * l * l + 43.0 * l * l * l) / l12;
  u = (lub_sub) {0.0, g2, g3, 0};
  out[4] = g2 * ls2 + g3 * S1 * ls2 + 4.0 * g3 / s
    + lub_series (fYB, 1, mmax, 1, s, l, &u, 1.0);
  out[5] = (-g2 * ls1 - g3 * S1 * ls1
            + lub_series (fYB, 2, mmax, 0, s, l, &u, 1.0)) * (-4.0 / l12);
  /* XC: a_m = -l^2 / (1 + l) / (m (m - 1)) from m = 2 */
  {
    const double c = l * l / l1;
    double v11 = c * (0.5 * ls1 + ls2 / s) + 1.0;
    double v12 = -c * (0.5 * ls2 + (ls1 - 2.0) / s);
    const double s1l = s * l1, s21 = 2.0 / s;
    double tm = 1.0, t2 = 1.0;
    for (int m = 1; m <= mmax; m++)
      {
        tm *= s1l;
        t2 *= s21;
        if (m < 2) continue;
        const double a = -c / (double) (m * (m - 1));
        if (m % 2 == 0) v11 += fXC[m] / tm + a * t2;
        else v12 += fXC[m] / tm + a * t2;
      }
    out[6] = v11;
    out[7] = v12 * (-8.0 / l13);
  }
  /* YC */
  g2 = 0.4 * l / l1;
  g3 = 0.008 * (8.0 + 6.0 * l + 33.0 * l * l) / l1;
  g4 = 0.8 * l * l / l14;
  g5 = yc12_g5 * 0.016 * l * (43.0 - 24.0 * l + 43.0 * l * l) / l14;
  u = (lub_sub) {0.0, g2, g3, 0};
  out[8] = -g2 * ls1 - g3 * S1 * ls1 + fYC[0]
    + lub_series (fYC, 2, mmax, 0, s, l, &u, 1.0);
  u = (lub_sub) {0.0, g4, g5, 0};
  out[9] = (g4 * ls2 + g5 * S1 * ls2 + 4.0 * g5 / s
            + lub_series (fYC, 1, mmax, 1, s, l, &u, 1.0)) * (8.0 / l13);
  if (version == 1) return;

  /* the G, H, M forms in terms of s^2 - 4 */
  const double T1 = 2.0 * s / (s * s - 4.0), T3 = (s * s - 4.0) / 4.0;
  const double T4 = 4.0 / (s * s - 4.0);
  const double ls5 = log (s * s / (s * s - 4.0));
  /* XG */
  g1 = 3.0 * l * l / l13;
  g2 = 0.3 * l * (1.0 + 12.0 * l - 4.0 * l * l) / l13;
  g3 = (5.0 + 181.0 * l - 453.0 * l * l + 566.0 * l * l * l
        - 65.0 * l * l * l * l) / (140.0 * l13);
  u = (lub_sub) {g1, g2, g3, 1};
  out[10] = g1 * T1 + (g2 + g3 * T3) * ls2 - g3 * s
    + lub_series (fXG, 1, mmax, 1, s, l, &u, 1.0);
  out[11] = (-g1 * T4 - (g2 + g3 * T3) * ls5 + g3
             + lub_series (fXG, 2, mmax, 0, s, l, &u, -1.0)) * (4.0 / l12);
  /* YG */
  g2 = 0.1 * l * (4.0 - l + 7.0 * l * l) / l13;
  g3 = (32.0 - 179.0 * l + 532.0 * l * l - 356.0 * l * l * l
        + 221.0 * l * l * l * l) / (500.0 * l13);
  u = (lub_sub) {0.0, g2, g3, 1};
  out[12] = (g2 + g3 * T3) * ls2 - g3 * s
    + lub_series (fYG, 1, mmax, 1, s, l, &u, 1.0);
  out[13] = (-(g2 + g3 * T3) * ls5 + g3
             + lub_series (fYG, 2, mmax, 0, s, l, &u, -1.0)) * (4.0 / l12);
  /* YH */
  g2 = 0.1 * l * (2.0 - l) / l12;
  g3 = (16.0 - 61.0 * l + 180.0 * l * l + 2.0 * l * l * l) / (500.0 * l12);
  g5 = 0.05 * l * l * (1.0 + 7.0 * l) / l12;
  g6 = l * (43.0 + 147.0 * l - 185.0 * l * l + 221.0 * l * l * l) / (1000.0 * l12);
  u = (lub_sub) {0.0, g2, g3, 1};
  out[14] = (g2 + g3 * T3) * ls5 - g3
    + lub_series (fYH, 2, mmax, 0, s, l, &u, 1.0);
  u = (lub_sub) {0.0, g5, g6, 1};
  out[15] = ((g5 + g6 * T3) * ls2 - g6 * s
             + lub_series (fYH, 1, mmax, 1, s, l, &u, 1.0)) * (8.0 / l13);
  /* XM */
  g1 = 1.2 * l * l / l13;
  g2 = 0.12 * l * (1.0 + 17.0 * l - 9.0 * l * l) / l13;
  g3 = (5.0 + 272.0 * l - 831.0 * l * l + 1322.0 * l * l * l
        - 415.0 * l * l * l * l) / (350.0 * l13);
  g4 = g1 * l;
  g5 = 0.12 * l * l * (-4.0 + 17.0 * l - 4.0 * l * l) / l13;
  g6 = l * (-65.0 + 832.0 * l - 1041.0 * l * l + 832.0 * l * l * l
            - 65.0 * l * l * l * l) / (350.0 * l13);
  u = (lub_sub) {g1, g2, g3, 1};
  out[16] = g1 * T4 + (g2 + g3 * T3) * ls5 - g3 + 1.0
    + lub_series (fXM, 2, mmax, 0, s, l, &u, 1.0);
  u = (lub_sub) {g4, g5, g6, 1};
  out[17] = (g4 * T1 + (g5 + g6 * T3) * ls2 - g6 * s
             + lub_series (fXM, 1, mmax, 1, s, l, &u, 1.0)) * (8.0 / l13);
  /* YM */
  g2 = 0.24 * l * (1.0 - l + 4.0 * l * l) / l13;
  g3 = (24.0 - 201.0 * l + 882.0 * l * l - 1182.0 * l * l * l
        + 591.0 * l * l * l * l) / (625.0 * l13);
  g5 = 0.06 * l * l * (7.0 - 10.0 * l + 7.0 * l * l) / l13;
  g6 = 3.0 * l * (221.0 - 728.0 * l + 1902.0 * l * l - 728.0 * l * l * l
                  + 221.0 * l * l * l * l) / (2500.0 * l13);
  u = (lub_sub) {0.0, g2, g3, 1};
  out[18] = (g2 + g3 * T3) * ls5 - g3 + 1.0
    + lub_series (fYM, 2, mmax, 0, s, l, &u, 1.0);
  u = (lub_sub) {0.0, g5, g6, 1};
  out[19] = ((g5 + g6 * T3) * ls2 - g6 * s
             + lub_series (fYM, 1, mmax, 1, s, l, &u, 1.0)) * (8.0 / l13);
  /* ZM */
  g3 = -0.3 * l * l * (1.0 + l * l) / l13;
  u = (lub_sub) {0.0, 0.0, g3, 1};
  out[20] = g3 * T3 * ls5 - g3 + 1.0
    + lub_series (fZM, 2, mmax, 0, s, l, &u, 1.0);
  out[21] = (-g3 * T3 * ls2 + g3 * s
             + lub_series (fZM, 1, mmax, 1, s, l, &u, -1.0)) * (8.0 / l13);
}

/* twobody_scale (twobody.c:2709-2772, flag 1: dimensional, radius a1) and
 * twobody_scale_SD (2646-2700, flag 2, a1 = 1) */
void
orc_twobody_scale (int version, double *two, double a1, double l, int sd)
{
  const double l1 = 1.0 + l, lA = l1 / 2.0, lB = l1 * l1 / 6.0;
  const double lC = l1 * l1 * l1 / 6.0, lM = 5.0 / 6.0 * lC;
  const double a = sd ? 1.0 : a1, a2 = a * a, a3 = a2 * a;
  if (!sd) { two[0] *= a; two[2] *= a; }
  two[1] *= a * lA;
  two[3] *= a * lA;
  if (version == 0) return;
  two[4] *= a2 * 2.0 / 3.0; two[5] *= a2 * lB;
  two[6] *= a3 * 4.0 / 3.0; two[7] *= a3 * lC;
  two[8] *= a3 * 4.0 / 3.0; two[9] *= a3 * lC;
  if (version == 1) return;
  two[10] *= a2 * 2.0 / 3.0; two[11] *= a2 * lB;
  two[12] *= a2 * 2.0 / 3.0; two[13] *= a2 * lB;
  two[14] *= a3 * 4.0 / 3.0; two[15] *= a3 * lC;
  for (int k = 16; k < 22; k += 2)
    {
      two[k] *= a3 * 10.0 / 9.0;
      two[k + 1] *= a3 * lM;
    }
}

/* twobody_scalars_res (twobody.c:2786-2814) */
void
orc_twobody_res (int version, double r, double aa, double ab, int n,
                 int flag_lub, int flag_scale, double *res)
{
  const double l = ab / aa, s = 2.0 * r / (aa + ab);
  if (n % 2) n--;
  double *f = malloc (sizeof (double) * 11 * (size_t) (n + 1));
  for (int k = 0; k < 11; k++) orc_twobody_table (k, n, l, f + (size_t) k * (n + 1));
  for (int k = 0; k < 22; k++) res[k] = 0.0;
  if (flag_lub) orc_twobody_lub (version, f, n + 1, n - 1, l, s, 1.0, res);
  else orc_twobody_far (version, f, n, l, s, res);
  free (f);
  if (flag_scale == 1) orc_twobody_scale (version, res, aa, l, 0);
  else if (flag_scale == 2) orc_twobody_scale (version, res, aa, l, 1);
}

/* scalar_two_body_res (two-body-res.c:31): equal spheres, the lubrication
 * form at lambda = 1 with terms m <= 50 and the YC12 g5 coefficient of
 * Jeffrey & Onishi (twice twobody.c's); dimensional form at a = 1 */
void
orc_two_body_res (double s, double *res)
{
  const int n = 50;
  double f[11 * 51];
  for (int k = 0; k < 11; k++) orc_twobody_table (k, n, 1.0, f + k * 51);
  orc_twobody_lub (2, f, 51, n, 1.0, s, 2.0, res);
  orc_twobody_scale (2, res, 1.0, 1.0, 0);
}

/* ------------------------------------------------------------------ */
/* inverse of the two-body far-field mobility                           */
/* ------------------------------------------------------------------ */

/* Gauss-Jordan with partial pivoting, in place, row-major n x n */
static int
orc_inv (int n, double *a)
{
  int piv[8];
  double b[64];
  for (int i = 0; i < n * n; i++) b[i] = 0.0;
  for (int i = 0; i < n; i++) b[i * n + i] = 1.0;
  for (int c = 0; c < n; c++)
    {
      int p = c;
      for (int r = c + 1; r < n; r++)
        if (fabs (a[r * n + c]) > fabs (a[p * n + c])) p = r;
      piv[c] = p;
      if (a[p * n + c] == 0.0) return -1;
      for (int k = 0; k < n; k++)
        {
          double t = a[c * n + k]; a[c * n + k] = a[p * n + k]; a[p * n + k] = t;
          t = b[c * n + k]; b[c * n + k] = b[p * n + k]; b[p * n + k] = t;
        }
      const double d = 1.0 / a[c * n + c];
      for (int k = 0; k < n; k++) { a[c * n + k] *= d; b[c * n + k] *= d; }
      for (int r = 0; r < n; r++)
        if (r != c)
          {
            const double g = a[r * n + c];
            if (g == 0.0) continue;
            for (int k = 0; k < n; k++)
              {
                a[r * n + k] -= g * a[c * n + k];
                b[r * n + k] -= g * b[c * n + k];
              }
          }
    }
  (void) piv;
  for (int i = 0; i < n * n; i++) a[i] = b[i];
  return 0;
}

/* scalars_nonewald_poly_full (non-ewald.c:230-318): the RPY far-field
 * mobility of spheres aa, ab at distance r, dimensional, as (11, 12, 21,
 * 22) quadruples per function */
static void
orc_nonewald_full (double r, double aa, double ab, double *m)
{
  for (int k = 0; k < 44; k++) m[k] = 0.0;
  const double r2 = r * r, aa2 = aa * aa, ab2 = ab * ab, aa3 = aa2 * aa;
  const double ab3 = ab2 * ab, sq = aa2 + ab2;
  double c12;
  m[0] = 1.0 / aa; m[3] = 1.0 / ab;
  m[1] = m[2] = (1.5 - 0.5 * sq / r2) / r;
  m[4] = 1.0 / aa; m[7] = 1.0 / ab;
  m[5] = m[6] = (0.75 + 0.25 * sq / r2) / r;
  m[9] = m[10] = -0.75 / r2;
  m[12] = 0.75 / aa3; m[15] = 0.75 / ab3;
  m[13] = m[14] = 0.75 / (r2 * r);
  m[16] = 0.75 / aa3; m[19] = 0.75 / ab3;
  m[17] = m[18] = -0.375 / (r2 * r);
  m[21] = (2.25 - 13.5 * (aa2 / 10.0 + ab2 / 6.0) / r2) / r2;
  m[22] = (2.25 - 13.5 * (ab2 / 10.0 + aa2 / 6.0) / r2) / r2;
  m[25] = 4.5 * (aa2 / 10.0 + ab2 / 6.0) / (r2 * r2);
  m[26] = 4.5 * (ab2 / 10.0 + aa2 / 6.0) / (r2 * r2);
  m[29] = m[30] = -1.125 / (r2 * r);
  m[32] = 0.9 / aa3; m[35] = 0.9 / ab3;
  c12 = (-4.5 + 5.4 * sq / r2) / (r2 * r);
  m[33] = m[34] = c12;
  m[36] = 0.9 / aa3; m[39] = 0.9 / ab3;
  m[37] = m[38] = (2.25 - 3.6 * sq / r2) / (r2 * r);
  m[40] = 0.9 / aa3; m[43] = 0.9 / ab3;
  m[41] = m[42] = 0.9 * sq / (r2 * r2 * r);
}

/* scalars_minv_fts_poly (minv-poly.c:267-560): the resistance scalars of
 * the inverse of the two-body far-field mobility, dimensional.  The
 * blocks decouple by azimuthal order: the rotation about and the m = 2
 * straining along the line of centres (2 x 2 each), the transverse
 * translation / rotation / straining (6 x 6) and the axial translation
 * with the two axisymmetric strainings (6 x 6); the signs and factors of
 * the entries are those of the block forms (f.c, ft.c, fts.c) in that
 * basis.  version 0: XA, YA only (2 x 2 each); 1: adds YB, XC, YC. */
void
orc_minv_poly (int version, double r, double aa, double ab, double *out)
{
  double m[44];
  orc_nonewald_full (r, aa, ab, m);
  for (int k = 0; k < 44; k++) out[k] = 0.0;
#define Q4(k) m[4 * (k)], m[4 * (k) + 1], m[4 * (k) + 2], m[4 * (k) + 3]
  if (version < 2)
    {
      double a[4] = {Q4 (0)};
      orc_inv (2, a);
      for (int k = 0; k < 4; k++) out[k] = a[k];
      if (version == 0)
        {
          double y[4] = {Q4 (1)};
          orc_inv (2, y);
          for (int k = 0; k < 4; k++) out[4 + k] = y[k];
          return;
        }
      double c[4] = {Q4 (3)};
      orc_inv (2, c);
      for (int k = 0; k < 4; k++) out[12 + k] = c[k];
      /* transverse translation and rotation (scalars_minv_ft_poly,
       * minv-poly.c:113-258): unknowns (U1, U2, O1, O2) */
      double t[16] = {
        m[4], m[5], -m[8], m[10],
        m[6], m[7], -m[9], m[11],
        -m[8], -m[9], m[16], m[17],
        m[10], m[11], m[18], m[19]};
      orc_inv (4, t);
      out[4] = t[0]; out[5] = t[1]; out[6] = t[4]; out[7] = t[5];
      out[8] = -t[8]; out[9] = -t[9]; out[10] = t[12]; out[11] = t[13];
      out[16] = t[10]; out[17] = t[11]; out[18] = t[14]; out[19] = t[15];
      return;
    }
  /* XC, ZM */
  {
    double a[4] = {Q4 (3)};
    orc_inv (2, a);
    for (int k = 0; k < 4; k++) out[12 + k] = a[k];
    double z[4] = {Q4 (10)};
    orc_inv (2, z);
    for (int k = 0; k < 4; k++) out[40 + k] = z[k];
  }
  /* transverse modes: (U1, U2, O1, O2, S1, S2) */
  {
    const double *ya = m + 4, *yb = m + 8, *yc = m + 16, *yg = m + 24;
    const double *yh = m + 28, *ym = m + 36;
    double a[36] = {
      ya[0], ya[1], -yb[0], yb[2], 2.0 * yg[0], -2.0 * yg[2],
      ya[2], ya[3], -yb[1], yb[3], 2.0 * yg[1], -2.0 * yg[3],
      -yb[0], -yb[1], yc[0], yc[1], 2.0 * yh[0], 2.0 * yh[2],
      yb[2], yb[3], yc[2], yc[3], 2.0 * yh[1], 2.0 * yh[3],
      yg[0], yg[1], yh[0], yh[1], ym[0], ym[1],
      -yg[2], -yg[3], yh[2], yh[3], ym[2], ym[3]};
    orc_inv (6, a);
    out[4] = a[0]; out[5] = a[1]; out[6] = a[6]; out[7] = a[7];
    out[8] = -a[12]; out[9] = -a[13]; out[10] = a[18]; out[11] = a[19];
    out[16] = a[14]; out[17] = a[15]; out[18] = a[20]; out[19] = a[21];
    out[24] = a[24]; out[25] = a[25]; out[26] = -a[30]; out[27] = -a[31];
    out[28] = a[26]; out[29] = a[27]; out[30] = a[32]; out[31] = a[33];
    out[36] = a[28]; out[37] = a[29]; out[38] = a[34]; out[39] = a[35];
  }
  /* axial modes: (U1, U2, and the two axisymmetric strain combinations of
   * each sphere, XM +- ZM) */
  {
    const double *xa = m, *xg = m + 20, *xm = m + 32, *zm = m + 40;
    double p[4], q[4];
    for (int k = 0; k < 4; k++)
      {
        p[k] = 0.5 * (xm[k] + zm[k]);
        q[k] = 0.5 * (xm[k] - zm[k]);
      }
    double a[36] = {
      xa[0], xa[1], -xg[0], xg[2], -xg[0], xg[2],
      xa[2], xa[3], -xg[1], xg[3], -xg[1], xg[3],
      -xg[0] / 3.0, -xg[1] / 3.0, p[0], p[1], q[0], q[1],
      xg[2] / 3.0, xg[3] / 3.0, p[2], p[3], q[2], q[3],
      -xg[0] / 3.0, -xg[1] / 3.0, q[0], q[1], p[0], p[1],
      xg[2] / 3.0, xg[3] / 3.0, q[2], q[3], p[2], p[3]};
    orc_inv (6, a);
    out[0] = a[0]; out[1] = a[1]; out[2] = a[6]; out[3] = a[7];
    out[20] = -3.0 * a[12]; out[21] = -3.0 * a[13];
    out[22] = 3.0 * a[18]; out[23] = 3.0 * a[19];
    out[32] = 2.0 * a[14] - out[40]; out[33] = 2.0 * a[15] - out[41];
    out[34] = 2.0 * a[20] - out[42]; out[35] = 2.0 * a[21] - out[43];
  }
#undef Q4
}

/* scalar_minv_fts (fts.c:1490): equal spheres a = 1, (11, 12) pairs */
void
orc_minv_fts (double s, double *res)
{
  double m[44];
  orc_minv_poly (2, s, 1.0, 1.0, m);
  for (int k = 0; k < 11; k++)
    {
      res[2 * k] = m[4 * k];
      res[2 * k + 1] = m[4 * k + 1];
    }
}

/* ------------------------------------------------------------------ */
/* the lubrication correction                                           */
/* ------------------------------------------------------------------ */

/* matrix_fts_self_atimes (fts.c:1024-1400): the block with the i == j
 * signs applied to the extracted-form vector */
void
orc_block_atimes_self (int version, const double *x, double *y, double ex,
                       double ey, double ez, const double *s)
{
  orc_block_atimes_sign (version, x, y, ex, ey, ez, s, 1);
}

static void
orc_lub_pair (const orc_sys *sys, double lubmin2, int nmax, int flag_lub,
              int i, int j, const double *pi, const double *pj,
              const double *uoe, double *fts)
{
  const int v = sys->version, d = v == 0 ? 3 : (v == 1 ? 6 : 11);
  double xx = pj[0] - pi[0], yy = pj[1] - pi[1], zz = pj[2] - pi[2];
  double r2 = xx * xx + yy * yy + zz * zz;
  double s11[11] = {0}, s12[11] = {0}, s21[11] = {0}, s22[11] = {0};
  if (sys->a == NULL)
    {
      /* calc_lub_fts_2b (fts.c:1758-1869): r clamped to lubmin */
      if (r2 < lubmin2) r2 = lubmin2;
      /* the far-field inverse of the version (scalar_minv_f, f.c:183;
       * scalar_minv_ft, ft.c:414; scalar_minv_fts, fts.c:1490) */
      double rr = sqrt (r2), res[22], inf[44];
      orc_two_body_res (rr, res);
      orc_minv_poly (v, rr, 1.0, 1.0, inf);
      for (int k = 0; k < 11; k++)
        {
          s11[k] = s22[k] = res[2 * k] - inf[4 * k];
          s12[k] = s21[k] = res[2 * k + 1] - inf[4 * k + 1];
        }
    }
  else
    {
      /* calc_lub_fts_2b_poly (fts.c:1991-2147): s = 2 r / (a1 + a2)
       * clamped to lubmin, scalars_lub_poly_full (minv-poly.c:759) in the
       * dimensional form, then scalars_res_poly_scale_SD (565-660) */
      const double a1 = sys->a[i], a2 = sys->a[j];
      const double rs = 0.25 * (a1 + a2) * (a1 + a2);
      if (r2 / rs < lubmin2) r2 = rs * lubmin2;
      double rr = sqrt (r2), r12[22], r21[22], inf[44], lub[44];
      orc_twobody_res (v, rr, a1, a2, nmax, flag_lub, 1, r12);
      orc_twobody_res (v, rr, a2, a1, nmax, flag_lub, 1, r21);
      orc_minv_poly (v, rr, a1, a2, inf);
      for (int k = 0; k < 11; k++)
        {
          lub[4 * k] = r12[2 * k] - inf[4 * k];
          lub[4 * k + 1] = r12[2 * k + 1] - inf[4 * k + 1];
          lub[4 * k + 2] = r21[2 * k + 1] - inf[4 * k + 2];
          lub[4 * k + 3] = r21[2 * k] - inf[4 * k + 3];
        }
      /* SD scaling: divided by a^1 (A), a^2 (B, G), a^3 (C, H, M) of the
       * sphere whose force the scalar gives */
      static const int pw[11] = {1, 1, 2, 3, 3, 2, 2, 3, 3, 3, 3};
      for (int k = 0; k < 11; k++)
        {
          const double d1 = pow (a1, pw[k]), d2 = pow (a2, pw[k]);
          s11[k] = lub[4 * k] / d1;
          s12[k] = lub[4 * k + 1] / d1;
          s21[k] = lub[4 * k + 2] / d2;
          s22[k] = lub[4 * k + 3] / d2;
        }
    }
  /* the direction uses the (possibly clamped) distance, as the reference */
  const double rc = sqrt (r2);
  const double ex = xx / rc, ey = yy / rc, ez = zz / rc;
  orc_block_atimes_self (v, uoe + (size_t) d * i, fts + (size_t) d * i, ex, ey, ez, s11);
  orc_block_atimes (v, uoe + (size_t) d * j, fts + (size_t) d * i, ex, ey, ez, s12);
  orc_block_atimes_self (v, uoe + (size_t) d * j, fts + (size_t) d * j, -ex, -ey, -ez, s22);
  orc_block_atimes (v, uoe + (size_t) d * i, fts + (size_t) d * j, -ex, -ey, -ez, s21);
}

/* calc_lub_3fts (lub.c:576-768): pairs j >= i over the image cells
 * -imax..imax of each direction (set_imax_lub_periodic, lub.c:129-162),
 * cond_lub / cond_lub_poly (lub.c:42-127): 0 < r and r^2 (scaled s^2 for
 * unequal spheres) < lubmax^2, no limit for lubmax <= 0 */
void
orc_lub_3fts (const orc_sys *sys, double lubmin2, double lubmax, int nmax,
              int flag_lub, const double *uoe, double *fts)
{
  const int np = sys->np, d = orc_dim_of (sys->version);
  for (int k = 0; k < np * d; k++) fts[k] = 0.0;
  const double lm2 = lubmax * lubmax;
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
              const double *pi = sys->pos + 3 * i;
              double pj[3] = {sys->pos[3 * j] + sys->lx * ix,
                              sys->pos[3 * j + 1] + sys->ly * iy,
                              sys->pos[3 * j + 2] + sys->lz * iz};
              if (sys->shear_mode == 1) pj[0] += iy * sys->shear_shift;
              else if (sys->shear_mode == 2) pj[0] += iz * sys->shear_shift;
              const double dx = pi[0] - pj[0], dy = pi[1] - pj[1], dz = pi[2] - pj[2];
              double r2 = dx * dx + dy * dy + dz * dz;
              if (r2 == 0.0) continue;
              if (sys->a)
                r2 *= 4.0 / ((sys->a[i] + sys->a[j]) * (sys->a[i] + sys->a[j]));
              if (lm2 > 0.0 && !(r2 < lm2)) continue;
              orc_lub_pair (sys, lubmin2, nmax, flag_lub, i, j, pi, pj, uoe, fts);
            }
}
