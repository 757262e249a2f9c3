// lub_math.h -- two-body lubrication correction of the resistance, shared by
// the device kernels and the host exports (FP64 throughout).
//
// Scalar order: two[22] = (11, 12) pairs of XA YA YB XC YC XG YG YH XM YM ZM
// (twobody.c:644-2814); four-fold sets [44] = (11, 12, 21, 22) per function
// (minv-poly.c).  A sphere pair is (1, 2) with lambda = a2 / a1 and scaled
// distance s = 2 r / (a1 + a2).
//
//  * lub_form / far_form: the Jeffrey & Onishi (1984) / Jeffrey (1992)
//    resistance functions from the coefficients f_k(lambda): the 1/s series
//    (far form, twobody_far_with_f) or the near-contact asymptotic form plus
//    the remaining series (lub form, twobody_lub_with_f).  The f values come
//    from a functor fv(family, k), so the device can evaluate the
//    polynomials in lambda on the fly.
//  * minv: the resistance scalars of the inverse of the two-body far-field
//    (RPY) mobility, mode by mode (scalars_minv_*_poly, minv-poly.c).
//  * apply_block: y += B(e; s) T(x), the pair block of fts.c / ft.c / f.c
//    with the pair or the self (i == j) signs.
#pragma once
#include <hip/hip_runtime.h>

#define LUB_HD __host__ __device__ inline

namespace shk {
namespace lub {

// families in the order of struct twobody_f
enum { XA = 0, YA, YB, XC, YC, XG, YG, YH, XM, YM, ZM, NFAM };

LUB_HD int nfam_of (int version) { return version == 0 ? 2 : (version == 1 ? 5 : 11); }

// ---------------------------------------------------------------------------
// sums of f_m t^m over m in [m0, m1], even m into se and odd m into so.
// The functor evaluates four consecutive rows together (fv.four), which
// lets it interleave their recurrences; t^m by repeated products.
// ---------------------------------------------------------------------------
template <class FV>
LUB_HD void fsums (const FV &fv, int fam, int m0, int m1, double t, double &se, double &so)
{
  se = so = 0.0;
  double tm = 1.0;
  for (int m = 0; m < m0; m++) tm *= t;
  int m = m0;
#pragma unroll 1
  for (; m + 3 <= m1; m += 4)
    {
      double f[4];
      fv.four (fam, m, f);
      const double t0 = tm, t1 = t0 * t, t2 = t1 * t, t3 = t2 * t;
      tm = t3 * t;
      if (m & 1)
        {
          so += f[0] * t0; se += f[1] * t1;
          so += f[2] * t2; se += f[3] * t3;
        }
      else
        {
          se += f[0] * t0; so += f[1] * t1;
          se += f[2] * t2; so += f[3] * t3;
        }
    }
#pragma unroll 1
  for (; m <= m1; m++)
    {
      const double fm = fv (fam, m);
      if (m & 1) so += fm * tm;
      else se += fm * tm;
      tm *= t;
    }
}

// ---------------------------------------------------------------------------
// far form: the 1/(s (1 + l)) series, even k for one of (11, 12), odd for the
// other, the 12 part times its prefactor (n even: k < n)
// ---------------------------------------------------------------------------
template <class FV>
LUB_HD void far_form (int version, const FV &fv, int n, double l, double s, double *two)
{
  const double l1 = 1.0 + l, l12 = l1 * l1, l13 = l12 * l1, t = 1.0 / (s * l1);
  // which of (11, 12) takes the odd terms, and the 12 prefactor
  const bool odd11[NFAM] = {false, false, true, false, false, true, true, false, false,
                            false, false};
  const double pre[NFAM] = {-2.0 / l1,  -2.0 / l1,  -4.0 / l12, -8.0 / l13,
                            8.0 / l13,  -4.0 / l12, -4.0 / l12, 8.0 / l13,
                            8.0 / l13,  8.0 / l13,  -8.0 / l13};
  const int nf = nfam_of (version);
  for (int f = 0; f < nf; f++)
    {
      double ev, od;
      fsums (fv, f, 0, n - 1, t, ev, od);
      two[2 * f] = odd11[f] ? od : ev;
      two[2 * f + 1] = pre[f] * (odd11[f] ? ev : od);
    }
}

// ---------------------------------------------------------------------------
// lubrication form
// ---------------------------------------------------------------------------
// The remaining series of one function: the terms m = 1 .. mmax of one
// parity, f_m (s (1 + l))^-m + a_m (2 / s)^m with a_m the m-th coefficient
// of the subtracted near-field part, a_m = -c1 - 2 c2 / m + 4 c3 / d_m,
// d_m = m (m - 2) (-2 m at m = 2, as twobody.c:1055-1062) or m (m + 2) for
// the G, H, M forms.  sf: sign of each term.  fsum: the f_m part of that
// parity (fsums over m = 1 .. mmax).
LUB_HD double rest (double fsum, int mmax, int parity, int mlo, double s, double c1, double c2,
                    double c3, bool plus2, double sf)
{
  const double u = 2.0 / s;
  double um = 1.0, acc = 0.0;
  for (int m = 1; m <= mmax; m++)
    {
      um *= u;
      if ((m & 1) != parity || m < mlo) continue;
      const double dm = plus2 ? double (m * (m + 2)) : (m == 2 ? -4.0 : double (m * (m - 2)));
      const double am = -c1 - 2.0 * c2 / m + 4.0 * c3 / dm;
      acc += am * um;
    }
  return sf * (fsum + acc);
}

// twobody_lub_with_f (twobody.c:1977-2644) with terms m <= mmax; yc12_g5
// scales the g5 coefficient of YC12 (1: twobody.c, 2: the equal-sphere
// series of two-body-res.c, whose coefficient is Jeffrey & Onishi's)
template <class FV>
LUB_HD void lub_form (int version, const FV &fv, int mmax, double l, double s, double yc12_g5,
                      double *two)
{
  const double l1 = 1.0 + l, l12 = l1 * l1, l13 = l12 * l1, l14 = l12 * l12;
  const double ll = l * l, l3 = ll * l, l4 = ll * ll;
  const double S1 = 1.0 - 4.0 / (s * s);
  const double tt = 1.0 / (s * l1);
  double fe, fo;
  const double ls1 = log (S1), ls2 = log ((s + 2.0) / (s - 2.0));
  double g1, g2, g3, g4, g5, g6;
  // XA
  g1 = 2.0 * ll / l13;
  g2 = 0.2 * l * (1.0 + 7.0 * l + ll) / l13;
  g3 = (1.0 + 18.0 * l - 29.0 * ll + 18.0 * l3 + l4) / (42.0 * l13);
  fsums (fv, XA, 1, mmax, tt, fe, fo);
  two[0] = g1 / S1 - g2 * ls1 - g3 * S1 * ls1 + fv (XA, 0) - g1
    + rest (fe, mmax, 0, 2, s, g1, g2, g3, false, 1.0);
  two[1] = -2.0 / l1 * (2.0 / s * g1 / S1 + g2 * ls2 + g3 * S1 * ls2 + 4.0 * g3 / s
                        + rest (fo, mmax, 1, 1, s, g1, g2, g3, false, 1.0));
  // YA
  g2 = 4.0 / 15.0 * l * (2.0 + l + 2.0 * ll) / l13;
  g3 = 2.0 / 375.0 * (16.0 - 45.0 * l + 58.0 * ll - 45.0 * l3 + 16.0 * l4) / l13;
  fsums (fv, YA, 1, mmax, tt, fe, fo);
  two[2] = -g2 * ls1 - g3 * S1 * ls1 + fv (YA, 0)
    + rest (fe, mmax, 0, 2, s, 0.0, g2, g3, false, 1.0);
  two[3] = -2.0 / l1 * (g2 * ls2 + g3 * S1 * ls2 + 4.0 * g3 / s
                        + rest (fo, mmax, 1, 1, s, 0.0, g2, g3, false, 1.0));
  if (version == 0) return;
  // YB
  g2 = -0.2 * l * (4.0 + l) / l12;
  g3 = -0.004 * (32.0 - 33.0 * l + 83.0 * ll + 43.0 * l3) / l12;
  fsums (fv, YB, 1, mmax, tt, fe, fo);
  two[4] = g2 * ls2 + g3 * S1 * ls2 + 4.0 * g3 / s
    + rest (fo, mmax, 1, 1, s, 0.0, g2, g3, false, 1.0);
  two[5] = -4.0 / l12 * (-g2 * ls1 - g3 * S1 * ls1
                         + rest (fe, mmax, 0, 2, s, 0.0, g2, g3, false, 1.0));
  // XC: a_m = -l^2 / ((1 + l) m (m - 1)) from m = 2
  {
    const double c = ll / l1, u = 2.0 / s;
    double v11 = c * (0.5 * ls1 + ls2 / s) + 1.0;
    double v12 = -c * (0.5 * ls2 + (ls1 - 2.0) / s);
    double a11 = 0.0, a12 = 0.0, um = u;
    for (int m = 2; m <= mmax; m++)
      {
        um *= u;
        const double term = -c / double (m * (m - 1)) * um;
        if (m & 1) a12 += term;
        else a11 += term;
      }
    fsums (fv, XC, 2, mmax, tt, fe, fo);
    two[6] = v11 + (fe + a11);
    two[7] = -8.0 / l13 * (v12 + (fo + a12));
  }
  // YC
  g2 = 0.4 * l / l1;
  g3 = 0.008 * (8.0 + 6.0 * l + 33.0 * ll) / l1;
  g4 = 0.8 * ll / l14;
  g5 = yc12_g5 * 0.016 * l * (43.0 - 24.0 * l + 43.0 * ll) / l14;
  fsums (fv, YC, 1, mmax, tt, fe, fo);
  two[8] = -g2 * ls1 - g3 * S1 * ls1 + fv (YC, 0)
    + rest (fe, mmax, 0, 2, s, 0.0, g2, g3, false, 1.0);
  two[9] = 8.0 / l13 * (g4 * ls2 + g5 * S1 * ls2 + 4.0 * g5 / s
                        + rest (fo, mmax, 1, 1, s, 0.0, g4, g5, false, 1.0));
  if (version == 1) return;
  // G, H, M: forms in s^2 - 4
  const double q = s * s - 4.0;
  const double T1 = 2.0 * s / q, T3 = 0.25 * q, T4 = 4.0 / q, ls5 = log (s * s / q);
  // XG
  g1 = 3.0 * ll / l13;
  g2 = 0.3 * l * (1.0 + 12.0 * l - 4.0 * ll) / l13;
  g3 = (5.0 + 181.0 * l - 453.0 * ll + 566.0 * l3 - 65.0 * l4) / (140.0 * l13);
  fsums (fv, XG, 1, mmax, tt, fe, fo);
  two[10] = g1 * T1 + (g2 + g3 * T3) * ls2 - g3 * s
    + rest (fo, mmax, 1, 1, s, g1, g2, g3, true, 1.0);
  two[11] = 4.0 / l12 * (-g1 * T4 - (g2 + g3 * T3) * ls5 + g3
                         + rest (fe, mmax, 0, 2, s, g1, g2, g3, true, -1.0));
  // YG
  g2 = 0.1 * l * (4.0 - l + 7.0 * ll) / l13;
  g3 = (32.0 - 179.0 * l + 532.0 * ll - 356.0 * l3 + 221.0 * l4) / (500.0 * l13);
  fsums (fv, YG, 1, mmax, tt, fe, fo);
  two[12] = (g2 + g3 * T3) * ls2 - g3 * s
    + rest (fo, mmax, 1, 1, s, 0.0, g2, g3, true, 1.0);
  two[13] = 4.0 / l12 * (-(g2 + g3 * T3) * ls5 + g3
                         + rest (fe, mmax, 0, 2, s, 0.0, g2, g3, true, -1.0));
  // YH
  g2 = 0.1 * l * (2.0 - l) / l12;
  g3 = (16.0 - 61.0 * l + 180.0 * ll + 2.0 * l3) / (500.0 * l12);
  g5 = 0.05 * ll * (1.0 + 7.0 * l) / l12;
  g6 = l * (43.0 + 147.0 * l - 185.0 * ll + 221.0 * l3) / (1000.0 * l12);
  fsums (fv, YH, 1, mmax, tt, fe, fo);
  two[14] = (g2 + g3 * T3) * ls5 - g3
    + rest (fe, mmax, 0, 2, s, 0.0, g2, g3, true, 1.0);
  two[15] = 8.0 / l13 * ((g5 + g6 * T3) * ls2 - g6 * s
                         + rest (fo, mmax, 1, 1, s, 0.0, g5, g6, true, 1.0));
  // XM
  g1 = 1.2 * ll / l13;
  g2 = 0.12 * l * (1.0 + 17.0 * l - 9.0 * ll) / l13;
  g3 = (5.0 + 272.0 * l - 831.0 * ll + 1322.0 * l3 - 415.0 * l4) / (350.0 * l13);
  g4 = g1 * l;
  g5 = 0.12 * ll * (-4.0 + 17.0 * l - 4.0 * ll) / l13;
  g6 = l * (-65.0 + 832.0 * l - 1041.0 * ll + 832.0 * l3 - 65.0 * l4) / (350.0 * l13);
  fsums (fv, XM, 1, mmax, tt, fe, fo);
  two[16] = g1 * T4 + (g2 + g3 * T3) * ls5 - g3 + 1.0
    + rest (fe, mmax, 0, 2, s, g1, g2, g3, true, 1.0);
  two[17] = 8.0 / l13 * (g4 * T1 + (g5 + g6 * T3) * ls2 - g6 * s
                         + rest (fo, mmax, 1, 1, s, g4, g5, g6, true, 1.0));
  // YM
  g2 = 0.24 * l * (1.0 - l + 4.0 * ll) / l13;
  g3 = (24.0 - 201.0 * l + 882.0 * ll - 1182.0 * l3 + 591.0 * l4) / (625.0 * l13);
  g5 = 0.06 * ll * (7.0 - 10.0 * l + 7.0 * ll) / l13;
  g6 = 3.0 * l * (221.0 - 728.0 * l + 1902.0 * ll - 728.0 * l3 + 221.0 * l4) / (2500.0 * l13);
  fsums (fv, YM, 1, mmax, tt, fe, fo);
  two[18] = (g2 + g3 * T3) * ls5 - g3 + 1.0
    + rest (fe, mmax, 0, 2, s, 0.0, g2, g3, true, 1.0);
  two[19] = 8.0 / l13 * ((g5 + g6 * T3) * ls2 - g6 * s
                         + rest (fo, mmax, 1, 1, s, 0.0, g5, g6, true, 1.0));
  // ZM
  g3 = -0.3 * ll * (1.0 + ll) / l13;
  fsums (fv, ZM, 1, mmax, tt, fe, fo);
  two[20] = g3 * T3 * ls5 - g3 + 1.0 + rest (fe, mmax, 0, 2, s, 0.0, 0.0, g3, true, 1.0);
  two[21] = 8.0 / l13 * (-g3 * T3 * ls2 + g3 * s
                         + rest (fo, mmax, 1, 1, s, 0.0, 0.0, g3, true, -1.0));
}

// twobody_scale (sd = false: dimensional with radius a1, twobody.c:2709-2772)
// and twobody_scale_SD (sd = true, 2646-2700)
LUB_HD void scale (int version, double *two, double a1, double l, bool sd)
{
  const double l1 = 1.0 + l, pA = 0.5 * l1, pB = l1 * l1 / 6.0, pC = l1 * l1 * l1 / 6.0;
  const double pM = 5.0 / 6.0 * pC;
  const double a = sd ? 1.0 : a1, a2 = a * a, a3 = a2 * a;
  if (!sd) { two[0] *= a; two[2] *= a; }
  two[1] *= a * pA;
  two[3] *= a * pA;
  if (version == 0) return;
  two[4] *= a2 * (2.0 / 3.0); two[5] *= a2 * pB;
  two[6] *= a3 * (4.0 / 3.0); two[7] *= a3 * pC;
  two[8] *= a3 * (4.0 / 3.0); two[9] *= a3 * pC;
  if (version == 1) return;
  two[10] *= a2 * (2.0 / 3.0); two[11] *= a2 * pB;
  two[12] *= a2 * (2.0 / 3.0); two[13] *= a2 * pB;
  two[14] *= a3 * (4.0 / 3.0); two[15] *= a3 * pC;
  for (int k = 16; k < 22; k += 2)
    {
      two[k] *= a3 * (10.0 / 9.0);
      two[k + 1] *= a3 * pM;
    }
}

// powers of the radius dividing each function in the SD form
// (scalars_res_poly_scale_SD, minv-poly.c:565-660)
LUB_HD int sd_power (int f) { return (f == XA || f == YA) ? 1 : ((f == YB || f == XG || f == YG) ? 2 : 3); }

// ---------------------------------------------------------------------------
// far-field inverse
// ---------------------------------------------------------------------------
// in-place inverse of the n x n row-major a (n <= 6), Gauss-Jordan with
// partial pivoting; returns false for a singular matrix
LUB_HD bool invert (int n, double *a)
{
  int perm[6];
  for (int c = 0; c < n; c++) perm[c] = c;
  for (int c = 0; c < n; c++)
    {
      int p = c;
      double best = fabs (a[c * n + c]);
      for (int r = c + 1; r < n; r++)
        if (fabs (a[r * n + c]) > best) { best = fabs (a[r * n + c]); p = r; }
      if (best == 0.0) return false;
      if (p != c)
        {
          for (int k = 0; k < n; k++)
            {
              const double t = a[c * n + k];
              a[c * n + k] = a[p * n + k];
              a[p * n + k] = t;
            }
          const int t = perm[c]; perm[c] = perm[p]; perm[p] = t;
        }
      const double d = 1.0 / a[c * n + c];
      a[c * n + c] = 1.0;
      for (int k = 0; k < n; k++) a[c * n + k] *= d;
      for (int r = 0; r < n; r++)
        {
          if (r == c) continue;
          const double g = a[r * n + c];
          a[r * n + c] = 0.0;
          for (int k = 0; k < n; k++) a[r * n + k] -= g * a[c * n + k];
        }
    }
  // the row swaps permute the columns of the inverse
  for (int c = n - 1; c >= 0; c--)
    {
      int j = c;
      while (perm[j] != c) j++;
      if (j != c)
        {
          for (int r = 0; r < n; r++)
            {
              const double t = a[r * n + c];
              a[r * n + c] = a[r * n + j];
              a[r * n + j] = t;
            }
          const int t = perm[c]; perm[c] = perm[j]; perm[j] = t;
        }
    }
  return true;
}

// scalars_nonewald_poly_full (non-ewald.c:230-318): the dimensional RPY
// mobility of the pair, (11, 12, 21, 22) per function
LUB_HD void mobility44 (double r, double aa, double ab, double *m)
{
  for (int k = 0; k < 44; k++) m[k] = 0.0;
  const double r2 = r * r, ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir, ir4 = ir2 * ir2;
  const double a2 = aa * aa, b2 = ab * ab, sq = a2 + b2;
  const double ia = 1.0 / aa, ib = 1.0 / ab, ia3 = ia * ia * ia, ib3 = ib * ib * ib;
  m[0] = m[4] = ia;
  m[3] = m[7] = ib;
  m[1] = m[2] = (1.5 - 0.5 * sq / r2) * ir;
  m[5] = m[6] = (0.75 + 0.25 * sq / r2) * ir;
  m[9] = m[10] = -0.75 * ir2;
  m[12] = m[16] = 0.75 * ia3;
  m[15] = m[19] = 0.75 * ib3;
  m[13] = m[14] = 0.75 * ir3;
  m[17] = m[18] = -0.375 * ir3;
  const double qa = a2 / 10.0 + b2 / 6.0, qb = b2 / 10.0 + a2 / 6.0;
  m[21] = (2.25 - 13.5 * qa / r2) * ir2;
  m[22] = (2.25 - 13.5 * qb / r2) * ir2;
  m[25] = 4.5 * qa * ir4;
  m[26] = 4.5 * qb * ir4;
  m[29] = m[30] = -1.125 * ir3;
  m[32] = m[36] = m[40] = 0.9 * ia3;
  m[35] = m[39] = m[43] = 0.9 * ib3;
  m[33] = m[34] = (-4.5 + 5.4 * sq / r2) * ir3;
  m[37] = m[38] = (2.25 - 3.6 * sq / r2) * ir3;
  m[41] = m[42] = 0.9 * sq * ir4 * ir;
}

// scalars_minv_f/ft/fts_poly (minv-poly.c:39-560): out[44], dimensional.
// The pair problem separates by azimuthal order about the line of centres;
// each block below is the reciprocal two-sphere mobility in that mode, with
// the signs and factors the block forms give in the mode basis.
LUB_HD void minv (int version, double r, double aa, double ab, double *out)
{
  double m[44];
  mobility44 (r, aa, ab, m);
  for (int k = 0; k < 44; k++) out[k] = 0.0;
  auto inv2 = [&] (int f, int dst) {
    const double *q = m + 4 * f;
    const double det = q[0] * q[3] - q[1] * q[2];
    out[dst] = q[3] / det;
    out[dst + 1] = -q[1] / det;
    out[dst + 2] = -q[2] / det;
    out[dst + 3] = q[0] / det;
  };
  if (version == 0)
    {
      inv2 (XA, 0);
      inv2 (YA, 4);
      return;
    }
  if (version == 1)
    {
      inv2 (XA, 0);
      inv2 (XC, 12);
      // transverse (U1, U2, O1, O2)
      double a[16] = {m[4], m[5], -m[8], m[10],
                      m[6], m[7], -m[9], m[11],
                      -m[8], -m[9], m[16], m[17],
                      m[10], m[11], m[18], m[19]};
      invert (4, a);
      out[4] = a[0]; out[5] = a[1]; out[6] = a[4]; out[7] = a[5];
      out[8] = -a[8]; out[9] = -a[9]; out[10] = a[12]; out[11] = a[13];
      out[16] = a[10]; out[17] = a[11]; out[18] = a[14]; out[19] = a[15];
      return;
    }
  inv2 (XC, 12);
  inv2 (ZM, 40);
  {
    // transverse (U1, U2, O1, O2, S1, S2)
    const double *ya = m + 4, *yb = m + 8, *yc = m + 16, *yg = m + 24, *yh = m + 28,
                 *ym = m + 36;
    double a[36] = {ya[0],  ya[1],  -yb[0], yb[2], 2.0 * yg[0], -2.0 * yg[2],
                    ya[2],  ya[3],  -yb[1], yb[3], 2.0 * yg[1], -2.0 * yg[3],
                    -yb[0], -yb[1], yc[0],  yc[1], 2.0 * yh[0], 2.0 * yh[2],
                    yb[2],  yb[3],  yc[2],  yc[3], 2.0 * yh[1], 2.0 * yh[3],
                    yg[0],  yg[1],  yh[0],  yh[1], ym[0],       ym[1],
                    -yg[2], -yg[3], yh[2],  yh[3], ym[2],       ym[3]};
    invert (6, a);
    out[4] = a[0]; out[5] = a[1]; out[6] = a[6]; out[7] = a[7];
    out[8] = -a[12]; out[9] = -a[13]; out[10] = a[18]; out[11] = a[19];
    out[16] = a[14]; out[17] = a[15]; out[18] = a[20]; out[19] = a[21];
    out[24] = a[24]; out[25] = a[25]; out[26] = -a[30]; out[27] = -a[31];
    out[28] = a[26]; out[29] = a[27]; out[30] = a[32]; out[31] = a[33];
    out[36] = a[28]; out[37] = a[29]; out[38] = a[34]; out[39] = a[35];
  }
  {
    // axial: (U1, U2) and the axisymmetric strain pair (XM + ZM, XM - ZM)/2
    const double *xa = m, *xg = m + 20, *xm = m + 32, *zm = m + 40;
    double p[4], q[4];
    for (int k = 0; k < 4; k++)
      {
        p[k] = 0.5 * (xm[k] + zm[k]);
        q[k] = 0.5 * (xm[k] - zm[k]);
      }
    double a[36] = {xa[0],        xa[1],        -xg[0], xg[2], -xg[0], xg[2],
                    xa[2],        xa[3],        -xg[1], xg[3], -xg[1], xg[3],
                    -xg[0] / 3.0, -xg[1] / 3.0, p[0],   p[1],  q[0],   q[1],
                    xg[2] / 3.0,  xg[3] / 3.0,  p[2],   p[3],  q[2],   q[3],
                    -xg[0] / 3.0, -xg[1] / 3.0, q[0],   q[1],  p[0],   p[1],
                    xg[2] / 3.0,  xg[3] / 3.0,  q[2],   q[3],  p[2],   p[3]};
    invert (6, a);
    out[0] = a[0]; out[1] = a[1]; out[2] = a[6]; out[3] = a[7];
    out[20] = -3.0 * a[12]; out[21] = -3.0 * a[13];
    out[22] = 3.0 * a[18]; out[23] = 3.0 * a[19];
    out[32] = 2.0 * a[14] - out[40]; out[33] = 2.0 * a[15] - out[41];
    out[34] = 2.0 * a[20] - out[42]; out[35] = 2.0 * a[21] - out[43];
  }
}

// ---------------------------------------------------------------------------
// pair block applied to a vector
// ---------------------------------------------------------------------------
// y += B(e; s) T(x) with s = (xa ya yb xc yc xg yg yh xm ym zm), T the
// extracted-form transform of the strain part; self: the i == j signs
// (B-tilde = -B, G-tilde = +G^T; fts.c:73-111).  Rows and columns follow the
// reference's extracted 5-vector order (xx, xy, xz, yz, yy).
template <int D>
LUB_HD void apply_block (bool self, const double *x, double *y, double ex, double ey, double ez,
                         const double *s)
{
  const double e[3] = {ex, ey, ez};
  double z[11];
  for (int k = 0; k < D; k++) z[k] = x[k];
  if (D == 11)
    {
      z[6] = 2.0 * x[6] + x[10];
      z[7] = 2.0 * x[7];
      z[8] = 2.0 * x[8];
      z[9] = 2.0 * x[9];
      z[10] = 2.0 * x[10] + x[6];
    }
  // A: ya I + (xa - ya) e e
  {
    const double ed = e[0] * z[0] + e[1] * z[1] + e[2] * z[2];
    for (int p = 0; p < 3; p++) y[p] += s[1] * z[p] + (s[0] - s[1]) * e[p] * ed;
  }
  if (D == 3) return;
  // B (omega <- F): yb eps_pqk e_k; B-tilde (U <- T) with the block's sign
  {
    const double b = s[2], tb = self ? -b : b;
    // (eps . e) v = e x v with the sign of the block's layout:
    //   (B v)_p = b * (v x e)_p
    const double *F = z, *T = z + 3;
    y[3] += b * (F[1] * e[2] - F[2] * e[1]);
    y[4] += b * (F[2] * e[0] - F[0] * e[2]);
    y[5] += b * (F[0] * e[1] - F[1] * e[0]);
    y[0] += tb * (T[1] * e[2] - T[2] * e[1]);
    y[1] += tb * (T[2] * e[0] - T[0] * e[2]);
    y[2] += tb * (T[0] * e[1] - T[1] * e[0]);
    // C: yc I + (xc - yc) e e
    const double ed = e[0] * T[0] + e[1] * T[1] + e[2] * T[2];
    for (int p = 0; p < 3; p++) y[3 + p] += s[4] * T[p] + (s[3] - s[4]) * e[p] * ed;
  }
  if (D == 6) return;
  {
    const int P[5] = {0, 0, 0, 1, 1}, Q[5] = {0, 1, 2, 2, 1};
    const double xg = s[5], yg = s[6], h = s[7];
    const double gsign = self ? 1.0 : -1.0;
    const double *F = z, *T = z + 3, *S = z + 6;
    // G_{r,k}: rows the strain components, columns F
    for (int r = 0; r < 5; r++)
      {
        const int p = P[r], q = Q[r];
        double gF = 0.0, hT = 0.0;
        for (int k = 0; k < 3; k++)
          {
            const double g = -xg / 3.0 * (p == q) * e[k] + yg * (e[p] * (q == k) + e[q] * (p == k))
              + (xg - 2.0 * yg) * e[p] * e[q] * e[k];
            gF += g * F[k];
            y[k] += gsign * g * S[r];
          }
        y[6 + r] += gF;
        (void) hT;
      }
    // H (strain <- omega) and H^T (omega <- strain)
    const double hxx = h * e[0] * e[0], hxy = h * e[0] * e[1], hxz = h * e[0] * e[2];
    const double hyy = h * e[1] * e[1], hyz = h * e[1] * e[2], hzz = h * e[2] * e[2];
    const double H[5][3] = {{0.0, 2.0 * hxz, -2.0 * hxy},
                            {-hxz, hyz, hxx - hyy},
                            {hxy, -hxx + hzz, -hyz},
                            {hyy - hzz, -hxy, hxz},
                            {-2.0 * hyz, 0.0, 2.0 * hxy}};
    for (int r = 0; r < 5; r++)
      for (int k = 0; k < 3; k++)
        {
          y[6 + r] += H[r][k] * T[k];
          y[3 + k] += H[r][k] * S[r];
        }
    // M
    const double xm = s[8], ym = s[9], zm = s[10];
    const double m1 = 1.5 * xm - 2.0 * ym + 0.5 * zm, m2 = -0.5 * xm + 0.5 * zm;
    const double m3 = xm / 6.0 - 0.5 * zm, m4 = 0.5 * ym - 0.5 * zm, m5 = 0.5 * zm;
    for (int r = 0; r < 5; r++)
      {
        const int p = P[r], q = Q[r];
        double acc = 0.0;
        for (int c = 0; c < 5; c++)
          {
            const int t = P[c], u = Q[c];
            const double dpq = p == q, dtu = t == u, dqt = q == t, dqu = q == u;
            const double dpt = p == t, dpu = p == u;
            acc += (m1 * e[p] * e[q] * e[t] * e[u] + m2 * (dpq * e[t] * e[u] + e[p] * e[q] * dtu)
                    + m3 * dpq * dtu
                    + m4 * (e[p] * dqt * e[u] + e[p] * dqu * e[t] + e[q] * dpt * e[u]
                            + e[q] * dpu * e[t])
                    + m5 * (dpt * dqu + dqt * dpu)) * S[c];
          }
        y[6 + r] += acc;
      }
  }
}

// the raw block B(e; s) (rows: U, O, E components; columns: F, T, S in the
// extracted form), as matrix_fts_ij / matrix_lub_fts_2b write it: y += B
// T(x) is apply_block.  self: the i == j signs.
template <int D>
LUB_HD void block_dense (bool self, double ex, double ey, double ez, const double *s,
                         double *B /* [D * D] row-major */)
{
  for (int k = 0; k < D * D; k++) B[k] = 0.0;
  const double e[3] = {ex, ey, ez};
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++)
      B[p * D + q] = (p == q ? s[1] : 0.0) + (s[0] - s[1]) * e[p] * e[q];
  if (D == 3) return;
  // (B v)_p = b (v x e)_p: row p, column q of the cross-product matrix
  const double X[3][3] = {{0.0, e[2], -e[1]}, {-e[2], 0.0, e[0]}, {e[1], -e[0], 0.0}};
  const double tb = self ? -s[2] : s[2];
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++)
      {
        B[(3 + p) * D + q] += s[2] * X[p][q];
        B[p * D + 3 + q] += tb * X[p][q];
        B[(3 + p) * D + 3 + q] += (p == q ? s[4] : 0.0) + (s[3] - s[4]) * e[p] * e[q];
      }
  if (D == 6) return;
  const int P[5] = {0, 0, 0, 1, 1}, Q[5] = {0, 1, 2, 2, 1};
  const double xg = s[5], yg = s[6], h = s[7];
  const double gsign = self ? 1.0 : -1.0;
  for (int r = 0; r < 5; r++)
    for (int k = 0; k < 3; k++)
      {
        const int p = P[r], q = Q[r];
        const double g = -xg / 3.0 * (p == q) * e[k] + yg * (e[p] * (q == k) + e[q] * (p == k))
          + (xg - 2.0 * yg) * e[p] * e[q] * e[k];
        B[(6 + r) * D + k] += g;
        B[k * D + 6 + r] += gsign * g;
      }
  const double hxx = h * e[0] * e[0], hxy = h * e[0] * e[1], hxz = h * e[0] * e[2];
  const double hyy = h * e[1] * e[1], hyz = h * e[1] * e[2], hzz = h * e[2] * e[2];
  const double H[5][3] = {{0.0, 2.0 * hxz, -2.0 * hxy},
                          {-hxz, hyz, hxx - hyy},
                          {hxy, -hxx + hzz, -hyz},
                          {hyy - hzz, -hxy, hxz},
                          {-2.0 * hyz, 0.0, 2.0 * hxy}};
  for (int r = 0; r < 5; r++)
    for (int k = 0; k < 3; k++)
      {
        B[(6 + r) * D + 3 + k] += H[r][k];
        B[(3 + k) * D + 6 + r] += H[r][k];
      }
  const double xm = s[8], ym = s[9], zm = s[10];
  const double m1 = 1.5 * xm - 2.0 * ym + 0.5 * zm, m2 = -0.5 * xm + 0.5 * zm;
  const double m3 = xm / 6.0 - 0.5 * zm, m4 = 0.5 * ym - 0.5 * zm, m5 = 0.5 * zm;
  for (int r = 0; r < 5; r++)
    for (int c = 0; c < 5; c++)
      {
        const int p = P[r], q = Q[r], t = P[c], u = Q[c];
        const double dpq = p == q, dtu = t == u, dqt = q == t, dqu = q == u;
        const double dpt = p == t, dpu = p == u;
        B[(6 + r) * D + 6 + c] += m1 * e[p] * e[q] * e[t] * e[u]
          + m2 * (dpq * e[t] * e[u] + e[p] * e[q] * dtu) + m3 * dpq * dtu
          + m4 * (e[p] * dqt * e[u] + e[p] * dqu * e[t] + e[q] * dpt * e[u] + e[q] * dpu * e[t])
          + m5 * (dpt * dqu + dqt * dpu);
      }
}

}  // namespace lub
}  // namespace shk
