// pair_math.h -- Ewald/RPY pair scalars and the image-summed pair-moment
// representation of the 11x11 FTS mobility block, shared by every kernel
// (real-space matvec, dense build, adjoint) in FP64 and FP32.
//
// Scalar functions: restated from ewald.c:48-186 (scalars_ewald_real),
// ewald.c:203-359 (scalars_ewald_real_poly), ewald.c:382-420 (SD scaling),
// non-ewald.c:49-90 / 110-145 (open-BC RPY).  Scalar order everywhere:
//   s[0..10] = xa ya yb xc yc xg yg yh xm ym zm.
//
// Moment representation.  The pair block of fts.c:45-610 is linear in a
// small set of tensors of the direction e, each multiplied by a scalar
// function of r.  Summing a pair's periodic images therefore reduces to
// accumulating these tensor moments (68 for FTS, 7 for F, 17 for FT); the
// 11x11 block is assembled once per pair from the moments and then applied
// to as many right-hand sides as the launch carries.
#pragma once
#include <hip/hip_runtime.h>

#define SHK_HD __host__ __device__ __forceinline__

namespace shk {

template <int V> struct Dim;
template <> struct Dim<0> { static constexpr int D = 3;  static constexpr int NMOM = 7;  };
template <> struct Dim<1> { static constexpr int D = 6;  static constexpr int NMOM = 17; };
template <> struct Dim<2> { static constexpr int D = 11; static constexpr int NMOM = 68; };

// ---------------------------------------------------------------------------
// math wrappers (device libm / host libm)
// ---------------------------------------------------------------------------
SHK_HD double m_erfc (double x) { return erfc (x); }
SHK_HD float  m_erfc (float x)  { return erfcf (x); }
SHK_HD double m_exp (double x)  { return exp (x); }
SHK_HD float  m_exp (float x)   { return expf (x); }
SHK_HD double m_sqrt (double x) { return sqrt (x); }
SHK_HD float  m_sqrt (float x)  { return sqrtf (x); }
SHK_HD double m_rsqrt (double x) { return 1.0 / sqrt (x); }
SHK_HD float  m_rsqrt (float x)  { return rsqrtf (x); }

template <typename T> SHK_HD T inv_sqrt_pi () { return T (0.56418958354775628695); }

// erfc(x) and exp(-x2) (x2 = x^2, x >= 0).  FP64: libm.  FP32 on the device:
// exp by the hardware exp2, erfc by the Chebyshev fit of Numerical Recipes
// in C (2nd ed., sec. 6.2, erfcc: fractional error below 1.2e-7 for x >= 0),
// erfc(x) = t exp(-x^2) exp(P(t)), t = 1 / (1 + x/2), sharing exp(-x^2).
SHK_HD void erfc_exp (double x, double x2, double &ec, double &e)
{
  ec = erfc (x);
  e = exp (-x2);
}
SHK_HD void erfc_exp (float x, float x2, float &ec, float &e)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const float l2e = 1.44269504088896341f;
  e = __builtin_amdgcn_exp2f (-x2 * l2e);
  const float t = __builtin_amdgcn_rcpf (1.f + 0.5f * x);
  const float pt = -1.26551223f + t * (1.00002368f + t * (0.37409196f + t * (0.09678418f
    + t * (-0.18628806f + t * (0.27886807f + t * (-1.13520398f + t * (1.48851587f
    + t * (-0.82215223f + t * 0.17087277f))))))));
  ec = t * e * __builtin_amdgcn_exp2f (pt * l2e);
#else
  ec = erfcf (x);
  e = expf (-x2);
#endif
}

// ---------------------------------------------------------------------------
// real-space Ewald scalars, monodisperse (ewald.c:48-186)
// ---------------------------------------------------------------------------
// `u` is 1/r, supplied by the caller (the pair kernels have it already):
// every division of the reference's expressions becomes a product with a
// power of u, which keeps the scalar evaluation to one reciprocal.
template <int V, typename T>
SHK_HD void scalars_real_u (T xi, T r, T u, T *o)
{
  const T xr = xi * r, x2 = xr * xr;
  const T u2 = u * u, u3 = u2 * u;
  T ec, ex;
  erfc_exp (xr, x2, ec, ex);
  ex *= xi * inv_sqrt_pi<T> ();
  o[1] = (T (0.75) + T (0.5) * u2) * u * ec
    + ((T (1) + x2 * (T (14) + T (4) * x2 * (T (-5) + x2))) * u2 - T (4.5) + T (3) * x2) * ex;
  const T a2 = (T (0.75) - T (1.5) * u2) * u * ec
    + ((T (-3) + x2 * (T (-2) + T (4) * x2 * (T (4) - x2))) * u2 + T (1.5) - T (3) * x2) * ex;
  o[0] = a2 + o[1];
  if (V == 0) return;
  o[2] = T (-0.75) * u2 * ec - T (1.5) * (T (1) + x2 * (T (-6) + x2 * T (2))) * u * ex;
  o[4] = T (-0.375) * u3 * ec
    - T (0.75) * (T (1) + x2 * (T (14) + x2 * (T (-20) + x2 * T (4)))) * u2 * ex;
  const T c2 = T (1.125) * u3 * ec
    - T (0.75) * (T (-3) + x2 * (T (-2) + x2 * (T (16) + x2 * T (-4)))) * u2 * ex;
  o[3] = c2 + o[4];
  if (V == 1) return;
  o[5] = (T (2.25) - T (3.6) * u2) * u2 * ec
    + (T (-1.5) * (T (-3) + x2 * T (6))
       - T (0.8) * (T (9) + x2 * (T (6) + x2 * (T (-48) + x2 * T (12)))) * u2) * u * ex;
  o[6] = T (1.2) * u2 * u2 * ec
    + (T (-3) * (x2 * (T (2) - x2))
       - T (0.8) * (T (-3) + x2 * (T (-2) + x2 * (T (-26) + x2 * (T (26) + x2 * T (-4))))) * u2)
      * u * ex;
  o[7] = T (-1.125) * u3 * ec
    + T (1.5) * (T (-1.5) + x2 * (T (-1) + x2 * (T (8) + x2 * T (-2)))) * u2 * ex;
  o[8] = (T (-4.5) + T (10.8) * u2) * u3 * ec
    + (T (1.5) * (T (-6) + x2 * (T (-12) + x2 * T (12)))
       + T (1.2) * (T (18) + x2 * (T (12) + x2 * (T (30) + x2 * (T (-66) + x2 * T (12))))) * u2)
      * u2 * ex;
  o[9] = (T (2.25) - T (7.2) * u2) * u3 * ec
    + (T (-1.5) * (T (-3) + x2 * (T (6) + x2 * (T (-12) + x2 * T (4))))
       - T (1.2) * (T (12) + x2 * (T (8) + x2 * (T (-22) + x2 * (T (58) + x2 * (T (-34) + x2 * T (4))))))
         * u2) * u2 * ex;
  o[10] = T (1.8) * u2 * u3 * ec
    + (T (-1.5) * (x2 * (T (8) + x2 * T (-4)))
       - T (1.2) * (T (-3) + x2 * (T (-2) + x2 * (T (-26) + x2 * (T (26) + x2 * T (-4))))) * u2)
      * u2 * ex;
}

template <int V, typename T>
SHK_HD void scalars_real (T xi, T r, T *o)
{
  scalars_real_u<V, T> (xi, r, T (1) / r, o);
}

// F-version real-space scalars xa, ya (ewald.c:80-99) and their radial
// derivatives, for the adjoint; `u` is 1/r.  With z = (xi r)^2,
// E = xi/sqrt(pi) exp(-z), d erfc(xi r)/dr = -2E, dE/dr = -2 xi^2 r E:
//   ya = (3/4 u + 1/2 u^3) erfc + (Pa(z) u^2 - 4.5 + 3z) E,
//   a2 = (3/4 u - 3/2 u^3) erfc + (Pb(z) u^2 + 1.5 - 3z) E,  xa = a2 + ya,
// Pa = 1 + 14z - 20z^2 + 4z^3, Pb = -3 - 2z + 16z^2 - 4z^3.
template <typename T>
SHK_HD void scalars_real_f_deriv (T xi, T r, T u, T *xa, T *ya, T *dxa, T *dya)
{
  const T xr = xi * r, z = xr * xr;
  const T u2 = u * u, u3 = u2 * u, u4 = u2 * u2;
  const T ec = m_erfc (xr);
  const T E = xi * inv_sqrt_pi<T> () * m_exp (-z);
  const T dz = T (2) * xi * xi * r;                 // dz/dr
  const T pa = T (1) + z * (T (14) + z * (T (-20) + z * T (4)));
  const T dpa = T (14) + z * (T (-40) + z * T (12));
  const T pb = T (-3) + z * (T (-2) + z * (T (16) + z * T (-4)));
  const T dpb = T (-2) + z * (T (32) + z * T (-12));
  const T ca = T (0.75) * u + T (0.5) * u3;          // erfc coefficients
  const T cb = T (0.75) * u - T (1.5) * u3;
  const T qa = pa * u2 - T (4.5) + T (3) * z;        // E coefficients
  const T qb = pb * u2 + T (1.5) - T (3) * z;
  const T ya_ = ca * ec + qa * E;
  const T a2 = cb * ec + qb * E;
  const T dca = T (-0.75) * u2 - T (1.5) * u4;
  const T dcb = T (-0.75) * u2 + T (4.5) * u4;
  const T dqa = dpa * dz * u2 - T (2) * pa * u3 + T (3) * dz;
  const T dqb = dpb * dz * u2 - T (2) * pb * u3 - T (3) * dz;
  const T dya_ = dca * ec - T (2) * ca * E + (dqa - qa * dz) * E;
  const T da2 = dcb * ec - T (2) * cb * E + (dqb - qb * dz) * E;
  *ya = ya_;
  *xa = a2 + ya_;
  *dya = dya_;
  *dxa = da2 + dya_;
}

// ---------------------------------------------------------------------------
// real-space Ewald scalars, unequal spheres: ewald.c:203-359, followed by
// the SD scaling of ewald.c:382-420 when SD is true
// ---------------------------------------------------------------------------
// `u` is 1/r (one reciprocal per pair-image, as in scalars_real_u)
template <int V, typename T, bool SD = true>
SHK_HD void scalars_real_poly_u (T xi, T r, T u, T aa, T ab, T *o)
{
  const T sa = SD ? aa : T (1);
  const T x = xi * r, x2 = x * x;
  const T u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u3 * u2;
  const T aa2 = aa * aa, ab2 = ab * ab;
  T ec, ex;
  erfc_exp (x, x2, ec, ex);
  ex *= xi * inv_sqrt_pi<T> ();
  const T sixth = T (1.0 / 6.0), tenth = T (0.1);
  const T a0 = T (2) * (T (-3) + x2 * T (2)) * ex + ec * u;
  const T b0 = T (2) * (T (1) + x2 * T (-2)) * ex + ec * u;
  const T a2 = T (4) * u2 * (T (1) + x2 * (T (14) + x2 * (T (-20) + x2 * T (4)))) * ex + T (2) * u3 * ec;
  const T b2 = T (-4) * u2 * (T (3) + x2 * (T (2) + x2 * (T (-16) + x2 * T (4)))) * ex - T (6) * u3 * ec;
  o[1] = T (0.75) * (a0 + a2 * (aa2 + ab2) * sixth);
  o[0] = T (0.75) * (b0 + b2 * (aa2 + ab2) * sixth) + o[1];
  o[0] *= sa; o[1] *= sa;
  if (V == 0) return;
  const T ad0 = T (-2) * u * (T (1) + x2 * (T (-10) + x2 * T (4))) * ex - ec * u2;
  const T add0 = T (4) * u2 * (T (1) + x2 * (T (6) + x2 * (T (-16) + x2 * T (4)))) * ex + T (2) * u3 * ec;
  const T bd0 = T (-2) * u * (T (1) + x2 * (T (6) + x2 * T (-4))) * ex - ec * u2;
  const T sa2 = sa * sa, sa3 = sa2 * sa;
  const T b0u = b0 * u;
  o[2] = T (-0.375) * (-ad0 + b0u) * sa2;
  const T c1 = T (0.1875) * (-add0 + (-ad0 + bd0) * u);
  const T c2 = T (0.1875) * (add0 + (-ad0 - bd0 + T (2) * b0u) * u);
  o[4] = c1 * sa3;
  o[3] = (c2 + c1) * sa3;
  if (V == 1) return;
  const T ad2 = T (-4) * u3 * (T (3) + x2 * (T (2) + x2 * (T (68) + x2 * (T (-56) + x2 * T (8))))) * ex
    - T (6) * u4 * ec;
  const T add2 = T (16) * u4
    * (T (3) + x2 * (T (2) + x2 * (T (-16) + x2 * (T (76) + x2 * (T (-38) + x2 * T (4)))))) * ex
    + T (24) * u5 * ec;
  const T bd2 = T (4) * u3 * (T (9) + x2 * (T (6) + x2 * (T (36) + x2 * (T (-48) + x2 * T (8))))) * ex
    + T (18) * u4 * ec;
  const T b2u = b2 * u;
  const T wg = aa2 * tenth + ab2 * sixth;
  const T g10 = T (-0.75) * b0u;
  const T g12 = T (-0.75) * b2u * wg;
  const T g20 = T (-0.375) * (ad0 + b0u);
  const T g22 = T (-0.375) * (ad2 + b2u) * wg;
  o[5] = T (-3) * (g10 + g12) * sa2;
  o[6] = (g20 + g22) * sa2;
  o[7] = T (-0.1875) * (add0 + (-ad0 - bd0 + T (2) * b0u) * u) * sa3;
  const T wm = (aa2 + ab2) * tenth;
  const T m50 = T (-0.375) * (ad0 + b0u) * u;
  const T m52 = T (-0.375) * (ad2 + b2u) * u * wm;
  const T zm = T (2) * (m50 + m52);
  const T m30 = T (-0.75) * b0 * u2;
  const T m32 = T (-0.75) * b2 * u2 * wm;
  const T xm = T (6) * (m30 + m32) + T (3) * zm;
  const T m40 = T (-0.1875) * (add0 + (-ad0 + T (3) * bd0 - T (6) * b0u) * u);
  const T m42 = T (-0.1875) * (add2 + (-ad2 + T (3) * bd2 - T (6) * b2u) * u) * wm;
  const T ym = T (2) * (m40 + m42) + zm;
  o[8] = xm * sa3; o[9] = ym * sa3; o[10] = zm * sa3;
}

template <int V, typename T, bool SD = true>
SHK_HD void scalars_real_poly_sd (T xi, T r, T aa, T ab, T *o)
{
  scalars_real_poly_u<V, T, SD> (xi, r, T (1) / r, aa, ab, o);
}

// open-BC RPY scalars, SD form (non-ewald.c:49-90, 110-145 + scaling)
template <int V, typename T>
SHK_HD void scalars_rpy_sd (T r, T aa, T ab, bool poly, T *o)
{
  T r2 = r * r;
  T s2 = poly ? (aa * aa + ab * ab) : T (2);
  T wg = poly ? (aa * aa / T (10) + ab * ab / T (6)) : T (4) / T (15);
  T sa = poly ? aa : T (1);
  o[0] = (T (1.5) - s2 * T (0.5) / r2) / r * sa;
  o[1] = (T (0.75) + s2 * T (0.25) / r2) / r * sa;
  if (V == 0) return;
  T sa2 = sa * sa, sa3 = sa2 * sa;
  o[2] = T (-0.75) / r2 * sa2;
  o[3] = T (0.75) / r2 / r * sa3;
  o[4] = T (-0.375) / r2 / r * sa3;
  if (V == 1) return;
  o[5] = (T (2.25) - T (13.5) * wg / r2) / r2 * sa2;
  o[6] = T (4.5) * wg / r2 / r2 * sa2;
  o[7] = T (-1.125) / r2 / r * sa3;
  o[8] = (T (-4.5) + T (5.4) * s2 / r2) / r2 / r * sa3;
  o[9] = (T (2.25) - T (3.6) * s2 / r2) / r2 / r * sa3;
  o[10] = T (0.9) * s2 / r2 / r2 / r * sa3;
}

// ---------------------------------------------------------------------------
// pair moments
// ---------------------------------------------------------------------------
// Layout of the moment vector (FTS; F uses [0,7), FT uses [0,17)):
//   0      SA   = sum a1                      a1 = ya, a2 = xa - ya
//   1..6   SAe  = sum a2 e(x)e   (xx yy zz xy yz xz)
//   7..9   SB   = sum b1 e                    b1 = yb
//   10     SC   = sum c1                      c1 = yc, c2 = xc - yc
//   11..16 SCe  = sum c2 e(x)e
//   17..19 SG1  = sum g1 e                    g1 = -xg/3
//   20..22 SG2  = sum g2 e                    g2 = yg
//   23..32 SG3  = sum g3 eee  (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz)
//   33..38 SH   = sum h1 e(x)e                h1 = yh
//   39..53 SM1  = sum mm1 eeee (xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz
//                                xyzz xzzz yyyy yyyz yyzz yzzz zzzz)
//   54..59 SM2  = sum mm2 e(x)e
//   60     SM3  = sum mm3
//   61..66 SM4  = sum mm4 e(x)e
//   67     SM5  = sum mm5
// with g3 = xg - 2yg and mm1..mm5 of fts.c:705-709.
enum {
  M_SA = 0, M_SAe = 1, M_SB = 7, M_SC = 10, M_SCe = 11, M_SG1 = 17, M_SG2 = 20,
  M_SG3 = 23, M_SH = 33, M_SM1 = 39, M_SM2 = 54, M_SM3 = 60, M_SM4 = 61, M_SM5 = 67
};

// index of symmetric pair (p,q) in (xx yy zz xy yz xz)
SHK_HD constexpr int sym2 (int p, int q)
{
  return p == q ? p : ((p + q == 1) ? 3 : ((p + q == 3) ? 4 : 5));
}
// index of sorted triple in (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz)
SHK_HD constexpr int sym3s (int a, int b, int c)
{
  // a <= b <= c
  return a == 0 ? (b == 0 ? c : (b == 1 ? 2 + c : 5))
                : (a == 1 ? (b == 1 ? 5 + c : 8) : 9);
}
SHK_HD constexpr int sym3 (int p, int q, int k)
{
  int a = p < q ? p : q, b = p < q ? q : p;
  // insert k
  return k <= a ? sym3s (k, a, b) : (k <= b ? sym3s (a, k, b) : sym3s (a, b, k));
}
// index of sorted quadruple in (xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz
//                                xyzz xzzz yyyy yyyz yyzz yzzz zzzz)
SHK_HD constexpr int sym4s (int a, int b, int c, int d)
{
  // count of y and z
  return 0 * a + 0 * b
    + ((a == 0 && b == 0 && c == 0) ? d
       : (a == 0 && b == 0) ? (c == 1 ? 2 + d : 5)
       : (a == 0) ? (b == 1 ? (c == 1 ? 5 + d : 8) : 9)
       : (a == 1 && b == 1 && c == 1) ? 9 + d
       : (a == 1 && b == 1) ? 12
       : (a == 1) ? 13 : 14);
}
SHK_HD constexpr int sort_min (int a, int b) { return a < b ? a : b; }
SHK_HD constexpr int sort_max (int a, int b) { return a < b ? b : a; }
SHK_HD constexpr int sym4 (int p, int q, int t, int u)
{
  // sorting network for 4 elements
  return sym4s (sort_min (sort_min (p, q), sort_min (t, u)),
                sort_min (sort_max (sort_min (p, q), sort_min (t, u)),
                          sort_min (sort_max (p, q), sort_max (t, u))),
                sort_max (sort_max (sort_min (p, q), sort_min (t, u)),
                          sort_min (sort_max (p, q), sort_max (t, u))),
                sort_max (sort_max (p, q), sort_max (t, u)));
}

// extracted-form index r -> (p,q): xx xy xz yz yy
SHK_HD constexpr int EP (int r) { return r < 3 ? 0 : 1; }
SHK_HD constexpr int EQ (int r) { return r == 0 ? 0 : (r == 1 ? 1 : (r == 2 ? 2 : (r == 3 ? 2 : 1))); }

// Accumulate one contribution with direction (ex,ey,ez) and SD scalars s.
template <int V, typename T>
SHK_HD void moments_add (T *M, T ex, T ey, T ez, const T *s)
{
  T e2[6] = {ex * ex, ey * ey, ez * ez, ex * ey, ey * ez, ez * ex};
  T a1 = s[1], a2 = s[0] - s[1];
  M[M_SA] += a1;
#pragma unroll
  for (int k = 0; k < 6; k++) M[M_SAe + k] += a2 * e2[k];
  if (V == 0) return;
  T b1 = s[2];
  M[M_SB + 0] += b1 * ex; M[M_SB + 1] += b1 * ey; M[M_SB + 2] += b1 * ez;
  T c1 = s[4], c2 = s[3] - s[4];
  M[M_SC] += c1;
#pragma unroll
  for (int k = 0; k < 6; k++) M[M_SCe + k] += c2 * e2[k];
  if (V == 1) return;
  T g1 = T (-1.0 / 3.0) * s[5], g2 = s[6], g3 = s[5] - T (2) * s[6];
  M[M_SG1 + 0] += g1 * ex; M[M_SG1 + 1] += g1 * ey; M[M_SG1 + 2] += g1 * ez;
  M[M_SG2 + 0] += g2 * ex; M[M_SG2 + 1] += g2 * ey; M[M_SG2 + 2] += g2 * ez;
  T g3x = g3 * ex, g3y = g3 * ey, g3z = g3 * ez;
  M[M_SG3 + 0] += g3x * e2[0];  // xxx
  M[M_SG3 + 1] += g3x * e2[3];  // xxy
  M[M_SG3 + 2] += g3x * e2[5];  // xxz
  M[M_SG3 + 3] += g3x * e2[1];  // xyy
  M[M_SG3 + 4] += g3x * e2[4];  // xyz
  M[M_SG3 + 5] += g3x * e2[2];  // xzz
  M[M_SG3 + 6] += g3y * e2[1];  // yyy
  M[M_SG3 + 7] += g3y * e2[4];  // yyz
  M[M_SG3 + 8] += g3y * e2[2];  // yzz
  M[M_SG3 + 9] += g3z * e2[2];  // zzz
  T h1 = s[7];
#pragma unroll
  for (int k = 0; k < 6; k++) M[M_SH + k] += h1 * e2[k];
  T xm = s[8], ym = s[9], zm = s[10];
  T mm1 = T (1.5) * xm - T (2) * ym + T (0.5) * zm;
  T mm2 = T (-0.5) * xm + T (0.5) * zm;
  T mm3 = T (1.0 / 6.0) * xm - T (0.5) * zm;
  T mm4 = T (0.5) * ym - T (0.5) * zm;
  T mm5 = T (0.5) * zm;
  T m1xx = mm1 * e2[0], m1xy = mm1 * e2[3], m1xz = mm1 * e2[5];
  T m1yy = mm1 * e2[1], m1yz = mm1 * e2[4], m1zz = mm1 * e2[2];
  M[M_SM1 + 0] += m1xx * e2[0];   // xxxx
  M[M_SM1 + 1] += m1xx * e2[3];   // xxxy
  M[M_SM1 + 2] += m1xx * e2[5];   // xxxz
  M[M_SM1 + 3] += m1xx * e2[1];   // xxyy
  M[M_SM1 + 4] += m1xx * e2[4];   // xxyz
  M[M_SM1 + 5] += m1xx * e2[2];   // xxzz
  M[M_SM1 + 6] += m1xy * e2[1];   // xyyy
  M[M_SM1 + 7] += m1xy * e2[4];   // xyyz
  M[M_SM1 + 8] += m1xy * e2[2];   // xyzz
  M[M_SM1 + 9] += m1xz * e2[2];   // xzzz
  M[M_SM1 + 10] += m1yy * e2[1];  // yyyy
  M[M_SM1 + 11] += m1yy * e2[4];  // yyyz
  M[M_SM1 + 12] += m1yy * e2[2];  // yyzz
  M[M_SM1 + 13] += m1yz * e2[2];  // yzzz
  M[M_SM1 + 14] += m1zz * e2[2];  // zzzz
  (void) m1xz; (void) m1yz;
#pragma unroll
  for (int k = 0; k < 6; k++) M[M_SM2 + k] += mm2 * e2[k];
  M[M_SM3] += mm3;
#pragma unroll
  for (int k = 0; k < 6; k++) M[M_SM4 + k] += mm4 * e2[k];
  M[M_SM5] += mm5;
}

// Block entry (row p, column q) from moments, non-self convention of the
// ATIMES routines (fts.c:625-1010); `self` selects the i == j tilde rules of
// fts.c:73-111 (only the dense matrix uses them).  Rows/columns: 0-2 U/F,
// 3-5 omega/T, 6-10 E/S in extracted order (xx xy xz yz yy); S columns act on
// the transformed vector (fts.c:660-670).
template <int V, typename T, typename MA>
SHK_HD T block_entry (const MA &M, int p, int q, bool self)
{
  // A: 0-2 x 0-2
  if (p < 3 && q < 3)
    return (p == q ? M[M_SA] : T (0)) + M[M_SAe + sym2 (p, q)];
  if (V == 0) return T (0);
  // B(omega,F) = eps_pqk SB_k ; B-tilde(U,T) = +/- same pattern
  if (p >= 3 && p < 6 && q < 3)
    {
      int a = p - 3, b = q;
      if (a == b) return T (0);
      int k = 3 - a - b;
      T sgn = ((a + 1) % 3 == b) ? T (1) : T (-1);
      return sgn * M[M_SB + k];
    }
  if (p < 3 && q >= 3 && q < 6)
    {
      int a = p, b = q - 3;
      if (a == b) return T (0);
      int k = 3 - a - b;
      T sgn = ((a + 1) % 3 == b) ? T (1) : T (-1);
      return (self ? -sgn : sgn) * M[M_SB + k];
    }
  if (p >= 3 && p < 6 && q >= 3 && q < 6)
    return (p == q ? M[M_SC] : T (0)) + M[M_SCe + sym2 (p - 3, q - 3)];
  if (V == 1) return T (0);
  // G(E,F): G_{rk} = SG1_k d_pq + SG2_p d_qk + SG2_q d_pk + SG3_pqk
  if (p >= 6 && q < 3)
    {
      int r = p - 6, k = q, a = EP (r), b = EQ (r);
      return (a == b ? M[M_SG1 + k] : T (0)) + (b == k ? M[M_SG2 + a] : T (0))
        + (a == k ? M[M_SG2 + b] : T (0)) + M[M_SG3 + sym3 (a, b, k)];
    }
  if (p < 3 && q >= 6)
    {
      int r = q - 6, k = p, a = EP (r), b = EQ (r);
      T g = (a == b ? M[M_SG1 + k] : T (0)) + (b == k ? M[M_SG2 + a] : T (0))
        + (a == k ? M[M_SG2 + b] : T (0)) + M[M_SG3 + sym3 (a, b, k)];
      return self ? g : -g;
    }
  // H(E,T): h1 (e_p eps_qkl e_l + e_q eps_pkl e_l)
  if ((p >= 6 && q >= 3 && q < 6) || (p >= 3 && p < 6 && q >= 6))
    {
      int r = p >= 6 ? p - 6 : q - 6;
      int k = p >= 6 ? q - 3 : p - 3;
      int a = EP (r), b = EQ (r);
      T h = T (0);
#pragma unroll
      for (int l = 0; l < 3; l++)
        {
          // eps_{b k l}
          if (b != k && k != l && b != l)
            {
              T e = ((b + 1) % 3 == k) ? T (1) : T (-1);
              h += e * M[M_SH + sym2 (a, l)];
            }
          if (a != k && k != l && a != l)
            {
              T e = ((a + 1) % 3 == k) ? T (1) : T (-1);
              h += e * M[M_SH + sym2 (b, l)];
            }
        }
      return h;
    }
  // M(E,S)
  {
    int r = p - 6, c = q - 6;
    int a = EP (r), b = EQ (r), t = EP (c), u = EQ (c);
    T m = M[M_SM1 + sym4 (a, b, t, u)];
    if (a == b) m += M[M_SM2 + sym2 (t, u)];
    if (t == u) m += M[M_SM2 + sym2 (a, b)];
    if (a == b && t == u) m += M[M_SM3];
    if (b == t) m += M[M_SM4 + sym2 (a, u)];
    if (b == u) m += M[M_SM4 + sym2 (a, t)];
    if (a == t) m += M[M_SM4 + sym2 (b, u)];
    if (a == u) m += M[M_SM4 + sym2 (b, t)];
    if (a == t && b == u) m += M[M_SM5];
    if (b == t && a == u) m += M[M_SM5];
    return m;
  }
}

// f(r, q, entry) for the block entries r < D, Q0 <= q < Q0 + NQ, in row
// order, by a template recursion: every moment index is a compile-time
// constant after inlining, so a moment array stays in registers
template <int V, int Q0, int NQ, int E> struct BlockEach {
  template <typename T, typename F>
  SHK_HD static __attribute__ ((always_inline)) void run (const T *M, bool self, F &f)
  {
    BlockEach<V, Q0, NQ, E - 1>::run (M, self, f);
    constexpr int r = (E - 1) / NQ, q = Q0 + (E - 1) % NQ;
    f (r, q, block_entry<V, T> (M, r, q, self));
  }
};
template <int V, int Q0, int NQ> struct BlockEach<V, Q0, NQ, 0> {
  template <typename T, typename F>
  SHK_HD static __attribute__ ((always_inline)) void run (const T *, bool, F &) {}
};

// extracted-form S transform of fts.c:660-670 applied to one 11-vector
template <typename T>
SHK_HD void s_transform (T *z)
{
  T s6 = z[6], s10 = z[10];
  z[6] = T (2) * s6 + s10;
  z[7] = T (2) * z[7];
  z[8] = T (2) * z[8];
  z[9] = T (2) * z[9];
  z[10] = T (2) * s10 + s6;
}

// ---------------------------------------------------------------------------
// Direct contraction of one pair-image block with one transformed 11-vector
// z (acc += block(e, s) . z, non-self convention), written with vector
// identities instead of the 121 entries:
//   U += a1 F + a2 e(e.F) + b1 (T x e) - [g1 e tr + 2 g2 (Z e) + g3 e Q]
//   W += b1 (F x e) + c1 T + c2 e(e.T) + 2 h1 e x (Z e)
//   E_ab += g1 d_ab (e.F) + g2 (e_a F_b + e_b F_a) + g3 e_a e_b (e.F)
//         + h1 (e_a w_b + e_b w_a)                       (w = T x e)
//         + mm1 e_a e_b Q + mm2 (d_ab Q + e_a e_b tr) + mm3 d_ab tr
//         + 2 mm4 (e_a (Ze)_b + e_b (Ze)_a) + 2 mm5 Z_ab
// with Z the symmetric 3x3 whose diagonal is (z6, z10, 0) and off-diagonal
// xy, xz, yz are z7/2, z8/2, z9/2 (the extracted, transformed stresslet),
// Q = e.Z.e and tr = z6 + z10.  Same operator as fts.c:625-1010.
// ---------------------------------------------------------------------------
template <int V, typename T>
SHK_HD void contract_item (T ex, T ey, T ez, const T *s, const T *z, T *acc)
{
  const T eF = ex * z[0] + ey * z[1] + ez * z[2];
  const T a1 = s[1], a2 = s[0] - s[1];
  acc[0] += a1 * z[0] + a2 * ex * eF;
  acc[1] += a1 * z[1] + a2 * ey * eF;
  acc[2] += a1 * z[2] + a2 * ez * eF;
  if (V == 0) return;
  const T b1 = s[2];
  // T x e and F x e
  const T tex = z[4] * ez - z[5] * ey, tey = z[5] * ex - z[3] * ez, tez = z[3] * ey - z[4] * ex;
  const T fex = z[1] * ez - z[2] * ey, fey = z[2] * ex - z[0] * ez, fez = z[0] * ey - z[1] * ex;
  acc[0] += b1 * tex; acc[1] += b1 * tey; acc[2] += b1 * tez;
  const T eT = ex * z[3] + ey * z[4] + ez * z[5];
  const T c1 = s[4], c2 = s[3] - s[4];
  acc[3] += b1 * fex + c1 * z[3] + c2 * ex * eT;
  acc[4] += b1 * fey + c1 * z[4] + c2 * ey * eT;
  acc[5] += b1 * fez + c1 * z[5] + c2 * ez * eT;
  if (V == 1) return;
  const T zxx = z[6], zxy = T (0.5) * z[7], zxz = T (0.5) * z[8], zyz = T (0.5) * z[9];
  const T zyy = z[10];
  const T Zx = zxx * ex + zxy * ey + zxz * ez;
  const T Zy = zxy * ex + zyy * ey + zyz * ez;
  const T Zz = zxz * ex + zyz * ey;
  const T Q = Zx * ex + Zy * ey + Zz * ez;
  const T tr = zxx + zyy;
  const T g1 = T (-1.0 / 3.0) * s[5], g2 = s[6], g3 = s[5] - T (2) * s[6];
  const T gu = g1 * tr + g3 * Q;
  acc[0] -= gu * ex + T (2) * g2 * Zx;
  acc[1] -= gu * ey + T (2) * g2 * Zy;
  acc[2] -= gu * ez + T (2) * g2 * Zz;
  const T h2 = T (2) * s[7];
  acc[3] += h2 * (ey * Zz - ez * Zy);
  acc[4] += h2 * (ez * Zx - ex * Zz);
  acc[5] += h2 * (ex * Zy - ey * Zx);
  const T xm = s[8], ym = s[9], zm = s[10];
  const T mm1 = T (1.5) * xm - T (2) * ym + T (0.5) * zm;
  const T mm2 = T (-0.5) * xm + T (0.5) * zm;
  const T mm3 = T (1.0 / 6.0) * xm - T (0.5) * zm;
  const T mm4x2 = ym - zm;          // 2 mm4
  const T mm5x2 = zm;               // 2 mm5
  const T h1 = s[7];
  // coefficient of e_a e_b and of d_ab
  const T cee = g3 * eF + mm1 * Q + mm2 * tr;
  const T cdd = g1 * eF + mm2 * Q + mm3 * tr;
  // vector multiplying e_b in the symmetric (e_a v_b + e_b v_a) part
  const T vx = g2 * z[0] + h1 * tex + mm4x2 * Zx;
  const T vy = g2 * z[1] + h1 * tey + mm4x2 * Zy;
  const T vz = g2 * z[2] + h1 * tez + mm4x2 * Zz;
  acc[6] += cee * ex * ex + cdd + T (2) * ex * vx + mm5x2 * zxx;   // xx
  acc[7] += cee * ex * ey + ex * vy + ey * vx + mm5x2 * zxy;       // xy
  acc[8] += cee * ex * ez + ex * vz + ez * vx + mm5x2 * zxz;       // xz
  acc[9] += cee * ey * ez + ey * vz + ez * vy + mm5x2 * zyz;       // yz
  acc[10] += cee * ey * ey + cdd + T (2) * ey * vy + mm5x2 * zyy;  // yy
}

// contract_item with the scalars known to be zero left out: bit t of M set
// when s[t] may be nonzero; every product with a scalar outside M is
// skipped (the terms it would add are zero)
template <int V, typename T, unsigned M>
SHK_HD void contract_item_m (T ex, T ey, T ez, const T *s, const T *z, T *acc)
{
  constexpr bool SA = (M & 0x3u) != 0, SB = (M & 0x4u) != 0, SC = (M & 0x18u) != 0;
  constexpr bool SG = (M & 0x60u) != 0, SH = (M & 0x80u) != 0, SM = (M & 0x700u) != 0;
  const T s0 = (M & 0x1u) ? s[0] : T (0), s1 = (M & 0x2u) ? s[1] : T (0);
  const T s3 = (M & 0x8u) ? s[3] : T (0), s4 = (M & 0x10u) ? s[4] : T (0);
  const T s5 = (M & 0x20u) ? s[5] : T (0), s6 = (M & 0x40u) ? s[6] : T (0);
  const T s8 = (M & 0x100u) ? s[8] : T (0), s9 = (M & 0x200u) ? s[9] : T (0);
  const T s10 = (M & 0x400u) ? s[10] : T (0);
  const T eF = ex * z[0] + ey * z[1] + ez * z[2];
  if (SA)
    {
      const T a1 = s1, a2 = s0 - s1;
      acc[0] += a1 * z[0] + a2 * ex * eF;
      acc[1] += a1 * z[1] + a2 * ey * eF;
      acc[2] += a1 * z[2] + a2 * ez * eF;
    }
  if (V == 0) return;
  const T tex = z[4] * ez - z[5] * ey, tey = z[5] * ex - z[3] * ez, tez = z[3] * ey - z[4] * ex;
  if (SB)
    {
      const T b1 = s[2];
      const T fex = z[1] * ez - z[2] * ey, fey = z[2] * ex - z[0] * ez, fez = z[0] * ey - z[1] * ex;
      acc[0] += b1 * tex; acc[1] += b1 * tey; acc[2] += b1 * tez;
      acc[3] += b1 * fex; acc[4] += b1 * fey; acc[5] += b1 * fez;
    }
  if (SC)
    {
      const T eT = ex * z[3] + ey * z[4] + ez * z[5];
      const T c1 = s4, c2 = s3 - s4;
      acc[3] += c1 * z[3] + c2 * ex * eT;
      acc[4] += c1 * z[4] + c2 * ey * eT;
      acc[5] += c1 * z[5] + c2 * ez * eT;
    }
  if (V == 1) return;
  if (!(SG || SH || SM)) return;
  const T zxx = z[6], zxy = T (0.5) * z[7], zxz = T (0.5) * z[8], zyz = T (0.5) * z[9];
  const T zyy = z[10];
  const T Zx = zxx * ex + zxy * ey + zxz * ez;
  const T Zy = zxy * ex + zyy * ey + zyz * ez;
  const T Zz = zxz * ex + zyz * ey;
  const T Q = Zx * ex + Zy * ey + Zz * ez;
  const T tr = zxx + zyy;
  const T g1 = T (-1.0 / 3.0) * s5, g2 = s6, g3 = s5 - T (2) * s6;
  if (SG)
    {
      const T gu = g1 * tr + g3 * Q;
      acc[0] -= gu * ex + T (2) * g2 * Zx;
      acc[1] -= gu * ey + T (2) * g2 * Zy;
      acc[2] -= gu * ez + T (2) * g2 * Zz;
    }
  if (SH)
    {
      const T h2 = T (2) * s[7];
      acc[3] += h2 * (ey * Zz - ez * Zy);
      acc[4] += h2 * (ez * Zx - ex * Zz);
      acc[5] += h2 * (ex * Zy - ey * Zx);
    }
  const T mm1 = T (1.5) * s8 - T (2) * s9 + T (0.5) * s10;
  const T mm2 = T (-0.5) * s8 + T (0.5) * s10;
  const T mm3 = T (1.0 / 6.0) * s8 - T (0.5) * s10;
  const T mm4x2 = s9 - s10;
  const T mm5x2 = s10;
  const T h1 = SH ? s[7] : T (0);
  T cee = T (0), cdd = T (0), vx = T (0), vy = T (0), vz = T (0);
  if (SG)
    {
      cee = g3 * eF; cdd = g1 * eF;
      vx = g2 * z[0]; vy = g2 * z[1]; vz = g2 * z[2];
    }
  if (SH)
    {
      vx += h1 * tex; vy += h1 * tey; vz += h1 * tez;
    }
  if (SM)
    {
      cee += mm1 * Q + mm2 * tr;
      cdd += mm2 * Q + mm3 * tr;
      vx += mm4x2 * Zx; vy += mm4x2 * Zy; vz += mm4x2 * Zz;
    }
  acc[6] += cee * ex * ex + cdd + T (2) * ex * vx;
  acc[7] += cee * ex * ey + ex * vy + ey * vx;
  acc[8] += cee * ex * ez + ex * vz + ez * vx;
  acc[9] += cee * ey * ez + ey * vz + ez * vy;
  acc[10] += cee * ey * ey + cdd + T (2) * ey * vy;
  if (M & 0x400u)
    {
      acc[6] += mm5x2 * zxx; acc[7] += mm5x2 * zxy; acc[8] += mm5x2 * zxz;
      acc[9] += mm5x2 * zyz; acc[10] += mm5x2 * zyy;
    }
}

}  // namespace shk
