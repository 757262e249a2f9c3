/* oracle.c -- FP64 CPU restatement of libstokes' Ewald mobility hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Pinning (how each part is checked, see tests/test_oracle.py):
 *  - orc_set_xi / table: against the compiled reference (oracle/_ref) with
 *    identical inputs; the stopping criterion is re-evaluated in the test.
 *  - scalars, blocks, operators: against oracle/_ref on F/FT/FTS, mono,
 *    poly, shear and open-BC cases; golden fixtures in tests/golden/;
 *    plus the reference's own invariants (atimes == extmat.x, mono == poly
 *    with a = 1, shear shift of +-L, xi independence) and the closed-form
 *    Hasimoto self mobility of a simple cubic array.
 *  - Chebyshev: closed-form function fits and z.z = y.A^-1.y on the
 *    reference's n=10 test matrix (test/check-brownian.c:106).
 *  - MT19937: the published first outputs of init_genrand(5489) and the
 *    compiled reference KIrand.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* set-up                                                               */
/* ------------------------------------------------------------------ */

void
orc_sys_init (orc_sys *s)
{
  memset (s, 0, sizeof (*s));
}

void
orc_sys_free_tables (orc_sys *s)
{
  free (s->rlx); free (s->rly); free (s->rlz);
  free (s->rmx); free (s->rmy); free (s->rmz);
  s->rlx = s->rly = s->rlz = NULL;
  s->rmx = s->rmy = s->rmz = NULL;
  s->nr = 0;
}

/* stokes.c:606-651 (stokes_set_l): pivol and box lengths. */
void
orc_set_l (orc_sys *s, double lx, double ly, double lz)
{
  s->pivol = M_PI / lx / ly / lz;
  s->lx = lx; s->ly = ly; s->lz = lz;
}

/* stokes.c:765-776 (xi_by_tratio). */
double
orc_xi_by_tratio (const orc_sys *s, double tratio)
{
  return pow (tratio, 1.0 / 6.0) * sqrt (M_PI)
    / pow (s->lx * s->ly * s->lz, 1.0 / 3.0);
}

/* stokes.c:654-682 (ewald_real): real-space tail estimate, safety 1.15. */
double
orc_ewald_real_bound (double r, double xi)
{
  double xr = xi * r / 1.15;
  double xr2 = xr * xr;
  double r2 = r * r;
  return erfc (xr) * (0.75 + 1.5 / r2) / r
    + exp (-xr2) * xi / sqrt (M_PI)
      * (4.5 + 3.0 * xr2 + (3.0 + xr2 * (14.0 + xr2 * (4.0 + xr2))) / r2);
}

/* stokes.c:684-700 (ewald_recip): reciprocal tail estimate, safety 1.15. */
double
orc_ewald_recip_bound (double k, double xi)
{
  double kx = k / xi / 1.15;
  double kx2 = kx * kx;
  double k2 = k * k;
  return 6.0 * M_PI * exp (-kx2 / 4.0) * (1.0 + k2 / 3.0) / k2
    * (1.0 + kx2 * (1.0 / 4.0 + kx2 / 8.0));
}

static void
orc_table_push (orc_sys *s, double x, double y, double z, int m1, int m2, int m3)
{
  int i = s->nr++;
  s->rlx = realloc (s->rlx, sizeof (double) * s->nr);
  s->rly = realloc (s->rly, sizeof (double) * s->nr);
  s->rlz = realloc (s->rlz, sizeof (double) * s->nr);
  s->rmx = realloc (s->rmx, sizeof (int) * s->nr);
  s->rmy = realloc (s->rmy, sizeof (int) * s->nr);
  s->rmz = realloc (s->rmz, sizeof (int) * s->nr);
  s->rlx[i] = x; s->rly[i] = y; s->rlz[i] = z;
  s->rmx[i] = m1; s->rmy[i] = m2; s->rmz[i] = m3;
}

/* stokes.c:702-762 (stokes_set_xi) and the real-space part of
 * stokes.c:173-233 (stokes_ewald_table_make). */
void
orc_set_xi (orc_sys *s, double xi, double ewald_eps)
{
  s->ewald_eps = ewald_eps;
  s->xi = xi;
  s->xi2 = xi * xi;
  s->xia2 = s->xi2;
  s->xiaspi = xi / sqrt (M_PI);

  s->self_a = 1.0  - s->xiaspi * (6.0 - 40.0 / 3.0 * s->xia2);
  s->self_c = 0.75 - s->xiaspi * s->xia2 * 10.0;
  s->self_m = 0.9  - s->xiaspi * s->xia2 * (12.0 - 30.24 * s->xia2);

  const double shrink = 0.99;
  double lmin = s->lx, lmax = s->lx;
  if (lmin > s->ly) lmin = s->ly;
  if (lmin > s->lz) lmin = s->lz;
  if (lmax < s->ly) lmax = s->ly;
  if (lmax < s->lz) lmax = s->lz;
  double dr = lmin;
  double dk = 2.0 * M_PI / lmax;

  double rmax = 10.0 * sqrt (-log (ewald_eps)) / xi;
  while (orc_ewald_real_bound (rmax * shrink, xi)
         + 6.0 * orc_ewald_real_bound (rmax * shrink + dr, xi) < ewald_eps)
    rmax *= shrink;
  s->rmax2 = rmax * rmax;
  s->rmaxx = (int) (rmax / s->lx) + 1;
  s->rmaxy = (int) (rmax / s->ly) + 1;
  s->rmaxz = (int) (rmax / s->lz) + 1;

  double kmax = 10.0 * sqrt (-log (ewald_eps)) * 2.0 * xi;
  while (orc_ewald_recip_bound (kmax * shrink, xi)
         + 6.0 * orc_ewald_recip_bound (kmax * shrink + dk, xi) < ewald_eps)
    kmax *= shrink;
  s->kmax = kmax;
  s->kmaxx = (int) (kmax * s->lx / 2.0 / M_PI);
  s->kmaxy = (int) (kmax * s->ly / 2.0 / M_PI);
  s->kmaxz = (int) (kmax * s->lz / 2.0 / M_PI);

  /* image table: cells whose half-cell-shifted centre lies within
   * rmax + |box diagonal| of the origin */
  orc_sys_free_tables (s);
  double rl = sqrt (s->lx * s->lx + s->ly * s->ly + s->lz * s->lz);
  for (int m1 = -s->rmaxx; m1 <= s->rmaxx; m1++)
    {
      double x = s->lx * (double) m1, xx = x;
      if (m1 > 0) xx -= 0.5 * s->lx;
      if (m1 < 0) xx += 0.5 * s->lx;
      for (int m2 = -s->rmaxy; m2 <= s->rmaxy; m2++)
        {
          double y = s->ly * (double) m2, yy = y;
          if (m2 > 0) yy -= 0.5 * s->ly;
          if (m2 < 0) yy += 0.5 * s->ly;
          for (int m3 = -s->rmaxz; m3 <= s->rmaxz; m3++)
            {
              double z = s->lz * (double) m3, zz = z;
              if (m3 > 0) zz -= 0.5 * s->lz;
              if (m3 < 0) zz += 0.5 * s->lz;
              if (sqrt (xx * xx + yy * yy + zz * zz) <= rmax + rl)
                orc_table_push (s, x, y, z, m1, m2, m3);
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* pair scalars                                                         */
/* ------------------------------------------------------------------ */

/* ewald.c:48-186 (scalars_ewald_real). */
void
orc_scalars_ewald_real (int version, double xi, double r, double *o)
{
  for (int k = 2; k < 11; k++) o[k] = 0.0;
  double xr = xi * r, x2 = xr * xr;
  double s = r, s2 = r * r;
  double ec = erfc (xr);
  double ex = xi / sqrt (M_PI) * exp (-x2);

  o[1] = (0.75 + 0.5 / s2) / s * ec
    + ((1.0 + x2 * (14.0 + 4.0 * x2 * (-5.0 + x2))) / s2 - 4.5 + 3.0 * x2) * ex;
  double a2 = (0.75 - 1.5 / s2) / s * ec
    + ((-3.0 + x2 * (-2.0 + 4.0 * x2 * (4.0 - x2))) / s2 + 1.5 - 3.0 * x2) * ex;
  o[0] = a2 + o[1];
  if (version == 0) return;

  o[2] = -0.75 / s2 * ec
    - 1.5 * (1.0 + x2 * (-6.0 + x2 * 2.0)) / s * ex;
  o[4] = -3.0 / 8.0 / s2 / s * ec
    - 0.75 * (1.0 + x2 * (14.0 + x2 * (-20.0 + x2 * 4.0))) / s2 * ex;
  double c2 = 9.0 / 8.0 / s2 / s * ec
    - 0.75 * (-3.0 + x2 * (-2.0 + x2 * (16.0 + x2 * (-4.0)))) / s2 * ex;
  o[3] = c2 + o[4];
  if (version == 1) return;

  o[5] = (2.25 - 3.6 / s2) / s2 * ec
    + (-1.5 * (-3.0 + x2 * 6.0)
       - 0.8 * (9.0 + x2 * (6.0 + x2 * (-48.0 + x2 * 12.0))) / s2) / s * ex;
  o[6] = 1.2 / s2 / s2 * ec
    + (-3.0 * (x2 * (2.0 + x2 * (-1.0)))
       - 0.8 * (-3.0 + x2 * (-2.0 + x2 * (-26.0 + x2 * (26.0 + x2 * (-4.0)))))
         / s2) / s * ex;
  o[7] = -9.0 / 8.0 / s2 / s * ec
    + 1.5 * (-1.5 + x2 * (-1.0 + x2 * (8.0 + x2 * (-2.0)))) / s2 * ex;
  o[8] = (-4.5 + 10.8 / s2) / s / s2 * ec
    + (1.5 * (-6.0 + x2 * (-12.0 + x2 * 12.0))
       + 1.2 * (18.0 + x2 * (12.0 + x2 * (30.0 + x2 * (-66.0 + x2 * 12.0))))
         / s2) / s2 * ex;
  o[9] = (2.25 - 7.2 / s2) / s / s2 * ec
    + (-1.5 * (-3.0 + x2 * (6.0 + x2 * (-12.0 + x2 * 4.0)))
       - 1.2 * (12.0 + x2 * (8.0 + x2 * (-22.0 + x2 * (58.0 + x2 * (-34.0
                                                                  + x2 * 4.0)))))
         / s2) / s2 * ex;
  o[10] = 1.8 / s2 / s / s2 * ec
    + (-1.5 * (0.0 + x2 * (8.0 + x2 * (-4.0)))
       - 1.2 * (-3.0 + x2 * (-2.0 + x2 * (-26.0 + x2 * (26.0 + x2 * (-4.0)))))
         / s2) / s2 * ex;
}

/* ewald.c:203-359 (scalars_ewald_real_poly), dimensional form. */
void
orc_scalars_ewald_real_poly (int version, double xi, double r,
                             double aa, double ab, double *o)
{
  for (int k = 2; k < 11; k++) o[k] = 0.0;
  double x = xi * r, x2 = x * x;
  double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r3 * r2;
  double aa2 = aa * aa, ab2 = ab * ab;
  double ec = erfc (x);
  double ex = xi / sqrt (M_PI) * exp (-x2);

  double a0 = 2.0 * (-3.0 + x2 * 2.0) * ex + ec / r;
  double ad0 = -2.0 / r * (1.0 + x2 * (-10.0 + x2 * 4.0)) * ex - ec / r2;
  double add0 = 4.0 / r2 * (1.0 + x2 * (6.0 + x2 * (-16.0 + x2 * 4.0))) * ex
    + 2.0 / r3 * ec;
  double b0 = 2.0 * (1.0 + x2 * (-2.0)) * ex + ec / r;
  double bd0 = -2.0 / r * (1.0 + x2 * (6.0 + x2 * (-4.0))) * ex - ec / r2;

  double a2 = 4.0 / r2 * (1.0 + x2 * (14.0 + x2 * (-20.0 + x2 * 4.0))) * ex
    + 2.0 / r3 * ec;
  double ad2 = -4.0 / r3
    * (3.0 + x2 * (2.0 + x2 * (68.0 + x2 * (-56.0 + x2 * 8.0)))) * ex
    - 6.0 / r4 * ec;
  double add2 = 16.0 / r4
    * (3.0 + x2 * (2.0 + x2 * (-16.0 + x2 * (76.0 + x2 * (-38.0 + x2 * 4.0)))))
    * ex + 24.0 / r5 * ec;
  double b2 = -4.0 / r2 * (3.0 + x2 * (2.0 + x2 * (-16.0 + x2 * 4.0))) * ex
    - 6.0 / r3 * ec;
  double bd2 = 4.0 / r3
    * (9.0 + x2 * (6.0 + x2 * (36.0 + x2 * (-48.0 + x2 * 8.0)))) * ex
    + 18.0 / r4 * ec;

  double a10 = a0;
  double a12 = a2 * (aa2 + ab2) / 6.0;
  double a20 = b0;
  double a22 = b2 * (aa2 + ab2) / 6.0;
  o[1] = 0.75 * (a10 + a12);
  o[0] = 0.75 * (a20 + a22) + o[1];
  if (version == 0) return;

  o[2] = -0.375 * (-ad0 + b0 / r);
  double c1 = 0.1875 * (-add0 + (-ad0 + bd0) / r);
  double c2 = 0.1875 * (add0 + (-ad0 - bd0 + 2.0 * b0 / r) / r);
  o[4] = c1;
  o[3] = c2 + o[4];
  if (version == 1) return;

  double wg = aa2 / 10.0 + ab2 / 6.0;
  double g10 = -0.75 / r * b0;
  double g12 = -0.75 / r * b2 * wg;
  double g20 = -0.375 * (ad0 + b0 / r);
  double g22 = -0.375 * (ad2 + b2 / r) * wg;
  o[5] = -3.0 * (g10 + g12);
  o[6] = g20 + g22;

  o[7] = -0.1875 * (add0 + (-ad0 - bd0 + 2.0 * b0 / r) / r);

  double wm = (aa2 + ab2) / 10.0;
  double m50 = -0.375 * (ad0 + b0 / r) / r;
  double m52 = -0.375 * (ad2 + b2 / r) / r * wm;
  o[10] = 2.0 * (m50 + m52);
  double m30 = -0.75 * b0 / r2;
  double m32 = -0.75 * b2 / r2 * wm;
  o[8] = 6.0 * (m30 + m32) + 3.0 * o[10];
  double m40 = -0.1875 * (add0 + (-ad0 + 3.0 * bd0 - 6.0 * b0 / r) / r);
  double m42 = -0.1875 * (add2 + (-ad2 + 3.0 * bd2 - 6.0 * b2 / r) / r) * wm;
  o[9] = 2.0 * (m40 + m42) + o[10];
}

/* ewald.c:382-420 (scalars_mob_poly_scale_SD_) and
 * non-ewald.c (scalars_mob_poly_scale_SD): dimensional -> SD form. */
void
orc_scalars_scale_SD (int version, double a1, double *s)
{
  s[0] *= a1; s[1] *= a1;
  if (version == 0) return;
  double a12 = a1 * a1, a13 = a12 * a1;
  s[2] *= a12;
  s[3] *= a13; s[4] *= a13;
  if (version == 1) return;
  s[5] *= a12; s[6] *= a12;
  s[7] *= a13;
  s[8] *= a13; s[9] *= a13; s[10] *= a13;
}

/* non-ewald.c:49-90 (scalars_nonewald): open-BC RPY. */
void
orc_scalars_nonewald (int version, double r, double *o)
{
  for (int k = 0; k < 11; k++) o[k] = 0.0;
  double r2 = r * r;
  o[0] = (1.5 - 1.0 / r2) / r;
  o[1] = (0.75 + 0.5 / r2) / r;
  if (version == 0) return;
  o[2] = -0.75 / r2;
  o[3] = 0.75 / r2 / r;
  o[4] = -0.375 / r2 / r;
  if (version == 1) return;
  o[5] = (2.25 - 3.6 / r2) / r2;
  o[6] = 1.2 / r2 / r2;
  o[7] = -1.125 / r2 / r;
  o[8] = (-4.5 + 10.8 / r2) / r2 / r;
  o[9] = (2.25 - 7.2 / r2) / r2 / r;
  o[10] = 1.8 / r2 / r2 / r;
}

/* non-ewald.c:110-145 (scalars_nonewald_poly). */
void
orc_scalars_nonewald_poly (int version, double r, double aa, double ab,
                           double *o)
{
  for (int k = 0; k < 11; k++) o[k] = 0.0;
  double r2 = r * r, aa2 = aa * aa, ab2 = ab * ab;
  o[0] = (1.5 - (aa2 + ab2) * 0.5 / r2) / r;
  o[1] = (0.75 + (aa2 + ab2) * 0.25 / r2) / r;
  if (version == 0) return;
  o[2] = -0.75 / r2;
  o[3] = 0.75 / r2 / r;
  o[4] = -3.0 / 8.0 / r2 / r;
  if (version == 1) return;
  o[5] = (2.25 - 13.5 * (aa2 / 10.0 + ab2 / 6.0) / r2) / r2;
  o[6] = 4.5 * (aa2 / 10.0 + ab2 / 6.0) / r2 / r2;
  o[7] = -9.0 / 8.0 / r2 / r;
  o[8] = (-4.5 + 5.4 * (aa2 + ab2) / r2) / r2 / r;
  o[9] = (2.25 - 3.6 * (aa2 + ab2) / r2) / r2 / r;
  o[10] = 0.9 * (aa2 + ab2) / r2 / r2 / r;
}

/* ------------------------------------------------------------------ */
/* 11x11 pair block                                                     */
/* ------------------------------------------------------------------ */

/* Dense pair block Bk[11][11] (row = output UOE component, column =
 * input FTS component, extracted-form S columns) for the non-self sign
 * convention, built from f.c:72-102 (A), ft.c:100-171 (B, C) and
 * fts.c:136-610 (G, GT, H, HT, M).  self != 0 applies the i == j tilde
 * rules of fts.c:73-111 / ft.c:66-86 (B-tilde = -B, G-tilde = +GT). */
static void
orc_block_dense (int version, double ex, double ey, double ez,
                 const double *s, int self, double B[11][11])
{
  memset (B, 0, sizeof (double) * 121);
  double xa = s[0], ya = s[1];
  double exx = ex * ex, eyy = ey * ey, ezz = ez * ez;
  double exy = ex * ey, eyz = ey * ez, exz = ez * ex;

  /* A (f.c:72-102) */
  double a1 = ya, a2 = xa - ya;
  B[0][0] = a1 + a2 * exx; B[1][1] = a1 + a2 * eyy; B[2][2] = a1 + a2 * ezz;
  B[0][1] = B[1][0] = a2 * exy;
  B[1][2] = B[2][1] = a2 * eyz;
  B[2][0] = B[0][2] = a2 * exz;
  if (version == 0) return;

  /* B (ft.c:100-123) at (omega, F); B-tilde at (U, T) */
  double b1 = s[2];
  double bx = b1 * ex, by = b1 * ey, bz = b1 * ez;
  double Bm[3][3] = {{0, bz, -by}, {-bz, 0, bx}, {by, -bx, 0}};
  double tsign = self ? -1.0 : 1.0;
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++)
      {
        B[3 + p][q] += Bm[p][q];
        B[p][3 + q] += tsign * Bm[p][q];
      }
  /* C (ft.c:135-171) */
  double c1 = s[4], c2 = s[3] - s[4];
  B[3][3] += c1 + c2 * exx; B[4][4] += c1 + c2 * eyy; B[5][5] += c1 + c2 * ezz;
  B[3][4] += c2 * exy; B[4][3] += c2 * exy;
  B[4][5] += c2 * eyz; B[5][4] += c2 * eyz;
  B[5][3] += c2 * exz; B[3][5] += c2 * exz;
  if (version == 1) return;

  /* G (fts.c:136-232): rows E(xx,xy,xz,yz,yy), columns F(x,y,z) */
  double xg = s[5], yg = s[6];
  double e[3] = {ex, ey, ez};
  static const int P[5] = {0, 0, 0, 1, 1}, Q[5] = {0, 1, 2, 2, 1};
  double G[5][3];
  for (int r = 0; r < 5; r++)
    for (int k = 0; k < 3; k++)
      {
        int p = P[r], q = Q[r];
        double d_pq = (p == q), d_qk = (q == k), d_pk = (p == k);
        G[r][k] = -1.0 / 3.0 * xg * d_pq * e[k]
          + yg * (e[p] * d_qk + e[q] * d_pk)
          + (xg - 2.0 * yg) * e[p] * e[q] * e[k];
      }
  /* G-tilde at (U, S): -GT for i != j, +GT for i == j (fts.c:90-105) */
  double gsign = self ? 1.0 : -1.0;
  for (int r = 0; r < 5; r++)
    for (int k = 0; k < 3; k++)
      {
        B[6 + r][k] += G[r][k];
        B[k][6 + r] += gsign * G[r][k];
      }
  /* H (fts.c:335-360), HT at (omega, S) (fts.c:374-410) */
  double h = s[7];
  double hxx = h * exx, hxy = h * exy, hxz = h * exz;
  double hyy = h * eyy, hyz = h * eyz, hzz = h * ezz;
  double H[5][3] = {
    {0.0, 2.0 * hxz, -2.0 * hxy},
    {-hxz, hyz, hxx - hyy},
    {hxy, -hxx + hzz, -hyz},
    {hyy - hzz, -hxy, hxz},
    {-2.0 * hyz, 0.0, 2.0 * hxy}};
  for (int r = 0; r < 5; r++)
    for (int k = 0; k < 3; k++)
      {
        B[6 + r][3 + k] += H[r][k];
        B[3 + k][6 + r] += H[r][k];
      }
  /* M (fts.c:420-610): M_{pq,tu} = mm1 e_p e_q e_t e_u
   *   + mm2 (d_pq e_t e_u + e_p e_q d_tu) + mm3 d_pq d_tu
   *   + mm4 (e_p d_qt e_u + e_p d_qu e_t + e_q d_pt e_u + e_q d_pu e_t)
   *   + mm5 (d_pt d_qu + d_qt d_pu)   on extracted rows/columns. */
  double xm = s[8], ym = s[9], zm = s[10];
  double mm1 = 1.5 * xm - 2.0 * ym + 0.5 * zm;
  double mm2 = -0.5 * xm + 0.5 * zm;
  double mm3 = 1.0 / 6.0 * xm - 0.5 * zm;
  double mm4 = 0.5 * ym - 0.5 * zm;
  double mm5 = 0.5 * zm;
  for (int r = 0; r < 5; r++)
    for (int c = 0; c < 5; c++)
      {
        int p = P[r], q = Q[r], t = P[c], u = Q[c];
        double d_pq = (p == q), d_tu = (t == u);
        double d_qt = (q == t), d_qu = (q == u), d_pt = (p == t), d_pu = (p == u);
        B[6 + r][6 + c] += mm1 * e[p] * e[q] * e[t] * e[u]
          + mm2 * (d_pq * e[t] * e[u] + e[p] * e[q] * d_tu)
          + mm3 * d_pq * d_tu
          + mm4 * (e[p] * d_qt * e[u] + e[p] * d_qu * e[t]
                   + e[q] * d_pt * e[u] + e[q] * d_pu * e[t])
          + mm5 * (d_pt * d_qu + d_qt * d_pu);
      }
}

/* fts.c:625-1010 (matrix_fts_atimes), ft.c:179-258, f.c:116-149:
 * y += block(e, s) . T(x), T the extracted-form S transform
 * (fts.c:660-670).  The ATIMES routines use the i != j convention for
 * every pair, which is what orc_block_dense(self=0) stores. */
void
orc_block_atimes (int version, const double *x, double *y,
                  double ex, double ey, double ez, const double *s)
{
  orc_block_atimes_sign (version, x, y, ex, ey, ez, s, 0);
}

/* the same with the i == j signs when self != 0 (matrix_fts_self_atimes,
 * fts.c:1024-1400, and its F / FT forms) */
void
orc_block_atimes_sign (int version, const double *x, double *y,
                       double ex, double ey, double ez, const double *s,
                       int self)
{
  int d = version == 0 ? 3 : (version == 1 ? 6 : 11);
  double B[11][11];
  orc_block_dense (version, ex, ey, ez, s, self, B);
  double z[11];
  for (int k = 0; k < d; k++) z[k] = x[k];
  if (version == 2)
    {
      z[6] = 2.0 * x[6] + x[10];
      z[7] = 2.0 * x[7];
      z[8] = 2.0 * x[8];
      z[9] = 2.0 * x[9];
      z[10] = 2.0 * x[10] + x[6];
    }
  for (int p = 0; p < d; p++)
    {
      double acc = 0.0;
      for (int q = 0; q < d; q++) acc += B[p][q] * z[q];
      y[p] += acc;
    }
}

/* fts.c:45-123 (matrix_fts_ij), ft.c:44-93, f.c:43-57: add the pair block
 * at (i, j) of the dense (n x n) row-major matrix. */
void
orc_block_ij (int version, int i, int j, double ex, double ey, double ez,
              const double *s, int n, double *mat)
{
  int d = version == 0 ? 3 : (version == 1 ? 6 : 11);
  double B[11][11];
  orc_block_dense (version, ex, ey, ez, s, i == j, B);
  for (int p = 0; p < d; p++)
    for (int q = 0; q < d; q++)
      mat[(size_t) (i * d + p) * n + (size_t) (j * d + q)] += B[p][q];
}

/* ------------------------------------------------------------------ */
/* operators                                                            */
/* ------------------------------------------------------------------ */

int
orc_dim_of (int version)
{
  return version == 0 ? 3 : (version == 1 ? 6 : 11);
}

static int
orc_dim (int version)
{
  return orc_dim_of (version);
}

/* self scalars; ewald.c:540-643 (mono: stokes.c:714-716; poly 591-598) */
static void
orc_self_scalars (const orc_sys *s, int i, double *sc)
{
  for (int k = 0; k < 11; k++) sc[k] = 0.0;
  double sa = s->self_a, scc = s->self_c, sm = s->self_m;
  if (s->a != NULL)
    {
      double a = s->a[i], a2 = a * a, a3 = a2 * a;
      sa = 1.0 - a * s->xiaspi * (6.0 - 40.0 / 3.0 * s->xia2 * a2);
      scc = 0.75 - a3 * s->xiaspi * s->xia2 * 10.0;
      sm = 0.9 - a3 * s->xiaspi * s->xia2 * (12.0 - 30.24 * s->xia2 * a2);
    }
  sc[0] = sc[1] = sa;
  sc[3] = sc[4] = scc;
  sc[8] = sc[9] = sc[10] = sm;
}

/* ewald.c:700-737: scalars of one pair-image, SD form */
static void
orc_real_scalars (const orc_sys *s, int i, int j, double r, double *sc)
{
  if (s->a == NULL)
    orc_scalars_ewald_real (s->version, s->xi, r, sc);
  else
    {
      orc_scalars_ewald_real_poly (s->version, s->xi, r, s->a[i], s->a[j], sc);
      orc_scalars_scale_SD (s->version, s->a[i], sc);
    }
}

/* ewald.c:978-1025: reciprocal scalars (ya yb yc yg yh ym) of k for (i,j) */
static void
orc_recip_scalars (const orc_sys *s, int i, int j, double k, double kk,
                   double kexp, double *y6)
{
  if (s->a == NULL)
    {
      y6[0] = 6.0 * (1.0 - kk / 3.0) * kexp;
      y6[1] = 3.0 * k * kexp;
      y6[2] = 3.0 / 2.0 * kk * kexp;
      y6[3] = 3.0 * (1.0 - 4.0 / 15.0 * kk) * k * kexp;
      y6[4] = 3.0 / 2.0 * kk * kexp;
      y6[5] = 3.0 * (1.0 - kk / 5.0) * kk * kexp;
    }
  else
    {
      double aa = s->a[i], aa2 = aa * aa, aa3 = aa2 * aa;
      double a2a = aa2, a2b = s->a[j] * s->a[j];
      y6[0] = aa * 6.0 * (1.0 - kk * (a2a + a2b) / 6.0) * kexp;
      y6[1] = aa2 * 3.0 * k * kexp;
      y6[2] = aa3 * 3.0 / 2.0 * kk * kexp;
      y6[3] = aa2 * 3.0 * (1.0 - kk * (a2a / 10.0 + a2b / 6.0)) * k * kexp;
      y6[4] = aa3 * 3.0 / 2.0 * kk * kexp;
      y6[5] = aa3 * 3.0 * (1.0 - kk * (a2a + a2b) / 10.0) * kk * kexp;
    }
}

/* Visitor shared by the matvec and the dense build: every contribution
 * (i, j, e, scalars) of ewald.c:519-1081 / 1092-1677 in reference order. */
typedef void (*orc_visit_fn) (void *ctx, int i, int j, double ex, double ey,
                              double ez, const double *sc);

static void
orc_ewald_visit (const orc_sys *s, int i0, int i1, orc_visit_fn visit,
                 void *ctx)
{
  int np = s->np;
  double sc[11];
  /* self, ewald.c:540-643 */
  for (int i = i0; i < i1; i++)
    {
      orc_self_scalars (s, i, sc);
      visit (ctx, i, i, 0.0, 0.0, 0.0, sc);
    }
  /* real space, ewald.c:649-789 */
  for (int i = i0; i < i1; i++)
    {
      double ai = s->a ? s->a[i] : 1.0;
      for (int j = 0; j < np; j++)
        {
          double aj = s->a ? s->a[j] : 1.0;
          for (int m = 0; m < s->nr; m++)
            {
              double xx = s->pos[3 * j] - s->pos[3 * i] + s->rlx[m];
              double yy = s->pos[3 * j + 1] - s->pos[3 * i + 1] + s->rly[m];
              double zz = s->pos[3 * j + 2] - s->pos[3 * i + 2] + s->rlz[m];
              if (s->shear_mode == 1)
                xx += (double) s->rmy[m] * s->shear_shift;
              else if (s->shear_mode == 2)
                xx += (double) s->rmz[m] * s->shear_shift;
              double rr = xx * xx + yy * yy + zz * zz;
              if (rr == 0.0) continue;
              if (rr > s->rmax2) continue;
              double r = sqrt (rr);
              double rmin = (ai + aj) * s->rmin;
              if (r < rmin) r = rmin;
              orc_real_scalars (s, i, j, r, sc);
              visit (ctx, i, j, xx / r, yy / r, zz / r, sc);
            }
        }
    }
  /* reciprocal space, ewald.c:923-1075 */
  double k0x = 2.0 * M_PI / s->lx;
  double k0y = 2.0 * M_PI / s->ly;
  double k0z = 2.0 * M_PI / s->lz;
  double sl = s->shear_shift / s->lx;
  for (int m1 = -s->kmaxx; m1 <= s->kmaxx; m1++)
    {
      double k1 = k0x * (double) m1;
      for (int m2 = -s->kmaxy; m2 <= s->kmaxy; m2++)
        {
          double k2 = k0y * (double) m2;
          for (int m3 = -s->kmaxz; m3 <= s->kmaxz; m3++)
            {
              double k3 = k0z * (double) m3;
              if (m1 == 0 && m2 == 0 && m3 == 0) continue;
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
              double ex = k1 / k, ey = k2 / k, ez = k3 / k;
              for (int i = i0; i < i1; i++)
                for (int j = 0; j < np; j++)
                  {
                    double y6[6];
                    orc_recip_scalars (s, i, j, k, kk, kexp, y6);
                    double xx = s->pos[3 * j] - s->pos[3 * i];
                    double yy = s->pos[3 * j + 1] - s->pos[3 * i + 1];
                    double zz = s->pos[3 * j + 2] - s->pos[3 * i + 2];
                    double ph = k1 * xx + k2 * yy + k3 * zz;
                    double cf = cos (ph);
                    double sf = (s->version == 0) ? 0.0 : -sin (ph);
                    /* ewald.c:1037-1068: (xa ya yb xc yc xg yg yh xm ym zm) */
                    sc[0] = 0.0; sc[1] = cf * y6[0];
                    sc[2] = sf * y6[1];
                    sc[3] = 0.0; sc[4] = cf * y6[2];
                    sc[5] = 0.0;
                    sc[6] = (s->fix_recip_g ? -sf : sf) * y6[3];
                    sc[7] = cf * y6[4];
                    sc[8] = 0.0; sc[9] = cf * y6[5]; sc[10] = 0.0;
                    visit (ctx, i, j, ex, ey, ez, sc);
                  }
            }
        }
    }
}

typedef struct { int version; const double *x; double *y; } orc_mv_ctx;

static void
orc_mv_visit (void *c, int i, int j, double ex, double ey, double ez,
              const double *sc)
{
  orc_mv_ctx *m = (orc_mv_ctx *) c;
  int d = orc_dim (m->version);
  orc_block_atimes (m->version, m->x + j * d, m->y + i * d, ex, ey, ez, sc);
}

/* ewald.c:519-1081 (atimes_ewald_3all), rows [i0, i1) of y only. */
void
orc_atimes_ewald_rows (const orc_sys *s, const double *x, double *y,
                       int i0, int i1)
{
  int d = orc_dim (s->version);
  if (i1 < 0 || i1 > s->np) i1 = s->np;
  for (int k = i0 * d; k < i1 * d; k++) y[k] = 0.0;
  orc_mv_ctx c = {s->version, x, y};
  orc_ewald_visit (s, i0, i1, orc_mv_visit, &c);
}

void
orc_atimes_ewald (const orc_sys *s, const double *x, double *y)
{
  orc_atimes_ewald_rows (s, x, y, 0, s->np);
}

typedef struct { int version, n; double *mat; } orc_mat_ctx;

static void
orc_mat_visit (void *c, int i, int j, double ex, double ey, double ez,
               const double *sc)
{
  orc_mat_ctx *m = (orc_mat_ctx *) c;
  orc_block_ij (m->version, i, j, ex, ey, ez, sc, m->n, m->mat);
}

/* ewald.c:1092-1677 (make_matrix_mob_ewald_3all). */
void
orc_matrix_ewald (const orc_sys *s, double *mat)
{
  int n = orc_dim (s->version) * s->np;
  memset (mat, 0, sizeof (double) * (size_t) n * (size_t) n);
  orc_mat_ctx c = {s->version, n, mat};
  orc_ewald_visit (s, 0, s->np, orc_mat_visit, &c);
}

/* non-ewald.c:330-493 (atimes_nonewald_3all) / 494- (make_matrix_...). */
static void
orc_nonewald_visit (const orc_sys *s, orc_visit_fn visit, void *ctx)
{
  double sc[11];
  for (int i = 0; i < s->np; i++)
    {
      for (int k = 0; k < 11; k++) sc[k] = 0.0;
      sc[0] = sc[1] = 1.0;
      sc[3] = sc[4] = 0.75;
      sc[8] = sc[9] = sc[10] = 0.9;
      visit (ctx, i, i, 0.0, 0.0, 0.0, sc);
    }
  for (int i = 0; i < s->np; i++)
    {
      double ai = s->a ? s->a[i] : 1.0;
      for (int j = 0; j < s->np; j++)
        {
          if (i == j) continue;
          double aj = s->a ? s->a[j] : 1.0;
          double xx = s->pos[3 * j] - s->pos[3 * i];
          double yy = s->pos[3 * j + 1] - s->pos[3 * i + 1];
          double zz = s->pos[3 * j + 2] - s->pos[3 * i + 2];
          double r = sqrt (xx * xx + yy * yy + zz * zz);
          double rmin = (ai + aj) * s->rmin;
          if (r < rmin) r = rmin;
          if (s->a == NULL)
            orc_scalars_nonewald (s->version, r, sc);
          else
            {
              orc_scalars_nonewald_poly (s->version, r, ai, aj, sc);
              orc_scalars_scale_SD (s->version, ai, sc);
            }
          visit (ctx, i, j, xx / r, yy / r, zz / r, sc);
        }
    }
}

void
orc_atimes_nonewald (const orc_sys *s, const double *x, double *y)
{
  int n = orc_dim (s->version) * s->np;
  for (int k = 0; k < n; k++) y[k] = 0.0;
  orc_mv_ctx c = {s->version, x, y};
  orc_nonewald_visit (s, orc_mv_visit, &c);
}

void
orc_matrix_nonewald (const orc_sys *s, double *mat)
{
  int n = orc_dim (s->version) * s->np;
  memset (mat, 0, sizeof (double) * (size_t) n * (size_t) n);
  orc_mat_ctx c = {s->version, n, mat};
  orc_nonewald_visit (s, orc_mat_visit, &c);
}

/* ewald.c:432-447 (atimes_3all): periodic dispatch. */
void
orc_atimes_3all (const orc_sys *s, const double *x, double *y)
{
  if (s->periodic) orc_atimes_ewald (s, x, y);
  else orc_atimes_nonewald (s, x, y);
}

/* matrix.c:447-478 (multiply_extmat_with_extvec_3fts). */
void
orc_extmat_fts (int np, const double *m, const double *x, double *y)
{
  int n = np * 11;
  for (int i = 0; i < n; i++)
    {
      double acc = 0.0;
      for (int j = 0; j < np; j++)
        {
          const double *row = m + (size_t) i * n + j * 11;
          const double *xj = x + j * 11;
          for (int q = 0; q < 6; q++) acc += row[q] * xj[q];
          acc += row[6] * (2.0 * xj[6] + xj[10]);
          acc += row[7] * 2.0 * xj[7];
          acc += row[8] * 2.0 * xj[8];
          acc += row[9] * 2.0 * xj[9];
          acc += row[10] * (2.0 * xj[10] + xj[6]);
        }
      y[i] = acc;
    }
}

/* matrix.c:365-397 (dot_prod_matrix), built-in GEMV order. */
void
orc_dot_prod_matrix (const double *m, int n1, int n2, const double *x,
                     double *y)
{
  for (int i = 0; i < n1; i++)
    {
      double acc = 0.0;
      for (int j = 0; j < n2; j++) acc += m[(size_t) i * n2 + j] * x[j];
      y[i] = acc;
    }
}

/* ------------------------------------------------------------------ */
/* Chebyshev                                                            */
/* ------------------------------------------------------------------ */

double
orc_func (int func_id, double x)
{
  if (func_id == 0) return 1.0 / sqrt (x);   /* brownian.c:282-286 */
  if (func_id == 1) return sqrt (x);
  if (func_id == 2) return cos (x);
  if (func_id == 3) return sin (x);
  return 0.0;
}

/* chebyshev.c:33-61 (chebyshev_coef). */
void
orc_cheb_coef (int ncheb, int func_id, double x0, double x1, double *a)
{
  double h = 0.5 * (x1 - x0), c = 0.5 * (x1 + x0);
  for (int j = 0; j < ncheb; j++)
    {
      double acc = 0.0;
      for (int k = 1; k <= ncheb; k++)
        {
          double t = M_PI * ((double) k - 0.5) / (double) ncheb;
          acc += orc_func (func_id, cos (t) * h + c) * cos ((double) j * t);
        }
      a[j] = 2.0 / (double) ncheb * acc;
    }
  a[0] *= 0.5;
}

/* chebyshev.c:76-107 (chebyshev_eval, Clenshaw). */
double
orc_cheb_eval (int ncheb, const double *a, double x0, double x1, double x)
{
  double h = 0.5 * (x1 - x0), c = 0.5 * (x1 + x0);
  double t = (x - c) / h;
  double b1 = 0.0, b0 = 0.0;   /* b_{j+1}, b_{j+2} */
  for (int j = ncheb - 1; j >= 1; j--)
    {
      double b = 2.0 * t * b1 - b0 + a[j];
      b0 = b1; b1 = b;
    }
  return t * b1 - b0 + a[0];
}

/* chebyshev.c:163-228 (chebyshev_eval_atimes): Clenshaw on an operator. */
void
orc_cheb_eval_atimes (int ncheb, const double *a, int n, const double *y,
                      double *z, double e0, double e1, orc_atimes_fn atimes,
                      void *ud)
{
  double da = 2.0 / (e1 - e0);
  double db = -(e1 + e0) / (e1 - e0);
  double *x = malloc (sizeof (double) * n);
  double *b1 = calloc (n, sizeof (double));   /* b_{j+1} */
  double *b0 = calloc (n, sizeof (double));   /* b_{j+2} */
  double *bn = malloc (sizeof (double) * n);
  for (int j = ncheb - 1; j >= 1; j--)
    {
      atimes (n, b1, x, ud);
      for (int k = 0; k < n; k++)
        bn[k] = 2.0 * da * x[k] + 2.0 * db * b1[k] - b0[k] + a[j] * y[k];
      double *t = b0; b0 = b1; b1 = bn; bn = t;
    }
  atimes (n, b1, x, ud);
  for (int k = 0; k < n; k++)
    z[k] = da * x[k] + db * b1[k] - b0[k] + a[0] * y[k];
  free (x); free (b1); free (b0); free (bn);
}

/* chebyshev.c:241-269 */
double
orc_cheb_error_minvsqrt (int n, const double *y, const double *z,
                         orc_atimes_fn atimes, void *ud)
{
  double *x = malloc (sizeof (double) * n);
  atimes (n, z, x, ud);
  double yy = 0.0, zz = 0.0;
  for (int k = 0; k < n; k++) { yy += y[k] * y[k]; zz += z[k] * x[k]; }
  free (x);
  return fabs (zz - yy) / zz;
}

/* chebyshev.c:282-310 */
double
orc_cheb_error_Rsqrt (int n, const double *y, const double *z,
                      orc_atimes_fn atimes, void *ud)
{
  double *x = malloc (sizeof (double) * n);
  atimes (n, y, x, ud);
  double yy = 0.0, zz = 0.0;
  for (int k = 0; k < n; k++) { yy += y[k] * x[k]; zz += z[k] * z[k]; }
  free (x);
  return fabs (yy - zz) / yy;
}

void
orc_atimes_sys_cb (int n, const double *x, double *y, void *ud)
{
  (void) n;
  orc_atimes_3all ((const orc_sys *) ud, x, y);
}

void
orc_atimes_dense_cb (int n, const double *x, double *y, void *ud)
{
  const orc_densemat *d = (const orc_densemat *) ud;
  orc_dot_prod_matrix (d->m, n, n, x, y);
}

/* ------------------------------------------------------------------ */
/* MT19937 (Matsumoto & Nishimura 1998) + Marsaglia polar Gaussian      */
/* ------------------------------------------------------------------ */

/* KIrand.c:137-153 (init_genrand) */
void
orc_rng_seed (orc_rng *r, unsigned long s)
{
  r->mt[0] = s & 0xffffffffUL;
  for (int i = 1; i < 624; i++)
    r->mt[i] = (1812433253UL * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30))
                + (unsigned long) i) & 0xffffffffUL;
  r->mti = 624;
  r->has_saved = 0;
  r->saved = 0.0;
}

/* KIrand.c:201-250 (genrand_int32) */
unsigned long
orc_rng_int32 (orc_rng *r)
{
  const unsigned long upper = 0x80000000UL, lower = 0x7fffffffUL;
  const unsigned long A = 0x9908b0dfUL;
  if (r->mti >= 624)
    {
      for (int k = 0; k < 624; k++)
        {
          unsigned long y = (r->mt[k] & upper) | (r->mt[(k + 1) % 624] & lower);
          r->mt[k] = r->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1UL) ? A : 0UL);
        }
      r->mti = 0;
    }
  unsigned long y = r->mt[r->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680UL;
  y ^= (y << 15) & 0xefc60000UL;
  y ^= (y >> 18);
  return y & 0xffffffffUL;
}

/* KIrand.c genrand_real2: [0,1) */
double
orc_rng_real2 (orc_rng *r)
{
  return (double) orc_rng_int32 (r) * (1.0 / 4294967296.0);
}

/* KIrand.c:108-135 (KIrand_Gaussian), Marsaglia polar method. */
double
orc_rng_gaussian (orc_rng *r)
{
  if (r->has_saved)
    {
      r->has_saved = 0;
      return r->saved;
    }
  double x, y, q;
  do
    {
      x = 2.0 * orc_rng_real2 (r) - 1.0;
      y = 2.0 * orc_rng_real2 (r) - 1.0;
      q = x * x + y * y;
    }
  while (q > 1.0 || q == 0.0);
  q = sqrt (-2.0 * log (q) / q);
  r->saved = y * q;
  r->has_saved = 1;
  return x * q;
}

/* brownian.c:1697-1706: z = M^{-1/2} y by Chebyshev with fixed bounds. */
void
orc_bd_noise_F (const orc_sys *s, int ncheb, const double *acoef,
                double e0, double e1, const double *y, double *z)
{
  int n = 3 * s->np;
  orc_cheb_eval_atimes (ncheb, acoef, n, y, z, e0, e1, orc_atimes_sys_cb,
                        (void *) s);
}

/* brownian.c:299-375 (check_overlap): number of (pair, image) entries with
 * r^2 < rmin (a_i + a_j)^2 over pairs i <= j, the primary cell for i != j
 * and the 26 neighbouring cells of the raw separation when periodic. */
int
orc_check_overlap (const orc_sys *s, const double *pos, double rmin)
{
  return orc_check_overlap_n (s, pos, s->np, rmin);
}

/* the same over the first n particles (check_overlap loops over the
 * mobile particles only, brownian.c:305-310) */
int
orc_check_overlap_n (const orc_sys *s, const double *pos, int n, double rmin)
{
  int flag = 0;
  for (int i = 0; i < n; i++)
    for (int j = i; j < n; j++)
      {
        double x = pos[3 * i] - pos[3 * j];
        double y = pos[3 * i + 1] - pos[3 * j + 1];
        double z = pos[3 * i + 2] - pos[3 * j + 2];
        double a2 = s->a ? (s->a[i] + s->a[j]) * (s->a[i] + s->a[j]) : 4.0;
        if (i != j && x * x + y * y + z * z < rmin * a2) flag++;
        if (!s->periodic) continue;
        for (int mx = -1; mx <= 1; mx++)
          for (int my = -1; my <= 1; my++)
            for (int mz = -1; mz <= 1; mz++)
              {
                if (mx == 0 && my == 0 && mz == 0) continue;
                double xx = x + mx * s->lx, yy = y + my * s->ly, zz = z + mz * s->lz;
                if (s->shear_mode == 1) xx += my * s->shear_shift;
                else if (s->shear_mode == 2) xx += mz * s->shear_shift;
                if (xx * xx + yy * yy + zz * zz < rmin * a2) flag++;
              }
      }
  return flag;
}

/* brownian.c:2177-2187 */
static double
orc_bd_fact (double peclet, double dt)
{
  return peclet > 0.0 ? sqrt (2.0 / (peclet * dt)) : 0.0;
}

/* ode-quaternion.c:58-79 (quaternion_Winv) and 89-105 (quaternion_dQdt):
 * dQ/dt = 1/2 W(Q)^{-1} (O, 0), W(Q)^{-1} = W(Q)^T / |Q|^2 */
void
orc_quat_dqdt (const double *Q, const double *O, double *dQdt)
{
  double den = Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3];
  double W[16] = {
    +Q[3] / den, -Q[2] / den, +Q[1] / den, +Q[0] / den,
    +Q[2] / den, +Q[3] / den, -Q[0] / den, +Q[1] / den,
    -Q[1] / den, +Q[0] / den, +Q[3] / den, +Q[2] / den,
    -Q[0] / den, -Q[1] / den, -Q[2] / den, +Q[3] / den };
  for (int i = 0; i < 4; i++)
    {
      dQdt[i] = 0.0;
      for (int j = 0; j < 3; j++) dQdt[i] += 0.5 * W[i * 4 + j] * O[j];
    }
}

/* The FTS mobility problem with E = 0 restricted to (F, T) -> (U, O) of
 * the nm mobile particles, the np - nm fixed ones at rest (U = O = E = 0):
 * BD_atimes_mob_FU, brownian.c:597-729, which solves it with GMRES through
 * solve_mob_3fts_0 (nm = np) or solve_mix_3fts_0.  The true operator M
 * (11 np)^2 is the extracted-form matrix (orc_matrix_ewald / _nonewald)
 * times the S transform of fts.c:660-670 (columns 6 and 10 of a particle:
 * 2 m_6 + m_10 and 2 m_10 + m_6, columns 7-9: 2 m_c); with K the known
 * forces (F, T of the mobile particles), X the unknowns (S of the mobile,
 * all 11 of the fixed particles) and R the rows with known zero (E of the
 * mobile, all 11 of the fixed particles): M[R, X] xi = -M[R, K] ft by
 * Gaussian elimination with partial pivoting, (U, O) = M[UO, K] ft +
 * M[UO, X] xi.  Exact to rounding: the limit of the reference's
 * iteration.  O(np^3); small systems only. */
static void orc_mix_known (const orc_sys *s, int nm, int D, int qs,
                           const double *ft, const double *vr, double *uo);

void
orc_fts_mix_ft (const orc_sys *s, int nm, const double *ft, double *uo)
{
  orc_mix_known (s, nm, 11, 6, ft, NULL, uo);
}

/* The same elimination for D values per particle (3 F, 6 FT, 11 FTS), the
 * first qs of a mobile particle known (forces), the rest of it and all D
 * of a fixed particle unknown, with zero velocity on the unknowns' rows.
 * F / FT (D = qs): the mixed problem of solve_mix_3f / solve_mix_3ft
 * (ewald-3f.c:269-330, ewald-3ft.c) with the fixed particles at rest, U_m
 * out; the extracted form is the true operator there (no S transform). */
static void
orc_mix_known (const orc_sys *s, int nm, int D, int qs, const double *ft,
               const double *vr, double *uo)
{
  /* vr: the known values on the unknowns' rows (velocities of the fixed
   * particles, strains), in the order of the unknowns; NULL: zero */
  const int np = s->np, n = D * np;
  const int nk = qs * nm, nx = n - nk;
  int *K = malloc (sizeof (int) * (nk > 0 ? nk : 1));
  int *X = malloc (sizeof (int) * (nx > 0 ? nx : 1));
  for (int i = 0, a = 0, b = 0; i < np; i++)
    for (int q = 0; q < D; q++)
      {
        if (i < nm && q < qs) K[a++] = D * i + q;
        else X[b++] = D * i + q;
      }
  /* the rows of known zero velocity are the unknowns' indices */
  const int *R = X;
  double *M = malloc (sizeof (double) * (size_t) n * n);
  double *A = malloc (sizeof (double) * (size_t) nx * nx);
  double *b = malloc (sizeof (double) * nx);
  double *S = malloc (sizeof (double) * nx);
  if (s->periodic) orc_matrix_ewald (s, M);
  else orc_matrix_nonewald (s, M);
  for (int r = 0; r < n && D == 11; r++)
    for (int p = 0; p < np; p++)
      {
        double *m = M + (size_t) r * n + 11 * p + 6;
        double m6 = m[0], m10 = m[4];
        m[0] = 2.0 * m6 + m10;
        m[1] *= 2.0;
        m[2] *= 2.0;
        m[3] *= 2.0;
        m[4] = 2.0 * m10 + m6;
      }
  for (int a = 0; a < nx; a++)
    {
      double acc = 0.0;
      for (int c = 0; c < nk; c++) acc += M[(size_t) R[a] * n + K[c]] * ft[c];
      b[a] = (vr ? vr[a] : 0.0) - acc;
      for (int c = 0; c < nx; c++) A[(size_t) a * nx + c] = M[(size_t) R[a] * n + X[c]];
    }
  for (int k = 0; k < nx; k++)
    {
      int p = k;
      for (int r = k + 1; r < nx; r++)
        if (fabs (A[(size_t) r * nx + k]) > fabs (A[(size_t) p * nx + k])) p = r;
      if (p != k)
        {
          for (int c = 0; c < nx; c++)
            {
              double t = A[(size_t) k * nx + c];
              A[(size_t) k * nx + c] = A[(size_t) p * nx + c];
              A[(size_t) p * nx + c] = t;
            }
          double t = b[k]; b[k] = b[p]; b[p] = t;
        }
      for (int r = k + 1; r < nx; r++)
        {
          double f = A[(size_t) r * nx + k] / A[(size_t) k * nx + k];
          for (int c = k; c < nx; c++) A[(size_t) r * nx + c] -= f * A[(size_t) k * nx + c];
          b[r] -= f * b[k];
        }
    }
  for (int k = nx - 1; k >= 0; k--)
    {
      double acc = b[k];
      for (int c = k + 1; c < nx; c++) acc -= A[(size_t) k * nx + c] * S[c];
      S[k] = acc / A[(size_t) k * nx + k];
    }
  for (int r = 0; r < nk; r++)
    {
      double acc = 0.0;
      for (int c = 0; c < nk; c++) acc += M[(size_t) K[r] * n + K[c]] * ft[c];
      for (int c = 0; c < nx; c++) acc += M[(size_t) K[r] * n + X[c]] * S[c];
      uo[r] = acc;
    }
  free (K); free (X); free (M); free (A); free (b); free (S);
}

void
orc_fts_mob_ft (const orc_sys *s, const double *ft, double *uo)
{
  orc_fts_mix_ft (s, s->np, ft, uo);
}

/* the mobility of BD_atimes_mob_FU (brownian.c:494-739), every particle
 * mobile, equal radii (the Global / Particle scalings are identities) */
void
orc_bd_op_cb (int n, const double *x, double *y, void *ud)
{
  (void) n;
  const orc_sys *s = (const orc_sys *) ud;
  if (s->version == 2) orc_fts_mob_ft (s, x, y);
  else orc_atimes_3all (s, x, y);
}

/* the same with the particles nm..np-1 fixed: the noise operator of F / FT
 * zero-pads the forces of the fixed particles and keeps the mobile rows
 * (brownian.c:521-547, 562-588), that of FTS is the mixed problem with
 * those particles at rest (650-729); the drift (solve_mix_3all) holds them
 * at rest in every version */
typedef struct { const orc_sys *s; int nm; } orc_bd_ctx;

static void orc_bd_mob (int n, const double *x, double *y, void *ud,
                        int scaled);

/* the noise operator (BD_atimes_mob_FU, with its Global scaling) */
static void
orc_bd_op_nm_cb (int n, const double *x, double *y, void *ud)
{
  orc_bd_mob (n, x, y, ud, 1);
}

/* the drift product (solve_mix_3all: no scaling) */
static void
orc_bd_drift (int n, const double *x, double *y, void *ud)
{
  orc_bd_mob (n, x, y, ud, 0);
}

/* the imposed flow at position p: w[0..2] = Ui + Oi x p + Ei . p, w[3..5]
 * = Oi, w[6..10] = Ei (Ei extracted: xx xy xz yz yy; f.c:569-660,
 * ft.c:901-955, fts.c:2303-2320) */
static void
orc_flow_at (const orc_sys *s, const double *p, double *w)
{
  const double *E = s->Ei, Ezz = -s->Ei[0] - s->Ei[4];
  w[0] = s->Ui[0] + s->Oi[1] * p[2] - s->Oi[2] * p[1]
    + E[0] * p[0] + E[1] * p[1] + E[2] * p[2];
  w[1] = s->Ui[1] + s->Oi[2] * p[0] - s->Oi[0] * p[2]
    + E[1] * p[0] + E[4] * p[1] + E[3] * p[2];
  w[2] = s->Ui[2] + s->Oi[0] * p[1] - s->Oi[1] * p[0]
    + E[2] * p[0] + E[3] * p[1] + Ezz * p[2];
  for (int q = 0; q < 3; q++) w[3 + q] = s->Oi[q];
  for (int q = 0; q < 5; q++) w[6 + q] = E[q];
}

static int
orc_has_flow (const orc_sys *s)
{
  for (int k = 0; k < 3; k++)
    if (s->Ui[k] != 0.0 || s->Oi[k] != 0.0) return 1;
  for (int k = 0; k < 5; k++)
    if (s->Ei[k] != 0.0) return 1;
  return 0;
}

/* the drift in an imposed flow: solve_mix_3all in the labo frame
 * (ewald-3f.c:269-321, ewald-3ft.c:323-382, ewald-3fts.c:603-667 with the
 * BD's zero labo-frame E and fixed-particle velocities): known values to
 * the rest frame (mobile E = -Ei; fixed particle k: -(u_inf, Oi, Ei) at the
 * position of particle k, as the reference's shift takes it), the mixed
 * problem by elimination, the mobile velocities back with the flow at
 * their positions */
static void
orc_bd_drift_flow (const orc_sys *s, int nm, const double *x, double *y)
{
  const int D = s->version == 0 ? 3 : (s->version == 1 ? 6 : 11);
  const int qs = s->version == 0 ? 3 : 6, np = s->np;
  const int nx = D * np - qs * nm;
  double *vr = malloc (sizeof (double) * (nx > 0 ? nx : 1));
  double *uo = malloc (sizeof (double) * qs * nm);
  for (int i = 0, a = 0; i < np; i++)
    for (int q = 0; q < D; q++)
      {
        if (i < nm && q < qs) continue;
        double w[11];
        orc_flow_at (s, s->pos + 3 * (i < nm ? i : i - nm), w);
        vr[a++] = -w[q];
      }
  orc_mix_known (s, nm, D, qs, x, vr, uo);
  for (int i = 0; i < nm; i++)
    {
      double w[11];
      orc_flow_at (s, s->pos + 3 * i, w);
      for (int q = 0; q < qs; q++) y[qs * i + q] = uo[qs * i + q] + w[q];
    }
  free (vr);
  free (uo);
}

static void
orc_bd_mob (int n, const double *x, double *y, void *ud, int scaled)
{
  const orc_bd_ctx *c = (const orc_bd_ctx *) ud;
  const orc_sys *s = c->s;
  const int d = s->version == 0 ? 3 : 6;
  if (!scaled && orc_has_flow (s))
    {
      orc_bd_drift_flow (s, c->nm, x, y);
      return;
    }
  if (c->nm == s->np) orc_bd_op_cb (n, x, y, (void *) s);
  else if (s->version == 2) orc_fts_mix_ft (s, c->nm, x, y);
  else if (!scaled)
    /* the drift (solve_mix_3f / 3ft): fixed particles at rest */
    orc_mix_known (s, c->nm, d, d, x, NULL, y);
  else
    {
      const int nf = d * s->np, nmd = d * c->nm;
      double *f = calloc ((size_t) nf, sizeof (double));
      double *u = malloc (sizeof (double) * nf);
      for (int i = 0; i < nmd; i++) f[i] = x[i];
      orc_atimes_3all (s, f, u);
      for (int i = 0; i < nmd; i++) y[i] = u[i];
      free (f); free (u);
    }
  /* Particle-wise to Global scaling (BD_scale_P_to_G_uo, brownian.c:430-480):
   * U / a, O / a^3 */
  if (scaled && s->a)
    for (int i = 0; i < c->nm; i++)
      for (int q = 0; q < d; q++)
        y[d * i + q] /= q < 3 ? s->a[i] : s->a[i] * s->a[i] * s->a[i];
}

/* One BD step, every particle mobile, no interaction forces, the drift in
 * the system's imposed flow (Ui, Oi, Ei; zero: a fluid at rest): scheme 0 the mid-point algorithm (BD_evolve_mid,
 * brownian.c:2352-2560), 1 Banchio-Brady 2003 with step parameter bb_n
 * (BD_evolve_BB03, 2579-2920), 2 Ball-Melrose 1997 (BD_evolve_BM97,
 * 2970-3190), with BD_calc_FB's atimes route (1545-1874, spectral bounds
 * and coefficients given, so the refit of 1707-1727 does not happen),
 * BD_calc_forces (2273-2311) and evolve_Euler_3all (1889-1930).  The
 * mobility is orc_bd_op_cb: one atimes_3all for F (brownian.c:512-520) and
 * FT (559-587), the E = 0 mobility problem for FTS (597-646).  For FT /
 * FTS the noise is particle-wise (F, T) (the reference reorders it to
 * [F; T] and back, 1851-1871 and solve_mix_3all), f = F + fact z_F, t = T
 * + fact z_T.
 *   mid:  u0 = M(x) f, x1 = x + dt/2 u0 | u = M(x1) f, x' = x + dt u
 *   BM97: u0 = M(x) f, x1 = x + dt u0   | u = (u0 + M(x1) f) / 2, x' = x + dt u
 *   BB03: u0 = M(x) fact z, x1 = x + dt/n u0 | uB = u0 + n/2 (M(x1) fact z
 *         - u0), u = M(x) F + uB (the deterministic product back at x and
 *         the shift of t: BD_add_FP (BD, x, f) moves the system to x before
 *         it, brownian.c:2209-2212, 2800-2816), x' = x + dt u
 * with the overlap checks after x1 (stage 1) and x' (stage 2).
 *
 * Particles nm..np-1 are fixed (BD_atimes_mob_FU's zero padding for F /
 * FT, the mixed problem for FTS; the overlap check over the mobile ones).
 * Unequal radii: the noise operator carries BD_atimes_mob_FU's Global
 * scaling (U / a, O / a^3); the drift products do not.
 * draws[ndraw][n] (n = 3 nm or 6 nm) are the Gaussian vectors BD_calc_FB
 * consumes, one per (re)start of the noise.  fix_fact != 0 recomputes fact
 * after a redraw resets dt (the reference keeps the halved-dt value).  x
 * [3 nm] (and q [4 nm] for FT / FTS when not NULL) are updated in place on
 * success; returns the dt taken, or -1 if the draws ran out. */
double
orc_bd_step_nm (orc_sys *s, int nm, int scheme, double bb_n, double *x,
                double *q, const double *F, const double *T, double peclet,
                double dt, double dt_lim, double rmin_ol, int ncheb,
                const double *acoef, double e0, double e1, int ndraw,
                const double *draws, int fix_fact, int *used, int *halvings,
                double *err)
{
  const int np = nm, ver = s->version;
  const int n = np * (ver == 0 ? 3 : 6), n3 = 3 * np;
  orc_bd_ctx ctx = {s, nm};
  /* positions of all s->np particles: the mobile ones from x or x_mid,
   * the fixed ones as the system holds them (stokes_set_pos_mobile) */
  double *pfull = malloc (sizeof (double) * 3 * (size_t) s->np);
  for (int i = 3 * nm; i < 3 * s->np; i++) pfull[i] = s->pos[i];
#define SETPOS(src) do { for (int i_ = 0; i_ < n3; i_++) pfull[i_] = (src)[i_]; s->pos = pfull; } while (0)
  double *z = malloc (sizeof (double) * n);
  double *f = malloc (sizeof (double) * n);
  double *u = malloc (sizeof (double) * n);
  double *u0 = malloc (sizeof (double) * n);
  double *xmid = malloc (sizeof (double) * n3);
  double dt_local = dt;
  double fact = orc_bd_fact (peclet, dt_local);
  int k = 0, nh = 0, stage;
  double ret = -1.0;
  const double *pos0 = s->pos;
#define UIDX(i) (ver == 0 ? (i) : 6 * ((i) / 3) + ((i) % 3))
  /* f = (with_ext ? F, T : 0) + zfac * fact z */
#define FORCES(with_ext, zfac)                                              \
  for (int i = 0; i < n; i++)                                               \
    {                                                                       \
      const double *ext = ver == 0 ? F : ((i % 6) < 3 ? F : T);             \
      int ie = ver == 0 ? i : 3 * (i / 6) + (i % 3);                        \
      f[i] = ((with_ext) && ext ? ext[ie] : 0.0) + (zfac) * fact * z[i];    \
    }

  /* sheared systems: the cell shift at time tt (stokes_get_shear_shift,
   * stokes.c:554-587: s0 + L rate (tt - t0), L = ly or lz, brought into
   * [-lx/2, lx/2)), set where the reference calls stokes_set_shear_shift
   * (brownian.c:2393, 2468; 2630, 2741, 2805; 3017, 3092) */
#define SHIFT(tt) do { if (s->shear_mode) {                                 \
    double sh_ = s->bd_s0 + (s->shear_mode == 1 ? s->ly : s->lz)            \
      * s->shear_rate * ((tt) - s->bd_t0);                                  \
    while (sh_ < -0.5 * s->lx) sh_ += s->lx;                                \
    while (sh_ >= 0.5 * s->lx) sh_ -= s->lx;                                \
    s->shear_shift = sh_; } } while (0)
redo:                              /* REDO */
  SETPOS (x);
  SHIFT (s->bd_t);
redo_fb:                           /* REDO_FB */
  if (k >= ndraw) goto done;
  orc_cheb_eval_atimes (ncheb, acoef, n, draws + (size_t) k * n, z, e0, e1,
                        orc_bd_op_nm_cb, (void *) &ctx);
  *err = orc_cheb_error_minvsqrt (n, draws + (size_t) k * n, z,
                                  orc_bd_op_nm_cb, (void *) &ctx);
  k++;
redo_scale:                        /* REDO_scale */
  /* the positions are not set here (REDO_scale, brownian.c:2400, 2637,
   * 3025): after an overlap in the second stage with dt / 2 still above
   * dt_lim (brownian.c:2523, 3162) the mid-point and BM97 systems keep the
   * intermediate configuration of the failed attempt, and the first-stage
   * product of the retry is taken there; BB03 has moved back to x for its
   * deterministic product (2808), so its retries start from x; the noise
   * is kept (computed where it was drawn) */
  for (int i = 0; i < n3; i++) xmid[i] = x[i];
  if (scheme == 1)
    {
      FORCES (0, 1.0);
    }
  else
    {
      FORCES (1, 1.0);
    }
  orc_bd_drift (n, f, u0, (void *) &ctx);
  {
    const double h = scheme == 0 ? dt_local * 0.5
      : (scheme == 1 ? dt_local / bb_n : dt_local);
    for (int i = 0; i < n3; i++) xmid[i] += h * u0[UIDX (i)];
  }
  stage = 1;
  if (s->rmin == 0.0 && orc_check_overlap_n (s, xmid, nm, rmin_ol) > 0)
    goto overlap;
  SETPOS (xmid);
  SHIFT (s->bd_t + (scheme == 0 ? dt_local * 0.5
                    : (scheme == 1 ? dt_local / bb_n : dt_local)));
  orc_bd_drift (n, f, u, (void *) &ctx);
  if (scheme == 2)
    for (int i = 0; i < n; i++) u[i] = 0.5 * (u0[i] + u[i]);
  else if (scheme == 1)
    {
      const double bbf = 0.5 * bb_n;
      for (int i = 0; i < n; i++) u0[i] = u0[i] + bbf * (u[i] - u0[i]);
      FORCES (1, 0.0);
      SHIFT (s->bd_t);                         /* brownian.c:2805 */
      SETPOS (x);                              /* BD_add_FP, 2808 */
      orc_bd_drift (n, f, u, (void *) &ctx);
      for (int i = 0; i < n; i++) u[i] += u0[i];
    }
  for (int i = 0; i < n3; i++) xmid[i] = x[i] + dt_local * u[UIDX (i)];
  stage = 2;
  if (s->rmin == 0.0 && orc_check_overlap_n (s, xmid, nm, rmin_ol) > 0)
    goto overlap;
  for (int i = 0; i < n3; i++) x[i] = xmid[i];
  if (ver > 0 && q)
    for (int i = 0; i < np; i++)
      {
        double dq[4];
        orc_quat_dqdt (q + 4 * i, u + 6 * i + 3, dq);
        for (int c = 0; c < 4; c++) q[4 * i + c] += dt_local * dq[c];
      }
  ret = dt_local;
  goto done;
overlap:
  dt_local *= 0.5;
  nh++;
  if (dt_local < dt_lim)
    {
      dt_local = dt;
      if (fix_fact) fact = orc_bd_fact (peclet, dt_local);
      if (stage == 1) goto redo_fb;
      goto redo;
    }
  fact = orc_bd_fact (peclet, dt_local);
  goto redo_scale;
done:
#undef UIDX
#undef FORCES
#undef SETPOS
#undef SHIFT
  s->pos = pos0;
  free (pfull);
  *used = k;
  *halvings = nh;
  free (z); free (f); free (u); free (u0); free (xmid);
  return ret;
}

/* every particle mobile */
double
orc_bd_step (orc_sys *s, int scheme, double bb_n, double *x, double *q,
             const double *F, const double *T, double peclet, double dt,
             double dt_lim, double rmin_ol, int ncheb, const double *acoef,
             double e0, double e1, int ndraw, const double *draws,
             int fix_fact, int *used, int *halvings, double *err)
{
  return orc_bd_step_nm (s, s->np, scheme, bb_n, x, q, F, T, peclet, dt,
                         dt_lim, rmin_ol, ncheb, acoef, e0, e1, ndraw, draws,
                         fix_fact, used, halvings, err);
}

/* the mid-point step (orc_bd_step, scheme 0) */
double
orc_bd_mid (orc_sys *s, double *x, double *q, const double *F,
            const double *T, double peclet, double dt, double dt_lim,
            double rmin_ol, int ncheb, const double *acoef, double e0,
            double e1, int ndraw, const double *draws, int fix_fact,
            int *used, int *halvings, double *err)
{
  return orc_bd_step (s, 0, 0.0, x, q, F, T, peclet, dt, dt_lim, rmin_ol,
                      ncheb, acoef, e0, e1, ndraw, draws, fix_fact, used,
                      halvings, err);
}

/* the F-version step (orc_bd_mid with no torques or orientations) */
double
orc_bd_mid_F (orc_sys *s, double *x, const double *F, double peclet,
              double dt, double dt_lim, double rmin_ol, int ncheb,
              const double *acoef, double e0, double e1, int ndraw,
              const double *draws, int fix_fact, int *used, int *halvings,
              double *err)
{
  return orc_bd_mid (s, x, NULL, F, NULL, peclet, dt, dt_lim, rmin_ol, ncheb,
                     acoef, e0, e1, ndraw, draws, fix_fact, used, halvings,
                     err);
}
