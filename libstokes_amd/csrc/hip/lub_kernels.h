// lub_kernels.h -- the lubrication correction y = L x on the device (FP64).
//
// Directed pairs: particle i takes, for every particle j and image cell m in
// -imax..imax with 0 < r_ij+m and r (or s = 2 r / (a_i + a_j)) < lubmax,
//   y_i += self(x_i; e; S11) + pair(x_j; e; S12),  e = (x_j + L m - x_i) / r,
// with S11, S12 the correction scalars of the pair seen from i (lambda =
// a_j / a_i).  This is calc_lub_3fts's loop (lub.c:576-768: each unordered
// pair once, both spheres updated) written per receiving particle; a
// periodic self image (i == j, m != 0) counts twice, as the reference adds
// both of its halves to the same sphere.  Launches: count per i, exclusive
// scan, fill the pair list in a fixed order, evaluate one pair per thread,
// gather per (i, component) in list order: no atomics, reproducible.
#pragma once
#include "lub_math.h"

namespace shk {

struct LubGeom {
  int np, periodic, poly;
  double lx, ly, lz;
  double s2x, s3x;          // shear: x shift per y-image / z-image
  int ix, iy, iz;           // image range
  double lubmax2;           // 0: no limit
};

// image cell code: 5 bits per direction, offset 15
SHK_HD int lub_img_code (int a, int b, int c) { return (a + 15) | ((b + 15) << 5) | ((c + 15) << 10); }

__device__ inline bool lub_hit (const LubGeom &g, const double4 &pi, const double4 &pj, int a,
                                int b, int c, double *d)
{
  d[0] = pj.x + g.lx * a + g.s2x * b + g.s3x * c - pi.x;
  d[1] = pj.y + g.ly * b - pi.y;
  d[2] = pj.z + g.lz * c - pi.z;
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (r2 == 0.0) return false;
  if (g.poly) r2 *= 4.0 / ((pi.w + pj.w) * (pi.w + pj.w));
  return g.lubmax2 <= 0.0 || r2 < g.lubmax2;
}

// one wave per receiving particle i, lanes over j; FILL = 0 counts, 1
// writes the (j, cell) records from off[i] in (j chunk, cell, lane) order
template <int FILL>
__global__ __launch_bounds__ (256) void
k_lub_pairs (const double4 *__restrict__ pos, LubGeom g, int *__restrict__ cnt,
             const int *__restrict__ off, int2 *__restrict__ list)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= g.np) return;
  const double4 pi = pos[i];
  int base = FILL ? off[i] : 0;
  for (int j0 = 0; j0 < g.np; j0 += 64)
    {
      const int j = j0 + lane;
      const bool valid = j < g.np;
      const double4 pj = pos[valid ? j : i];
      for (int a = -g.ix; a <= g.ix; a++)
        for (int b = -g.iy; b <= g.iy; b++)
          for (int c = -g.iz; c <= g.iz; c++)
            {
              double d[3];
              const bool hit = valid && lub_hit (g, pi, pj, a, b, c, d);
              const unsigned long long m = __ballot (hit);
              if (FILL && hit)
                {
                  const unsigned long long below = lane ? (m & ((~0ull) >> (64 - lane))) : 0ull;
                  list[base + __popcll (below)] = make_int2 (j, lub_img_code (a, b, c));
                }
              base += __popcll (m);
            }
    }
  if (!FILL && lane == 0) cnt[i] = base;
}

// exclusive scan of cnt[0..n) into off[0..n] (one workgroup; n moderate)
__global__ __launch_bounds__ (1024) void
k_lub_scan (const int *__restrict__ cnt, int n, int *__restrict__ off)
{
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024, lo = t * per, hi = min (n, lo + per);
  int s = 0;
  for (int k = lo; k < hi; k++) s += cnt[k];
  part[t] = s;
  __syncthreads ();
  if (t == 0)
    {
      int run = 0;
      for (int k = 0; k < 1024; k++)
        {
          const int v = part[k];
          part[k] = run;
          run += v;
        }
    }
  __syncthreads ();
  int run = part[t];
  for (int k = lo; k < hi; k++)
    {
      off[k] = run;
      run += cnt[k];
    }
  if (hi == n && lo < hi) off[n] = run;
  if (n == 0 && t == 0) off[0] = 0;
}

// f values of the device coefficient table at one lambda: f_k = sum_q c_q l^q
struct DevPolyF {
  const double *c;   // [fam][k (N + 2) + q]
  int N;
  double l;
  __device__ double operator() (int fam, int k) const
  {
    const double *p = c + ((size_t) fam * (N + 1) + k) * (N + 2);
    double v = 0.0;
    for (int q = k + 1; q >= 0; q--) v = v * l + p[q];
    return v;
  }
};

// f values at lambda = 1 tabulated [fam][k], k <= 50
struct DevTabF {
  const double *t;
  __device__ double operator() (int fam, int k) const { return t[fam * 51 + k]; }
};

struct LubEval {
  double lubmin2;
  int nmax, flag_lub;
  const double *coef;   // poly: coefficient table
  int N;
  const double *f1;     // equal spheres: f_k(1), k <= 50
};

// the correction scalars of directed pair p (S11, S12 seen from the
// receiving particle i), the unit vector e and the weight (2 for a periodic
// self image)
template <int V>
__device__ inline void lub_pair (const double4 *__restrict__ pos, const LubGeom &g,
                                 const LubEval &ev, const int *__restrict__ off,
                                 const int2 *__restrict__ list, int p, int &i, int &j, double *e,
                                 double *s11, double *s12, double &w)
{
  // the receiving particle: the last i with off[i] <= p
  int lo = 0, hi = g.np;
  while (hi - lo > 1)
    {
      const int mid = (lo + hi) >> 1;
      if (off[mid] <= p) lo = mid;
      else hi = mid;
    }
  i = lo;
  const int2 rec = list[p];
  j = rec.x;
  const int a = (rec.y & 31) - 15, b = ((rec.y >> 5) & 31) - 15, c = ((rec.y >> 10) & 31) - 15;
  const double4 pi = pos[i], pj = pos[j];
  double d[3];
  d[0] = pj.x + g.lx * a + g.s2x * b + g.s3x * c - pi.x;
  d[1] = pj.y + g.ly * b - pi.y;
  d[2] = pj.z + g.lz * c - pi.z;
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double two[22], m[44];
  for (int k = 0; k < 22; k++) two[k] = 0.0;
  if (!g.poly)
    {
      // calc_lub_fts_2b (fts.c:1758-1869): r clamped, the equal-sphere series
      if (r2 < ev.lubmin2) r2 = ev.lubmin2;
      const double r = sqrt (r2);
      lub::lub_form (V, DevTabF{ev.f1}, 50, 1.0, r, 2.0, two);
      lub::scale (V, two, 1.0, 1.0, false);
      lub::minv (V, r, 1.0, 1.0, m);
      for (int f = 0; f < lub::nfam_of (V); f++)
        {
          s11[f] = two[2 * f] - m[4 * f];
          s12[f] = two[2 * f + 1] - m[4 * f + 1];
        }
    }
  else
    {
      // calc_lub_fts_2b_poly (fts.c:1991-2147) seen from i: lambda = a_j / a_i
      const double ai = pi.w, aj = pj.w, rs = 0.25 * (ai + aj) * (ai + aj);
      if (r2 / rs < ev.lubmin2) r2 = rs * ev.lubmin2;
      const double r = sqrt (r2), l = aj / ai, s = 2.0 * r / (ai + aj);
      const int n = ev.nmax % 2 ? ev.nmax - 1 : ev.nmax;
      const DevPolyF fv{ev.coef, ev.N, l};
      if (ev.flag_lub) lub::lub_form (V, fv, n - 1, l, s, 1.0, two);
      else lub::far_form (V, fv, n, l, s, two);
      lub::scale (V, two, ai, l, false);
      lub::minv (V, r, ai, aj, m);
      for (int f = 0; f < lub::nfam_of (V); f++)
        {
          const int pw = lub::sd_power (f);
          const double dv = pw == 1 ? ai : (pw == 2 ? ai * ai : ai * ai * ai);
          s11[f] = (two[2 * f] - m[4 * f]) / dv;
          s12[f] = (two[2 * f + 1] - m[4 * f + 1]) / dv;
        }
    }
  for (int f = lub::nfam_of (V); f < 11; f++) s11[f] = s12[f] = 0.0;
  const double r = sqrt (r2);
  e[0] = d[0] / r;
  e[1] = d[1] / r;
  e[2] = d[2] / r;
  w = (i == j) ? 2.0 : 1.0;
}

template <int V>
__global__ __launch_bounds__ (128) void
k_lub_eval (const double4 *__restrict__ pos, LubGeom g, LubEval ev, const int *__restrict__ off,
            const int2 *__restrict__ list, int npairs, const double *__restrict__ x,
            double *__restrict__ val)
{
  constexpr int D = Dim<V>::D;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  int i, j;
  double e[3], s11[11], s12[11], w;
  lub_pair<V> (pos, g, ev, off, list, p, i, j, e, s11, s12, w);
  double y[11];
  for (int k = 0; k < D; k++) y[k] = 0.0;
  lub::apply_block<D> (true, x + (size_t) D * i, y, e[0], e[1], e[2], s11);
  lub::apply_block<D> (false, x + (size_t) D * j, y, e[0], e[1], e[2], s12);
  double *o = val + (size_t) p * D;
  for (int k = 0; k < D; k++) o[k] = w * y[k];
}

// per pair: e, w, S11, S12 (26 values) for the dense build
template <int V>
__global__ __launch_bounds__ (128) void
k_lub_scal (const double4 *__restrict__ pos, LubGeom g, LubEval ev, const int *__restrict__ off,
            const int2 *__restrict__ list, int npairs, double *__restrict__ sc)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  int i, j;
  double e[3], s11[11], s12[11], w;
  lub_pair<V> (pos, g, ev, off, list, p, i, j, e, s11, s12, w);
  double *o = sc + (size_t) p * 26;
  o[0] = e[0]; o[1] = e[1]; o[2] = e[2]; o[3] = w;
  for (int k = 0; k < 11; k++) { o[4 + k] = s11[k]; o[15 + k] = s12[k]; }
}

// the dense extracted-form matrix (make_matrix_lub_3fts, lub-matrix.c:
// 373-507, zeroed beforehand): the workgroup of particle i adds, pair by pair
// in list order, the self block at (i, i) and the pair block at (i, j); only
// this workgroup writes row block i
template <int V>
__global__ __launch_bounds__ (128) void
k_lub_matrix (const int *__restrict__ off, const int2 *__restrict__ list,
              const double *__restrict__ sc, int np, double *__restrict__ mat)
{
  constexpr int D = Dim<V>::D;
  __shared__ double Bs[D * D], Bp[D * D];
  const int i = blockIdx.x, t = threadIdx.x;
  const size_t n = (size_t) np * D;
  const int r = t / D, c = t - r * D;
  double acc = 0.0;
  for (int p = off[i]; p < off[i + 1]; p++)
    {
      const double *o = sc + (size_t) p * 26;
      __syncthreads ();
      if (t == 0)
        {
          lub::block_dense<D> (true, o[0], o[1], o[2], o + 4, Bs);
          lub::block_dense<D> (false, o[0], o[1], o[2], o + 15, Bp);
        }
      __syncthreads ();
      if (t < D * D)
        {
          const double w = o[3];
          acc += w * Bs[t];
          double *q = mat + ((size_t) i * D + r) * n + (size_t) list[p].x * D + c;
          *q += w * Bp[t];
        }
    }
  if (t < D * D) mat[((size_t) i * D + r) * n + (size_t) i * D + c] += acc;
}

// y[i][q] = sum over i's pairs in list order
__global__ void k_lub_gather (const int *__restrict__ off, const double *__restrict__ val,
                              int np, int D, double *__restrict__ y)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * D) return;
  const int i = t / D, q = t - i * D;
  double acc = 0.0;
  for (int p = off[i]; p < off[i + 1]; p++) acc += val[(size_t) p * D + q];
  y[t] = acc;
}

}  // namespace shk
