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
  int minimg;               // the cut-off is at most half of every box length
                            // (unsheared): only the nearest image can be near
};

// the nearest-image search applies: periodic and unsheared, a cut-off
// (lubmax, in r or in s = 2 r / (a_i + a_j) units: at most lubmax a_max in
// r) of at most half of every box length
inline int lub_minimg (const LubGeom &g, double lubmax, double amax)
{
  if (!g.periodic || g.s2x != 0.0 || g.s3x != 0.0 || !(lubmax > 0.0) || !(amax > 0.0)) return 0;
  const double rcut = g.poly ? lubmax * amax : lubmax;
  const double lmin = g.lx < g.ly ? (g.lx < g.lz ? g.lx : g.lz) : (g.ly < g.lz ? g.ly : g.lz);
  return rcut <= 0.5 * lmin ? 1 : 0;
}

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
// writes the (j, cell) records from off[i] in (j chunk, cell, lane) order,
// those with an index below cap
template <int FILL>
__device__ __forceinline__ void
lub_pairs_wave (int i, int lane, const double4 *__restrict__ pos, const LubGeom &g,
                int *__restrict__ cnt, const int *__restrict__ off, int2 *__restrict__ list,
                int cap)
{
  if (i >= g.np) return;
  const double4 pi = pos[i];
  int base = FILL ? off[i] : 0;
  if (g.minimg)
    {
      // every image but the nearest is at least half a box length away in
      // one direction, beyond the cut-off: one test per pair, the nearest
      // image when it is inside the enumerated cells
      for (int j0 = 0; j0 < g.np; j0 += 64)
        {
          const int j = j0 + lane;
          const bool valid = j < g.np;
          const double4 pj = pos[valid ? j : i];
          const int a = -(int) rint ((pj.x - pi.x) / g.lx);
          const int b = -(int) rint ((pj.y - pi.y) / g.ly);
          const int c = -(int) rint ((pj.z - pi.z) / g.lz);
          double d[3];
          const bool hit = valid && abs (a) <= g.ix && abs (b) <= g.iy && abs (c) <= g.iz
            && lub_hit (g, pi, pj, a, b, c, d);
          const unsigned long long m = __ballot (hit);
          if (FILL && hit)
            {
              const unsigned long long below = lane ? (m & ((~0ull) >> (64 - lane))) : 0ull;
              const int at = base + __popcll (below);
              if (at < cap) list[at] = make_int2 (j, lub_img_code (a, b, c));
            }
          base += __popcll (m);
        }
      if (!FILL && lane == 0) cnt[i] = base;
      return;
    }
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
                  const int at = base + __popcll (below);
                  if (at < cap) list[at] = make_int2 (j, lub_img_code (a, b, c));
                }
              base += __popcll (m);
            }
    }
  if (!FILL && lane == 0) cnt[i] = base;
}

template <int FILL>
__global__ __launch_bounds__ (256) void
k_lub_pairs (const double4 *__restrict__ pos, LubGeom g, int *__restrict__ cnt,
             const int *__restrict__ off, int2 *__restrict__ list)
{
  lub_pairs_wave<FILL> (blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63, pos, g, cnt, off,
                        list, 0x7fffffff);
}

// exclusive scan of cnt[0..n) into off[0..n] (one workgroup; n moderate)
static __global__ __launch_bounds__ (1024) void
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
// by Horner's rule from q = k + 1 down; the loop takes eight coefficients
// per step once the remaining count is a multiple of eight, so that the
// loads of a step can be issued together (same order of operations as the
// plain loop)
struct DevPolyF {
  const double *c;   // [fam][k (N + 2) + q]
  int N;
  double l;
  __device__ double operator() (int fam, int k) const
  {
    const double *p = c + ((size_t) fam * (N + 1) + k) * (N + 2);
    double v = 0.0;
    int q = k + 1;
    for (; (q + 1) & 7; q--) v = v * l + p[q];
    if (q < 7) return v;
    // the next step's eight coefficients are loaded before this step's
    // products (the prefetch index stays within the row: q - 8 >= 7 or 7)
    double n7 = p[q], n6 = p[q - 1], n5 = p[q - 2], n4 = p[q - 3];
    double n3 = p[q - 4], n2 = p[q - 5], n1 = p[q - 6], n0 = p[q - 7];
    for (; q >= 7; q -= 8)
      {
        const double c7 = n7, c6 = n6, c5 = n5, c4 = n4, c3 = n3, c2 = n2, c1 = n1, c0 = n0;
        const int qn = q - 8 >= 7 ? q - 8 : 7;
        n7 = p[qn]; n6 = p[qn - 1]; n5 = p[qn - 2]; n4 = p[qn - 3];
        n3 = p[qn - 4]; n2 = p[qn - 5]; n1 = p[qn - 6]; n0 = p[qn - 7];
        v = v * l + c7;
        v = v * l + c6;
        v = v * l + c5;
        v = v * l + c4;
        v = v * l + c3;
        v = v * l + c2;
        v = v * l + c1;
        v = v * l + c0;
      }
    return v;
  }
  // f_k .. f_k+3 together (k + 3 <= N): four Horner recurrences side by
  // side from the top degree of row k + 3, the shorter rows reading their
  // zero coefficients above degree k + 1 (0 l + 0 = 0 exactly, so each
  // value is the one operator() gives)
  __device__ void four (int fam, int k, double *f) const
  {
    const size_t R = (size_t) (N + 2);
    horner4 (c + ((size_t) fam * (N + 1) + k) * R, R, k + 4, l, f);
  }
  // rows p0, p0 + R, p0 + 2R, p0 + 3R from degree qtop down
  __device__ static void horner4 (const double *p0, size_t R, int qtop, double l, double *f)
  {
    const double *p1 = p0 + R, *p2 = p1 + R, *p3 = p2 + R;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    int q = qtop;
    for (; (q + 1) & 3; q--)
      {
        v0 = v0 * l + p0[q]; v1 = v1 * l + p1[q];
        v2 = v2 * l + p2[q]; v3 = v3 * l + p3[q];
      }
    for (; q >= 3; q -= 4)
      {
        const double a0 = p0[q], a1 = p0[q - 1], a2 = p0[q - 2], a3 = p0[q - 3];
        const double b0 = p1[q], b1 = p1[q - 1], b2 = p1[q - 2], b3 = p1[q - 3];
        const double c0 = p2[q], c1 = p2[q - 1], c2 = p2[q - 2], c3 = p2[q - 3];
        const double d0 = p3[q], d1 = p3[q - 1], d2 = p3[q - 2], d3 = p3[q - 3];
        v0 = v0 * l + a0; v1 = v1 * l + b0; v2 = v2 * l + c0; v3 = v3 * l + d0;
        v0 = v0 * l + a1; v1 = v1 * l + b1; v2 = v2 * l + c1; v3 = v3 * l + d1;
        v0 = v0 * l + a2; v1 = v1 * l + b2; v2 = v2 * l + c2; v3 = v3 * l + d2;
        v0 = v0 * l + a3; v1 = v1 * l + b3; v2 = v2 * l + c3; v3 = v3 * l + d3;
      }
    f[0] = v0; f[1] = v1; f[2] = v2; f[3] = v3;
  }
};

// The same values with the four rows of each step staged in LDS by the whole
// workgroup: every thread of the block must make the same sequence of
// four() calls (uniform control flow; a thread without a pair of its own
// evaluates a neighbour's), so the rows are read from memory once per block
// and each thread's recurrences read them back as broadcasts.  sh holds
// 4 (N + 2) doubles.
struct LdsPolyF {
  const double *c;
  int N;
  double l;
  double *sh;
  __device__ double operator() (int fam, int k) const { return DevPolyF{c, N, l} (fam, k); }
  __device__ void four (int fam, int k, double *f) const
  {
    const int R = N + 2;
    const double *g = c + ((size_t) fam * (N + 1) + k) * R;   // rows k .. k + 3, contiguous
    __syncthreads ();
    for (int t = threadIdx.x; t < 4 * R; t += blockDim.x) sh[t] = g[t];
    __syncthreads ();
    DevPolyF::horner4 (sh, (size_t) R, k + 4, l, f);
  }
};

// f values at lambda = 1 tabulated [fam][k], k <= 50
struct DevTabF {
  const double *t;
  __device__ double operator() (int fam, int k) const { return t[fam * 51 + k]; }
  __device__ void four (int fam, int k, double *f) const
  {
    for (int i = 0; i < 4; i++) f[i] = t[fam * 51 + k + i];
  }
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
// self image).  sh != nullptr: the coefficient rows go through that LDS
// buffer (LdsPolyF; block-uniform calls)
template <int V>
__device__ inline void lub_pair (const double4 *__restrict__ pos, const LubGeom &g,
                                 const LubEval &ev, const int *__restrict__ off,
                                 const int2 *__restrict__ list, int p, int &i, int &j, double *e,
                                 double *s11, double *s12, double &w, double *sh = nullptr)
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
      if (sh)
        {
          const LdsPolyF fv{ev.coef, ev.N, l, sh};
          if (ev.flag_lub) lub::lub_form (V, fv, n - 1, l, s, 1.0, two);
          else lub::far_form (V, fv, n, l, s, two);
        }
      else
        {
          const DevPolyF fv{ev.coef, ev.N, l};
          if (ev.flag_lub) lub::lub_form (V, fv, n - 1, l, s, 1.0, two);
          else lub::far_form (V, fv, n, l, s, two);
        }
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

// directed pair p: its D values into val[p] (store false: evaluate only);
// pos, off and x of the pair's own system (off[0..np], pair indices p in
// [off[0], off[np])); sh: as for lub_pair
template <int V>
__device__ __forceinline__ void
lub_eval_item (int p, const double4 *__restrict__ pos, const LubGeom &g, const LubEval &ev,
               const int *__restrict__ off, const int2 *__restrict__ list,
               const double *__restrict__ x, double *__restrict__ val, double *sh = nullptr,
               bool store = true)
{
  constexpr int D = Dim<V>::D;
  int i, j;
  double e[3], s11[11], s12[11], w;
  lub_pair<V> (pos, g, ev, off, list, p, i, j, e, s11, s12, w, sh);
  if (!store) return;
  double y[11];
  for (int k = 0; k < D; k++) y[k] = 0.0;
  lub::apply_block<D> (true, x + (size_t) D * i, y, e[0], e[1], e[2], s11);
  lub::apply_block<D> (false, x + (size_t) D * j, y, e[0], e[1], e[2], s12);
  double *o = val + (size_t) p * D;
  for (int k = 0; k < D; k++) o[k] = w * y[k];
}

// POLY: the kind of radii as a compile-time constant.  Unequal spheres
// stage the coefficient rows in LDS (dynamic, 4 (N + 2) doubles): every
// thread of the block runs the evaluation, those past the last pair on the
// last pair without storing.
template <int V, int POLY>
__global__ __launch_bounds__ (128) void
k_lub_eval (const double4 *__restrict__ pos, LubGeom g, LubEval ev, const int *__restrict__ off,
            const int2 *__restrict__ list, int npairs, const double *__restrict__ x,
            double *__restrict__ val)
{
  extern __shared__ double lub_sh[];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  g.poly = POLY;
  if (POLY)
    lub_eval_item<V> (p < npairs ? p : npairs - 1, pos, g, ev, off, list, x, val, lub_sh,
                      p < npairs);
  else if (p < npairs)
    lub_eval_item<V> (p, pos, g, ev, off, list, x, val);
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

// y[i][q] = sum over i's pairs in list order (pairs below cap: a list
// that overflowed leaves y incomplete, never reads past it)
static __global__ void k_lub_gather (const int *__restrict__ off, const double *__restrict__ val,
                                     int np, int D, int cap, double *__restrict__ y)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * D) return;
  const int i = t / D, q = t - i * D;
  double acc = 0.0;
  const int p1 = min (off[i + 1], cap);
  for (int p = off[i]; p < p1; p++) acc += val[(size_t) p * D + q];
  y[t] = acc;
}

}  // namespace shk
