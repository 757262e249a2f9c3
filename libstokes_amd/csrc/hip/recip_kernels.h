// recip_kernels.h -- reciprocal-space Ewald sum as structure factors
// (ewald.c:923-1075 recast; every k of the half space carries the weight of
// k and -k).
//
//   k_recip_sf     C_k = sum_j cos(k.x_j) z_j, S_k = sum_j sin(k.x_j) z_j
//                  (and the a_j^2-weighted sums for unequal spheres); one
//                  thread per k, the j range split over blockIdx.y, the j
//                  tile staged in LDS; partial sums per split.
//   k_recip_blocks per k: sum the splits in a fixed order and apply the
//                  even / odd k-blocks: P = Ac C - As S, Q = Ac S + As C.
//   k_recip_back   y_i += sum_k a_i^p [cos(k.x_i) P + sin(k.x_i) Q]; one
//                  thread per i, the k range split over blockIdx.y.
//   k_combine      y_i = self_i z_i + real partials + reciprocal partials,
//                  summed in a fixed order (bitwise deterministic).
#pragma once
#include <hip/hip_runtime.h>
#include "pair_math.h"
#include "ewald_kernels.h"
#include "real_kernels.h"

namespace shk {

template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (256) void
k_recip_sf (const double4 *__restrict__ pos, const double *__restrict__ x,
            const double4 *__restrict__ kv, int np, int nk, int jper, T *__restrict__ part)
{
  constexpr int D = Dim<V>::D;
  constexpr int NS = (NPOW > 1) ? 4 : 2;
  __shared__ double4 sp[256];
  __shared__ T sz[256][D];
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int j0 = blockIdx.y * jper;
  const int j1 = min (np, j0 + jper);
  const double4 kk = kv[min (k, nk - 1)];
  T acc[NS * D];
#pragma unroll
  for (int q = 0; q < NS * D; q++) acc[q] = T (0);
  for (int jb = j0; jb < j1; jb += 256)
    {
      const int j = jb + threadIdx.x;
      if (j < j1)
        {
          sp[threadIdx.x] = pos[j];
          T zz[D];
          load_z<V, T> (x, j, zz);
#pragma unroll
          for (int q = 0; q < D; q++) sz[threadIdx.x][q] = zz[q];
        }
      __syncthreads ();
      const int n = min (256, j1 - jb);
      for (int t = 0; t < n; t++)
        {
          const double4 pj = sp[t];
          T c, s;
          phase_cs (kk.x * pj.x + kk.y * pj.y + kk.z * pj.z, c, s);
#pragma unroll
          for (int q = 0; q < D; q++)
            {
              const T zq = sz[t][q];
              acc[q] += c * zq;
              acc[D + q] += s * zq;
            }
          if (NPOW > 1)
            {
              const T a2 = T (pj.w * pj.w);
              const T c2 = a2 * c, s2 = a2 * s;
#pragma unroll
              for (int q = 0; q < D; q++)
                {
                  const T zq = sz[t][q];
                  acc[2 * D + q] += c2 * zq;
                  acc[3 * D + q] += s2 * zq;
                }
            }
        }
      __syncthreads ();
    }
  if (k < nk)
    {
      T *out = part + ((size_t) blockIdx.y * nk + k) * (NS * D);
#pragma unroll
      for (int q = 0; q < NS * D; q++) out[q] = acc[q];
    }
}

// one thread per k: the splits summed in a fixed order, then P and Q of
// every power of a_i (unequal spheres: five, each with the scalars of its
// power only)
template <int NPOW> __host__ __device__ inline long recip_blocks_threads (int nk)
{
  return (long) nk;
}

template <int V, typename T, int PW>
__device__ __forceinline__ void recip_blocks_pw (const double4 &kk, const double *w,
                                                 const double *CS, T *out)
{
  constexpr int D = Dim<V>::D;
  double P[D], Q[D];
  kblock_pq_poly<V, PW> (kk, w, CS, P, Q);
#pragma unroll
  for (int q = 0; q < D; q++) { out[q] = T (P[q]); out[D + q] = T (Q[q]); }
}

template <int V, typename T, int NPOW>
__device__ __forceinline__ void
recip_blocks_item (int t, const double4 *__restrict__ kv, const double *__restrict__ ks,
                   int nk, int nsplit, const T *__restrict__ part, T *__restrict__ PQ)
{
  constexpr int D = Dim<V>::D;
  constexpr int NS = (NPOW > 1) ? 4 : 2;
  const int k = t;
  if (k >= nk) return;
  double CS[NS * D];
#pragma unroll
  for (int q = 0; q < NS * D; q++) CS[q] = 0.0;
  // part[(split * NS*D + q) * nk + k] (k fastest: coalesced)
  for (int sp = 0; sp < nsplit; sp++)
    {
      const T *src = part + (size_t) sp * (NS * D) * nk + k;
#pragma unroll
      for (int q = 0; q < NS * D; q++) CS[q] += double (src[(size_t) q * nk]);
    }
  const double4 kk = kv[k];
  const double *w = ks + (size_t) k * 8;
  if (NPOW == 1)
    {
      double P[D], Q[D];
      kblock_pq_mono<V> (kk, w, CS, P, Q);
      T *out = PQ + (size_t) k * 2 * D;
#pragma unroll
      for (int q = 0; q < D; q++) { out[q] = T (P[q]); out[D + q] = T (Q[q]); }
      return;
    }
  T *out = PQ + (size_t) k * NPOW * 2 * D;
  recip_blocks_pw<V, T, 1> (kk, w, CS, out);
  recip_blocks_pw<V, T, 2> (kk, w, CS, out + 2 * D);
  recip_blocks_pw<V, T, 3> (kk, w, CS, out + 4 * D);
  recip_blocks_pw<V, T, 4> (kk, w, CS, out + 6 * D);
  recip_blocks_pw<V, T, 5> (kk, w, CS, out + 8 * D);
}

template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (128) void
k_recip_blocks (const double4 *__restrict__ kv, const double *__restrict__ ks, int nk,
                int nsplit, double gsign, const T *__restrict__ part, T *__restrict__ PQ)
{
  recip_blocks_item<V, T, NPOW> (blockIdx.x * blockDim.x + threadIdx.x, kv, ks, nk, nsplit,
                                 part, PQ);
}

template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (256) void
k_recip_back (const double4 *__restrict__ pos, const double4 *__restrict__ kv,
              const T *__restrict__ PQ, int np, int nk, int kper, T *__restrict__ part)
{
  constexpr int D = Dim<V>::D;
  constexpr int KT = 32;
  __shared__ double4 skv[KT];
  __shared__ T spq[KT][NPOW * 2 * D];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int k0 = blockIdx.y * kper;
  const int k1 = min (nk, k0 + kper);
  const double4 pi = pos[min (i, np - 1)];
  T apow[NPOW];
  apow[0] = NPOW > 1 ? T (pi.w) : T (1);
#pragma unroll
  for (int t = 1; t < NPOW; t++) apow[t] = apow[t - 1] * T (pi.w);
  T acc[D];
#pragma unroll
  for (int q = 0; q < D; q++) acc[q] = T (0);
  for (int kb = k0; kb < k1; kb += KT)
    {
      const int n = min (KT, k1 - kb);
      for (int e = threadIdx.x; e < n * (NPOW * 2 * D); e += 256)
        {
          const int kt = e / (NPOW * 2 * D), q = e - kt * (NPOW * 2 * D);
          spq[kt][q] = PQ[(size_t) (kb + kt) * (NPOW * 2 * D) + q];
        }
      if (threadIdx.x < n) skv[threadIdx.x] = kv[kb + threadIdx.x];
      __syncthreads ();
      for (int t = 0; t < n; t++)
        {
          const double4 kk = skv[t];
          T c, s;
          phase_cs (kk.x * pi.x + kk.y * pi.y + kk.z * pi.z, c, s);
#pragma unroll
          for (int pw = 0; pw < NPOW; pw++)
            {
              const T cc = apow[pw] * c, ss = apow[pw] * s;
              const T *P = spq[t] + pw * 2 * D;
#pragma unroll
              for (int q = 0; q < D; q++) acc[q] += cc * P[q] + ss * P[D + q];
            }
        }
      __syncthreads ();
    }
  if (i < np)
    {
      T *out = part + ((size_t) blockIdx.y * np + i) * D;
#pragma unroll
      for (int q = 0; q < D; q++) out[q] = acc[q];
    }
}

// COMBINE_LANES lanes per output component (i, q): lane g sums the partial
// chunks g, g + 16, ...; the lanes then combine by a fixed butterfly, so
// the summation order (and the result) does not depend on scheduling
constexpr int COMBINE_LANES = 16;

// thread gt of np * D * COMBINE_LANES (all lanes of a wave take part)
template <int V, typename T>
__device__ __forceinline__ void
combine_item (size_t gt, const double4 *__restrict__ pos, const double *__restrict__ x,
              const double *__restrict__ rpart, int nreal, const T *__restrict__ kpart,
              int nrecip, const DevParams &p, int periodic, double *__restrict__ y)
{
  constexpr int D = Dim<V>::D;
  const size_t t = gt / COMBINE_LANES;
  const int g = (int) (gt % COMBINE_LANES);
  const size_t n = (size_t) p.np * D;
  const bool live = t < n;
  const size_t tt = live ? t : 0;
  double acc = 0.0;
  if (live)
    {
      for (int ch = g; ch < nreal; ch += COMBINE_LANES) acc += rpart[(size_t) ch * n + tt];
      for (int ch = g; ch < nrecip; ch += COMBINE_LANES)
        acc += double (kpart[(size_t) ch * n + tt]);
    }
#pragma unroll
  for (int off = COMBINE_LANES / 2; off > 0; off >>= 1) acc += __shfl_xor (acc, off, 64);
  if (live && g == 0)
    {
      const int i = (int) (tt / D), q = (int) (tt - (size_t) i * D);
      y[tt] = self_diag (p, pos[i].w, periodic != 0, V, q) * x[tt] + acc;
    }
}

template <int V, typename T>
__global__ __launch_bounds__ (256) void
k_combine (const double4 *__restrict__ pos, const double *__restrict__ x,
           const double *__restrict__ rpart, int nreal, const T *__restrict__ kpart,
           int nrecip, DevParams p, int periodic, double *__restrict__ y)
{
  combine_item<V, T> ((size_t) blockIdx.x * blockDim.x + threadIdx.x, pos, x, rpart, nreal,
                      kpart, nrecip, p, periodic, y);
}

// ---- separable phases ---------------------------------------------------
//
// For k = m1 b1 + m2 b2 + m3 b3 on the (possibly sheared) reciprocal
// lattice, k.x = m1 th1 + m2 th2 + m3 th3 with th2 = k0y y, th3 = k0z z and
// th1 = k0x x - s1 k0y y - s2 k0z z (s1 or s2 the shear shift / lx of mode 1
// or 2).  Every particle gets a table of e^{i m th} for m1 in [0, K1],
// m2 in [-K2, K2], m3 in [-K3, K3] (FP64 sincos, stored in T); the
// structure factors then need two complex products per (k, j) instead of a
// sincos, and the back-projection walks each lattice row (m1, m2 fixed, m3
// consecutive) by multiplying with e^{i th3} in FP64.

template <typename T> struct Cplx { T re, im; };

struct PhaseBasis {
  double k0x, k0y, k0z, s1, s2;
  int K1, K2, K3, W;           // W = (K1 + 1) + (2 K2 + 1) + (2 K3 + 1)
};

__device__ __forceinline__ void phase_angles (const PhaseBasis &b, const double4 p, double &t1,
                                              double &t2, double &t3)
{
  t2 = b.k0y * p.y;
  t3 = b.k0z * p.z;
  t1 = b.k0x * p.x - b.s1 * t2 - b.s2 * t3;
}

// tab[j ld + e], e < W (ld >= W: row stride); thread t of np * W
template <typename T>
__device__ __forceinline__ void
phase_tab_item (size_t t, const double4 *__restrict__ pos, int np, const PhaseBasis &b, int ld,
                Cplx<T> *__restrict__ tab)
{
  if (t >= (size_t) np * b.W) return;
  const int j = (int) (t / b.W), e = (int) (t - (size_t) j * b.W);
  double t1, t2, t3;
  phase_angles (b, pos[j], t1, t2, t3);
  double th;
  int m;
  if (e <= b.K1) { m = e; th = t1; }
  else if (e <= b.K1 + 2 * b.K2 + 1) { m = e - (b.K1 + 1) - b.K2; th = t2; }
  else { m = e - (b.K1 + 2 * b.K2 + 2) - b.K3; th = t3; }
  double s, c;
  sincos (double (m) * th, &s, &c);
  tab[(size_t) j * ld + e] = Cplx<T>{T (c), T (s)};
}

template <typename T>
__global__ void k_phase_tab (const double4 *__restrict__ pos, int np, PhaseBasis b, int ld,
                             Cplx<T> *__restrict__ tab)
{
  phase_tab_item<T> ((size_t) blockIdx.x * blockDim.x + threadIdx.x, pos, np, b, ld, tab);
}

template <typename T>
__device__ __forceinline__ Cplx<T> cmul (const Cplx<T> a, const Cplx<T> b)
{
  return Cplx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// structure factors with tabulated phases.  A block is 64 k (one per lane)
// x 4 waves: the block's j range goes in tiles of JT particles (tables, z
// and a_j^2 in dynamic LDS), wave w takes every fourth particle of a tile,
// and the four partial sums are added in LDS in wave order at the end --
// four times the waves of a thread-per-k layout for the same partial
// buffer, and a fixed summation order.
// z rows padded to a multiple of 4 for vector LDS reads
template <int V> struct ZPad { static constexpr int P = (Dim<V>::D + 3) & ~3; };
template <typename T> struct Quad;
template <> struct Quad<float> { using type = float4; };
template <> struct Quad<double> { using type = double4; };

// block (bx, by): k block bx of 64, particle split by
template <int V, typename T, int NPOW>
__device__ __forceinline__ void
recip_sf_tab_block (int bx, int by, const double4 *__restrict__ pos, const double *__restrict__ x,
                    const int4 *__restrict__ km, const Cplx<T> *__restrict__ tab,
                    const PhaseBasis &b, int np, int nk, int jper, int JT, T *__restrict__ part)
{
  constexpr int D = Dim<V>::D;
  constexpr int DP = ZPad<V>::P;
  constexpr int NS = (NPOW > 1) ? 4 : 2;
  extern __shared__ __align__ (32) unsigned char smem_sf[];
  // [JT][DP] z rows, [JT] a_j^2, then [JT][W] phase tables (JT a multiple of 4)
  T *sz = (T *) smem_sf;
  T *sa2 = sz + (size_t) JT * DP;
  Cplx<T> *stab = (Cplx<T> *) (sa2 + JT);
  T *red = (T *) smem_sf;                        // reused after the j loop
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int k = bx * 64 + lane;
  const int j0 = by * jper;
  const int j1 = min (np, j0 + jper);
  const int4 m = km[min (k, nk - 1)];
  const int o1 = m.x, o2 = b.K1 + 1 + b.K2 + m.y, o3 = b.K1 + 2 * b.K2 + 2 + b.K3 + m.z;
  T acc[NS * D];
#pragma unroll
  for (int q = 0; q < NS * D; q++) acc[q] = T (0);
  for (int jb = j0; jb < j1; jb += JT)
    {
      const int n = min (JT, j1 - jb);
      const Cplx<T> *src = tab + (size_t) jb * b.W;
      for (int e = threadIdx.x; e < n * b.W; e += 256) stab[e] = src[e];
      if (threadIdx.x < n)
        {
          T zz[D];
          load_z<V, T> (x, jb + threadIdx.x, zz);
#pragma unroll
          for (int q = 0; q < DP; q++) sz[threadIdx.x * DP + q] = q < D ? zz[q] : T (0);
          if (NPOW > 1)
            {
              const double a = pos[jb + threadIdx.x].w;
              sa2[threadIdx.x] = T (a * a);
            }
        }
      __syncthreads ();
      for (int t = wave; t < n; t += 4)
        {
          const Cplx<T> *row = stab + (size_t) t * b.W;
          const Cplx<T> e = cmul (cmul (row[o1], row[o2]), row[o3]);
          const T c = e.re, s = e.im;
          T zq[DP];
          const typename Quad<T>::type *zv = (const typename Quad<T>::type *) (sz + t * DP);
#pragma unroll
          for (int v = 0; v < DP / 4; v++)
            {
              const typename Quad<T>::type z4 = zv[v];
              zq[4 * v] = z4.x; zq[4 * v + 1] = z4.y; zq[4 * v + 2] = z4.z; zq[4 * v + 3] = z4.w;
            }
#pragma unroll
          for (int q = 0; q < D; q++)
            {
              acc[q] += c * zq[q];
              acc[D + q] += s * zq[q];
            }
          if (NPOW > 1)
            {
              const T c2 = sa2[t] * c, s2 = sa2[t] * s;
#pragma unroll
              for (int q = 0; q < D; q++)
                {
                  acc[2 * D + q] += c2 * zq[q];
                  acc[3 * D + q] += s2 * zq[q];
                }
            }
        }
      __syncthreads ();
    }
  // waves 1..3 hand their sums to wave 0, D components at a time (fixed
  // order); the region holds 3 x 64 x D values
#pragma unroll
  for (int c = 0; c < NS; c++)
    {
      if (wave > 0)
#pragma unroll
        for (int q = 0; q < D; q++) red[((wave - 1) * D + q) * 64 + lane] = acc[c * D + q];
      __syncthreads ();
      if (wave == 0)
#pragma unroll
        for (int q = 0; q < D; q++)
          acc[c * D + q] += red[q * 64 + lane] + red[(D + q) * 64 + lane]
            + red[(2 * D + q) * 64 + lane];
      __syncthreads ();
    }
  if (wave == 0 && k < nk)
    {
      // part[(split * NS*D + q) * nk + k]: consecutive lanes, consecutive k
      T *out = part + (size_t) by * (NS * D) * nk + k;
#pragma unroll
      for (int q = 0; q < NS * D; q++) out[(size_t) q * nk] = acc[q];
    }
}

template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (256) void
k_recip_sf_tab (const double4 *__restrict__ pos, const double *__restrict__ x,
                const int4 *__restrict__ km, const Cplx<T> *__restrict__ tab, PhaseBasis b,
                int np, int nk, int jper, int JT, T *__restrict__ part)
{
  recip_sf_tab_block<V, T, NPOW> (blockIdx.x, blockIdx.y, pos, x, km, tab, b, np, nk, jper, JT,
                                  part);
}

// back-projection walking lattice rows.  A block is 64 i (one per lane) x
// 4 waves; the block's k range [k0, k1) is cut into four consecutive
// quarters, one per wave, each with its own LDS tiles.  km[k].w != 0 marks
// the first k of a row (direct phase); the other k rotate the FP64 phase by
// e^{i th3}.  The four partial sums are added in LDS in wave order.
// block (bx, by): particle block bx of 64, k split by
template <int V, typename T, int NPOW>
__device__ __forceinline__ void
recip_back_rows_block (int bx, int by, const double4 *__restrict__ pos,
                       const double4 *__restrict__ kv, const int4 *__restrict__ km,
                       const PhaseBasis &b, const T *__restrict__ PQ, int np, int nk, int kper,
                       T *__restrict__ part)
{
  constexpr int D = Dim<V>::D;
  constexpr int KT = NPOW > 1 ? 8 : 16;          // LDS stays under 64 KB
  constexpr int PW = NPOW * 2 * D;
  __shared__ double4 skv[4][KT];
  __shared__ int srow[4][KT];
  __shared__ T spq[4][KT][PW];
  __shared__ T red[3][D][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = bx * 64 + lane;
  const int kb0 = by * kper;
  const int kb1 = min (nk, kb0 + kper);
  const int kq = (kb1 - kb0 + 3) / 4;
  const int k0 = min (kb1, kb0 + wave * kq);
  const int k1 = min (kb1, k0 + kq);
  const double4 pi = pos[min (i, np - 1)];
  double t1, t2, t3;
  phase_angles (b, pi, t1, t2, t3);
  double rs, rc;
  sincos (t3, &rs, &rc);
  double pc = 1.0, ps = 0.0;
  T apow[NPOW];
  apow[0] = NPOW > 1 ? T (pi.w) : T (1);
#pragma unroll
  for (int t = 1; t < NPOW; t++) apow[t] = apow[t - 1] * T (pi.w);
  T acc[D];
#pragma unroll
  for (int q = 0; q < D; q++) acc[q] = T (0);
  for (int kb = k0; kb < k1; kb += KT)
    {
      const int n = min (KT, k1 - kb);
      for (int e = lane; e < n * PW; e += 64)
        {
          const int kt = e / PW, q = e - kt * PW;
          spq[wave][kt][q] = PQ[(size_t) (kb + kt) * PW + q];
        }
      if (lane < n)
        {
          skv[wave][lane] = kv[kb + lane];
          srow[wave][lane] = km[kb + lane].w;
        }
      wave_sync ();
      for (int t = 0; t < n; t++)
        {
          if (srow[wave][t] || kb + t == k0)
            {
              const double4 kk = skv[wave][t];
              T c0, s0;
              phase_cs (kk.x * pi.x + kk.y * pi.y + kk.z * pi.z, c0, s0);
              pc = double (c0);
              ps = double (s0);
            }
          else
            {
              const double nc = pc * rc - ps * rs;
              ps = ps * rc + pc * rs;
              pc = nc;
            }
          const T c = T (pc), s = T (ps);
#pragma unroll
          for (int pw = 0; pw < NPOW; pw++)
            {
              const T cc = apow[pw] * c, ss = apow[pw] * s;
              const T *P = spq[wave][t] + pw * 2 * D;
#pragma unroll
              for (int q = 0; q < D; q++) acc[q] += cc * P[q] + ss * P[D + q];
            }
        }
      wave_sync ();
    }
  if (wave > 0)
#pragma unroll
    for (int q = 0; q < D; q++) red[wave - 1][q][lane] = acc[q];
  __syncthreads ();
  if (wave == 0 && i < np)
    {
      T *out = part + ((size_t) by * np + i) * D;
#pragma unroll
      for (int q = 0; q < D; q++) out[q] = acc[q] + red[0][q][lane] + red[1][q][lane] + red[2][q][lane];
    }
}

template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (256) void
k_recip_back_rows (const double4 *__restrict__ pos, const double4 *__restrict__ kv,
                   const int4 *__restrict__ km, PhaseBasis b, const T *__restrict__ PQ, int np,
                   int nk, int kper, T *__restrict__ part)
{
  recip_back_rows_block<V, T, NPOW> (blockIdx.x, blockIdx.y, pos, kv, km, b, PQ, np, nk, kper,
                                     part);
}

}  // namespace shk
