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

template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (128) void
k_recip_blocks (const double4 *__restrict__ kv, const double *__restrict__ ks, int nk,
                int nsplit, double gsign, const T *__restrict__ part, T *__restrict__ PQ)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  constexpr int NS = (NPOW > 1) ? 4 : 2;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  double CS[NS * D];
#pragma unroll
  for (int q = 0; q < NS * D; q++) CS[q] = 0.0;
  for (int sp = 0; sp < nsplit; sp++)
    {
      const T *src = part + ((size_t) sp * nk + k) * (NS * D);
#pragma unroll
      for (int q = 0; q < NS * D; q++) CS[q] += double (src[q]);
    }
  const double4 kk = kv[k];
  const double *w = ks + (size_t) k * 8;
  double Mc[NM], Ms[NM];
  if (NPOW == 1)
    {
      double P[D], Q[D];
#pragma unroll
      for (int q = 0; q < D; q++) P[q] = Q[q] = 0.0;
      kblock_moments<V, double> (kk, w, Mc, Ms);
      kblock_apply<V> (Mc, Ms, CS, CS + D, P, Q);
      T *out = PQ + (size_t) k * 2 * D;
#pragma unroll
      for (int q = 0; q < D; q++) { out[q] = T (P[q]); out[D + q] = T (Q[q]); }
      return;
    }
  const double ik = 1.0 / kk.w;
  const double ex = kk.x * ik, ey = kk.y * ik, ez = kk.z * ik;
  for (int pw = 1; pw <= NPOW; pw++)
    {
      double sc1[11], ss1[11], sc2[11], ss2[11];
      poly_kscalars (pw, kk.w, w[6], w[7], w[3], sc1, ss1, sc2, ss2);
      double P[D], Q[D];
#pragma unroll
      for (int q = 0; q < D; q++) P[q] = Q[q] = 0.0;
#pragma unroll
      for (int t = 0; t < NM; t++) Mc[t] = Ms[t] = 0.0;
      moments_add<V, double> (Mc, ex, ey, ez, sc1);
      moments_add<V, double> (Ms, ex, ey, ez, ss1);
      kblock_apply<V> (Mc, Ms, CS, CS + D, P, Q);
#pragma unroll
      for (int t = 0; t < NM; t++) Mc[t] = Ms[t] = 0.0;
      moments_add<V, double> (Mc, ex, ey, ez, sc2);
      moments_add<V, double> (Ms, ex, ey, ez, ss2);
      kblock_apply<V> (Mc, Ms, CS + 2 * D, CS + 3 * D, P, Q);
      T *out = PQ + (((size_t) k * NPOW + (pw - 1)) * 2) * D;
#pragma unroll
      for (int q = 0; q < D; q++) { out[q] = T (P[q]); out[D + q] = T (Q[q]); }
    }
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

// one thread per output component (i, q)
template <int V, typename T>
__global__ __launch_bounds__ (256) void
k_combine (const double4 *__restrict__ pos, const double *__restrict__ x,
           const double *__restrict__ rpart, int nreal, const T *__restrict__ kpart,
           int nrecip, DevParams p, int periodic, double *__restrict__ y)
{
  constexpr int D = Dim<V>::D;
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t) p.np * D;
  if (t >= n) return;
  const int i = (int) (t / D), q = (int) (t - (size_t) i * D);
  double acc = self_diag (p, pos[i].w, periodic != 0, V, q) * x[t];
  for (int ch = 0; ch < nreal; ch++) acc += rpart[(size_t) ch * n + t];
  for (int ch = 0; ch < nrecip; ch++) acc += double (kpart[(size_t) ch * n + t]);
  y[t] = acc;
}

}  // namespace shk
