// cache_kernels.h -- the real-space part of the operator held as a dense
// matrix for a fixed configuration, for callers that apply M(x) many times
// at the same positions (the Chebyshev recursion of a Brownian step: 32
// products at x).
//
//   k_real_dense  Mr[(i D + r) ND + j D + q] = the image-summed pair block
//                 (i, j) of the real-space sum (ewald.c:649-789; self-images
//                 of i included, the self term not), one wave per (i, 64 j),
//                 lane = j: a per-lane candidate mask, then the lane's own
//                 images, moments in registers, entries through BlockEach
//   k_z_prep      z = the (S-transformed) vector in the matrix precision
//   k_real_gemv   y_real = Mr z, one wave per row, summed in FP64 in a fixed
//                 lane order
#pragma once
#include <hip/hip_runtime.h>
#include "pair_math.h"
#include "ewald_kernels.h"
#include "real_kernels.h"

namespace shk {

template <int V, typename T>
__global__ __launch_bounds__ (256) void
k_real_dense (const double4 *__restrict__ pos, const typename Vec4<T>::type *__restrict__ cand,
              const int4 *__restrict__ candm, const unsigned *__restrict__ tbits, RealParams p,
              T *__restrict__ Mr)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nj = (p.np + 63) / 64;
  const long gw = (long) blockIdx.x * 4 + wave;
  const int i = (int) (gw / nj), jc = (int) (gw - (long) i * nj);
  if (i >= p.np) return;
  const int j = jc * 64 + lane;
  const bool valid = j < p.np;
  const double4 pi = pos[i];
  const double4 pj = pos[valid ? j : i];
  double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
  const double n3 = rint (dz / p.lz);
  dx -= n3 * p.s3x; dz -= n3 * p.lz;
  const double n2 = rint (dy / p.ly);
  dx -= n2 * p.s2x; dy -= n2 * p.ly;
  const double n1 = rint (dx / p.lx);
  dx -= n1 * p.lx;
  const int w1 = int (n1), w2 = int (n2), w3 = int (n3);
  const T mdx = T (dx), mdy = T (dy), mdz = T (dz);
  const T ai = T (pi.w), aj = T (pj.w), xi = T (p.xi), rmax2 = T (p.rmax2);
  const T rmn = (ai + aj) * T (p.rmin);
  T M[NM];
#pragma unroll
  for (int t = 0; t < NM; t++) M[t] = T (0);
  for (int c0 = 0; c0 < p.ncand; c0 += 64)
    {
      const int cn = min (64, p.ncand - c0);
      unsigned long long mask = 0ull;
#pragma unroll 8
      for (int c = 0; c < cn; c++)
        {
          const typename Vec4<T>::type v = cand[c0 + c];
          const T ex = mdx + v.x, ey = mdy + v.y, ez = mdz + v.z;
          const T rr = ex * ex + ey * ey + ez * ez;
          const bool hit = valid && rr != T (0) && rr <= rmax2;
          mask |= (unsigned long long) hit << c;
        }
      while (mask)
        {
          const int c = __builtin_ctzll (mask);
          mask &= mask - 1ull;
          const int4 mm = candm[c0 + c];
          if (!in_table (p, tbits, mm.x - w1, mm.y - w2, mm.z - w3)) continue;
          const typename Vec4<T>::type v = cand[c0 + c];
          const T ex = mdx + v.x, ey = mdy + v.y, ez = mdz + v.z;
          const T rr = ex * ex + ey * ey + ez * ez;
          T iu = m_rsqrt (rr);
          T r = rr * iu;
          if (r < rmn)
            {
              r = rmn;
              iu = T (1) / rmn;
            }
          T sc[11];
          if (p.poly) scalars_real_poly_u<V, T> (xi, r, iu, ai, aj, sc);
          else scalars_real_u<V, T> (xi, r, iu, sc);
          moments_add<V, T> (M, ex * iu, ey * iu, ez * iu, sc);
        }
    }
  if (!valid) return;
  const size_t ND = (size_t) p.np * D;
  T *blk = Mr + (size_t) i * D * ND + (size_t) j * D;
  auto put = [&] (int r, int q, T v) { blk[(size_t) r * ND + q] = v; };
  BlockEach<V, 0, D, D * D>::run (M, false, put);
}

template <int V, typename T>
__global__ void k_z_prep (const double *__restrict__ x, int np, T *__restrict__ z)
{
  constexpr int D = Dim<V>::D;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  T zz[D];
  load_z<V, T> (x, j, zz);
#pragma unroll
  for (int q = 0; q < D; q++) z[(size_t) j * D + q] = zz[q];
}

template <typename T>
__global__ __launch_bounds__ (256) void
k_real_gemv (const T *__restrict__ Mr, const T *__restrict__ z, int ND, double *__restrict__ out)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  if (row >= ND) return;
  const T *m = Mr + (size_t) row * ND;
  double acc = 0.0;
  for (int c = lane; c < ND; c += 64) acc += double (m[c]) * double (z[c]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor (acc, off, 64);
  if (lane == 0) out[row] = acc;
}

}  // namespace shk
