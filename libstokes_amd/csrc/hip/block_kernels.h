// block_kernels.h -- the mobility operator on a block of B right-hand sides
// (the Chebyshev noise block of the Brownian step, batched matvecs).
//
// Internal layout: XT[j][q][b] = T(x_b)_{jq} (S-transformed, column index
// fastest) so that lanes indexed by the column read consecutive addresses.
//
//   k_block_prep      column-major x (n x B) -> XT
//   k_real_block      one wave per (i, 256 j): lanes over pairs accumulate
//                     the image-summed moments of their pair (wrapped
//                     candidates + reference-table bitmap, as in
//                     real_kernels.h), write the 11x11 block to LDS, then
//                     lanes over columns apply the 16 blocks of a group to
//                     their column: Y_i[:, b] += sum_j B_ij XT_j[:, b]
//   k_recip_sf_block  C, S structure factors per (k, b); 32 k x 8 columns
//                     per workgroup, phases of a (32 k x 32 j) tile shared
//                     through LDS
//   k_recip_blk_block per (k, b): P, Q from the split partial sums
//   k_recip_back_block per (i, b): sum_k a_i^p [cos P + sin Q]
//   k_combine_block   self + real + reciprocal partials -> column-major y
#pragma once
#include <hip/hip_runtime.h>
#include "pair_math.h"
#include "ewald_kernels.h"
#include "real_kernels.h"

namespace shk {

template <int V, typename T>
__global__ void k_block_prep (const double *__restrict__ x, int np, int B,
                              T *__restrict__ XT)
{
  constexpr int D = Dim<V>::D;
  const size_t n = (size_t) np * D;
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) np * B) return;
  const int b = (int) (t % B);
  const int j = (int) (t / B);
  T z[D];
#pragma unroll
  for (int q = 0; q < D; q++) z[q] = T (x[(size_t) b * n + (size_t) j * D + q]);
  if (V == 2) s_transform (z);
#pragma unroll
  for (int q = 0; q < D; q++) XT[((size_t) j * D + q) * B + b] = z[q];
}

template <int V, typename T> struct BlockLds {
  static constexpr int D = Dim<V>::D;
  T blk[16][D * D];
};

// JW = source particles per wave (multiple of 64); XT rows of zs values
// (particle j at XT + j zs, component q of column b at q B + b)
template <int V, typename T, int JW>
__global__ __launch_bounds__ (256) void
k_real_block (const double4 *__restrict__ pos, const T *__restrict__ XT, int B, int zs,
              const typename Vec4<T>::type *__restrict__ cand,
              const int4 *__restrict__ candm, const unsigned *__restrict__ tbits,
              RealParams p, T *__restrict__ partial)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  __shared__ BlockLds<V, T> lds_all[4];
  __shared__ unsigned bits_lds[1024];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  BlockLds<V, T> &L = lds_all[wave];
  const int nj = (p.np + JW - 1) / JW;
  const int gw = blockIdx.x * 4 + wave;
  const int i = gw / nj, jc = gw - i * nj;
  const bool bits_in_lds = p.tbits_words <= 1024;
  if (bits_in_lds)
    for (int w = threadIdx.x; w < p.tbits_words; w += blockDim.x) bits_lds[w] = tbits[w];
  __syncthreads ();
  const unsigned *bits = bits_in_lds ? bits_lds : tbits;
  if (i >= p.np) return;
  const double4 pi = pos[i];
  const T ai = T (pi.w), xi = T (p.xi), rmax2 = T (p.rmax2);

  // column accumulators: lanes own columns b = lane, lane + 64, ...
  constexpr int MAXC = 2;       // up to 128 columns per launch
  T acc[MAXC][D];
#pragma unroll
  for (int c = 0; c < MAXC; c++)
#pragma unroll
    for (int q = 0; q < D; q++) acc[c][q] = T (0);

  for (int j0 = jc * JW; j0 < min (p.np, (jc + 1) * JW); j0 += 64)
    {
      // phase A: lane = pair (i, j0 + lane) -> moments
      const int j = j0 + lane;
      const bool valid = j < p.np;
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
      const T aj = T (pj.w);
      const T rmn = (ai + aj) * T (p.rmin);
      T M[NM];
#pragma unroll
      for (int t = 0; t < NM; t++) M[t] = T (0);
      if (valid)
        for (int c = 0; c < p.ncand; c++)
          {
            const typename Vec4<T>::type v = cand[c];
            const T ex = mdx + v.x, ey = mdy + v.y, ez = mdz + v.z;
            const T rr = ex * ex + ey * ey + ez * ez;
            if (rr == T (0) || rr > rmax2) continue;
            const int4 mm = candm[c];
            if (!in_table (p, bits, mm.x - w1, mm.y - w2, mm.z - w3)) continue;
            T ir = m_rsqrt (rr);
            T r = rr * ir;
            if (r < rmn)
              {
                r = rmn;
                ir = T (1) / rmn;
              }
            T s[11];
            if (p.poly) scalars_real_poly_u<V, T> (xi, r, ir, ai, aj, s);
            else scalars_real_u<V, T> (xi, r, ir, s);
            moments_add<V, T> (M, ex * ir, ey * ir, ez * ir, s);
          }
      // phases B / C in groups of 16 pairs
#pragma unroll 1
      for (int g = 0; g < 4; g++)
        {
          if ((lane >> 4) == g)
            {
              T *dst = L.blk[lane & 15];
              auto put = [&] (int r, int cq, T v) { dst[r * D + cq] = v; };
              BlockEach<V, 0, D, D * D>::run (M, false, put);
            }
          wave_sync ();
          const int nvalid = min (16, p.np - (j0 + g * 16));
#pragma unroll
          for (int c = 0; c < MAXC; c++)
            {
              const int b = lane + 64 * c;
              if (b >= B) break;
              for (int t = 0; t < nvalid; t++)
                {
                  const int jj = j0 + g * 16 + t;
                  T z[D];
#pragma unroll
                  for (int q = 0; q < D; q++) z[q] = XT[(size_t) jj * zs + (size_t) q * B + b];
#pragma unroll
                  for (int r = 0; r < D; r++)
                    {
                      T sacc = T (0);
#pragma unroll
                      for (int q = 0; q < D; q++) sacc += L.blk[t][r * D + q] * z[q];
                      acc[c][r] += sacc;
                    }
                }
            }
          wave_sync ();
        }
    }
#pragma unroll
  for (int c = 0; c < MAXC; c++)
    {
      const int b = lane + 64 * c;
      if (b >= B) break;
#pragma unroll
      for (int q = 0; q < D; q++)
        partial[(((size_t) jc * p.np + i) * D + q) * B + b] = acc[c][q];
    }
}

// structure factors, 32 k x 8 columns per workgroup; part[sj][k][b][NS*D]
template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (256) void
k_recip_sf_block (const double4 *__restrict__ pos, const T *__restrict__ XT, int B,
                  const double4 *__restrict__ kv, int np, int nk, int jper,
                  T *__restrict__ part)
{
  constexpr int D = Dim<V>::D;
  constexpr int NS = (NPOW > 1) ? 4 : 2;
  constexpr int JT = 32;
  __shared__ T ph[32][JT][2];
  __shared__ T a2s[JT];
  const int tk = threadIdx.x >> 3, tb = threadIdx.x & 7;
  const int k = blockIdx.x * 32 + tk;
  const int b = blockIdx.y * 8 + tb;
  const int j0 = blockIdx.z * jper;
  const int j1 = min (np, j0 + jper);
  T acc[NS * D];
#pragma unroll
  for (int q = 0; q < NS * D; q++) acc[q] = T (0);
  for (int jb = j0; jb < j1; jb += JT)
    {
      const int n = min (JT, j1 - jb);
      // phases of the (32 k x n j) tile: 4 per thread
      for (int e = threadIdx.x; e < 32 * JT; e += 256)
        {
          const int kk = e / JT, jj = e - kk * JT;
          const int kg = blockIdx.x * 32 + kk;
          if (jj < n && kg < nk)
            {
              const double4 kvv = kv[kg];
              const double4 pj = pos[jb + jj];
              T c, s;
              phase_cs (kvv.x * pj.x + kvv.y * pj.y + kvv.z * pj.z, c, s);
              ph[kk][jj][0] = c;
              ph[kk][jj][1] = s;
            }
        }
      if (NPOW > 1 && threadIdx.x < n)
        {
          const double aw = pos[jb + threadIdx.x].w;
          a2s[threadIdx.x] = T (aw * aw);
        }
      __syncthreads ();
      if (k < nk && b < B)
        for (int t = 0; t < n; t++)
          {
            const T c = ph[tk][t][0], s = ph[tk][t][1];
            const T *zz = XT + (size_t) (jb + t) * D * B + b;
#pragma unroll
            for (int q = 0; q < D; q++)
              {
                const T zq = zz[(size_t) q * B];
                acc[q] += c * zq;
                acc[D + q] += s * zq;
                if (NPOW > 1)
                  {
                    acc[2 * D + q] += a2s[t] * c * zq;
                    acc[3 * D + q] += a2s[t] * s * zq;
                  }
              }
          }
      __syncthreads ();
    }
  if (k < nk && b < B)
    {
      T *out = part + (((size_t) blockIdx.z * nk + k) * B + b) * (NS * D);
#pragma unroll
      for (int q = 0; q < NS * D; q++) out[q] = acc[q];
    }
}

// P, Q per (k, b): PQ[k][b][NPOW][2][D]
template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (128) void
k_recip_blk_block (const double4 *__restrict__ kv, const double *__restrict__ ks, int nk,
                   int B, int nsplit, double gsign, const T *__restrict__ part,
                   T *__restrict__ PQ)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  constexpr int NS = (NPOW > 1) ? 4 : 2;
  // one thread per (k, power of a_i, b)
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) nk * NPOW * B) return;
  const int k = (int) (t / ((size_t) NPOW * B));
  const int r = (int) (t - (size_t) k * NPOW * B), pw = r / B + 1, b = r - (pw - 1) * B;
  double CS[NS * D];
#pragma unroll
  for (int q = 0; q < NS * D; q++) CS[q] = 0.0;
  for (int sp = 0; sp < nsplit; sp++)
    {
      const T *src = part + (((size_t) sp * nk + k) * B + b) * (NS * D);
#pragma unroll
      for (int q = 0; q < NS * D; q++) CS[q] += double (src[q]);
    }
  const double4 kk = kv[k];
  const double *w = ks + (size_t) k * 8;
  T *o = PQ + ((size_t) k * B + b) * (NPOW * 2 * D) + (pw - 1) * 2 * D;
  double P[D], Q[D];
  if (NPOW == 1) kblock_pq_mono<V> (kk, w, CS, P, Q);
  else kblock_pq_poly_rt<V> (pw, kk, w, CS, P, Q);
#pragma unroll
  for (int q = 0; q < D; q++) { o[q] = T (P[q]); o[D + q] = T (Q[q]); }
}

// back-projection, 32 i x 8 columns per workgroup; part[sk][i][q][b]
template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (256) void
k_recip_back_block (const double4 *__restrict__ pos, const double4 *__restrict__ kv,
                    const T *__restrict__ PQ, int np, int nk, int B, int kper,
                    T *__restrict__ part)
{
  constexpr int D = Dim<V>::D;
  constexpr int KT = 32;
  __shared__ T ph[32][KT][2];
  const int ti = threadIdx.x >> 3, tb = threadIdx.x & 7;
  const int i = blockIdx.x * 32 + ti;
  const int b = blockIdx.y * 8 + tb;
  const int k0 = blockIdx.z * kper;
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
      for (int e = threadIdx.x; e < 32 * KT; e += 256)
        {
          const int ii = e / KT, kk = e - ii * KT;
          const int ig = blockIdx.x * 32 + ii;
          if (kk < n && ig < np)
            {
              const double4 kvv = kv[kb + kk];
              const double4 pp = pos[ig];
              T c, s;
              phase_cs (kvv.x * pp.x + kvv.y * pp.y + kvv.z * pp.z, c, s);
              ph[ii][kk][0] = c;
              ph[ii][kk][1] = s;
            }
        }
      __syncthreads ();
      if (i < np && b < B)
        for (int t = 0; t < n; t++)
          {
            const T c = ph[ti][t][0], s = ph[ti][t][1];
            const T *P = PQ + ((size_t) (kb + t) * B + b) * (NPOW * 2 * D);
#pragma unroll
            for (int pw = 0; pw < NPOW; pw++)
              {
                const T cc = apow[pw] * c, ss = apow[pw] * s;
#pragma unroll
                for (int q = 0; q < D; q++)
                  acc[q] += cc * P[pw * 2 * D + q] + ss * P[pw * 2 * D + D + q];
              }
          }
      __syncthreads ();
    }
  if (i < np && b < B)
    {
#pragma unroll
      for (int q = 0; q < D; q++)
        part[(((size_t) blockIdx.z * np + i) * D + q) * B + b] = acc[q];
    }
}

// y = self x + real partials + reciprocal partials, summed in that order
// (partials in chunk order).  x and y hold the columns one after another
// and the partials [chunk][i][q][b] have b fastest, so a workgroup takes a
// tile of CB_I consecutive (i, q) rows and CB_B columns through LDS: the x
// tile read along (i, q), the sums formed along b, y written along (i, q);
// every global access is coalesced.
constexpr int CB_I = 16, CB_B = 64;         // CB_I CB_B a multiple of 256

template <int V, typename T>
__global__ __launch_bounds__ (256) void
k_combine_block (const double4 *__restrict__ pos, const double *__restrict__ x, int B,
                 const T *__restrict__ rpart, int nreal, const T *__restrict__ kpart,
                 int nrecip, DevParams p, int periodic, double *__restrict__ y)
{
  constexpr int D = Dim<V>::D;
  __shared__ double xt[CB_B][CB_I + 1];
  __shared__ double yt[CB_I][CB_B + 1];
  const size_t n = (size_t) p.np * D;
  const size_t iq0 = (size_t) blockIdx.x * CB_I;
  const int b0 = blockIdx.y * CB_B;
  const size_t stride = n * B;
  for (int e = threadIdx.x; e < CB_I * CB_B; e += 256)
    {
      const int il = e % CB_I, bl = e / CB_I;
      const size_t iq = iq0 + il;
      const int b = b0 + bl;
      xt[bl][il] = (iq < n && b < B) ? x[(size_t) b * n + iq] : 0.0;
    }
  __syncthreads ();
  // CB_I CB_B / 256 elements per thread, their partial loads interleaved
#pragma unroll
  for (int u = 0; u < CB_I * CB_B / 256; u++)
    {
      const int e = threadIdx.x + 256 * u;
      const int bl = e % CB_B, il = e / CB_B;
      const size_t iq = iq0 + il;
      const int b = b0 + bl;
      if (iq < n && b < B)
        {
          const int i = (int) (iq / D), q = (int) (iq - (size_t) i * D);
          const size_t t = iq * B + b;
          double acc = self_diag (p, pos[i].w, periodic != 0, V, q) * xt[bl][il];
#pragma unroll 4
          for (int ch = 0; ch < nreal; ch++) acc += double (rpart[(size_t) ch * stride + t]);
#pragma unroll 4
          for (int ch = 0; ch < nrecip; ch++) acc += double (kpart[(size_t) ch * stride + t]);
          yt[il][bl] = acc;
        }
    }
  __syncthreads ();
  for (int e = threadIdx.x; e < CB_I * CB_B; e += 256)
    {
      const int il = e % CB_I, bl = e / CB_I;
      const size_t iq = iq0 + il;
      const int b = b0 + bl;
      if (iq < n && b < B) y[(size_t) b * n + iq] = yt[il][bl];
    }
}

}  // namespace shk
