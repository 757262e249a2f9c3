// matrix_kernels.h -- the dense extracted-form FP64 mobility matrix of equal
// spheres in a periodic box whose image table is symmetric under m -> -m
// (make_matrix_mob_ewald_3all, ewald.c:1092-1677).
//
// The block of a pair is a fixed linear map of its moments (pair_math.h);
// the moments of (j, i) are those of (i, j) with the odd ones negated, in
// real space (the images of (j, i) are those of (i, j) reflected) and in
// reciprocal space (cos k.(x_i - x_j) is even, sin odd, and the odd
// moments come with the odd scalars yb, yg).  So one thread per pair
// i <= j sums the moments once and writes both blocks.
//
//   k_kmom_weights  per k: W_k[t] = the moments of k^ with the scalars of k
//                   (ya, yb, yc, yg, yh, ym of the k table); the moments of
//                   a pair from k are then cos(k.d) W_k[t] for even t and
//                   -sin(k.d) W_k[t] for odd t (moments_add is linear in
//                   its scalars)
//   k_matrix_sym    workgroup 16 i x 16 j (j-tile >= i-tile): per pair the
//                   image masks over the staged candidates, then the
//                   real-space moments (pair_images_add), the self term on
//                   the diagonal, then the k
//                   sum from the phase table E (cos and sin of k.x per
//                   particle, k.d = k.x_j - k.x_i by two products) and W,
//                   both staged in LDS 32 k at a time; blocks (i, j) and
//                   (j, i) written to the row-major (D np)^2 matrix.
#pragma once
#include <hip/hip_runtime.h>
#include "pair_math.h"
#include "ewald_kernels.h"
#include "real_kernels.h"
#include "mfma_kernels.h"

namespace shk {

template <int V>
__global__ void k_kmom_weights (const double4 *__restrict__ kv, const double *__restrict__ ks,
                                int nk, double *__restrict__ W)
{
  constexpr int NM = Dim<V>::NMOM;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  const double4 kk = kv[k];
  const double *w = ks + (size_t) k * 8;
  const double ik = 1.0 / kk.w;
  const double sc[11] = {0.0, w[0], w[1], 0.0, w[2], 0.0, w[3], w[4], 0.0, w[5], 0.0};
  double M[NM];
#pragma unroll
  for (int t = 0; t < NM; t++) M[t] = 0.0;
  moments_add<V, double> (M, kk.x * ik, kk.y * ik, kk.z * ik, sc);
#pragma unroll
  for (int t = 0; t < NM; t++) W[(size_t) k * NM + t] = M[t];
}

constexpr int MS_KC = 32;     // k per LDS stage of k_matrix_sym

template <int V>
__global__ __launch_bounds__ (256) __attribute__ ((amdgpu_waves_per_eu (2))) void
k_matrix_sym (const double4 *__restrict__ pos, const double4 *__restrict__ cand,
              const int4 *__restrict__ candm, const unsigned *__restrict__ tbits, RealParams rp,
              DevParams q, const double *__restrict__ E, const double *__restrict__ W, int nk,
              double *__restrict__ mat)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  if (blockIdx.y < blockIdx.x) return;
  __shared__ double4 cand_l[REAL_PC_CAND];
  __shared__ unsigned long long tmask_l[REAL_PC_CAND / 64][27];
  __shared__ double ph_l[MS_KC][4][16];        // cos x_i, sin x_i, cos x_j, sin x_j
  __shared__ double w_l[MS_KC][NM];
  stage_candidates<double4> (rp, cand, candm, tbits, cand_l, tmask_l);
  __syncthreads ();
  const int np = rp.np;
  const int jl = threadIdx.x & 15, il = threadIdx.x >> 4;
  const int i = blockIdx.x * 16 + il, j = blockIdx.y * 16 + jl;
  const bool valid = i <= j && j < np;
  // the image masks first, the moments after them
  PairSep sp;
  unsigned long long mk[REAL_NCH] = {};
  double4 pi, pj;
  if (valid)
    {
      pi = pos[i]; pj = pos[j];
      sp = pair_sep (pi, pj, rp);
      pair_masks<double> (sp, rp, cand_l, candm, tbits, tmask_l, mk);
    }
  double M[NM];
#pragma unroll
  for (int t = 0; t < NM; t++) M[t] = 0.0;
  if (valid)
    {
      if (i == j)
        {
          double s[11];
          self_scalars<double> (q, pi.w, true, s);
          moments_add<V, double> (M, 0.0, 0.0, 0.0, s);
        }
      pair_images_add<V, double> (sp, rp.xi, (pi.w + pj.w) * rp.rmin, cand_l, mk, M);
    }
  for (int k0 = 0; k0 < nk; k0 += MS_KC)
    {
      const int kn = min (MS_KC, nk - k0);
      __syncthreads ();
      for (int e = threadIdx.x; e < MS_KC * 64; e += 256)
        {
          const int kk = e >> 6, r = e & 63, which = r >> 4, l = r & 15;
          const int idx = (which < 2 ? blockIdx.x : blockIdx.y) * 16 + l;
          const int row = (which & 1) * nk + k0 + kk;
          ph_l[kk][which][l] = (kk < kn && idx < np) ? E[(size_t) row * np + idx] : 0.0;
        }
      for (int e = threadIdx.x; e < kn * NM; e += 256) w_l[e / NM][e % NM] = W[(size_t) k0 * NM + e];
      __syncthreads ();
      if (valid)
        for (int kk = 0; kk < kn; kk++)
          {
            const double ci = ph_l[kk][0][il], si = ph_l[kk][1][il];
            const double cj = ph_l[kk][2][jl], sj = ph_l[kk][3][jl];
            const double c = ci * cj + si * sj;          // cos k.(x_j - x_i)
            const double sf = -(ci * sj - si * cj);      // -sin k.(x_j - x_i)
            const double *w = w_l[kk];
#pragma unroll
            for (int t = 0; t < NM; t++)
              M[t] += (moment_parity<V> (t) < 0.f ? sf : c) * w[t];
          }
    }
  if (!valid) return;
  const size_t n = (size_t) np * D;
  {
    double *blk = mat + (size_t) i * D * n + (size_t) j * D;
    auto put = [&] (int r, int c, double v) { blk[(size_t) r * n + c] = v; };
    BlockEach<V, 0, D, D * D>::run (M, i == j, put);
  }
  if (i == j) return;
#pragma unroll
  for (int t = 0; t < NM; t++)
    if (moment_parity<V> (t) < 0.f) M[t] = -M[t];
  double *blk = mat + (size_t) j * D * n + (size_t) i * D;
  auto put = [&] (int r, int c, double v) { blk[(size_t) r * n + c] = v; };
  BlockEach<V, 0, D, D * D>::run (M, false, put);
}

}  // namespace shk
