// adj_kernels.h -- adjoint of the F-version Ewald mobility product with
// respect to the positions: grad_i = d(g . M(x) f)/dx_i, equal spheres.
//
// M(x) = self + sum over pair-images of T(d) + reciprocal sum, d = x_j -
// x_i + image.  T(d) = c(r) d d^T + ya(r) I with c = (xa - ya) / r_c^2,
// r_c = max(|d|, 2 r_min) (the reference clamps r but keeps e = d / r_c,
// ewald.c:688-700).  With h(d) = g_i.T f_j + g_j.T f_i and the image set
// symmetric, grad_i = -sum_{j, images} grad_d h(d), where for a.T b
//   grad_d = c' (a.d)(b.d) d/r + c [(b.d) a + (a.d) b] + ya' (a.b) d/r,
// c' = (xa' - ya') / r^2 - 2 (xa - ya) / r^3 unclamped and 0 clamped.
// The pair scalars are recomputed, not stored.
//
// Reciprocal part (ewald.c:923-1075, F block ya(k) (I - k^ k^)):
//   grad_i = sum_k ya_k k [g_i.P (c_i S_f - s_i C_f) + f_i.P (c_i S_g - s_i C_g)]
// with c_i, s_i = cos, sin(k.x_i), C_v = sum_j cos(k.x_j) v_j,
// S_v = sum_j sin(k.x_j) v_j and P = I - k^ k^; the structure factors of
// (f, g) come from k_recip_sf on the 6-vector (f_j, g_j).
#pragma once
#include <hip/hip_runtime.h>
#include "pair_math.h"
#include "ewald_kernels.h"
#include "real_kernels.h"

namespace shk {

template <typename T> struct AdjLds {
  T d[64][3];
  T fg[64][6];   // f_j, g_j
  int queue[128];
};

template <typename T>
__device__ __forceinline__ void adj_item (const RealParams &p, T xi, const T *fi, const T *gi,
                                          const AdjLds<T> &L, int src, T dx, T dy, T dz, T rr,
                                          T *acc)
{
  T ir = m_rsqrt (rr);
  const T r = rr * ir;
  const T rmn = T (2) * T (p.rmin);
  T rc = r, irc = ir;
  const bool clamped = r < rmn;
  if (clamped)
    {
      rc = rmn;
      irc = T (1) / rmn;
    }
  T xa, ya, dxa, dya;
  scalars_real_f_deriv<T> (xi, rc, irc, &xa, &ya, &dxa, &dya);
  const T irc2 = irc * irc;
  const T c = (xa - ya) * irc2;
  T dc = (dxa - dya) * irc2 - T (2) * (xa - ya) * irc2 * irc;
  if (clamped)
    {
      dc = T (0);
      dya = T (0);
    }
  const T *fj = L.fg[src], *gj = L.fg[src] + 3;
  // pair (g_i, f_j) and pair (g_j, f_i)
  const T ad1 = gi[0] * dx + gi[1] * dy + gi[2] * dz;
  const T bd1 = fj[0] * dx + fj[1] * dy + fj[2] * dz;
  const T ab1 = gi[0] * fj[0] + gi[1] * fj[1] + gi[2] * fj[2];
  const T ad2 = gj[0] * dx + gj[1] * dy + gj[2] * dz;
  const T bd2 = fi[0] * dx + fi[1] * dy + fi[2] * dz;
  const T ab2 = gj[0] * fi[0] + gj[1] * fi[1] + gj[2] * fi[2];
  // radial part along d / r
  const T rad = (dc * (ad1 * bd1 + ad2 * bd2) + dya * (ab1 + ab2)) * ir;
  // grad_i -= grad_d h
  acc[0] -= rad * dx + c * (bd1 * gi[0] + ad1 * fj[0] + bd2 * gj[0] + ad2 * fi[0]);
  acc[1] -= rad * dy + c * (bd1 * gi[1] + ad1 * fj[1] + bd2 * gj[1] + ad2 * fi[1]);
  acc[2] -= rad * dz + c * (bd1 * gi[2] + ad1 * fj[2] + bd2 * gj[2] + ad2 * fi[2]);
}

// same enumeration as k_real_compact (minimum-image wrap, sorted lattice
// candidates, core without tests, compacted shell, reference-table
// membership); fg holds (f_j, g_j) per particle, partial[(jc*np + i)*3]
template <typename T>
__global__ __launch_bounds__ (256) void
k_real_adj (const double4 *__restrict__ pos, const double *__restrict__ fg,
            const typename Vec4<T>::type *__restrict__ cand, const int4 *__restrict__ candm,
            const unsigned *__restrict__ tbits, RealParams p, double *__restrict__ partial)
{
  __shared__ AdjLds<T> lds_all[4];
  __shared__ unsigned bits_lds[1024];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  AdjLds<T> &L = lds_all[wave];
  const int nj = (p.np + 63) / 64;
  const int gw = blockIdx.x * 4 + wave;
  const int i = gw / nj, jc = gw - i * nj;

  const bool bits_in_lds = p.tbits_words <= 1024;
  if (bits_in_lds)
    for (int w = threadIdx.x; w < p.tbits_words; w += blockDim.x) bits_lds[w] = tbits[w];
  __syncthreads ();
  const unsigned *bits = bits_in_lds ? bits_lds : tbits;
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
  L.d[lane][0] = T (dx); L.d[lane][1] = T (dy); L.d[lane][2] = T (dz);
  const int jj = valid ? j : 0;
#pragma unroll
  for (int q = 0; q < 6; q++) L.fg[lane][q] = valid ? T (fg[6 * (size_t) jj + q]) : T (0);
  wave_sync ();

  T fi[3], gi[3];
#pragma unroll
  for (int q = 0; q < 3; q++)
    {
      fi[q] = T (fg[6 * (size_t) i + q]);
      gi[q] = T (fg[6 * (size_t) i + 3 + q]);
    }
  const T xi = T (p.xi), rmax2 = T (p.rmax2);
  const T mdx = L.d[lane][0], mdy = L.d[lane][1], mdz = L.d[lane][2];
  T acc[3] = {T (0), T (0), T (0)};

  for (int c = 0; c < p.ncore; c++)
    {
      const typename Vec4<T>::type v = cand[c];
      const int4 mm = candm[c];
      const T ddx = mdx + v.x, ddy = mdy + v.y, ddz = mdz + v.z;
      const T rr = ddx * ddx + ddy * ddy + ddz * ddz;
      if (valid && rr != T (0) && rr <= rmax2
          && in_table (p, bits, mm.x - w1, mm.y - w2, mm.z - w3))
        adj_item<T> (p, xi, fi, gi, L, lane, ddx, ddy, ddz, rr, acc);
    }

  int count = 0;
  for (int c = p.ncore; c < p.ncand; c++)
    {
      const typename Vec4<T>::type v = cand[c];
      const int4 mm = candm[c];
      const T ddx = mdx + v.x, ddy = mdy + v.y, ddz = mdz + v.z;
      const T rr = ddx * ddx + ddy * ddy + ddz * ddz;
      const bool hit = valid && rr != T (0) && rr <= rmax2
        && in_table (p, bits, mm.x - w1, mm.y - w2, mm.z - w3);
      const unsigned long long bal = __ballot (hit);
      if (bal == 0ull) continue;
      const int below = __popcll (bal & ((1ull << lane) - 1ull));
      if (hit) L.queue[count + below] = (c << 6) | lane;
      count += __popcll (bal);
      wave_sync ();
      if (count >= 64)
        {
          const int it = L.queue[lane];
          const int src = it & 63, cc = it >> 6;
          const typename Vec4<T>::type w = cand[cc];
          const T ex = L.d[src][0] + w.x, ey = L.d[src][1] + w.y, ez = L.d[src][2] + w.z;
          adj_item<T> (p, xi, fi, gi, L, src, ex, ey, ez, ex * ex + ey * ey + ez * ez, acc);
          wave_sync ();
          count -= 64;
          if (lane < count) L.queue[lane] = L.queue[64 + lane];
          wave_sync ();
        }
    }
  if (lane < count)
    {
      const int it = L.queue[lane];
      const int src = it & 63, cc = it >> 6;
      const typename Vec4<T>::type w = cand[cc];
      const T ex = L.d[src][0] + w.x, ey = L.d[src][1] + w.y, ez = L.d[src][2] + w.z;
      adj_item<T> (p, xi, fi, gi, L, src, ex, ey, ez, ex * ex + ey * ey + ez * ez, acc);
    }

#pragma unroll
  for (int q = 0; q < 3; q++)
    {
      T v = acc[q];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor (v, off, 64);
      acc[q] = v;
    }
  if (lane == 0)
    {
      double *out = partial + ((size_t) jc * p.np + i) * 3;
#pragma unroll
      for (int q = 0; q < 3; q++) out[q] = double (acc[q]);
    }
}

// per k: the split partial sums of the (f, g) structure factors, summed in
// a fixed order: CS[k] = (C_f, C_g, S_f, S_g) (12 values)
template <typename T>
__global__ void k_adj_sf_sum (const T *__restrict__ part, int nk, int nsplit, T *__restrict__ CS)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  double acc[12];
#pragma unroll
  for (int q = 0; q < 12; q++) acc[q] = 0.0;
  for (int sp = 0; sp < nsplit; sp++)
    {
      const T *src = part + ((size_t) sp * nk + k) * 12;
#pragma unroll
      for (int q = 0; q < 12; q++) acc[q] += double (src[q]);
    }
#pragma unroll
  for (int q = 0; q < 12; q++) CS[(size_t) k * 12 + q] = T (acc[q]);
}

// grad_i (reciprocal) -- thread per i, k split over blockIdx.y, k tiles in
// LDS; out[(split*np + i)*3]
template <typename T>
__global__ __launch_bounds__ (256) void
k_adj_back (const double4 *__restrict__ pos, const double *__restrict__ fg,
            const double4 *__restrict__ kv, const double *__restrict__ ks,
            const T *__restrict__ CS, int np, int nk, int kper, T *__restrict__ out)
{
  constexpr int KT = 32;
  __shared__ double4 skv[KT];
  __shared__ T sw[KT];
  __shared__ T scs[KT][12];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int k0 = blockIdx.y * kper;
  const int k1 = min (nk, k0 + kper);
  const int ii = min (i, np - 1);
  const double4 pi = pos[ii];
  T fi[3], gi[3];
#pragma unroll
  for (int q = 0; q < 3; q++)
    {
      fi[q] = T (fg[6 * (size_t) ii + q]);
      gi[q] = T (fg[6 * (size_t) ii + 3 + q]);
    }
  T acc[3] = {T (0), T (0), T (0)};
  for (int kb = k0; kb < k1; kb += KT)
    {
      const int n = min (KT, k1 - kb);
      for (int e = threadIdx.x; e < n * 12; e += 256)
        scs[e / 12][e % 12] = CS[(size_t) (kb + e / 12) * 12 + e % 12];
      if (threadIdx.x < n)
        {
          skv[threadIdx.x] = kv[kb + threadIdx.x];
          sw[threadIdx.x] = T (ks[(size_t) (kb + threadIdx.x) * 8]);   // 2 ya_k
        }
      __syncthreads ();
      for (int t = 0; t < n; t++)
        {
          const double4 kk = skv[t];
          T c, s;
          phase_cs (kk.x * pi.x + kk.y * pi.y + kk.z * pi.z, c, s);
          const T *cs = scs[t];
          // vf = c S_f - s C_f, vg = c S_g - s C_g
          T vf[3], vg[3];
#pragma unroll
          for (int q = 0; q < 3; q++)
            {
              vf[q] = c * cs[6 + q] - s * cs[q];
              vg[q] = c * cs[9 + q] - s * cs[3 + q];
            }
          const T ik = T (1) / T (kk.w);
          const T ex = T (kk.x) * ik, ey = T (kk.y) * ik, ez = T (kk.z) * ik;
          const T ke_g = ex * gi[0] + ey * gi[1] + ez * gi[2];
          const T ke_f = ex * fi[0] + ey * fi[1] + ez * fi[2];
          const T ke_vf = ex * vf[0] + ey * vf[1] + ez * vf[2];
          const T ke_vg = ex * vg[0] + ey * vg[1] + ez * vg[2];
          const T t1 = gi[0] * vf[0] + gi[1] * vf[1] + gi[2] * vf[2] - ke_g * ke_vf;
          const T t2 = fi[0] * vg[0] + fi[1] * vg[1] + fi[2] * vg[2] - ke_f * ke_vg;
          const T wgt = sw[t] * (t1 + t2);
          acc[0] += wgt * T (kk.x);
          acc[1] += wgt * T (kk.y);
          acc[2] += wgt * T (kk.z);
        }
      __syncthreads ();
    }
  if (i < np)
    {
      T *o = out + ((size_t) blockIdx.y * np + i) * 3;
#pragma unroll
      for (int q = 0; q < 3; q++) o[q] = acc[q];
    }
}

// dx_i = sum of real partials + sum of reciprocal partials (fixed order)
template <typename T>
__global__ void k_adj_combine (const double *__restrict__ rpart, int nreal,
                               const T *__restrict__ kpart, int nrecip, int np,
                               double *__restrict__ dx)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t) np * 3;
  if (t >= n) return;
  double acc = 0.0;
  for (int ch = 0; ch < nreal; ch++) acc += rpart[(size_t) ch * n + t];
  for (int ch = 0; ch < nrecip; ch++) acc += double (kpart[(size_t) ch * n + t]);
  dx[t] = acc;
}

// (f, g) interleaved per particle
__global__ void k_adj_pack (const double *__restrict__ f, const double *__restrict__ g, int np,
                            double *__restrict__ fg)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int q = 0; q < 3; q++)
    {
      fg[6 * (size_t) i + q] = f[3 * (size_t) i + q];
      fg[6 * (size_t) i + 3 + q] = g[3 * (size_t) i + q];
    }
}

}  // namespace shk
