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
//                 lane order (k_real_gemv16: 16-byte loads when the row
//                 length allows)
#pragma once
#include <hip/hip_runtime.h>
#include "pair_math.h"
#include "ewald_kernels.h"
#include "real_kernels.h"
#include "mfma_kernels.h"

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

// 16-byte vectors of the matrix precision
typedef float f32x4 __attribute__ ((ext_vector_type (4)));
typedef double f64x2 __attribute__ ((ext_vector_type (2)));
template <typename T> struct Vec16;
template <> struct Vec16<float> {
  using type = f32x4;
  static constexpr int W = 4;
  __device__ static double dot (const f32x4 &m, const f32x4 &z)
  {
    return double (m.x) * double (z.x) + double (m.y) * double (z.y)
      + double (m.z) * double (z.z) + double (m.w) * double (z.w);
  }
};
template <> struct Vec16<double> {
  using type = f64x2;
  static constexpr int W = 2;
  __device__ static double dot (const f64x2 &m, const f64x2 &z)
  {
    return m.x * z.x + m.y * z.y;
  }
};

// y_real = Mr z for ND a multiple of the 16-byte width W: one wave per
// row, each lane reads 16-byte vectors (four in flight per lane), FP64
// accumulation, fixed lane order.  XIN: z is the caller's FP64 vector x
// itself (F and FT, where the S transform is the identity), rounded to T
// as k_z_prep would
template <typename T, bool XIN>
__global__ __launch_bounds__ (256) void
k_real_gemv16 (const T *__restrict__ Mr, const void *__restrict__ zv, int ND,
               double *__restrict__ out)
{
  using V16 = typename Vec16<T>::type;
  constexpr int W = Vec16<T>::W;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  if (row >= ND) return;
  const V16 *m = (const V16 *) (Mr + (size_t) row * ND);
  auto zat = [&] (int c) -> V16 {
    if constexpr (XIN)
      {
        const double *xd = (const double *) zv + (size_t) c * W;
        V16 v;
#pragma unroll
        for (int t = 0; t < W; t++) v[t] = T (xd[t]);
        return v;
      }
    else
      return ((const V16 *) zv)[c];
  };
  const int nv = ND / W;
  double acc = 0.0;
  int c = lane;
  for (; c + 192 < nv; c += 256)
    {
      const V16 m0 = __builtin_nontemporal_load (m + c);
      const V16 m1 = __builtin_nontemporal_load (m + c + 64);
      const V16 m2 = __builtin_nontemporal_load (m + c + 128);
      const V16 m3 = __builtin_nontemporal_load (m + c + 192);
      acc += Vec16<T>::dot (m0, zat (c)) + Vec16<T>::dot (m1, zat (c + 64))
        + Vec16<T>::dot (m2, zat (c + 128)) + Vec16<T>::dot (m3, zat (c + 192));
    }
  for (; c < nv; c += 64) acc += Vec16<T>::dot (m[c], zat (c));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor (acc, off, 64);
  if (lane == 0) out[row] = acc;
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

// Y_s = Mr[:, K_s] Z[K_s, :] for K slice s of the FP32 cached matrix
// (row-major ND x ND) and a block Z (row-major ND x B, B <= 16 CT): split K
// so that a 64-column block still fills the chip; the slices' partials
// [s][row][b] are summed in slice order by the combine (deterministic).
// Workgroup = 4 waves, each on RT tiles of 16 rows (RT = 1: 64 rows and
// about 768 workgroups measured fastest for a 6144^2 matrix and 64 columns,
// ahead of more rows per workgroup and longer or shorter K slices);
// per chunk of 64 k the Z rows go to LDS once for all of them (two LDS
// buffers: the next chunk's A and Z load while the current one multiplies); each lane
// reads 4 consecutive k of each of its rows (16-byte loads), and MFMA m of
// a sub-step takes k = kc + 16 sub + 4 (lane / 16) + m in its K slot
// lane / 16, in A (row lane % 16) and B (column lane % 16) alike; one LDS
// read gives a lane its B operands of all CT column tiles, each serving the
// RT row tiles.
__host__ __device__ constexpr int gemm_rt (int) { return 1; }
template <int CT, bool VEC>
__global__ __launch_bounds__ (256) void
k_cache_gemm (const float *__restrict__ Mr, const float *__restrict__ Z, int ND, int B,
              int kslice, float *__restrict__ part)
{
  constexpr int RT = gemm_rt (CT);
  constexpr int ZS = 16 * CT + 4;
  constexpr int ZQ = 64 * 4 * CT / 256;           // float4 of Z per thread per chunk
  __shared__ __align__ (16) float zl[2][64 * ZS];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = blockIdx.x * (64 * RT) + wave * (16 * RT);
  const int k_begin = blockIdx.y * kslice;
  const int k_end = min (ND, k_begin + kslice);
  mf4 acc[RT][CT];
  const float *arow[RT];
#pragma unroll
  for (int rt = 0; rt < RT; rt++)
    {
      const int row = min (r0 + 16 * rt + li, ND - 1);
      arow[rt] = Mr + (size_t) row * ND;
#pragma unroll
      for (int ct = 0; ct < CT; ct++) acc[rt][ct] = mf4{0.f, 0.f, 0.f, 0.f};
    }
  // A of one chunk (4 sub-steps x RT rows x 4 k per lane)
  float a[2][4][RT][4];
  auto load_a = [&] (int kc, float (&dst)[4][RT][4]) {
#pragma unroll
    for (int sub = 0; sub < 4; sub++)
      {
        const int kb = kc + 16 * sub + 4 * lk;
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
          {
            if (VEC && kb + 3 < k_end)
              {
                const f32x4 v = *(const f32x4 *) (arow[rt] + kb);
                dst[sub][rt][0] = v.x; dst[sub][rt][1] = v.y;
                dst[sub][rt][2] = v.z; dst[sub][rt][3] = v.w;
              }
            else
              {
#pragma unroll
                for (int m = 0; m < 4; m++)
                  dst[sub][rt][m] = kb + m < k_end ? arow[rt][kb + m] : 0.f;
              }
          }
      }
  };
  // Z rows of a chunk: thread t holds float4 q = t + 256 h, row q / (4 CT),
  // columns 4 (q % (4 CT)) .. +3 (zero past k_end or B)
  f32x4 zr4[ZQ];
  auto load_z = [&] (int kc) {
#pragma unroll
    for (int h = 0; h < ZQ; h++)
      {
        const int q = threadIdx.x + 256 * h;
        const int kr = q / (4 * CT), c = 4 * (q - kr * 4 * CT);
        const int k = kc + kr;
        f32x4 v = {0.f, 0.f, 0.f, 0.f};
        if (k < k_end)
          {
            if ((B & 3) == 0)
              {
                if (c < B) v = *(const f32x4 *) (Z + (size_t) k * B + c);
              }
            else
              {
                const float *zp = Z + (size_t) k * B;
                if (c < B) v.x = zp[c];
                if (c + 1 < B) v.y = zp[c + 1];
                if (c + 2 < B) v.z = zp[c + 2];
                if (c + 3 < B) v.w = zp[c + 3];
              }
          }
        zr4[h] = v;
      }
  };
  // LDS row kr: column c = 16 ct + l at l CT + ct (a lane's CT column tiles
  // side by side: one LDS read per k)
  auto store_z = [&] (float *buf) {
#pragma unroll
    for (int h = 0; h < ZQ; h++)
      {
        const int q = threadIdx.x + 256 * h;
        const int kr = q / (4 * CT), c = 4 * (q - kr * 4 * CT);
        float *zrow = buf + kr * ZS + (c & 15) * CT + (c >> 4);
        zrow[0] = zr4[h].x; zrow[CT] = zr4[h].y; zrow[2 * CT] = zr4[h].z; zrow[3 * CT] = zr4[h].w;
      }
  };
  // one chunk from LDS buffer zb and registers acur; the next chunk's A and
  // Z are loaded meanwhile and its Z stored to the other buffer
  auto chunk = [&] (int kc, float (&acur)[4][RT][4], float (&anext)[4][RT][4],
                    const float *zb, float *znext) {
    const bool more = kc + 64 < k_end;
    if (more)
      {
        load_a (kc + 64, anext);
        load_z (kc + 64);
      }
    // the chunk's B operands first (one LDS latency per chunk), then its
    // 16 RT CT matrix-core products
    float bz[4][4][CT];
#pragma unroll
    for (int sub = 0; sub < 4; sub++)
#pragma unroll
      for (int m = 0; m < 4; m++)
        {
          const float *zr = zb + (16 * sub + 4 * lk + m) * ZS + li * CT;
          if constexpr (CT >= 4)
            {
#pragma unroll
              for (int h = 0; h < CT / 4; h++)
                {
                  const f32x4 v = *(const f32x4 *) (zr + 4 * h);
                  bz[sub][m][4 * h] = v.x; bz[sub][m][4 * h + 1] = v.y;
                  bz[sub][m][4 * h + 2] = v.z; bz[sub][m][4 * h + 3] = v.w;
                }
            }
          else
            {
#pragma unroll
              for (int ct = 0; ct < CT; ct++) bz[sub][m][ct] = zr[ct];
            }
        }
#pragma unroll
    for (int sub = 0; sub < 4; sub++)
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int ct = 0; ct < CT; ct++)
#pragma unroll
          for (int rt = 0; rt < RT; rt++)
            acc[rt][ct] = mfma4 (acur[sub][rt][m], bz[sub][m][ct], acc[rt][ct]);
    if (more) store_z (znext);
    __syncthreads ();
  };
  if (k_begin < k_end)
    {
      load_a (k_begin, a[0]);
      load_z (k_begin);
      store_z (zl[0]);
      __syncthreads ();
    }
  for (int kc = k_begin; kc < k_end; kc += 128)
    {
      chunk (kc, a[0], a[1], zl[0], zl[1]);
      if (kc + 64 < k_end) chunk (kc + 64, a[1], a[0], zl[1], zl[0]);
    }
#pragma unroll
  for (int rt = 0; rt < RT; rt++)
#pragma unroll
    for (int ct = 0; ct < CT; ct++)
#pragma unroll
      for (int e = 0; e < 4; e++)
        {
          const int r = r0 + 16 * rt + 4 * lk + e, c = 16 * ct + li;
          if (r < ND && c < B) part[((size_t) blockIdx.y * ND + r) * B + c] = acc[rt][ct][e];
        }
}

}  // namespace shk
