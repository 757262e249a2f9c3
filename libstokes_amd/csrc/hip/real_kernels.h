// real_kernels.h -- real-space Ewald pair sum, compacted form (ewald.c:649-789).
//
// One wavefront owns one output particle i and 64 source particles j (one
// per lane).  Each lane wraps d0 = x_j - x_i into the primary cell of the
// (possibly sheared) lattice; the images that can lie inside r_max are then
// the lattice vectors v with |v| <= r_max + R_half (R_half: half the cell
// diagonal), sorted by |v| on the host.  The first `ncore` of them
// (|v| <= r_max - R_half) are inside r_max for every lane and are summed in
// lock-step.  The rest (the shell) are tested lane by lane; hits are
// compacted through a per-wave LDS queue and evaluated 64 at a time, so no
// lane idles on a miss.  The reference's own image table decides
// membership as well (bitmap test on m_ref = m - n_wrap), so the image set
// is the reference's even where its table is incomplete (large shear
// shifts).  Every hit of the wave adds to the same y_i: each lane keeps a
// running 11-vector, reduced across the wave at the end, written per
// (j-chunk, i) and summed in a fixed order later: bitwise deterministic.
#pragma once
#include <hip/hip_runtime.h>
#include "pair_math.h"
#include "ewald_kernels.h"

namespace shk {

struct RealParams {
  int np, ncand, ncore, poly;
  double xi, rmax2, rmin;
  double lx, ly, lz, s2x, s3x;   // lattice a1 = (lx,0,0), a2 = (s2x,ly,0), a3 = (s3x,0,lz)
  int tx, ty, tz;                // table half-ranges rmaxx/y/z
  int tbits_words;
};

__device__ __forceinline__ void wave_sync ()
{
  __builtin_amdgcn_fence (__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier ();
  __builtin_amdgcn_fence (__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T> struct Vec4;
template <> struct Vec4<double> { using type = double4; };
template <> struct Vec4<float> { using type = float4; };

template <int V, typename T> struct RealLds {
  static constexpr int D = Dim<V>::D;
  T d[64][3];
  T z[64][D];
  T aj[64];
  int queue[128];
};

template <int V, typename T>
__device__ __forceinline__ void real_item (const RealParams &p, T ai, T xi,
                                           const RealLds<V, T> &L, int src, T dx,
                                           T dy, T dz, T rr, T *acc)
{
  // one reciprocal square root per pair-image: u = 1/r, r = rr u
  T ir = m_rsqrt (rr);
  T r = rr * ir;
  const T aj = L.aj[src];
  const T rmn = (ai + aj) * T (p.rmin);
  if (r < rmn)
    {
      r = rmn;
      ir = T (1) / rmn;
    }
  T s[11];
  if (p.poly) scalars_real_poly_u<V, T> (xi, r, ir, ai, aj, s);
  else scalars_real_u<V, T> (xi, r, ir, s);
  contract_item<V, T> (dx * ir, dy * ir, dz * ir, s, L.z[src], acc);
}

// table membership of reference image m (bitmap over the reference's table)
__device__ __forceinline__ bool in_table (const RealParams &p, const unsigned *bits, int m1,
                                          int m2, int m3)
{
  if (m1 < -p.tx || m1 > p.tx || m2 < -p.ty || m2 > p.ty || m3 < -p.tz || m3 > p.tz)
    return false;
  const int idx = ((m1 + p.tx) * (2 * p.ty + 1) + (m2 + p.ty)) * (2 * p.tz + 1) + (m3 + p.tz);
  return (bits[idx >> 5] >> (idx & 31)) & 1u;
}

// The pair images of output particle i and the 64 sources of chunk jc (one
// per lane): loads the chunk into the wave's LDS, then the core and shell
// loops add every image inside r_max to acc (per lane, not yet reduced).
// Every lane of the wave calls it with the same (i, jc).
template <int V, typename T>
__device__ __forceinline__ void
real_chunk (const double4 *__restrict__ pos, const double *__restrict__ x,
            const typename Vec4<T>::type *__restrict__ cand, const int4 *__restrict__ candm,
            const unsigned *bits, const RealParams &p, int i, int jc, RealLds<V, T> &L,
            int lane, T *acc)
{
  constexpr int D = Dim<V>::D;
  const int j = jc * 64 + lane;
  const bool valid = j < p.np;
  const double4 pi = pos[i];
  const double4 pj = pos[valid ? j : i];
  // wrap into the primary cell of the lattice: z, then y, then x
  double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
  const double n3 = rint (dz / p.lz);
  dx -= n3 * p.s3x; dz -= n3 * p.lz;
  const double n2 = rint (dy / p.ly);
  dx -= n2 * p.s2x; dy -= n2 * p.ly;
  const double n1 = rint (dx / p.lx);
  dx -= n1 * p.lx;
  const int w1 = int (n1), w2 = int (n2), w3 = int (n3);
  wave_sync ();                   // the previous chunk's LDS reads are done
  L.d[lane][0] = T (dx); L.d[lane][1] = T (dy); L.d[lane][2] = T (dz);
  {
    T zz[D];
    load_z<V, T> (x, valid ? j : 0, zz);
#pragma unroll
    for (int q = 0; q < D; q++) L.z[lane][q] = zz[q];
  }
  L.aj[lane] = T (pj.w);
  wave_sync ();

  const T ai = T (pi.w), xi = T (p.xi), rmax2 = T (p.rmax2);
  const T mdx = L.d[lane][0], mdy = L.d[lane][1], mdz = L.d[lane][2];

  // core: inside r_max for every lane
  for (int c = 0; c < p.ncore; c++)
    {
      const typename Vec4<T>::type v = cand[c];
      const int4 mm = candm[c];
      const T ddx = mdx + v.x, ddy = mdy + v.y, ddz = mdz + v.z;
      const T rr = ddx * ddx + ddy * ddy + ddz * ddz;
      if (valid && rr != T (0) && rr <= rmax2
          && in_table (p, bits, mm.x - w1, mm.y - w2, mm.z - w3))
        real_item<V, T> (p, ai, xi, L, lane, ddx, ddy, ddz, rr, acc);
    }

  // shell: test, compact hits into the queue, evaluate 64 at a time
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
          real_item<V, T> (p, ai, xi, L, src, ex, ey, ez, ex * ex + ey * ey + ez * ez, acc);
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
      real_item<V, T> (p, ai, xi, L, src, ex, ey, ez, ex * ex + ey * ey + ez * ez, acc);
    }
}

// sum of acc over the wave (fixed butterfly), into every lane
template <int D, typename T>
__device__ __forceinline__ void wave_reduce (T *acc)
{
#pragma unroll
  for (int q = 0; q < D; q++)
    {
      T v = acc[q];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor (v, off, 64);
      acc[q] = v;
    }
}

template <int V, typename T>
__global__ __launch_bounds__ (256) void
k_real_compact (const double4 *__restrict__ pos, const double *__restrict__ x,
                const typename Vec4<T>::type *__restrict__ cand,
                const int4 *__restrict__ candm, const unsigned *__restrict__ tbits,
                RealParams p, double *__restrict__ partial)
{
  constexpr int D = Dim<V>::D;
  __shared__ RealLds<V, T> lds_all[4];
  __shared__ unsigned bits_lds[1024];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  RealLds<V, T> &L = lds_all[wave];
  const int nj = (p.np + 63) / 64;
  const int gw = blockIdx.x * 4 + wave;
  const int i = gw / nj, jc = gw - i * nj;

  // table bitmap to LDS (whole block; every wave passes this barrier)
  const bool bits_in_lds = p.tbits_words <= 1024;
  if (bits_in_lds)
    for (int w = threadIdx.x; w < p.tbits_words; w += blockDim.x) bits_lds[w] = tbits[w];
  __syncthreads ();
  const unsigned *bits = bits_in_lds ? bits_lds : tbits;
  if (i >= p.np) return;

  T acc[D];
#pragma unroll
  for (int q = 0; q < D; q++) acc[q] = T (0);
  real_chunk<V, T> (pos, x, cand, candm, bits, p, i, jc, L, lane, acc);
  wave_reduce<D> (acc);
  if (lane == 0)
    {
      double *out = partial + ((size_t) jc * p.np + i) * D;
#pragma unroll
      for (int q = 0; q < D; q++) out[q] = double (acc[q]);
    }
}

}  // namespace shk
