// mfma_kernels.h -- reciprocal-space Ewald sum (ewald.c:923-1075 recast) for
// a block of B right-hand sides as FP32 matrix-core contractions
// (v_mfma_f32_16x16x4_f32, FP32 products accumulated in FP32).
//
// With the phases e^{i k.x_j} of the half space as a complex nk x np matrix
// E and the (S-transformed) forces as an np x C matrix Z (C = D B columns,
// twice that with the a_j^2-weighted copy for unequal spheres):
//
//   k_sf_mfma    Cf = Re(E) Z, Sf = Im(E) Z              (nk x C, twice)
//   k_blk_mfma   per (k, b): P, Q from Cf, Sf by the FP64 k-blocks
//   k_back_mfma  y = sum_p a^p [Re(E)^T P_p + Im(E)^T Q_p]  (np x D B)
//
// and the real-space pair sum as Y = A Z with A the (D np) x (D np) matrix
// of the image-summed pair blocks, built on the fly 16 x 4 particles at a
// time (k_real_mfma).
//
// The phase matrix never goes to memory: the lanes that feed an A operand
// build their entry from the per-particle tables e^{i m th} (two complex
// products, as in k_recip_sf_tab), held entry-major in LDS.  Operand layouts of the 16x16x4 FP32
// instruction: A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, and
// D[i][j] in element i % 4 of lane j + 16 (i / 4).
//
// A workgroup is four waves on one 64 x 64 output tile; each wave takes
// every fourth quad of the reduction index, and the four partial tiles are
// added in LDS in wave order (deterministic).
#pragma once
#include <hip/hip_runtime.h>
#include "pair_math.h"
#include "ewald_kernels.h"
#include "recip_kernels.h"
#include "real_kernels.h"

namespace shk {

typedef float mf4 __attribute__ ((ext_vector_type (4)));

__device__ __forceinline__ mf4 mfma4 (float a, float b, mf4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32 (a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int4 phase_offsets (const PhaseBasis &b, const int4 m)
{
  return make_int4 (m.x, b.K1 + 1 + b.K2 + m.y, b.K1 + 2 * b.K2 + 2 + b.K3 + m.z, 0);
}

// column-major x (n x B) -> Zt[j zs + (h D + q) B + b] (h = 1: a_j^2 z,
// unequal spheres; zs >= NH D B, a multiple of 4 for 16-byte rows); rows
// np <= j < npad are zero
template <int V, int NH>
__global__ void k_block_prep_f (const double *__restrict__ x, const double4 *__restrict__ pos,
                                int np, int npad, int B, int zs, float *__restrict__ Zt)
{
  constexpr int D = Dim<V>::D;
  const size_t n = (size_t) np * D;
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) npad * B) return;
  const int b = (int) (t % B);
  const int j = (int) (t / B);
  float z[D];
#pragma unroll
  for (int q = 0; q < D; q++) z[q] = j < np ? float (x[(size_t) b * n + (size_t) j * D + q]) : 0.f;
  if (V == 2) s_transform (z);
  float *row = Zt + (size_t) j * zs + b;
#pragma unroll
  for (int q = 0; q < D; q++) row[(size_t) q * B] = z[q];
  if (NH > 1)
    {
      const double a = j < np ? pos[j].w : 0.0;
      const float a2 = float (a * a);
#pragma unroll
      for (int q = 0; q < D; q++) row[(size_t) (D + q) * B] = a2 * z[q];
    }
}

// phase tables in two layouts: tab[j LD + e] (particle-major, for the
// structure factors) and tabT[e npad + j] (entry-major, for the
// back-projection), e^{i m th_j} of entry e as in k_phase_tab; zero for
// np <= j < npad
template <int LD>
__global__ void k_phase_tab2 (const double4 *__restrict__ pos, int np, int npad, PhaseBasis b,
                              Cplx<float> *__restrict__ tab, Cplx<float> *__restrict__ tabT)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) npad * b.W) return;
  const int e = (int) (t / npad), j = (int) (t - (size_t) e * npad);
  Cplx<float> v{0.f, 0.f};
  if (j < np)
    {
      double t1, t2, t3;
      phase_angles (b, pos[j], t1, t2, t3);
      double th;
      int m;
      if (e <= b.K1) { m = e; th = t1; }
      else if (e <= b.K1 + 2 * b.K2 + 1) { m = e - (b.K1 + 1) - b.K2; th = t2; }
      else { m = e - (b.K1 + 2 * b.K2 + 2) - b.K3; th = t3; }
      double sn, cs;
      sincos (double (m) * th, &sn, &cs);
      v = Cplx<float>{float (cs), float (sn)};
    }
  tabT[t] = v;
  tab[(size_t) j * LD + e] = v;
}

// 4 waves x 4 x 4 tiles; 64 KB of LDS for the end-of-block reduction
constexpr int MF_RED_BYTES = 4 * 64 * 64 * 4;

typedef __attribute__ ((address_space (3))) void lds_void;

// asynchronous copy global -> LDS of nv 16-byte vectors by the 256 threads
// of a block (global_load_lds_dwordx4: lane t of a wave writes 16 bytes at
// the wave's LDS base + 16 t).  src16(v) gives the global address of
// vector v; dst is 16-byte aligned.  Complete after s_waitcnt vmcnt(0) and
// a barrier.
template <typename F>
__device__ __forceinline__ void async_copy16 (void *dst, int nv, F src16)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int v0 = wave * 64; v0 < nv; v0 += 256)
    if (v0 + lane < nv)
      __builtin_amdgcn_global_load_lds (src16 (v0 + lane), (lds_void *) ((char *) dst + 16 * v0), 16,
                                        0, 0);
}

__device__ __forceinline__ void async_wait_all ()
{
  __builtin_amdgcn_s_waitcnt (0);
  __syncthreads ();
}

// LDS read of a complex entry at a byte address
__device__ __forceinline__ Cplx<float> lds_c (const unsigned char *base, int byte)
{
  return *(const Cplx<float> *) (base + byte);
}

// structure factors: part[((sj nk + k) 2 + s) cz + c], s = 0 cos, 1 sin.
// Zt rows of zs floats (16-byte rows, npad rows, zero past np), tables
// tab[j LD + e].  grid (ceil(cz/64), ceil(nk/64), splits of j, each a
// multiple of 64 particles): the column blocks of one k block run side by
// side and share the tables through L2.  The particles go through LDS in
// stages of SP (32 for LD = 96, 16 for LD = 128 and 160), double-buffered: the
// next stage's tables ([SP][LD]) and z slice ([SP][64]) are copied by
// direct global -> LDS loads while the current one is multiplied, and one
// barrier per stage hands the buffers over.  Wave w takes quads w, w + 4 of
// a stage, whose LDS addresses differ from the first by constants, so the
// inner loop spends its vector instructions on the phase products only.
// The 16 lanes of a row read consecutive entries (consecutive m3): no bank
// conflicts.
template <int LD> struct SfStage {
  static constexpr int SP = LD <= 96 ? 32 : 16;             // particles per stage
  static constexpr int TB = SP * LD * 8;                    // table bytes of a stage
  static constexpr int BYTES = TB + SP * 64 * 4;            // + the z slice
};

template <int LD>
__global__ __launch_bounds__ (256, 2) void
k_sf_mfma (const float *__restrict__ Zt, int cz, int zs, const int4 *__restrict__ km,
           const Cplx<float> *__restrict__ tab, int npad, PhaseBasis b, int nk, int jper,
           float *__restrict__ part)
{
  extern __shared__ __align__ (16) unsigned char smem_mf[];
  constexpr int SP = SfStage<LD>::SP, TB = SfStage<LD>::TB, SB = SfStage<LD>::BYTES;
  float *red = (float *) smem_mf;                          // [4][64][64], after the loop
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ir = lane & 15, kq = lane >> 4;
  const int cbase = blockIdx.x * 64, kbase = blockIdx.y * 64;
  const int j0 = blockIdx.z * jper, j1 = min (npad, j0 + jper);
  // byte addresses of the lane's table entries for its first quad
  int ax[4], ay[4], az[4];
#pragma unroll
  for (int kt = 0; kt < 4; kt++)
    {
      const int4 o = phase_offsets (b, km[min (kbase + 16 * kt + ir, nk - 1)]);
      const int rb = (4 * wave + kq) * LD;
      ax[kt] = (rb + o.x) * 8;
      ay[kt] = (rb + o.y) * 8;
      az[kt] = (rb + o.z) * 8;
    }
  const int zb = TB + ((4 * wave + kq) * 64 + ir) * 4;
  mf4 accc[4][4], accs[4][4];
#pragma unroll
  for (int kt = 0; kt < 4; kt++)
#pragma unroll
    for (int ct = 0; ct < 4; ct++)
      {
        accc[kt][ct] = mf4{0.f, 0.f, 0.f, 0.f};
        accs[kt][ct] = mf4{0.f, 0.f, 0.f, 0.f};
      }
  const int ncv = min (16, (zs - cbase) >> 2);           // 16-byte vectors per z row
  auto stage = [&] (int jb, int buf) {
    unsigned char *dst = smem_mf + buf * SB;
    const Cplx<float> *ts = tab + (size_t) jb * LD;
    async_copy16 (dst, SP * LD / 2, [&] (int v) { return (const void *) (ts + 2 * v); });
    async_copy16 (dst + TB, SP * 16, [&] (int v) -> const void * {
      const int row = v >> 4, cv = v & 15;
      return Zt + (size_t) (jb + row) * zs + cbase + 4 * min (cv, ncv - 1);
    });
  };
  stage (j0, 0);
  async_wait_all ();
  int buf = 0;
  for (int jb = j0; jb < j1; jb += SP, buf ^= 1)
    {
      if (jb + SP < j1) stage (jb + SP, buf ^ 1);
      const unsigned char *sb = smem_mf + buf * SB;
#pragma unroll
      for (int u = 0; u < SP / 16; u++)
        {
          float ac[4], as[4], bz[4];
#pragma unroll
          for (int kt = 0; kt < 4; kt++)
            {
              const Cplx<float> e = cmul (cmul (lds_c (sb, ax[kt] + 128 * LD * u),
                                                lds_c (sb, ay[kt] + 128 * LD * u)),
                                          lds_c (sb, az[kt] + 128 * LD * u));
              ac[kt] = e.re;
              as[kt] = e.im;
            }
#pragma unroll
          for (int ct = 0; ct < 4; ct++)
            bz[ct] = *(const float *) (sb + zb + 4096 * u + 64 * ct);
#pragma unroll
          for (int kt = 0; kt < 4; kt++)
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
              {
                accc[kt][ct] = mfma4 (ac[kt], bz[ct], accc[kt][ct]);
                accs[kt][ct] = mfma4 (as[kt], bz[ct], accs[kt][ct]);
              }
        }
      async_wait_all ();
    }
#pragma unroll
  for (int s = 0; s < 2; s++)
    {
#pragma unroll
      for (int kt = 0; kt < 4; kt++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
          for (int r = 0; r < 4; r++)
            red[(wave * 64 + 16 * kt + 4 * kq + r) * 64 + 16 * ct + ir]
              = s ? accs[kt][ct][r] : accc[kt][ct][r];
      __syncthreads ();
      for (int e = threadIdx.x; e < 64 * 64; e += 256)
        {
          const int k = kbase + (e >> 6), c = cbase + (e & 63);
          const float v = red[e] + red[4096 + e] + red[8192 + e] + red[12288 + e];
          if (k < nk && c < cz) part[(((size_t) blockIdx.z * nk + k) * 2 + s) * cz + c] = v;
        }
      __syncthreads ();
    }
}

// P, Q per (k, power, b) from the split structure factors (FP64 k-blocks,
// as k_recip_blocks); PQ[((k NPOW + p) 2 + s) Cp + q B + b], Cp >= D B;
// rows nk <= k < nkpad are zero
template <int V, int NPOW>
__global__ __launch_bounds__ (128) void
k_blk_mfma (const double4 *__restrict__ kv, const double *__restrict__ ks, int nk, int nkpad,
            int B, int nsplit, int Cp, const float *__restrict__ part, float *__restrict__ PQ)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  constexpr int NS = (NPOW > 1) ? 4 : 2;
  constexpr int NH = NS / 2;
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) nkpad * NPOW * B) return;
  const int k = (int) (t / ((size_t) NPOW * B));
  const int r = (int) (t - (size_t) k * NPOW * B), pw = r / B + 1, b = r - (pw - 1) * B;
  const size_t cz = (size_t) NH * D * B, C = (size_t) Cp;
  float *out = PQ + ((size_t) k * NPOW + (pw - 1)) * 2 * C + b;
  if (k >= nk)
    {
      // rows past nk: zero P, Q (the back-projection runs whole quads)
#pragma unroll
      for (int q = 0; q < D; q++) out[(size_t) q * B] = out[C + (size_t) q * B] = 0.f;
      return;
    }
  // CS: [cos z | sin z | cos a^2 z | sin a^2 z], D each
  double CS[NS * D];
#pragma unroll
  for (int q = 0; q < NS * D; q++) CS[q] = 0.0;
  for (int sp = 0; sp < nsplit; sp++)
#pragma unroll
    for (int s = 0; s < 2; s++)
      {
        const float *src = part + (((size_t) sp * nk + k) * 2 + s) * cz + b;
#pragma unroll
        for (int h = 0; h < NH; h++)
#pragma unroll
          for (int q = 0; q < D; q++)
            CS[(2 * h + s) * D + q] += double (src[(size_t) (h * D + q) * B]);
      }
  const double4 kk = kv[k];
  const double *w = ks + (size_t) k * 8;
  double P[D], Q[D];
  if (NPOW == 1) kblock_pq_mono<V> (kk, w, CS, P, Q);
  else kblock_pq_poly_rt<V> (pw, kk, w, CS, P, Q);
#pragma unroll
  for (int q = 0; q < D; q++)
    {
      out[(size_t) q * B] = float (P[q]);
      out[C + (size_t) q * B] = float (Q[q]);
    }
}

// back-projection: part[(sk np + i) C + c] = sum_k sum_p a_i^p [cos P_p +
// sin Q_p] (a_i^p = 1 for equal spheres, p = 1..NPOW otherwise).  PQ rows
// of Cp floats (Cp a multiple of 64), zero from nk to the end of the last
// split, so every quad is whole; tables transposed, tabT[e npad + j].  grid
// (ceil(C/64), ceil(np/64), splits of kper k).  The block's 64 particle
// tables sit in LDS ([entry][64]) for the whole k range; the lattice
// offsets and the B operands (P, Q) come straight from memory, one quad
// ahead.
template <int NPOW>
__global__ __launch_bounds__ (256, 2) void
k_back_mfma (const double4 *__restrict__ pos, const int4 *__restrict__ km,
             const Cplx<float> *__restrict__ tabT, int npad, PhaseBasis b,
             const float *__restrict__ PQ, int C, int Cp, int np, int nk, int kper,
             float *__restrict__ part)
{
  extern __shared__ __align__ (16) unsigned char smem_mf[];
  constexpr int PR = NPOW * 2;                            // P, Q rows per k
  float *red = (float *) smem_mf;                         // [4][64][64], after the loop
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ir = lane & 15, kq = lane >> 4;
  const int cbase = blockIdx.x * 64, ibase = blockIdx.y * 64;
  const int k0 = blockIdx.z * kper, k1 = k0 + kper;
  async_copy16 (smem_mf, b.W * 32, [&] (int v) {
    return (const void *) (tabT + (size_t) (v >> 5) * npad + ibase + 2 * (v & 31));
  });
  float apow[4][NPOW];
#pragma unroll
  for (int it = 0; it < 4; it++)
    {
      const float a = NPOW > 1 ? float (pos[min (ibase + 16 * it + ir, np - 1)].w) : 1.f;
      apow[it][0] = a;
#pragma unroll
      for (int p = 1; p < NPOW; p++) apow[it][p] = apow[it][p - 1] * a;
    }
  const float *pq = PQ + cbase + ir;
  mf4 acc[4][4];
#pragma unroll
  for (int it = 0; it < 4; it++)
#pragma unroll
    for (int ct = 0; ct < 4; ct++) acc[it][ct] = mf4{0.f, 0.f, 0.f, 0.f};
  async_wait_all ();
  const int nq = kper / 4;
  // operands of quad q: table byte offsets of k = k0 + 4 q + kq and its
  // P, Q rows
  auto load = [&] (int q, int4 &oo, float (&v)[NPOW][2][4]) {
    const int k = k0 + 4 * min (q, nq - 1) + kq;
    const int4 m = phase_offsets (b, km[min (k, nk - 1)]);
    oo = make_int4 (m.x * 512 + ir * 8, m.y * 512 + ir * 8, m.z * 512 + ir * 8, 0);
    const float *r = pq + (size_t) k * PR * Cp;
#pragma unroll
    for (int p = 0; p < NPOW; p++)
#pragma unroll
      for (int s = 0; s < 2; s++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++) v[p][s][ct] = r[(2 * p + s) * Cp + 16 * ct];
  };
  int4 o;
  float bpq[NPOW][2][4];
  load (wave, o, bpq);
  for (int qd = wave; qd < nq; qd += 4)
    {
      int4 on;
      float bn[NPOW][2][4];
      load (qd + 4, on, bn);
      float ac[4], as[4];
#pragma unroll
      for (int it = 0; it < 4; it++)
        {
          const Cplx<float> e = cmul (cmul (lds_c (smem_mf, o.x + 128 * it),
                                            lds_c (smem_mf, o.y + 128 * it)),
                                      lds_c (smem_mf, o.z + 128 * it));
          ac[it] = e.re;
          as[it] = e.im;
        }
#pragma unroll
      for (int p = 0; p < NPOW; p++)
#pragma unroll
        for (int it = 0; it < 4; it++)
          {
            const float cc = NPOW > 1 ? apow[it][p] * ac[it] : ac[it];
            const float ss = NPOW > 1 ? apow[it][p] * as[it] : as[it];
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
              {
                acc[it][ct] = mfma4 (cc, bpq[p][0][ct], acc[it][ct]);
                acc[it][ct] = mfma4 (ss, bpq[p][1][ct], acc[it][ct]);
              }
          }
      o = on;
#pragma unroll
      for (int p = 0; p < NPOW; p++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
          for (int ct = 0; ct < 4; ct++) bpq[p][s][ct] = bn[p][s][ct];
    }
  __syncthreads ();
#pragma unroll
  for (int it = 0; it < 4; it++)
#pragma unroll
    for (int ct = 0; ct < 4; ct++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        red[(wave * 64 + 16 * it + 4 * kq + r) * 64 + 16 * ct + ir] = acc[it][ct][r];
  __syncthreads ();
  for (int e = threadIdx.x; e < 64 * 64; e += 256)
    {
      const int i = ibase + (e >> 6), c = cbase + (e & 63);
      const float v = red[e] + red[4096 + e] + red[8192 + e] + red[12288 + e];
      if (i < np && c < C) part[((size_t) blockIdx.z * np + i) * C + c] = v;
    }
}

// real-space pair sum for a block of columns.  One wave per (16 target
// particles i, range of source particles j, 16 CT columns).  Per chunk of 4
// source particles:
//   phase A: lane = pair (i_l = lane % 16, j_l = lane / 16) sums the
//     moments of its pair over the images inside r_max -- each lane first
//     marks its images among the next 64 candidates in a bit mask, then
//     evaluates its own marked images, so a lane idles only while another
//     still has images left;
//   phase B: Y (16 D x 16 CT) += A (16 D x 4 D) Z (4 D x 16 CT) on the
//     matrix cores, with row R = D i_l + r and reduction index K = 4 q + j_l,
//     so quad u of the reduction is component q = u and the lane's k index
//     is its source particle.  The block matrix goes through LDS in rounds
//     of QR components (every lane writes its entries of those columns),
//     which keeps the LDS of a wave near 15 KB.
// partial[((split np + i) D + q) B + b]; XT rows of zs values.
constexpr int REAL_QR = 3;

template <int V, int CT, int RND> struct RealRounds {
  static constexpr int D = Dim<V>::D, Q0 = RND * REAL_QR;
  static constexpr int NQ = (D - Q0) < REAL_QR ? (D - Q0) : REAL_QR;
  static constexpr int LDA = 4 * REAL_QR + 1;
  static constexpr int NCOL = 16 * CT;
  __device__ __forceinline__ static void run (const float *M, float *blk, const float *zs_l,
                                              int ir, int kq, mf4 (&acc)[Dim<V>::D][CT])
  {
    float *dst = blk + ir * D * LDA + kq;
    auto put = [&] (int r, int q, float v) { dst[r * LDA + 4 * (q - Q0)] = v; };
    BlockEach<V, Q0, NQ, D * NQ>::run (M, false, put);
    __syncthreads ();
#pragma unroll
    for (int u = 0; u < NQ; u++)
      {
        float bz[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ct++) bz[ct] = zs_l[(4 * (Q0 + u) + kq) * NCOL + 16 * ct + ir];
#pragma unroll
        for (int t = 0; t < D; t++)
          {
            const float a = blk[(16 * t + ir) * LDA + 4 * u + kq];
#pragma unroll
            for (int ct = 0; ct < CT; ct++) acc[t][ct] = mfma4 (a, bz[ct], acc[t][ct]);
          }
      }
    __syncthreads ();
    if constexpr (Q0 + NQ < D) RealRounds<V, CT, RND + 1>::run (M, blk, zs_l, ir, kq, acc);
  }
};

template <int V, int CT>
__global__ __launch_bounds__ (64) __attribute__ ((amdgpu_waves_per_eu (2))) void
k_real_mfma (const double4 *__restrict__ pos, const float *__restrict__ XT, int B, int zs,
             const float4 *__restrict__ cand, const int4 *__restrict__ candm,
             const unsigned *__restrict__ tbits, RealParams p, int jper,
             float *__restrict__ partial)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  constexpr int LDA = 4 * REAL_QR + 1;
  constexpr int NCOL = 16 * CT;                 // columns of the z slice
  constexpr int ZE = 4 * D * NCOL;              // z slice entries
  __shared__ float blk[16 * D * LDA];
  __shared__ float zs_l[ZE];                    // [K = 4 q + j_l][c]
  __shared__ unsigned bits_l[64];
  const int lane = threadIdx.x;
  const int ir = lane & 15, kq = lane >> 4;
  const int ibase = blockIdx.x * 16;
  const int j0 = blockIdx.y * jper, j1 = min (p.np, j0 + jper);
  const int cbase = blockIdx.z * NCOL;
  const bool bits_in_lds = p.tbits_words <= 64;
  if (bits_in_lds && lane < p.tbits_words) bits_l[lane] = tbits[lane];
  __syncthreads ();
  const unsigned *bits = bits_in_lds ? bits_l : tbits;
  const int ia = ibase + ir;
  const bool ivalid = ia < p.np;
  const double4 pi = pos[ivalid ? ia : 0];
  const float ai = float (pi.w), xi = float (p.xi), rmax2 = float (p.rmax2);
  mf4 acc[D][CT];
#pragma unroll
  for (int t = 0; t < D; t++)
#pragma unroll
    for (int ct = 0; ct < CT; ct++) acc[t][ct] = mf4{0.f, 0.f, 0.f, 0.f};

  for (int jb = j0; jb < j1; jb += 4)
    {
      // z slice of the chunk (read after the first barrier of phase B)
      for (int e = lane; e < ZE; e += 64)
        {
          const int K = e / NCOL, c = e - K * NCOL;
          const int q = K >> 2, jz = jb + (K & 3), b = cbase + c;
          const bool ok = jz < j1 && b < B;
          const float v = XT[(size_t) (ok ? jz : 0) * zs + (ok ? q * B + b : 0)];
          zs_l[e] = ok ? v : 0.f;
        }
      // ---- phase A
      const int j = jb + kq;
      const bool valid = ivalid && j < j1;
      const double4 pj = pos[valid ? j : 0];
      double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
      const double n3 = rint (dz / p.lz);
      dx -= n3 * p.s3x; dz -= n3 * p.lz;
      const double n2 = rint (dy / p.ly);
      dx -= n2 * p.s2x; dy -= n2 * p.ly;
      const double n1 = rint (dx / p.lx);
      dx -= n1 * p.lx;
      const int w1 = int (n1), w2 = int (n2), w3 = int (n3);
      const float mdx = float (dx), mdy = float (dy), mdz = float (dz);
      const float aj = float (pj.w);
      const float rmn = (ai + aj) * float (p.rmin);
      float M[NM];
#pragma unroll
      for (int t = 0; t < NM; t++) M[t] = 0.f;
      for (int c0 = 0; c0 < p.ncand; c0 += 64)
        {
          const int cn = min (64, p.ncand - c0);
          unsigned long long mask = 0ull;
#pragma unroll 8
          for (int c = 0; c < cn; c++)
            {
              const float4 v = cand[c0 + c];
              const float ex = mdx + v.x, ey = mdy + v.y, ez = mdz + v.z;
              const float rr = ex * ex + ey * ey + ez * ez;
              const bool hit = valid && rr != 0.f && rr <= rmax2;
              mask |= (unsigned long long) hit << c;
            }
          // the reference-table test only for the images inside r_max
          while (mask)
            {
              const int c = __builtin_ctzll (mask);
              mask &= mask - 1ull;
              const int4 mm = candm[c0 + c];
              if (!in_table (p, bits, mm.x - w1, mm.y - w2, mm.z - w3)) continue;
              const float4 v = cand[c0 + c];
              const float ex = mdx + v.x, ey = mdy + v.y, ez = mdz + v.z;
              const float rr = ex * ex + ey * ey + ez * ez;
              float iu = m_rsqrt (rr);
              float r = rr * iu;
              if (r < rmn)
                {
                  r = rmn;
                  iu = 1.f / rmn;
                }
              float sc[11];
              if (p.poly) scalars_real_poly_u<V, float> (xi, r, iu, ai, aj, sc);
              else scalars_real_u<V, float> (xi, r, iu, sc);
              moments_add<V, float> (M, ex * iu, ey * iu, ez * iu, sc);
            }
        }
      // ---- phase B
      RealRounds<V, CT, 0>::run (M, blk, zs_l, ir, kq, acc);
    }
  // rows R = 16 t + 4 (lane / 16) + e of the (16 D)-row tile
#pragma unroll
  for (int t = 0; t < D; t++)
#pragma unroll
    for (int e = 0; e < 4; e++)
      {
        const int R = 16 * t + 4 * kq + e, il = R / D, r = R - il * D;
        const int i = ibase + il;
        if (i < p.np)
#pragma unroll
          for (int ct = 0; ct < CT; ct++)
            {
              const int b = cbase + 16 * ct + ir;
              if (b < B) partial[(((size_t) blockIdx.y * p.np + i) * D + r) * B + b] = acc[t][ct][e];
            }
      }
}

// v from the lane given by the quad permutation CTRL (DPP quad_perm)
template <int CTRL>
__device__ __forceinline__ float dpp_quad (float v)
{
  return __builtin_bit_cast (float, __builtin_amdgcn_update_dpp (0, __builtin_bit_cast (int, v), CTRL,
                                                                 0xF, 0xF, true));
}

template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_quad64 (unsigned long long v)
{
  const unsigned lo = (unsigned) v, hi = (unsigned) (v >> 32);
  const unsigned plo = (unsigned) __builtin_amdgcn_update_dpp (0, (int) lo, CTRL, 0xF, 0xF, true);
  const unsigned phi = (unsigned) __builtin_amdgcn_update_dpp (0, (int) hi, CTRL, 0xF, 0xF, true);
  return ((unsigned long long) phi << 32) | plo;
}

// candidate images held in LDS by k_real_mfma_pc
constexpr int REAL_PC_CAND = 256;

// real-space pair sum for 32 NC columns by producer and consumer waves: a
// workgroup of NP producer and NC consumer waves on 16 target particles i, a
// range of source particles j and 32 columns per consumer wave.  Per chunk of
// 4 source particles (64 pairs, pair P = i_l + 16 j_l):
//   producers: the NP adjacent lanes l of producer wave w with l / NP = P -
//     w 64/NP share pair P: they test every NP-th candidate image each, OR
//     the hit masks, and lane phase ph = l % NP evaluates the images of rank
//     ph, ph + NP, ... (so the images of a pair are spread evenly over its
//     lanes), sums their moments; the NP partial sums are added by quad
//     permutations (the same sum in every lane of the group) and written to
//     LDS as mom[m][P];
//   consumers: contract the 16 D x 4 D block matrix of the chunk with their
//     32 columns on the matrix cores, rows component-major (row tile r =
//     component, row i_l), reduction index 4 q + j_l: lane i_l + 16 j_l
//     supplies the entry (r, q) of its own pair from the moments in LDS.
// Producers fill chunk c while consumers contract chunk c - 1 (two moment
// buffers, one barrier per chunk), so the vector work of the image sums and
// the matrix-core work run side by side on every SIMD, and the image sums are
// evaluated once for all 32 NC columns.  partial[((split np + i) D + r) B +
// b]; XT rows of zs values.
template <int V, int NP, int NC>
__global__ __launch_bounds__ (64 * (NP + NC)) __attribute__ ((amdgpu_waves_per_eu (2))) void
k_real_mfma_pc (const double4 *__restrict__ pos, const float *__restrict__ XT, int B, int zs,
                const float4 *__restrict__ cand, const int4 *__restrict__ candm,
                const unsigned *__restrict__ tbits, RealParams p, int jper,
                float *__restrict__ partial)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  constexpr int PW = 64 / NP;                   // pairs per producer wave
  static_assert (NP == 1 || NP == 2 || NP == 4, "phases are lanes of a quad");
  __shared__ float mom[2][NM * 64];
  __shared__ unsigned bits_l[64];
  __shared__ float4 cand_l[REAL_PC_CAND];
  __shared__ int4 candm_l[REAL_PC_CAND];
  // reference-table masks of the candidates, per 64-candidate chunk and
  // wrap offset (w1, w2, w3) in {-1, 0, 1}^3
  __shared__ unsigned long long tmask_l[REAL_PC_CAND / 64][27];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ibase = blockIdx.x * 16;
  const int j0 = blockIdx.y * jper, j1 = min (p.np, j0 + jper);
  const int nch = (j1 - j0 + 3) >> 2;
  const bool bits_in_lds = p.tbits_words <= 64;
  if (bits_in_lds && threadIdx.x < p.tbits_words) bits_l[threadIdx.x] = tbits[threadIdx.x];
  // the candidate images in LDS when they fit (the host keeps ncand within
  // REAL_PC_CAND for this kernel)
  for (int c = threadIdx.x; c < p.ncand; c += 64 * (NP + NC))
    {
      cand_l[c] = cand[c];
      candm_l[c] = candm[c];
    }
  for (int e = threadIdx.x; e < ((p.ncand + 63) >> 6) * 27; e += 64 * (NP + NC))
    {
      const int ch = e / 27, w = e - 27 * ch;
      const int a = w / 9 - 1, b = (w / 3) % 3 - 1, c3 = w % 3 - 1;
      const int cn = min (64, p.ncand - 64 * ch);
      unsigned long long m = 0ull;
      for (int c = 0; c < cn; c++)
        {
          const int4 mm = candm[64 * ch + c];
          if (in_table (p, tbits, mm.x - a, mm.y - b, mm.z - c3)) m |= 1ull << c;
        }
      tmask_l[ch][w] = m;
    }
  __syncthreads ();
  if (wave < NP)
    {
      // ---- producer
      const unsigned *bits = bits_in_lds ? bits_l : tbits;
      const int P = wave * PW + lane / NP, ph = lane % NP;
      const int il = P & 15, jl = P >> 4;
      const int ia = ibase + il;
      const bool ivalid = ia < p.np;
      const double4 pi = pos[ivalid ? ia : 0];
      const float ai = float (pi.w), xi = float (p.xi), rmax2 = float (p.rmax2);
      // the source particle of the next chunk, loaded one chunk ahead
      double4 pjn = pos[ivalid && j0 + jl < j1 ? j0 + jl : 0];
      for (int it = 0; it <= nch; it++)
        {
          if (it < nch)
            {
              const int j = j0 + 4 * it + jl;
              const bool valid = ivalid && j < j1;
              const double4 pj = pjn;
              {
                const int jn = j + 4;
                pjn = pos[ivalid && jn < j1 ? jn : 0];
              }
              double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
              const double n3 = rint (dz / p.lz);
              dx -= n3 * p.s3x; dz -= n3 * p.lz;
              const double n2 = rint (dy / p.ly);
              dx -= n2 * p.s2x; dy -= n2 * p.ly;
              const double n1 = rint (dx / p.lx);
              dx -= n1 * p.lx;
              const int w1 = int (n1), w2 = int (n2), w3 = int (n3);
              const int wtab = (abs (w1) <= 1 && abs (w2) <= 1 && abs (w3) <= 1)
                ? (w1 + 1) * 9 + (w2 + 1) * 3 + (w3 + 1) : -1;
              const float mdx = float (dx), mdy = float (dy), mdz = float (dz);
              const float aj = float (pj.w);
              const float rmn = (ai + aj) * float (p.rmin);
              float M[NM];
#pragma unroll
              for (int t = 0; t < NM; t++) M[t] = 0.f;
              for (int c0 = 0; c0 < p.ncand; c0 += 64)
                {
                  // the images of the pair among the next 64 candidates
                  // (inside r_max and in the reference table): each lane
                  // of the quad tests every NP-th, the quad ORs the masks
                  const int cn = min (64, p.ncand - c0);
                  unsigned long long mask = 0ull;
#pragma unroll 4
                  for (int c = ph; c < cn; c += NP)
                    {
                      const float4 v = cand_l[c0 + c];
                      const float ex = mdx + v.x, ey = mdy + v.y, ez = mdz + v.z;
                      const float rr = ex * ex + ey * ey + ez * ez;
                      const bool hit = valid && rr != 0.f && rr <= rmax2;
                      mask |= (unsigned long long) hit << c;
                    }
                  if (wtab >= 0)
                    mask &= tmask_l[c0 >> 6][wtab];
                  else
                    {
                      // wrap offsets beyond one cell: the table test image by image
                      unsigned long long m2 = mask;
                      while (m2)
                        {
                          const int c = __builtin_ctzll (m2);
                          m2 &= m2 - 1ull;
                          const int4 mm = candm_l[c0 + c];
                          if (!in_table (p, bits, mm.x - w1, mm.y - w2, mm.z - w3))
                            mask &= ~(1ull << c);
                        }
                    }
                  if (NP >= 2) mask |= dpp_quad64<0xB1> (mask);
                  if (NP >= 4) mask |= dpp_quad64<0x4E> (mask);
                  // this lane's share: the images of rank ph, ph + NP, ...
                  unsigned long long mine = 0ull;
                  for (int k = 0; mask; k++)
                    {
                      const unsigned long long b = mask & (~mask + 1ull);
                      if ((k & (NP - 1)) == ph) mine |= b;
                      mask ^= b;
                    }
                  while (mine)
                    {
                      const int c = __builtin_ctzll (mine);
                      mine &= mine - 1ull;
                      const float4 v = cand_l[c0 + c];
                      const float ex = mdx + v.x, ey = mdy + v.y, ez = mdz + v.z;
                      const float rr = ex * ex + ey * ey + ez * ez;
                      float iu = m_rsqrt (rr);
                      float r = rr * iu;
                      if (r < rmn)
                        {
                          r = rmn;
                          iu = 1.f / rmn;
                        }
                      float sc[11];
                      if (p.poly) scalars_real_poly_u<V, float> (xi, r, iu, ai, aj, sc);
                      else scalars_real_u<V, float> (xi, r, iu, sc);
                      moments_add<V, float> (M, ex * iu, ey * iu, ez * iu, sc);
                    }
                }
              // the NP phases of a pair are adjacent lanes: pairwise sums
              // by quad permutations (commutative, so every lane of the
              // group holds the same value)
#pragma unroll
              for (int t = 0; t < NM; t++)
                {
                  float v = M[t];
                  if (NP >= 2) v += dpp_quad<0xB1> (v);      // lane ^ 1
                  if (NP >= 4) v += dpp_quad<0x4E> (v);      // lane ^ 2
                  M[t] = v;
                }
              if (ph == 0)
                {
                  float *mw = mom[it & 1];
#pragma unroll
                  for (int t = 0; t < NM; t++) mw[t * 64 + P] = M[t];
                }
            }
          __syncthreads ();
        }
      return;
    }
  // ---- consumer: lane ir + 16 kq holds pair (ir, kq) of a chunk
  const int ir = lane & 15, kq = lane >> 4;
  const int cbase = blockIdx.z * 32 * NC + 32 * (wave - NP);
  mf4 acc[D][2];
#pragma unroll
  for (int t = 0; t < D; t++)
#pragma unroll
    for (int ct = 0; ct < 2; ct++) acc[t][ct] = mf4{0.f, 0.f, 0.f, 0.f};
  // z[j][q][cbase + 16 ct + ir] of the next chunk, loaded one chunk ahead
  float bzn[D][2];
  auto load_z = [&] (int chunk) {
    const int jz = j0 + 4 * chunk + kq;
#pragma unroll
    for (int ct = 0; ct < 2; ct++)
      {
        const int b = cbase + 16 * ct + ir;
        const bool ok = jz < j1 && b < B;
        const float *src = XT + (size_t) (ok ? jz : 0) * zs + (ok ? b : 0);
#pragma unroll
        for (int q = 0; q < D; q++) bzn[q][ct] = ok ? src[(size_t) q * B] : 0.f;
      }
  };
  for (int it = 0; it <= nch; it++)
    {
      float bz[D][2];
#pragma unroll
      for (int q = 0; q < D; q++)
#pragma unroll
        for (int ct = 0; ct < 2; ct++) bz[q][ct] = bzn[q][ct];
      if (it < nch) load_z (it);
      if (it >= 1)
        {
          // moment t of this lane's pair at mr[64 t], read from LDS as needed
          struct LdsM {
            const float *b;
            __device__ float operator[] (int t) const { return b[64 * t]; }
          } Ml{mom[(it - 1) & 1] + lane};
          // q outer: D independent accumulator pairs between dependent issues
#pragma unroll
          for (int q = 0; q < D; q++)
#pragma unroll
            for (int r = 0; r < D; r++)
              {
                const float a = block_entry<V, float> (Ml, r, q, false);
                acc[r][0] = mfma4 (a, bz[q][0], acc[r][0]);
                acc[r][1] = mfma4 (a, bz[q][1], acc[r][1]);
              }
        }
      __syncthreads ();
    }
  // row i_l = 4 (lane / 16) + e of row tile r, column 16 ct + lane % 16
#pragma unroll
  for (int r = 0; r < D; r++)
#pragma unroll
    for (int e = 0; e < 4; e++)
      {
        const int i = ibase + 4 * kq + e;
        if (i < p.np)
#pragma unroll
          for (int ct = 0; ct < 2; ct++)
            {
              const int b = cbase + 16 * ct + ir;
              if (b < B) partial[(((size_t) blockIdx.y * p.np + i) * D + r) * B + b] = acc[r][ct][e];
            }
      }
}

// ---------------------------------------------------------------------------
// Equal spheres with a table symmetric under m -> -m: the image sums of the
// pair (j, i) are those of (i, j) with e -> -e, so the moments of odd order
// in e change sign and the rest are equal.  k_real_moments evaluates each
// pair i <= j once (one lane per pair, every image) and writes both orders
// into mt[m][i / 16][j][i % 16] (npad x npad per moment, zero where i or j
// >= np), and k_real_mfma_mom contracts the blocks with the columns on the
// matrix cores, reading the moments of its pairs from mt.
// ---------------------------------------------------------------------------

// sign of moment t of the pair seen the other way round
template <int V> __device__ __forceinline__ float moment_parity (int t)
{
  return ((t >= M_SB && t < M_SC) || (t >= M_SG1 && t < M_SH)) ? -1.f : 1.f;
}

// the workgroup's copy of the lattice candidates v with |v|^2 / 2 in w,
// padded to a multiple of 32 with vectors beyond reach, and, per chunk of
// 64 candidates and wrap offset (w1, w2, w3) in {-1, 0, 1}^3, the bit mask
// of those whose reference image m - w lies in the table (every thread
// calls it; a barrier before use)
template <typename T4>
__device__ __forceinline__ void
stage_candidates (const RealParams &p, const T4 *__restrict__ cand, const int4 *__restrict__ candm,
                  const unsigned *__restrict__ tbits, T4 *cand_l,
                  unsigned long long (*tmask_l)[27])
{
  const int n32 = (p.ncand + 31) & ~31;
  for (int c = threadIdx.x; c < n32; c += blockDim.x)
    {
      T4 v;
      if (c < p.ncand)
        {
          v = cand[c];
          v.w = 0.5 * ((double) v.x * v.x + (double) v.y * v.y + (double) v.z * v.z);
        }
      else
        v.x = v.w = 1e30, v.y = v.z = 0.0;
      cand_l[c] = v;
    }
  for (int e = threadIdx.x; e < ((p.ncand + 63) >> 6) * 27; e += blockDim.x)
    {
      const int ch = e / 27, w = e - 27 * ch;
      const int a = w / 9 - 1, b = (w / 3) % 3 - 1, c3 = w % 3 - 1;
      const int cn = min (64, p.ncand - 64 * ch);
      unsigned long long m = 0ull;
      for (int c = 0; c < cn; c++)
        {
          const int4 mm = candm[64 * ch + c];
          if (in_table (p, tbits, mm.x - a, mm.y - b, mm.z - c3)) m |= 1ull << c;
        }
      tmask_l[ch][w] = m;
    }
}

// the separation of the equal-sphere pair (pi, pj) wrapped into the primary
// cell, as the reference's minimum image, and the index of its wrap offset
// among the 27 reference-table masks of stage_candidates (-1 beyond them)
struct PairSep {
  double dx, dy, dz;
  int w1, w2, w3, wtab;
};

__device__ __forceinline__ PairSep pair_sep (const double4 pi, const double4 pj,
                                             const RealParams &p)
{
  PairSep s;
  double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
  const double n3 = rint (dz / p.lz);
  dx -= n3 * p.s3x; dz -= n3 * p.lz;
  const double n2 = rint (dy / p.ly);
  dx -= n2 * p.s2x; dy -= n2 * p.ly;
  const double n1 = rint (dx / p.lx);
  dx -= n1 * p.lx;
  s.dx = dx; s.dy = dy; s.dz = dz;
  s.w1 = int (n1); s.w2 = int (n2); s.w3 = int (n3);
  s.wtab = (abs (s.w1) <= 1 && abs (s.w2) <= 1 && abs (s.w3) <= 1)
    ? (s.w1 + 1) * 9 + (s.w2 + 1) * 3 + (s.w3 + 1) : -1;
  return s;
}

// 0 < t <= h for t, h >= 0 or t negative: one unsigned compare of the bit
// patterns (t = +0 wraps to the largest pattern, a negative t has its sign
// bit set; both compare false)
__device__ __forceinline__ bool in_open_closed (float t, float h)
{
  return __float_as_uint (t) - 1u < __float_as_uint (h);
}
__device__ __forceinline__ bool in_open_closed (double t, double h)
{
  return (unsigned long long) __double_as_longlong (t) - 1ull
         < (unsigned long long) __double_as_longlong (h);
}

// bit c set: candidate c0 + c is an image of the pair inside r_max (r > 0)
// that the reference's table holds.  With d the wrapped separation,
// |d + v|^2 / 2 = d.v + |v|^2 / 2 + |d|^2 / 2 (three FMAs and an add per
// candidate); for v = 0 it is |d|^2 / 2 itself, and for v != 0 the
// wrapped d keeps |d + v| >= L / 2, so only images at r_max can round to
// the other side.  32 candidates per unrolled group (the staged list is
// padded to a multiple of 32).
template <typename T>
__device__ __forceinline__ unsigned long long
pair_chunk_mask (const PairSep &s, T mdx, T mdy, T mdz, int c0, const RealParams &p,
                 const typename Vec4<T>::type *cand_l, const int4 *__restrict__ candm,
                 const unsigned *__restrict__ tbits, const unsigned long long (*tmask_l)[27])
{
  using T4 = typename Vec4<T>::type;
  const T h2 = T (0.5 * p.rmax2);
  const T d2h = T (0.5) * (mdx * mdx + mdy * mdy + mdz * mdz);
  const int cn = min (64, p.ncand - c0);
  unsigned long long mask = 0ull;
  for (int g = 0; g < cn; g += 32)
    {
      unsigned bits = 0u;
#pragma unroll
      for (int b = 0; b < 32; b++)
        {
          const T4 v = cand_l[c0 + g + b];
          const T t = fma (mdx, v.x, fma (mdy, v.y, fma (mdz, v.z, v.w + d2h)));
          bits |= in_open_closed (t, h2) ? 1u << b : 0u;
        }
      mask |= (unsigned long long) bits << g;
    }
  if (s.wtab >= 0)
    mask &= tmask_l[c0 >> 6][s.wtab];
  else
    {
      unsigned long long m2 = mask;
      while (m2)
        {
          const int c = __builtin_ctzll (m2);
          m2 &= m2 - 1ull;
          const int4 mm = candm[c0 + c];
          if (!in_table (p, tbits, mm.x - s.w1, mm.y - s.w2, mm.z - s.w3)) mask &= ~(1ull << c);
        }
    }
  return mask;
}

// M += the moments of one image (separation (mdx, mdy, mdz) + v, clamped
// at rmn like the reference's minimum distance)
template <int V, typename T>
__device__ __forceinline__ void
image_moments_add (T mdx, T mdy, T mdz, const typename Vec4<T>::type v, T xi, T rmn, T *M)
{
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
  scalars_real_u<V, T> (xi, r, iu, sc);
  moments_add<V, T> (M, ex * iu, ey * iu, ez * iu, sc);
}

constexpr int REAL_NCH = REAL_PC_CAND / 64;      // candidate chunks of 64

// the image masks of the pair, one per chunk of 64 candidates (zero past
// the last chunk)
template <typename T>
__device__ __forceinline__ void
pair_masks (const PairSep &s, const RealParams &p, const typename Vec4<T>::type *cand_l,
            const int4 *__restrict__ candm, const unsigned *__restrict__ tbits,
            const unsigned long long (*tmask_l)[27], unsigned long long (&mk)[REAL_NCH])
{
#pragma unroll
  for (int ch = 0; ch < REAL_NCH; ch++)
    mk[ch] = 64 * ch < p.ncand
      ? pair_chunk_mask<T> (s, T (s.dx), T (s.dy), T (s.dz), 64 * ch, p, cand_l, candm, tbits,
                            tmask_l)
      : 0ull;
}

// M += the moments of the images selected by the masks, in candidate order,
// one loop over the set bits of all chunks
template <int V, typename T>
__device__ __forceinline__ void
pair_images_add (const PairSep &s, T xi, T rmn, const typename Vec4<T>::type *cand_l,
                 const unsigned long long (&mk)[REAL_NCH], T *M)
{
  static_assert (REAL_NCH == 4, "four candidate chunks");
  const T mdx = T (s.dx), mdy = T (s.dy), mdz = T (s.dz);
  unsigned long long m0 = mk[0], m1 = mk[1], m2 = mk[2], m3 = mk[3];
  int base = 0;
  while (true)
    {
      while (m0 == 0ull && (m1 | m2 | m3) != 0ull)
        {
          m0 = m1; m1 = m2; m2 = m3; m3 = 0ull;
          base += 64;
        }
      if (m0 == 0ull) break;
      const int c = base + __builtin_ctzll (m0);
      m0 &= m0 - 1ull;
      image_moments_add<V, T> (mdx, mdy, mdz, cand_l[c], xi, rmn, M);
    }
}

// the 256 threads of a workgroup rank their values cnt (0 <= cnt <= 256)
// by a counting sort in LDS; returns the thread index holding the value of
// rank threadIdx.x (ties in an arbitrary order).  bin_l: REAL_PC_CAND + 2
// ints zeroed before the call (and a barrier); perm_l: 256 entries.  Every
// thread calls it.
__device__ __forceinline__ int rank_by_count (int cnt, int *bin_l, unsigned short *perm_l)
{
  const int slot = atomicAdd (&bin_l[cnt + 1], 1);
  __syncthreads ();
  // prefix sums over the counts (wave 0: PB bins per lane)
  if (threadIdx.x < 64)
    {
      constexpr int PB = (REAL_PC_CAND + 2 + 63) / 64;
      const int b0 = threadIdx.x * PB;
      int v[PB], sum = 0;
#pragma unroll
      for (int u = 0; u < PB; u++)
        {
          v[u] = (b0 + u < REAL_PC_CAND + 2) ? bin_l[b0 + u] : 0;
          sum += v[u];
        }
      int incl = sum;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1)
        {
          const int o = __shfl_up (incl, off, 64);
          if ((int) threadIdx.x >= off) incl += o;
        }
      int run = incl - sum;
#pragma unroll
      for (int u = 0; u < PB; u++)
        {
          run += v[u];
          if (b0 + u < REAL_PC_CAND + 2) bin_l[b0 + u] = run;
        }
    }
  __syncthreads ();
  // bin_l[c] now counts the values below c
  perm_l[bin_l[cnt] + slot] = (unsigned short) threadIdx.x;
  __syncthreads ();
  return perm_l[threadIdx.x];
}

// workgroup: the 16 i x 16 j pairs of one tile (blockIdx.x <= blockIdx.y;
// the others return at once), in three phases.  (1) thread q = 16 jl + il
// finds the images of the pair (16 bx + il, 16 by + jl) (one mask per 64
// candidates, kept in LDS) and counts them; (2) the pairs are ranked by
// count (a counting sort in LDS), so that the 64 lanes of a wave take pairs
// of nearly equal image counts; (3) thread k evaluates the moments of the
// k-th pair in that order, one loop over the set bits of all its masks.
// Each pair's sum runs over its images in candidate order whichever lane
// takes it, so the results do not depend on the ranking.  The moments go
// out through LDS in rounds of MR: (i, j) to mt[m][i / 16][j][i % 16] and
// (j, i), odd moments negated, to mt[m][j / 16][i][j % 16], both as 256
// consecutive elements per tile and moment.
template <int V, typename T>
__global__ __launch_bounds__ (256) __attribute__ ((amdgpu_waves_per_eu (sizeof (T) == 4 ? 3 : 2))) void
k_real_moments (const double4 *__restrict__ pos, const typename Vec4<T>::type *__restrict__ cand,
                const int4 *__restrict__ candm, const unsigned *__restrict__ tbits, RealParams p,
                int npad, T *__restrict__ mt)
{
  constexpr int NM = Dim<V>::NMOM;
  constexpr int NCH = REAL_PC_CAND / 64;
  using T4 = typename Vec4<T>::type;
  if (blockIdx.y < blockIdx.x) return;
  __shared__ T4 cand_l[REAL_PC_CAND];
  __shared__ unsigned long long tmask_l[NCH][27];
  __shared__ unsigned long long mk_l[NCH][256];
  __shared__ int bin_l[REAL_PC_CAND + 2];
  __shared__ unsigned short perm_l[256];
  constexpr int MR = 17;
  __shared__ T tr_l[MR][16 * 17];
  stage_candidates<T4> (p, cand, candm, tbits, cand_l, tmask_l);
  for (int e = threadIdx.x; e < REAL_PC_CAND + 2; e += 256) bin_l[e] = 0;
  __syncthreads ();
  const int nch = (p.ncand + 63) >> 6;
  const int bx = blockIdx.x, by = blockIdx.y;
  // (1) masks and image count of pair q = threadIdx.x
  int cnt = 0;
  {
    const int q = threadIdx.x, i = bx * 16 + (q & 15), j = by * 16 + (q >> 4);
    const bool valid = i <= j && j < p.np;
    PairSep s;
    if (valid) s = pair_sep (pos[i], pos[j], p);
    for (int ch = 0; ch < nch; ch++)
      {
        unsigned long long m = 0ull;
        if (valid)
          m = pair_chunk_mask<T> (s, T (s.dx), T (s.dy), T (s.dz), 64 * ch, p, cand_l, candm,
                                  tbits, tmask_l);
        mk_l[ch][q] = m;
        cnt += __builtin_popcountll (m);
      }
  }
  // (2) rank by count
  const int q = rank_by_count (cnt, bin_l, perm_l);
  // (3) moments of the pair ranked threadIdx.x
  const int il = q & 15, jl = q >> 4;
  const int i = bx * 16 + il, j = by * 16 + jl;
  T M[NM];
#pragma unroll
  for (int t = 0; t < NM; t++) M[t] = T (0);
  if (i <= j && j < p.np)
    {
      const double4 pi = pos[i], pj = pos[j];
      const PairSep s = pair_sep (pi, pj, p);
      const T mdx = T (s.dx), mdy = T (s.dy), mdz = T (s.dz);
      const T xi = T (p.xi), rmn = (T (pi.w) + T (pj.w)) * T (p.rmin);
      int ch = 0;
      unsigned long long m = nch > 0 ? mk_l[0][q] : 0ull;
      while (true)
        {
          while (m == 0ull && ++ch < nch) m = mk_l[ch][q];
          if (m == 0ull) break;
          const int c = 64 * ch + __builtin_ctzll (m);
          m &= m - 1ull;
          image_moments_add<V, T> (mdx, mdy, mdz, cand_l[c], xi, rmn, M);
        }
    }
  // out through LDS: (i, j) for i <= j (zeros for the padding pairs) and
  // (j, i) for i < j with the odd moments negated
  const size_t ms = (size_t) npad * npad;
  T *od = mt + (size_t) bx * npad * 16 + (size_t) by * 256;
  T *ob = mt + (size_t) by * npad * 16 + (size_t) bx * 256;
  const bool diag = bx == by;
  const int e = threadIdx.x;
  // direct element e: pair (e % 16, e / 16); transposed element e: pair
  // (e / 16, e % 16) seen from its j
  const bool wd = !diag || (e & 15) <= (e >> 4);
  const bool wt = !diag || (e >> 4) < (e & 15);
  T *pd = od + e, *pt = ob + e;
  const T *sd = &tr_l[0][(e & 15) * 17 + (e >> 4)], *st = &tr_l[0][(e >> 4) * 17 + (e & 15)];
#pragma unroll
  for (int t0 = 0; t0 < NM; t0 += MR)
    {
#pragma unroll
      for (int u = 0; u < MR; u++)
        if (t0 + u < NM) tr_l[u][il * 17 + jl] = M[t0 + u];
      __syncthreads ();
      if (wd)
#pragma unroll
        for (int u = 0; u < MR; u++)
          if (t0 + u < NM) pd[(size_t) (t0 + u) * ms] = sd[u * 16 * 17];
      if (wt)
#pragma unroll
        for (int u = 0; u < MR; u++)
          if (t0 + u < NM) pt[(size_t) (t0 + u) * ms] = T (moment_parity<V> (t0 + u)) * st[u * 16 * 17];
      __syncthreads ();
    }
}

// the block contraction of k_real_mfma_pc's consumer waves with the moments
// read from mt: NC waves of 32 columns per workgroup share the moments of a
// chunk of JC sources (8 for four waves, 4 for two: KS = JC / 4 MFMA
// k-steps; two workgroups per CU either way) through LDS.  The next
// chunk's moments go straight from mt to the other LDS buffer
// (global_load_lds_dwordx4, 512 contiguous bytes per moment, no registers),
// each k-step's z rows are reloaded as soon as its products have issued,
// and one barrier per chunk hands the buffers over.  jper and j0 are
// multiples of 8 and npad of 64, so a chunk never reads past row npad of
// mt; its rows at or past j1 hold finite moments (zeros past np) and meet
// zero z rows.

template <int V, int NC>
__global__ __launch_bounds__ (64 * NC) __attribute__ ((amdgpu_waves_per_eu (2))) void
k_real_mfma_mom (const float *__restrict__ mt, int npad, const float *__restrict__ XT, int B,
                 int zs, int np, int jper, float *__restrict__ partial)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  constexpr int JC = NC >= 4 ? 8 : 4, KS = JC / 4;
  constexpr int VPM = JC * 16 / 4;              // 16-byte vectors per moment and chunk
  __shared__ __attribute__ ((aligned (16))) float mom[2][NM * JC * 16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ir = lane & 15, kq = lane >> 4;
  const int ibase = blockIdx.x * 16;
  const int j0 = blockIdx.y * jper, j1 = min (np, j0 + jper);
  const int cbase = blockIdx.z * 32 * NC + 32 * wave;
  const size_t ms = (size_t) npad * npad;
  // moment t of the pair (ibase + i, j) at mrow[t ms + 16 j + i]
  const float *mrow = mt + (size_t) blockIdx.x * npad * 16;
  auto load_mom = [&] (int jb, int buf) {
    const float *src = mrow + (size_t) jb * 16;
    char *dst = (char *) mom[buf];
    for (int v0 = wave * 64; v0 < NM * VPM; v0 += 64 * NC)
      {
        const int v = v0 + lane;
        if (v < NM * VPM)
          __builtin_amdgcn_global_load_lds (src + (size_t) (v / VPM) * ms + 4 * (v % VPM),
                                            (lds_void *) (dst + 16 * v0), 16, 0, 0);
      }
  };
  mf4 acc[D][2];
#pragma unroll
  for (int t = 0; t < D; t++)
#pragma unroll
    for (int ct = 0; ct < 2; ct++) acc[t][ct] = mf4{0.f, 0.f, 0.f, 0.f};
  // z[j][q][cbase + 16 ct + ir] of k-step s (j = jb + 4 s + kq)
  float bz[KS][D][2];
  auto load_z = [&] (int jb, int s) {
    const int jz = jb + 4 * s + kq;
#pragma unroll
    for (int ct = 0; ct < 2; ct++)
      {
        const int b = cbase + 16 * ct + ir;
        const bool ok = jz < j1 && b < B;
        const float *src = XT + (size_t) (ok ? jz : 0) * zs + (ok ? b : 0);
#pragma unroll
        for (int q = 0; q < D; q++) bz[s][q][ct] = ok ? src[(size_t) q * B] : 0.f;
      }
  };
  load_mom (j0, 0);
#pragma unroll
  for (int s = 0; s < KS; s++) load_z (j0, s);
  async_wait_all ();
  int buf = 0;
  for (int jb = j0; jb < j1; jb += JC, buf ^= 1)
    {
      const bool more = jb + JC < j1;
      if (more) load_mom (jb + JC, buf ^ 1);
#pragma unroll
      for (int s = 0; s < KS; s++)
        {
          struct LdsM {
            const float *b;
            __device__ float operator[] (int t) const { return b[JC * 16 * t]; }
          } Ml{mom[buf] + 64 * s + lane};
#pragma unroll
          for (int q = 0; q < D; q++)
#pragma unroll
            for (int r = 0; r < D; r++)
              {
                const float a = block_entry<V, float> (Ml, r, q, false);
                acc[r][0] = mfma4 (a, bz[s][q][0], acc[r][0]);
                acc[r][1] = mfma4 (a, bz[s][q][1], acc[r][1]);
              }
          if (more) load_z (jb + JC, s);
        }
      async_wait_all ();
    }
#pragma unroll
  for (int r = 0; r < D; r++)
#pragma unroll
    for (int e = 0; e < 4; e++)
      {
        const int i = ibase + 4 * kq + e;
        if (i < np)
#pragma unroll
          for (int ct = 0; ct < 2; ct++)
            {
              const int b = cbase + 16 * ct + ir;
              if (b < B) partial[(((size_t) blockIdx.y * np + i) * D + r) * B + b] = acc[r][ct][e];
            }
      }
}

typedef double md4 __attribute__ ((ext_vector_type (4)));

__device__ __forceinline__ md4 mfma4d (double a, double b, md4 c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64 (a, b, c, 0, 0, 0);
}

// FP64 form of k_real_mfma_mom: v_mfma_f64_16x16x4f64 (A[i][k] in lane i +
// 16 k, B[k][j] in lane j + 16 k, D[i][j] in element i / 4 of lane j + 16
// (i % 4)), 16 columns per wave, NC waves per workgroup; chunks of JC
// sources (8 for eight waves, whose registers allow one workgroup per CU
// anyway, 4 otherwise) copied from mt to LDS by global_load_lds_dwordx4
// (jper and j0 multiples of 8, npad of 64: a chunk never reads past row
// npad; rows at or past j1 hold finite moments and meet zero z rows)
template <int V, int NC>
__global__ __launch_bounds__ (64 * NC) __attribute__ ((amdgpu_waves_per_eu (2))) void
k_real_mfma_mom64 (const double *__restrict__ mt, int npad, const double *__restrict__ XT,
                   int B, int zs, int np, int jper, double *__restrict__ partial)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  constexpr int JC = NC >= 8 ? 8 : 4, KS = JC / 4;
  constexpr int VPM = JC * 16 / 2;              // 16-byte vectors per moment and chunk
  __shared__ __attribute__ ((aligned (16))) double mom[2][NM * JC * 16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ir = lane & 15, kq = lane >> 4;
  const int ibase = blockIdx.x * 16;
  const int j0 = blockIdx.y * jper, j1 = min (np, j0 + jper);
  const int cbase = blockIdx.z * 16 * NC + 16 * wave;
  const size_t ms = (size_t) npad * npad;
  const double *mrow = mt + (size_t) blockIdx.x * npad * 16;
  auto load_mom = [&] (int jb, int buf) {
    const double *src = mrow + (size_t) jb * 16;
    char *dst = (char *) mom[buf];
    for (int v0 = wave * 64; v0 < NM * VPM; v0 += 64 * NC)
      {
        const int v = v0 + lane;
        if (v < NM * VPM)
          __builtin_amdgcn_global_load_lds (src + (size_t) (v / VPM) * ms + 2 * (v % VPM),
                                            (lds_void *) (dst + 16 * v0), 16, 0, 0);
      }
  };
  double bz[KS][D];
  auto load_z = [&] (int jb, int s) {
    const int jz = jb + 4 * s + kq, b = cbase + ir;
    const bool ok = jz < j1 && b < B;
    const double *src = XT + (size_t) (ok ? jz : 0) * zs + (ok ? b : 0);
#pragma unroll
    for (int q = 0; q < D; q++) bz[s][q] = ok ? src[(size_t) q * B] : 0.0;
  };
  md4 acc[D];
#pragma unroll
  for (int t = 0; t < D; t++) acc[t] = md4{0.0, 0.0, 0.0, 0.0};
  load_mom (j0, 0);
#pragma unroll
  for (int s = 0; s < KS; s++) load_z (j0, s);
  async_wait_all ();
  int buf = 0;
  for (int jb = j0; jb < j1; jb += JC, buf ^= 1)
    {
      const bool more = jb + JC < j1;
      if (more) load_mom (jb + JC, buf ^ 1);
#pragma unroll
      for (int s = 0; s < KS; s++)
        {
          struct LdsM {
            const double *b;
            __device__ double operator[] (int t) const { return b[JC * 16 * t]; }
          } Ml{mom[buf] + 64 * s + lane};
#pragma unroll
          for (int q = 0; q < D; q++)
#pragma unroll
            for (int r = 0; r < D; r++)
              acc[r] = mfma4d (block_entry<V, double> (Ml, r, q, false), bz[s][q], acc[r]);
          if (more) load_z (jb + JC, s);
        }
      async_wait_all ();
    }
#pragma unroll
  for (int r = 0; r < D; r++)
#pragma unroll
    for (int e = 0; e < 4; e++)
      {
        const int i = ibase + kq + 4 * e, b = cbase + ir;
        if (i < np && b < B) partial[(((size_t) blockIdx.y * np + i) * D + r) * B + b] = acc[r][e];
      }
}

}  // namespace shk
