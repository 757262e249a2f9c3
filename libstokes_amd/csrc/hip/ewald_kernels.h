// ewald_kernels.h -- device kernels of the Ewald/RPY mobility operator.
//
//   k_pairs       pair sum with per-pair moment accumulation and one 11x11
//                 contraction per pair: the open-boundary RPY operator
//                 (non-ewald.c:330-493), KIND 1.
//   kblock_*      per-k reciprocal blocks shared with recip_kernels.h.
//   k_matrix      dense (11N)^2 extracted-form matrix, ewald.c:1092-1677.
//   k_gemv_rows   y = mat . T(x), matrix.c:447-478.
// The periodic matvec lives in real_kernels.h and recip_kernels.h.
#pragma once
#include <hip/hip_runtime.h>
#include "pair_math.h"

namespace shk {

struct DevParams {
  int np, nr, nk;
  int poly;
  double xi, rmax2, rmin;
  double self_a, self_c, self_m, xiaspi, xia2;
  double gsign;   // sign of the reciprocal G term (+1: as the reference)
};

// self scalars of particle i (ewald.c:540-643)
template <typename T>
__device__ __forceinline__ void self_scalars (const DevParams &p, double ai,
                                              bool periodic, T *s)
{
#pragma unroll
  for (int k = 0; k < 11; k++) s[k] = T (0);
  double sa, sc, sm;
  if (!periodic)
    { sa = 1.0; sc = 0.75; sm = 0.9; }
  else if (!p.poly)
    { sa = p.self_a; sc = p.self_c; sm = p.self_m; }
  else
    {
      double a = ai, a2 = a * a, a3 = a2 * a;
      sa = 1.0 - a * p.xiaspi * (6.0 - 40.0 / 3.0 * p.xia2 * a2);
      sc = 0.75 - a3 * p.xiaspi * p.xia2 * 10.0;
      sm = 0.9 - a3 * p.xiaspi * p.xia2 * (12.0 - 30.24 * p.xia2 * a2);
    }
  s[0] = s[1] = T (sa);
  s[3] = s[4] = T (sc);
  s[8] = s[9] = s[10] = T (sm);
}

// The self block (e = 0) acts on the extracted vector as a diagonal: with
// a1 = self_a, c1 = self_c and mm3 = -self_m/3, mm5 = self_m/2 the
// transformed stresslet gives E = self_m S component by component
// (ewald.c:540-643 with fts.c:625-1010 at e = 0).
__device__ __forceinline__ double self_diag (const DevParams &p, double ai, bool periodic,
                                             int V, int q)
{
  double s[11];
  self_scalars<double> (p, ai, periodic, s);
  if (q < 3) return s[1];
  if (q < 6 && V > 0) return s[4];
  return s[10];
}

template <int V, typename T>
__device__ __forceinline__ void load_z (const double *x, int j, T *z)
{
  constexpr int D = Dim<V>::D;
#pragma unroll
  for (int q = 0; q < D; q++) z[q] = T (x[(size_t) j * D + q]);
  if (V == 2) s_transform (z);
}

// acc[p] += sum_q block(M)[p][q] * z[q]
template <int V, typename T, typename A>
__device__ __forceinline__ void block_apply (const T *M, const T *z, A *acc)
{
  constexpr int D = Dim<V>::D;
#pragma unroll
  for (int p = 0; p < D; p++)
    {
      T s = T (0);
#pragma unroll
      for (int q = 0; q < D; q++) s += block_entry<V, T> (M, p, q, false) * z[q];
      acc[p] += A (s);
    }
}

// image vector (the shear fold rm*shift is already in .x)
template <typename T> struct Img;
template <> struct Img<double> { using type = double4; };
template <> struct Img<float> { using type = float4; };

// KIND: 0 Ewald, 1 open-BC RPY
template <int V, typename T, int KIND>
__global__ __launch_bounds__ (256) void
k_pairs (const double4 *__restrict__ pos, const double *__restrict__ x,
         const typename Img<T>::type *__restrict__ img, DevParams p,
         int jchunk, double *__restrict__ partial)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
  const int i = blockIdx.x * 16 + ti;
  const int jbeg = blockIdx.y * jchunk;
  const int jend = min (p.np, jbeg + jchunk);
  T acc[D];
#pragma unroll
  for (int q = 0; q < D; q++) acc[q] = T (0);
  double4 pi = pos[min (i, p.np - 1)];
  const T rmax2 = T (p.rmax2);
  const T xi = T (p.xi);

  for (int jb = jbeg; jb < jend; jb += 16)
    {
      const int j = jb + tj;
      if (i >= p.np || j >= jend) continue;
      double4 pj = pos[j];
      T M[NM];
#pragma unroll
      for (int k = 0; k < NM; k++) M[k] = T (0);
      const double d0x = pj.x - pi.x, d0y = pj.y - pi.y, d0z = pj.z - pi.z;
      const T rminij = T ((pi.w + pj.w) * p.rmin);
      if (KIND == 0)
        {
          const T dx0 = T (d0x), dy0 = T (d0y), dz0 = T (d0z);
          for (int m = 0; m < p.nr; m++)
            {
              const typename Img<T>::type im = img[m];
              const T dx = dx0 + T (im.x), dy = dy0 + T (im.y), dz = dz0 + T (im.z);
              const T rr = dx * dx + dy * dy + dz * dz;
              if (rr == T (0) || rr > rmax2) continue;
              T r = m_sqrt (rr);
              if (r < rminij) r = rminij;
              const T ir = T (1) / r;
              T s[11];
              if (p.poly)
                scalars_real_poly_sd<V, T> (xi, r, T (pi.w), T (pj.w), s);
              else
                scalars_real<V, T> (xi, r, s);
              moments_add<V, T> (M, dx * ir, dy * ir, dz * ir, s);
            }
        }
      else
        {
          if (i == j) continue;
          const T dx = T (d0x), dy = T (d0y), dz = T (d0z);
          T r = m_sqrt (dx * dx + dy * dy + dz * dz);
          if (r < rminij) r = rminij;
          const T ir = T (1) / r;
          T s[11];
          scalars_rpy_sd<V, T> (r, T (pi.w), T (pj.w), p.poly != 0, s);
          moments_add<V, T> (M, dx * ir, dy * ir, dz * ir, s);
        }
      T z[D];
      load_z<V, T> (x, j, z);
      block_apply<V, T, T> (M, z, acc);
    }
#pragma unroll
  for (int q = 0; q < D; q++)
    {
      T v = acc[q];
      v += __shfl_xor (v, 8, 16);
      v += __shfl_xor (v, 4, 16);
      v += __shfl_xor (v, 2, 16);
      v += __shfl_xor (v, 1, 16);
      acc[q] = v;
    }
  if (tj == 0 && i < p.np)
    {
      double *out = partial + ((size_t) blockIdx.y * p.np + i) * D;
#pragma unroll
      for (int q = 0; q < D; q++) out[q] = double (acc[q]);
    }
}

// phase -> (cos, sin) in precision T; FP32 reduces the argument in FP64
__device__ __forceinline__ void phase_cs (double ph, double &c, double &s)
{
  sincos (ph, &s, &c);
}
__device__ __forceinline__ void phase_cs (double ph, float &c, float &s)
{
  const double inv2pi = 0.15915494309189533577, twopi = 6.28318530717958647692;
  double red = ph - twopi * rint (ph * inv2pi);
  sincosf (float (red), &s, &c);
}

// Polydisperse reciprocal scalars (ewald.c:985-1025) expanded in powers of
// the radii: every scalar of pair (i, j) is a sum of c(k) a_i^p a_j^q with
// q in {0, 2}.  For power p of a_i this returns the paired-k (x2) scalars
// acting on the q = 0 sums (sc1 even / ss1 odd) and on the q = 2 sums
// (sc2 / ss2), in the order xa ya yb xc yc xg yg yh xm ym zm.  The G
// terms take their own (signed) kexp, kexpg: sign * kexp, plus the
// reference-sign correction of another splitting where one is applied.
template <int P>
__device__ __forceinline__ void poly_kscalars (double k, double kk, double kexp, double kexpg,
                                               double *sc1, double *ss1, double *sc2, double *ss2)
{
#pragma unroll
  for (int t = 0; t < 11; t++) sc1[t] = ss1[t] = sc2[t] = ss2[t] = 0.0;
  const double e = 2.0 * kexp, eg = 2.0 * kexpg;
  if (P == 1)                                  // ya: 6 a_i - kk a_i a_j^2
    {
      sc1[1] = 6.0 * e;
      sc2[1] = -kk * e;
    }
  else if (P == 2)                             // yb; yg: 3k a_i^2 - 0.5 kk k a_i^2 a_j^2
    {
      ss1[2] = 3.0 * k * e;
      ss1[6] = 3.0 * k * eg;
      ss2[6] = -0.5 * kk * k * eg;
    }
  else if (P == 3)                             // ya, yc, yh, ym
    {
      sc1[1] = -kk * e;
      sc1[4] = 1.5 * kk * e;
      sc1[7] = 1.5 * kk * e;
      sc1[9] = 3.0 * kk * e;
      sc2[9] = -0.3 * kk * kk * e;
    }
  else if (P == 4)                             // yg: -0.3 kk k a_i^4
    ss1[6] = -0.3 * kk * k * eg;
  else                                         // ym: -0.3 kk^2 a_i^5
    sc1[9] = -0.3 * kk * kk * e;
}

// P += Ac C - As S, Q += Ac S + As C with the even / odd blocks Ac =
// block(e, sc), As = block(e, ss) of one k: four direct contractions
// (contract_item, the block is linear in its scalars), no moment arrays
template <int V>
__device__ __forceinline__ void kblock_apply_s (double ex, double ey, double ez, const double *sc,
                                                const double *ss, const double *C,
                                                const double *S, double *P, double *Q)
{
  double nss[11];
#pragma unroll
  for (int t = 0; t < 11; t++) nss[t] = -ss[t];
  contract_item<V, double> (ex, ey, ez, sc, C, P);
  contract_item<V, double> (ex, ey, ez, nss, S, P);
  contract_item<V, double> (ex, ey, ez, sc, S, Q);
  contract_item<V, double> (ex, ey, ez, ss, C, Q);
}

// kblock_apply_s for scalar sets whose possibly nonzero entries are the
// bits of MC (sc) and MS (ss): the contractions skip the zero scalars
template <int V, unsigned MC, unsigned MS>
__device__ __forceinline__ void kblock_apply_m (double ex, double ey, double ez, const double *sc,
                                                const double *ss, const double *C,
                                                const double *S, double *P, double *Q)
{
  double nss[11];
#pragma unroll
  for (int t = 0; t < 11; t++) nss[t] = -ss[t];
  if (MC) contract_item_m<V, double, MC> (ex, ey, ez, sc, C, P);
  if (MS) contract_item_m<V, double, MS> (ex, ey, ez, nss, S, P);
  if (MC) contract_item_m<V, double, MC> (ex, ey, ez, sc, S, Q);
  if (MS) contract_item_m<V, double, MS> (ex, ey, ez, ss, C, Q);
}

// the nonzero scalars of the reciprocal blocks: equal spheres ya, yc, yh,
// ym (even) and yb, yg (odd); unequal spheres per power of a_i
// (poly_kscalars): sc1, ss1, sc2, ss2
constexpr unsigned KM_MONO_C = 0x292u, KM_MONO_S = 0x44u;
template <int P> struct KMask;
template <> struct KMask<1> { static constexpr unsigned c1 = 0x2u, s1 = 0u, c2 = 0x2u, s2 = 0u; };
template <> struct KMask<2> { static constexpr unsigned c1 = 0u, s1 = 0x44u, c2 = 0u, s2 = 0x40u; };
template <> struct KMask<3> { static constexpr unsigned c1 = 0x292u, s1 = 0u, c2 = 0x200u, s2 = 0u; };
template <> struct KMask<4> { static constexpr unsigned c1 = 0u, s1 = 0x40u, c2 = 0u, s2 = 0u; };
template <> struct KMask<5> { static constexpr unsigned c1 = 0x200u, s1 = 0u, c2 = 0u, s2 = 0u; };

// P, Q (D each) of one k from its structure factors CS = [C | S] for equal
// spheres: the paired-k weights w of the k table (ya yb yc yg yh ym)
template <int V>
__device__ __forceinline__ void kblock_pq_mono (const double4 kvec, const double *w,
                                                const double *CS, double *P, double *Q)
{
  constexpr int D = Dim<V>::D;
  const double ik = 1.0 / kvec.w;
  const double ex = kvec.x * ik, ey = kvec.y * ik, ez = kvec.z * ik;
  const double sc[11] = {0.0, w[0], 0.0, 0.0, w[2], 0.0, 0.0, w[4], 0.0, w[5], 0.0};
  const double ss[11] = {0.0, 0.0, w[1], 0.0, 0.0, 0.0, w[3], 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < D; q++) P[q] = Q[q] = 0.0;
  kblock_apply_m<V, KM_MONO_C, KM_MONO_S> (ex, ey, ez, sc, ss, CS, CS + D, P, Q);
}

// P, Q of power PW (1..5) of a_i for unequal spheres: CS = [C | S | C2 | S2]
// with the a_j^2-weighted sums second (PW a compile-time constant, so the
// zero scalars of poly_kscalars fold away)
template <int V, int PW>
__device__ __forceinline__ void kblock_pq_poly (const double4 kvec, const double *w,
                                                const double *CS, double *P, double *Q)
{
  constexpr int D = Dim<V>::D;
  const double ik = 1.0 / kvec.w;
  const double ex = kvec.x * ik, ey = kvec.y * ik, ez = kvec.z * ik;
  double sc1[11], ss1[11], sc2[11], ss2[11];
  poly_kscalars<PW> (kvec.w, w[6], w[7], w[3], sc1, ss1, sc2, ss2);
#pragma unroll
  for (int q = 0; q < D; q++) P[q] = Q[q] = 0.0;
  kblock_apply_m<V, KMask<PW>::c1, KMask<PW>::s1> (ex, ey, ez, sc1, ss1, CS, CS + D, P, Q);
  kblock_apply_m<V, KMask<PW>::c2, KMask<PW>::s2> (ex, ey, ez, sc2, ss2, CS + 2 * D, CS + 3 * D,
                                                    P, Q);
}

// the same for a power known only at run time (1..5)
template <int V>
__device__ __forceinline__ void kblock_pq_poly_rt (int pw, const double4 kvec, const double *w,
                                                   const double *CS, double *P, double *Q)
{
  switch (pw)
    {
    case 1: kblock_pq_poly<V, 1> (kvec, w, CS, P, Q); break;
    case 2: kblock_pq_poly<V, 2> (kvec, w, CS, P, Q); break;
    case 3: kblock_pq_poly<V, 3> (kvec, w, CS, P, Q); break;
    case 4: kblock_pq_poly<V, 4> (kvec, w, CS, P, Q); break;
    default: kblock_pq_poly<V, 5> (kvec, w, CS, P, Q); break;
    }
}

// Dense build: one thread per (i, j) block.
template <int V, int KIND>
__global__ __launch_bounds__ (64) void
k_matrix (const double4 *__restrict__ pos, const double4 *__restrict__ img,
          const double4 *__restrict__ kv, const double *__restrict__ ks,
          DevParams p, double *__restrict__ mat)
{
  constexpr int D = Dim<V>::D, NM = Dim<V>::NMOM;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= p.np) return;
  const double4 pi = pos[i], pj = pos[j];
  double M[NM];
#pragma unroll
  for (int k = 0; k < NM; k++) M[k] = 0.0;
  double s[11];
  if (i == j)
    {
      self_scalars<double> (p, pi.w, KIND == 0, s);
      moments_add<V, double> (M, 0.0, 0.0, 0.0, s);
    }
  const double d0x = pj.x - pi.x, d0y = pj.y - pi.y, d0z = pj.z - pi.z;
  const double rminij = (pi.w + pj.w) * p.rmin;
  if (KIND == 0)
    {
      for (int m = 0; m < p.nr; m++)
        {
          const double4 im = img[m];
          const double dx = d0x + im.x, dy = d0y + im.y, dz = d0z + im.z;
          const double rr = dx * dx + dy * dy + dz * dz;
          if (rr == 0.0 || rr > p.rmax2) continue;
          double r = sqrt (rr);
          if (r < rminij) r = rminij;
          if (p.poly) scalars_real_poly_sd<V, double> (p.xi, r, pi.w, pj.w, s);
          else scalars_real<V, double> (p.xi, r, s);
          moments_add<V, double> (M, dx / r, dy / r, dz / r, s);
        }
      for (int k = 0; k < p.nk; k++)
        {
          const double4 kk = kv[k];
          const double *w = ks + (size_t) k * 8;
          double sf, cf;
          sincos (kk.x * d0x + kk.y * d0y + kk.z * d0z, &sf, &cf);
          sf = -sf;
          double ya = w[0], yb = w[1], yc = w[2], yg = w[3], yh = w[4], ym = w[5];
          if (p.poly)
            {
              // ewald.c:985-1025, SD form: a_i powers and a_i^2, a_j^2 terms
              const double aa = pi.w, aa2 = aa * aa, aa3 = aa2 * aa, ab2 = pj.w * pj.w;
              const double kk2 = w[6], kexp2 = 2.0 * w[7], kn = kk.w;
              ya = aa * 6.0 * (1.0 - kk2 * (aa2 + ab2) / 6.0) * kexp2;
              yb = aa2 * 3.0 * kn * kexp2;
              yc = aa3 * 1.5 * kk2 * kexp2;
              yg = aa2 * 3.0 * (1.0 - kk2 * (aa2 / 10.0 + ab2 / 6.0)) * kn * 2.0 * w[3];
              yh = aa3 * 1.5 * kk2 * kexp2;
              ym = aa3 * 3.0 * (1.0 - kk2 * (aa2 + ab2) / 10.0) * kk2 * kexp2;
            }
          const double ik = 1.0 / kk.w;
          double sc[11] = {0.0, cf * ya, sf * yb, 0.0, cf * yc, 0.0, sf * yg, cf * yh, 0.0, cf * ym, 0.0};
          moments_add<V, double> (M, kk.x * ik, kk.y * ik, kk.z * ik, sc);
        }
    }
  else if (i != j)
    {
      double r = sqrt (d0x * d0x + d0y * d0y + d0z * d0z);
      if (r < rminij) r = rminij;
      scalars_rpy_sd<V, double> (r, pi.w, pj.w, p.poly != 0, s);
      moments_add<V, double> (M, d0x / r, d0y / r, d0z / r, s);
    }
  const size_t n = (size_t) p.np * D;
  const bool self = (i == j);
  double *blk0 = mat + (size_t) i * D * n + (size_t) j * D;
  auto put = [&] (int r, int c, double v) { blk0[(size_t) r * n + c] = v; };
  BlockEach<V, 0, D, D * D>::run (M, self, put);
}

// y[row] = sum_col mat[row][col] * z[col], one wave per row
static __global__ __launch_bounds__ (256) void
k_gemv_rows (const double *__restrict__ mat, const double *__restrict__ z,
             size_t n, double *__restrict__ y)
{
  const size_t row = (size_t) blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double *mr = mat + row * n;
  double acc = 0.0;
  for (size_t c = lane; c < n; c += 64) acc += mr[c] * z[c];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor (acc, off, 64);
  if (lane == 0) y[row] = acc;
}

// z = T(x) per particle (FTS S transform); copies for F/FT
template <int V>
__global__ void k_transform (const double *__restrict__ x, int np, double *__restrict__ z)
{
  constexpr int D = Dim<V>::D;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double v[D];
  load_z<V, double> (x, i, v);
#pragma unroll
  for (int q = 0; q < D; q++) z[(size_t) i * D + q] = v[q];
}

}  // namespace shk
