// dgemm_kernels.h -- C = A B in FP64 on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64), row-major operands with leading dimensions.
//
// A workgroup of four waves owns a 64 x 64 tile of C; wave w computes the
// 32 x 32 quarter (rows 32 (w / 2), columns 32 (w % 2)) as 2 x 2 MFMA tiles
// of 16 x 16.  K goes through LDS 16 at a time, double-buffered: the next
// step's A (64 x 16) and B (16 x 64) slices are read into registers before
// the current step's 16 products per wave, then written to the other LDS
// buffer (one barrier per step).  Both slices are held k-major
// (As[k][row], Bs[k][col]) so that the 16 lanes of an operand row read
// consecutive doubles.  Split-K: workgroup z takes K range [z kper, (z + 1)
// kper) and writes its tile to C + z cz; the slices are summed in slice
// order by k_dgemm_sum (no atomics: bitwise repeatable).
//
// Operand layouts of the instruction (checked in-bounds before use):
// A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in element
// i / 4 of lane j + 16 (i % 4).
#pragma once
#include <hip/hip_runtime.h>

namespace shk {

typedef double gd4 __attribute__ ((ext_vector_type (4)));

constexpr int DG_LD = 66;          // LDS row stride of a k slice (doubles)

__global__ __launch_bounds__ (256) void
k_dgemm (const double *__restrict__ A, int lda, const double *__restrict__ B, int ldb,
         double *__restrict__ C, int ldc, size_t cz, int M, int N, int K, int kper)
{
  __shared__ double As[2][16 * DG_LD];
  __shared__ double Bs[2][16 * DG_LD];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int ir = lane & 15, kq = lane >> 4;
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int kb = blockIdx.z * kper, ke = min (K, kb + kper);
  const int wr = 32 * (wave >> 1), wc = 32 * (wave & 1);
  // this thread's share of a step: A row t / 4, k 4 (t % 4) .. + 3; B row
  // t / 16, columns 4 (t % 16) .. + 3
  const int ar = t >> 2, ak = 4 * (t & 3);
  const int bk = t >> 4, bc = 4 * (t & 15);
  double ra[4], rb[4];
  auto load = [&] (int k0) {
    const int gr = r0 + ar;
#pragma unroll
    for (int u = 0; u < 4; u++)
      {
        const int gk = k0 + ak + u;
        ra[u] = (gr < M && gk < ke) ? A[(size_t) gr * lda + gk] : 0.0;
      }
    const int gk = k0 + bk;
#pragma unroll
    for (int u = 0; u < 4; u++)
      {
        const int gc = c0 + bc + u;
        rb[u] = (gk < ke && gc < N) ? B[(size_t) gk * ldb + gc] : 0.0;
      }
  };
  auto store = [&] (int buf) {
#pragma unroll
    for (int u = 0; u < 4; u++) As[buf][(ak + u) * DG_LD + ar] = ra[u];
#pragma unroll
    for (int u = 0; u < 4; u++) Bs[buf][bk * DG_LD + bc + u] = rb[u];
  };
  gd4 acc[2][2];
#pragma unroll
  for (int mi = 0; mi < 2; mi++)
#pragma unroll
    for (int nj = 0; nj < 2; nj++) acc[mi][nj] = gd4{0.0, 0.0, 0.0, 0.0};
  if (kb < ke)
    {
      load (kb);
      store (0);
      __syncthreads ();
      int buf = 0;
      for (int k0 = kb; k0 < ke; k0 += 16, buf ^= 1)
        {
          const bool more = k0 + 16 < ke;
          if (more) load (k0 + 16);
          const double *as = As[buf], *bs = Bs[buf];
#pragma unroll
          for (int ks = 0; ks < 4; ks++)
            {
              const int kk = (4 * ks + kq) * DG_LD;
              const double a0 = as[kk + wr + ir], a1 = as[kk + wr + 16 + ir];
              const double b0 = bs[kk + wc + ir], b1 = bs[kk + wc + 16 + ir];
              acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64 (a0, b0, acc[0][0], 0, 0, 0);
              acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64 (a0, b1, acc[0][1], 0, 0, 0);
              acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64 (a1, b0, acc[1][0], 0, 0, 0);
              acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64 (a1, b1, acc[1][1], 0, 0, 0);
            }
          if (more) store (buf ^ 1);
          __syncthreads ();
        }
    }
  double *out = C + (size_t) blockIdx.z * cz;
#pragma unroll
  for (int mi = 0; mi < 2; mi++)
#pragma unroll
    for (int nj = 0; nj < 2; nj++)
#pragma unroll
      for (int e = 0; e < 4; e++)
        {
          const int r = r0 + wr + 16 * mi + kq + 4 * e, c = c0 + wc + 16 * nj + ir;
          if (r < M && c < N) out[(size_t) r * ldc + c] = acc[mi][nj][e];
        }
}

// C[r][c] = sum over z < nz of P[z cz + r ldp + c] in slice order
__global__ void k_dgemm_sum (const double *__restrict__ P, int ldp, size_t cz, int nz, int M,
                             int N, double *__restrict__ C, int ldc)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) M * N) return;
  const int r = (int) (t / N), c = (int) (t - (size_t) r * N);
  double s = 0.0;
  for (int z = 0; z < nz; z++) s += P[(size_t) z * cz + (size_t) r * ldp + c];
  C[(size_t) r * ldc + c] = s;
}

}  // namespace shk
