// linalg.hip -- small dense helpers for the device GMRES (Arnoldi basis
// products) and the FTS mobility-problem operator pieces
// (ewald-3fts.c:140-253).
#include <hip/hip_runtime.h>
#include "../shk.h"

namespace {

// out[c] = sum_k V[c n + k] w[k] for c < m: one block per column, fixed
// tree order (deterministic)
__global__ __launch_bounds__ (256) void
k_vtw (const double *__restrict__ V, const double *__restrict__ w, size_t n,
       double *__restrict__ out)
{
  __shared__ double red[256];
  const double *v = V + (size_t) blockIdx.x * n;
  double acc = 0.0;
  for (size_t k = threadIdx.x; k < n; k += 256) acc += v[k] * w[k];
  red[threadIdx.x] = acc;
  __syncthreads ();
  for (int s = 128; s > 0; s >>= 1)
    {
      if ((int) threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads ();
    }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// w[k] += sign * sum_c V[c n + k] h[c]
__global__ void k_vh (const double *__restrict__ V, const double *__restrict__ h, int m,
                      size_t n, double sign, double *__restrict__ w)
{
  const size_t k = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double acc = 0.0;
  for (int c = 0; c < m; c++) acc += V[(size_t) c * n + k] * h[c];
  w[k] += sign * acc;
}

// y = a x + b y
__global__ void k_axpby (double *__restrict__ y, const double *__restrict__ x, double a,
                         double b, size_t n)
{
  const size_t k = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[k] = a * x[k] + b * y[k];
}

// per particle of D values, mobile when i < nm; the first qs values of a
// mobile particle are its known forces (F; F, T), the rest of them and
// every value of a fixed particle its known velocities: lo = the known
// forces (0 elsewhere), hi = the rest (0 at the known forces); either may
// be NULL
__global__ void k_mix_split (const double *__restrict__ x, int np, int nm, int D, int qs,
                             double *__restrict__ lo, double *__restrict__ hi)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) np * D) return;
  const int q = (int) (t % D);
  const bool kf = (int) (t / D) < nm && q < qs;
  const double v = x[t];
  if (lo) lo[t] = kf ? v : 0.0;
  if (hi) hi[t] = kf ? 0.0 : v;
}

// y = lo(a) - z (mode 0) or y = z - hi(a) (mode 1), per particle of D
__global__ void k_mix_combine (const double *__restrict__ a, const double *__restrict__ z,
                               int np, int nm, int D, int qs, int mode, double *__restrict__ y)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) np * D) return;
  const int q = (int) (t % D);
  const bool kf = (int) (t / D) < nm && q < qs;
  if (mode == 0) y[t] = (kf ? a[t] : 0.0) - z[t];
  else y[t] = z[t] - (kf ? 0.0 : a[t]);
}

// (F, T) per particle of 6 -> (F, T, 0) per particle of 11 (mode 0), or
// the first 6 of 11 back (mode 1): the FT view of the FTS mobility problem
// with E = 0 (BD_atimes_mob_FU, brownian.c:597-634)
__global__ void k_ft_fts (const double *__restrict__ x, int np, int mode,
                          double *__restrict__ y)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) np * 11) return;
  const size_t i = t / 11;
  const int q = (int) (t % 11);
  if (mode == 0) y[t] = q < 6 ? x[i * 6 + q] : 0.0;
  else if (q < 6) y[i * 6 + q] = x[t];
}

// y[k] *= s[k]
__global__ void k_dmul (double *__restrict__ y, const double *__restrict__ sv, size_t n)
{
  const size_t k = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[k] *= sv[k];
}

}  // namespace

#define CHK(call)                                                           \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      {                                                                     \
        shk_set_error ("%s: %s", #call, hipGetErrorString (e_));            \
        return SHK_EHIP;                                                    \
      }                                                                     \
  } while (0)

extern "C" int shk_dev_axpby (shk_dev *d, double *y, const double *x, double a, double b,
                              size_t len)
{
  if (len == 0) return SHK_OK;
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_axpby, dim3 ((unsigned) ((len + 255) / 256)), dim3 (256), 0, s, y, x, a,
                      b, len);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_vtw (shk_dev *d, const double *V, int m, size_t n, const double *w,
                            double *out_dev)
{
  if (m <= 0) return SHK_OK;
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_vtw, dim3 (m), dim3 (256), 0, s, V, w, n, out_dev);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_vh (shk_dev *d, const double *V, const double *h_dev, int m, size_t n,
                           double sign, double *w)
{
  if (m <= 0) return SHK_OK;
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_vh, dim3 ((unsigned) ((n + 255) / 256)), dim3 (256), 0, s, V, h_dev, m,
                      n, sign, w);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_mix_split (shk_dev *d, const double *x, int np, int nm, int D, int qs,
                                  double *lo, double *hi)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  const size_t n = (size_t) np * D;
  hipLaunchKernelGGL (k_mix_split, dim3 ((unsigned) ((n + 255) / 256)), dim3 (256), 0, s, x, np,
                      nm, D, qs, lo, hi);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_mix_combine (shk_dev *d, const double *a, const double *z, int np,
                                    int nm, int D, int qs, int mode, double *y)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  const size_t n = (size_t) np * D;
  hipLaunchKernelGGL (k_mix_combine, dim3 ((unsigned) ((n + 255) / 256)), dim3 (256), 0, s, a,
                      z, np, nm, D, qs, mode, y);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_fts_split (shk_dev *d, const double *x, int np, int nm, double *lo,
                                  double *hi)
{
  return shk_dev_mix_split (d, x, np, nm, 11, 6, lo, hi);
}

extern "C" int shk_dev_fts_combine (shk_dev *d, const double *a, const double *z, int np,
                                    int nm, int mode, double *y)
{
  return shk_dev_mix_combine (d, a, z, np, nm, 11, 6, mode, y);
}

extern "C" int shk_dev_ft_fts (shk_dev *d, const double *x, int np, int mode, double *y)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  const size_t n = (size_t) np * 11;
  hipLaunchKernelGGL (k_ft_fts, dim3 ((unsigned) ((n + 255) / 256)), dim3 (256), 0, s, x, np,
                      mode, y);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_dmul (shk_dev *d, double *y, const double *sv, size_t n)
{
  if (n == 0) return SHK_OK;
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_dmul, dim3 ((unsigned) ((n + 255) / 256)), dim3 (256), 0, s, y, sv, n);
  CHK (hipGetLastError ());
  return SHK_OK;
}
