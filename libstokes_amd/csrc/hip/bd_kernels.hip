// bd_kernels.hip -- device pieces of the Brownian-dynamics mid-point step
// (brownian.c:2352-2560, F, FT and FTS versions): Gaussian noise, forces F + fact z
// (brownian.c:2273-2311), Euler updates (brownian.c:1889-1930), the
// Chebyshev accuracy dots (chebyshev.c:241-269) and the overlap count
// (brownian.c:299-375), plus hipGraph capture / replay of a step.
//
// Gaussian noise: Philox4x32-10 counter-based generator (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011) and
// the Box-Muller transform.  Element e of draw c depends only on
// (seed, c, e); the draw index lives in device memory, so a captured graph
// produces a fresh draw on every replay.
//
// Step scalars (fact, dt) and the status block (dots, overlap counts) are
// device-resident structs read / written by the kernels, so one captured
// graph serves every dt the host may pick.
#include <hip/hip_runtime.h>
#include <cmath>
#include "../shk.h"

namespace {

struct u4 { unsigned x, y, z, w; };

__device__ __forceinline__ void mulhilo (unsigned a, unsigned b, unsigned &hi, unsigned &lo)
{
  const unsigned long long p = (unsigned long long) a * b;
  hi = (unsigned) (p >> 32);
  lo = (unsigned) p;
}

// Philox4x32, 10 rounds, with the multipliers and Weyl key increments of
// the published generator
__device__ __forceinline__ u4 philox10 (u4 c, unsigned k0, unsigned k1)
{
  const unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const unsigned W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++)
    {
      unsigned h0, l0, h1, l1;
      mulhilo (M0, c.x, h0, l0);
      mulhilo (M1, c.z, h1, l1);
      c = u4{h1 ^ c.y ^ k0, l1, h0 ^ c.w ^ k1, l0};
      k0 += W0;
      k1 += W1;
    }
  return c;
}

// 53-bit uniform in (0, 1] from two 32-bit words
__device__ __forceinline__ double u01 (unsigned hi, unsigned lo)
{
  const unsigned long long v = ((unsigned long long) hi << 21) | (unsigned long long) (lo >> 11);
  return ((double) v + 1.0) * (1.0 / 9007199254740992.0);
}

// two normals per thread
__global__ void k_gauss (unsigned long long seed, const unsigned long long *__restrict__ counter,
                         size_t n, double *__restrict__ out)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * t >= n) return;
  const unsigned long long c = *counter;
  const u4 ctr{(unsigned) t, (unsigned) (t >> 32), (unsigned) c, (unsigned) (c >> 32)};
  const u4 r = philox10 (ctr, (unsigned) seed, (unsigned) (seed >> 32));
  const double a = u01 (r.x, r.y), b = u01 (r.z, r.w);
  const double rad = sqrt (-2.0 * log (a));
  double s, co;
  sincospi (2.0 * b, &s, &co);
  out[2 * t] = rad * co;
  if (2 * t + 1 < n) out[2 * t + 1] = rad * s;
}

__global__ void k_counter_add (unsigned long long *counter, unsigned long long inc)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter += inc;
}

__global__ void k_status_clear (shk_bd_status *st)
{
  if (threadIdx.x == 0 && blockIdx.x == 0)
    {
      st->yy = 0.0;
      st->zmz = 0.0;
      st->overlap[0] = 0;
      st->overlap[1] = 0;
    }
}

// f = F + fact z (F or z may be NULL)
__global__ void k_bd_force (double *__restrict__ f, const double *__restrict__ F,
                            const double *__restrict__ z, const shk_bd_scalars *__restrict__ sc,
                            size_t n)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double fz = z ? sc->fact * z[i] : 0.0;
  f[i] = F ? F[i] + fz : fz;
}

// xo = xi + (frac dt) u
__global__ void k_bd_euler (double *__restrict__ xo, const double *__restrict__ xi,
                            const double *__restrict__ u, const shk_bd_scalars *__restrict__ sc,
                            double frac, size_t n)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  xo[i] = xi[i] + (frac * sc->dt) * u[i];
}

// FT / FTS forces per particle of 6: (F + fact z_F, T + fact z_T), F and T
// of 3 per particle (either may be NULL), z particle-wise of 6
// (BD_calc_forces, brownian.c:2273-2311, with BD_calc_FB's particle-wise
// noise, 1851-1871)
__global__ void k_bd_force_ft (double *__restrict__ f, const double *__restrict__ F,
                               const double *__restrict__ T, const double *__restrict__ z,
                               const shk_bd_scalars *__restrict__ sc, int np)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) np * 6) return;
  const size_t i = t / 6;
  const int q = (int) (t % 6);
  const double *ext = q < 3 ? F : T;
  const double fz = z ? sc->fact * z[t] : 0.0;
  f[t] = ext ? ext[i * 3 + (q % 3)] + fz : fz;
}

// FT / FTS Euler update (evolve_Euler_3all, brownian.c:1889-1930): xo = xi
// + (frac dt) U and, when qi is given, qo = qi + (frac dt) dQ/dt with
// dQ/dt = 1/2 W(q)^{-1} (O, 0) (quaternion_dQdt, ode-quaternion.c:89-105;
// W(q)^{-1} = W(q)^T / |q|^2, quaternion_Winv, :58-79); u holds (U, O) per
// particle of 6
__global__ void k_bd_euler_ft (double *__restrict__ xo, const double *__restrict__ xi,
                               double *__restrict__ qo, const double *__restrict__ qi,
                               const double *__restrict__ u,
                               const shk_bd_scalars *__restrict__ sc, double frac, int np)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double h = frac * sc->dt;
  const double *ui = u + (size_t) i * 6;
  for (int a = 0; a < 3; a++) xo[i * 3 + a] = xi[i * 3 + a] + h * ui[a];
  if (!qi) return;
  const double q0 = qi[i * 4], q1 = qi[i * 4 + 1], q2 = qi[i * 4 + 2], q3 = qi[i * 4 + 3];
  const double o0 = ui[3], o1 = ui[4], o2 = ui[5];
  const double hd = 0.5 / (q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
  const double d0 = hd * (q3 * o0 - q2 * o1 + q1 * o2);
  const double d1 = hd * (q2 * o0 + q3 * o1 - q0 * o2);
  const double d2 = hd * (-q1 * o0 + q0 * o1 + q3 * o2);
  const double d3 = hd * (-q0 * o0 - q1 * o1 - q2 * o2);
  qo[i * 4] = q0 + h * d0;
  qo[i * 4 + 1] = q1 + h * d1;
  qo[i * 4 + 2] = q2 + h * d2;
  qo[i * 4 + 3] = q3 + h * d3;
}

// y.y and z.w in one block (fixed summation order: deterministic)
__global__ __launch_bounds__ (1024) void
k_bd_dots (const double *__restrict__ y, const double *__restrict__ z,
           const double *__restrict__ w, size_t n, shk_bd_status *st)
{
  __shared__ double ry[1024], rz[1024];
  double ay = 0.0, az = 0.0;
  for (size_t k = threadIdx.x; k < n; k += 1024)
    {
      ay += y[k] * y[k];
      az += z[k] * w[k];
    }
  ry[threadIdx.x] = ay;
  rz[threadIdx.x] = az;
  __syncthreads ();
  for (int s = 512; s > 0; s >>= 1)
    {
      if ((int) threadIdx.x < s)
        {
          ry[threadIdx.x] += ry[threadIdx.x + s];
          rz[threadIdx.x] += rz[threadIdx.x + s];
        }
      __syncthreads ();
    }
  if (threadIdx.x == 0)
    {
      st->yy = ry[0];
      st->zmz = rz[0];
    }
}

// overlapping (pair, image) count of brownian.c:299-375: pairs i <= j, the
// primary cell for i != j and, in a periodic box, the 26 neighbouring cells
// of the raw separation (a sheared box shifts x by the y or z cell index);
// a pair-image overlaps when r^2 < rfac (a_i + a_j)^2
__global__ __launch_bounds__ (256) void
k_bd_overlap (const double *__restrict__ pos, const double *__restrict__ a, int np, int periodic,
              double lx, double ly, double lz, int shear_mode, double shift, double rfac,
              unsigned long long *__restrict__ count)
{
  const int i = blockIdx.y;
  const int j = i + blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  const double x = pos[3 * i] - pos[3 * j];
  const double y = pos[3 * i + 1] - pos[3 * j + 1];
  const double z = pos[3 * i + 2] - pos[3 * j + 2];
  const double s = a ? a[i] + a[j] : 2.0;
  const double lim = rfac * (s * s);
  unsigned cnt = 0;
  if (i != j) cnt += (x * x + y * y + z * z) < lim;
  if (periodic)
    for (int m = 0; m < 27; m++)
      {
        if (m == 13) continue;   // (0, 0, 0)
        const int mx = m / 9 - 1, my = (m / 3) % 3 - 1, mz = m % 3 - 1;
        double xx = x + mx * lx;
        const double yy = y + my * ly, zz = z + mz * lz;
        if (shear_mode == 1) xx += my * shift;
        else if (shear_mode == 2) xx += mz * shift;
        cnt += (xx * xx + yy * yy + zz * zz) < lim;
      }
  if (cnt) atomicAdd (count, (unsigned long long) cnt);
}

}  // namespace

#define CHK(call)                                                           \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) return shk_set_hip_error (#call, e_);             \
  } while (0)

static int shk_set_hip_error (const char *what, hipError_t e)
{
  shk_set_error ("%s: %s", what, hipGetErrorString (e));
  return SHK_EHIP;
}

static unsigned blocks_for (size_t n, unsigned bs) { return (unsigned) ((n + bs - 1) / bs); }

extern "C" int shk_dev_gauss (shk_dev *d, unsigned long long seed,
                              const unsigned long long *counter_dev, size_t n, double *out)
{
  if (n == 0) return SHK_OK;
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_gauss, dim3 (blocks_for ((n + 1) / 2, 256)), dim3 (256), 0, s, seed,
                      counter_dev, n, out);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_counter_add (shk_dev *d, unsigned long long *counter_dev,
                                    unsigned long long inc)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_counter_add, dim3 (1), dim3 (64), 0, s, counter_dev, inc);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_bd_status_clear (shk_dev *d, shk_bd_status *st)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_status_clear, dim3 (1), dim3 (64), 0, s, st);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_bd_force (shk_dev *d, double *f, const double *F, const double *z,
                                 const shk_bd_scalars *sc, size_t n)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_bd_force, dim3 (blocks_for (n, 256)), dim3 (256), 0, s, f, F, z, sc, n);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_bd_euler (shk_dev *d, double *xo, const double *xi, const double *u,
                                 const shk_bd_scalars *sc, double frac, size_t n)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_bd_euler, dim3 (blocks_for (n, 256)), dim3 (256), 0, s, xo, xi, u, sc,
                      frac, n);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_bd_force_ft (shk_dev *d, double *f, const double *F, const double *T,
                                    const double *z, const shk_bd_scalars *sc, int np)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_bd_force_ft, dim3 (blocks_for ((size_t) np * 6, 256)), dim3 (256), 0, s,
                      f, F, T, z, sc, np);
  CHK (hipGetLastError ());
  return SHK_OK;
}

// the imposed flow at particle i (shift_rest_to_labo_U / _O, f.c:621-660,
// ft.c:935-955; the strain part of shift_labo_to_rest_E, fts.c:2303-2320),
// added with sign to v[D i + q] for q in [qlo, qhi)
__global__ void k_bd_imposed (double *__restrict__ v, const double *__restrict__ p, int n,
                              int D, int qlo, int qhi, shk_flow fl, double sign)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
  const double *E = fl.E;
  const double Ezz = -E[0] - E[4];
  double w[11];
  w[0] = fl.U[0] + fl.O[1] * z - fl.O[2] * y + E[0] * x + E[1] * y + E[2] * z;
  w[1] = fl.U[1] + fl.O[2] * x - fl.O[0] * z + E[1] * x + E[4] * y + E[3] * z;
  w[2] = fl.U[2] + fl.O[0] * y - fl.O[1] * x + E[2] * x + E[3] * y + Ezz * z;
  for (int q = 0; q < 3; q++) w[3 + q] = fl.O[q];
  for (int q = 0; q < 5; q++) w[6 + q] = E[q];
  double *vi = v + (size_t) i * D;
  for (int q = qlo; q < qhi; q++) vi[q] += sign * w[q];
}

extern "C" int shk_dev_bd_imposed (shk_dev *d, double *v, const double *p, int n, int D,
                                   int qlo, int qhi, const shk_flow *fl, double sign)
{
  if (n <= 0) return SHK_OK;
  if (qlo < 0 || qhi > D || qhi > 11) return SHK_EARG;
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_bd_imposed, dim3 (blocks_for ((size_t) n, 256)), dim3 (256), 0, s, v, p,
                      n, D, qlo, qhi, *fl, sign);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_bd_euler_ft (shk_dev *d, double *xo, const double *xi, double *qo,
                                    const double *qi, const double *u, const shk_bd_scalars *sc,
                                    double frac, int np)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_bd_euler_ft, dim3 (blocks_for ((size_t) np, 256)), dim3 (256), 0, s, xo,
                      xi, qo, qi, u, sc, frac, np);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_bd_dots (shk_dev *d, const double *y, const double *z, const double *w,
                                size_t n, shk_bd_status *st)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_bd_dots, dim3 (1), dim3 (1024), 0, s, y, z, w, n, st);
  CHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_bd_overlap (shk_dev *d, const double *pos, const double *a, int np,
                                   int periodic, double lx, double ly, double lz,
                                   int shear_mode, double shift, double rfac,
                                   unsigned long long *count)
{
  if (np <= 0 || np > 65535)
    {
      shk_set_error ("overlap check: np = %d outside [1, 65535]", np);
      return SHK_EARG;
    }
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipLaunchKernelGGL (k_bd_overlap, dim3 (blocks_for ((size_t) np, 256), np), dim3 (256), 0, s,
                      pos, a, np, periodic, lx, ly, lz, shear_mode, shift, rfac, count);
  CHK (hipGetLastError ());
  return SHK_OK;
}

// ---- graph capture of a step ------------------------------------------------

extern "C" int shk_dev_capture_begin (shk_dev *d)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  CHK (hipStreamBeginCapture (s, hipStreamCaptureModeThreadLocal));
  return SHK_OK;
}

extern "C" int shk_dev_capture_end (shk_dev *d, void **exec)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipGraph_t g = nullptr;
  CHK (hipStreamEndCapture (s, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t rc = hipGraphInstantiate (&e, g, nullptr, nullptr, 0);
  (void) hipGraphDestroy (g);
  CHK (rc);
  *exec = (void *) e;
  return SHK_OK;
}

// end a capture that failed part-way (the partial graph is discarded)
extern "C" void shk_dev_capture_abort (shk_dev *d)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture (s, &g) == hipSuccess && g) (void) hipGraphDestroy (g);
  (void) hipGetLastError ();
}

extern "C" int shk_dev_graph_launch (shk_dev *d, void *exec)
{
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  CHK (hipGraphLaunch ((hipGraphExec_t) exec, s));
  return SHK_OK;
}

extern "C" void shk_dev_graph_destroy (void *exec)
{
  if (exec) (void) hipGraphExecDestroy ((hipGraphExec_t) exec);
}

extern "C" int shk_host_alloc (void **p, size_t bytes)
{
  CHK (hipHostMalloc (p, bytes, hipHostMallocDefault));
  return SHK_OK;
}

extern "C" void shk_host_free (void *p)
{
  if (p) (void) hipHostFree (p);
}
