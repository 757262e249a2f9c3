// shk_dev.hip -- device context, buffers and kernel launches behind the
// internal C ABI of shk.h.  One context per struct stokes, one HIP stream
// per context (default priority), deterministic reduction order.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <vector>
#include <algorithm>
#include <mutex>
#include <type_traits>
#include "../shk.h"
#include "ewald_kernels.h"
#include "real_kernels.h"
#include "recip_kernels.h"
#include "block_kernels.h"
#include "mfma_kernels.h"
#include "cache_kernels.h"
#include "adj_kernels.h"
#include "lub_kernels.h"
#include "dgemm_kernels.h"
#include "matrix_kernels.h"

using namespace shk;

static thread_local char g_err[512] = "";

static void set_err (const char *fmt, ...)
{
  va_list ap;
  va_start (ap, fmt);
  vsnprintf (g_err, sizeof (g_err), fmt, ap);
  va_end (ap);
}

extern "C" const char *shk_last_error (void) { return g_err; }

extern "C" void shk_set_error (const char *fmt, ...)
{
  va_list ap;
  va_start (ap, fmt);
  vsnprintf (g_err, sizeof (g_err), fmt, ap);
  va_end (ap);
}

#define HIPCHK(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      {                                                                       \
        set_err ("%s:%d %s: %s", __FILE__, __LINE__, #call,                   \
                 hipGetErrorString (e_));                                     \
        return SHK_EHIP;                                                      \
      }                                                                       \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct shk_dev {
  hipStream_t stream = nullptr;
  shk_params prm{};
  int have_sys = 0;
  DevBuf pos, img, imgf, kv, ks, partial, pq, xs, ys, zt;
  DevBuf cand, candf, candm, tbits, kpart, kpq, kback;
  DevBuf adjfg, adjcs, adjpart;
  DevBuf lubcoef, lubf1, lubcnt, luboff, lublist, lubval;   // lubrication
  DevBuf momt;                    // real-space pair moments (equal spheres)
  DevBuf eph, gpart;              // FP64 phase matrices; split-K GEMM partials
  int tsym = 0;                   // reference table symmetric under m -> -m
  int lub_N = -1;                 // order of the uploaded coefficient table
  int lub_f1 = 0;                 // f_k(1) table uploaded
  DevBuf km, ptab;                // k lattice indices, per-particle phase tables
  unsigned long pos_gen = 0;      // positions uploads so far
  unsigned long dense_gen = 0;    // positions of the per-call dense matrix
  int dense_ok = 0;               // rcache holds that matrix (FP32)
  int dense_blocks = 0;           // allow the per-call dense path
  unsigned long ptab_gen = 0;     // positions the single-column table was built
  int ptab_kind = 0;              // for (1 FP32, 2 FP64; 0 none)
  DevBuf rcache, rz;              // dense real-space matrix, its vector(s)
  int rc_use = 0, rc_prec = 0;    // products use rcache (built in rc_prec)
  PhaseBasis basis{};
  int ncand = 0, ncore = 0, tbits_words = 0;
  int tx = 0, ty = 0, tz = 0;
  double s2x = 0.0, s3x = 0.0;
  int timing = 0;
  int ncu = 256;                  // compute units of the device
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // FP64 block products: the reciprocal part on a second stream beside
  // the real-space part (forked after the block is prepared, joined before
  // the combine)
  int overlap = 1;
  hipStream_t side = nullptr;
  hipEvent_t evf = nullptr, evj = nullptr;
  double last_ms[3] = {0, 0, 0};
  unsigned long generation = 0;   // bumped whenever a work buffer moves
  // allocation-free products: while carve_base is set, the scratch buffers
  // below are carved from the per-GPU array instead of allocated, and any
  // other buffer that would have to grow is refused
  unsigned char *carve_base = nullptr;
  size_t carve_cap = 0, carve_off = 0;
  DevBuf carve_saved[6];
};

// the work buffers of the single-column product that hold nothing between
// calls (every entry read is written earlier in the same call)
static DevBuf *scratch_buf (shk_dev *d, int k)
{
  DevBuf *t[6] = {&d->partial, &d->rz, &d->kpart, &d->kpq, &d->kback, &d->ptab};
  return t[k];
}

static int ensure (shk_dev *d, DevBuf &b, size_t bytes)
{
  if (d->carve_base)
    {
      for (int k = 0; k < 6; k++)
        if (&b == scratch_buf (d, k))
          {
            const size_t off = (d->carve_off + 255) & ~(size_t) 255;
            if (off + bytes > d->carve_cap)
              {
                set_err ("allocation-free product: this system needs more than the %zu bytes"
                         " of scratch the library holds per GPU", d->carve_cap);
                return SHK_ENOMEM;
              }
            b.p = d->carve_base + off;
            b.bytes = bytes;
            d->carve_off = off + bytes;
            return SHK_OK;
          }
      if (b.bytes >= bytes && b.p) return SHK_OK;
      set_err ("allocation-free product: a device buffer of this system is not set up"
               " (run one ordinary product on it first)");
      return SHK_EARG;
    }
  if (b.bytes >= bytes && b.p) return SHK_OK;
  if (b.p) { HIPCHK (hipFree (b.p)); b.p = nullptr; b.bytes = 0; }
  if (bytes == 0) return SHK_OK;
  hipError_t e = hipMalloc (&b.p, bytes);
  if (e != hipSuccess)
    {
      set_err ("hipMalloc(%zu): %s", bytes, hipGetErrorString (e));
      return SHK_ENOMEM;
    }
  b.bytes = bytes;
  d->generation++;
  return SHK_OK;
}

static void release (DevBuf &b)
{
  if (b.p) (void) hipFree (b.p);
  b.p = nullptr;
  b.bytes = 0;
}

extern "C" shk_dev *shk_dev_create (void)
{
  int ndev = 0;
  if (hipGetDeviceCount (&ndev) != hipSuccess || ndev < 1)
    {
      set_err ("no HIP device available");
      return nullptr;
    }
  shk_dev *d = new shk_dev ();
  int dev = 0, ncu = 0;
  if (hipGetDevice (&dev) == hipSuccess
      && hipDeviceGetAttribute (&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess
      && ncu > 0)
    d->ncu = ncu;
  if (hipStreamCreateWithFlags (&d->stream, hipStreamNonBlocking) != hipSuccess
      || hipStreamCreateWithFlags (&d->side, hipStreamNonBlocking) != hipSuccess
      || hipEventCreateWithFlags (&d->evf, hipEventDisableTiming) != hipSuccess
      || hipEventCreateWithFlags (&d->evj, hipEventDisableTiming) != hipSuccess)
    {
      set_err ("hipStreamCreate failed");
      shk_dev_destroy (d);
      return nullptr;
    }
  for (auto &e : d->ev)
    if (hipEventCreate (&e) != hipSuccess)
      {
        set_err ("hipEventCreate failed");
        shk_dev_destroy (d);
        return nullptr;
      }
  return d;
}

extern "C" void shk_dev_destroy (shk_dev *d)
{
  if (!d) return;
  if (d->stream) (void) hipStreamSynchronize (d->stream);
  for (DevBuf *b : {&d->pos, &d->img, &d->imgf, &d->kv, &d->ks, &d->partial,
                    &d->pq, &d->xs, &d->ys, &d->zt, &d->cand, &d->candf, &d->candm,
                    &d->tbits, &d->kpart, &d->kpq, &d->kback, &d->adjfg, &d->adjcs,
                    &d->adjpart, &d->km, &d->ptab, &d->rcache, &d->rz, &d->lubcoef,
                    &d->lubf1, &d->lubcnt, &d->luboff, &d->lublist, &d->lubval,
                    &d->momt, &d->eph, &d->gpart})
    release (*b);
  if (d->side) (void) hipStreamSynchronize (d->side);
  for (auto &e : d->ev)
    if (e) (void) hipEventDestroy (e);
  if (d->evf) (void) hipEventDestroy (d->evf);
  if (d->evj) (void) hipEventDestroy (d->evj);
  if (d->side) (void) hipStreamDestroy (d->side);
  if (d->stream) (void) hipStreamDestroy (d->stream);
  delete d;
}

extern "C" void *shk_dev_stream (shk_dev *d) { return d ? (void *) d->stream : nullptr; }

extern "C" unsigned long shk_dev_generation (shk_dev *d) { return d ? d->generation : 0; }

extern "C" void *shk_dev_swap_stream (shk_dev *d, void *stream)
{
  void *old = (void *) d->stream;
  d->stream = (hipStream_t) stream;
  return old;
}

extern "C" int shk_dev_malloc (shk_dev *d, void **p, size_t bytes)
{
  (void) d;
  hipError_t e = hipMalloc (p, bytes);
  if (e != hipSuccess)
    {
      set_err ("hipMalloc(%zu): %s", bytes, hipGetErrorString (e));
      return SHK_ENOMEM;
    }
  return SHK_OK;
}

extern "C" int shk_dev_free (shk_dev *d, void *p)
{
  if (d && d->stream) HIPCHK (hipStreamSynchronize (d->stream));
  if (p) HIPCHK (hipFree (p));
  return SHK_OK;
}

extern "C" int shk_dev_h2d (shk_dev *d, void *dst, const void *src, size_t bytes)
{
  HIPCHK (hipMemcpyAsync (dst, src, bytes, hipMemcpyHostToDevice, d->stream));
  return SHK_OK;
}

extern "C" int shk_dev_d2h (shk_dev *d, void *dst, const void *src, size_t bytes)
{
  HIPCHK (hipMemcpyAsync (dst, src, bytes, hipMemcpyDeviceToHost, d->stream));
  HIPCHK (hipStreamSynchronize (d->stream));
  return SHK_OK;
}

extern "C" int shk_dev_d2h_async (shk_dev *d, void *dst, const void *src, size_t bytes)
{
  HIPCHK (hipMemcpyAsync (dst, src, bytes, hipMemcpyDeviceToHost, d->stream));
  return SHK_OK;
}

extern "C" int shk_dev_h2d_async (shk_dev *d, void *dst, const void *src, size_t bytes)
{
  HIPCHK (hipMemcpyAsync (dst, src, bytes, hipMemcpyHostToDevice, d->stream));
  return SHK_OK;
}

extern "C" int shk_dev_d2d (shk_dev *d, void *dst, const void *src, size_t bytes)
{
  HIPCHK (hipMemcpyAsync (dst, src, bytes, hipMemcpyDeviceToDevice, d->stream));
  return SHK_OK;
}

extern "C" int shk_dev_sync (shk_dev *d)
{
  HIPCHK (hipStreamSynchronize (d->stream));
  return SHK_OK;
}

extern "C" int shk_device_sync (void)
{
  HIPCHK (hipDeviceSynchronize ());
  return SHK_OK;
}

extern "C" void shk_dev_timing (shk_dev *d, int enable) { d->timing = enable; }

extern "C" double shk_dev_last_ms (shk_dev *d, int which)
{
  if (which < 0 || which > 2) return -1.0;
  return d->last_ms[which];
}

// positions AoS [3np] + radii -> double4 (x, y, z, a)
// (keep_a: leave the radius component as it is)
__global__ void k_pack_pos (const double *__restrict__ pos, const double *__restrict__ a,
                            int np, int keep_a, double4 *__restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double r = keep_a ? out[i].w : (a ? a[i] : 1.0);
  out[i] = make_double4 (pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], r);
}

__global__ void k_img_to_f32 (const double4 *__restrict__ in, int n, float4 *__restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 v = in[i];
  out[i] = make_float4 (float (v.x), float (v.y), float (v.z), float (v.w));
}

static DevParams dev_params (const shk_params &p)
{
  DevParams q;
  q.np = p.np; q.nr = p.nr; q.nk = p.nk; q.poly = p.poly;
  q.xi = p.xi; q.rmax2 = p.rmax2; q.rmin = p.rmin;
  q.self_a = p.self_a; q.self_c = p.self_c; q.self_m = p.self_m;
  q.xiaspi = p.xiaspi; q.xia2 = p.xia2;
  q.gsign = p.gsign;
  return q;
}

static int upload_pos (shk_dev *d, const double *pos, const double *a, int on_device,
                       int keep_a)
{
  const int np = d->prm.np;
  int rc = ensure (d, d->pos, sizeof (double4) * np);
  if (rc) return rc;
  d->pos_gen++;
  // staging for AoS positions and radii
  rc = ensure (d, d->zt, sizeof (double) * 4 * (size_t) np);
  if (rc) return rc;
  double *st = (double *) d->zt.p;
  if (on_device)
    HIPCHK (hipMemcpyAsync (st, pos, sizeof (double) * 3 * np, hipMemcpyDeviceToDevice, d->stream));
  else
    HIPCHK (hipMemcpyAsync (st, pos, sizeof (double) * 3 * np, hipMemcpyHostToDevice, d->stream));
  double *ad = nullptr;
  if (a)
    {
      ad = st + 3 * (size_t) np;
      HIPCHK (hipMemcpyAsync (ad, a, sizeof (double) * np, hipMemcpyHostToDevice, d->stream));
    }
  hipLaunchKernelGGL (k_pack_pos, dim3 ((np + 255) / 256), dim3 (256), 0, d->stream,
                      st, ad, np, keep_a, (double4 *) d->pos.p);
  HIPCHK (hipGetLastError ());
  // the staging buffer is reused by later calls on the same stream: order is kept
  return SHK_OK;
}

static int upload (shk_dev *d, DevBuf &b, const void *src, size_t bytes)
{
  int rc = ensure (d, b, bytes > 0 ? bytes : 16);
  if (rc) return rc;
  if (bytes) HIPCHK (hipMemcpyAsync (b.p, src, bytes, hipMemcpyHostToDevice, d->stream));
  return SHK_OK;
}

__global__ void k_d4_to_f4 (const double4 *__restrict__ in, int n, float4 *__restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 v = in[i];
  out[i] = make_float4 (float (v.x), float (v.y), float (v.z), float (v.w));
}

extern "C" int shk_dev_set_system (shk_dev *d, const shk_params *p, const double *pos,
                                   const double *a, const shk_tables *t)
{
  if (!d || !p || p->np <= 0 || p->version < 0 || p->version > 2)
    {
      set_err ("shk_dev_set_system: invalid arguments");
      return SHK_EARG;
    }
  d->prm = *p;
  d->rc_use = 0;                  // a cached real-space matrix is stale now
  d->ptab_kind = 0;               // so are the phase tables (k basis)
  d->dense_ok = 0;
  int rc = upload_pos (d, pos, a, 0, 0);
  if (rc) return rc;
  if (p->periodic)
    {
      if (!t) { set_err ("shk_dev_set_system: periodic system without tables"); return SHK_EARG; }
      if ((rc = upload (d, d->img, t->img, sizeof (double4) * p->nr))) return rc;
      if ((rc = upload (d, d->cand, t->cand, sizeof (double4) * t->ncand))) return rc;
      if ((rc = upload (d, d->candm, t->candm, sizeof (int4) * t->ncand))) return rc;
      if ((rc = upload (d, d->tbits, t->tbits, sizeof (unsigned) * t->tbits_words))) return rc;
      if ((rc = upload (d, d->kv, t->kv, sizeof (double4) * p->nk))) return rc;
      if ((rc = upload (d, d->ks, t->ks, sizeof (double) * 8 * p->nk))) return rc;
      if (p->nk > 0 && !t->km)
        {
          set_err ("shk_dev_set_system: k table without lattice indices");
          return SHK_EARG;
        }
      if ((rc = upload (d, d->km, t->km, sizeof (int4) * p->nk))) return rc;
      const double twopi = 6.28318530717958647692;
      d->basis.k0x = twopi / p->lx;
      d->basis.k0y = twopi / p->ly;
      d->basis.k0z = twopi / p->lz;
      d->basis.s1 = t->kbasis_s1;
      d->basis.s2 = t->kbasis_s2;
      d->basis.K1 = t->kmax1;
      d->basis.K2 = t->kmax2;
      d->basis.K3 = t->kmax3;
      d->basis.W = (t->kmax1 + 1) + (2 * t->kmax2 + 1) + (2 * t->kmax3 + 1);
      if ((rc = ensure (d, d->candf, sizeof (float4) * (t->ncand > 0 ? t->ncand : 1)))) return rc;
      if (t->ncand > 0)
        {
          hipLaunchKernelGGL (k_d4_to_f4, dim3 ((t->ncand + 255) / 256), dim3 (256), 0, d->stream,
                              (const double4 *) d->cand.p, t->ncand, (float4 *) d->candf.p);
          HIPCHK (hipGetLastError ());
        }
      d->ncand = t->ncand; d->ncore = t->ncore; d->tbits_words = t->tbits_words;
      d->tx = t->tx; d->ty = t->ty; d->tz = t->tz;
      // the image set of (j, i) is that of (i, j) negated when the table
      // holds m and -m alike
      d->tsym = 1;
      {
        const int ny = 2 * t->ty + 1, nz = 2 * t->tz + 1;
        const long ntot = (long) (2 * t->tx + 1) * ny * nz;
        for (long idx = 0; idx < ntot && d->tsym; idx++)
          {
            const long jdx = ntot - 1 - idx;      // (m1, m2, m3) -> (-m1, -m2, -m3)
            const int a = (t->tbits[idx >> 5] >> (idx & 31)) & 1;
            const int b = (t->tbits[jdx >> 5] >> (jdx & 31)) & 1;
            if (a != b) d->tsym = 0;
          }
      }
      d->s2x = t->s2x; d->s3x = t->s3x;
    }
  // the host arrays may be freed by the caller right after return
  HIPCHK (hipStreamSynchronize (d->stream));
  d->have_sys = 1;
  return SHK_OK;
}

extern "C" int shk_dev_set_pos (shk_dev *d, const double *pos, int on_device)
{
  if (!d || !d->have_sys) { set_err ("shk_dev_set_pos: no system"); return SHK_EARG; }
  int rc = upload_pos (d, pos, nullptr, on_device, 1);
  if (rc) return rc;
  if (!on_device) HIPCHK (hipStreamSynchronize (d->stream));
  return SHK_OK;
}

static int dimv (int v) { return v == 0 ? 3 : (v == 1 ? 6 : 11); }

// j-chunking of the pair kernel: enough workgroups to fill 256 CUs
static void pair_grid (int np, int &nit, int &nchunk, int &jchunk)
{
  nit = (np + 15) / 16;
  int njt = (np + 15) / 16;
  nchunk = (2048 + nit - 1) / nit;
  if (nchunk > njt) nchunk = njt;
  if (nchunk < 1) nchunk = 1;
  int jt_per = (njt + nchunk - 1) / nchunk;
  jchunk = jt_per * 16;
  nchunk = (np + jchunk - 1) / jchunk;
}

template <int V, typename T, int KIND>
static void launch_pairs (shk_dev *d, const DevParams &q, const double *x, int nit,
                          int nchunk, int jchunk)
{
  const void *imgp = std::is_same<T, double>::value ? d->img.p : d->imgf.p;
  hipLaunchKernelGGL ((k_pairs<V, T, KIND>), dim3 (nit, nchunk), dim3 (256), 0, d->stream,
                      (const double4 *) d->pos.p, x,
                      (const typename Img<T>::type *) imgp, q, jchunk,
                      (double *) d->partial.p);
}

// reciprocal-space launch shape: enough workgroups for 256 CUs, bounded
// partial-sum buffers
static void recip_shape (int np, int nk, int &sj, int &jper, int &sk, int &kper)
{
  const int nkt = (nk + 255) / 256;
  sj = 4096 / nkt;
  if (sj > 64) sj = 64;
  if (sj > (np + 63) / 64) sj = (np + 63) / 64;
  if (sj < 1) sj = 1;
  jper = (np + sj - 1) / sj;
  sj = (np + jper - 1) / jper;
  const int nit = (np + 255) / 256;
  sk = 4096 / nit;
  if (sk > 128) sk = 128;
  if (sk > (nk + 31) / 32) sk = (nk + 31) / 32;
  if (sk < 1) sk = 1;
  kper = (nk + sk - 1) / sk;
  sk = (nk + kper - 1) / kper;
}

// particles per LDS tile of the table-phase structure-factor kernel (the
// tile region also holds the 3 x 64 x D end-of-block reduction)
template <int V, typename T>
static int sf_tile (const PhaseBasis &b)
{
  const size_t per = sizeof (Cplx<T>) * b.W + sizeof (T) * (ZPad<V>::P + 1);
  int jt = (int) ((40 * 1024) / per) & ~3;
  return jt > 64 ? 64 : (jt < 4 ? 4 : jt);
}

template <int V, typename T, int NPOW>
static size_t sf_lds (const PhaseBasis &b, int jt)
{
  const size_t tile = sizeof (Cplx<T>) * (size_t) jt * b.W
    + sizeof (T) * (size_t) jt * (ZPad<V>::P + 1);
  const size_t red = sizeof (T) * 3 * 64 * Dim<V>::D;
  return tile > red ? tile : red;
}

// splits of the reciprocal kernels: about 8192 waves in flight, at least 16
// particles / 16 k per block
static void recip_shape_tab (int np, int nk, int &sj, int &jper, int &sk, int &kper)
{
  const int nkb = (nk + 63) / 64;
  sj = (8192 + 4 * nkb - 1) / (4 * nkb);
  if (sj > 64) sj = 64;
  if (sj > (np + 15) / 16) sj = (np + 15) / 16;
  if (sj < 1) sj = 1;
  jper = (np + sj - 1) / sj;
  sj = (np + jper - 1) / jper;
  const int nib = (np + 63) / 64;
  sk = (8192 + 4 * nib - 1) / (4 * nib);
  if (sk > 256) sk = 256;
  if (sk > (nk + 15) / 16) sk = (nk + 15) / 16;
  if (sk < 1) sk = 1;
  kper = (nk + sk - 1) / sk;
  sk = (nk + kper - 1) / kper;
}

template <int V, typename T, int NPOW>
static int recip_launch (shk_dev *d, const double *x, int &nrecip, const void *&kpart)
{
  constexpr int D = Dim<V>::D;
  constexpr int NS = NPOW > 1 ? 4 : 2;
  const shk_params &p = d->prm;
  const int nk = p.nk;
  int sj, jper, sk, kper;
  recip_shape_tab (p.np, nk, sj, jper, sk, kper);
  int rc;
  if ((rc = ensure (d, d->kpart, sizeof (T) * (size_t) sj * nk * NS * D))) return rc;
  if ((rc = ensure (d, d->kpq, sizeof (T) * (size_t) nk * NPOW * 2 * D))) return rc;
  if ((rc = ensure (d, d->kback, sizeof (T) * (size_t) sk * p.np * D))) return rc;
  const PhaseBasis &b = d->basis;
  const void *ptab_before = d->ptab.p;
  if ((rc = ensure (d, d->ptab, sizeof (Cplx<T>) * (size_t) p.np * b.W))) return rc;
  const double4 *pos4 = (const double4 *) d->pos.p;
  const double4 *kv = (const double4 *) d->kv.p;
  const int4 *km = (const int4 *) d->km.p;
  // the tables depend on the positions only: kept across products at one
  // configuration (stream order keeps a captured graph consistent, since
  // every positions upload is recorded before the products that follow it)
  const int kind = std::is_same<T, float>::value ? 1 : 2;
  if (d->ptab.p != ptab_before || d->ptab_kind != kind || d->ptab_gen != d->pos_gen)
    {
      hipLaunchKernelGGL ((k_phase_tab<T>), dim3 ((unsigned) (((size_t) p.np * b.W + 255) / 256)),
                          dim3 (256), 0, d->stream, pos4, p.np, b, b.W, (Cplx<T> *) d->ptab.p);
      HIPCHK (hipGetLastError ());
      d->ptab_kind = kind;
      d->ptab_gen = d->pos_gen;
    }
  const int jt = sf_tile<V, T> (b);
  const size_t lds = sf_lds<V, T, NPOW> (b, jt);
  hipLaunchKernelGGL ((k_recip_sf_tab<V, T, NPOW>), dim3 ((nk + 63) / 64, sj), dim3 (256), lds,
                      d->stream, pos4, x, km, (const Cplx<T> *) d->ptab.p, b, p.np, nk, jper, jt,
                      (T *) d->kpart.p);
  HIPCHK (hipGetLastError ());
  hipLaunchKernelGGL ((k_recip_blocks<V, T, NPOW>),
                      dim3 ((unsigned) ((recip_blocks_threads<NPOW> (nk) + 127) / 128)), dim3 (128), 0,
                      d->stream, kv, (const double *) d->ks.p, nk, sj, p.gsign,
                      (const T *) d->kpart.p, (T *) d->kpq.p);
  HIPCHK (hipGetLastError ());
  hipLaunchKernelGGL ((k_recip_back_rows<V, T, NPOW>), dim3 ((p.np + 63) / 64, sk), dim3 (256),
                      0, d->stream, pos4, kv, km, b, (const T *) d->kpq.p, p.np, nk, kper,
                      (T *) d->kback.p);
  HIPCHK (hipGetLastError ());
  nrecip = sk;
  kpart = d->kback.p;
  return SHK_OK;
}

template <int V, typename T>
static int atimes_vt (shk_dev *d, const double *x, double *y)
{
  constexpr int D = Dim<V>::D;
  const shk_params &p = d->prm;
  DevParams q = dev_params (p);
  const double4 *pos4 = (const double4 *) d->pos.p;
  int rc, nreal = 0, nrecip = 0;
  const void *kpart = nullptr;
  if (d->timing) HIPCHK (hipEventRecord (d->ev[0], d->stream));
  const int want_prec = std::is_same<T, float>::value ? SHK_MIXED : SHK_FP64;
  if (p.periodic && d->rc_use && d->rc_prec == want_prec)
    {
      // real space from the cached dense matrix of these positions
      const size_t ND = (size_t) p.np * D;
      nreal = 1;
      if ((rc = ensure (d, d->partial, sizeof (double) * ND))) return rc;
      if ((rc = ensure (d, d->rz, sizeof (T) * ND))) return rc;
      const dim3 grows ((unsigned) ((ND + 3) / 4));
      const bool w16 = ND % Vec16<T>::W == 0;
      if (V < 2 && w16)
        // F, FT: no S transform, the kernel reads x itself
        hipLaunchKernelGGL ((k_real_gemv16<T, true>), grows, dim3 (256), 0, d->stream,
                            (const T *) d->rcache.p, (const void *) x, (int) ND,
                            (double *) d->partial.p);
      else
        {
          hipLaunchKernelGGL ((k_z_prep<V, T>), dim3 ((unsigned) ((p.np + 255) / 256)),
                              dim3 (256), 0, d->stream, x, p.np, (T *) d->rz.p);
          if (w16)
            hipLaunchKernelGGL ((k_real_gemv16<T, false>), grows, dim3 (256), 0, d->stream,
                                (const T *) d->rcache.p, (const void *) d->rz.p, (int) ND,
                                (double *) d->partial.p);
          else
            hipLaunchKernelGGL ((k_real_gemv<T>), grows, dim3 (256), 0, d->stream,
                                (const T *) d->rcache.p, (const T *) d->rz.p, (int) ND,
                                (double *) d->partial.p);
        }
      HIPCHK (hipGetLastError ());
      if (d->timing) HIPCHK (hipEventRecord (d->ev[1], d->stream));
      if (p.nk > 0)
        {
          rc = p.poly ? recip_launch<V, T, 5> (d, x, nrecip, kpart)
                      : recip_launch<V, T, 1> (d, x, nrecip, kpart);
          if (rc) return rc;
        }
    }
  else if (p.periodic)
    {
      const int nj = (p.np + 63) / 64;
      nreal = nj;
      if ((rc = ensure (d, d->partial, sizeof (double) * (size_t) nj * p.np * D))) return rc;
      RealParams rp;
      rp.np = p.np; rp.ncand = d->ncand; rp.ncore = d->ncore; rp.poly = p.poly;
      rp.xi = p.xi; rp.rmax2 = p.rmax2; rp.rmin = p.rmin;
      rp.lx = p.lx; rp.ly = p.ly; rp.lz = p.lz; rp.s2x = d->s2x; rp.s3x = d->s3x;
      rp.tx = d->tx; rp.ty = d->ty; rp.tz = d->tz; rp.tbits_words = d->tbits_words;
      const long waves = (long) p.np * nj;
      const void *cand = std::is_same<T, double>::value ? d->cand.p : d->candf.p;
      hipLaunchKernelGGL ((k_real_compact<V, T>), dim3 ((unsigned) ((waves + 3) / 4)), dim3 (256),
                          0, d->stream, pos4, x, (const typename Vec4<T>::type *) cand,
                          (const int4 *) d->candm.p, (const unsigned *) d->tbits.p, rp,
                          (double *) d->partial.p);
      HIPCHK (hipGetLastError ());
      if (d->timing) HIPCHK (hipEventRecord (d->ev[1], d->stream));
      if (p.nk > 0)
        {
          rc = p.poly ? recip_launch<V, T, 5> (d, x, nrecip, kpart)
                      : recip_launch<V, T, 1> (d, x, nrecip, kpart);
          if (rc) return rc;
        }
    }
  else
    {
      int nit, nchunk, jchunk;
      pair_grid (p.np, nit, nchunk, jchunk);
      nreal = nchunk;
      if ((rc = ensure (d, d->partial, sizeof (double) * (size_t) nchunk * p.np * D))) return rc;
      launch_pairs<V, T, 1> (d, q, x, nit, nchunk, jchunk);
      HIPCHK (hipGetLastError ());
      if (d->timing) HIPCHK (hipEventRecord (d->ev[1], d->stream));
    }
  hipLaunchKernelGGL ((k_combine<V, T>),
                      dim3 ((unsigned) (((size_t) p.np * D * COMBINE_LANES + 255) / 256)),
                      dim3 (256), 0, d->stream,
                      pos4, x, (const double *) d->partial.p, nreal, (const T *) kpart, nrecip, q,
                      p.periodic, y);
  HIPCHK (hipGetLastError ());
  if (d->timing)
    {
      HIPCHK (hipEventRecord (d->ev[2], d->stream));
      HIPCHK (hipEventSynchronize (d->ev[2]));
      float ta = 0, tb = 0;
      HIPCHK (hipEventElapsedTime (&ta, d->ev[0], d->ev[1]));
      HIPCHK (hipEventElapsedTime (&tb, d->ev[1], d->ev[2]));
      d->last_ms[0] = ta;
      d->last_ms[1] = tb;
      d->last_ms[2] = ta + tb;
    }
  return SHK_OK;
}

template <int V>
static int atimes_v (shk_dev *d, const double *x, double *y, int prec)
{
  return prec == SHK_MIXED ? atimes_vt<V, float> (d, x, y) : atimes_vt<V, double> (d, x, y);
}

// ---- block of right-hand sides (periodic) --------------------------------

template <int V, typename T, int NPOW>
static int recip_block_launch (shk_dev *d, const T *XT, int B, int &nrecip,
                               const void *&kpart)
{
  constexpr int D = Dim<V>::D;
  constexpr int NS = NPOW > 1 ? 4 : 2;
  const shk_params &p = d->prm;
  const int nk = p.nk, nbg = (B + 7) / 8;
  int sj = 1024 / (((nk + 31) / 32) * nbg);
  if (sj > 8) sj = 8;
  if (sj < 1) sj = 1;
  int jper = (p.np + sj - 1) / sj;
  sj = (p.np + jper - 1) / jper;
  int sk = 1024 / (((p.np + 31) / 32) * nbg);
  if (sk > 16) sk = 16;
  if (sk > (nk + 31) / 32) sk = (nk + 31) / 32;
  if (sk < 1) sk = 1;
  int kper = (nk + sk - 1) / sk;
  sk = (nk + kper - 1) / kper;
  int rc;
  if ((rc = ensure (d, d->kpart, sizeof (T) * (size_t) sj * nk * B * NS * D))) return rc;
  if ((rc = ensure (d, d->kpq, sizeof (T) * (size_t) nk * B * NPOW * 2 * D))) return rc;
  if ((rc = ensure (d, d->kback, sizeof (T) * (size_t) sk * p.np * D * B))) return rc;
  const double4 *pos4 = (const double4 *) d->pos.p;
  const double4 *kv = (const double4 *) d->kv.p;
  hipLaunchKernelGGL ((k_recip_sf_block<V, T, NPOW>), dim3 ((nk + 31) / 32, nbg, sj), dim3 (256),
                      0, d->stream, pos4, XT, B, kv, p.np, nk, jper, (T *) d->kpart.p);
  HIPCHK (hipGetLastError ());
  const size_t nkb = (size_t) nk * NPOW * B;
  hipLaunchKernelGGL ((k_recip_blk_block<V, T, NPOW>), dim3 ((unsigned) ((nkb + 127) / 128)),
                      dim3 (128), 0, d->stream, kv, (const double *) d->ks.p, nk, B, sj, p.gsign,
                      (const T *) d->kpart.p, (T *) d->kpq.p);
  HIPCHK (hipGetLastError ());
  hipLaunchKernelGGL ((k_recip_back_block<V, T, NPOW>), dim3 ((p.np + 31) / 32, nbg, sk),
                      dim3 (256), 0, d->stream, pos4, kv, (const T *) d->kpq.p, p.np, nk, B, kper,
                      (T *) d->kback.p);
  HIPCHK (hipGetLastError ());
  nrecip = sk;
  kpart = d->kback.p;
  return SHK_OK;
}

// number of splits s (1 <= s <= smax) of a launch with `base` workgroups
// per split when `slots` workgroups are resident at once: the smallest s
// whose grid fills at least 90 % of the rounds of slots it takes (a last
// round left mostly empty costs a whole round), else the best filling
static int balanced_splits (long base, long slots, int smax)
{
  int best = smax;
  double best_eff = -1.0;
  for (int s = 1; s <= smax; s++)
    {
      const long w = base * s;
      const long rounds = (w + slots - 1) / slots;
      const double eff = (double) w / (double) (rounds * slots);
      if (eff >= 0.9) return s;
      if (eff > best_eff)
        {
          best_eff = eff;
          best = s;
        }
    }
  return best;
}

// FP32 matrix-core reciprocal sums for a block (MIXED precision); Zt rows
// of zs floats, npad = np rounded up to 64 rows (zero past np)
template <int V, int NPOW>
static int recip_block_launch_mfma (shk_dev *d, const float *Zt, int zs, int npad, int B,
                                    int &nrecip, const void *&kpart)
{
  constexpr int D = Dim<V>::D;
  constexpr int NH = NPOW > 1 ? 2 : 1;
  constexpr int PR = NPOW * 2;
  const shk_params &p = d->prm;
  const int nk = p.nk, np = p.np;
  const int cz = NH * D * B, C = D * B;
  const PhaseBasis &b = d->basis;
  // LDS: two stages of the structure factors' tables and z slices (LD = 96,
  // 128 or 160 >= W entries per particle), the back-projection's W x 64
  // tables (at most 80 KB), the 64 KB end-of-block reduction aliased over
  // them
  const int LD = b.W <= 96 ? 96 : (b.W <= 128 ? 128 : 160);
  if (b.W > 160)
    {
      set_err ("batched MIXED reciprocal sum: phase tables of %d entries (more than 160: "
               "k_max spans more than 31 lattice steps) exceed LDS",
               b.W);
      return SHK_EARG;
    }
  size_t lds_sf = 2 * (size_t) (LD == 96 ? SfStage<96>::BYTES
                                         : (LD == 128 ? SfStage<128>::BYTES : SfStage<160>::BYTES)),
         lds_bk = (size_t) b.W * 512;
  if (lds_sf < (size_t) MF_RED_BYTES) lds_sf = MF_RED_BYTES;
  if (lds_bk < (size_t) MF_RED_BYTES) lds_bk = MF_RED_BYTES;
  // structure factors: splits of j, multiples of 64, filling the rounds of
  // two resident workgroups per CU
  const int nkb = (nk + 63) / 64, ncz = (cz + 63) / 64;
  const long slots = 2L * d->ncu;
  int sj = balanced_splits ((long) nkb * ncz, slots, std::max (1, std::min (npad / 64, 64)));
  int jper = (npad + sj - 1) / sj;
  jper = (jper + 63) & ~63;
  sj = (npad + jper - 1) / jper;
  // back-projection: splits of kper k (a multiple of 4); P, Q rows padded
  // with zeros to sk kper, columns to a multiple of 64
  const int nib = (np + 63) / 64, ncc = (C + 63) / 64, Cp = ncc * 64;
  int sk = balanced_splits ((long) nib * ncc, slots, std::max (1, std::min ((nk + 63) / 64, 64)));
  int kper = (nk + sk - 1) / sk;
  kper = (kper + 3) & ~3;
  sk = (nk + kper - 1) / kper;
  const int nkpad = sk * kper;
  int rc;
  if ((rc = ensure (d, d->kpart, sizeof (float) * (size_t) sj * nk * 2 * cz))) return rc;
  if ((rc = ensure (d, d->kpq, sizeof (float) * (size_t) nkpad * PR * Cp))) return rc;
  if ((rc = ensure (d, d->kback, sizeof (float) * (size_t) sk * np * C))) return rc;
  if ((rc = ensure (d, d->ptab, sizeof (Cplx<float>) * (size_t) npad * (b.W + LD)))) return rc;
  d->ptab_kind = 0;               // the single-column tables are overwritten
  const double4 *pos4 = (const double4 *) d->pos.p;
  const double4 *kv = (const double4 *) d->kv.p;
  const int4 *km = (const int4 *) d->km.p;
  Cplx<float> *tab = (Cplx<float> *) d->ptab.p, *tabT = tab + (size_t) npad * LD;
  const dim3 gt ((unsigned) (((size_t) npad * b.W + 255) / 256));
  if (LD == 96)
    hipLaunchKernelGGL ((k_phase_tab2<96>), gt, dim3 (256), 0, d->stream, pos4, np, npad, b, tab,
                        tabT);
  else if (LD == 128)
    hipLaunchKernelGGL ((k_phase_tab2<128>), gt, dim3 (256), 0, d->stream, pos4, np, npad, b, tab,
                        tabT);
  else
    hipLaunchKernelGGL ((k_phase_tab2<160>), gt, dim3 (256), 0, d->stream, pos4, np, npad, b, tab,
                        tabT);
  HIPCHK (hipGetLastError ());
  static bool attr_sf = false, attr_bk = false;
  if (!attr_sf)
    {
      HIPCHK (hipFuncSetAttribute ((const void *) k_sf_mfma<96>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 80 * 1024));
      HIPCHK (hipFuncSetAttribute ((const void *) k_sf_mfma<128>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 80 * 1024));
      HIPCHK (hipFuncSetAttribute ((const void *) k_sf_mfma<160>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 80 * 1024));
      attr_sf = true;
    }
  if (!attr_bk)
    {
      HIPCHK (hipFuncSetAttribute ((const void *) k_back_mfma<NPOW>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 80 * 1024));
      attr_bk = true;
    }
  if (LD == 96)
    hipLaunchKernelGGL ((k_sf_mfma<96>), dim3 (ncz, nkb, sj), dim3 (256), lds_sf, d->stream, Zt,
                        cz, zs, km, tab, npad, b, nk, jper, (float *) d->kpart.p);
  else if (LD == 128)
    hipLaunchKernelGGL ((k_sf_mfma<128>), dim3 (ncz, nkb, sj), dim3 (256), lds_sf, d->stream, Zt,
                        cz, zs, km, tab, npad, b, nk, jper, (float *) d->kpart.p);
  else
    hipLaunchKernelGGL ((k_sf_mfma<160>), dim3 (ncz, nkb, sj), dim3 (256), lds_sf, d->stream, Zt,
                        cz, zs, km, tab, npad, b, nk, jper, (float *) d->kpart.p);
  HIPCHK (hipGetLastError ());
  const size_t nkB = (size_t) nkpad * NPOW * B;
  hipLaunchKernelGGL ((k_blk_mfma<V, NPOW>), dim3 ((unsigned) ((nkB + 127) / 128)), dim3 (128),
                      0, d->stream, kv, (const double *) d->ks.p, nk, nkpad, B, sj, Cp,
                      (const float *) d->kpart.p, (float *) d->kpq.p);
  HIPCHK (hipGetLastError ());
  hipLaunchKernelGGL ((k_back_mfma<NPOW>), dim3 (ncc, nib, sk), dim3 (256), lds_bk, d->stream,
                      pos4, km, tabT, npad, b, (const float *) d->kpq.p, C, Cp, np, nk, kper,
                      (float *) d->kback.p);
  HIPCHK (hipGetLastError ());
  nrecip = sk;
  kpart = d->kback.p;
  return SHK_OK;
}

// C = A B in FP64 on the matrix cores (k_dgemm), row-major; K split so that
// about 512 workgroups run (slices of at least 256), the slices summed in
// order (k_dgemm_sum) through d->gpart
static int dgemm (shk_dev *d, int M, int N, int K, const double *A, int lda, const double *B,
                  int ldb, double *C, int ldc)
{
  const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
  int sz = 1;
  while (tiles * sz < 512 && K / (2 * sz) >= 256) sz *= 2;
  int kper = ((K + sz - 1) / sz + 15) & ~15;
  sz = (K + kper - 1) / kper;
  const dim3 g ((unsigned) ((N + 63) / 64), (unsigned) ((M + 63) / 64), (unsigned) sz);
  if (sz == 1)
    hipLaunchKernelGGL (k_dgemm, g, dim3 (256), 0, d->stream, A, lda, B, ldb, C, ldc, (size_t) 0,
                        M, N, K, kper);
  else
    {
      const size_t cz = (size_t) M * N;
      int rc = ensure (d, d->gpart, sizeof (double) * cz * sz);
      if (rc) return rc;
      double *P = (double *) d->gpart.p;
      hipLaunchKernelGGL (k_dgemm, g, dim3 (256), 0, d->stream, A, lda, B, ldb, P, N, cz, M, N, K,
                          kper);
      hipLaunchKernelGGL (k_dgemm_sum, dim3 ((unsigned) ((cz + 255) / 256)), dim3 (256), 0,
                          d->stream, (const double *) P, N, cz, sz, M, N, C, ldc);
    }
  HIPCHK (hipGetLastError ());
  return SHK_OK;
}

// the phase matrix of the half-space k: E[k][j] = cos k.x_j, E[nk + k][j]
// = sin k.x_j (2 nk x np), or transposed, ET[j][k], ET[j][nk + k]
// (np x 2 nk); FP64 sincos of the FP64 phase
template <bool TRANS>
__global__ void k_phase_E (const double4 *__restrict__ pos, const double4 *__restrict__ kv,
                           int nk, int np, double *__restrict__ E)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) nk * np) return;
  int k, j;
  if (TRANS) { j = (int) (t / nk); k = (int) (t - (size_t) j * nk); }
  else { k = (int) (t / np); j = (int) (t - (size_t) k * np); }
  const double4 kk = kv[k], pj = pos[j];
  double sn, cs;
  sincos (kk.x * pj.x + kk.y * pj.y + kk.z * pj.z, &sn, &cs);
  if (TRANS)
    {
      E[(size_t) j * 2 * nk + k] = cs;
      E[(size_t) j * 2 * nk + nk + k] = sn;
    }
  else
    {
      E[(size_t) k * np + j] = cs;
      E[(size_t) (nk + k) * np + j] = sn;
    }
}

// P, Q of the equal-sphere k-blocks from SF = [C; S] (2 nk x D B rows,
// column q B + b) into PQ = [P; Q] in the same layout; thread per (k, b)
template <int V>
__global__ __launch_bounds__ (128) void
k_blk64 (const double4 *__restrict__ kv, const double *__restrict__ ks, int nk, int B,
         const double *__restrict__ SF, double *__restrict__ PQ)
{
  constexpr int D = Dim<V>::D;
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) nk * B) return;
  const int k = (int) (t / B), b = (int) (t - (size_t) k * B);
  const size_t C = (size_t) D * B;
  double CS[2 * D];
#pragma unroll
  for (int q = 0; q < D; q++)
    {
      CS[q] = SF[(size_t) k * C + (size_t) q * B + b];
      CS[D + q] = SF[(size_t) (nk + k) * C + (size_t) q * B + b];
    }
  double P[D], Q[D];
  kblock_pq_mono<V> (kv[k], ks + (size_t) k * 8, CS, P, Q);
#pragma unroll
  for (int q = 0; q < D; q++)
    {
      PQ[(size_t) k * C + (size_t) q * B + b] = P[q];
      PQ[(size_t) (nk + k) * C + (size_t) q * B + b] = Q[q];
    }
}

// FP64 reciprocal sums of a block of equal spheres on the matrix cores: the
// structure factors SF = E Z (2 nk x D B), the k-blocks, the
// back-projection Y = E^T PQ (np x D B, the layout of one reciprocal
// partial); Z = XT (np rows of D B)
template <int V>
static int recip_block_launch_dgemm (shk_dev *d, const double *Z, int B, int &nrecip,
                                     const void *&kpart)
{
  constexpr int D = Dim<V>::D;
  const shk_params &p = d->prm;
  const int nk = p.nk, np = p.np, C = D * B;
  int rc;
  if ((rc = ensure (d, d->eph, sizeof (double) * 4 * (size_t) nk * np))) return rc;
  if ((rc = ensure (d, d->kpart, sizeof (double) * 2 * (size_t) nk * C))) return rc;
  if ((rc = ensure (d, d->kpq, sizeof (double) * 2 * (size_t) nk * C))) return rc;
  if ((rc = ensure (d, d->kback, sizeof (double) * (size_t) np * C))) return rc;
  const double4 *pos4 = (const double4 *) d->pos.p;
  const double4 *kv = (const double4 *) d->kv.p;
  double *E = (double *) d->eph.p, *ET = E + 2 * (size_t) nk * np;
  const dim3 ge ((unsigned) (((size_t) nk * np + 255) / 256));
  hipLaunchKernelGGL ((k_phase_E<false>), ge, dim3 (256), 0, d->stream, pos4, kv, nk, np, E);
  hipLaunchKernelGGL ((k_phase_E<true>), ge, dim3 (256), 0, d->stream, pos4, kv, nk, np, ET);
  HIPCHK (hipGetLastError ());
  double *SF = (double *) d->kpart.p, *PQ = (double *) d->kpq.p, *Y = (double *) d->kback.p;
  if ((rc = dgemm (d, 2 * nk, C, np, E, np, Z, C, SF, C))) return rc;
  hipLaunchKernelGGL ((k_blk64<V>), dim3 ((unsigned) (((size_t) nk * B + 127) / 128)), dim3 (128),
                      0, d->stream, kv, (const double *) d->ks.p, nk, B, (const double *) SF, PQ);
  HIPCHK (hipGetLastError ());
  if ((rc = dgemm (d, np, C, 2 * nk, ET, 2 * nk, PQ, C, Y, C))) return rc;
  nrecip = 1;
  kpart = d->kback.p;
  return SHK_OK;
}

template <int V, typename T>
static int atimes_block_vt (shk_dev *d, const double *x, double *y, int B, bool dense)
{
  constexpr int D = Dim<V>::D;
  constexpr int JW = 256;
  const shk_params &p = d->prm;
  DevParams q = dev_params (p);
  const double4 *pos4 = (const double4 *) d->pos.p;
  int rc, nrecip = 0;
  const void *kpart = nullptr;
  constexpr bool MF = std::is_same<T, float>::value;
  const int NH = MF && p.poly ? 2 : 1;
  // rows of the transposed block: 16-byte multiples on the FP32 path
  const int zs = MF ? (NH * D * B + 3) & ~3 : D * B;
  const int npad = MF ? (p.np + 63) & ~63 : p.np;     // zero rows past np
  if ((rc = ensure (d, d->xs, sizeof (T) * (size_t) npad * zs))) return rc;
  // real-space splits: JW sources per wave (FP64), or for the matrix-core
  // path (MIXED) one wave per 16 targets x jper sources x 16 CT columns,
  // about 2048 waves (two per SIMD)
  // 16 CT columns per wave: the smallest power of two covering B, at most
  // 2 / 4 / 8 tiles for FTS / FT / F (the accumulators of D row tiles)
  constexpr int CTMAX = V == 2 ? 2 : (V == 1 ? 4 : 8);
  int CT = 1;
  while (CT < CTMAX && 16 * CT < B) CT *= 2;
  const int nib = (p.np + 15) / 16, ncg = (B + 16 * CT - 1) / (16 * CT);
  int nj = (p.np + JW - 1) / JW, jper = 0;
  // FTS above 32 columns: workgroups of NWG producer waves (image sums) and
  // NWG consumer waves of 32 columns each (k_real_mfma_pc), 32 NWG columns
  // per workgroup; one workgroup per CU
  const int NWG = B > 64 ? 4 : 2;
  const bool wg = MF && V == 2 && B > 32 && d->ncand <= REAL_PC_CAND;
  const int ncw = wg ? (B + 32 * NWG - 1) / (32 * NWG) : ncg;
  // equal spheres with a symmetric image table: each pair's moments once
  // (k_real_moments, both orders into momt), then NWG waves of 32 columns
  // per workgroup contract them (k_real_mfma_mom); momt holds
  // NMOM npad^2 floats, used up to 2 GiB
  const bool mom = wg && !p.poly && d->tsym
                   && (double) Dim<V>::NMOM * npad * npad * 4.0 <= 2.0 * 1024 * 1024 * 1024;
  // the same for FP64 blocks of 16 columns or more: moments in FP64, then
  // NC64 waves of 16 columns per workgroup on the FP64 matrix cores
  // (k_real_mfma_mom64); momt holds NMOM npm^2 doubles, used up to 2 GiB
  const int npm = (p.np + 63) & ~63;
  const int NC64 = B > 64 ? 8 : (B > 32 ? 4 : 2);
  const int ncw64 = (B + 16 * NC64 - 1) / (16 * NC64);
  const bool mom64 = !MF && V == 2 && B >= 16 && !p.poly && d->tsym && p.periodic
                     && d->ncand <= REAL_PC_CAND
                     && (double) Dim<V>::NMOM * npm * npm * 8.0 <= 2.0 * 1024 * 1024 * 1024;
  if (MF || mom64)
    {
      // about 2048 waves, or two workgroups per CU of the producer/consumer form
      const int wpb = mom64 ? NC64 * ncw64 : (mom ? NWG * ncw : (wg ? 4 * ncw : ncg));
      nj = (2048 + nib * wpb - 1) / (nib * wpb);
      if (nj > (p.np + 15) / 16) nj = (p.np + 15) / 16;
      if (nj < 1) nj = 1;
      jper = (p.np + nj - 1) / nj;
      // multiples of 4 sources (8 for the moment contractions' chunks)
      const int ja = (mom || mom64) ? 7 : 3;
      jper = (jper + ja) & ~ja;
      nj = (p.np + jper - 1) / jper;
    }
  if ((rc = ensure (d, d->partial, sizeof (T) * (size_t) nj * p.np * D * B))) return rc;
  T *XT = (T *) d->xs.p;
  const size_t nb = (size_t) p.np * B;
  if (!MF)
    hipLaunchKernelGGL ((k_block_prep<V, T>), dim3 ((unsigned) ((nb + 255) / 256)), dim3 (256),
                        0, d->stream, x, p.np, B, XT);
  else if (NH == 2)
    hipLaunchKernelGGL ((k_block_prep_f<V, 2>), dim3 ((unsigned) (((size_t) npad * B + 255) / 256)),
                        dim3 (256), 0, d->stream, x, pos4, p.np, npad, B, zs, (float *) XT);
  else
    hipLaunchKernelGGL ((k_block_prep_f<V, 1>), dim3 ((unsigned) (((size_t) npad * B + 255) / 256)),
                        dim3 (256), 0, d->stream, x, pos4, p.np, npad, B, zs, (float *) XT);
  HIPCHK (hipGetLastError ());
  RealParams rp;
  rp.np = p.np; rp.ncand = d->ncand; rp.ncore = d->ncore; rp.poly = p.poly;
  rp.xi = p.xi; rp.rmax2 = p.rmax2; rp.rmin = p.rmin;
  rp.lx = p.lx; rp.ly = p.ly; rp.lz = p.lz; rp.s2x = d->s2x; rp.s3x = d->s3x;
  rp.tx = d->tx; rp.ty = d->ty; rp.tz = d->tz; rp.tbits_words = d->tbits_words;
  const long waves = (long) p.np * nj;
  const void *cand = std::is_same<T, double>::value ? d->cand.p : d->candf.p;
  // the reciprocal part reads only the prepared block and the positions and
  // writes its own buffers: FP64 blocks run it on the side stream beside the
  // real-space part (N = 1024 FTS, 128 columns: 1.33 -> 1.30 ms per call);
  // MIXED blocks keep the two apart, the matrix-core structure factors
  // slow the real-space kernels' balanced rounds more than they gain
  // (0.68 -> 0.80 ms).  Not when timed (the events then bracket each part
  // alone), nor with the dense real-space matrix (its FP64 GEMM shares the
  // split-K buffer with the reciprocal GEMMs), nor while carving from the
  // per-GPU scratch
  const bool fork = !MF && p.nk > 0 && d->overlap && !d->timing && !dense && !d->carve_base;
  if (fork)
    {
      HIPCHK (hipEventRecord (d->evf, d->stream));
      HIPCHK (hipStreamWaitEvent (d->side, d->evf, 0));
      hipStream_t s_main = d->stream;
      d->stream = d->side;
      if (MF)
        rc = p.poly ? recip_block_launch_mfma<V, 5> (d, (const float *) XT, zs, npad, B, nrecip,
                                                      kpart)
                    : recip_block_launch_mfma<V, 1> (d, (const float *) XT, zs, npad, B, nrecip,
                                                      kpart);
      else if (!p.poly)
        rc = recip_block_launch_dgemm<V> (d, (const double *) XT, B, nrecip, kpart);
      else
        rc = recip_block_launch<V, T, 5> (d, XT, B, nrecip, kpart);
      d->stream = s_main;
      if (rc) return rc;
      HIPCHK (hipEventRecord (d->evj, d->side));
    }
  // timing: ev[0] -> ev[1] brackets the real-space kernel alone
  if (d->timing) HIPCHK (hipEventRecord (d->ev[0], d->stream));
  if (dense)
    {
      // real space from the cached dense matrix: Y (ND x B, row-major) =
      // Mr (ND x ND, row-major) Z (ND x B, row-major).  MIXED: split-K
      // matrix-core GEMM (k_cache_gemm), one partial per K slice; FP64:
      // column-major Y^T = Z^T Mr^T, one rocBLAS GEMM without atomics
      const int ND = p.np * D;
      const T *Z = XT;
      if constexpr (MF)
        {
          if ((rc = ensure (d, d->rz, sizeof (T) * (size_t) ND * B))) return rc;
          hipLaunchKernelGGL ((k_block_prep<V, T>), dim3 ((unsigned) ((nb + 255) / 256)),
                              dim3 (256), 0, d->stream, x, p.np, B, (T *) d->rz.p);
          HIPCHK (hipGetLastError ());
          Z = (const T *) d->rz.p;
          // K slices of a multiple of 64 so that about 768 workgroups run
          const int ct = (B + 15) / 16;
          const int rows_wg = 64 * gemm_rt (ct > 4 ? 8 : 4);
          const int nrg = (ND + rows_wg - 1) / rows_wg;
          int ns = (768 + nrg / 2) / nrg;
          if (ns > (ND + 255) / 256) ns = (ND + 255) / 256;
          if (ns < 1) ns = 1;
          int ks = (ND + ns - 1) / ns;
          ks = (ks + 63) & ~63;
          ns = (ND + ks - 1) / ks;
          if ((rc = ensure (d, d->partial, sizeof (T) * (size_t) ns * ND * B))) return rc;
          const dim3 g ((unsigned) nrg, (unsigned) ns);
          const bool vec = ND % 4 == 0;
          const float *mr = (const float *) d->rcache.p;
          float *pf = (float *) d->partial.p;
#define CACHE_GEMM(CT)                                                                  \
  do {                                                                                \
      if (vec)                                                                        \
        hipLaunchKernelGGL ((k_cache_gemm<CT, true>), g, dim3 (256), 0, d->stream, mr, \
                            (const float *) Z, ND, B, ks, pf);                        \
      else                                                                            \
        hipLaunchKernelGGL ((k_cache_gemm<CT, false>), g, dim3 (256), 0, d->stream,   \
                            mr, (const float *) Z, ND, B, ks, pf);                    \
  } while (0)
          if (ct <= 1) CACHE_GEMM (1);
          else if (ct <= 2) CACHE_GEMM (2);
          else if (ct <= 4) CACHE_GEMM (4);
          else CACHE_GEMM (8);
#undef CACHE_GEMM
          HIPCHK (hipGetLastError ());
          nj = ns;
        }
      else
        {
          // FP64: Y = Mr Z on the FP64 matrix cores (rows of Mr are the
          // output components, Z the transposed block)
          if ((rc = dgemm (d, (int) ND, B, (int) ND, (const double *) d->rcache.p, (int) ND,
                           (const double *) Z, B, (double *) d->partial.p, B)))
            return rc;
          nj = 1;
        }
    }
  else if (mom64)
    {
      const size_t mbytes = sizeof (double) * Dim<V>::NMOM * (size_t) npm * npm;
      if ((rc = ensure (d, d->momt, mbytes))) return rc;
      double *mt = (double *) d->momt.p;
      const int nt = npm / 16;
      hipLaunchKernelGGL ((k_real_moments<V, double>), dim3 (nt, nt), dim3 (256), 0, d->stream,
                          pos4, (const double4 *) d->cand.p, (const int4 *) d->candm.p,
                          (const unsigned *) d->tbits.p, rp, npm, mt);
      HIPCHK (hipGetLastError ());
      const dim3 g (nib, nj, ncw64);
      double *pd = (double *) d->partial.p;
      const double *xd = (const double *) XT;
      const int zd = D * B;
      if (NC64 == 8)
        hipLaunchKernelGGL ((k_real_mfma_mom64<V, 8>), g, dim3 (512), 0, d->stream,
                            (const double *) mt, npm, xd, B, zd, p.np, jper, pd);
      else if (NC64 == 4)
        hipLaunchKernelGGL ((k_real_mfma_mom64<V, 4>), g, dim3 (256), 0, d->stream,
                            (const double *) mt, npm, xd, B, zd, p.np, jper, pd);
      else
        hipLaunchKernelGGL ((k_real_mfma_mom64<V, 2>), g, dim3 (128), 0, d->stream,
                            (const double *) mt, npm, xd, B, zd, p.np, jper, pd);
    }
  else if (mom)
    {
      const size_t mbytes = sizeof (float) * Dim<V>::NMOM * (size_t) npad * npad;
      if ((rc = ensure (d, d->momt, mbytes))) return rc;
      float *mt = (float *) d->momt.p;
      const int nt = npad / 16;
      hipLaunchKernelGGL ((k_real_moments<V, float>), dim3 (nt, nt), dim3 (256), 0, d->stream, pos4,
                          (const float4 *) d->candf.p, (const int4 *) d->candm.p,
                          (const unsigned *) d->tbits.p, rp, npad, mt);
      HIPCHK (hipGetLastError ());
      const dim3 g (nib, nj, ncw);
      float *pf = (float *) d->partial.p;
      if (NWG == 4)
        hipLaunchKernelGGL ((k_real_mfma_mom<V, 4>), g, dim3 (256), 0, d->stream,
                            (const float *) mt, npad, (const float *) XT, B, zs, p.np, jper, pf);
      else
        hipLaunchKernelGGL ((k_real_mfma_mom<V, 2>), g, dim3 (128), 0, d->stream,
                            (const float *) mt, npad, (const float *) XT, B, zs, p.np, jper, pf);
    }
  else if (wg)
    {
      const dim3 g (nib, nj, ncw);
      const float4 *cf = (const float4 *) d->candf.p;
      const int4 *cm = (const int4 *) d->candm.p;
      const unsigned *tb = (const unsigned *) d->tbits.p;
      const float *xf = (const float *) XT;
      float *pf = (float *) d->partial.p;
      if (NWG == 4)
        hipLaunchKernelGGL ((k_real_mfma_pc<V, 4, 4>), g, dim3 (512), 0, d->stream, pos4, xf, B,
                            zs, cf, cm, tb, rp, jper, pf);
      else
        hipLaunchKernelGGL ((k_real_mfma_pc<V, 2, 2>), g, dim3 (256), 0, d->stream, pos4, xf, B,
                            zs, cf, cm, tb, rp, jper, pf);
    }
  else if (MF)
    {
      const dim3 g (nib, nj, ncg);
      const float4 *cf = (const float4 *) d->candf.p;
      const int4 *cm = (const int4 *) d->candm.p;
      const unsigned *tb = (const unsigned *) d->tbits.p;
      const float *xf = (const float *) XT;
      float *pf = (float *) d->partial.p;
      if (CT == 1)
        hipLaunchKernelGGL ((k_real_mfma<V, 1>), g, dim3 (64), 0, d->stream, pos4, xf, B, zs, cf,
                            cm, tb, rp, jper, pf);
      else if (CT == 2)
        hipLaunchKernelGGL ((k_real_mfma<V, (CTMAX < 2 ? CTMAX : 2)>), g, dim3 (64), 0, d->stream,
                            pos4, xf, B, zs, cf, cm, tb, rp, jper, pf);
      else if (CT == 4)
        hipLaunchKernelGGL ((k_real_mfma<V, (CTMAX < 4 ? CTMAX : 4)>), g, dim3 (64), 0, d->stream,
                            pos4, xf, B, zs, cf, cm, tb, rp, jper, pf);
      else
        hipLaunchKernelGGL ((k_real_mfma<V, CTMAX>), g, dim3 (64), 0, d->stream, pos4, xf, B, zs,
                            cf, cm, tb, rp, jper, pf);
    }
  else
    hipLaunchKernelGGL ((k_real_block<V, T, JW>), dim3 ((unsigned) ((waves + 3) / 4)), dim3 (256),
                        0, d->stream, pos4, (const T *) XT, B, zs,
                        (const typename Vec4<T>::type *) cand,
                        (const int4 *) d->candm.p, (const unsigned *) d->tbits.p, rp,
                        (T *) d->partial.p);
  HIPCHK (hipGetLastError ());
  if (d->timing) HIPCHK (hipEventRecord (d->ev[1], d->stream));
  if (fork)
    HIPCHK (hipStreamWaitEvent (d->stream, d->evj, 0));
  else if (p.nk > 0)
    {
      if (MF)
        rc = p.poly ? recip_block_launch_mfma<V, 5> (d, (const float *) XT, zs, npad, B, nrecip,
                                                      kpart)
                    : recip_block_launch_mfma<V, 1> (d, (const float *) XT, zs, npad, B, nrecip,
                                                      kpart);
      else if (!p.poly)
        rc = recip_block_launch_dgemm<V> (d, (const double *) XT, B, nrecip, kpart);
      else
        rc = recip_block_launch<V, T, 5> (d, XT, B, nrecip, kpart);
      if (rc) return rc;
    }
  hipLaunchKernelGGL ((k_combine_block<V, T>),
                      dim3 ((unsigned) (((size_t) p.np * D + CB_I - 1) / CB_I),
                            (unsigned) ((B + CB_B - 1) / CB_B)),
                      dim3 (256), 0, d->stream, pos4, x, B, (const T *) d->partial.p, nj,
                      (const T *) kpart, nrecip, q, 1, y);
  HIPCHK (hipGetLastError ());
  if (d->timing)
    {
      HIPCHK (hipEventRecord (d->ev[2], d->stream));
      HIPCHK (hipEventSynchronize (d->ev[2]));
      float ta = 0, tb = 0;
      HIPCHK (hipEventElapsedTime (&ta, d->ev[0], d->ev[1]));
      HIPCHK (hipEventElapsedTime (&tb, d->ev[1], d->ev[2]));
      d->last_ms[0] = ta;
      d->last_ms[1] = tb;
      d->last_ms[2] = ta + tb;
    }
  return SHK_OK;
}

template <int V, typename T> static int real_cache_build (shk_dev *d);

// MIXED blocks wider than one wave's column tiles (32 / 64 / 128 columns for
// FTS / FT / F) evaluate every pair's scalars once per column group in the
// matrix-free kernel.  With dense blocks allowed (shk_dev_set_dense_blocks)
// the real-space blocks are instead written once as a dense FP32 matrix for
// this call (rebuilt after every positions upload, reused by the further
// 128-column passes of the call) and multiplied by the split-K GEMM, when
// the matrix stays within 4 GiB
template <int V>
static int atimes_block_v (shk_dev *d, const double *x, double *y, int B, int prec)
{
  constexpr int D = Dim<V>::D;
  constexpr int CTMAX = V == 2 ? 2 : (V == 1 ? 4 : 8);
  const shk_params &p = d->prm;
  const bool explicit_cache = d->rc_use && d->rc_prec == prec;
  const double nd = (double) p.np * D;
  const bool auto_dense = d->dense_blocks && !d->rc_use && prec == SHK_MIXED && p.periodic
                          && B > 16 * CTMAX && nd * nd * 4.0 <= 4.0 * 1024.0 * 1024.0 * 1024.0;
  if (auto_dense && !(d->dense_ok && d->dense_gen == d->pos_gen))
    {
      int rc = real_cache_build<V, float> (d);
      if (rc) return rc;
      d->dense_ok = 1;
      d->dense_gen = d->pos_gen;
    }
  const bool dense = explicit_cache || auto_dense;
  return prec == SHK_MIXED ? atimes_block_vt<V, float> (d, x, y, B, dense)
                           : atimes_block_vt<V, double> (d, x, y, B, dense);
}

template <int V, typename T>
static int real_cache_build (shk_dev *d)
{
  constexpr int D = Dim<V>::D;
  const shk_params &p = d->prm;
  const size_t ND = (size_t) p.np * D;
  int rc;
  if ((rc = ensure (d, d->rcache, sizeof (T) * ND * ND))) return rc;
  RealParams rp;
  rp.np = p.np; rp.ncand = d->ncand; rp.ncore = d->ncore; rp.poly = p.poly;
  rp.xi = p.xi; rp.rmax2 = p.rmax2; rp.rmin = p.rmin;
  rp.lx = p.lx; rp.ly = p.ly; rp.lz = p.lz; rp.s2x = d->s2x; rp.s3x = d->s3x;
  rp.tx = d->tx; rp.ty = d->ty; rp.tz = d->tz; rp.tbits_words = d->tbits_words;
  const long waves = (long) p.np * ((p.np + 63) / 64);
  const void *cand = std::is_same<T, double>::value ? d->cand.p : d->candf.p;
  hipLaunchKernelGGL ((k_real_dense<V, T>), dim3 ((unsigned) ((waves + 3) / 4)), dim3 (256), 0,
                      d->stream, (const double4 *) d->pos.p,
                      (const typename Vec4<T>::type *) cand, (const int4 *) d->candm.p,
                      (const unsigned *) d->tbits.p, rp, (T *) d->rcache.p);
  HIPCHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_set_overlap (shk_dev *d, int on)
{
  if (!d) { set_err ("shk_dev_set_overlap: no device context"); return SHK_EARG; }
  d->overlap = on != 0;
  return SHK_OK;
}

extern "C" int shk_dev_set_dense_blocks (shk_dev *d, int on)
{
  if (!d) { set_err ("shk_dev_set_dense_blocks: no device context"); return SHK_EARG; }
  d->dense_blocks = on != 0;
  return SHK_OK;
}

extern "C" int shk_dev_real_cache (shk_dev *d, int on, int prec)
{
  if (d && !on)
    {
      d->rc_use = 0;
      return SHK_OK;
    }
  if (!d || !d->have_sys) { set_err ("shk_dev_real_cache: no system"); return SHK_EARG; }
  const shk_params &p = d->prm;
  if (!p.periodic) { set_err ("shk_dev_real_cache: periodic systems only"); return SHK_EARG; }
  d->dense_ok = 0;                // the buffer now holds the explicit cache
  int rc;
  if (p.version == 0)
    rc = prec == SHK_MIXED ? real_cache_build<0, float> (d) : real_cache_build<0, double> (d);
  else if (p.version == 1)
    rc = prec == SHK_MIXED ? real_cache_build<1, float> (d) : real_cache_build<1, double> (d);
  else
    rc = prec == SHK_MIXED ? real_cache_build<2, float> (d) : real_cache_build<2, double> (d);
  if (rc) return rc;
  d->rc_use = 1;
  d->rc_prec = prec;
  return SHK_OK;
}

// ---- allocation-free products -------------------------------------------
// Scratch of the single-column product for callers that cannot allocate:
// one array per GPU (each GPU that runs such a product holds its own copy
// until the process ends).  Host code reaches it only through the address
// hipGetSymbolAddress gives on each device, never through its name.  All
// such products on one GPU share it: calls that overlap on one GPU race.
static __device__ __align__ (256) unsigned char scratch[SHK_NOALLOC_BYTES];

static std::mutex g_scratch_lock;
static unsigned char *g_scratch_addr[SHK_MAX_DEVICES];

static int scratch_address (unsigned char **out)
{
  int dev = -1;
  HIPCHK (hipGetDevice (&dev));
  if (dev < 0 || dev >= SHK_MAX_DEVICES)
    {
      set_err ("allocation-free product: device %d beyond the %d the library tracks", dev,
               SHK_MAX_DEVICES);
      return SHK_EARG;
    }
  std::lock_guard<std::mutex> g (g_scratch_lock);
  if (!g_scratch_addr[dev])
    {
      void *p = nullptr;
      hipError_t e = hipGetSymbolAddress (&p, HIP_SYMBOL (scratch));
      if (e != hipSuccess || !p)
        {
          set_err ("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString (e));
          return SHK_EHIP;
        }
      g_scratch_addr[dev] = (unsigned char *) p;
    }
  *out = g_scratch_addr[dev];
  return SHK_OK;
}

extern "C" int shk_dev_noalloc_begin (shk_dev *d)
{
  if (!d || !d->have_sys) { set_err ("allocation-free product: no system"); return SHK_EARG; }
  unsigned char *base = nullptr;
  int rc = scratch_address (&base);
  if (rc) return rc;
  for (int k = 0; k < 6; k++) d->carve_saved[k] = *scratch_buf (d, k);
  d->carve_base = base;
  d->carve_cap = SHK_NOALLOC_BYTES;
  d->carve_off = 0;
  d->ptab_kind = 0;          // the phase tables are rebuilt in the scratch
  return SHK_OK;
}

extern "C" void shk_dev_noalloc_end (shk_dev *d)
{
  if (!d || !d->carve_base) return;
  for (int k = 0; k < 6; k++) *scratch_buf (d, k) = d->carve_saved[k];
  d->carve_base = nullptr;
  d->carve_cap = d->carve_off = 0;
  d->ptab_kind = 0;          // the owned table no longer matches the positions
}

extern "C" int shk_dev_atimes (shk_dev *d, const double *x, double *y, int nrhs, int prec)
{
  if (!d || !d->have_sys) { set_err ("shk_dev_atimes: no system"); return SHK_EARG; }
  if (nrhs < 1) { set_err ("shk_dev_atimes: nrhs < 1"); return SHK_EARG; }
  const shk_params &p = d->prm;
  const size_t n = (size_t) p.np * dimv (p.version);
  if (nrhs >= 2 && p.periodic)
    {
      // block path, at most 128 columns per pass
      for (int c0 = 0; c0 < nrhs; c0 += 128)
        {
          const int B = nrhs - c0 < 128 ? nrhs - c0 : 128;
          int rc;
          if (p.version == 0) rc = atimes_block_v<0> (d, x + c0 * n, y + c0 * n, B, prec);
          else if (p.version == 1) rc = atimes_block_v<1> (d, x + c0 * n, y + c0 * n, B, prec);
          else rc = atimes_block_v<2> (d, x + c0 * n, y + c0 * n, B, prec);
          if (rc) return rc;
        }
      return SHK_OK;
    }
  for (int c = 0; c < nrhs; c++)
    {
      int rc;
      if (p.version == 0) rc = atimes_v<0> (d, x + c * n, y + c * n, prec);
      else if (p.version == 1) rc = atimes_v<1> (d, x + c * n, y + c * n, prec);
      else rc = atimes_v<2> (d, x + c * n, y + c * n, prec);
      if (rc) return rc;
    }
  return SHK_OK;
}

// ---- lubrication ----------------------------------------------------------

// the directed pair list of the current positions (count, scan, fill) and
// the geometry / evaluation parameters of the lubrication kernels
static int lub_setup (shk_dev *d, int version, const shk_lub *lp, LubGeom &g, LubEval &ev,
                      int &total)
{
  if (!d || !d->have_sys) { set_err ("lubrication: no system"); return SHK_EARG; }
  if (version < 0 || version > 2) { set_err ("lubrication: version %d", version); return SHK_EARG; }
  const shk_params &p = d->prm;
  const int np = p.np;
  g.np = np;
  g.periodic = p.periodic;
  g.poly = p.poly;
  g.lx = p.lx; g.ly = p.ly; g.lz = p.lz;
  g.s2x = p.periodic ? d->s2x : 0.0;
  g.s3x = p.periodic ? d->s3x : 0.0;
  g.lubmax2 = lp->lubmax * lp->lubmax;
  g.ix = g.iy = g.iz = 0;
  if (p.periodic)
    {
      // set_imax_lub_periodic (lub.c:129-162): lubmax, or the largest box
      // length when there is no limit
      double lm = lp->lubmax;
      if (lm <= 0.0) lm = std::max (p.lx, std::max (p.ly, p.lz));
      g.ix = (int) (lm / p.lx) + 1;
      g.iy = (int) (lm / p.ly) + 1;
      g.iz = (int) (lm / p.lz) + 1;
      if (g.ix > 15 || g.iy > 15 || g.iz > 15)
        {
          set_err ("lubrication: more than 15 image cells per direction (lubmax %g, box %g %g %g)",
                   lp->lubmax, p.lx, p.ly, p.lz);
          return SHK_EUNSUP;
        }
    }
  g.minimg = lub_minimg (g, lp->lubmax, lp->amax);
  ev.lubmin2 = lp->lubmin2;
  ev.nmax = lp->nmax;
  ev.flag_lub = lp->flag_lub;
  ev.N = lp->N;
  ev.coef = nullptr;
  ev.f1 = nullptr;
  int rc;
  if (p.poly)
    {
      if (!lp->coef || lp->N < lp->nmax) { set_err ("lubrication: no coefficient table"); return SHK_EARG; }
      if (d->lub_N != lp->N)
        {
          const size_t per = (size_t) (lp->N + 1) * (lp->N + 2);
          if ((rc = ensure (d, d->lubcoef, sizeof (double) * 11 * per))) return rc;
          for (int f = 0; f < 11; f++)
            HIPCHK (hipMemcpyAsync ((double *) d->lubcoef.p + f * per, lp->coef[f],
                                    sizeof (double) * per, hipMemcpyHostToDevice, d->stream));
          d->lub_N = lp->N;
        }
      ev.coef = (const double *) d->lubcoef.p;
    }
  else
    {
      if (!d->lub_f1)
        {
          if (!lp->f1) { set_err ("lubrication: no f(1) table"); return SHK_EARG; }
          if ((rc = ensure (d, d->lubf1, sizeof (double) * 11 * 51))) return rc;
          HIPCHK (hipMemcpyAsync (d->lubf1.p, lp->f1, sizeof (double) * 11 * 51,
                                  hipMemcpyHostToDevice, d->stream));
          d->lub_f1 = 1;
        }
      ev.f1 = (const double *) d->lubf1.p;
    }
  if ((rc = ensure (d, d->lubcnt, sizeof (int) * (size_t) (np > 0 ? np : 1)))) return rc;
  if ((rc = ensure (d, d->luboff, sizeof (int) * (size_t) (np + 1)))) return rc;
  const double4 *pos4 = (const double4 *) d->pos.p;
  const dim3 waves ((unsigned) ((np + 3) / 4));
  hipLaunchKernelGGL ((k_lub_pairs<0>), waves, dim3 (256), 0, d->stream, pos4, g,
                      (int *) d->lubcnt.p, (const int *) nullptr, (int2 *) nullptr);
  HIPCHK (hipGetLastError ());
  hipLaunchKernelGGL (k_lub_scan, dim3 (1), dim3 (1024), 0, d->stream, (const int *) d->lubcnt.p,
                      np, (int *) d->luboff.p);
  HIPCHK (hipGetLastError ());
  total = 0;
  HIPCHK (hipMemcpyAsync (&total, (int *) d->luboff.p + np, sizeof (int), hipMemcpyDeviceToHost,
                          d->stream));
  HIPCHK (hipStreamSynchronize (d->stream));
  if (total < 0) { set_err ("lubrication: pair count overflow"); return SHK_EARG; }
  if ((rc = ensure (d, d->lublist, sizeof (int2) * (size_t) (total > 0 ? total : 1)))) return rc;
  if (total > 0)
    {
      hipLaunchKernelGGL ((k_lub_pairs<1>), waves, dim3 (256), 0, d->stream, pos4, g,
                          (int *) nullptr, (const int *) d->luboff.p, (int2 *) d->lublist.p);
      HIPCHK (hipGetLastError ());
    }
  return SHK_OK;
}

extern "C" int shk_dev_lub (shk_dev *d, int version, const shk_lub *lp, const double *x,
                            double *y)
{
  LubGeom g;
  LubEval ev;
  int total = 0, rc = lub_setup (d, version, lp, g, ev, total);
  if (rc) return rc;
  const int np = g.np, D = dimv (version);
  if ((rc = ensure (d, d->lubval, sizeof (double) * (size_t) D * (total > 0 ? total : 1))))
    return rc;
  const double4 *pos4 = (const double4 *) d->pos.p;
  if (total > 0)
    {
      const dim3 pg ((unsigned) ((total + 127) / 128));
      const int *lo = (const int *) d->luboff.p;
      const int2 *li = (const int2 *) d->lublist.p;
      double *val = (double *) d->lubval.p;
      const size_t lsh = sizeof (double) * 4 * (size_t) (ev.N + 2);
#define LUB_EVAL(VV, PP)                                                                  \
      hipLaunchKernelGGL ((k_lub_eval<VV, PP>), pg, dim3 (128), PP ? lsh : 0, d->stream, pos4, \
                          g, ev, lo, li, total, x, val)
      if (version == 0) { if (g.poly) LUB_EVAL (0, 1); else LUB_EVAL (0, 0); }
      else if (version == 1) { if (g.poly) LUB_EVAL (1, 1); else LUB_EVAL (1, 0); }
      else { if (g.poly) LUB_EVAL (2, 1); else LUB_EVAL (2, 0); }
#undef LUB_EVAL
      HIPCHK (hipGetLastError ());
    }
  const size_t nd = (size_t) np * D;
  hipLaunchKernelGGL (k_lub_gather, dim3 ((unsigned) ((nd + 255) / 256)), dim3 (256), 0, d->stream,
                      (const int *) d->luboff.p, (const double *) d->lubval.p, np, D,
                      0x7fffffff, y);
  HIPCHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_lub_matrix (shk_dev *d, int version, const shk_lub *lp, double *mat)
{
  LubGeom g;
  LubEval ev;
  int total = 0, rc = lub_setup (d, version, lp, g, ev, total);
  if (rc) return rc;
  const int np = g.np, D = dimv (version);
  const size_t n = (size_t) np * D;
  HIPCHK (hipMemsetAsync (mat, 0, sizeof (double) * n * n, d->stream));
  if (total == 0) return SHK_OK;
  if ((rc = ensure (d, d->lubval, sizeof (double) * 26 * (size_t) total))) return rc;
  const double4 *pos4 = (const double4 *) d->pos.p;
  const dim3 pg ((unsigned) ((total + 127) / 128));
  const int *off = (const int *) d->luboff.p;
  const int2 *list = (const int2 *) d->lublist.p;
  double *sc = (double *) d->lubval.p;
  if (version == 0)
    {
      hipLaunchKernelGGL ((k_lub_scal<0>), pg, dim3 (128), 0, d->stream, pos4, g, ev, off, list,
                          total, sc);
      hipLaunchKernelGGL ((k_lub_matrix<0>), dim3 (np), dim3 (128), 0, d->stream, off, list,
                          (const double *) sc, np, mat);
    }
  else if (version == 1)
    {
      hipLaunchKernelGGL ((k_lub_scal<1>), pg, dim3 (128), 0, d->stream, pos4, g, ev, off, list,
                          total, sc);
      hipLaunchKernelGGL ((k_lub_matrix<1>), dim3 (np), dim3 (128), 0, d->stream, off, list,
                          (const double *) sc, np, mat);
    }
  else
    {
      hipLaunchKernelGGL ((k_lub_scal<2>), pg, dim3 (128), 0, d->stream, pos4, g, ev, off, list,
                          total, sc);
      hipLaunchKernelGGL ((k_lub_matrix<2>), dim3 (np), dim3 (128), 0, d->stream, off, list,
                          (const double *) sc, np, mat);
    }
  HIPCHK (hipGetLastError ());
  return SHK_OK;
}

template <int V>
static int matrix_v (shk_dev *d, double *mat)
{
  const shk_params &p = d->prm;
  DevParams q = dev_params (p);
  dim3 grid ((p.np + 63) / 64, p.np);
  if (d->timing) HIPCHK (hipEventRecord (d->ev[0], d->stream));
  if (p.periodic && !p.poly && d->tsym && d->ncand <= REAL_PC_CAND)
    {
      // each pair i <= j once: real-space moments over the candidates, the
      // k sum from the phase table and the per-k moment weights, both
      // blocks written
      const int nk = p.nk, np = p.np;
      int rc;
      if ((rc = ensure (d, d->eph, sizeof (double) * 2 * (size_t) (nk > 0 ? nk : 1) * np)))
        return rc;
      if ((rc = ensure (d, d->kpart, sizeof (double) * Dim<V>::NMOM * (size_t) (nk > 0 ? nk : 1))))
        return rc;
      const double4 *pos4 = (const double4 *) d->pos.p;
      const double4 *kv = (const double4 *) d->kv.p;
      if (nk > 0)
        {
          hipLaunchKernelGGL ((k_phase_E<false>), dim3 ((unsigned) (((size_t) nk * np + 255) / 256)),
                              dim3 (256), 0, d->stream, pos4, kv, nk, np, (double *) d->eph.p);
          hipLaunchKernelGGL ((k_kmom_weights<V>), dim3 ((nk + 127) / 128), dim3 (128), 0,
                              d->stream, kv, (const double *) d->ks.p, nk, (double *) d->kpart.p);
        }
      RealParams rp;
      rp.np = p.np; rp.ncand = d->ncand; rp.ncore = d->ncore; rp.poly = 0;
      rp.xi = p.xi; rp.rmax2 = p.rmax2; rp.rmin = p.rmin;
      rp.lx = p.lx; rp.ly = p.ly; rp.lz = p.lz; rp.s2x = d->s2x; rp.s3x = d->s3x;
      rp.tx = d->tx; rp.ty = d->ty; rp.tz = d->tz; rp.tbits_words = d->tbits_words;
      const int nt = (np + 15) / 16;
      hipLaunchKernelGGL ((k_matrix_sym<V>), dim3 (nt, nt), dim3 (256), 0, d->stream, pos4,
                          (const double4 *) d->cand.p, (const int4 *) d->candm.p,
                          (const unsigned *) d->tbits.p, rp, q, (const double *) d->eph.p,
                          (const double *) d->kpart.p, nk, mat);
    }
  else if (p.periodic)
    hipLaunchKernelGGL ((k_matrix<V, 0>), grid, dim3 (64), 0, d->stream,
                        (const double4 *) d->pos.p, (const double4 *) d->img.p,
                        (const double4 *) d->kv.p, (const double *) d->ks.p, q, mat);
  else
    hipLaunchKernelGGL ((k_matrix<V, 1>), grid, dim3 (64), 0, d->stream,
                        (const double4 *) d->pos.p, (const double4 *) d->img.p,
                        (const double4 *) d->kv.p, (const double *) d->ks.p, q, mat);
  HIPCHK (hipGetLastError ());
  if (d->timing)
    {
      HIPCHK (hipEventRecord (d->ev[1], d->stream));
      HIPCHK (hipEventSynchronize (d->ev[1]));
      float a = 0;
      HIPCHK (hipEventElapsedTime (&a, d->ev[0], d->ev[1]));
      d->last_ms[0] = a; d->last_ms[1] = 0; d->last_ms[2] = a;
    }
  return SHK_OK;
}

extern "C" int shk_dev_matrix (shk_dev *d, double *mat)
{
  if (!d || !d->have_sys) { set_err ("shk_dev_matrix: no system"); return SHK_EARG; }
  if (d->prm.version == 0) return matrix_v<0> (d, mat);
  if (d->prm.version == 1) return matrix_v<1> (d, mat);
  return matrix_v<2> (d, mat);
}

extern "C" int shk_dev_extmat (shk_dev *d, const double *mat, const double *x, double *y,
                               int nrhs)
{
  if (!d || !d->have_sys) { set_err ("shk_dev_extmat: no system"); return SHK_EARG; }
  const shk_params &p = d->prm;
  const int D = dimv (p.version);
  const size_t n = (size_t) p.np * D;
  int rc = ensure (d, d->zt, sizeof (double) * (n > 4 * (size_t) p.np ? n : 4 * (size_t) p.np));
  if (rc) return rc;
  for (int c = 0; c < nrhs; c++)
    {
      const double *xc = x + c * n;
      double *zc = (double *) d->zt.p;
      if (p.version == 2)
        hipLaunchKernelGGL (k_transform<2>, dim3 ((p.np + 255) / 256), dim3 (256), 0, d->stream,
                            xc, p.np, zc);
      else
        HIPCHK (hipMemcpyAsync (zc, xc, sizeof (double) * n, hipMemcpyDeviceToDevice, d->stream));
      hipLaunchKernelGGL (k_gemv_rows, dim3 ((unsigned) ((n + 3) / 4)), dim3 (256), 0, d->stream,
                          mat, zc, n, y + c * n);
      HIPCHK (hipGetLastError ());
    }
  return SHK_OK;
}

__global__ void k_axpby4 (double *__restrict__ b2, const double *__restrict__ x,
                          const double *__restrict__ b1, const double *__restrict__ b0,
                          const double *__restrict__ y, double ca, double cb, double cy,
                          size_t len)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  b2[i] = ca * x[i] + cb * b1[i] - b0[i] + cy * y[i];
}

extern "C" int shk_dev_axpby4 (shk_dev *d, double *b2, const double *x, const double *b1,
                               const double *b0, const double *y, double ca, double cb,
                               double cy, size_t len)
{
  hipLaunchKernelGGL (k_axpby4, dim3 ((unsigned) ((len + 255) / 256)), dim3 (256), 0,
                      d->stream, b2, x, b1, b0, y, ca, cb, cy, len);
  HIPCHK (hipGetLastError ());
  return SHK_OK;
}

__global__ void k_coldot (const double *__restrict__ u, const double *__restrict__ v, int n,
                          double *__restrict__ out)
{
  __shared__ double red[256];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x)
    acc += u[(size_t) c * n + k] * v[(size_t) c * n + k];
  red[threadIdx.x] = acc;
  __syncthreads ();
  for (int st = blockDim.x / 2; st > 0; st >>= 1)
    {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads ();
    }
  if (threadIdx.x == 0) out[c] = red[0];
}

extern "C" int shk_dev_coldot (shk_dev *d, const double *u, const double *v, int n, int m,
                               double *out_host)
{
  int rc = ensure (d, d->ys, sizeof (double) * (m > 0 ? m : 1));
  if (rc) return rc;
  hipLaunchKernelGGL (k_coldot, dim3 (m), dim3 (256), 0, d->stream, u, v, n, (double *) d->ys.p);
  HIPCHK (hipGetLastError ());
  HIPCHK (hipMemcpyAsync (out_host, d->ys.p, sizeof (double) * m, hipMemcpyDeviceToHost,
                          d->stream));
  HIPCHK (hipStreamSynchronize (d->stream));
  return SHK_OK;
}

// y[row] = sum_col mat[row][col] * x[col] for a general (n1 x n2) matrix
__global__ __launch_bounds__ (256) void
k_gemv_general (const double *__restrict__ mat, const double *__restrict__ x, int n1, int n2,
                double *__restrict__ y)
{
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n1) return;
  const double *mr = mat + (size_t) row * n2;
  double acc = 0.0;
  for (int c = lane; c < n2; c += 64) acc += mr[c] * x[c];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor (acc, off, 64);
  if (lane == 0) y[row] = acc;
}

extern "C" int shk_dev_gemv (shk_dev *d, const double *mat, int n1, int n2, const double *x,
                             int extracted, double *y)
{
  if (n1 <= 0 || n2 <= 0) { set_err ("shk_dev_gemv: empty matrix"); return SHK_EARG; }
  const double *xx = x;
  if (extracted)
    {
      if (n1 != n2 || n2 % 11) { set_err ("shk_dev_gemv: extracted form needs 11np"); return SHK_EARG; }
      int rc = ensure (d, d->zt, sizeof (double) * (size_t) n2);
      if (rc) return rc;
      hipLaunchKernelGGL (k_transform<2>, dim3 ((n2 / 11 + 255) / 256), dim3 (256), 0, d->stream,
                          x, n2 / 11, (double *) d->zt.p);
      HIPCHK (hipGetLastError ());
      xx = (const double *) d->zt.p;
    }
  hipLaunchKernelGGL (k_gemv_general, dim3 ((n1 + 3) / 4), dim3 (256), 0, d->stream, mat, xx,
                      n1, n2, y);
  HIPCHK (hipGetLastError ());
  return SHK_OK;
}

__global__ void k_scal (double *__restrict__ dst, const double *__restrict__ src, double alpha,
                        size_t len)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  dst[i] = src ? alpha * src[i] : 0.0;
}

extern "C" int shk_dev_scal (shk_dev *d, double *dst, const double *src, double alpha,
                             size_t len)
{
  if (len == 0) return SHK_OK;
  hipLaunchKernelGGL (k_scal, dim3 ((unsigned) ((len + 255) / 256)), dim3 (256), 0, d->stream,
                      dst, src, alpha, len);
  HIPCHK (hipGetLastError ());
  return SHK_OK;
}

// plain streaming copy of known size: the reference point that converts
// hardware fetch/write counters into bytes for the roofline report
__global__ __launch_bounds__ (256) void
k_copy_calibrate (const double4 *__restrict__ src, double4 *__restrict__ dst, size_t n4)
{
  for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (size_t) gridDim.x * blockDim.x)
    dst[i] = src[i];
}

extern "C" int shk_dev_copy_calibrate (shk_dev *d, void *dst, const void *src, size_t bytes)
{
  const size_t n4 = bytes / sizeof (double4);
  hipLaunchKernelGGL (k_copy_calibrate, dim3 (2048), dim3 (256), 0, d->stream,
                      (const double4 *) src, (double4 *) dst, n4);
  HIPCHK (hipGetLastError ());
  HIPCHK (hipStreamSynchronize (d->stream));
  return SHK_OK;
}

// ---- adjoint of the F-version product with respect to the positions ----

template <typename T>
static int adjoint_t (shk_dev *d, const double *f, const double *g, double *dx)
{
  const shk_params &p = d->prm;
  const int np = p.np, nk = p.nk;
  int rc;
  if ((rc = ensure (d, d->adjfg, sizeof (double) * 6 * (size_t) np))) return rc;
  hipLaunchKernelGGL (k_adj_pack, dim3 ((np + 255) / 256), dim3 (256), 0, d->stream, f, g, np,
                      (double *) d->adjfg.p);
  HIPCHK (hipGetLastError ());
  const double *fg = (const double *) d->adjfg.p;
  const double4 *pos4 = (const double4 *) d->pos.p;
  // real space
  const int nj = (np + 63) / 64;
  if ((rc = ensure (d, d->partial, sizeof (double) * (size_t) nj * np * 3))) return rc;
  RealParams rp;
  rp.np = np; rp.ncand = d->ncand; rp.ncore = d->ncore; rp.poly = 0;
  rp.xi = p.xi; rp.rmax2 = p.rmax2; rp.rmin = p.rmin;
  rp.lx = p.lx; rp.ly = p.ly; rp.lz = p.lz; rp.s2x = d->s2x; rp.s3x = d->s3x;
  rp.tx = d->tx; rp.ty = d->ty; rp.tz = d->tz; rp.tbits_words = d->tbits_words;
  const long waves = (long) np * nj;
  const void *cand = std::is_same<T, double>::value ? d->cand.p : d->candf.p;
  hipLaunchKernelGGL ((k_real_adj<T>), dim3 ((unsigned) ((waves + 3) / 4)), dim3 (256), 0,
                      d->stream, pos4, fg, (const typename Vec4<T>::type *) cand,
                      (const int4 *) d->candm.p, (const unsigned *) d->tbits.p, rp,
                      (double *) d->partial.p);
  HIPCHK (hipGetLastError ());
  // reciprocal space
  int nrecip = 0;
  if (nk > 0)
    {
      int sj, jper, sk, kper;
      recip_shape (np, nk, sj, jper, sk, kper);
      if ((rc = ensure (d, d->kpart, sizeof (T) * (size_t) sj * nk * 12))) return rc;
      if ((rc = ensure (d, d->adjcs, sizeof (T) * (size_t) nk * 12))) return rc;
      if ((rc = ensure (d, d->adjpart, sizeof (T) * (size_t) sk * np * 3))) return rc;
      const double4 *kv = (const double4 *) d->kv.p;
      hipLaunchKernelGGL ((k_recip_sf<1, T, 1>), dim3 ((nk + 255) / 256, sj), dim3 (256), 0,
                          d->stream, pos4, fg, kv, np, nk, jper, (T *) d->kpart.p);
      HIPCHK (hipGetLastError ());
      hipLaunchKernelGGL ((k_adj_sf_sum<T>), dim3 ((nk + 127) / 128), dim3 (128), 0, d->stream,
                          (const T *) d->kpart.p, nk, sj, (T *) d->adjcs.p);
      HIPCHK (hipGetLastError ());
      hipLaunchKernelGGL ((k_adj_back<T>), dim3 ((np + 255) / 256, sk), dim3 (256), 0, d->stream,
                          pos4, fg, kv, (const double *) d->ks.p, (const T *) d->adjcs.p, np, nk,
                          kper, (T *) d->adjpart.p);
      HIPCHK (hipGetLastError ());
      nrecip = sk;
    }
  hipLaunchKernelGGL ((k_adj_combine<T>), dim3 ((np * 3 + 255) / 256), dim3 (256), 0, d->stream,
                      (const double *) d->partial.p, nj, (const T *) d->adjpart.p, nrecip, np,
                      dx);
  HIPCHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_dev_adjoint_x (shk_dev *d, const double *f, const double *g, double *dx,
                                  int prec)
{
  if (!d || !d->have_sys) { set_err ("shk_dev_adjoint_x: no system"); return SHK_EARG; }
  const shk_params &p = d->prm;
  if (p.version != 0 || !p.periodic || p.poly)
    {
      set_err ("adjoint: implemented for the periodic F version with equal spheres");
      return SHK_EUNSUP;
    }
  return prec == SHK_MIXED ? adjoint_t<float> (d, f, g, dx) : adjoint_t<double> (d, f, g, dx);
}
