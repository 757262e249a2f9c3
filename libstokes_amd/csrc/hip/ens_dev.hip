// ens_dev.hip -- the operator of many independent periodic systems (boxes of
// uneven size, radii and Ewald splitting) as one launch per phase.
//
// Every phase of the single-system product (real-space pair sum, phase
// tables, structure factors, k-blocks, back-projection, combine; the FP64
// lubrication correction: near-pair count, scan, fill, evaluate, gather) is
// one kernel over the work items of all boxes.  A box owns a contiguous
// range of the phase's workgroups (start[s] .. start[s + 1] for the box at
// position s of the launch order); a workgroup finds its box by binary
// search over start, so no workgroup straddles two boxes and the per-box
// tile shapes (64-particle chunks per real-space wave, particle and k splits
// of the reciprocal kernels) are free.  Boxes are ordered by decreasing
// work, so the smallest boxes fill the tail of each launch.  The kernels
// are the single-system kernels' bodies, reading the box's descriptor
// (EnsBox) and its slice of the concatenated tables and work buffers; the
// summation order of every output is fixed, so a box's result does not
// depend on the other boxes or on scheduling.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <vector>
#include <algorithm>
#include <numeric>
#include <type_traits>
#include "../shk.h"
#include "ewald_kernels.h"
#include "real_kernels.h"
#include "recip_kernels.h"
#include "lub_kernels.h"

using namespace shk;

#define ECHK(call)                                                            \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      {                                                                       \
        shk_set_error ("%s:%d %s: %s", __FILE__, __LINE__, #call,             \
                       hipGetErrorString (e_));                               \
        return SHK_EHIP;                                                      \
      }                                                                       \
  } while (0)

namespace {

// one box as the batched kernels see it
struct EnsBox {
  int np, off;                 // particles; first particle in the packed arrays
  int nj, jw, njw;             // 64-particle chunks, chunks per wave, waves per row
  int cand_off, tbits_off;     // slices of the concatenated lattice tables
  int nk, k_off;               // k vectors and their slice
  int sj, jper, jt;            // structure factors: particle splits, per split, LDS tile
  int sk, kper, nkb, nib;      // back-projection: k splits, per split; blocks of 64
  long long ptab_off;          // Cplx<T> elements
  long long part_off;          // doubles: real-space partials [njw][np][D]
  long long kpart_off, kpq_off, kback_off;   // T elements
  RealParams rp;
  DevParams q;
  PhaseBasis b;
  LubGeom lg;
  LubEval ev;
};

enum { PH_REAL, PH_TAB, PH_SF, PH_BLK, PH_BACK, PH_COMB, PH_LUBW, NPH };

__device__ __forceinline__ int ens_find (const int *__restrict__ start, int nbox, int blk)
{
  int lo = 0, hi = nbox;
  while (hi - lo > 1)
    {
      const int mid = (lo + hi) >> 1;
      if (start[mid] <= blk) lo = mid;
      else hi = mid;
    }
  return lo;
}

struct EnsMap {
  const int *start;            // [NPH][nbox + 1]
  const int *order;            // launch position -> box
  int nbox;
};

__device__ __forceinline__ int ens_box (const EnsMap &m, int ph, int blk, int &local)
{
  const int *st = m.start + (size_t) ph * (m.nbox + 1);
  const int s = ens_find (st, m.nbox, blk);
  local = blk - st[s];
  return m.order[s];
}

__global__ void k_ens_pack (const double *__restrict__ pos, const double *__restrict__ rad, int n,
                            double4 *__restrict__ out)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[t] = make_double4 (pos[3 * t], pos[3 * t + 1], pos[3 * t + 2], rad[t]);
}

template <int V, typename T>
__global__ __launch_bounds__ (256) void
k_ens_real (EnsMap m, const EnsBox *__restrict__ boxes, const double4 *__restrict__ pos,
            const double *__restrict__ x, const typename Vec4<T>::type *__restrict__ cand,
            const int4 *__restrict__ candm, const unsigned *__restrict__ tbits,
            double *__restrict__ partial)
{
  constexpr int D = Dim<V>::D;
  __shared__ RealLds<V, T> lds_all[4];
  __shared__ unsigned bits_lds[1024];
  int local;
  const EnsBox &B = boxes[ens_box (m, PH_REAL, blockIdx.x, local)];
  const RealParams rp = B.rp;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned *gbits = tbits + B.tbits_off;
  const bool bits_in_lds = rp.tbits_words <= 1024;
  if (bits_in_lds)
    for (int w = threadIdx.x; w < rp.tbits_words; w += blockDim.x) bits_lds[w] = gbits[w];
  __syncthreads ();
  const unsigned *bits = bits_in_lds ? bits_lds : gbits;
  const int njw = B.njw, jw = B.jw, nj = B.nj;
  const int gw = local * 4 + wave;
  const int i = gw / njw, g = gw - i * njw;
  if (i >= rp.np) return;
  const int off = B.off;
  T acc[D];
#pragma unroll
  for (int q = 0; q < D; q++) acc[q] = T (0);
  const int jc1 = min (nj, (g + 1) * jw);
  for (int jc = g * jw; jc < jc1; jc++)
    real_chunk<V, T> (pos + off, x + (size_t) D * off, cand + B.cand_off, candm + B.cand_off,
                      bits, rp, i, jc, lds_all[wave], lane, acc);
  wave_reduce<D> (acc);
  if (lane == 0)
    {
      double *out = partial + B.part_off + ((size_t) g * rp.np + i) * D;
#pragma unroll
      for (int q = 0; q < D; q++) out[q] = double (acc[q]);
    }
}

template <typename T>
__global__ __launch_bounds__ (256) void
k_ens_phase_tab (EnsMap m, const EnsBox *__restrict__ boxes, const double4 *__restrict__ pos,
                 Cplx<T> *__restrict__ tab)
{
  int local;
  const EnsBox &B = boxes[ens_box (m, PH_TAB, blockIdx.x, local)];
  const PhaseBasis b = B.b;
  phase_tab_item<T> ((size_t) local * 256 + threadIdx.x, pos + B.off, B.np, b, b.W,
                     tab + B.ptab_off);
}

template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (256) void
k_ens_sf (EnsMap m, const EnsBox *__restrict__ boxes, const double4 *__restrict__ pos,
          const double *__restrict__ x, const int4 *__restrict__ km,
          const Cplx<T> *__restrict__ tab, T *__restrict__ kpart)
{
  int local;
  const EnsBox &B = boxes[ens_box (m, PH_SF, blockIdx.x, local)];
  const PhaseBasis b = B.b;
  const int bx = local % B.nkb, by = local / B.nkb;
  recip_sf_tab_block<V, T, NPOW> (bx, by, pos + B.off, x + (size_t) Dim<V>::D * B.off,
                                  km + B.k_off, tab + B.ptab_off, b, B.np, B.nk, B.jper, B.jt,
                                  kpart + B.kpart_off);
}

template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (128) void
k_ens_blocks (EnsMap m, const EnsBox *__restrict__ boxes, const double4 *__restrict__ kv,
              const double *__restrict__ ks, const T *__restrict__ kpart, T *__restrict__ kpq)
{
  int local;
  const EnsBox &B = boxes[ens_box (m, PH_BLK, blockIdx.x, local)];
  recip_blocks_item<V, T, NPOW> (local * 128 + threadIdx.x, kv + B.k_off,
                                 ks + (size_t) 8 * B.k_off, B.nk, B.sj, kpart + B.kpart_off,
                                 kpq + B.kpq_off);
}

template <int V, typename T, int NPOW>
__global__ __launch_bounds__ (256) void
k_ens_back (EnsMap m, const EnsBox *__restrict__ boxes, const double4 *__restrict__ pos,
            const double4 *__restrict__ kv, const int4 *__restrict__ km,
            const T *__restrict__ kpq, T *__restrict__ kback)
{
  int local;
  const EnsBox &B = boxes[ens_box (m, PH_BACK, blockIdx.x, local)];
  const PhaseBasis b = B.b;
  const int bx = local % B.nib, by = local / B.nib;
  recip_back_rows_block<V, T, NPOW> (bx, by, pos + B.off, kv + B.k_off, km + B.k_off, b,
                                     kpq + B.kpq_off, B.np, B.nk, B.kper, kback + B.kback_off);
}

template <int V, typename T>
__global__ __launch_bounds__ (256) void
k_ens_combine (EnsMap m, const EnsBox *__restrict__ boxes, const double4 *__restrict__ pos,
               const double *__restrict__ x, const double *__restrict__ partial,
               const T *__restrict__ kback, double *__restrict__ y)
{
  constexpr int D = Dim<V>::D;
  int local;
  const EnsBox &B = boxes[ens_box (m, PH_COMB, blockIdx.x, local)];
  const DevParams q = B.q;
  combine_item<V, T> ((size_t) local * 256 + threadIdx.x, pos + B.off, x + (size_t) D * B.off,
                      partial + B.part_off, B.njw, kback + B.kback_off, B.nk > 0 ? B.sk : 0, q,
                      1, y + (size_t) D * B.off);
}

// near pairs of the lubrication correction: count (FILL 0) or fill (FILL 1)
// the directed pairs of every receiving particle, one wave per particle;
// cnt and off are indexed by the packed particle index, list entries hold
// box-local j
template <int FILL>
__global__ __launch_bounds__ (256) void
k_ens_lub_pairs (EnsMap m, const EnsBox *__restrict__ boxes, const double4 *__restrict__ pos,
                 int *__restrict__ cnt, const int *__restrict__ loff, int2 *__restrict__ list,
                 int cap)
{
  int local;
  const EnsBox &B = boxes[ens_box (m, PH_LUBW, blockIdx.x, local)];
  const LubGeom g = B.lg;
  lub_pairs_wave<FILL> (local * 4 + (threadIdx.x >> 6), threadIdx.x & 63, pos + B.off, g,
                        FILL ? nullptr : cnt + B.off, FILL ? loff + B.off : nullptr, list, cap);
}

// one directed pair per thread, over min(total, cap) pairs; the receiving
// particle by binary search over loff, its box over the box offsets.  The
// evaluation parameters (clamp, series order, form, coefficient tables) are
// the same for every box and come in as kernel arguments, and the kind of
// radii is a template argument: the series loops then run in uniform
// control flow with wave-uniform coefficient addresses (scalar loads).
template <int V, int POLY>
__global__ __launch_bounds__ (128) void
k_ens_lub_eval (const EnsBox *__restrict__ boxes, const int *__restrict__ box_off, int nbox,
                const double4 *__restrict__ pos, const int *__restrict__ loff, int ntot,
                const int2 *__restrict__ list, int cap, LubEval ev, const double *__restrict__ x,
                double *__restrict__ val)
{
  extern __shared__ double lub_sh[];
  const int total = min (loff[ntot], cap);
  if (total <= 0) return;                      // the whole grid alike
  const int p0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x * blockDim.x >= total) return;  // whole blocks past the last pair
  // threads past the last pair evaluate it without storing (the block's
  // coefficient staging needs every thread)
  const int p = p0 < total ? p0 : total - 1;
  int lo = 0, hi = ntot;
  while (hi - lo > 1)
    {
      const int mid = (lo + hi) >> 1;
      if (loff[mid] <= p) lo = mid;
      else hi = mid;
    }
  const int b = ens_find (box_off, nbox, lo);
  const EnsBox &B = boxes[b];
  LubGeom g = B.lg;
  g.poly = POLY;
  lub_eval_item<V> (p, pos + B.off, g, ev, loff + B.off, list, x + (size_t) Dim<V>::D * B.off,
                    val, POLY ? lub_sh : nullptr, p0 < total);
}

struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
};

int grow (Buf &b, size_t bytes)
{
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return SHK_OK;
  if (b.p) (void) hipFree (b.p);
  b.p = nullptr;
  b.bytes = 0;
  if (hipMalloc (&b.p, bytes) != hipSuccess)
    {
      (void) hipGetLastError ();
      shk_set_error ("ensemble: device allocation of %zu bytes failed", bytes);
      return SHK_ENOMEM;
    }
  b.bytes = bytes;
  return SHK_OK;
}

void drop (Buf &b)
{
  if (b.p) (void) hipFree (b.p);
  b.p = nullptr;
  b.bytes = 0;
}

int dimv (int v) { return v == 0 ? 3 : (v == 1 ? 6 : 11); }

}  // namespace

struct shk_ens {
  hipStream_t s = nullptr, s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int nbox = 0, ntot = 0, version = 2, prec = SHK_MIXED, poly = 0;
  int have = 0, have_lub = 0;
  std::vector<EnsBox> hb;          // host copies of the descriptors
  std::vector<int> hstart;         // [NPH][nbox + 1]
  std::vector<int> hbox_off;       // particle offsets, nbox + 1
  long long nblk[NPH] = {};
  size_t sf_lds = 0;
  Buf boxes, start, order, box_off;
  Buf rad, pos4;
  Buf cand, candf, candm, tbits, kv, ks, km;
  Buf ptab, partial, kpart, kpq, kback;
  Buf lubcoef, lubf1, lubcnt, luboff, lublist, lubval;
  int lub_cap = 0;
  int timed_lub = 0;               // the last timed launch ran the lubrication
  int *lub_total_host = nullptr;   // pinned: pair total of the last launch
  unsigned long generation = 0;
};

extern "C" shk_ens *shk_ens_create (void)
{
  shk_ens *e = new (std::nothrow) shk_ens;
  if (!e) { shk_set_error ("ensemble: out of host memory"); return nullptr; }
  if (hipStreamCreateWithFlags (&e->s, hipStreamNonBlocking) != hipSuccess
      || hipStreamCreateWithFlags (&e->s2, hipStreamNonBlocking) != hipSuccess
      || hipEventCreateWithFlags (&e->fork, hipEventDisableTiming) != hipSuccess
      || hipEventCreateWithFlags (&e->join, hipEventDisableTiming) != hipSuccess
      || hipHostMalloc ((void **) &e->lub_total_host, sizeof (int), 0) != hipSuccess)
    {
      (void) hipGetLastError ();
      shk_set_error ("ensemble: no GPU stream");
      shk_ens_destroy (e);
      return nullptr;
    }
  for (int k = 0; k < 5; k++)
    if (hipEventCreate (&e->ev[k]) != hipSuccess)
      {
        (void) hipGetLastError ();
        shk_set_error ("ensemble: no GPU event");
        shk_ens_destroy (e);
        return nullptr;
      }
  *e->lub_total_host = 0;
  return e;
}

extern "C" void shk_ens_destroy (shk_ens *e)
{
  if (!e) return;
  if (e->s) (void) hipStreamSynchronize (e->s);
  if (e->s2) (void) hipStreamSynchronize (e->s2);
  Buf *all[] = {&e->boxes, &e->start, &e->order, &e->box_off, &e->rad, &e->pos4, &e->cand,
                &e->candf, &e->candm, &e->tbits, &e->kv, &e->ks, &e->km, &e->ptab,
                &e->partial, &e->kpart, &e->kpq, &e->kback, &e->lubcoef, &e->lubf1,
                &e->lubcnt, &e->luboff, &e->lublist, &e->lubval};
  for (Buf *b : all) drop (*b);
  if (e->lub_total_host) (void) hipHostFree (e->lub_total_host);
  for (int k = 0; k < 5; k++)
    if (e->ev[k]) (void) hipEventDestroy (e->ev[k]);
  if (e->fork) (void) hipEventDestroy (e->fork);
  if (e->join) (void) hipEventDestroy (e->join);
  if (e->s) (void) hipStreamDestroy (e->s);
  if (e->s2) (void) hipStreamDestroy (e->s2);
  delete e;
}

extern "C" void *shk_ens_stream (shk_ens *e) { return e ? (void *) e->s : nullptr; }
extern "C" unsigned long shk_ens_generation (shk_ens *e) { return e ? e->generation : 0; }

template <typename T> static size_t sf_lds_of (int D, int DP, const PhaseBasis &b, int jt)
{
  const size_t tile = sizeof (Cplx<T>) * (size_t) jt * b.W + sizeof (T) * (size_t) jt * (DP + 1);
  const size_t red = sizeof (T) * 3 * 64 * D;
  return tile > red ? tile : red;
}

template <typename T> static int sf_jt_of (int DP, const PhaseBasis &b)
{
  const size_t per = sizeof (Cplx<T>) * b.W + sizeof (T) * (DP + 1);
  int jt = (int) ((40 * 1024) / per) & ~3;
  return jt > 64 ? 64 : (jt < 4 ? 4 : jt);
}

template <typename Tb> static int upload_vec (hipStream_t s, Buf &b, const std::vector<Tb> &v)
{
  int rc = grow (b, sizeof (Tb) * v.size ());
  if (rc) return rc;
  if (!v.empty ()) ECHK (hipMemcpyAsync (b.p, v.data (), sizeof (Tb) * v.size (),
                                         hipMemcpyHostToDevice, s));
  return SHK_OK;
}

extern "C" int shk_ens_set (shk_ens *e, int nbox, const shk_ens_box *in, int prec)
{
  if (!e || nbox <= 0 || !in) { shk_set_error ("ensemble: no boxes"); return SHK_EARG; }
  if (prec != SHK_FP64 && prec != SHK_MIXED) { shk_set_error ("ensemble: precision %d", prec); return SHK_EARG; }
  const int version = in[0].p.version, poly = in[0].p.poly;
  const int D = dimv (version);
  const int DP = (D + 3) & ~3;
  const bool f32 = prec == SHK_MIXED;
  const size_t tsz = f32 ? sizeof (float) : sizeof (double);
  std::vector<EnsBox> hb (nbox);
  std::vector<double> cand, kv, ks, rad;
  std::vector<int> candm, km;
  std::vector<unsigned> tbits;
  long long ptab = 0, part = 0, kpart = 0, kpq = 0, kback = 0;
  size_t lds = 0;
  int off = 0;
  const int NPOW = poly ? 5 : 1, NS = poly ? 4 : 2;
  for (int b = 0; b < nbox; b++)
    {
      const shk_params &p = in[b].p;
      const shk_tables &t = in[b].t;
      if (p.version != version || p.poly != poly || !p.periodic || p.np <= 0)
        {
          shk_set_error ("ensemble box %d: every box must be periodic with the version and"
                         " radii kind of box 0", b);
          return SHK_EUNSUP;
        }
      if (t.tbits_words < 0 || t.ncand < 0 || p.nk < 0 || (p.nk > 0 && !t.km))
        {
          shk_set_error ("ensemble box %d: lattice tables missing", b);
          return SHK_EARG;
        }
      EnsBox &B = hb[b];
      memset (&B, 0, sizeof (B));
      B.np = p.np;
      B.off = off;
      B.nj = (p.np + 63) / 64;
      B.jw = std::max (1, std::min (in[b].jw, B.nj));
      B.njw = (B.nj + B.jw - 1) / B.jw;
      B.cand_off = (int) (cand.size () / 4);
      B.tbits_off = (int) tbits.size ();
      B.nk = p.nk;
      B.k_off = (int) (kv.size () / 4);
      RealParams &rp = B.rp;
      rp.np = p.np; rp.ncand = t.ncand; rp.ncore = t.ncore; rp.poly = p.poly;
      rp.xi = p.xi; rp.rmax2 = p.rmax2; rp.rmin = p.rmin;
      rp.lx = p.lx; rp.ly = p.ly; rp.lz = p.lz; rp.s2x = t.s2x; rp.s3x = t.s3x;
      rp.tx = t.tx; rp.ty = t.ty; rp.tz = t.tz; rp.tbits_words = t.tbits_words;
      DevParams &q = B.q;
      q.np = p.np; q.nr = p.nr; q.nk = p.nk; q.poly = p.poly;
      q.xi = p.xi; q.rmax2 = p.rmax2; q.rmin = p.rmin;
      q.self_a = p.self_a; q.self_c = p.self_c; q.self_m = p.self_m;
      q.xiaspi = p.xiaspi; q.xia2 = p.xia2; q.gsign = p.gsign;
      const double twopi = 6.28318530717958647692;
      B.b.k0x = twopi / p.lx; B.b.k0y = twopi / p.ly; B.b.k0z = twopi / p.lz;
      B.b.s1 = t.kbasis_s1; B.b.s2 = t.kbasis_s2;
      B.b.K1 = t.kmax1; B.b.K2 = t.kmax2; B.b.K3 = t.kmax3;
      B.b.W = (t.kmax1 + 1) + (2 * t.kmax2 + 1) + (2 * t.kmax3 + 1);
      cand.insert (cand.end (), t.cand, t.cand + 4 * (size_t) t.ncand);
      candm.insert (candm.end (), t.candm, t.candm + 4 * (size_t) t.ncand);
      tbits.insert (tbits.end (), t.tbits, t.tbits + t.tbits_words);
      if (p.nk > 0)
        {
          kv.insert (kv.end (), t.kv, t.kv + 4 * (size_t) p.nk);
          ks.insert (ks.end (), t.ks, t.ks + 8 * (size_t) p.nk);
          km.insert (km.end (), t.km, t.km + 4 * (size_t) p.nk);
        }
      // splits of the reciprocal kernels
      B.sj = std::max (1, std::min (in[b].sj, (p.np + 15) / 16));
      B.jper = (p.np + B.sj - 1) / B.sj;
      B.sj = (p.np + B.jper - 1) / B.jper;
      B.sk = p.nk > 0 ? std::max (1, std::min (in[b].sk, (p.nk + 15) / 16)) : 1;
      B.kper = p.nk > 0 ? (p.nk + B.sk - 1) / B.sk : 0;
      if (p.nk > 0) B.sk = (p.nk + B.kper - 1) / B.kper;
      B.nkb = (p.nk + 63) / 64;
      B.nib = (p.np + 63) / 64;
      B.jt = f32 ? sf_jt_of<float> (DP, B.b) : sf_jt_of<double> (DP, B.b);
      const size_t l = f32 ? sf_lds_of<float> (D, DP, B.b, B.jt) : sf_lds_of<double> (D, DP, B.b, B.jt);
      if (p.nk > 0) lds = std::max (lds, l);
      B.ptab_off = ptab;  ptab += p.nk > 0 ? (long long) p.np * B.b.W : 0;
      B.part_off = part;  part += (long long) B.njw * p.np * D;
      B.kpart_off = kpart; kpart += (long long) B.sj * p.nk * NS * D;
      B.kpq_off = kpq;    kpq += (long long) p.nk * NPOW * 2 * D;
      B.kback_off = kback; kback += p.nk > 0 ? (long long) B.sk * p.np * D : 0;
      for (int j = 0; j < p.np; j++) rad.push_back (in[b].a ? in[b].a[j] : 1.0);
      off += p.np;
    }
  if (lds > 64 * 1024) { shk_set_error ("ensemble: structure-factor tile exceeds LDS"); return SHK_EUNSUP; }
  // launch order: decreasing work (np^2 x candidates), ties by box index
  std::vector<int> order (nbox);
  std::iota (order.begin (), order.end (), 0);
  std::stable_sort (order.begin (), order.end (), [&] (int a, int c) {
    const double wa = (double) hb[a].np * hb[a].np * (hb[a].rp.ncand + 1);
    const double wc = (double) hb[c].np * hb[c].np * (hb[c].rp.ncand + 1);
    return wa > wc;
  });
  std::vector<int> start ((size_t) NPH * (nbox + 1));
  for (int ph = 0; ph < NPH; ph++)
    {
      long long acc = 0;
      for (int s = 0; s < nbox; s++)
        {
          const EnsBox &B = hb[order[s]];
          start[(size_t) ph * (nbox + 1) + s] = (int) acc;
          long long n = 0;
          switch (ph)
            {
            case PH_REAL: n = ((long long) B.np * B.njw + 3) / 4; break;
            case PH_TAB: n = B.nk > 0 ? ((long long) B.np * B.b.W + 255) / 256 : 0; break;
            case PH_SF: n = B.nk > 0 ? (long long) B.nkb * B.sj : 0; break;
            case PH_BLK:
              n = B.nk > 0 ? (poly ? (recip_blocks_threads<5> (B.nk) + 127) / 128
                                   : (recip_blocks_threads<1> (B.nk) + 127) / 128) : 0;
              break;
            case PH_BACK: n = B.nk > 0 ? (long long) B.nib * B.sk : 0; break;
            case PH_COMB: n = ((long long) B.np * D * COMBINE_LANES + 255) / 256; break;
            case PH_LUBW: n = (B.np + 3) / 4; break;
            }
          acc += n;
        }
      if (acc >= 0x7fffffffLL) { shk_set_error ("ensemble: too many workgroups"); return SHK_EUNSUP; }
      start[(size_t) ph * (nbox + 1) + nbox] = (int) acc;
      e->nblk[ph] = acc;
    }
  std::vector<int> box_off (nbox + 1);
  for (int b = 0; b < nbox; b++) box_off[b] = hb[b].off;
  box_off[nbox] = off;
  int rc;
  ECHK (hipStreamSynchronize (e->s));
  ECHK (hipStreamSynchronize (e->s2));
  void *old[] = {e->pos4.p, e->ptab.p, e->partial.p, e->kpart.p, e->kpq.p, e->kback.p,
                 e->boxes.p, e->start.p, e->order.p, e->box_off.p, e->rad.p, e->cand.p,
                 e->candf.p, e->candm.p, e->tbits.p, e->kv.p, e->ks.p, e->km.p};
  if ((rc = upload_vec (e->s, e->start, start))) return rc;
  if ((rc = upload_vec (e->s, e->order, order))) return rc;
  if ((rc = upload_vec (e->s, e->box_off, box_off))) return rc;
  if ((rc = upload_vec (e->s, e->rad, rad))) return rc;
  if ((rc = upload_vec (e->s, e->cand, cand))) return rc;
  if ((rc = upload_vec (e->s, e->candm, candm))) return rc;
  if ((rc = upload_vec (e->s, e->tbits, tbits))) return rc;
  if ((rc = upload_vec (e->s, e->kv, kv))) return rc;
  if ((rc = upload_vec (e->s, e->ks, ks))) return rc;
  if ((rc = upload_vec (e->s, e->km, km))) return rc;
  std::vector<float> candf (cand.begin (), cand.end ());
  if ((rc = upload_vec (e->s, e->candf, candf))) return rc;
  if ((rc = grow (e->pos4, sizeof (double4) * (size_t) off))) return rc;
  if ((rc = grow (e->ptab, (f32 ? sizeof (Cplx<float>) : sizeof (Cplx<double>)) * (size_t) ptab))) return rc;
  if ((rc = grow (e->partial, sizeof (double) * (size_t) part))) return rc;
  if ((rc = grow (e->kpart, tsz * (size_t) kpart))) return rc;
  if ((rc = grow (e->kpq, tsz * (size_t) kpq))) return rc;
  if ((rc = grow (e->kback, tsz * (size_t) kback))) return rc;
  // lubrication pointers are filled in by shk_ens_set_lub
  for (EnsBox &B : hb)
    {
      B.ev.coef = (const double *) e->lubcoef.p;
      B.ev.f1 = (const double *) e->lubf1.p;
    }
  if ((rc = upload_vec (e->s, e->boxes, hb))) return rc;
  ECHK (hipStreamSynchronize (e->s));
  void *now[] = {e->pos4.p, e->ptab.p, e->partial.p, e->kpart.p, e->kpq.p, e->kback.p,
                 e->boxes.p, e->start.p, e->order.p, e->box_off.p, e->rad.p, e->cand.p,
                 e->candf.p, e->candm.p, e->tbits.p, e->kv.p, e->ks.p, e->km.p};
  if (memcmp (old, now, sizeof (old)) != 0) e->generation++;
  e->hb = std::move (hb);
  e->hstart = std::move (start);
  e->hbox_off = std::move (box_off);
  e->nbox = nbox;
  e->ntot = off;
  e->version = version;
  e->poly = poly;
  e->prec = prec;
  e->sf_lds = lds;
  e->have = 1;
  e->have_lub = 0;
  return SHK_OK;
}

extern "C" int shk_ens_set_lub (shk_ens *e, const shk_lub *lp)
{
  if (!e || !e->have || !lp) { shk_set_error ("ensemble: no boxes"); return SHK_EARG; }
  int rc;
  void *old[] = {e->lubcoef.p, e->lubf1.p, e->lubcnt.p, e->luboff.p};
  for (int b = 0; b < e->nbox; b++)
    {
      EnsBox &B = e->hb[b];
      const shk_lub &L = lp[b];
      if (e->poly && (!L.coef || L.N < L.nmax || L.N != lp[0].N))
        { shk_set_error ("ensemble box %d: lubrication coefficient table missing", b); return SHK_EARG; }
      if (!e->poly && !L.f1) { shk_set_error ("ensemble box %d: no f(1) table", b); return SHK_EARG; }
      if (L.lubmin2 != lp[0].lubmin2 || L.nmax != lp[0].nmax || L.flag_lub != lp[0].flag_lub)
        {
          shk_set_error ("ensemble box %d: lubmin2, twobody_nmax and twobody_lub must be those"
                         " of box 0", b);
          return SHK_EUNSUP;
        }
      LubGeom &g = B.lg;
      g.np = B.np;
      g.periodic = 1;
      g.poly = e->poly;
      g.lx = B.rp.lx; g.ly = B.rp.ly; g.lz = B.rp.lz;
      g.s2x = B.rp.s2x; g.s3x = B.rp.s3x;
      g.lubmax2 = L.lubmax * L.lubmax;
      // set_imax_lub_periodic (lub.c:129-162)
      double lm = L.lubmax;
      if (lm <= 0.0) lm = std::max (g.lx, std::max (g.ly, g.lz));
      g.ix = (int) (lm / g.lx) + 1;
      g.iy = (int) (lm / g.ly) + 1;
      g.iz = (int) (lm / g.lz) + 1;
      if (g.ix > 15 || g.iy > 15 || g.iz > 15)
        {
          shk_set_error ("ensemble box %d: more than 15 lubrication image cells per direction", b);
          return SHK_EUNSUP;
        }
      g.minimg = lub_minimg (g, L.lubmax, L.amax);
      B.ev.lubmin2 = L.lubmin2;
      B.ev.nmax = L.nmax;
      B.ev.flag_lub = L.flag_lub;
      B.ev.N = L.N;
    }
  ECHK (hipStreamSynchronize (e->s));
  if (e->poly)
    {
      const int N = lp[0].N;
      const size_t per = (size_t) (N + 1) * (N + 2);
      if ((rc = grow (e->lubcoef, sizeof (double) * 11 * per))) return rc;
      for (int f = 0; f < 11; f++)
        ECHK (hipMemcpyAsync ((double *) e->lubcoef.p + f * per, lp[0].coef[f],
                              sizeof (double) * per, hipMemcpyHostToDevice, e->s));
    }
  else
    {
      if ((rc = grow (e->lubf1, sizeof (double) * 11 * 51))) return rc;
      ECHK (hipMemcpyAsync (e->lubf1.p, lp[0].f1, sizeof (double) * 11 * 51,
                            hipMemcpyHostToDevice, e->s));
    }
  if ((rc = grow (e->lubcnt, sizeof (int) * (size_t) e->ntot))) return rc;
  if ((rc = grow (e->luboff, sizeof (int) * (size_t) (e->ntot + 1)))) return rc;
  for (EnsBox &B : e->hb)
    {
      B.ev.coef = (const double *) e->lubcoef.p;
      B.ev.f1 = (const double *) e->lubf1.p;
    }
  if ((rc = upload_vec (e->s, e->boxes, e->hb))) return rc;
  ECHK (hipStreamSynchronize (e->s));
  void *now[] = {e->lubcoef.p, e->lubf1.p, e->lubcnt.p, e->luboff.p};
  if (memcmp (old, now, sizeof (old)) != 0) e->generation++;
  e->have_lub = 1;
  return SHK_OK;
}

static EnsMap ens_map (shk_ens *e)
{
  return EnsMap{(const int *) e->start.p, (const int *) e->order.p, e->nbox};
}

static unsigned grid_of (long long n) { return (unsigned) (n > 0 ? n : 0); }

template <int V, typename T, int NPOW>
static int ens_recip (shk_ens *e, hipStream_t s, const double *x)
{
  const EnsMap m = ens_map (e);
  const EnsBox *bx = (const EnsBox *) e->boxes.p;
  const double4 *pos4 = (const double4 *) e->pos4.p;
  if (e->nblk[PH_TAB] > 0)
    hipLaunchKernelGGL ((k_ens_phase_tab<T>), dim3 (grid_of (e->nblk[PH_TAB])), dim3 (256), 0, s,
                        m, bx, pos4, (Cplx<T> *) e->ptab.p);
  if (e->nblk[PH_SF] > 0)
    hipLaunchKernelGGL ((k_ens_sf<V, T, NPOW>), dim3 (grid_of (e->nblk[PH_SF])), dim3 (256),
                        e->sf_lds, s, m, bx, pos4, x, (const int4 *) e->km.p,
                        (const Cplx<T> *) e->ptab.p, (T *) e->kpart.p);
  if (e->nblk[PH_BLK] > 0)
    hipLaunchKernelGGL ((k_ens_blocks<V, T, NPOW>), dim3 (grid_of (e->nblk[PH_BLK])), dim3 (128),
                        0, s, m, bx, (const double4 *) e->kv.p, (const double *) e->ks.p,
                        (const T *) e->kpart.p, (T *) e->kpq.p);
  if (e->nblk[PH_BACK] > 0)
    hipLaunchKernelGGL ((k_ens_back<V, T, NPOW>), dim3 (grid_of (e->nblk[PH_BACK])), dim3 (256),
                        0, s, m, bx, pos4, (const double4 *) e->kv.p, (const int4 *) e->km.p,
                        (const T *) e->kpq.p, (T *) e->kback.p);
  ECHK (hipGetLastError ());
  return SHK_OK;
}

template <int V, typename T>
static int ens_atimes_vt (shk_ens *e, const double *x, double *y, int timed)
{
  const EnsMap m = ens_map (e);
  const EnsBox *bx = (const EnsBox *) e->boxes.p;
  const double4 *pos4 = (const double4 *) e->pos4.p;
  hipStream_t s = e->s, sr = timed ? e->s : e->s2;
  if (!timed)
    {
      // the reciprocal phases on the second stream, beside the real space
      ECHK (hipEventRecord (e->fork, s));
      ECHK (hipStreamWaitEvent (sr, e->fork, 0));
    }
  else
    ECHK (hipEventRecord (e->ev[0], s));
  const void *cand = std::is_same<T, double>::value ? e->cand.p : e->candf.p;
  hipLaunchKernelGGL ((k_ens_real<V, T>), dim3 (grid_of (e->nblk[PH_REAL])), dim3 (256), 0, s, m,
                      bx, pos4, x, (const typename Vec4<T>::type *) cand,
                      (const int4 *) e->candm.p, (const unsigned *) e->tbits.p,
                      (double *) e->partial.p);
  ECHK (hipGetLastError ());
  if (timed) ECHK (hipEventRecord (e->ev[1], s));
  int rc = e->poly ? ens_recip<V, T, 5> (e, sr, x) : ens_recip<V, T, 1> (e, sr, x);
  if (rc) return rc;
  if (!timed)
    {
      ECHK (hipEventRecord (e->join, sr));
      ECHK (hipStreamWaitEvent (s, e->join, 0));
    }
  else
    ECHK (hipEventRecord (e->ev[2], s));
  hipLaunchKernelGGL ((k_ens_combine<V, T>), dim3 (grid_of (e->nblk[PH_COMB])), dim3 (256), 0, s,
                      m, bx, pos4, x, (const double *) e->partial.p, (const T *) e->kback.p, y);
  ECHK (hipGetLastError ());
  return SHK_OK;
}

template <int V>
static int ens_lub_v (shk_ens *e, const double *x, double *ylub)
{
  const EnsMap m = ens_map (e);
  const EnsBox *bx = (const EnsBox *) e->boxes.p;
  const double4 *pos4 = (const double4 *) e->pos4.p;
  hipStream_t s = e->s;
  const unsigned nw = grid_of (e->nblk[PH_LUBW]);
  hipLaunchKernelGGL ((k_ens_lub_pairs<0>), dim3 (nw), dim3 (256), 0, s, m, bx, pos4,
                      (int *) e->lubcnt.p, (const int *) nullptr, (int2 *) nullptr, 0);
  hipLaunchKernelGGL (k_lub_scan, dim3 (1), dim3 (1024), 0, s, (const int *) e->lubcnt.p, e->ntot,
                      (int *) e->luboff.p);
  // the pair total travels back with the launch; the host checks it against
  // the list capacity afterwards
  ECHK (hipMemcpyAsync (e->lub_total_host, (const int *) e->luboff.p + e->ntot, sizeof (int),
                        hipMemcpyDeviceToHost, s));
  if (e->lub_cap > 0)
    {
      hipLaunchKernelGGL ((k_ens_lub_pairs<1>), dim3 (nw), dim3 (256), 0, s, m, bx, pos4,
                          (int *) nullptr, (const int *) e->luboff.p, (int2 *) e->lublist.p,
                          e->lub_cap);
      const dim3 g ((unsigned) ((e->lub_cap + 127) / 128));
      const LubEval ev = e->hb[0].ev;
      if (e->poly)
        hipLaunchKernelGGL ((k_ens_lub_eval<V, 1>), g, dim3 (128),
                            sizeof (double) * 4 * (size_t) (ev.N + 2), s, bx,
                            (const int *) e->box_off.p, e->nbox, pos4, (const int *) e->luboff.p,
                            e->ntot, (const int2 *) e->lublist.p, e->lub_cap, ev, x,
                            (double *) e->lubval.p);
      else
        hipLaunchKernelGGL ((k_ens_lub_eval<V, 0>), g, dim3 (128), 0, s, bx,
                            (const int *) e->box_off.p, e->nbox, pos4, (const int *) e->luboff.p,
                            e->ntot, (const int2 *) e->lublist.p, e->lub_cap, ev, x,
                            (double *) e->lubval.p);
    }
  const size_t nd = (size_t) e->ntot * Dim<V>::D;
  hipLaunchKernelGGL (k_lub_gather, dim3 ((unsigned) ((nd + 255) / 256)), dim3 (256), 0, s,
                      (const int *) e->luboff.p, (const double *) e->lubval.p, e->ntot,
                      Dim<V>::D, e->lub_cap, ylub);
  ECHK (hipGetLastError ());
  return SHK_OK;
}

extern "C" int shk_ens_launch (shk_ens *e, const double *pos, const double *x, double *y,
                               double *ylub, int timed)
{
  if (!e || !e->have || !pos || !x) { shk_set_error ("ensemble: not set up"); return SHK_EARG; }
  if (ylub && (!e->have_lub || e->lub_cap <= 0))
    { shk_set_error ("ensemble: lubrication not prepared"); return SHK_EARG; }
  hipLaunchKernelGGL (k_ens_pack, dim3 ((unsigned) ((e->ntot + 255) / 256)), dim3 (256), 0, e->s,
                      pos, (const double *) e->rad.p, e->ntot, (double4 *) e->pos4.p);
  ECHK (hipGetLastError ());
  int rc = 0;
  if (timed && !y)
    for (int k = 0; k < 3; k++) ECHK (hipEventRecord (e->ev[k], e->s));
  if (y)
    {
      const bool f32 = e->prec == SHK_MIXED;
      switch (e->version)
        {
        case 0: rc = f32 ? ens_atimes_vt<0, float> (e, x, y, timed) : ens_atimes_vt<0, double> (e, x, y, timed); break;
        case 1: rc = f32 ? ens_atimes_vt<1, float> (e, x, y, timed) : ens_atimes_vt<1, double> (e, x, y, timed); break;
        default: rc = f32 ? ens_atimes_vt<2, float> (e, x, y, timed) : ens_atimes_vt<2, double> (e, x, y, timed); break;
        }
      if (rc) return rc;
    }
  if (timed) ECHK (hipEventRecord (e->ev[3], e->s));
  if (ylub)
    {
      switch (e->version)
        {
        case 0: rc = ens_lub_v<0> (e, x, ylub); break;
        case 1: rc = ens_lub_v<1> (e, x, ylub); break;
        default: rc = ens_lub_v<2> (e, x, ylub); break;
        }
    }
  if (timed) ECHK (hipEventRecord (e->ev[4], e->s));
  e->timed_lub = timed && ylub;
  return rc;
}

// the pair count of the current positions (eager, waits): sizes the list
// (capacity total * 5/4 + 1024) when it does not fit
extern "C" int shk_ens_lub_prepare (shk_ens *e, const double *pos, int *total_out)
{
  if (!e || !e->have_lub) { shk_set_error ("ensemble: lubrication not set up"); return SHK_EARG; }
  const EnsMap m = ens_map (e);
  const EnsBox *bx = (const EnsBox *) e->boxes.p;
  hipLaunchKernelGGL (k_ens_pack, dim3 ((unsigned) ((e->ntot + 255) / 256)), dim3 (256), 0, e->s,
                      pos, (const double *) e->rad.p, e->ntot, (double4 *) e->pos4.p);
  hipLaunchKernelGGL ((k_ens_lub_pairs<0>), dim3 (grid_of (e->nblk[PH_LUBW])), dim3 (256), 0,
                      e->s, m, bx, (const double4 *) e->pos4.p, (int *) e->lubcnt.p,
                      (const int *) nullptr, (int2 *) nullptr, 0);
  hipLaunchKernelGGL (k_lub_scan, dim3 (1), dim3 (1024), 0, e->s, (const int *) e->lubcnt.p,
                      e->ntot, (int *) e->luboff.p);
  ECHK (hipGetLastError ());
  int total = 0;
  ECHK (hipMemcpyAsync (&total, (const int *) e->luboff.p + e->ntot, sizeof (int),
                        hipMemcpyDeviceToHost, e->s));
  ECHK (hipStreamSynchronize (e->s));
  if (total < 0) { shk_set_error ("ensemble: lubrication pair count overflow"); return SHK_EARG; }
  if (total > e->lub_cap || e->lub_cap == 0)
    {
      const long long cap = (long long) total + total / 4 + 1024;
      if (cap > 0x7fffffffLL) { shk_set_error ("ensemble: too many lubrication pairs"); return SHK_EUNSUP; }
      const int D = dimv (e->version);
      int rc;
      if ((rc = grow (e->lublist, sizeof (int2) * (size_t) cap))) return rc;
      if ((rc = grow (e->lubval, sizeof (double) * D * (size_t) cap))) return rc;
      e->lub_cap = (int) cap;
      e->generation++;
    }
  if (total_out) *total_out = total;
  return SHK_OK;
}

// the pair total of the last launch (after the launch has completed)
extern "C" int shk_ens_lub_last_total (shk_ens *e) { return e ? *e->lub_total_host : -1; }
extern "C" int shk_ens_lub_capacity (shk_ens *e) { return e ? e->lub_cap : 0; }

// per-phase times (ms) of the last timed launch: real space, reciprocal
// space, combine, lubrication (0 .. 3; 0 for a part the launch skipped)
extern "C" int shk_ens_phase_ms (shk_ens *e, double *ms)
{
  if (!e) return SHK_EARG;
  ECHK (hipEventSynchronize (e->ev[4]));
  for (int k = 0; k < 4; k++)
    {
      float t = 0.0f;
      if (k < 3 || e->timed_lub) ECHK (hipEventElapsedTime (&t, e->ev[k], e->ev[k + 1]));
      ms[k] = t;
    }
  return SHK_OK;
}

// work counts of box b for the autotuner's model: real-space candidate
// tests (np x 64 nj x ncand) and k-space (k, particle) products
extern "C" int shk_ens_box_info (shk_ens *e, int b, int *info)
{
  if (!e || b < 0 || b >= e->nbox) return SHK_EARG;
  const EnsBox &B = e->hb[b];
  info[0] = B.np; info[1] = B.rp.ncand; info[2] = B.rp.ncore; info[3] = B.nk;
  info[4] = B.b.W; info[5] = B.jw; info[6] = B.sj; info[7] = B.sk;
  return SHK_OK;
}
