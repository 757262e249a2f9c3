// solve.hip -- dense FP64 solve of the mobility problem with given forces
// on some components and given velocities on the others: the FTS
// natural-mobility and fixed-particle problems of ewald-3fts-matrix.c:570-652
// and 1192-1290 (there: explicit inverse by LAPACK in solve_linear,
// matrix.c:266-312; here: LU factorisation and solve by rocSOLVER).
//
// With the true map A (generalised forces g -> generalised velocities v,
// v = A g) and the components split into K (force known) and H (velocity
// known):   A_HH g_H = v_H - A_HK g_K,   v_K = A_KK g_K + A_KH g_H.
// Both products with the off-diagonal blocks are taken as one full GEMV
// each on a vector with the unknown part zeroed, so only A_HH is gathered.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <vector>
#include "../shk.h"

namespace {

// mat := mat . T on every 11-column particle block, T (FTS) = (F, T, S) ->
// extracted form (trans_mat_ext2ext, ewald-3fts-matrix.c:383-437): column
// 6 <- 2 c6 + c10, columns 7-9 doubled, column 10 <- c6 + 2 c10
__global__ void k_trans_ext (double *__restrict__ mat, int n, int np)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) n * np) return;
  const size_t row = t / np;
  const int j = (int) (t - row * np);
  double *c = mat + row * n + 11 * (size_t) j;
  const double c6 = c[6], c10 = c[10];
  c[6] = 2.0 * c6 + c10;
  c[7] *= 2.0;
  c[8] *= 2.0;
  c[9] *= 2.0;
  c[10] = c6 + 2.0 * c10;
}

// G[r * nh + c] = A[h[r]][h[c]] (row-major; read column-major it is A_HH^T)
__global__ void k_gather_hh (const double *__restrict__ mat, size_t n, const int *__restrict__ h,
                             int nh, double *__restrict__ G)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= nh) return;
  G[(size_t) r * nh + c] = mat[(size_t) h[r] * n + h[c]];
}

// g = b on K, 0 on H
__global__ void k_known (const double *__restrict__ b, const unsigned char *__restrict__ isH,
                         int n, double *__restrict__ g)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) g[k] = isH[k] ? 0.0 : b[k];
}

// r[q] = b[h[q]] - w[h[q]]
__global__ void k_rhs_h (const double *__restrict__ b, const double *__restrict__ w,
                         const int *__restrict__ h, int nh, double *__restrict__ r)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nh) r[q] = b[h[q]] - w[h[q]];
}

// g[h[q]] = r[q]
__global__ void k_scatter_h (const double *__restrict__ r, const int *__restrict__ h, int nh,
                             double *__restrict__ g)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nh) g[h[q]] = r[q];
}

// x = g on H, v on K
__global__ void k_select (const double *__restrict__ v, const double *__restrict__ g,
                          const unsigned char *__restrict__ isH, int n, double *__restrict__ x)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = isH[k] ? g[k] : v[k];
}

struct Scratch {
  std::vector<void *> p;
  ~Scratch ()
  {
    for (void *q : p) (void) hipFree (q);
  }
  template <typename T> int get (T **out, size_t count)
  {
    void *q = nullptr;
    if (hipMalloc (&q, sizeof (T) * (count ? count : 1)) != hipSuccess)
      {
        shk_set_error ("solve: hipMalloc of %zu bytes failed", sizeof (T) * count);
        return SHK_ENOMEM;
      }
    p.push_back (q);
    *out = (T *) q;
    return SHK_OK;
  }
};

unsigned blocks (size_t n) { return (unsigned) ((n + 255) / 256); }

}  // namespace

#define HCHK(call)                                                          \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      {                                                                     \
        shk_set_error ("solve: %s: %s", #call, hipGetErrorString (e_));     \
        return SHK_EHIP;                                                    \
      }                                                                     \
  } while (0)

// mat [n x n] row-major, n = np * dim: the operator's dense matrix (extracted
// form for FTS; transformed here in place).  Particles [0, nm) are mobile:
// their force / torque components are known; their stresslet (FTS) is not.
// Particles [nm, np) are fixed: all their velocity components are known.
// b: known values per component (forces on K, velocities on H); x: the
// result (velocities on K, forces / stresslets on H).
extern "C" int shk_dev_solve_mixed (shk_dev *d, double *mat, int np, int nm, int version,
                                    const double *b, double *x)
{
  const int dim = version == 0 ? 3 : (version == 1 ? 6 : 11);
  const int n = np * dim;
  if (np <= 0 || nm < 0 || nm > np)
    {
      shk_set_error ("solve: invalid np %d / nm %d", np, nm);
      return SHK_EARG;
    }
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  // component partition
  std::vector<int> hidx;
  std::vector<unsigned char> isH (n, 0);
  for (int i = 0; i < np; i++)
    for (int q = 0; q < dim; q++)
      {
        const bool h = i >= nm || q >= 6;
        if (h)
          {
            isH[(size_t) i * dim + q] = 1;
            hidx.push_back (i * dim + q);
          }
      }
  const int nh = (int) hidx.size ();
  if (nh > 65535)
    {
      shk_set_error ("solve: %d known-velocity components exceed the gather grid", nh);
      return SHK_EUNSUP;
    }
  Scratch sc;
  double *g, *w, *v, *r = nullptr, *G = nullptr;
  int *hd = nullptr, *ipiv = nullptr, *info = nullptr;
  unsigned char *isHd;
  int rc;
  if ((rc = sc.get (&g, n)) || (rc = sc.get (&w, n)) || (rc = sc.get (&v, n))
      || (rc = sc.get (&isHd, n)))
    return rc;
  HCHK (hipMemcpyAsync (isHd, isH.data (), n, hipMemcpyHostToDevice, s));
  if (version == 2)
    {
      hipLaunchKernelGGL (k_trans_ext, dim3 (blocks ((size_t) n * np)), dim3 (256), 0, s, mat, n,
                          np);
      HCHK (hipGetLastError ());
    }
  hipLaunchKernelGGL (k_known, dim3 (blocks (n)), dim3 (256), 0, s, b, isHd, n, g);
  HCHK (hipGetLastError ());
  if (nh > 0)
    {
      if ((rc = sc.get (&r, nh)) || (rc = sc.get (&G, (size_t) nh * nh))
          || (rc = sc.get (&hd, nh)) || (rc = sc.get (&ipiv, nh)) || (rc = sc.get (&info, 1)))
        return rc;
      HCHK (hipMemcpyAsync (hd, hidx.data (), sizeof (int) * nh, hipMemcpyHostToDevice, s));
      // w = A g_K; r = v_H - (A g_K)_H
      if ((rc = shk_dev_gemv (d, mat, n, n, g, 0, w))) return rc;
      hipLaunchKernelGGL (k_rhs_h, dim3 (blocks (nh)), dim3 (256), 0, s, b, w, hd, nh, r);
      HCHK (hipGetLastError ());
      hipLaunchKernelGGL (k_gather_hh, dim3 (blocks (nh), nh), dim3 (256), 0, s, mat, (size_t) n,
                          hd, nh, G);
      HCHK (hipGetLastError ());
      rocblas_handle hb = nullptr;
      if (rocblas_create_handle (&hb) != rocblas_status_success)
        {
          shk_set_error ("solve: rocblas_create_handle failed");
          return SHK_EHIP;
        }
      rocblas_set_stream (hb, s);
      // G read column-major is A_HH^T: factor it, solve with the transpose
      rocblas_status st = rocsolver_dgetrf (hb, nh, nh, G, nh, ipiv, info);
      if (st == rocblas_status_success)
        st = rocsolver_dgetrs (hb, rocblas_operation_transpose, nh, 1, G, nh, ipiv, r, nh);
      int hinfo = 0;
      hipError_t he = hipMemcpyAsync (&hinfo, info, sizeof (int), hipMemcpyDeviceToHost, s);
      if (he == hipSuccess) he = hipStreamSynchronize (s);
      rocblas_destroy_handle (hb);
      if (st != rocblas_status_success || he != hipSuccess)
        {
          shk_set_error ("solve: rocsolver getrf/getrs failed (%d)", (int) st);
          return SHK_EHIP;
        }
      if (hinfo != 0)
        {
          shk_set_error ("solve: the known-velocity block is singular (info %d)", hinfo);
          return SHK_EARG;
        }
      hipLaunchKernelGGL (k_scatter_h, dim3 (blocks (nh)), dim3 (256), 0, s, r, hd, nh, g);
      HCHK (hipGetLastError ());
    }
  // v = A g (all components), x = g on H, v on K
  if ((rc = shk_dev_gemv (d, mat, n, n, g, 0, v))) return rc;
  hipLaunchKernelGGL (k_select, dim3 (blocks (n)), dim3 (256), 0, s, v, g, isHd, n, x);
  HCHK (hipGetLastError ());
  HCHK (hipStreamSynchronize (s));
  return SHK_OK;
}

// ---- dense lubricated solves ------------------------------------------------

namespace {

// C = I + B L on the known-force (U, O) columns, -B on the stresslet columns
__global__ void k_lub_assemble (const double *__restrict__ B, double *__restrict__ C, size_t n)
{
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const size_t row = t / n, col = t - row * n;
  if (col % 11 < 6) C[t] += (row == col) ? 1.0 : 0.0;
  else C[t] = -B[t];
}

// g = x on the components with (k % 11 >= 6) == hi, 0 elsewhere
__global__ void k_part (const double *__restrict__ x, int n, int hi, double *__restrict__ g)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) g[k] = ((k % 11 >= 6) == (hi != 0)) ? x[k] : 0.0;
}

// LU solve of the row-major n x n A (overwritten) for one right-hand side r
int lu_solve (hipStream_t s, double *A, int n, double *r, Scratch &sc)
{
  int *ipiv = nullptr, *info = nullptr;
  int rc;
  if ((rc = sc.get (&ipiv, n)) || (rc = sc.get (&info, 1))) return rc;
  rocblas_handle hb = nullptr;
  if (rocblas_create_handle (&hb) != rocblas_status_success)
    {
      shk_set_error ("solve: rocblas_create_handle failed");
      return SHK_EHIP;
    }
  rocblas_set_stream (hb, s);
  // A read column-major is A^T: factor it, solve with the transpose
  rocblas_status st = rocsolver_dgetrf (hb, n, n, A, n, ipiv, info);
  if (st == rocblas_status_success)
    st = rocsolver_dgetrs (hb, rocblas_operation_transpose, n, 1, A, n, ipiv, r, n);
  int hinfo = 0;
  hipError_t he = hipMemcpyAsync (&hinfo, info, sizeof (int), hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipStreamSynchronize (s);
  rocblas_destroy_handle (hb);
  if (st != rocblas_status_success || he != hipSuccess)
    {
      shk_set_error ("solve: rocsolver getrf/getrs failed (%d)", (int) st);
      return SHK_EHIP;
    }
  if (hinfo != 0)
    {
      shk_set_error ("solve: singular matrix (info %d)", hinfo);
      return SHK_EARG;
    }
  return SHK_OK;
}

}  // namespace

// solve_res_lub_3fts_matrix_0 (ewald-3fts-matrix.c:264-330): x = M^-1 (I +
// M L) b; solve_mob_lub_3fts_matrix_0 (731-838): with B = M T and C = I + (M
// T)(L T) the relation B (F, T, S) = C (U, O, E) solved for (U, O, S) given
// (F, T, E): [C_K | -B_H] (U, O, S) = B_K (F, T) - C_H E.  (There: the
// matrices split into blocks and a generalised linear solve with LAPACK's
// inverse; here one LU solve of the assembled system on the GPU.)
extern "C" int shk_dev_solve_lub (shk_dev *d, double *M, double *L, int np, int kind,
                                  const double *b, double *x)
{
  const int n = np * 11;
  if (np <= 0) { shk_set_error ("solve: np %d", np); return SHK_EARG; }
  hipStream_t s = (hipStream_t) shk_dev_stream (d);
  Scratch sc;
  double *g1, *g2, *w1, *w2, *r;
  int rc;
  if ((rc = sc.get (&g1, n)) || (rc = sc.get (&g2, n)) || (rc = sc.get (&w1, n))
      || (rc = sc.get (&w2, n)) || (rc = sc.get (&r, n)))
    return rc;
  const size_t nn = (size_t) n * n;
  hipLaunchKernelGGL (k_trans_ext, dim3 (blocks (nn / 11)), dim3 (256), 0, s, M, n, np);
  hipLaunchKernelGGL (k_trans_ext, dim3 (blocks (nn / 11)), dim3 (256), 0, s, L, n, np);
  HCHK (hipGetLastError ());
  if (kind == 0)
    {
      // r = b + B (L b); B g = r
      if ((rc = shk_dev_gemv (d, L, n, n, b, 0, w1))) return rc;
      if ((rc = shk_dev_gemv (d, M, n, n, w1, 0, w2))) return rc;
      HCHK (hipMemcpyAsync (r, b, sizeof (double) * n, hipMemcpyDeviceToDevice, s));
      if ((rc = shk_dev_axpby (d, r, w2, 1.0, 1.0, n))) return rc;
      if ((rc = lu_solve (s, M, n, r, sc))) return rc;
      HCHK (hipMemcpyAsync (x, r, sizeof (double) * n, hipMemcpyDeviceToDevice, s));
      HCHK (hipStreamSynchronize (s));
      return SHK_OK;
    }
  // r = B (F, T, 0) - (0, 0, E) - B L (0, 0, E)
  hipLaunchKernelGGL (k_part, dim3 (blocks (n)), dim3 (256), 0, s, b, n, 0, g1);
  hipLaunchKernelGGL (k_part, dim3 (blocks (n)), dim3 (256), 0, s, b, n, 1, g2);
  HCHK (hipGetLastError ());
  if ((rc = shk_dev_gemv (d, M, n, n, g1, 0, r))) return rc;
  if ((rc = shk_dev_axpby (d, r, g2, -1.0, 1.0, n))) return rc;
  if ((rc = shk_dev_gemv (d, L, n, n, g2, 0, w1))) return rc;
  if ((rc = shk_dev_gemv (d, M, n, n, w1, 0, w2))) return rc;
  if ((rc = shk_dev_axpby (d, r, w2, -1.0, 1.0, n))) return rc;
  // C = B L (row-major: column-major C^T = L^T B^T), then the assembly
  double *C;
  if ((rc = sc.get (&C, nn))) return rc;
  rocblas_handle hb = nullptr;
  if (rocblas_create_handle (&hb) != rocblas_status_success)
    {
      shk_set_error ("solve: rocblas_create_handle failed");
      return SHK_EHIP;
    }
  rocblas_set_stream (hb, s);
  const double one = 1.0, zero = 0.0;
  rocblas_status st = rocblas_dgemm (hb, rocblas_operation_none, rocblas_operation_none, n, n, n,
                                     &one, L, n, M, n, &zero, C, n);
  rocblas_destroy_handle (hb);
  if (st != rocblas_status_success)
    {
      shk_set_error ("solve: rocblas_dgemm failed (%d)", (int) st);
      return SHK_EHIP;
    }
  hipLaunchKernelGGL (k_lub_assemble, dim3 (blocks (nn)), dim3 (256), 0, s, (const double *) M, C,
                      nn > 0 ? (size_t) n : 0);
  HCHK (hipGetLastError ());
  if ((rc = lu_solve (s, C, n, r, sc))) return rc;
  HCHK (hipMemcpyAsync (x, r, sizeof (double) * n, hipMemcpyDeviceToDevice, s));
  HCHK (hipStreamSynchronize (s));
  return SHK_OK;
}
