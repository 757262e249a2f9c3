/* shk.h -- internal C ABI between the C host layer (csrc/host/ C files) and
 * the HIP layer (csrc/hip/ HIP files).  Plain C types only; every HIP-side entry
 * returns 0 on success or a negative SHK_E* code.
 */
#ifndef SHK_H
#define SHK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SHK_OK = 0,
  SHK_EHIP = -1,      /* HIP runtime error (message via shk_last_error) */
  SHK_EARG = -2,      /* invalid argument / size */
  SHK_ENOMEM = -3,    /* device allocation failed */
  SHK_EUNSUP = -4,    /* configuration not supported by the device path */
  SHK_ENODEV = -5     /* no usable GPU */
};

enum { SHK_FP64 = 0, SHK_MIXED = 1 };

/* Ewald / RPY parameters of one system, host-prepared.  The field meanings
 * follow struct stokes (stokes.h:27-154). */
typedef struct shk_params {
  int version;          /* 0 F, 1 FT, 2 FTS */
  int np;
  int periodic;         /* 0 open-BC RPY, 1 Ewald */
  int poly;             /* radii present */
  double lx, ly, lz, pivol;
  double xi, xi2, xiaspi, xia2;
  double rmax2, rmin;
  double self_a, self_c, self_m;
  int nr;               /* real-space images */
  int nk;               /* reciprocal half-space vectors (k and -k paired) */
  double gsign;         /* sign of the reciprocal G term (+1: as the reference) */
} shk_params;

typedef struct shk_dev shk_dev;

const char *shk_last_error (void);

shk_dev *shk_dev_create (void);
void     shk_dev_destroy (shk_dev *d);

/* Host-prepared lattice tables of a periodic system:
 *  img[nr*4]     reference image vectors (x + shear fold, y, z, 0)
 *  cand[ncand*4] lattice vectors v = m1 a1 + m2 a2 + m3 a3 (a1 = (lx,0,0),
 *                a2 = (s2x,ly,0), a3 = (s3x,0,lz)) with |v| <= r_max +
 *                half the cell diagonal, sorted by |v| (x, y, z, |v|); the
 *                first ncore have |v| <= r_max - half diagonal
 *  candm[ncand*4] their (m1, m2, m3, 0)
 *  tbits         bitmap of the reference image table over
 *                [-tx,tx] x [-ty,ty] x [-tz,tz] (z fastest)
 *  kv[nk*4]      half-space reciprocal vectors (k1, k2, k3, |k|)
 *  ks[nk*8]      2*(ya, yb, yc, yg, yh, ym) at unit radius, kk, kexp (the
 *                factor 2 accounts for the paired -k) */
typedef struct shk_tables {
  const double *img;
  const double *cand;
  const int *candm;
  int ncand, ncore;
  const unsigned *tbits;
  int tbits_words, tx, ty, tz;
  double s2x, s3x;
  const double *kv, *ks;
  const int *km;        /* (m1, m2, m3, first-of-row) per k */
  int kmax1, kmax2, kmax3;          /* max |m| over the k table */
  double kbasis_s1, kbasis_s2;      /* shear shift / lx of mode 1 / 2 */
} shk_tables;

/* Upload positions [3np] (AoS), radii [np] or NULL and the tables. */
int shk_dev_set_system (shk_dev *d, const shk_params *p, const double *pos,
                        const double *a, const shk_tables *t);
/* Replace only the positions (same np). */
int shk_dev_set_pos (shk_dev *d, const double *pos, int on_device);

/* y = M x for nrhs right-hand sides (column-major n x nrhs), device
 * pointers.  prec: SHK_FP64 or SHK_MIXED. */
int shk_dev_atimes (shk_dev *d, const double *x, double *y, int nrhs, int prec);

/* Allocation-free single-column products: between begin and end the
 * scratch of the product (chunk partials, reciprocal partials, P / Q,
 * phase tables) is carved from a per-GPU device array of
 * SHK_NOALLOC_BYTES (no hipMalloc / hipFree); a buffer of the system that
 * would have to grow is refused with SHK_EARG, a scratch need beyond the
 * array with SHK_ENOMEM.  end restores the context's own buffers. */
#define SHK_NOALLOC_BYTES ((size_t) 64 << 20)
#define SHK_MAX_DEVICES 64
int  shk_dev_noalloc_begin (shk_dev *d);
void shk_dev_noalloc_end (shk_dev *d);

/* Lubrication correction y = L x (lub.c:164-768) for version 0 / 1 / 2
 * (3 / 6 / 11 values per particle, device pointers) at the uploaded
 * positions (radii in the positions' fourth component when the system has
 * them).  coef[11] / N: the coefficient polynomials (unequal spheres;
 * uploaded on the first call), f1[11 * 51]: f_k(1), k <= 50 (equal
 * spheres).  lubmin2 / lubmax in r (equal) or s = 2 r / (a1 + a2) units. */
typedef struct shk_lub {
  double lubmin2, lubmax;
  int nmax, flag_lub;
  const double *const *coef;
  int N;
  const double *f1;
  double amax;              /* largest radius (1 for equal spheres; 0: unknown) */
} shk_lub;
int shk_dev_lub (shk_dev *d, int version, const shk_lub *lp, const double *x, double *y);
/* the dense extracted-form matrix of L (make_matrix_lub_3f/3ft/3fts,
 * lub-matrix.c), n x n row-major, overwritten */
int shk_dev_lub_matrix (shk_dev *d, int version, const shk_lub *lp, double *mat);
/* dense lubricated FTS solves from the dense extracted-form matrices M and
 * L (both transformed in place): kind 0 resistance, b = (U, O, E) -> x =
 * M^-1 (I + M L) b = (F, T, S); kind 1 mobility, b = (F, T, E) -> x = (U,
 * O, S) with (F, T, S) = (M^-1 + L) (U, O, E) */
int shk_dev_solve_lub (shk_dev *d, double *M, double *L, int np, int kind, const double *b,
                       double *x);

/* ---- ensembles: many periodic systems, one launch per phase ---------
 * (ens_dev.hip).  A box is given by its parameters and host tables (as
 * for shk_dev_set_system), its radii (NULL: 1) and its tile shape: jw
 * 64-particle chunks per real-space wave, sj particle splits of the
 * structure factors, sk k splits of the back-projection.  Every box must
 * be periodic with box 0's version and kind of radii. */
typedef struct shk_ens shk_ens;
typedef struct shk_ens_box {
  shk_params p;
  shk_tables t;
  const double *a;
  int jw, sj, sk;
} shk_ens_box;
shk_ens *shk_ens_create (void);
void shk_ens_destroy (shk_ens *e);
void *shk_ens_stream (shk_ens *e);
/* changes whenever a device buffer of e moves (captured graphs are stale) */
unsigned long shk_ens_generation (shk_ens *e);
/* upload the boxes (waits); prec SHK_FP64 or SHK_MIXED */
int shk_ens_set (shk_ens *e, int nbox, const shk_ens_box *boxes, int prec);
/* lubrication parameters per box (lp[nbox]; the coefficient table of box
 * 0 serves all boxes) */
int shk_ens_set_lub (shk_ens *e, const shk_lub *lp);
/* count the near pairs at the packed positions pos[3 ntot] (waits) and
 * size the pair list for them */
int shk_ens_lub_prepare (shk_ens *e, const double *pos, int *total);
/* launches only (may be captured): y = M x for every box (y may be NULL),
 * ylub = L x, the lubrication correction (may be NULL; needs
 * shk_ens_lub_prepare); timed: on one stream with events between the
 * phases (shk_ens_phase_ms) */
int shk_ens_launch (shk_ens *e, const double *pos, const double *x, double *y,
                    double *ylub, int timed);
/* pair total of the last completed launch, and the list capacity: a total
 * above the capacity means that launch's ylub is incomplete */
int shk_ens_lub_last_total (shk_ens *e);
int shk_ens_lub_capacity (shk_ens *e);
/* ms[0..3]: real space, reciprocal space, combine, lubrication of the last
 * timed launch (0 for a part it skipped) */
int shk_ens_phase_ms (shk_ens *e, double *ms);
/* info[8]: np, candidates, core candidates, nk, phase-table width, jw, sj, sk */
int shk_ens_box_info (shk_ens *e, int b, int *info);

/* dense (n x n) row-major extracted-form matrix, device pointer */
int shk_dev_matrix (shk_dev *d, double *mat);
/* y = mat . T(x) (FTS, extracted form) or mat . x (F/FT), device pointers */
int shk_dev_extmat (shk_dev *d, const double *mat, const double *x, double *y,
                    int nrhs);

/* y[n1] = mat[n1 x n2] . x (extracted != 0: FTS S transform of x first,
 * n1 = n2 = 11 np), device pointers */
int shk_dev_gemv (shk_dev *d, const double *mat, int n1, int n2,
                  const double *x, int extracted, double *y);

/* device memory and copies on the context stream */
int  shk_dev_malloc (shk_dev *d, void **p, size_t bytes);
int  shk_dev_free (shk_dev *d, void *p);
int  shk_dev_h2d (shk_dev *d, void *dst, const void *src, size_t bytes);
int  shk_dev_d2h (shk_dev *d, void *dst, const void *src, size_t bytes);
int  shk_dev_d2d (shk_dev *d, void *dst, const void *src, size_t bytes);
int  shk_dev_sync (shk_dev *d);
/* wait for all work on the GPU (every stream of every context) */
int  shk_device_sync (void);
void *shk_dev_stream (shk_dev *d);
/* changes whenever a work buffer of d is (re)allocated: a captured graph
 * holding its addresses is stale once this moves */
unsigned long shk_dev_generation (shk_dev *d);

/* per-part device timers of the last operator call (ms): 0 real space,
 * 1 reciprocal space, 2 whole call.  enable != 0 records events. */
void   shk_dev_timing (shk_dev *d, int enable);
double shk_dev_last_ms (shk_dev *d, int which);

/* Chebyshev block helpers (device): b2 = ca*x + cb*b1 - b0 + cy*y, all n*m */
int shk_dev_axpby4 (shk_dev *d, double *b2, const double *x, const double *b1,
                    const double *b0, const double *y, double ca, double cb,
                    double cy, size_t len);
/* plain device copy of `bytes` (a multiple of 32) by a streaming kernel */
int shk_dev_copy_calibrate (shk_dev *d, void *dst, const void *src,
                            size_t bytes);
/* dst = alpha * src (src == NULL: dst = 0), len doubles */
int shk_dev_scal (shk_dev *d, double *dst, const double *src, double alpha,
                  size_t len);
/* out[c] = sum_k u[k + c*n] * v[k + c*n] for c < m (column dots) */
int shk_dev_coldot (shk_dev *d, const double *u, const double *v, int n, int m,
                    double *out_host);

/* copies that do not wait (host side must be pinned for overlap) */
int shk_dev_d2h_async (shk_dev *d, void *dst, const void *src, size_t bytes);
int shk_dev_h2d_async (shk_dev *d, void *dst, const void *src, size_t bytes);
void shk_set_error (const char *fmt, ...);

/* dx = d(g . M(x) f)/dx (3 np) for the periodic F version with equal
 * spheres; f, g, dx device vectors */
int shk_dev_adjoint_x (shk_dev *d, const double *f, const double *g,
                       double *dx, int prec);

/* on: build the dense real-space matrix of the current positions in the
 * precision prec (periodic systems) and use it for the real-space part of
 * every single-rhs product until turned off; off: back to the pair kernels.
 * The caller keeps the positions fixed while it is on. */
/* allow (on != 0) the per-call dense FP32 real-space matrix for MIXED
 * blocks wider than one wave's column tiles */
int shk_dev_set_dense_blocks (shk_dev *d, int on);
/* on != 0 (the default): FP64 block products run their reciprocal part on
 * a second stream beside the real-space part */
int shk_dev_set_overlap (shk_dev *d, int on);
int shk_dev_real_cache (shk_dev *d, int on, int prec);

/* y = a x + b y (len doubles) */
int shk_dev_axpby (shk_dev *d, double *y, const double *x, double a, double b,
                   size_t len);
/* Arnoldi helpers (linalg.hip): out_dev[c] = V[:, c] . w (c < m, V column-
 * major n x m); w += sign V h (h_dev[m]) */
int shk_dev_vtw (shk_dev *d, const double *V, int m, size_t n, const double *w,
                 double *out_dev);
int shk_dev_vh (shk_dev *d, const double *V, const double *h_dev, int m,
                size_t n, double sign, double *w);
/* FTS mobility-problem pieces per particle of 11 (ewald-3fts.c:140-253):
 * lo = (x_0..5, 0), hi = (0, x_6..10) (either may be NULL);
 * y = lo(a) - z (mode 0) or z - hi(a) (mode 1) */
/* FTS mixed problems, particle i mobile when i < nm (11 values each):
 * split: lo = [(x_0..5, 0)_mobile, 0_fixed], hi = [(0, x_6..10)_mobile,
 * x_fixed] (either may be NULL); combine mode 0: y = [(a_0..5, 0)_mobile,
 * 0_fixed] - z; mode 1: y = z - [(0, a_6..10)_mobile, a_fixed] */
int shk_dev_fts_split (shk_dev *d, const double *x, int np, int nm, double *lo,
                       double *hi);
/* the same for D values per particle whose first qs are a mobile
 * particle's known forces (FTS: D = 11, qs = 6; F, FT: qs = D) */
int shk_dev_mix_split (shk_dev *d, const double *x, int np, int nm, int D, int qs,
                       double *lo, double *hi);
int shk_dev_mix_combine (shk_dev *d, const double *a, const double *z, int np,
                         int nm, int D, int qs, int mode, double *y);
int shk_dev_fts_combine (shk_dev *d, const double *a, const double *z, int np,
                         int nm, int mode, double *y);

/* Dense FP64 solve with known forces on the mobile particles' force /
 * torque components and known velocities elsewhere (solve.hip); mat is the
 * n x n operator matrix (extracted form for FTS, transformed in place), b
 * and x device vectors of n = np * dim; waits for the result. */
int shk_dev_solve_mixed (shk_dev *d, double *mat, int np, int nm, int version,
                         const double *b, double *x);

/* ---- Brownian-dynamics step pieces (bd_kernels.hip) ---------------- */

/* step scalars read by the kernels from device memory */
typedef struct shk_bd_scalars {
  double fact;     /* sqrt(2 / (Pe dt)) */
  double dt;       /* step length */
} shk_bd_scalars;

/* per-step results written by the kernels */
typedef struct shk_bd_status {
  double yy;                         /* y.y of the Gaussian vector */
  double zmz;                        /* z.(M z) of the Chebyshev result */
  unsigned long long overlap[2];     /* overlap counts at x_mid, x_new */
} shk_bd_status;

/* out[0..n) ~ N(0,1) from (seed, *counter_dev) */
int shk_dev_gauss (shk_dev *d, unsigned long long seed,
                   const unsigned long long *counter_dev, size_t n, double *out);
int shk_dev_counter_add (shk_dev *d, unsigned long long *counter_dev,
                         unsigned long long inc);
int shk_dev_bd_status_clear (shk_dev *d, shk_bd_status *st);
/* f = F + fact z (F or z may be NULL) */
int shk_dev_bd_force (shk_dev *d, double *f, const double *F, const double *z,
                      const shk_bd_scalars *sc, size_t n);
/* xo = xi + frac dt u */
int shk_dev_bd_euler (shk_dev *d, double *xo, const double *xi, const double *u,
                      const shk_bd_scalars *sc, double frac, size_t n);
/* FT / FTS, np particles: f[6 i + q] = (F | T)[3 i + q % 3] + fact z[6 i + q]
 * (F, T, z may be NULL) */
int shk_dev_bd_force_ft (shk_dev *d, double *f, const double *F, const double *T,
                         const double *z, const shk_bd_scalars *sc, int np);
/* FT / FTS: xo = xi + frac dt U, and qo = qi + frac dt dQ/dt(qi, O) when qi
 * is not NULL; u = (U, O) per particle of 6 */
int shk_dev_bd_euler_ft (shk_dev *d, double *xo, const double *xi, double *qo,
                         const double *qi, const double *u,
                         const shk_bd_scalars *sc, double frac, int np);
/* y[k] *= s[k] */
int shk_dev_dmul (shk_dev *d, double *y, const double *s, size_t n);
/* an imposed flow: Ui, Oi and Ei in the extracted form (xx xy xz yz yy) */
typedef struct shk_flow {
  double U[3], O[3], E[5];
} shk_flow;
/* v[D i + q] += sign * (the flow at particle i, positions p[3 i]) for q in
 * [qlo, qhi), i < n: q < 3 the velocity Ui + Oi x p + Ei . p, q < 6 Oi,
 * q < 11 Ei */
int shk_dev_bd_imposed (shk_dev *d, double *v, const double *p, int n, int D,
                        int qlo, int qhi, const shk_flow *fl, double sign);
/* (F, T) of 6 per particle -> (F, T, 0) of 11 (mode 0), or the first 6 of
 * 11 back (mode 1) */
int shk_dev_ft_fts (shk_dev *d, const double *x, int np, int mode, double *y);
/* st->yy = y.y, st->zmz = z.w */
int shk_dev_bd_dots (shk_dev *d, const double *y, const double *z,
                     const double *w, size_t n, shk_bd_status *st);
/* *count += overlapping (pair, image) entries, brownian.c:299-375 */
int shk_dev_bd_overlap (shk_dev *d, const double *pos, const double *a, int np,
                        int periodic, double lx, double ly, double lz,
                        int shear_mode, double shift, double rfac,
                        unsigned long long *count);

/* stream capture into an executable graph */
int  shk_dev_capture_begin (shk_dev *d);
int  shk_dev_capture_end (shk_dev *d, void **exec);
void shk_dev_capture_abort (shk_dev *d);
int  shk_dev_graph_launch (shk_dev *d, void *exec);
void shk_dev_graph_destroy (void *exec);
/* launches of d go to `stream` until swapped back; returns the previous */
void *shk_dev_swap_stream (shk_dev *d, void *stream);
/* plain streams, events and graphs (streams.hip) */
int  shk_stream_create (void **s);
void shk_stream_destroy (void *s);
int  shk_stream_sync (void *s);
int  shk_event_create (void **e);
void shk_event_destroy (void *e);
int  shk_event_record (void *e, void *s);
int  shk_stream_wait (void *s, void *e);
int  shk_stream_capture_begin (void *s);
int  shk_stream_capture_end (void *s, void **exec);
void shk_stream_capture_abort (void *s);
int  shk_graph_launch_on (void *exec, void *s);
/* pinned host memory */
int  shk_host_alloc (void **p, size_t bytes);
void shk_host_free (void *p);

#ifdef __cplusplus
}
#endif
#endif
