/* stokes_hip.h -- entry points of the MI355X libstokes operator that have
 * no counterpart in the reference API: int status codes, device-resident
 * buffers, several right-hand sides per launch, mixed precision, Lanczos
 * bounds and the device-resident Chebyshev recursion.
 *
 * Every function returns 0 on success or a negative STOKES_HIP_E* code;
 * stokes_hip_last_error() gives the message.  Pointers named *_dev are
 * device (HBM) pointers, e.g. from stokes_hip_malloc().  Vectors are laid
 * out as in the reference (particle-major, 3/6/11 components per particle,
 * extracted-form S/E); a block of nrhs vectors is stored column after
 * column (vector c at offset c*n).
 */
#ifndef LIBSTOKES_HIP_STOKES_HIP_H
#define LIBSTOKES_HIP_STOKES_HIP_H

#include <stddef.h>
#include "stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  STOKES_HIP_OK = 0,
  STOKES_HIP_EHIP = -1,
  STOKES_HIP_EARG = -2,
  STOKES_HIP_ENOMEM = -3,
  STOKES_HIP_EUNSUP = -4,
  STOKES_HIP_ENODEV = -5
};

/* arithmetic of the operator of sys:
 *  FP64  -- every part in FP64 (parity path, reference to ~1e-14)
 *  MIXED -- real-space pairs and reciprocal phases/sums in FP32 with FP64
 *           displacements, partial sums and self term */
enum { STOKES_HIP_FP64 = 0, STOKES_HIP_MIXED = 1 };

const char *stokes_hip_last_error (void);

int stokes_hip_set_precision (struct stokes *sys, int mode);

/* The reference sums the reciprocal part of the FTS G / G-tilde couplings
 * (F -> E, S -> U) with sin(k.(x_j - x_i)) where the real-space part
 * implies the opposite sign (ewald.c:1037-1068 pass sf * yg with
 * sf = -sin), so those two couplings change with xi.  By default the
 * library reproduces the reference; enable != 0 uses the other sign, which
 * makes the FTS sum independent of xi (to the Ewald tolerance). */
int stokes_hip_set_recip_g_fix (struct stokes *sys, int enable);
/* Compute the operator at the Ewald splitting xi_by_tratio(sys, tratio)
 * (stokes.c:765-776; r_max, k_max and the image table from the same
 * set-up rule and sys->ewald_eps) instead of sys->xi.  The Ewald sum does
 * not depend on the splitting, to the Ewald tolerance, except for the
 * reference's reciprocal G sign of the FTS version; that term is restored
 * at sys->xi (equal and unequal spheres).  Results therefore match the
 * reference at sys->xi
 * within the Ewald tolerance while the real/reciprocal balance is chosen
 * for speed.  tratio = 0 computes at sys->xi (the default). */
int stokes_hip_set_split (struct stokes *sys, double tratio);

/* enable: build the real-space part of M at the current positions as a
 * dense matrix ((D np)^2 values in the current precision) and use it for
 * the real-space sum of every product (one GEMV per column; several
 * columns per call: one GEMM, rocBLAS without atomics) until disabled or until
 * the system's parameters change (splitting, precision, box).  The caller
 * keeps the positions fixed meanwhile (products re-read them but the
 * matrix is not rebuilt).  Pays off where many products share one
 * configuration and the pair sum costs more than reading the matrix (the
 * F version: 9 values per pair).  No counterpart in the reference. */
int stokes_hip_set_real_cache (struct stokes *sys, int enable);
/* enable: MIXED products of more columns than one wave covers (32 FTS, 64
 * FT, 128 F) write the real-space blocks once per call as a dense FP32
 * matrix ((D np)^2 values, at most 4 GiB, rebuilt from the positions on
 * every call) and multiply it with the block on the matrix cores, instead
 * of evaluating every pair once per 32-column group matrix-free.  Off by
 * default (matrix-free).  No counterpart in the reference. */
int stokes_hip_set_dense_blocks (struct stokes *sys, int enable);
/* enable (the default): FP64 products of more than one column run the
 * reciprocal (wave-space) part on a second stream of the library beside the
 * real-space part, joined before the two are added; 0 runs them one after
 * the other on the library's stream.  MIXED products always run them one
 * after the other (faster there).  The results do not depend on it.  No
 * counterpart in the reference. */
int stokes_hip_set_overlap (struct stokes *sys, int enable);
/* time reps operator calls of nrhs columns at each candidate tratio (NULL:
 * 0 (= sys->xi), 10, 100, 300, 1000, 3000, 1e4) and keep the fastest as
 * the splitting of sys */
int stokes_hip_autotune (struct stokes *sys, int nrhs, const double *trs,
                         int ntr, int reps, double *best_tr, double *best_ms);
/* record per-part GPU times (ms) into sys->cpu1/cpu2/cpu3 */
int stokes_hip_set_timing (struct stokes *sys, int enable);

/* device memory on the GPU of the library; the copies wait for all queued
 * work first, and operator calls on device pointers are asynchronous
 * (stokes_hip_sync or a copy orders them) */
int stokes_hip_malloc (void **p, size_t bytes);
int stokes_hip_free (void *p);
int stokes_hip_memcpy_h2d (void *dst_dev, const void *src, size_t bytes);
int stokes_hip_memcpy_d2h (void *dst, const void *src_dev, size_t bytes);
/* wait for all work queued for sys */
int stokes_hip_sync (struct stokes *sys);
/* copy `bytes` (a multiple of 32) between device buffers with a plain
 * streaming kernel; a known-size reference for hardware byte counters */
int stokes_hip_copy_calibrate (void *dst_dev, const void *src_dev,
                               size_t bytes);

/* take positions [3*np] from device memory for the following calls
 * (instead of re-reading sys->pos on every call) until the next call of
 * this function or stokes_hip_use_host_pos() */
int stokes_hip_set_pos_dev (struct stokes *sys, const double *pos_dev);
int stokes_hip_use_host_pos (struct stokes *sys);

/* y = M x for nrhs vectors, device pointers */
int stokes_hip_atimes_dev (struct stokes *sys, int nrhs, const double *x_dev,
                           double *y_dev);
/* y = M x for one vector (device pointers) without allocating: no
 * hipMalloc / hipFree or allocating library call.  No counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  The product's scratch lives in one
 * 64 MiB device array per GPU, held until the process ends; a system whose
 * product needs more is refused (STOKES_HIP_ENOMEM).  Every such call on
 * one GPU shares that array, so calls that overlap on one GPU race.  The
 * system must have been set up on the device by an ordinary product (e.g.
 * stokes_hip_atimes_dev) with its current parameters; positions are read
 * as by stokes_hip_atimes_dev.  Equal spheres only (sys->a == NULL). */
int stokes_hip_atimes_dev_noalloc (struct stokes *sys, const double *x_dev,
                                   double *y_dev);
/* int-returning form of atimes_3all (host pointers) */
int stokes_hip_atimes (struct stokes *sys, int n, const double *x, double *y);

/* dense extracted-form matrix (n x n, row-major) into device memory */
int stokes_hip_make_matrix_dev (struct stokes *sys, double *mat_dev);
/* int-returning form of make_matrix_mob_3all (host pointer) */
int stokes_hip_make_matrix (struct stokes *sys, double *mat);
/* y = mat . T(x) for nrhs vectors (FTS) or mat . x (F/FT) */
int stokes_hip_extmat_dev (struct stokes *sys, const double *mat_dev,
                           int nrhs, const double *x_dev, double *y_dev);

/* Gradients of the scalar g . M(x) f for the periodic F version with equal
 * spheres (device vectors of 3 np): dx_dev = d/dx (adjoint kernels that
 * recompute the pair scalars and their radial derivatives; the real-space
 * sum of ewald.c:649-789 and the reciprocal sum of ewald.c:923-1075
 * differentiated), df_dev = d/df = M g.  Either output may be NULL.  Runs
 * in the precision of sys (stokes_hip_set_precision).  No counterpart in
 * the reference. */
int stokes_hip_adjoint_dev (struct stokes *sys, const double *g_dev,
                            const double *f_dev, double *dx_dev,
                            double *df_dev);

/* xa, ya of scalars_ewald_real for the F version (ewald.c:80-99) and
 * their derivatives in r: out = (xa, ya, dxa/dr, dya/dr), as the adjoint
 * evaluates them */
void stokes_hip_scalars_real_f_deriv (double xi, double r, double *out);

/* Dense FP64 mixed problem on the device: particles [0, nm) are mobile
 * with known force / torque (and, FTS, known strain E), particles
 * [nm, np) are fixed with known velocity, angular velocity (and E).
 * b_dev[np * dim] holds the known values per particle in the reference's
 * order (mobile: F T E, fixed: U O E), x_dev the unknowns (mobile: U O S,
 * fixed: F T S); dim = 3, 6, 11 for F, FT, FTS.  The same problem as
 * solve_mix_3fts_matrix_0 (ewald-3fts-matrix.c:1192-1290) for every
 * version; LU on the GPU in place of the explicit LAPACK inverse. */
int stokes_hip_solve_mix_dev (struct stokes *sys, const double *b_dev,
                              double *x_dev);

/* FTS natural mobility problem by GMRES on the device (the libiter route
 * of solve_mob_3fts_0, ewald-3fts.c:263-305): fte_dev[11 np] holds F, T, E
 * per particle, uos_dev[11 np] receives U, O, S; iterations and the final
 * relative residual |b - A x| / |b| are returned when the pointers are not
 * NULL.  Parameters from sys->it (stokes_set_iter). */
int stokes_hip_solve_mob_3fts_dev (struct stokes *sys, const double *fte_dev,
                                   double *uos_dev, int *iters,
                                   double *relres);

/* The same route for the mixed problem with the fixed particles
 * sys->nm..np-1 (solve_mix_3fts_0, ewald-3fts.c:531-593): b_dev[11 np]
 * holds F, T, E of the mobile and U, O, E of the fixed particles, x_dev
 * receives U, O, S and F, T, S (the layout of stokes_hip_solve_mix_dev). */
int stokes_hip_solve_mix_3fts_dev (struct stokes *sys, const double *b_dev,
                                   double *x_dev, int *iters, double *relres);

/* The resistance problem M (F, T, S) = (U, O, E) by GMRES
 * (solve_res_3fts_0, ewald-3fts.c:50-80): uoe_dev[11 np] in, fts_dev out. */
int stokes_hip_solve_res_3fts_dev (struct stokes *sys, const double *uoe_dev,
                                   double *fts_dev, int *iters,
                                   double *relres);

/* Chebyshev-Clenshaw z = f(M) y for nrhs vectors held on the device
 * (chebyshev.c:163-228 with the whole recursion on the GPU) */
int stokes_hip_cheb_eval_dev (struct stokes *sys, int ncheb, const double *a,
                              int nrhs, const double *y_dev, double *z_dev,
                              double eig0, double eig1);

/* FP64 Lanczos estimate of the extreme eigenvalues of the F-version (or
 * FT) mobility of sys; replaces the ARPACK bounds of dnaupd_c.c:909-1117
 * as used by brownian.c:1690-1699.  lmin/lmax are widened by the given
 * relative margin. */
int stokes_hip_lanczos_bounds (struct stokes *sys, int niter,
                               unsigned long seed, double margin,
                               double *lmin, double *lmax);

/* ---- ensembles of independent systems -----------------------------------
 * The operator of nbox periodic systems (uneven np, radii and box length;
 * one version, all with radii or all without) as one graph launch of
 * batched kernels: each phase of the operator (real-space pair sum, phase
 * tables, structure factors, k-blocks, back-projection, combine) and of the
 * FP64 lubrication correction (near-pair count, scan, fill, evaluate,
 * gather) is one kernel over the work items of every box.  Box b reads
 * positions at pos_dev + 3 off_b and x, y at x_dev / y_dev + dim off_b,
 * off_b the sum of np over the boxes before b (stokes_hip_ensemble_offset).
 * Positions are re-read on every launch; the other parameters of the
 * systems (box, radii, Ewald parameters, lubrication parameters) are those
 * of the first launch until stokes_hip_ensemble_refresh().  A member
 * system's splitting (stokes_hip_set_split), G-sign fix and, for box 0, its
 * precision are taken at creation.  Each box's result is that of its own
 * product to the precision's bar; calls on the member systems must not
 * overlap an ensemble launch.  No counterpart in the reference. */
typedef struct stokes_hip_ensemble stokes_hip_ensemble;
int stokes_hip_ensemble_create (struct stokes *const *sys, int nbox,
                                stokes_hip_ensemble **out);
/* y = M x for every box (the precision of the ensemble) */
int stokes_hip_ensemble_atimes (stokes_hip_ensemble *e, const double *pos_dev,
                                const double *x_dev, double *y_dev);
/* ylub = L x, the lubrication correction of calc_lub_3f / 3ft / 3fts
 * (lub.c:164-768, FP64) for every box with its own lubmin2, lubmax,
 * twobody_nmax and twobody_lub */
int stokes_hip_ensemble_lub (stokes_hip_ensemble *e, const double *pos_dev,
                             const double *x_dev, double *ylub_dev);
/* both in one launch */
int stokes_hip_ensemble_atimes_lub (stokes_hip_ensemble *e, const double *pos_dev,
                                    const double *x_dev, double *y_dev,
                                    double *ylub_dev);
int stokes_hip_ensemble_refresh (stokes_hip_ensemble *e);
/* STOKES_HIP_FP64 or STOKES_HIP_MIXED for every box */
int stokes_hip_ensemble_set_precision (stokes_hip_ensemble *e, int mode);
/* splitting tratio (0: the box's own xi) of box `box`, or of every box
 * when box = -1; results stay those of the box's own xi */
int stokes_hip_ensemble_set_split (stokes_hip_ensemble *e, int box, double tratio);
/* tile shape of the batched kernels: jw 64-particle chunks per real-space
 * wave, jper particles per structure-factor split, kper k vectors per
 * back-projection split */
int stokes_hip_ensemble_set_tiles (stokes_hip_ensemble *e, int jw, int jper, int kper);
int stokes_hip_ensemble_get_tiles (const stokes_hip_ensemble *e, int *jw, int *jper,
                                   int *kper);
/* the autotuner: times the batch at every candidate splitting trs[ntr]
 * (NULL: 10, 30, 100, 300, 1000, 3000) for all boxes alike, fits the
 * real-space and reciprocal phase times to the boxes' work at each
 * splitting, gives each box its cheapest splitting under that fit (kept
 * when the whole launch beats the best common splitting), then sweeps the
 * tile shape; best of reps timed launches each.  best_tr[nbox] (may be
 * NULL) receives the splittings kept.  Uses the member systems' host
 * positions. */
int stokes_hip_ensemble_autotune (stokes_hip_ensemble *e, const double *trs,
                                  int ntr, int reps, double *best_tr);
/* one eager launch on one stream with events between the phases (y_dev
 * or ylub_dev may be NULL to skip that product): ms[0] real space, ms[1]
 * reciprocal space, ms[2] combine, ms[3] lubrication (0 when skipped) */
int stokes_hip_ensemble_phase_ms (stokes_hip_ensemble *e, const double *pos_dev,
                                  const double *x_dev, double *y_dev, double *ylub_dev,
                                  double *ms);
size_t stokes_hip_ensemble_offset (const stokes_hip_ensemble *e, int box);
void stokes_hip_ensemble_destroy (stokes_hip_ensemble *e);

/* ---- Brownian dynamics ------------------------------------------------
 * The mid-point step of BD_evolve_mid (brownian.c:2352-2560) for the F, FT
 * and FTS versions (the BD_params fields it reads are the arguments of
 * stokes_hip_bd_create).  M is the mobility of
 * BD_atimes_mob_FU (brownian.c:494-739): F -> U (F), (F, T) -> (U, O)
 * particle-wise (FT), and for FTS the mobility problem with E = 0 solved
 * by GMRES for each product; the drift is that of solve_mix_3all in the
 * labo frame: with fixed particles it holds them at rest, and with an
 * imposed flow (sys->Ui, Oi, Ei, read at each step) it adds the flow at the
 * positions of each drift product (FTS: E = -Ei in the rest frame), the
 * noise staying in the rest frame.  For F and FT with every
 * particle mobile an attempt runs on the GPU as one captured graph (FTS,
 * and fixed particles: the same launches, eagerly, the GMRES steered from
 * the host): Gaussian draw y, z = M^{-1/2} y by the Chebyshev recursion
 * (BD_calc_FB, brownian.c:1545-1874), its accuracy |z.Mz - y.y| / z.Mz,
 * drift at x, Euler half step, overlap count, drift at the mid point,
 * Euler full step, overlap count.  The host then takes the reference's
 * decisions: refit the spectral bounds once when the accuracy exceeds eps
 * (brownian.c:1707-1727, FP64 Lanczos in place of ARPACK), halve dt on
 * overlap and retry with the same noise, or redraw the noise and restore
 * dt once dt falls below dt_lim (brownian.c:2435-2459, 2501-2525).  As in
 * the reference, a retry after an overlap in the second stage takes its
 * first-stage product at the failed attempt's intermediate configuration
 * (REDO_scale sets no positions, brownian.c:2400, 2523).
 *
 * Noise: Philox4x32-10 with Box-Muller (the reference draws from MT19937
 * with the polar method, KIrand.c:108-135), so trajectories follow the
 * reference's distribution, not its random stream. */
typedef struct stokes_hip_bd stokes_hip_bd;

typedef struct stokes_hip_bd_info {
  double dt;          /* step taken (the return value of BD_evolve_mid) */
  double err_cheb;    /* Chebyshev accuracy of the accepted noise */
  double eig[2];      /* spectral bounds in use */
  int attempts;       /* step graphs run */
  int halvings;       /* dt halvings after overlaps */
  int redraws;        /* noise redraws */
  int refits;         /* spectral-bound refits */
  unsigned long long overlaps[2];  /* overlap counts of the last attempt at
                                      x_mid and x_new */
} stokes_hip_bd_info;

/* Hold the real-space part of M(x) as a dense matrix ((D N)^2 values in
 * the system's precision, D = 3, 6, 11) for the products at x of each
 * step -- the Chebyshev recursion, its accuracy check and the drift at x,
 * 32 of the 33 products of a step (F, FT) -- and for the Lanczos products
 * of a refit; the
 * product at x_mid stays matrix-free.  On by default for periodic systems
 * whose matrix stays below 4 GiB in FP64.  No counterpart in the reference
 * (its products are matrix-free). */
int stokes_hip_bd_set_real_cache (stokes_hip_bd *bd, int enable);

/* sys: version 0 (F), 1 (FT) or 2 (FTS), set up as for atimes_3all (a
 * sheared system must be periodic and stepped by stokes_hip_bd_step_t); particles nm..np-1 (stokes_set_np) are fixed: the step
 * moves x [3 nm] (and q [4 nm]), the fixed positions are read from sys->pos
 * at each step.  Unequal radii (sys->a, read at creation): the noise
 * operator carries BD_atimes_mob_FU's Global scaling (U / a, O / a^3,
 * brownian.c:430-480) and the overlap test a_i + a_j.  F_dev: constant force [3 nm] (BD->F)
 * or NULL; it is read on
 * every step and must stay allocated until stokes_hip_bd_destroy.  peclet:
 * fact = sqrt(2 / (Pe dt)), 0 when Pe <= 0 (brownian.c:2177-2187).
 * n_cheb: order of the 1/sqrt(x) fit (BD->n_minv > 0).  eps: refit
 * threshold (BD->eps).  dt_lim, rmin_ol: BD->dt_lim and BD->rmin (pairs
 * with r^2 < rmin_ol (a_i+a_j)^2 overlap; checked only when sys->rmin is
 * 0, as in the reference).  seed: key of the noise generator.
 * FTS: the drift's known strain is the imposed flow's alone (E = -Ei in the
 * rest frame), i.e. the reference's BD->E (brownian.c:146, 2304, 2795) is
 * taken as zero; there is no entry point for a nonzero BD->E, so a caller
 * that needs one cannot use this integrator. */
int stokes_hip_bd_create (struct stokes *sys, const double *F_dev,
                          double peclet, int n_cheb, double eps,
                          double dt_lim, double rmin_ol,
                          unsigned long long seed, stokes_hip_bd **out);
/* spectral bounds of M for the Chebyshev fit (BD->eig_minv); without this
 * call the first step estimates them by Lanczos */
int stokes_hip_bd_set_bounds (stokes_hip_bd *bd, double lmin, double lmax);
/* enable != 0: recompute fact when a redraw restores dt (the reference
 * keeps the value of the halved dt, brownian.c:2446-2449, 2512-2515) */
int stokes_hip_bd_set_fact_fix (stokes_hip_bd *bd, int enable);
/* one step from the positions x_dev [3 nm] (updated in place), step dt */
int stokes_hip_bd_step (stokes_hip_bd *bd, double *x_dev, double dt,
                        stokes_hip_bd_info *info);
/* the same with the orientations (BD_evolve_mid's q, brownian.c:2352):
 * q_dev [4 np] quaternions updated in place by q += dt dQ/dt(q, O)
 * (evolve_Euler_3all, brownian.c:1889-1930, quaternion_dQdt,
 * ode-quaternion.c:89-105) for FT and FTS; NULL, or the F version,
 * tracks none */
int stokes_hip_bd_step_q (stokes_hip_bd *bd, double *x_dev, double *q_dev,
                          double dt, stokes_hip_bd_info *info);
/* the same at time t: a sheared system (sys->shear_mode 1 or 2, periodic)
 * takes its cell shift from t as BD_evolve_mid / BB03 / BM97 do
 * (stokes_set_shear_shift at t for the noise and the first stage, at t + h
 * for the second, h = dt / 2, dt / n or dt; BB03's deterministic product
 * at t with the intermediate positions, brownian.c:2393, 2468, 2741, 2805,
 * 3092), with the reference of stokes_hip_bd_set_shear_ref; a sheared
 * system must be stepped by this call.  The attempts of a sheared system
 * run eagerly (the shift changes between the products) */
int stokes_hip_bd_step_t (stokes_hip_bd *bd, double *x_dev, double *q_dev,
                          double t, double dt, stokes_hip_bd_info *info);
/* BD_set_shear_shift_ref (brownian.c:274-280): the cell shift s0 at time
 * t0 (default 0, 0) */
int stokes_hip_bd_set_shear_ref (stokes_hip_bd *bd, double t0, double s0);
/* draw the noise from the reference's own stream instead of the device
 * generator: MT19937 seeded as KIrand_init_genrand (KIrand.c:137-157) with
 * the polar method of KIrand_Gaussian (KIrand.c:108-135), the n = 3 nm or
 * 6 nm draws of an attempt made on the host in BD_calc_FB's order and
 * copied to the device by the attempt; trajectories then follow the
 * reference's random stream */
int stokes_hip_bd_set_rng_mt19937 (stokes_hip_bd *bd, unsigned long seed);
/* the time-stepping scheme of BD->scheme (brownian.h:77-85, the dispatch
 * of brownian.c:2108-2125): 0 the mid-point algorithm (BD_evolve_mid,
 * the default), 1 Banchio-Brady 2003 with step parameter bb_n = BD->BB_n
 * (BD_evolve_BB03, brownian.c:2579-2920; its deterministic velocity is
 * evaluated at the intermediate configuration, as the reference does,
 * 2770-2816), 2 Ball-Melrose 1997 (BD_evolve_BM97, 2970-3190) */
int stokes_hip_bd_set_scheme (stokes_hip_bd *bd, int scheme, double bb_n);
/* BD_params_init's flag_lub (brownian.c:154): with enable the drift of
 * every stage is the lubricated mixed problem of solve_mix_3all
 * (solve_mix_lub_3f / 3ft / 3fts, ode.c:41-160; stokes_hip_solve_mix_lub_dev
 * over the correction of calc_lub_3f / 3ft / 3fts with sys's lubrication
 * parameters), and the attempts run eagerly (GMRES steered from the host).
 * The noise stays M^-1/2 y, as in the reference, whose lubricated noise
 * branch (flag_lub_B, brownian.c:1734-1849) is never taken.  0 or a
 * STOKES_HIP_E* code */
int stokes_hip_bd_set_lub (stokes_hip_bd *bd, int enable);

/* FT / FTS: constant torques [3 nm] (BD->T, brownian.c:2296-2299) or NULL,
 * read on every step */
int stokes_hip_bd_set_torque (stokes_hip_bd *bd, const double *T_dev);
/* FTS: GMRES parameters of the noise products and of the Lanczos bounds
 * (brownian.c:597-600 fixes restart 20, max 2000, eps 1e-6, the default
 * here); the step's own solves take those of stokes_set_iter, as
 * solve_mix_3all does */
int stokes_hip_bd_set_fts_iter (stokes_hip_bd *bd, int restart, int max,
                                double eps);
/* copies of the Gaussian draw and of z = M^{-1/2} y of the last attempt
 * (device pointers, 3 np (F) or 6 np (FT, FTS, particle-wise (F, T)) each;
 * either may be NULL) */
int stokes_hip_bd_last_noise (stokes_hip_bd *bd, double *y_dev, double *z_dev);
void stokes_hip_bd_destroy (stokes_hip_bd *bd);

#ifdef __cplusplus
}
#endif
#endif
