/* oracle.h -- FP64 CPU restatement of libstokes' Ewald mobility hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library
 * in libstokes_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links, loads or calls it.
 *
 * Each function restates one routine of the reference (file:line cited at
 * the definition in oracle.c).  Arithmetic is plain FP64 in the reference's
 * loop order so that it can be checked against the compiled reference
 * (oracle/_ref) to rounding level.
 */
#ifndef ORACLE_H
#define ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* System description used by the oracle (its own type; the field meanings
 * follow struct stokes of stokes.h:27-154). */
typedef struct orc_sys {
  int version;          /* 0 = F, 1 = FT, 2 = FTS */
  int np;
  const double *pos;    /* [3*np], particle-major xyz */
  const double *a;      /* [np] radii or NULL (monodisperse a = 1) */
  double rmin;          /* r is clamped to rmin*(ai+aj) */
  int periodic;

  double lx, ly, lz, pivol;
  double xi, xi2, xiaspi, xia2, ewald_eps;
  double self_a, self_c, self_m;

  double rmax2; int rmaxx, rmaxy, rmaxz;
  double kmax;  int kmaxx, kmaxy, kmaxz;

  int nr;               /* real-space image table */
  double *rlx, *rly, *rlz;
  int *rmx, *rmy, *rmz;

  int shear_mode;       /* 0 none, 1 x-flow/y-grad, 2 x-flow/z-grad */
  double shear_shift;

  int fix_recip_g;      /* 1: reciprocal G / G-tilde with the sign that makes
                           the FTS sum xi-independent (0: as the reference) */
  double Ui[3], Oi[3], Ei[5];  /* imposed flow (stokes.h:116-118): read by
                                  the BD step's drift only */
  double shear_rate;    /* sheared BD steps: the shift at time tt is */
  double bd_t, bd_t0, bd_s0;   /* s0 + L rate (tt - t0), the step at bd_t */
} orc_sys;

/* set-up (stokes.c) */
void   orc_sys_init (orc_sys *s);
void   orc_sys_free_tables (orc_sys *s);
void   orc_set_l (orc_sys *s, double lx, double ly, double lz);
double orc_xi_by_tratio (const orc_sys *s, double tratio);
void   orc_set_xi (orc_sys *s, double xi, double ewald_eps);
double orc_ewald_real_bound (double r, double xi);
double orc_ewald_recip_bound (double k, double xi);

/* pair scalars (ewald.c, non-ewald.c); out[11] = xa ya yb xc yc xg yg yh xm ym zm */
void orc_scalars_ewald_real (int version, double xi, double r, double *out);
void orc_scalars_ewald_real_poly (int version, double xi, double r,
                                  double aa, double ab, double *out);
void orc_scalars_scale_SD (int version, double a1, double *s);
void orc_scalars_nonewald (int version, double r, double *out);
void orc_scalars_nonewald_poly (int version, double r, double aa, double ab,
                                double *out);

/* block kernels (f.c, ft.c, fts.c) */
void orc_block_atimes (int version, const double *x, double *y,
                       double ex, double ey, double ez, const double *s);
void orc_block_atimes_sign (int version, const double *x, double *y,
                            double ex, double ey, double ez, const double *s,
                            int self);
int  orc_dim_of (int version);
void orc_block_ij (int version, int i, int j,
                   double ex, double ey, double ez, const double *s,
                   int n, double *mat);

/* operators (ewald.c, non-ewald.c, matrix.c); the row range [i0,i1)
 * restricts the output particles (i1 < 0 means all) -- used for bounded
 * CPU timing samples. */
void orc_atimes_ewald_rows (const orc_sys *s, const double *x, double *y,
                            int i0, int i1);
void orc_atimes_ewald (const orc_sys *s, const double *x, double *y);
void orc_matrix_ewald (const orc_sys *s, double *mat);
void orc_atimes_nonewald (const orc_sys *s, const double *x, double *y);
void orc_matrix_nonewald (const orc_sys *s, double *mat);
void orc_atimes_3all (const orc_sys *s, const double *x, double *y);
void orc_extmat_fts (int np, const double *m, const double *x, double *y);
void orc_dot_prod_matrix (const double *m, int n1, int n2,
                          const double *x, double *y);

/* Chebyshev (chebyshev.c); func_id 0 = 1/sqrt(x), 1 = sqrt(x) */
double orc_func (int func_id, double x);
void   orc_cheb_coef (int ncheb, int func_id, double x0, double x1, double *a);
double orc_cheb_eval (int ncheb, const double *a, double x0, double x1, double x);
typedef void (*orc_atimes_fn) (int n, const double *x, double *y, void *ud);
void   orc_cheb_eval_atimes (int ncheb, const double *a, int n, const double *y,
                             double *z, double e0, double e1,
                             orc_atimes_fn atimes, void *ud);
double orc_cheb_error_minvsqrt (int n, const double *y, const double *z,
                                orc_atimes_fn atimes, void *ud);
double orc_cheb_error_Rsqrt (int n, const double *y, const double *z,
                             orc_atimes_fn atimes, void *ud);
/* operator adaptors: user data is an orc_sys* or an orc_densemat* */
void orc_atimes_sys_cb (int n, const double *x, double *y, void *ud);
typedef struct orc_densemat { int n; const double *m; } orc_densemat;
void orc_atimes_dense_cb (int n, const double *x, double *y, void *ud);

/* MT19937 + polar Gaussian (KIrand.c) */
typedef struct orc_rng {
  unsigned long mt[624];
  int mti;
  int has_saved;
  double saved;
} orc_rng;
void          orc_rng_seed (orc_rng *r, unsigned long s);
unsigned long orc_rng_int32 (orc_rng *r);
double        orc_rng_real2 (orc_rng *r);
double        orc_rng_gaussian (orc_rng *r);

/* Brownian force, F version, atimes route (brownian.c:1545-1874):
 * z = M^{-1/2} y by Chebyshev with given bounds, y ~ N(0,1) from rng. */
void orc_bd_noise_F (const orc_sys *s, int ncheb, const double *acoef,
                     double e0, double e1, const double *y, double *z);

/* overlapping (pair, image) count (brownian.c:299-375) */
int orc_check_overlap (const orc_sys *s, const double *pos, double rmin);

/* quaternion_dQdt (ode-quaternion.c:89-105) */
void orc_quat_dqdt (const double *Q, const double *O, double *dQdt);
/* FTS mobility problem with E = 0, (F, T) -> (U, O) particle-wise */
void orc_fts_mob_ft (const orc_sys *s, const double *ft, double *uo);
/* the BD mobility of the system's version (BD_atimes_mob_FU) */
void orc_bd_op_cb (int n, const double *x, double *y, void *ud);
/* overlap count over the first n particles */
int orc_check_overlap_n (const orc_sys *s, const double *pos, int n,
                         double rmin);
/* FTS mobility problem with E = 0, (F, T) -> (U, O) of the nm mobile
 * particles, particles nm..np-1 at rest */
void orc_fts_mix_ft (const orc_sys *s, int nm, const double *ft, double *uo);
/* orc_bd_step with particles nm..np-1 fixed */
double orc_bd_step_nm (orc_sys *s, int nm, int scheme, double bb_n, double *x,
                       double *q, const double *F, const double *T,
                       double peclet, double dt, double dt_lim, double rmin_ol,
                       int ncheb, const double *acoef, double e0, double e1,
                       int ndraw, const double *draws, int fix_fact, int *used,
                       int *halvings, double *err);
/* one BD step of scheme 0 (mid-point), 1 (BB03, step parameter bb_n) or 2
 * (BM97), F / FT / FTS (brownian.c:2352-3200); see oracle.c */
double orc_bd_step (orc_sys *s, int scheme, double bb_n, double *x, double *q,
                    const double *F, const double *T, double peclet,
                    double dt, double dt_lim, double rmin_ol, int ncheb,
                    const double *acoef, double e0, double e1, int ndraw,
                    const double *draws, int fix_fact, int *used,
                    int *halvings, double *err);
/* one mid-point BD step, F / FT / FTS (brownian.c:2352-2560); see oracle.c */
double orc_bd_mid (orc_sys *s, double *x, double *q, const double *F,
                   const double *T, double peclet, double dt, double dt_lim,
                   double rmin_ol, int ncheb, const double *acoef, double e0,
                   double e1, int ndraw, const double *draws, int fix_fact,
                   int *used, int *halvings, double *err);
/* one mid-point BD step, F version (brownian.c:2352-2560); see oracle.c */
double orc_bd_mid_F (orc_sys *s, double *x, const double *F, double peclet,
                     double dt, double dt_lim, double rmin_ol, int ncheb,
                     const double *acoef, double e0, double e1, int ndraw,
                     const double *draws, int fix_fact, int *used,
                     int *halvings, double *err);

/* lubrication (orc_lub.c): two-body coefficients, resistance scalars,
 * far-field inverse, calc_lub_3fts */
int  orc_twobody_init (int nmax);
void orc_twobody_table (int fam, int n, double l, double *f);
void orc_twobody_far (int version, const double *f, int n, double l, double s,
                      double *out);
void orc_twobody_lub (int version, const double *f, int N1, int mmax,
                      double l, double s, double yc12_g5, double *out);
void orc_twobody_scale (int version, double *two, double a1, double l, int sd);
void orc_twobody_res (int version, double r, double aa, double ab, int n,
                      int flag_lub, int flag_scale, double *res);
void orc_two_body_res (double s, double *res);
void orc_minv_poly (int version, double r, double aa, double ab, double *out);
void orc_minv_fts (double s, double *res);
void orc_block_atimes_self (int version, const double *x, double *y,
                            double ex, double ey, double ez, const double *s);
void orc_lub_3fts (const orc_sys *sys, double lubmin2, double lubmax, int nmax,
                   int flag_lub, const double *uoe, double *fts);

#ifdef __cplusplus
}
#endif
#endif
