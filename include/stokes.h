/* stokes.h -- system state of the MI355X libstokes operator (drop-in).
 *
 * struct stokes keeps the field order and types of the reference's
 * struct stokes (stokes.h:27-154, identical copy at libstokes.h:29-156) so
 * that code compiled against the reference header can hand its pointers
 * to this library unchanged (LP64).  Fields the hot path does not use
 * (lubrication tables, slip tables, iterative-solver handle) are kept as
 * opaque pointers so the byte layout stays the same.
 */
#ifndef LIBSTOKES_HIP_STOKES_H
#define LIBSTOKES_HIP_STOKES_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

struct twobody_f_list;        /* lubrication tables (not used here) */
struct twobody_slip_f_list;   /* slip tables (not used here) */
struct iter;                  /* iterative-solver parameters (not used here) */

struct stokes {
  int version;                /* 0 = F, 1 = FT, 2 = FTS */

  int np;                     /* number of all particles */
  int nm;                     /* number of mobile particles */
  double *pos;                /* [np*3] particle-major xyz */

  /* polydisperse parameters */
  double *a;                  /* [np] radii, NULL = monodisperse a = 1 */
  double rmin;                /* r is replaced by rmin*(ai+aj) when smaller */
  int twobody_nmax;
  int twobody_lub;
  int *poly_table;
  struct twobody_f_list *twobody_f_list;

  /* slip parameters (NULL = no slip) */
  double *slip;
  double *slip_a;
  double *slip_G32;
  double *slip_G30;
  double *slip_G52;
  int *slip_table;
  struct twobody_slip_f_list *twobody_slip_f_list;

  /* imposed flow */
  double Ui[3];
  double Oi[3];
  double Ei[5];

  /* simple shear: 0 none, 1 x-flow/y-gradient, 2 x-flow/z-gradient */
  int shear_mode;
  double shear_rate;
  double shear_shift;

  /* periodic parameters */
  int periodic;               /* 0 = open boundary, 1 = Ewald */

  double ewald_eps;
  double rmax2;
  int rmaxx, rmaxy, rmaxz;
  double kmax;
  int kmaxx, kmaxy, kmaxz;

  double xi, xi2, xiaspi, xia2;
  double pivol;

  double lx, ly, lz;
  int ilx[27], ily[27], ilz[27];

  double self_a;
  double self_c;
  double self_m;

  /* lattice-sum tables */
  int flag_table;
  int nr;                     /* real-space images */
  double *rlx;
  double *rly;
  double *rlz;
  int *rmx;
  int *rmy;
  int *rmz;
  int nk;                     /* reciprocal lattice points */
  double *ex;
  double *ey;
  double *ez;
  double *k;
  double *k1;
  double *k2;
  double *k3;
  double *ya;
  double *yb;
  double *yc;
  double *yg;
  double *yh;
  double *ym;

  /* lubrication */
  double lubmin2;
  double lubmax;

  /* timing of the last operator call: cpu2 real space, cpu3 reciprocal
   * space, cpu1 = cpu2 + cpu3 -- here in GPU milliseconds */
  double cpu1, cpu2, cpu3;

  struct iter *it;
};

/* replaces stokes.c:316-419 */
struct stokes *stokes_init (void);
/* replaces stokes.c:422-448; also releases the device context */
void stokes_free (struct stokes *sys);
/* replaces stokes.c:450-466 */
void stokes_set_np (struct stokes *sys, int np, int nm);
/* replaces stokes.c:469-492 */
void stokes_set_Ui (struct stokes *sys, double uix, double uiy, double uiz);
void stokes_set_Oi (struct stokes *sys, double oix, double oiy, double oiz);
void stokes_set_Ei (struct stokes *sys, double eixx, double eixy,
                    double eixz, double eiyz, double eiyy);
/* replaces stokes.c:506-547 */
void stokes_set_shear (struct stokes *sys, int shear_mode, double shear_rate);
/* replaces stokes.c:554-590 */
double stokes_get_shear_shift (struct stokes *sys, double t, double t0,
                               double s0);
/* replaces stokes.c:599-605 */
void stokes_set_shear_shift (struct stokes *sys, double t, double t0,
                             double s0);
/* replaces stokes.c:606-651 */
void stokes_set_l (struct stokes *sys, double lx, double ly, double lz);
/* replaces stokes.c:702-762 (and the table of stokes.c:173-310) */
void stokes_set_xi (struct stokes *sys, double xi, double ewald_eps);
/* replaces stokes.c:765-776 */
double xi_by_tratio (struct stokes *sys, double tratio);
/* replaces stokes.c:790-813: keeps solver name, max, restart, eps, debug
 * and out in sys->it (freed by stokes_free); the library's GMRES
 * (solve_mob_3fts_0) uses max, restart and eps */
void stokes_set_iter (struct stokes *sys, const char *solver, int max,
                      int restart, double eps, int debug, FILE *out);
/* replaces stokes.c:822-868 */
void stokes_set_pos (struct stokes *sys, const double *pos);
void stokes_set_pos_mobile (struct stokes *sys, const double *pos);
void stokes_set_pos_fixed (struct stokes *sys, const double *pos);
/* replaces stokes.c:902-959 for the mobility path: stores a[np] only (the
 * lubrication tables of the reference are not built) */
void stokes_set_radius (struct stokes *sys, const double *a);
/* replaces stokes.c:972-990 */
void stokes_unset_radius (struct stokes *sys);

#ifdef __cplusplus
}
#endif
#endif
