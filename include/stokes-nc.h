/* stokes-nc.h -- libstokes trajectory files (drop-in for the writing half
 * of the reference's stokes-nc.h).  The files are netCDF classic with
 * 64-bit offsets, written by the library itself (no netCDF library), with
 * the dimensions, variables and long_name attributes the reference's
 * stokes_nc_init defines.  struct stokes_nc is opaque here (the reference
 * exposes its netCDF ids).  Errors print a message and exit, as the
 * reference's stokes_nc_error does.  Not provided: the reading half
 * (stokes-nc-read.c), stokes_nc_print_actives, the BD parameter files
 * (stokes_nc_set_by_params and its readers) and stokes_nc_set_rng (which
 * takes the reference's struct KIrand; stokes_nc_set_rng_state writes the
 * same variables). */
#ifndef LIBSTOKES_HIP_STOKES_NC_H
#define LIBSTOKES_HIP_STOKES_NC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct stokes_nc;

/* replaces stokes-nc.c:1367-1443: a file with x[time][p][vec] only
 * (np mobile particles); an existing file is an error (NC_NOCLOBBER) */
struct stokes_nc *stokes_nc_x_init (const char *filename, int np);

/* replaces stokes-nc.c:1445-1598: np mobile and nf fixed particles;
 * version 0 F, 1 FT, 2 FTS; flag_poly: radii a (and af); flag_Q:
 * quaternions q; flag_it: imposed flow per step (Ui, Oi, Ei[time]) rather
 * than constant (Ui0, Oi0, Ei0); shear_mode 1 or 2: shear_rate and
 * shear_shift[time]; flag_rng: the random-generator state per step */
struct stokes_nc *stokes_nc_init (const char *filename, int np, int nf,
                                  int version, int flag_poly, int flag_Q,
                                  int flag_it, int shear_mode, int flag_rng);

/* replaces stokes-nc.c:1605-1619: close the file */
void stokes_nc_free (struct stokes_nc *nc);

/* replace stokes-nc.c:1625-2116: the time-independent entries (vectors of
 * 3, strains of 5 per particle; a, af one per particle) */
void stokes_nc_set_shear_rate (struct stokes_nc *nc, double shear_rate);
void stokes_nc_set_l (struct stokes_nc *nc, const double *l);
void stokes_nc_set_ui0 (struct stokes_nc *nc, const double *ui0);
void stokes_nc_set_oi0 (struct stokes_nc *nc, const double *oi0);
void stokes_nc_set_ei0 (struct stokes_nc *nc, const double *ei0);
void stokes_nc_set_x0 (struct stokes_nc *nc, const double *x0);
void stokes_nc_set_u0 (struct stokes_nc *nc, const double *u0);
void stokes_nc_set_o0 (struct stokes_nc *nc, const double *o0);
void stokes_nc_set_e0 (struct stokes_nc *nc, const double *e0);
void stokes_nc_set_f0 (struct stokes_nc *nc, const double *f0);
void stokes_nc_set_t0 (struct stokes_nc *nc, const double *t0);
void stokes_nc_set_s0 (struct stokes_nc *nc, const double *s0);
void stokes_nc_set_xf0 (struct stokes_nc *nc, const double *xf0);
void stokes_nc_set_uf0 (struct stokes_nc *nc, const double *uf0);
void stokes_nc_set_of0 (struct stokes_nc *nc, const double *of0);
void stokes_nc_set_ef0 (struct stokes_nc *nc, const double *ef0);
void stokes_nc_set_ff0 (struct stokes_nc *nc, const double *ff0);
void stokes_nc_set_tf0 (struct stokes_nc *nc, const double *tf0);
void stokes_nc_set_sf0 (struct stokes_nc *nc, const double *sf0);
void stokes_nc_set_a (struct stokes_nc *nc, const double *a);
void stokes_nc_set_af (struct stokes_nc *nc, const double *af);

/* replace stokes-nc.c:2118-2645: the entries of record step (steps past
 * the last written one are filled with the fill value first) */
void stokes_nc_set_time (struct stokes_nc *nc, int step, double time);
void stokes_nc_set_shear_shift (struct stokes_nc *nc, int step,
                                double shear_shift);
void stokes_nc_set_ui (struct stokes_nc *nc, int step, const double *ui);
void stokes_nc_set_oi (struct stokes_nc *nc, int step, const double *oi);
void stokes_nc_set_ei (struct stokes_nc *nc, int step, const double *ei);
void stokes_nc_set_x (struct stokes_nc *nc, int step, const double *x);
void stokes_nc_set_q (struct stokes_nc *nc, int step, const double *q);
void stokes_nc_set_u (struct stokes_nc *nc, int step, const double *u);
void stokes_nc_set_o (struct stokes_nc *nc, int step, const double *o);
void stokes_nc_set_e (struct stokes_nc *nc, int step, const double *e);
void stokes_nc_set_f (struct stokes_nc *nc, int step, const double *f);
void stokes_nc_set_t (struct stokes_nc *nc, int step, const double *t);
void stokes_nc_set_s (struct stokes_nc *nc, int step, const double *s);
void stokes_nc_set_xf (struct stokes_nc *nc, int step, const double *xf);
void stokes_nc_set_uf (struct stokes_nc *nc, int step, const double *uf);
void stokes_nc_set_of (struct stokes_nc *nc, int step, const double *of);
void stokes_nc_set_ef (struct stokes_nc *nc, int step, const double *ef);
void stokes_nc_set_ff (struct stokes_nc *nc, int step, const double *ff);
void stokes_nc_set_tf (struct stokes_nc *nc, int step, const double *tf);
void stokes_nc_set_sf (struct stokes_nc *nc, int step, const double *sf);

/* the random-generator entries of record step (mt[time][rng], mti,
 * mt_Ghs, mt_Gs of stokes-nc.c:1236-1253) from a raw MT19937 state: the
 * 624 words stored whole as their 32-bit patterns (the reference's
 * stokes_nc_set_rng, stokes-nc.c:2646-2729, keeps half of them on LP64),
 * the index, and the polar method's saved-value flag and value.  No
 * counterpart in the reference under this name */
void stokes_nc_set_rng_state (struct stokes_nc *nc, int step,
                              const uint32_t *mt, int mti,
                              int gauss_has_saved, double gauss_saved);

#ifdef __cplusplus
}
#endif
#endif
