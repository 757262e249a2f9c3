/* stokes-nc.h -- libstokes trajectory files (drop-in for the writing half
 * of the reference's stokes-nc.h).  The files are netCDF classic with
 * 64-bit offsets, written by the library itself (no netCDF library), with
 * the dimensions, variables and long_name attributes the reference's
 * stokes_nc_init defines, and read back by the functions of the
 * reference's stokes-nc-read.h (any classic or 64-bit-offset file).
 * Errors print a message and exit, as the reference's stokes_nc_error
 * does.  Not provided: stokes_nc_set_rng / stokes_nc_get_rng, which take
 * the reference's struct KIrand (stokes_nc_set_rng_state /
 * stokes_nc_get_rng_state use the same variables with a raw MT19937
 * state). */
#ifndef LIBSTOKES_HIP_STOKES_NC_H
#define LIBSTOKES_HIP_STOKES_NC_H

#include <stdint.h>
#include <stdio.h>
#include "stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

/* the reference's struct stokes_nc (stokes-nc.h:27-158): the same int
 * fields in the same order, so code that reads nc->np, nc->ntime or the
 * flags compiles and reads them unchanged; the *_id fields hold this
 * library's dimension and variable indices (no netCDF ids), id the file
 * descriptor.  priv: the library's own state */
struct stokes_nc {
  int id, np, p_dim, p_id, npf, pf_dim, pf_id, nvec, vec_dim, vec_id, nstt,
      stt_dim, stt_id, nquat, quat_dim, quat_id, nrng, rng_dim, rng_id,
      flag_rng, ntime, time_dim, time_id, l_id, ui0_id, oi0_id, ei0_id,
      ui_id, oi_id, ei_id, x0_id, u0_id, o0_id, e0_id, f0_id, t0_id, s0_id,
      xf0_id, uf0_id, of0_id, ef0_id, ff0_id, tf0_id, sf0_id, x_id, u_id,
      o_id, e_id, f_id, t_id, s_id, q_id, xf_id, uf_id, of_id, ef_id, ff_id,
      tf_id, sf_id, a_id, af_id, shear_mode_id, shear_rate_id,
      shear_shift_id, mt_id, mti_id, mt_Ghs_id, mt_Gs_id, flag_ui0,
      flag_oi0, flag_ei0, flag_ui, flag_oi, flag_ei, flag_x0, flag_u0,
      flag_o0, flag_e0, flag_f0, flag_t0, flag_s0, flag_xf0, flag_uf0,
      flag_of0, flag_ef0, flag_ff0, flag_tf0, flag_sf0, flag_x, flag_u,
      flag_o, flag_e, flag_f, flag_t, flag_s, flag_q, flag_xf, flag_uf,
      flag_of, flag_ef, flag_ff, flag_tf, flag_sf, flag_a, flag_af;
  void *priv;
};

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

/* ---- reading (the reference's stokes-nc-read.h) ---- */

/* replaces stokes-nc-read.c:741-745: open a file for reading; np, npf,
 * ntime and the flag_* fields describe it */
struct stokes_nc *stokes_nc_open (const char *filename);

/* replaces stokes-nc-read.c:750-754: open for appending records (the
 * setters then write past ntime) */
struct stokes_nc *stokes_nc_reopen (const char *filename);

/* replaces stokes-nc-read.c:763-898: Ui0, Oi0, Ei0, a, af or l (Ui, Oi, Ei:
 * their first record); another name prints a message and reads nothing */
void stokes_nc_get_array1d (const struct stokes_nc *nc, const char *name,
                            double *x);

/* replaces stokes-nc-read.c:902-1118: a time-independent [np or npf][3 or
 * 5] entry (x0, U0, ..., Sf0) */
void stokes_nc_get_data0 (const struct stokes_nc *nc, const char *name,
                          double *x);

/* replaces stokes-nc-read.c:1122-1356: entry name (x, q, U, ..., Sf) of
 * record step */
void stokes_nc_get_data (const struct stokes_nc *nc, const char *name,
                         int step, double *x);

/* replaces stokes-nc-read.c:1364-1385: time[ntime] */
void stokes_nc_get_time (const struct stokes_nc *nc, double *time);

/* replaces stokes-nc-read.c:1393-1413: time[step] */
double stokes_nc_get_time_step (const struct stokes_nc *nc, int step);

/* the random-generator state of record step as stokes_nc_set_rng_state
 * writes it (mt[624] as 32-bit words); no counterpart in the reference
 * under this name (stokes_nc_get_rng, stokes-nc-read.c:1419-1482, fills a
 * struct KIrand) */
void stokes_nc_get_rng_state (const struct stokes_nc *nc, int step,
                              uint32_t *mt, int *mti, int *gauss_has_saved,
                              double *gauss_saved);

/* ---- BD parameter files ---- */

/* replaces stokes-nc.c:52-115: the table of the defined entries and the
 * shear mode */
void stokes_nc_print_actives (struct stokes_nc *nc, FILE *out);

/* replaces stokes-nc.c:2731-2841: a new file (stokes_nc_init with the
 * constant imposed flow Ui0, Oi0, Ei0) holding the parameters of a BD run:
 * sys->version, np, nm and radii a (NULL: equal spheres), the lattice lat
 * when sys->periodic, the forces F (and torques T, strains E) of the nm
 * mobile particles, for np > nm the fixed particles' velocities uf (and
 * of, ef) and positions xf, and the shear rate for shear_mode != 0 */
struct stokes_nc *stokes_nc_set_by_params (const char *out_file, const struct stokes *sys,
                                           int flag_Q, const double *Ui, const double *Oi,
                                           const double *Ei, const double *F,
                                           const double *T, const double *E,
                                           const double *uf, const double *of,
                                           const double *ef, const double *xf,
                                           const double *lat, int shear_mode,
                                           double shear_rate, int flag_rng);

/* replaces stokes-nc.c:2869-3188: 0 when the file holds these parameters
 * (entries compared to within tiny, the expected entries defined), else
 * the number of mismatches (each printed to stderr) */
int stokes_nc_check_params (const struct stokes_nc *nc, const struct stokes *sys,
                            int flag_Q, const double *Ui, const double *Oi,
                            const double *Ei, const double *F, const double *T,
                            const double *E, const double *uf, const double *of,
                            const double *ef, const double *xf, const double *lat,
                            double tiny);

/* replaces stokes-nc.c:3191-3342: 0 F, 1 FT, 2 FTS from the entries the
 * file defines (0 when they match none) */
int stokes_nc_get_version (const struct stokes_nc *nc);

/* replaces stokes-nc.c:3369-3540: the parameters of stokes_nc_set_by_params
 * back; sys gets version, np / nm (stokes_set_np), radii, periodic and
 * the lattice (stokes_set_l), shear_mode and shear_rate, not the Ewald or
 * iteration parameters */
void stokes_nc_get_params (const struct stokes_nc *nc, struct stokes *sys, int *flag_Q,
                           double *Ui, double *Oi, double *Ei, double *F, double *T,
                           double *E, double *uf, double *of, double *ef, double *xf,
                           double *lat, int *shear_mode, double *shear_rate,
                           int *flag_rng);

#ifdef __cplusplus
}
#endif
#endif
