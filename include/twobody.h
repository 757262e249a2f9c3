/* twobody.h -- two-body resistance coefficients of the lubrication
 * correction (drop-in for the reference's twobody.h).
 *
 * f_k(lambda), k = 0..n, of the eleven resistance families for the radius
 * ratio lambda = a2 / a1.  The reference tabulates them as generated
 * polynomials (XA.c ... ZM.c); this library computes them once per process
 * from the Jeffrey & Onishi (1984) / Jeffrey (1992) recurrences.
 */
#ifndef LIBSTOKES_HIP_TWOBODY_H
#define LIBSTOKES_HIP_TWOBODY_H

#ifdef __cplusplus
extern "C" {
#endif

/* twobody.h:25-41 of the reference: the same layout */
struct twobody_f {
  int nmax;
  double lambda;
  double *XA, *YA, *YB, *XC, *YC, *XG, *YG, *YH, *XM, *YM, *ZM;
};

/* twobody.h:43-47 */
struct twobody_f_list {
  int n;
  double *l;
  struct twobody_f **f;
};

/* twobody.h:50-72 (XA.c ... ZM.c): f[0..n] at lambda = l */
void twobody_XA (int n, double l, double *f);
void twobody_YA (int n, double l, double *f);
void twobody_YB (int n, double l, double *f);
void twobody_XC (int n, double l, double *f);
void twobody_YC (int n, double l, double *f);
void twobody_XG (int n, double l, double *f);
void twobody_YG (int n, double l, double *f);
void twobody_YH (int n, double l, double *f);
void twobody_XM (int n, double l, double *f);
void twobody_YM (int n, double l, double *f);
void twobody_ZM (int n, double l, double *f);

/* twobody.c:30-79, 99-145 */
struct twobody_f *twobody_f_init (int nmax, double lambda);
void twobody_f_free (struct twobody_f *f);
struct twobody_f_list *twobody_f_list_init (void);
void twobody_f_list_append (struct twobody_f_list *list, int nmax, double lambda);
void twobody_f_list_free (struct twobody_f_list *list);

/* twobody.c:644-686 and 689-998 (far form: the 1/s series to order n) and
 * 1932-1975, 1977-2644 (lubrication form); f12 NULL computes the
 * coefficients; out[22] = (11, 12) of XA YA YB XC YC XG YG YH XM YM ZM in
 * the Jeffrey form.  version 0 (F), 1 (FT), 2 (FTS) */
void twobody_far (int version, int n, double l, double s, double *far);
void twobody_far_with_f (int version, struct twobody_f *f12, int n, double l,
                         double s, double *far);
void twobody_lub (int version, int n, double l, double s, double *lub);
void twobody_lub_with_f (int version, struct twobody_f *f12, int n, double l,
                         double s, double *lub);
/* twobody.c:2646-2700 (Stokesian-dynamics form), 2709-2772 (dimensional) */
void twobody_scale_SD (int version, double *two, double l);
void twobody_scale (int version, double *two, double a1, double l);
/* twobody.c:2786-2814: r = distance, aa / ab radii of spheres 1 / 2;
 * flag_lub 0 far, 1 lubrication form; flag_scale 0 Jeffrey, 1
 * dimensional, 2 SD */
void twobody_scalars_res (int version, double r, double aa, double ab,
                          struct twobody_f *f12, int n, int flag_lub,
                          int flag_scale, double *res);

#ifdef __cplusplus
}
#endif

#endif
