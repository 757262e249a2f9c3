// lub_export.hip -- host-callable lubrication scalars of the reference API
// (twobody.h, two-body-res.h, minv-poly.h, f.h / ft.h / fts.h).  They
// evaluate the templates of lub_math.h that the device kernels use.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include "lub_math.h"

extern "C" {
#include "twobody.h"
#include "lub.h"
}

using namespace shk::lub;

namespace {

// f values of a struct twobody_f
struct HostF {
  const double *p[NFAM];
  explicit HostF (const struct twobody_f *f)
  {
    p[XA] = f->XA; p[YA] = f->YA; p[YB] = f->YB; p[XC] = f->XC; p[YC] = f->YC;
    p[XG] = f->XG; p[YG] = f->YG; p[YH] = f->YH; p[XM] = f->XM; p[YM] = f->YM;
    p[ZM] = f->ZM;
  }
  double operator() (int fam, int k) const { return p[fam][k]; }
  void four (int fam, int k, double *f) const
  {
    for (int i = 0; i < 4; i++) f[i] = p[fam][k + i];
  }
};

// the reference's n: made even, the series to m < n
int even (int n) { return n % 2 ? n - 1 : n; }

// the table f12 or one made for (n, l)
struct FHold {
  struct twobody_f *own = nullptr;
  const struct twobody_f *f;
  FHold (struct twobody_f *f12, int n, double l)
  {
    if (f12) f = f12;
    else { own = twobody_f_init (n, l); f = own; }
  }
  ~FHold () { if (own) twobody_f_free (own); }
};

const int sd_pw[NFAM] = {1, 1, 2, 3, 3, 2, 2, 3, 3, 3, 3};

}  // namespace

extern "C" void
twobody_far_with_f (int version, struct twobody_f *f12, int n, double l, double s, double *far)
{
  n = even (n);
  FHold h (f12, n, l);
  for (int k = 0; k < 2 * nfam_of (version); k++) far[k] = 0.0;
  far_form (version, HostF (h.f), n, l, s, far);
}

extern "C" void
twobody_far (int version, int n, double l, double s, double *far)
{
  twobody_far_with_f (version, nullptr, n, l, s, far);
}

extern "C" void
twobody_lub_with_f (int version, struct twobody_f *f12, int n, double l, double s, double *lub)
{
  n = even (n);
  FHold h (f12, n, l);
  lub_form (version, HostF (h.f), n - 1, l, s, 1.0, lub);
}

extern "C" void
twobody_lub (int version, int n, double l, double s, double *lub)
{
  twobody_lub_with_f (version, nullptr, n, l, s, lub);
}

extern "C" void
twobody_scale_SD (int version, double *two, double l)
{
  scale (version, two, 1.0, l, true);
}

extern "C" void
twobody_scale (int version, double *two, double a1, double l)
{
  scale (version, two, a1, l, false);
}

extern "C" void
twobody_scalars_res (int version, double r, double aa, double ab, struct twobody_f *f12, int n,
                     int flag_lub, int flag_scale, double *res)
{
  const double l = ab / aa, s = 2.0 * r / (aa + ab);
  if (flag_lub == 0) twobody_far_with_f (version, f12, n, l, s, res);
  else twobody_lub_with_f (version, f12, n, l, s, res);
  if (flag_scale == 1) scale (version, res, aa, l, false);
  else if (flag_scale == 2) scale (version, res, aa, l, true);
}

// two-body-res.c:31: the lubrication form at lambda = 1 to m = 50 with the
// equal-sphere YC12 coefficient, dimensional at radius 1
extern "C" void
scalar_two_body_res (double s, double *res)
{
  struct twobody_f *f = twobody_f_init (50, 1.0);
  lub_form (2, HostF (f), 50, 1.0, s, 2.0, res);
  twobody_f_free (f);
  scale (2, res, 1.0, 1.0, false);
}

namespace {
void minv_pairs (int version, double s, double *out, int nf)
{
  double m[44];
  minv (version, s, 1.0, 1.0, m);
  for (int k = 0; k < nf; k++)
    {
      out[2 * k] = m[4 * k];
      out[2 * k + 1] = m[4 * k + 1];
    }
}
void minv_quads (int version, double r, double aa, double ab, double *out, int nf)
{
  double m[44];
  minv (version, r, aa, ab, m);
  for (int k = 0; k < 4 * nf; k++) out[k] = m[k];
}
}  // namespace

extern "C" void scalar_minv_f (double s, double *o) { minv_pairs (0, s, o, 2); }
extern "C" void scalar_minv_ft (double s, double *o) { minv_pairs (1, s, o, 5); }
extern "C" void scalar_minv_fts (double s, double *o) { minv_pairs (2, s, o, 11); }
extern "C" void scalars_minv_f_poly (double r, double aa, double ab, double *o)
{
  minv_quads (0, r, aa, ab, o, 2);
}
extern "C" void scalars_minv_ft_poly (double r, double aa, double ab, double *o)
{
  minv_quads (1, r, aa, ab, o, 5);
}
extern "C" void scalars_minv_fts_poly (double r, double aa, double ab, double *o)
{
  minv_quads (2, r, aa, ab, o, 11);
}

extern "C" void
scalars_res_poly_scale_SD (int version, double a1, double a2, double *scalar)
{
  const int nf = nfam_of (version);
  for (int f = 0; f < nf; f++)
    {
      const double d1 = sd_pw[f] == 1 ? a1 : (sd_pw[f] == 2 ? a1 * a1 : a1 * a1 * a1);
      const double d2 = sd_pw[f] == 1 ? a2 : (sd_pw[f] == 2 ? a2 * a2 : a2 * a2 * a2);
      scalar[4 * f] /= d1;
      scalar[4 * f + 1] /= d1;
      scalar[4 * f + 2] /= d2;
      scalar[4 * f + 3] /= d2;
    }
}

extern "C" void
scalars_lub_poly (int version, double r, double a1, double a2, struct twobody_f *f12, int n,
                  int flag_lub, double *lub)
{
  double res[22] = {0}, m[44];
  for (int k = 0; k < 22; k++) lub[k] = 0.0;
  twobody_scalars_res (version, r, a1, a2, f12, n, flag_lub, 1, res);
  minv (version, r, a1, a2, m);
  for (int f = 0; f < nfam_of (version); f++)
    {
      lub[2 * f] = res[2 * f] - m[4 * f];
      lub[2 * f + 1] = res[2 * f + 1] - m[4 * f + 1];
    }
}

extern "C" void
scalars_lub_poly_full (int version, double r, double a1, double a2, struct twobody_f *f12,
                       struct twobody_f *f21, int n, int flag_lub, double *lub)
{
  double r12[22] = {0}, r21[22] = {0}, m[44];
  for (int k = 0; k < 44; k++) lub[k] = 0.0;
  twobody_scalars_res (version, r, a1, a2, f12, n, flag_lub, 1, r12);
  twobody_scalars_res (version, r, a2, a1, f21, n, flag_lub, 1, r21);
  minv (version, r, a1, a2, m);
  for (int f = 0; f < nfam_of (version); f++)
    {
      lub[4 * f] = r12[2 * f] - m[4 * f];
      lub[4 * f + 1] = r12[2 * f + 1] - m[4 * f + 1];
      lub[4 * f + 2] = r21[2 * f + 1] - m[4 * f + 2];
      lub[4 * f + 3] = r21[2 * f] - m[4 * f + 3];
    }
}
