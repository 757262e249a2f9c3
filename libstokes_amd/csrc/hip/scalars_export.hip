// scalars_export.hip -- host-callable scalar functions of the reference API
// (ewald.h, non-ewald.h).  They evaluate the same templates as the device
// kernels (pair_math.h), so the exported values are those the GPU uses.
#include <hip/hip_runtime.h>
#include "pair_math.h"

namespace {
template <typename F>
void dispatch (int version, F f)
{
  if (version == 0) f (std::integral_constant<int, 0> ());
  else if (version == 1) f (std::integral_constant<int, 1> ());
  else f (std::integral_constant<int, 2> ());
}

void unpack (const double *o, double *xa, double *ya, double *yb, double *xc, double *yc,
             double *xg, double *yg, double *yh, double *xm, double *ym, double *zm)
{
  *xa = o[0]; *ya = o[1]; *yb = o[2]; *xc = o[3]; *yc = o[4];
  *xg = o[5]; *yg = o[6]; *yh = o[7]; *xm = o[8]; *ym = o[9]; *zm = o[10];
}
}  // namespace

extern "C" void
scalars_ewald_real (int version, double xi, double r, double *xa, double *ya, double *yb,
                    double *xc, double *yc, double *xg, double *yg, double *yh, double *xm,
                    double *ym, double *zm)
{
  double o[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  dispatch (version, [&] (auto v) { shk::scalars_real<decltype (v)::value, double> (xi, r, o); });
  unpack (o, xa, ya, yb, xc, yc, xg, yg, yh, xm, ym, zm);
}

extern "C" void
scalars_ewald_real_poly (int version, double xi, double r, double aa, double ab, double *xa,
                         double *ya, double *yb, double *xc, double *yc, double *xg, double *yg,
                         double *yh, double *xm, double *ym, double *zm)
{
  double o[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  dispatch (version, [&] (auto v) {
    shk::scalars_real_poly_sd<decltype (v)::value, double, false> (xi, r, aa, ab, o);
  });
  unpack (o, xa, ya, yb, xc, yc, xg, yg, yh, xm, ym, zm);
}

extern "C" void
scalars_mob_poly_scale_SD_ (int version, double a1, double *xa, double *ya, double *yb,
                            double *xc, double *yc, double *xg, double *yg, double *yh,
                            double *xm, double *ym, double *zm)
{
  // ewald.c:382-420
  *xa *= a1; *ya *= a1;
  if (version == 0) return;
  const double a12 = a1 * a1, a13 = a12 * a1;
  *yb *= a12;
  *xc *= a13; *yc *= a13;
  if (version == 1) return;
  *xg *= a12; *yg *= a12;
  *yh *= a13;
  *xm *= a13; *ym *= a13; *zm *= a13;
}

extern "C" void
scalars_nonewald (int version, double r, double *scalar)
{
  for (int k = 0; k < 11; k++) scalar[k] = 0.0;
  dispatch (version, [&] (auto v) {
    shk::scalars_rpy_sd<decltype (v)::value, double> (r, 1.0, 1.0, false, scalar);
  });
}

// F-version scalars with their radial derivatives (the adjoint's inputs):
// out = (xa, ya, dxa/dr, dya/dr)
extern "C" void
stokes_hip_scalars_real_f_deriv (double xi, double r, double *out)
{
  shk::scalars_real_f_deriv<double> (xi, r, 1.0 / r, out, out + 1, out + 2, out + 3);
}
