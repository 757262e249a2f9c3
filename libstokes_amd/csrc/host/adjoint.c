/* adjoint.c -- gradients of the F-version mobility product (host side).
 *
 * For a scalar g . M(x) f (g, f, x of 3 np): d/df = M^T g = M g (the F
 * mobility is symmetric) and d/dx by the adjoint kernels (adj_kernels.h),
 * which recompute the pair scalars and their radial derivatives instead of
 * storing them.  A loss L(y) with y = M(x) f back-propagates with g = dL/dy.
 */
#include "stokes_hip.h"
#include "stokes_internal.h"

int
stokes_hip_adjoint_dev (struct stokes *sys, const double *g_dev,
                        const double *f_dev, double *dx_dev, double *df_dev)
{
  if (sys == NULL || g_dev == NULL || (dx_dev != NULL && f_dev == NULL))
    {
      stokes_hip_set_error ("adjoint: needs g, and f for d/dx");
      return STOKES_HIP_EARG;
    }
  if (sys->version != 0)
    {
      stokes_hip_set_error ("adjoint: F version only (version %d)",
                            sys->version);
      return STOKES_HIP_EUNSUP;
    }
  stokes_ctx *c = stokes_hip_ctx (sys);
  if (!c) return STOKES_HIP_ENODEV;
  int rc = stokes_hip_prepare (c, c->pos_dev != NULL);
  if (rc) return rc;
  if (dx_dev)
    {
      rc = shk_dev_adjoint_x (c->dev, f_dev, g_dev, dx_dev, c->prec);
      if (rc)
        {
          stokes_hip_set_error ("adjoint: %s", shk_last_error ());
          return rc;
        }
    }
  if (df_dev)
    {
      rc = shk_dev_atimes (c->dev, g_dev, df_dev, 1, c->prec);
      if (rc)
        {
          stokes_hip_set_error ("adjoint: %s", shk_last_error ());
          return rc;
        }
    }
  return 0;
}
