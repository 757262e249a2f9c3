"""torch.autograd wrapper of the F-version mobility product y = M(x) f.

The forward pass is the library's operator (stokes_hip_atimes_dev), the
backward pass its adjoint (stokes_hip_adjoint_dev: d(g.M(x)f)/dx recomputed
from the pair scalars, and M^T g = M g).  Tensors are float64 vectors of
3N on the GPU; the arithmetic inside follows the precision mode of the
Stokes object (FP64, or FP32 pair and reciprocal sums for MIXED).  torch
is imported by the caller; this module only needs it at call time.
"""
import ctypes as C

import torch

from . import check


class MobilityProductF(torch.autograd.Function):
    """y = M(x) f for a periodic F-version system (equal spheres)."""

    @staticmethod
    def forward(ctx, x, f, stokes):
        if x.dtype != torch.float64 or f.dtype != torch.float64:
            raise TypeError("positions and forces must be float64")
        x = x.contiguous()
        f = f.contiguous()
        y = torch.empty_like(f)
        torch.cuda.synchronize()
        L = stokes.L
        check(L.stokes_hip_set_pos_dev(stokes.sys, C.c_void_p(x.data_ptr())),
              "set_pos_dev")
        check(L.stokes_hip_atimes_dev(stokes.sys, 1, C.c_void_p(f.data_ptr()),
                                      C.c_void_p(y.data_ptr())), "atimes_dev")
        stokes.sync()
        check(L.stokes_hip_use_host_pos(stokes.sys), "use_host_pos")
        ctx.save_for_backward(x, f)
        ctx.stokes = stokes
        return y

    @staticmethod
    def backward(ctx, gy):
        x, f = ctx.saved_tensors
        stokes = ctx.stokes
        gy = gy.contiguous()
        dx = torch.empty_like(x)
        df = torch.empty_like(f)
        torch.cuda.synchronize()
        L = stokes.L
        check(L.stokes_hip_set_pos_dev(stokes.sys, C.c_void_p(x.data_ptr())),
              "set_pos_dev")
        check(L.stokes_hip_adjoint_dev(stokes.sys, C.c_void_p(gy.data_ptr()),
                                       C.c_void_p(f.data_ptr()),
                                       C.c_void_p(dx.data_ptr()),
                                       C.c_void_p(df.data_ptr())), "adjoint")
        stokes.sync()
        check(L.stokes_hip_use_host_pos(stokes.sys), "use_host_pos")
        return dx, df, None


def mobility_product(x, f, stokes):
    """differentiable y = M(x) f"""
    return MobilityProductF.apply(x, f, stokes)
