"""torch.autograd wrapper of the F-version mobility product y = M(x) f.

The forward pass is the library's operator (stokes_hip_atimes_dev), the
backward pass its adjoint (stokes_hip_adjoint_dev: d(g.M(x)f)/dx recomputed
from the pair scalars, and M^T g = M g).  The tensors are float64 CPU
tensors of 3N; the work runs on the GPU through the library's own device
buffers, so torch's GPU runtime is never initialised in the process (torch
bundles a HIP runtime of its own, and two HIP runtimes do not share one
process).  The arithmetic follows the precision mode of the Stokes object
(FP64, or FP32 pair and reciprocal sums for MIXED).
"""
import numpy as np
import torch

from . import DeviceArray, check


class MobilityProductF(torch.autograd.Function):
    """y = M(x) f for a periodic F-version system (equal spheres)."""

    @staticmethod
    def forward(ctx, x, f, stokes):
        if x.dtype != torch.float64 or f.dtype != torch.float64:
            raise TypeError("positions and forces must be float64")
        if x.is_cuda or f.is_cuda:
            raise TypeError("pass CPU tensors (the library holds the device "
                            "copies)")
        xn = x.detach().contiguous().numpy()
        fn = f.detach().contiguous().numpy()
        xd = DeviceArray.from_host(xn)
        fd = DeviceArray.from_host(fn)
        yd = DeviceArray(fn.size)
        L = stokes.L
        check(L.stokes_hip_set_pos_dev(stokes.sys, xd.ptr), "set_pos_dev")
        check(L.stokes_hip_atimes_dev(stokes.sys, 1, fd.ptr, yd.ptr),
              "atimes_dev")
        y = yd.to_host()
        check(L.stokes_hip_use_host_pos(stokes.sys), "use_host_pos")
        ctx.save_for_backward(x, f)
        ctx.stokes = stokes
        return torch.from_numpy(y)

    @staticmethod
    def backward(ctx, gy):
        x, f = ctx.saved_tensors
        stokes = ctx.stokes
        xd = DeviceArray.from_host(x.detach().contiguous().numpy())
        fd = DeviceArray.from_host(f.detach().contiguous().numpy())
        gd = DeviceArray.from_host(
            np.ascontiguousarray(gy.detach().numpy(), dtype=np.float64))
        dxd = DeviceArray(xd.n)
        dfd = DeviceArray(fd.n)
        L = stokes.L
        check(L.stokes_hip_set_pos_dev(stokes.sys, xd.ptr), "set_pos_dev")
        check(L.stokes_hip_adjoint_dev(stokes.sys, gd.ptr, fd.ptr, dxd.ptr,
                                       dfd.ptr), "adjoint")
        dx, df = dxd.to_host(), dfd.to_host()
        check(L.stokes_hip_use_host_pos(stokes.sys), "use_host_pos")
        return torch.from_numpy(dx), torch.from_numpy(df), None


def mobility_product(x, f, stokes):
    """differentiable y = M(x) f"""
    return MobilityProductF.apply(x, f, stokes)
