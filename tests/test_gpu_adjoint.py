"""Adjoint of the F-version mobility product (stokes_hip_adjoint_dev) and
its torch.autograd wrapper, against central finite differences of the FP64
GPU operator (itself parity-pinned to the oracle).

* FP64 adjoint at N = 64: d(g.M(x)f)/dx against central differences
  (h = 1e-5) on every coordinate, 1e-7 relative; the same gradient from two
  Ewald splittings (the sum and hence its gradient do not depend on the
  splitting), 1e-10.
* The C5 workload: N = 4096, phi = 0.2, L = 1/2 |M(x) f - u_t|^2 with
  f, u_t ~ N(0,1): MIXED (FP32 pair and reciprocal sums) forward and
  backward through torch.autograd against FP64 central differences
  (h = 1e-6) on 32 random coordinates, 1e-4 relative (normwise over the 32).
"""
import numpy as np
import pytest

import libstokes_amd as ls
from libstokes_amd import workloads
from conftest import relerr

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _phi(s, pos, f, g):
    s.set_pos(pos.reshape(-1, 3))
    return g @ s.atimes(f)


def test_adjoint_fp64_vs_finite_differences_n64():
    n = 64
    pos, L = workloads.configuration(n, 0.1, seed=11)
    rng = np.random.default_rng(1)
    f, g = rng.standard_normal(3 * n), rng.standard_normal(3 * n)
    s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0)
    fd_, gd_ = ls.DeviceArray.from_host(f), ls.DeviceArray.from_host(g)
    dxd, dfd = ls.DeviceArray(3 * n), ls.DeviceArray(3 * n)
    s.adjoint_dev(gd_, fd_, dxd, dfd)
    dx, df = dxd.to_host(), dfd.to_host()
    assert relerr(df, s.atimes(g)) < 1e-14
    h = 1e-5
    x0 = pos.reshape(-1).copy()
    fd = np.empty(3 * n)
    for c in range(3 * n):
        xp, xm = x0.copy(), x0.copy()
        xp[c] += h
        xm[c] -= h
        fd[c] = (_phi(s, xp, f, g) - _phi(s, xm, f, g)) / (2 * h)
    e = relerr(dx, fd)
    _say(f"  FP64 adjoint vs central differences: {e:.2e}")
    assert e < 1e-7
    # another splitting: the same gradient
    s.set_pos(pos)
    s.set_split(1000.0)
    s.adjoint_dev(gd_, fd_, dxd, None)
    e2 = relerr(dxd.to_host(), dx)
    _say(f"  tr = 1 vs tr = 1000 gradient: {e2:.2e}")
    assert e2 < 1e-10
    s.close()


def test_adjoint_c5_mixed_autograd_vs_fp64_fd():
    import torch
    from libstokes_amd.autograd import mobility_product
    n = 4096
    pos, L = workloads.configuration(n, 0.2, seed=5)
    rng = np.random.default_rng(7)
    f = rng.standard_normal(3 * n)
    ut = rng.standard_normal(3 * n)
    smix = ls.Stokes(0, pos, box=(L, L, L), tr=1.0, precision=ls.MIXED)
    smix.set_split(300.0)
    x = torch.tensor(pos.reshape(-1), dtype=torch.float64, requires_grad=True)
    ft = torch.tensor(f, dtype=torch.float64, requires_grad=True)
    ut_t = torch.tensor(ut, dtype=torch.float64)
    y = mobility_product(x, ft, smix)
    loss = 0.5 * ((y - ut_t) ** 2).sum()
    loss.backward()
    gx = x.grad.detach().numpy()
    gf = ft.grad.detach().numpy()

    # FP64 references: dL/df = M r exactly, dL/dx by central differences
    s64 = ls.Stokes(0, pos, box=(L, L, L), tr=1.0)
    s64.set_split(300.0)
    r = s64.atimes(f) - ut
    ef = relerr(gf, s64.atimes(r))
    _say(f"  dL/df mixed vs FP64: {ef:.2e}")
    assert ef < 1e-5
    x0 = pos.reshape(-1).copy()

    def loss64(xx):
        s64.set_pos(xx.reshape(-1, 3))
        return 0.5 * np.sum((s64.atimes(f) - ut) ** 2)

    coords = rng.choice(3 * n, 32, replace=False)
    h = 1e-6
    fd = np.empty(32)
    for k, c in enumerate(coords):
        xp, xm = x0.copy(), x0.copy()
        xp[c] += h
        xm[c] -= h
        fd[k] = (loss64(xp) - loss64(xm)) / (2 * h)
    e = relerr(gx[coords], fd)
    _say(f"  C5 N={n}: dL/dx (mixed, autograd) vs FP64 central FD on 32 "
         f"coordinates: {e:.2e}")
    assert e < 1e-4
    smix.close()
    s64.close()
