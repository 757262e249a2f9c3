"""The benchmark workloads at their full sizes, on the GPU.

* Dense mobility at N = 1024 FTS, phi = 0.3 (make_matrix_mob_ewald_3all,
  ewald.c:1092-1677): the 11264^2 FP64 extracted-form matrix built on the
  device; mat . T(x) equals the matrix-free product (the invariant of the
  reference's check-ewald.c:46-197) and the oracle's rows.
* Brownian step at N = 2048 F, phi = 0.2, order-30 Chebyshev: Lanczos
  bounds found by the step itself enclose the spectrum of the device's own
  6144^2 FP64 matrix (numpy eigvalsh); the step's noise equals the
  Chebyshev recursion on that dense matrix, and its accuracy figure
  |z.Mz - y.y| / z.Mz (chebyshev.c:241-269) equals the one of that exact
  recursion and stays below the fit's worst case over the spectrum; the
  64-column noise block through the cached real-space matrix (split-K
  GEMM) equals the matrix-free block column by column.
* Batched ensemble of 256 polydisperse FTS boxes (N log-uniform on
  [50, 2000], radii log-uniform on [0.5, 2], phi = 0.25), MIXED in one
  graph launch: every box against its own FP64 product (1e-6 normwise),
  and the FP64 products of the smallest, the largest and three other boxes
  against the oracle's rows (1e-12).
"""
import math
import time

import numpy as np
import pytest

import libstokes_amd as ls
from libstokes_amd import workloads
import pyoracle as po
from conftest import relerr

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_dense_build_n1024_fts():
    n, phi = 1024, 0.3
    pos, L = workloads.configuration(n, phi, seed=31)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    d = 11 * n
    _say(f"dense build N={n} FTS: {d}^2 FP64 = {d * d * 8 / 1e9:.3f} GB")
    mat = ls.DeviceArray(d * d)
    t0 = time.perf_counter()
    s.matrix_dev(mat)
    s.sync()
    _say(f"  build {time.perf_counter() - t0:.3f} s (first call, incl. set-up)")
    x = np.random.default_rng(32).standard_normal(d)
    xd = ls.DeviceArray.from_host(x)
    y1, y2 = ls.DeviceArray(d), ls.DeviceArray(d)
    s.atimes_dev(xd, y1)
    s.extmat_dev(mat, xd, y2)
    ya, ym = y1.to_host(), y2.to_host()
    e_inv = relerr(ym, ya)
    rows = 4
    yo = po.System(2, pos, box=(L, L, L), tr=1.0).atimes(x, rows=(0, rows))
    e_rows = relerr(ym[:11 * rows], yo[:11 * rows])
    _say(f"  mat.T(x) vs atimes {e_inv:.2e}, rows 0..{rows - 1} vs oracle {e_rows:.2e}")
    assert e_inv <= 1e-12
    assert e_rows <= 1e-12
    mat.free()
    s.close()


@pytest.mark.parametrize("precision", [ls.FP64, ls.MIXED])
def test_bd_step_n2048_order30_own_bounds(precision):
    n, phi, nch, dt = 2048, 0.2, 30, 1e-3
    name = "MIXED" if precision == ls.MIXED else "FP64"
    _say(f"BD N={n} F phi={phi} order {nch} {name}")
    pos, L = workloads.configuration(n, phi, seed=41)
    s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0, precision=precision)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    bd = ls.BD(s, peclet=1.0, n_cheb=nch, eps=1e-2, dt_lim=1e-6, rmin_ol=1.0,
               seed=9)
    info = bd.step(xd, dt)
    e0, e1 = info.eig
    y, z = bd.last_noise()
    _say(f"  {info.summary()}")
    # the device's own FP64 matrix at the positions of the noise
    s64 = ls.Stokes(0, pos, box=(L, L, L), tr=1.0)
    M = s64.matrix()
    s64.close()
    M = 0.5 * (M + M.T)
    t0 = time.perf_counter()
    w = np.linalg.eigvalsh(M)
    _say(f"  eigvalsh of {3 * n}^2 in {time.perf_counter() - t0:.1f} s: "
         f"spectrum [{w[0]:.5g}, {w[-1]:.5g}], bounds [{e0:.5g}, {e1:.5g}]")
    assert e0 < w[0] and e1 > w[-1]
    assert e0 > 0.5 * w[0] and e1 < 2.0 * w[-1]
    a = po.cheb_coef(nch, 0, e0, e1)
    zd = po.cheb_eval_dense(a, M, y, e0, e1)
    ez = relerr(z, zd)
    zMz = zd @ (M @ zd)
    err = abs(zMz - y @ y) / zMz
    # the fit's worst case over the spectrum: |w f(w)^2 - 1| (z = f(M) y)
    fw = np.array([po.cheb_eval(a, e0, e1, v) for v in w])
    worst = np.max(np.abs(w * fw * fw - 1.0))
    _say(f"  z vs dense recursion {ez:.2e}; err_cheb step {info.err_cheb:.4e} "
         f"dense {err:.4e}, fit worst case {worst:.4e}")
    assert ez <= (1e-10 if precision == ls.FP64 else 1e-5)
    assert info.err_cheb == pytest.approx(
        err, rel=1e-6 if precision == ls.FP64 else 2e-2)
    assert err <= worst
    bd.close()
    # the 64-column noise block: cached real-space matrix vs matrix-free
    s.set_split(10.0)
    yb = np.random.default_rng(43).standard_normal(3 * n * 64)
    ybd = ls.DeviceArray.from_host(yb)
    zb = ls.DeviceArray(3 * n * 64)
    out = []
    for cache in (False, True):
        s.set_real_cache(cache)
        s.cheb_dev(a, ybd, zb, e0, e1, nrhs=64)
        s.sync()
        out.append(zb.to_host().reshape(64, 3 * n))
    s.set_real_cache(False)
    ec = max(relerr(out[1][c], out[0][c]) for c in range(64))
    e_dense = max(relerr(out[1][c], po.cheb_eval_dense(
        a, M, yb[c * 3 * n:(c + 1) * 3 * n], e0, e1)) for c in (0, 63))
    _say(f"  64-column block: cached vs matrix-free {ec:.2e} (max over columns),"
         f" vs dense recursion {e_dense:.2e}")
    assert ec <= (1e-12 if precision == ls.FP64 else 2e-6)
    assert e_dense <= (1e-10 if precision == ls.FP64 else 1e-5)
    s.close()


def test_ensemble_256_boxes_bench_size():
    nbox = 256
    rng = np.random.default_rng(0)
    ns = np.exp(rng.uniform(math.log(50), math.log(2000), nbox)).astype(int)
    t0 = time.perf_counter()
    boxes = [workloads.configuration_poly(int(nn), 0.25, seed=1000 + k)
             for k, nn in enumerate(ns)]
    _say(f"ensemble: {nbox} boxes, sum N = {ns.sum()}, configurations in "
         f"{time.perf_counter() - t0:.1f} s")
    systems = []
    for pos, a, L in boxes:
        s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, a=a, precision=ls.MIXED)
        s.set_split(1000.0)
        systems.append(s)
    ens = ls.Ensemble(systems)
    ens.autotune(trs=[100.0, 300.0, 1000.0, 3000.0], reps=1)
    off = ens.offsets
    tot = off[-1]
    x = np.random.default_rng(1).standard_normal(11 * tot)
    posd = ls.DeviceArray.from_host(
        np.concatenate([b[0].reshape(-1) for b in boxes]))
    xd = ls.DeviceArray.from_host(x)
    yd = ls.DeviceArray(11 * tot)
    ens.atimes_dev(posd, xd, yd)
    y = yd.to_host()
    worst = 0.0
    big, small = int(np.argmax(ns)), int(np.argmin(ns))
    sample = sorted({big, small, 7, 100, 200})
    for b, (pos, a, L) in enumerate(boxes):
        sl = slice(11 * off[b], 11 * off[b + 1])
        s64 = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, a=a)
        y64 = s64.atimes(x[sl])
        s64.close()
        e = relerr(y[sl], y64)
        worst = max(worst, e)
        assert e <= 1e-6, (b, ns[b], e)
        if b in sample:
            o = po.System(2, pos, box=(L, L, L), tr=1.0, a=a)
            yo = o.atimes(x[sl], rows=(0, 4))
            er = relerr(y64[:44], yo[:44])
            _say(f"  box {b} N={ns[b]}: MIXED vs FP64 {e:.2e}, FP64 rows 0..3 "
                 f"vs oracle {er:.2e}")
            assert er <= 1e-12
    _say(f"  all {nbox} boxes: worst MIXED vs FP64 {worst:.2e}")
    ens.close()
    for s in systems:
        s.close()
