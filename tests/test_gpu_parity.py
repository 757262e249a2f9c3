"""Parity of the HIP operator with the oracle and the reference's golden
vectors (GPU; every call goes through the C ABI of libstokes_hip.so).

Bars (normwise relative error ||y - y_ref|| / ||y_ref||):
  FP64 path   <= 1e-12 against the oracle and the golden vectors
  MIXED path  <= 1e-6  against the FP64 oracle
Full-size configurations are checked on a row sample of the oracle and by
size-independent properties (FP64 vs MIXED, translation invariance,
determinism, operator == dense matrix).
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import libstokes_amd as ls
from libstokes_amd import workloads
import pyoracle as po
from conftest import relerr, maxrel, ROOT

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOLMIX = 1e-6
FIX9_L = 5.733683


def _say(msg):
    print(msg, flush=True)


def _gpu(meta, data, precision=ls.FP64):
    k = meta["name"]
    a = data[f"{k}/a"] if meta["poly"] else None
    return ls.Stokes(meta["version"], data[f"{k}/pos"], box=meta["box"],
                     tr=meta["tr"], a=a, rmin=meta["rmin"],
                     shear_mode=meta["shear_mode"],
                     shear_shift=meta["shear_shift"], precision=precision)


def _oracle(meta, data):
    k = meta["name"]
    a = data[f"{k}/a"] if meta["poly"] else None
    return po.System(meta["version"], data[f"{k}/pos"], box=meta["box"],
                     tr=meta["tr"], a=a, rmin=meta["rmin"],
                     shear_mode=meta["shear_mode"],
                     shear_shift=meta["shear_shift"])


def test_golden_atimes_fp64(golden):
    data, cases = golden
    for meta in cases:
        k = meta["name"]
        _say(f"case {k}")
        s = _gpu(meta, data)
        y = s.atimes(data[f"{k}/x"])
        e_ref = relerr(y, data[f"{k}/y"])
        e_orc = relerr(y, _oracle(meta, data).atimes(data[f"{k}/x"]))
        s.close()
        assert e_orc <= TOL64, (k, e_orc)
        # N = 1 cases cancel to ~1e-1 of their terms (see test_oracle)
        assert e_ref <= (5e-12 if k.startswith("sc1") else TOL64), (k, e_ref)


def test_golden_matrix_fp64(golden):
    data, cases = golden
    for meta in cases:
        if not meta["matrix"]:
            continue
        k = meta["name"]
        _say(f"matrix {k}")
        s = _gpu(meta, data)
        m = s.matrix()
        s.close()
        assert maxrel(m, data[f"{k}/mat"]) <= TOL64, k


@pytest.mark.parametrize("version", [0, 1, 2])
def test_poly_random_radii(version):
    """polydisperse Ewald (radii 0.5..2, the ensemble workload's range):
    FP64 against the oracle, MIXED within 1e-6"""
    rng = np.random.default_rng(30 + version)
    n = 40
    a = np.exp(rng.uniform(np.log(0.5), np.log(2.0), n))
    L = (np.sum(4.0 / 3.0 * np.pi * a ** 3) / 0.25) ** (1.0 / 3.0)
    pos = rng.uniform(0, L, (n, 3))
    x = rng.standard_normal(n * ls.dim(version))
    ref = po.System(version, pos, box=(L,) * 3, tr=3.0, a=a, rmin=1.0).atimes(x)
    s = ls.Stokes(version, pos, box=(L,) * 3, tr=3.0, a=a, rmin=1.0)
    e64 = relerr(s.atimes(x), ref)
    s.set_precision(ls.MIXED)
    emix = relerr(s.atimes(x), ref)
    s.close()
    _say(f"poly v{version}: fp64 {e64:.2e} mixed {emix:.2e}")
    assert e64 <= TOL64
    assert emix <= TOLMIX


@pytest.mark.parametrize("version", [0, 1, 2])
def test_c1_n64_fp64_and_mixed(version):
    pos, L = workloads.configuration(64, 0.1, seed=11)
    x = np.random.default_rng(12).standard_normal(64 * ls.dim(version))
    ref = po.System(version, pos, box=(L,) * 3, tr=1.0).atimes(x)
    s = ls.Stokes(version, pos, box=(L,) * 3, tr=1.0)
    e64 = relerr(s.atimes(x), ref)
    s.set_precision(ls.MIXED)
    emix = relerr(s.atimes(x), ref)
    s.close()
    _say(f"N=64 v{version}: fp64 {e64:.2e} mixed {emix:.2e}")
    assert e64 <= TOL64
    assert emix <= TOLMIX


def test_recip_g_fix_mode():
    """corrected reciprocal G sign: GPU == oracle in that mode, and the FTS
    result no longer depends on xi (FP64 and MIXED)"""
    pos, L = workloads.configuration(64, 0.2, seed=15)
    x = np.random.default_rng(16).standard_normal(64 * 11)
    ys = {}
    for tr in (1.0, 100.0):
        s = ls.Stokes(2, pos, box=(L,) * 3, tr=tr)
        s.set_recip_g_fix(True)
        ys[tr] = s.atimes(x)
        s.set_precision(ls.MIXED)
        ys[(tr, "mix")] = s.atimes(x)
        s.close()
    ref = po.System(2, pos, box=(L,) * 3, tr=1.0, fix_recip_g=True).atimes(x)
    assert relerr(ys[1.0], ref) <= TOL64
    assert relerr(ys[100.0], ys[1.0]) <= 1e-10
    assert relerr(ys[(100.0, "mix")], ys[1.0]) <= TOLMIX
    assert relerr(ys[(1.0, "mix")], ys[1.0]) <= TOLMIX


def test_determinism_bitwise():
    pos, L = workloads.configuration(200, 0.2, seed=3)
    x = np.random.default_rng(4).standard_normal(200 * 11)
    s = ls.Stokes(2, pos, box=(L,) * 3, tr=1.0)
    y1 = s.atimes(x)
    y2 = s.atimes(x)
    s.close()
    assert np.array_equal(y1, y2)


def test_device_api_multi_rhs_and_extmat(golden):
    data, _ = golden
    pos = data["fix9_FTS/pos"]
    s = ls.Stokes(2, pos, box=(FIX9_L,) * 3, tr=1.0, shear_mode=2,
                  shear_shift=0.8)
    n = s.n
    X = np.random.default_rng(9).standard_normal((3, n))
    xd = ls.DeviceArray.from_host(X)
    yd = ls.DeviceArray(3 * n)
    s.atimes_dev(xd, yd, nrhs=3)
    Y = yd.to_host().reshape(3, n)
    o = po.System(2, pos, box=(FIX9_L,) * 3, tr=1.0, shear_mode=2,
                  shear_shift=0.8)
    for c in range(3):
        assert relerr(Y[c], o.atimes(X[c])) <= TOL64
    md = ls.DeviceArray(n * n)
    s.matrix_dev(md)
    zd = ls.DeviceArray(3 * n)
    s.extmat_dev(md, xd, zd, nrhs=3)
    Z = zd.to_host().reshape(3, n)
    assert relerr(Z, Y) <= TOL64
    for a in (xd, yd, md, zd):
        a.free()
    s.close()


@pytest.mark.parametrize("version,poly,nrhs", [
    (0, False, 70), (1, False, 9), (2, False, 5), (2, True, 3), (0, True, 16),
    (2, False, 64), (2, True, 40), (2, False, 130), (2, True, 129)])
def test_block_rhs(version, poly, nrhs):
    """several right-hand sides in one call (block kernels): every column
    equals the oracle's single matvec; MIXED within 1e-6"""
    rng = np.random.default_rng(40 + version + nrhs)
    n = 150
    a = np.exp(rng.uniform(np.log(0.6), np.log(1.6), n)) if poly else None
    pos, L = workloads.configuration(n, 0.15, seed=41)
    d = ls.dim(version)
    X = rng.standard_normal((nrhs, n * d))
    s = ls.Stokes(version, pos, box=(L,) * 3, tr=20.0, a=a, shear_mode=1,
                  shear_shift=0.3 * L)
    o = po.System(version, pos, box=(L,) * 3, tr=20.0, a=a, shear_mode=1,
                  shear_shift=0.3 * L)
    xd = ls.DeviceArray.from_host(X)
    yd = ls.DeviceArray(nrhs * n * d)
    s.atimes_dev(xd, yd, nrhs=nrhs)
    Y = yd.to_host().reshape(nrhs, n * d)
    s.set_precision(ls.MIXED)
    s.atimes_dev(xd, yd, nrhs=nrhs)
    Ym = yd.to_host().reshape(nrhs, n * d)
    for c in sorted({0, nrhs // 2, nrhs - 1}):
        yo = o.atimes(X[c])
        assert relerr(Y[c], yo) <= TOL64, c
        assert relerr(Ym[c], yo) <= TOLMIX, c
    xd.free()
    yd.free()
    s.close()


@pytest.mark.parametrize("nrhs", [16, 40, 130])
def test_block_pair_moments(nrhs):
    """equal spheres in an unsheared box (symmetric image table): the block
    path through the per-pair moments (each pair i <= j evaluated once, the
    other order by parity) and the matrix cores, FP64 (v_mfma_f64) and
    MIXED, every sampled column against the oracle's single matvec"""
    rng = np.random.default_rng(70 + nrhs)
    n = 200
    pos, L = workloads.configuration(n, 0.2, seed=71)
    X = rng.standard_normal((nrhs, n * 11))
    s = ls.Stokes(2, pos, box=(L,) * 3, tr=20.0)
    o = po.System(2, pos, box=(L,) * 3, tr=20.0)
    xd = ls.DeviceArray.from_host(X)
    yd = ls.DeviceArray(nrhs * n * 11)
    s.atimes_dev(xd, yd, nrhs=nrhs)
    Y = yd.to_host().reshape(nrhs, -1)
    s.set_precision(ls.MIXED)
    s.atimes_dev(xd, yd, nrhs=nrhs)
    Ym = yd.to_host().reshape(nrhs, -1)
    for c in sorted({0, 15, nrhs // 2, nrhs - 1}):
        yo = o.atimes(X[c])
        assert relerr(Y[c], yo) <= TOL64, c
        assert relerr(Ym[c], yo) <= TOLMIX, c
    xd.free()
    yd.free()
    s.close()


@pytest.mark.parametrize("tr,B,dense", [(3000.0, 32, False), (300.0, 7, False),
                                        (300.0, 128, False), (300.0, 128, True),
                                        (1000.0, 200, True)])
def test_block_mixed_full_size(tr, B, dense):
    """N = 1024, phi = 0.3 FTS: the MIXED block path (FP32 matrix-core
    sums; with dense blocks allowed, beyond 32 columns the real-space part
    as a per-call dense FP32 matrix times the block, 200 columns: two
    passes on one matrix) against the FP64 block path column by column,
    1e-6; bitwise repeatable"""
    n = 1024
    pos, L = workloads.configuration(n, 0.3, seed=1)
    X = np.random.default_rng(3).standard_normal(B * 11 * n)
    s = ls.Stokes(2, pos, box=(L,) * 3, tr=1.0)
    s.set_split(tr)
    s.set_dense_blocks(dense)
    xd = ls.DeviceArray.from_host(X)
    yd = ls.DeviceArray(X.size)
    s.atimes_dev(xd, yd, nrhs=B)
    Y = yd.to_host().reshape(B, -1)
    s.set_precision(ls.MIXED)
    s.atimes_dev(xd, yd, nrhs=B)
    Ym = yd.to_host().reshape(B, -1)
    s.atimes_dev(xd, yd, nrhs=B)
    Ym2 = yd.to_host().reshape(B, -1)
    errs = [relerr(Ym[c], Y[c]) for c in range(B)]
    _say(f"  N={n} tr={tr:g} B={B}: MIXED block vs FP64 block max "
         f"{max(errs):.2e}")
    assert max(errs) <= TOLMIX
    assert np.array_equal(Ym, Ym2)
    # one column against the single-vector FP64 path
    assert relerr(s_single(s, X[:11 * n]), Y[0]) <= TOL64
    xd.free()
    yd.free()
    s.close()


@pytest.mark.parametrize("tr", [3e3, 3e4, 1e5])
def test_block_mixed_phase_table_widths(tr):
    """MIXED block path at splittings whose phase tables hold 78 (tr = 3e3),
    113 (3e4) and 138 (1e5) entries per particle, i.e. the 96-, 128- and
    160-entry layouts of the matrix-core structure factors: every column
    against the FP64 block path, 1e-6, and column 0 of the FP64 block
    against the single-column FP64 path (other kernels; both are pinned to
    the oracle at smaller splittings, where the oracle runs in seconds);
    N = 100, phi = 0.05 FTS, 40 columns"""
    n, B = 100, 40
    pos, L = workloads.configuration(n, 0.05, seed=5)
    X = np.random.default_rng(9).standard_normal(B * 11 * n)
    s = ls.Stokes(2, pos, box=(L,) * 3, tr=tr)
    xd = ls.DeviceArray.from_host(X)
    yd = ls.DeviceArray(X.size)
    s.atimes_dev(xd, yd, nrhs=B)
    Y = yd.to_host().reshape(B, -1)
    s.set_precision(ls.MIXED)
    s.atimes_dev(xd, yd, nrhs=B)
    Ym = yd.to_host().reshape(B, -1)
    errs = [relerr(Ym[c], Y[c]) for c in range(B)]
    _say(f"  tr={tr:g}: MIXED block vs FP64 block max {max(errs):.2e}")
    assert max(errs) <= TOLMIX
    assert relerr(s_single(s, X[:11 * n]), Y[0]) <= TOL64
    xd.free()
    yd.free()
    s.close()


@pytest.mark.parametrize("version,prec,poly,B", [
    (2, "FP64", False, 128), (2, "FP64", False, 32), (1, "FP64", True, 16),
    (0, "FP64", True, 8), (2, "MIXED", False, 128)])
def test_block_overlap_matches_serial(version, prec, poly, B):
    """FP64 block products with the reciprocal part on the side stream
    (stokes_hip_set_overlap, the default) equal those with both parts on
    one stream, bit for bit, over three calls in a row (the side stream's
    buffers reused while the next call's real-space part runs); MIXED
    products, serial either way, likewise"""
    n = 96
    a = None
    if poly:
        pos, a, L = workloads.configuration_poly(n, 0.1, seed=4, a_lo=0.5, a_hi=1.5)
    else:
        pos, L = workloads.configuration(n, 0.2, seed=4)
    D = (3, 6, 11)[version]
    rng = np.random.default_rng(17)
    s = ls.Stokes(version, pos, a=a, box=(L,) * 3, tr=10.0)
    s.set_precision(getattr(ls, prec))
    xs = [rng.standard_normal(B * D * n) for _ in range(3)]
    out = {}
    for on in (False, True):
        s.set_overlap(on)
        ys = []
        for X in xs:
            xd = ls.DeviceArray.from_host(X)
            yd = ls.DeviceArray(X.size)
            s.atimes_dev(xd, yd, nrhs=B)
            ys.append(yd.to_host())
            xd.free()
            yd.free()
        out[on] = ys
    for k in range(3):
        d = np.abs(out[True][k] - out[False][k]).max()
        _say(f"  v={version} {prec} poly={poly} B={B} call {k}: |on - off| {d:.1e}")
        assert d == 0.0
    s.close()


def s_single(s, x):
    s.set_precision(ls.FP64)
    return s.atimes(x)


def test_open_boundary_fp64(golden):
    data, cases = golden
    for meta in cases:
        if not meta["name"].startswith("rpy_"):
            continue
        k = meta["name"]
        s = _gpu(meta, data)
        assert relerr(s.atimes(data[f"{k}/x"]), data[f"{k}/y"]) <= TOL64, k
        s.close()


def test_chebyshev_device_recursion(golden):
    """z = M^{-1/2} y by the device Clenshaw recursion (F version, the
    brownian.c route) against the oracle's recursion on the same bounds"""
    data, _ = golden
    pos = data["fix9_F/pos"]
    o = po.System(0, pos, box=(FIX9_L,) * 3, tr=1.0)
    ev = np.linalg.eigvalsh(o.matrix())
    e0, e1 = 0.95 * ev[0], 1.05 * ev[-1]
    a = po.cheb_coef(30, 0, e0, e1)
    y = np.random.default_rng(10).standard_normal((2, 27))
    s = ls.Stokes(0, pos, box=(FIX9_L,) * 3, tr=1.0)
    yd = ls.DeviceArray.from_host(y)
    zd = ls.DeviceArray(54)
    s.cheb_dev(a, yd, zd, e0, e1, nrhs=2)
    z = zd.to_host().reshape(2, 27)
    for c in range(2):
        zo = po.cheb_eval_sys(a, o, y[c], e0, e1)
        assert relerr(z[c], zo) <= TOL64
    # the Chebyshev value approximates M^{-1/2} y
    m = o.matrix()
    w, V = np.linalg.eigh(m)
    exact = V @ ((V.T @ y[0]) / np.sqrt(w))
    assert relerr(z[0], exact) < 1e-6
    s.close()


def test_lanczos_bounds(golden):
    data, _ = golden
    pos = data["fix9_F/pos"]
    o = po.System(0, pos, box=(FIX9_L,) * 3, tr=1.0)
    ev = np.linalg.eigvalsh(o.matrix())
    s = ls.Stokes(0, pos, box=(FIX9_L,) * 3, tr=1.0)
    lo, hi = s.lanczos_bounds(niter=27, seed=3, margin=0.0)
    s.close()
    assert lo == pytest.approx(ev[0], rel=1e-8)
    assert hi == pytest.approx(ev[-1], rel=1e-8)


def test_legacy_void_entry_points():
    """atimes_3all / make_matrix_mob_3all / atimes_3all_matrix /
    dot_prod_matrix / multiply_extmat_with_extvec_3fts /
    chebyshev_eval_atimes through their reference prototypes (run in a
    child process: on failure they exit(1) like the reference)."""
    code = r"""
import ctypes as C, sys, numpy as np
sys.path.insert(0, %r); sys.path.insert(0, %r)
import libstokes_amd as ls, pyoracle as po
pos = np.random.default_rng(1).uniform(0, 7, (6, 3)); L = 7.0
s = ls.Stokes(2, pos, box=(L, L, L), tr=2.0)
o = po.System(2, pos, box=(L, L, L), tr=2.0)
n = s.n; x = np.random.default_rng(2).standard_normal(n)
P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
lib = ls.lib(); y = np.zeros(n)
lib.atimes_3all(C.c_int(n), P(x), P(y), s.sys)
assert np.linalg.norm(y - o.atimes(x)) <= 1e-12 * np.linalg.norm(y)
m = np.zeros((n, n)); lib.make_matrix_mob_ewald_3all(s.sys, P(m))
assert np.abs(m - o.matrix()).max() <= 1e-12 * np.abs(m).max()
y2 = np.zeros(n); lib.atimes_3all_matrix(C.c_int(n), P(x), P(y2), s.sys)
assert np.linalg.norm(y2 - y) <= 1e-12 * np.linalg.norm(y)
y3 = np.zeros(n); lib.multiply_extmat_with_extvec_3fts(C.c_int(6), P(m), P(x), P(y3))
assert np.linalg.norm(y3 - po.extmat_fts(6, m, x)) <= 1e-13 * np.linalg.norm(y3)
y4 = np.zeros(n); lib.dot_prod_matrix(P(m), C.c_int(n), C.c_int(n), P(x), P(y4))
assert np.linalg.norm(y4 - m @ x) <= 1e-13 * np.linalg.norm(y4)
s0 = ls.Stokes(0, pos, box=(L, L, L), tr=2.0); o0 = po.System(0, pos, box=(L, L, L), tr=2.0)
ev = np.linalg.eigvalsh(o0.matrix()); a = po.cheb_coef(20, 0, 0.9*ev[0], 1.1*ev[-1])
yy = np.random.default_rng(3).standard_normal(18); zz = np.zeros(18)
lib.chebyshev_eval_atimes(C.c_int(20), P(a), C.c_int(18), P(yy), P(zz), C.c_double(0.9*ev[0]),
                          C.c_double(1.1*ev[-1]), lib.atimes_3all, s0.sys)
zo = po.cheb_eval_sys(a, o0, yy, 0.9*ev[0], 1.1*ev[-1])
assert np.linalg.norm(zz - zo) <= 1e-12 * np.linalg.norm(zo)
# chebyshev_error_minvsqrt / _Rsqrt (chebyshev.c:241-310) over atimes_3all,
# for the order-20 fit and for an unrelated z, against the oracle's
lib.chebyshev_error_minvsqrt.restype = C.c_double
lib.chebyshev_error_Rsqrt.restype = C.c_double
cb = C.cast(po.lib().orc_atimes_sys_cb, C.c_void_p)
for kind, fn, of in ((0, lib.chebyshev_error_minvsqrt, po.lib().orc_cheb_error_minvsqrt),
                     (1, lib.chebyshev_error_Rsqrt, po.lib().orc_cheb_error_Rsqrt)):
    for z in (zz, np.random.default_rng(4 + kind).standard_normal(18)):
        got = fn(C.c_int(18), P(yy), P(z), lib.atimes_3all, s0.sys)
        want = of(C.c_int(18), P(yy), P(z), cb, C.byref(o0.s))
        assert abs(got - want) <= 1e-10 * abs(want) + 1e-15, (kind, got, want)
        print("cheb error", kind, got, want)
# A24: with the timing split on, a legacy call writes cpu2 (real space),
# cpu3 (reciprocal space) and cpu1 = cpu2 + cpu3 (ewald.c:647, 791-794,
# 1077-1080)
s.set_timing(True)
s.s.cpu1 = s.s.cpu2 = s.s.cpu3 = -1.0
lib.atimes_3all(C.c_int(n), P(x), P(y), s.sys)
c1, c2, c3 = s.times()
print("timing", c1, c2, c3)
assert c2 > 0 and c3 > 0 and c1 == c2 + c3
print("legacy ok")
""" % (ROOT, os.path.join(ROOT, "oracle"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "legacy ok" in r.stdout


@pytest.mark.parametrize("n,phi,version", [(1024, 0.3, 2), (2048, 0.2, 0)])
def test_full_size_row_sample_and_properties(n, phi, version):
    """full workload sizes (N = 1024 FTS phi 0.3, N = 2048 F phi 0.2): FP64
    rows against the oracle's rows, MIXED against FP64, invariance under a
    common translation, determinism."""
    pos, L = workloads.configuration(n, phi, seed=21)
    d = ls.dim(version)
    x = np.random.default_rng(22).standard_normal(n * d)
    s = ls.Stokes(version, pos, box=(L,) * 3, tr=1.0)
    y64 = s.atimes(x)
    rows = 4
    yo = po.System(version, pos, box=(L,) * 3, tr=1.0).atimes(x, rows=(0, rows))
    e_rows = relerr(y64[:rows * d], yo[:rows * d])
    s.set_precision(ls.MIXED)
    ymix = s.atimes(x)
    e_mix = relerr(ymix, y64)
    s.set_precision(ls.FP64)
    s.set_pos(pos + np.array([0.37, -1.1, 2.3]))
    e_tr = relerr(s.atimes(x), y64)
    s.set_pos(pos)
    y64b = s.atimes(x)
    s.close()
    _say(f"N={n} v{version}: rows {e_rows:.2e} mixed {e_mix:.2e} translate {e_tr:.2e}")
    assert e_rows <= TOL64
    assert e_mix <= TOLMIX
    assert e_tr <= 1e-11
    assert np.array_equal(y64, y64b)


def test_block_mixed_dense_unequal_spheres():
    """unequal spheres, FTS, 48 columns: the per-call dense real-space path
    of the MIXED block against the FP64 block path (1e-6) and against the
    oracle on two columns (1e-6)"""
    n = 120
    pos, a, L = workloads.configuration_poly(n, 0.2, seed=21)
    B = 48
    X = np.random.default_rng(22).standard_normal(B * 11 * n)
    s = ls.Stokes(2, pos, box=(L,) * 3, tr=1.0, a=a)
    s.set_split(100.0)
    s.set_dense_blocks(True)
    xd = ls.DeviceArray.from_host(X)
    yd = ls.DeviceArray(X.size)
    s.atimes_dev(xd, yd, nrhs=B)
    Y = yd.to_host().reshape(B, -1)
    s.set_precision(ls.MIXED)
    s.atimes_dev(xd, yd, nrhs=B)
    Ym = yd.to_host().reshape(B, -1)
    errs = [relerr(Ym[c], Y[c]) for c in range(B)]
    _say(f"  unequal spheres N={n} B={B}: MIXED dense block vs FP64 block "
         f"max {max(errs):.2e}")
    assert max(errs) <= TOLMIX
    o = po.System(2, pos, box=(L, L, L), tr=1.0, a=a)
    for c in (0, B - 1):
        assert relerr(Ym[c], o.atimes(X[c * 11 * n:(c + 1) * 11 * n])) <= TOLMIX
    s.close()
