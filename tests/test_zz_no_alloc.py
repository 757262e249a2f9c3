"""The allocation-free single-column product (stokes_hip_atimes_dev_noalloc,
no counterpart in the reference) against the ordinary device product
(stokes_hip_atimes_dev) through the C ABI.

The two run the same kernels in the same order, the first with its scratch
in the library's per-GPU device array: results are expected to be equal;
the bar is the FP64 parity bar (normwise 1e-12) and, in MIXED precision,
the same against the ordinary MIXED product.  Two calls in a row, the
second on a smaller system; the error paths (no device set-up yet, a
scratch need beyond the array) are refused on the host before any launch.
"""
import ctypes as C

import numpy as np
import pytest

import libstokes_amd as ls
from libstokes_amd import workloads
from conftest import relerr

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _noalloc(s, xd, yd):
    f = ls.lib().stokes_hip_atimes_dev_noalloc
    f.restype = C.c_int
    f.argtypes = [type(s.sys), C.c_void_p, C.c_void_p]
    return f(s.sys, xd.ptr, yd.ptr)


def _case(version, n, phi, precision, seed):
    pos, L = workloads.configuration(n, phi, seed=seed)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0, precision=precision)
    x = np.random.default_rng(seed + 1).standard_normal(s.n)
    xd = ls.DeviceArray.from_host(x)
    yd = ls.DeviceArray(s.n)
    yn = ls.DeviceArray(s.n)
    return s, xd, yd, yn


def test_noalloc_matches_ordinary_product():
    cases = [(2, 1024, 0.3, ls.FP64, 5), (2, 64, 0.1, ls.FP64, 6),
             (2, 1024, 0.3, ls.MIXED, 7), (0, 2048, 0.2, ls.FP64, 8),
             (1, 27, 0.1, ls.MIXED, 9)]
    for version, n, phi, prec, seed in cases:
        _say(f"noalloc version={version} N={n} phi={phi} "
             f"{'MIXED' if prec == ls.MIXED else 'FP64'}")
        s, xd, yd, yn = _case(version, n, phi, prec, seed)
        s.atimes_dev(xd, yd)                 # the device set-up
        y0 = yd.to_host()
        ls.check(_noalloc(s, xd, yn), "atimes_dev_noalloc")
        y1 = yn.to_host()
        e = relerr(y1, y0)
        _say(f"  noalloc vs ordinary: rel err {e:.2e}")
        assert e <= 1e-12
        # the context's own buffers are back: the ordinary product again
        s.atimes_dev(xd, yd)
        assert relerr(yd.to_host(), y0) <= 1e-12
        # positions moved in place are read again, as by atimes_dev
        s.s.pos[0] += 0.01
        s.atimes_dev(xd, yd)
        ls.check(_noalloc(s, xd, yn), "atimes_dev_noalloc")
        assert relerr(yn.to_host(), yd.to_host()) <= 1e-12
        assert relerr(yn.to_host(), y0) > 1e-10
        s.close()


def test_noalloc_open_boundary():
    _say("noalloc open boundary FTS N=20")
    pos = np.random.default_rng(3).uniform(0, 12, (20, 3))
    s = ls.Stokes(2, pos)
    x = np.random.default_rng(4).standard_normal(s.n)
    xd = ls.DeviceArray.from_host(x)
    yd, yn = ls.DeviceArray(s.n), ls.DeviceArray(s.n)
    s.atimes_dev(xd, yd)
    ls.check(_noalloc(s, xd, yn), "atimes_dev_noalloc")
    assert relerr(yn.to_host(), yd.to_host()) <= 1e-12
    s.close()


def test_noalloc_refusals():
    # no device set-up yet: refused before any launch
    _say("noalloc refusals")
    s, xd, yd, yn = _case(0, 27, 0.1, ls.FP64, 11)
    rc = _noalloc(s, xd, yn)
    assert rc == -2
    assert b"set-up" in ls.lib().stokes_hip_last_error()
    s.atimes_dev(xd, yd)
    # a parameter changed since the set-up: refused
    s.L.stokes_set_xi(s.sys, s.s.xi * 1.5, 1e-12)
    assert _noalloc(s, xd, yn) == -2
    s.atimes_dev(xd, yd)
    ls.check(_noalloc(s, xd, yn), "atimes_dev_noalloc")
    assert relerr(yn.to_host(), yd.to_host()) <= 1e-12
    s.close()
    # a scratch need beyond the 64 MiB array: N = 8192 FTS needs 92 MB of
    # real-space chunk partials alone
    _say("noalloc refusal of N=8192 FTS")
    s, xd, yd, yn = _case(2, 8192, 0.1, ls.MIXED, 12)
    s.atimes_dev(xd, yd)
    rc = _noalloc(s, xd, yn)
    assert rc == -3
    assert b"scratch" in ls.lib().stokes_hip_last_error()
    # the ordinary product is unaffected
    y0 = yd.to_host()
    s.atimes_dev(xd, yd)
    assert relerr(yd.to_host(), y0) <= 1e-12
    s.close()
