"""The library's host exports of the lubrication API (include/twobody.h,
include/lub.h) against the compiled reference (oracle/_ref) and, where the
reference needs the absent LAPACK (the unequal-sphere far-field inverse),
against the oracle, which is pinned in test_oracle_lub.py.  Host code only
(no GPU): the same templates (lub_math.h) run in the device kernels.
"""
import ctypes as C

import numpy as np
import pytest

import libstokes_amd as ls
import pyoracle as po

pytestmark = pytest.mark.skipif(not po.ref_available(),
                                reason="oracle/_ref not built")
_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


def _lib():
    L = ls.lib()
    L.twobody_scalars_res.argtypes = [C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_void_p, C.c_int,
                                      C.c_int, C.c_int, _dp]
    for f in ("scalar_two_body_res", "scalar_minv_f", "scalar_minv_ft",
              "scalar_minv_fts"):
        getattr(L, f).argtypes = [C.c_double, _dp]
    for f in ("scalars_minv_f_poly", "scalars_minv_ft_poly",
              "scalars_minv_fts_poly"):
        getattr(L, f).argtypes = [C.c_double] * 3 + [_dp]
    L.scalars_lub_poly_full.argtypes = [C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_void_p, C.c_void_p,
                                        C.c_int, C.c_int, _dp]
    return L


def test_tables():
    L = _lib()
    worst = 0.0
    for k, name in enumerate(po.TWOBODY_FAMILIES):
        for lam in (0.25, 0.5, 1.0, 2.0, 4.0):
            a = np.zeros(101)
            getattr(L, "twobody_" + name)(100, C.c_double(lam), _p(a))
            b = po.twobody_table(k, 100, lam, reference=True)
            nz = b != 0.0
            assert np.all(a[~nz] == 0.0)
            worst = max(worst, float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz]))))
    assert worst < 1e-12, worst


def test_resistance_scalars():
    L = _lib()
    for aa, ab in ((1.0, 1.0), (1.0, 2.0), (2.0, 1.0), (0.5, 1.7)):
        for s in (2.0005, 2.05, 2.4, 3.0, 4.0):
            r = 0.5 * s * (aa + ab)
            for flag_lub in (0, 1):
                for flag_scale in (0, 1, 2):
                    for v in (0, 1, 2):
                        a = np.zeros(22)
                        L.twobody_scalars_res(v, r, aa, ab, None, 100,
                                              flag_lub, flag_scale, _p(a))
                        b = po.twobody_res(v, r, aa, ab, 100, flag_lub,
                                           flag_scale, reference=True)
                        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12)) < 1e-12


def test_equal_sphere_series_and_inverses():
    L = _lib()
    for s in (2.0001, 2.01, 2.5, 4.0):
        a = np.zeros(22)
        L.scalar_two_body_res(s, _p(a))
        b = po.two_body_res(s, reference=True)
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12)) < 1e-12
        for f, n, v in (("scalar_minv_f", 4, 0), ("scalar_minv_ft", 10, 1),
                        ("scalar_minv_fts", 22, 2)):
            a = np.zeros(n)
            getattr(L, f)(s, _p(a))
            w = po.minv_poly(v, s, 1.0, 1.0)
            want = np.array([w[4 * (k // 2) + (k % 2)] for k in range(n)])
            assert np.max(np.abs(a - want)) < 1e-13 * np.abs(want).max()
        a = np.zeros(22)
        L.scalar_minv_fts(s, _p(a))
        b = po.minv_fts(s, reference=True)
        assert np.max(np.abs(a - b)) < 1e-13 * np.abs(b).max()


def test_unequal_sphere_inverse_and_correction():
    L = _lib()
    for aa, ab, r in ((1.0, 2.0, 3.3), (0.5, 1.7, 2.4), (2.0, 0.6, 2.9)):
        for f, n, v in (("scalars_minv_f_poly", 8, 0),
                        ("scalars_minv_ft_poly", 20, 1),
                        ("scalars_minv_fts_poly", 44, 2)):
            a = np.zeros(n)
            getattr(L, f)(r, aa, ab, _p(a))
            b = po.minv_poly(v, r, aa, ab)[:n]
            assert np.max(np.abs(a - b)) < 1e-13 * np.abs(b).max()
        # scalars_lub_poly_full = res(12), res(21) - the far-field inverse
        lub = np.zeros(44)
        L.scalars_lub_poly_full(2, r, aa, ab, None, None, 100, 1, _p(lub))
        r12 = po.twobody_res(2, r, aa, ab, 100, 1, 1, reference=True)
        r21 = po.twobody_res(2, r, ab, aa, 100, 1, 1, reference=True)
        inf = po.minv_poly(2, r, aa, ab)
        want = np.zeros(44)
        want[0::4] = r12[0::2] - inf[0::4]
        want[1::4] = r12[1::2] - inf[1::4]
        want[2::4] = r21[1::2] - inf[2::4]
        want[3::4] = r21[0::2] - inf[3::4]
        # the correction is a difference of two close terms: the bar is
        # relative to the terms
        terms = np.zeros(44)
        terms[0::4] = np.maximum(np.abs(r12[0::2]), np.abs(inf[0::4]))
        terms[1::4] = np.maximum(np.abs(r12[1::2]), np.abs(inf[1::4]))
        terms[2::4] = np.maximum(np.abs(r21[1::2]), np.abs(inf[2::4]))
        terms[3::4] = np.maximum(np.abs(r21[0::2]), np.abs(inf[3::4]))
        assert np.max(np.abs(lub - want) / np.maximum(terms, 1e-10)) < 1e-12
