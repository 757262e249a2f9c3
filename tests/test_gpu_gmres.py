"""FTS natural mobility problem by GMRES on the device
(stokes_hip_solve_mob_3fts_dev, solve_mob_3fts_0) against the dense solve.

Unknowns (U, O, S) from known (F, T, E) for every particle; the same
problem the dense path (solve_mix with nm = np) solves by LU, and the numpy
restatement of ewald-3fts-matrix.c solves by explicit inversion.

Bars: relative residual below the requested eps (reported by the solver
from a fresh operator application); solution within 1e-10 of the
restatement at N = 64 (eps 1e-13) and within 1e-9 of the device LU at
N = 256 (eps 1e-12); MIXED operator with eps 1e-6 within 1e-5 of FP64.
"""
import ctypes as C

import numpy as np
import pytest

import libstokes_amd as ls
from libstokes_amd import workloads
import pyoracle as po
from conftest import relerr

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_gmres_matches_restatement_n64():
    n = 64
    pos, L = workloads.configuration(n, 0.1, seed=21)
    o = po.System(2, pos, box=(L, L, L), tr=1.0)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    b = np.random.default_rng(3).standard_normal(11 * n)
    s.set_iter(max_it=500, restart=30, eps=1e-13)
    x, it, rr = s.solve_mob_gmres(b)
    want = po.solve_mix_dense(o, n, b)
    e = relerr(x, want)
    _say(f"  GMRES N={n}: {it} iterations, residual {rr:.2e}, "
         f"vs restatement {e:.2e}")
    assert rr < 1e-13
    assert e < 1e-10
    # defining property with the oracle's operator: M (F T S) = (U O E)
    xb = x.reshape(n, 11)
    bb = b.reshape(n, 11)
    g = np.concatenate([bb[:, :6], xb[:, 6:]], axis=1).reshape(-1)
    v = np.concatenate([xb[:, :6], bb[:, 6:]], axis=1).reshape(-1)
    res = np.abs(o.atimes(g) - v).max() / np.abs(v).max()
    assert res < 1e-11
    s.close()


def test_gmres_matches_device_lu_n256_and_splitting():
    n = 256
    pos, L = workloads.configuration(n, 0.2, seed=4)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    b = np.random.default_rng(8).standard_normal(11 * n)
    want = s.solve_mix(b, nm=n)
    s.set_iter(max_it=2000, restart=40, eps=1e-12)
    x, it, rr = s.solve_mob_gmres(b)
    e = relerr(x, want)
    _say(f"  GMRES N={n} tr=1: {it} iterations, residual {rr:.2e}, "
         f"vs LU {e:.2e}")
    assert rr < 1e-12
    assert e < 1e-9
    s.set_split(1000.0)
    x2, it2, rr2 = s.solve_mob_gmres(b)
    e2 = relerr(x2, want)
    _say(f"  GMRES N={n} tr=1000: {it2} iterations, vs LU {e2:.2e}")
    assert e2 < 1e-9
    s.close()


def test_gmres_mixed_precision_operator():
    n = 256
    pos, L = workloads.configuration(n, 0.2, seed=5)
    s64 = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    b = np.random.default_rng(9).standard_normal(11 * n)
    want = s64.solve_mix(b, nm=n)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, precision=ls.MIXED)
    s.set_split(1000.0)
    s.set_iter(max_it=500, restart=30, eps=1e-6)
    x, it, rr = s.solve_mob_gmres(b)
    e = relerr(x, want)
    _say(f"  GMRES mixed N={n}: {it} iterations, residual {rr:.2e}, "
         f"vs FP64 LU {e:.2e}")
    assert rr < 1e-6
    assert e < 1e-5
    s.close()
    s64.close()


def test_reference_entry_point_solve_mob_3fts_0():
    n = 32
    pos, L = workloads.configuration(n, 0.15, seed=6)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    rng = np.random.default_rng(4)
    f, t, e = (rng.standard_normal(n * k) for k in (3, 3, 5))
    b = np.concatenate([f.reshape(n, 3), t.reshape(n, 3), e.reshape(n, 5)],
                       axis=1).reshape(-1)
    want = s.solve_mix(b, nm=n).reshape(n, 11)
    s.set_iter(max_it=1000, restart=20, eps=1e-12)
    u, om, st = np.zeros(3 * n), np.zeros(3 * n), np.zeros(5 * n)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    # iter NULL: the parameters kept by stokes_set_iter
    s.L.solve_mob_3fts_0(s.sys, None, P(f), P(t), P(e), P(u), P(om), P(st))
    got = np.concatenate([u.reshape(n, 3), om.reshape(n, 3),
                          st.reshape(n, 5)], axis=1)
    err = relerr(got, want)
    _say(f"  solve_mob_3fts_0 N={n} vs LU {err:.2e}")
    assert err < 1e-9
    # zero right-hand side: zero solution, no iterations
    x, it, rr = s.solve_mob_gmres(np.zeros(11 * n))
    assert it == 0 and not np.any(x)
    s.close()
