"""Dense FP64 mixed-problem solves on the GPU (stokes_hip_solve_mix_dev,
solve_mob_3fts_matrix_0, solve_mix_3fts_matrix_0) against the numpy
restatement of ewald-3fts-matrix.c:570-652 / 1192-1290 and against the
defining property v = M g checked with the oracle's operator.

Bars: the solution within 1e-10 (normwise) of the restatement, which
inverts explicitly where the device factors; residual of v = M g within
1e-11 of max |v|.
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


def _check(version, n, nm, phi=0.1, seed=1):
    pos, L = workloads.configuration(n, phi, seed=seed)
    o = po.System(version, pos, box=(L, L, L), tr=1.0)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0)
    d = ls.dim(version)
    b = np.random.default_rng(seed).standard_normal(n * d)
    x = s.solve_mix(b, nm=nm)
    want = po.solve_mix_dense(o, nm, b)
    e = relerr(x, want)
    isH = np.array([i >= nm or q >= 6 for i in range(n) for q in range(d)])
    g = np.where(isH, x, b)
    v = np.where(isH, b, x)
    res = np.abs(o.atimes(g) - v).max() / np.abs(v).max()
    _say(f"  version {version} N={n} nm={nm}: vs restatement {e:.2e}, "
         f"residual {res:.2e}")
    assert e < 1e-10
    assert res < 1e-11
    s.close()


@pytest.mark.parametrize("version,n,nm", [(2, 27, 27), (2, 27, 20),
                                          (1, 27, 22), (0, 27, 20),
                                          (0, 27, 27)])
def test_solve_mix_matches_restatement(version, n, nm):
    _check(version, n, nm)


def test_solve_mob_fts_n256():
    _check(2, 256, 256, phi=0.2, seed=3)


def test_reference_void_entry_points():
    n, nm = 16, 12
    pos, L = workloads.configuration(n, 0.1, seed=5)
    o = po.System(2, pos, box=(L, L, L), tr=1.0)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    rng = np.random.default_rng(2)
    f, t, e = (rng.standard_normal(n * k) for k in (3, 3, 5))
    u, om, st = np.zeros(3 * n), np.zeros(3 * n), np.zeros(5 * n)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    s.s.nm = n
    s.L.solve_mob_3fts_matrix_0(s.sys, P(f), P(t), P(e), P(u), P(om), P(st))
    b = np.concatenate([f.reshape(n, 3), t.reshape(n, 3), e.reshape(n, 5)],
                       axis=1).reshape(-1)
    want = po.solve_mix_dense(o, n, b).reshape(n, 11)
    got = np.concatenate([u.reshape(n, 3), om.reshape(n, 3),
                          st.reshape(n, 5)], axis=1)
    assert relerr(got, want) < 1e-10
    # fixed particles [nm, n)
    nf = n - nm
    s.s.nm = nm
    uf, of_, ef = (rng.standard_normal(nf * k) for k in (3, 3, 5))
    ff, tf, sf = np.zeros(3 * nf), np.zeros(3 * nf), np.zeros(5 * nf)
    u, om, st = np.zeros(3 * nm), np.zeros(3 * nm), np.zeros(5 * nm)
    s.L.solve_mix_3fts_matrix_0(s.sys, P(f[:3 * nm]), P(t[:3 * nm]),
                                P(e[:5 * nm]), P(uf), P(of_), P(ef), P(u),
                                P(om), P(st), P(ff), P(tf), P(sf))
    b = np.concatenate([
        np.concatenate([f[:3 * nm].reshape(nm, 3), t[:3 * nm].reshape(nm, 3),
                        e[:5 * nm].reshape(nm, 5)], axis=1).reshape(-1),
        np.concatenate([uf.reshape(nf, 3), of_.reshape(nf, 3),
                        ef.reshape(nf, 5)], axis=1).reshape(-1)])
    want = po.solve_mix_dense(o, nm, b).reshape(n, 11)
    got = np.concatenate([
        np.concatenate([u.reshape(nm, 3), om.reshape(nm, 3),
                        st.reshape(nm, 5)], axis=1),
        np.concatenate([ff.reshape(nf, 3), tf.reshape(nf, 3),
                        sf.reshape(nf, 5)], axis=1)])
    assert relerr(got, want) < 1e-10
    s.close()
