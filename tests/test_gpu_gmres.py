"""F, FT and FTS problems by GMRES on the device (ewald-3f.c, ewald-3ft.c,
ewald-3fts.c): for FTS the natural mobility problem
(stokes_hip_solve_mob_3fts_dev, solve_mob_3fts_0), the mixed problem with
fixed particles (stokes_hip_solve_mix_3fts_dev, solve_mix_3fts_0), the
resistance problem (stokes_hip_solve_res_3fts_dev, solve_res_3fts_0) and the
labo-frame wrappers; for F and FT the reference entry points solve_res_3f(t)
(_0), solve_mob_3f(t) and solve_mix_3f(t) against numpy's dense solves of the
oracle's operator matrix.

Mobility / mixed: the same problems the dense path (solve_mix) solves by LU
and the numpy restatement of ewald-3fts-matrix.c by explicit inversion.
Resistance: numpy's dense solve of the oracle's operator assembled column
by column (N = 16) and the defining property M x = b with the oracle's
operator (N = 64).

Bars: relative residual below the requested eps (reported by the solver
from a fresh operator application); solutions within 1e-10 of the
restatements (eps 1e-13) and within 1e-9 of the device LU at N = 256
(eps 1e-12); MIXED operator with eps 1e-6 within 1e-5 of FP64; the
labo-frame wrappers equal the fluid-rest solves with the shifts written
out in numpy to 1e-12.
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


def _operator_matrix(o, n):
    """the oracle's FTS operator as a dense matrix, column by column"""
    m = np.empty((11 * n, 11 * n))
    for c in range(11 * n):
        e = np.zeros(11 * n)
        e[c] = 1.0
        m[:, c] = o.atimes(e)
    return m


def test_gmres_resistance_problem():
    """M (F, T, S) = (U, O, E): against numpy's dense solve of the oracle's
    operator (N = 16), and the defining property at N = 64"""
    n = 16
    pos, L = workloads.configuration(n, 0.1, seed=31)
    o = po.System(2, pos, box=(L, L, L), tr=1.0)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    b = np.random.default_rng(6).standard_normal(11 * n)
    s.set_iter(max_it=1000, restart=40, eps=1e-13)
    x, it, rr = s.solve_gmres(b, kind="res")
    want = np.linalg.solve(_operator_matrix(o, n), b)
    e = relerr(x, want)
    _say(f"  resistance N={n}: {it} iterations, residual {rr:.2e}, "
         f"vs dense solve {e:.2e}")
    assert rr < 1e-13
    assert e < 1e-10
    s.close()
    n = 64
    pos, L = workloads.configuration(n, 0.15, seed=32)
    o = po.System(2, pos, box=(L, L, L), tr=1.0)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    b = np.random.default_rng(7).standard_normal(11 * n)
    s.set_iter(max_it=1000, restart=40, eps=1e-13)
    x, it, rr = s.solve_gmres(b, kind="res")
    res = relerr(o.atimes(x), b)
    _say(f"  resistance N={n}: {it} iterations, |M x - b| / |b| (oracle) "
         f"{res:.2e}")
    assert res < 1e-11
    s.close()


def test_gmres_mixed_problem_with_fixed_particles():
    n, nm = 64, 48
    pos, L = workloads.configuration(n, 0.1, seed=33)
    o = po.System(2, pos, box=(L, L, L), tr=1.0)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    b = np.random.default_rng(8).standard_normal(11 * n)
    want = po.solve_mix_dense(o, nm, b)
    lu = s.solve_mix(b, nm=nm)
    s.set_iter(max_it=1000, restart=30, eps=1e-13)
    x, it, rr = s.solve_gmres(b, kind="mix", nm=nm)
    e = relerr(x, want)
    _say(f"  mixed N={n} nm={nm}: {it} iterations, residual {rr:.2e}, "
         f"vs restatement {e:.2e}, vs device LU {relerr(x, lu):.2e}")
    assert rr < 1e-13
    assert e < 1e-10
    assert relerr(x, lu) < 1e-10
    s.close()


def _imposed_u(pos, Ui, Oi, Ei):
    E = np.array([[Ei[0], Ei[1], Ei[2]], [Ei[1], Ei[4], Ei[3]],
                  [Ei[2], Ei[3], -Ei[0] - Ei[4]]])
    return Ui + np.cross(Oi, pos) + pos @ E.T


def test_labo_frame_wrappers():
    """solve_res_3fts / solve_mob_3fts / solve_mix_3fts (labo frame) equal the
    fluid-rest solve with the imposed-flow shifts written out in numpy; the
    mixed one shifts the fixed particles with the positions of particles
    0..nf-1, as ewald-3fts.c:652 does"""
    n, nm = 24, 18
    nf = n - nm
    pos, L = workloads.configuration(n, 0.1, seed=34)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    s.set_iter(max_it=1000, restart=30, eps=1e-13)
    Ui, Oi = np.array([0.3, -0.2, 0.1]), np.array([0.05, 0.1, -0.07])
    Ei = np.array([0.2, 0.05, -0.1, 0.03, -0.15])
    s.L.stokes_set_Ui(s.sys, *Ui)
    s.L.stokes_set_Oi(s.sys, *Oi)
    s.L.stokes_set_Ei(s.sys, *Ei)
    rng = np.random.default_rng(9)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    pk = lambda a, b, c, k: np.concatenate([a.reshape(k, 3), b.reshape(k, 3),
                                            c.reshape(k, 5)], 1).reshape(-1)
    # resistance, labo frame
    u, om, e = (rng.standard_normal(n * k) for k in (3, 3, 5))
    f, t, st = np.zeros(3 * n), np.zeros(3 * n), np.zeros(5 * n)
    s.L.solve_res_3fts(s.sys, P(u), P(om), P(e), P(f), P(t), P(st))
    u0 = u - _imposed_u(pos, Ui, Oi, Ei).reshape(-1)
    o0 = om - np.tile(Oi, n)
    e0 = e - np.tile(Ei, n)
    x, _, _ = s.solve_gmres(pk(u0, o0, e0, n), kind="res")
    assert relerr(pk(f, t, st, n), x) < 1e-12
    # mobility, labo frame
    f, t, e = (rng.standard_normal(n * k) for k in (3, 3, 5))
    u, om, st = np.zeros(3 * n), np.zeros(3 * n), np.zeros(5 * n)
    s.s.nm = n
    s.L.solve_mob_3fts(s.sys, P(f), P(t), P(e), P(u), P(om), P(st))
    x, _, _ = s.solve_gmres(pk(f, t, e - np.tile(Ei, n), n), kind="mix", nm=n)
    xr = x.reshape(n, 11)
    u_want = xr[:, :3] + _imposed_u(pos, Ui, Oi, Ei)
    o_want = xr[:, 3:6] + Oi
    got = np.concatenate([u.reshape(n, 3), om.reshape(n, 3), st.reshape(n, 5)], 1)
    want = np.concatenate([u_want, o_want, xr[:, 6:]], 1)
    assert relerr(got, want) < 1e-12
    # mixed, labo frame
    s.s.nm = nm
    f, t, e = (rng.standard_normal(nm * k) for k in (3, 3, 5))
    uf, of, ef = (rng.standard_normal(nf * k) for k in (3, 3, 5))
    u, om, st = np.zeros(3 * nm), np.zeros(3 * nm), np.zeros(5 * nm)
    ff, tf, sf = np.zeros(3 * nf), np.zeros(3 * nf), np.zeros(5 * nf)
    s.L.solve_mix_3fts(s.sys, P(f), P(t), P(e), P(uf), P(of), P(ef), P(u),
                       P(om), P(st), P(ff), P(tf), P(sf))
    uf0 = uf - _imposed_u(pos[:nf], Ui, Oi, Ei).reshape(-1)
    b = np.concatenate([pk(f, t, e - np.tile(Ei, nm), nm),
                        pk(uf0, of - np.tile(Oi, nf), ef - np.tile(Ei, nf), nf)])
    x, _, _ = s.solve_gmres(b, kind="mix", nm=nm)
    xr = x.reshape(n, 11)
    want = np.concatenate([
        (np.concatenate([xr[:nm, :3] + _imposed_u(pos[:nm], Ui, Oi, Ei),
                         xr[:nm, 3:6] + Oi, xr[:nm, 6:]], 1)).reshape(-1),
        xr[nm:].reshape(-1)])
    got = np.concatenate([pk(u, om, st, nm), pk(ff, tf, sf, nf)])
    err = relerr(got, want)
    _say(f"  labo-frame wrappers: mixed {err:.2e}")
    assert err < 1e-12
    s.close()


@pytest.mark.parametrize("version", [0, 1])
def test_f_ft_entry_points(version):
    """solve_res_3f(t)(_0), solve_mob_3f(t) and solve_mix_3f(t) (ewald-3f.c,
    ewald-3ft.c) against numpy's dense solves of the oracle's operator
    matrix, in the labo frame with an imposed flow; the mixed problem holds
    the fixed particles' velocities and shifts them with the positions of
    particles 0..nf-1 (ewald-3f.c:301, ewald-3ft.c:359-360)"""
    n, nm = 32, 24
    nf, d = n - nm, (3 if version == 0 else 6)
    _say(f"F / FT entry points version={version} N={n} nm={nm}")
    pos, L = workloads.configuration(n, 0.1, seed=35 + version)
    o = po.System(version, pos, box=(L, L, L), tr=1.0)
    M = o.matrix()
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0)
    s.set_iter(max_it=1000, restart=30, eps=1e-13)
    Ui, Oi = np.array([0.3, -0.2, 0.1]), np.array([0.05, 0.1, -0.07])
    Ei = np.array([0.2, 0.05, -0.1, 0.03, -0.15])
    s.L.stokes_set_Ui(s.sys, *Ui)
    s.L.stokes_set_Oi(s.sys, *Oi)
    s.L.stokes_set_Ei(s.sys, *Ei)
    rng = np.random.default_rng(10 + version)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))

    def pk(a, b, k):        # (a_i, b_i) per particle, or a alone for F
        if version == 0:
            return a.copy()
        return np.concatenate([a.reshape(k, 3), b.reshape(k, 3)], 1).reshape(-1)

    def imp(k):             # imposed (U, O) of particles 0..k-1
        return pk(_imposed_u(pos[:k], Ui, Oi, Ei).reshape(-1), np.tile(Oi, k), k)

    u, om = rng.standard_normal(3 * n), rng.standard_normal(3 * n)
    # resistance: fluid-rest frame and labo frame
    f, t = np.zeros(3 * n), np.zeros(3 * n)
    if version == 0:
        s.L.solve_res_3f_0(s.sys, P(u), P(f))
    else:
        s.L.solve_res_3ft_0(s.sys, P(u), P(om), P(f), P(t))
    want = np.linalg.solve(M, pk(u, om, n))
    e0 = relerr(pk(f, t, n), want)
    if version == 0:
        s.L.solve_res_3f(s.sys, P(u), P(f))
    else:
        s.L.solve_res_3ft(s.sys, P(u), P(om), P(f), P(t))
    want = np.linalg.solve(M, pk(u, om, n) - imp(n))
    e1 = relerr(pk(f, t, n), want)
    # mobility, labo frame
    f, t = rng.standard_normal(3 * n), rng.standard_normal(3 * n)
    uu, oo = np.zeros(3 * n), np.zeros(3 * n)
    if version == 0:
        s.L.solve_mob_3f(s.sys, P(f), P(uu))
    else:
        s.L.solve_mob_3ft(s.sys, P(f), P(t), P(uu), P(oo))
    e2 = relerr(pk(uu, oo, n), M @ pk(f, t, n) + imp(n))
    # mixed, labo frame
    s.s.nm = nm
    f, t = rng.standard_normal(3 * nm), rng.standard_normal(3 * nm)
    uf, of = rng.standard_normal(3 * nf), rng.standard_normal(3 * nf)
    uu, oo = np.zeros(3 * nm), np.zeros(3 * nm)
    ff, tf = np.zeros(3 * nf), np.zeros(3 * nf)
    if version == 0:
        s.L.solve_mix_3f(s.sys, P(f), P(uf), P(uu), P(ff))
    else:
        s.L.solve_mix_3ft(s.sys, P(f), P(t), P(uf), P(of), P(uu), P(oo),
                          P(ff), P(tf))
    m = d * nm
    fm = pk(f, t, nm)
    uf0 = pk(uf, of, nf) - imp(nf)
    Ff = np.linalg.solve(M[m:, m:], uf0 - M[m:, :m] @ fm)
    um = M[:m, :m] @ fm + M[:m, m:] @ Ff + imp(nm)
    e3 = relerr(np.concatenate([pk(uu, oo, nm), pk(ff, tf, nf)]),
                np.concatenate([um, Ff]))
    _say(f"  res_0 {e0:.2e}, res {e1:.2e}, mob {e2:.2e}, mix {e3:.2e}")
    assert max(e0, e1, e3) < 1e-10
    assert e2 < 1e-12
    s.close()
