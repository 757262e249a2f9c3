"""The dense lubrication matrix and the lubricated FTS solves on the GPU
(FP64): make_matrix_lub_3f / 3ft / 3fts and stokes_hip_make_matrix_lub_dev
(lub-matrix.c:76-507); solve_res_lub_3fts(_0) (ewald-3fts.c:676-770) and
solve_mob_lub_3fts(_0) (ewald-3fts.c:776-1000) by GMRES on the device; their
dense forms solve_res_lub_3fts_matrix(_0) (ewald-3fts-matrix.c:264-380) and
solve_mob_lub_3fts_matrix(_0) (ewald-3fts-matrix.c:731-880).

Oracle: the linear maps of the oracle's operators: M, the Ewald mobility
(the oracle's extracted-form matrix times T), and L, calc_lub_3fts applied
column by column (the compiled reference's matrix_lub_*_2b pair routines
times T equal it, tests/test_oracle_lub.py).  Resistance: (F, T, S) =
M^-1 (I + M L) (U, O, E); mobility: the unknowns (U, O, S) of
M (F, T, S) = (I + M L) (U, O, E); both by numpy's dense solves.

Bars: the device matrix within 1e-12 of the oracle's map (normwise); dense
device solves within 1e-10 of numpy's; GMRES (eps 1e-12) with a residual
below eps and within 1e-9; the labo-frame host wrappers equal the
fluid-rest device solves with the imposed-flow shifts written out in numpy
to 1e-12.  Equal spheres (periodic, open, a box of 2.6 radii where every
sphere lubricates with its own images) and unequal spheres (radii on
[0.5, 2]); N = 256 dense against GMRES.
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


def _case(version, case):
    a = None
    if case == "self_image":
        pos = np.array([[0.3, 0.2, 0.1], [1.55, 1.4, 1.35]])
        L = 2.6
    elif case == "poly":
        pos, a, L = workloads.configuration_poly(12, 0.25, seed=5, gap=1.01,
                                                 a_lo=0.5, a_hi=2.0)
    else:
        pos, L = workloads.configuration(16, 0.2, seed=3, gap=2.02)
    kw = {} if case == "open" else {"box": (L, L, L), "tr": 1.0}
    s = ls.Stokes(version, pos, a=a, **kw)
    o = po.System(version, pos, a=a, **kw)
    s.set_lub(4.0 + 1e-12, 3.0, 100, 1)
    return pos, s, o


def _maps(o, npart):
    """(M, L): the oracle's linear maps"""
    M = po.trans_mat_ext2ext(npart, o.matrix())
    return M, o.lub_columns()


def _mob_restated(M, Lm, b, npart):
    """unknowns (U, O, S) of M g = (I + M L) v with g = (F, T, S) and
    v = (U, O, E), b = (F, T, E) per particle"""
    n = 11 * npart
    P = np.eye(n) + M @ Lm
    q = np.arange(n) % 11
    K, H = np.flatnonzero(q < 6), np.flatnonzero(q >= 6)
    A = np.concatenate([M[:, H], -P[:, K]], axis=1)
    rhs = P[:, H] @ b[H] - M[:, K] @ b[K]
    z = np.linalg.solve(A, rhs)
    x = np.empty(n)
    x[H] = z[:H.size]
    x[K] = z[H.size:]
    return x


@pytest.mark.parametrize("version,case", [(2, "periodic"), (2, "open"),
                                          (2, "poly"), (2, "self_image"),
                                          (1, "periodic"), (0, "poly")])
def test_lub_matrix(version, case):
    pos, s, o = _case(version, case)
    m = s.lub_matrix()
    got = po.ext_to_map(version, len(pos), m)
    want = o.lub_columns()
    e = relerr(got, want)
    _say(f"lub matrix v{version} {case}: rel err {e:.2e}")
    assert np.linalg.norm(want) > 0 and e <= 1e-12
    # the drop-in host-pointer form
    name = ("make_matrix_lub_3f", "make_matrix_lub_3ft",
            "make_matrix_lub_3fts")[version]
    h = np.zeros((s.n, s.n))
    getattr(s.L, name)(s.sys, h.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(h, m)
    s.close()


@pytest.mark.parametrize("case", ["periodic", "poly", "open"])
def test_solve_res_lub(case):
    pos, s, o = _case(2, case)
    npart = len(pos)
    M, Lm = _maps(o, npart)
    b = np.random.default_rng(11).standard_normal(s.n)
    want = np.linalg.solve(M, b + M @ (Lm @ b))
    xd, _, _ = s.solve_lub(b, "res", matrix=True)
    ed = relerr(xd, want)
    s.set_iter(max_it=2000, restart=40, eps=1e-12)
    xg, it, rr = s.solve_lub(b, "res")
    eg = relerr(xg, want)
    _say(f"res lub {case}: dense {ed:.2e}, GMRES {it} it residual {rr:.2e} "
         f"err {eg:.2e}")
    assert ed <= 1e-10
    assert rr < 1e-12 and eg <= 1e-9
    s.close()


@pytest.mark.parametrize("case", ["periodic", "poly", "open"])
def test_solve_mob_lub(case):
    pos, s, o = _case(2, case)
    npart = len(pos)
    M, Lm = _maps(o, npart)
    b = np.random.default_rng(12).standard_normal(s.n)
    want = _mob_restated(M, Lm, b, npart)
    xd, _, _ = s.solve_lub(b, "mob", matrix=True)
    ed = relerr(xd, want)
    s.set_iter(max_it=2000, restart=40, eps=1e-12)
    xg, it, rr = s.solve_lub(b, "mob")
    eg = relerr(xg, want)
    _say(f"mob lub {case}: dense {ed:.2e}, GMRES {it} it residual {rr:.2e} "
         f"err {eg:.2e}")
    assert ed <= 1e-10
    assert rr < 1e-12 and eg <= 1e-9
    # the defining property with the oracle's operators, matrix-free
    xr, br = xd.reshape(npart, 11), b.reshape(npart, 11)
    g = np.concatenate([br[:, :6], xr[:, 6:]], 1).reshape(-1)
    v = np.concatenate([xr[:, :6], br[:, 6:]], 1).reshape(-1)
    lhs = o.atimes(g)
    rhs = v + o.atimes(o.lub(v))
    assert relerr(lhs, rhs) <= 1e-10
    s.close()


def test_dense_against_gmres_n256():
    n = 256
    pos, L = workloads.configuration(n, 0.3, seed=17, gap=2.01)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    s.set_lub(4.0 + 1e-12, 3.0, 100, 1)
    s.set_iter(max_it=4000, restart=60, eps=1e-12)
    rng = np.random.default_rng(13)
    for kind in ("res", "mob"):
        b = rng.standard_normal(s.n)
        xd, _, _ = s.solve_lub(b, kind, matrix=True)
        xg, it, rr = s.solve_lub(b, kind)
        e = relerr(xg, xd)
        _say(f"N={n} {kind} lub: GMRES {it} it residual {rr:.2e}, "
             f"vs dense {e:.2e}")
        assert rr < 1e-12 and e <= 1e-8
    s.close()


def _imposed_u(pos, Ui, Oi, Ei):
    E = np.array([[Ei[0], Ei[1], Ei[2]], [Ei[1], Ei[4], Ei[3]],
                  [Ei[2], Ei[3], -Ei[0] - Ei[4]]])
    return Ui + np.cross(Oi, pos) + pos @ E.T


@pytest.mark.parametrize("matrix", [False, True])
def test_labo_frame_wrappers(matrix):
    """solve_res_lub_3fts(_matrix) and solve_mob_lub_3fts(_matrix) (labo
    frame, host arrays) equal the device fluid-rest solves with the shifts
    written out in numpy; the _0 forms equal the device solves"""
    n = 16
    pos, L = workloads.configuration(n, 0.2, seed=19, gap=2.02)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    s.set_lub(4.0 + 1e-12, 3.0, 100, 1)
    s.set_iter(max_it=2000, restart=40, eps=1e-13)
    Ui, Oi = np.array([0.3, -0.2, 0.1]), np.array([0.05, 0.1, -0.07])
    Ei = np.array([0.2, 0.05, -0.1, 0.03, -0.15])
    s.L.stokes_set_Ui(s.sys, *Ui)
    s.L.stokes_set_Oi(s.sys, *Oi)
    s.L.stokes_set_Ei(s.sys, *Ei)
    rng = np.random.default_rng(23)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    pk = lambda a, b, c: np.concatenate([a.reshape(n, 3), b.reshape(n, 3),
                                         c.reshape(n, 5)], 1).reshape(-1)
    sfx = "_matrix" if matrix else ""
    # resistance
    u, om, e = (rng.standard_normal(n * k) for k in (3, 3, 5))
    f, t, st = np.zeros(3 * n), np.zeros(3 * n), np.zeros(5 * n)
    getattr(s.L, "solve_res_lub_3fts" + sfx)(s.sys, P(u), P(om), P(e), P(f),
                                             P(t), P(st))
    b0 = pk(u - _imposed_u(pos, Ui, Oi, Ei).reshape(-1), om - np.tile(Oi, n),
            e - np.tile(Ei, n))
    x, _, _ = s.solve_lub(b0, "res", matrix=matrix)
    assert relerr(pk(f, t, st), x) < 1e-12
    f0, t0, s0 = np.zeros(3 * n), np.zeros(3 * n), np.zeros(5 * n)
    xr = b0.reshape(n, 11)
    getattr(s.L, "solve_res_lub_3fts" + sfx + "_0")(
        s.sys, P(np.ascontiguousarray(xr[:, :3])),
        P(np.ascontiguousarray(xr[:, 3:6])),
        P(np.ascontiguousarray(xr[:, 6:])), P(f0), P(t0), P(s0))
    assert np.array_equal(pk(f0, t0, s0), x)
    # mobility
    f, t, e = (rng.standard_normal(n * k) for k in (3, 3, 5))
    u, om, st = np.zeros(3 * n), np.zeros(3 * n), np.zeros(5 * n)
    getattr(s.L, "solve_mob_lub_3fts" + sfx)(s.sys, P(f), P(t), P(e), P(u),
                                             P(om), P(st))
    x, _, _ = s.solve_lub(pk(f, t, e - np.tile(Ei, n)), "mob", matrix=matrix)
    xr = x.reshape(n, 11)
    want = np.concatenate([xr[:, :3] + _imposed_u(pos, Ui, Oi, Ei),
                           xr[:, 3:6] + Oi, xr[:, 6:]], 1)
    err = relerr(pk(u, om, st), want.reshape(-1))
    _say(f"labo-frame lubricated wrappers matrix={matrix}: mob {err:.2e}")
    assert err < 1e-12
    s.close()
