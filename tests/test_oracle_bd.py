"""Pins of the oracle's Brownian-dynamics step (F version, mid-point
scheme, brownian.c:2352-2560) and of its overlap count (brownian.c:299-375).

* the overlap count against a brute-force enumeration of the 27 cells;
* one sphere in a periodic box, where M = mu I: the step must equal
  x + dt (mu F + fact sqrt(mu) y) with mu the self-mobility (Hasimoto's
  closed form to 1e-6, the operator's own value to 1e-10);
* eight spheres: the step against the same composition built from the
  compiled reference's own chebyshev_eval_atimes and atimes_3all;
* the overlap branches: halving dt keeps the draw, a redraw below dt_lim
  takes the next draw and restores dt, with the reference's stale fact.
"""
import math

import numpy as np
import pytest

import pyoracle as po


def _brute_overlap(pos, L, rmin, periodic=True, shear_mode=0, shift=0.0):
    pos = pos.reshape(-1, 3)
    n = len(pos)
    cnt = 0
    cells = [(0, 0, 0)]
    if periodic:
        cells = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1)
                 for c in (-1, 0, 1)]
    for i in range(n):
        for j in range(i, n):
            d = pos[i] - pos[j]
            for (a, b, c) in cells:
                if i == j and (a, b, c) == (0, 0, 0):
                    continue
                v = d + np.array([a * L, b * L, c * L])
                if shear_mode == 1:
                    v[0] += b * shift
                elif shear_mode == 2:
                    v[0] += c * shift
                cnt += v @ v < rmin * 4.0
    return cnt


@pytest.mark.parametrize("shear_mode", [0, 1, 2])
def test_overlap_count_brute_force(shear_mode):
    rng = np.random.default_rng(5)
    L = 6.0
    pos = rng.uniform(0, L, (12, 3))
    pos[3] = pos[2] + [1.0, 0.5, 0.2]          # planted primary overlap
    pos[7] = pos[6] + [L - 0.9, 0.0, 0.3]      # overlap through a cell wall
    s = po.System(0, pos, box=(L, L, L), tr=1.0, shear_mode=shear_mode,
                  shear_shift=0.7 if shear_mode else 0.0)
    for rmin in (1.0, 0.5):
        want = _brute_overlap(pos, L, rmin, True, shear_mode,
                              0.7 if shear_mode else 0.0)
        assert s.check_overlap(pos, rmin) == want
    assert s.check_overlap(pos, 1.0) >= 2


def test_bd_single_sphere_closed_form():
    L = 12.0
    s = po.System(0, np.zeros((1, 3)), box=(L, L, L), tr=1.0)
    mu = s.atimes(np.array([1.0, 0.0, 0.0]))[0]
    x = 1.0 / L
    assert mu == pytest.approx(1 - 2.837297479 * x + 4 * math.pi / 3 * x**3,
                               rel=1e-6)
    F = np.array([0.3, -1.2, 0.5])
    y = np.array([0.7, 1.1, -0.4])
    pe, dt, nch = 2.0, 1e-2, 30
    e0, e1 = 0.8 * mu, 1.2 * mu
    a = po.cheb_coef(nch, 0, e0, e1)
    x0 = np.array([0.5, 1.0, -2.0])
    xn, dtt, used, nh, err = s.bd_mid_F(x0, y[None], pe, dt, nch, a, e0, e1,
                                        F=F)
    fact = math.sqrt(2 / (pe * dt))
    want = x0 + dt * (mu * F + fact * math.sqrt(mu) * y)
    assert dtt == dt and used == 1 and nh == 0
    assert err < 1e-12
    assert np.abs(xn - want).max() < 1e-10 * np.abs(want - x0).max()


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
def test_bd_step_vs_reference_composition():
    """brownian.c:2352-2560 composed from the compiled reference's own
    chebyshev_eval_atimes and atimes_3all (no overlap in this step)"""
    L = 8.0
    pos = np.array([[1, 1, 1], [4, 1.2, 0.8], [1.1, 4.2, 1], [4, 4, 1.3],
                    [1, 1.1, 4], [4.2, 1, 4], [1, 4, 4.1], [4, 4.1, 4.2]],
                   dtype=np.float64) + 0.3
    s = po.System(0, pos, box=(L, L, L), tr=1.0)
    n = pos.size
    rng = np.random.default_rng(11)
    y = rng.standard_normal(n)
    F = rng.standard_normal(n)
    pe, dt, nch = 1.0, 1e-3, 30
    M = s.matrix()
    w = np.linalg.eigvalsh(0.5 * (M + M.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    a = po.cheb_coef(nch, 0, e0, e1)
    xn, dtt, used, nh, err = s.bd_mid_F(pos, y[None], pe, dt, nch, a, e0, e1,
                                        F=F)
    assert used == 1 and nh == 0 and dtt == dt

    import ctypes as C
    R = po.ref()

    def ref_sys(p):
        t = po.System(0, p, box=(L, L, L), tr=1.0)
        return t, R.ref_sys_new(C.byref(t.s))

    t0, h0 = ref_sys(pos)
    z = np.zeros(n)
    R.ref_cheb_eval_sys(nch, po._p(a), h0, po._p(y), po._p(z), e0, e1)
    fact = math.sqrt(2 / (pe * dt))
    f = F + fact * z
    u = np.zeros(n)
    R.ref_atimes(h0, n, po._p(f), po._p(u))
    xmid = pos.reshape(-1) + 0.5 * dt * u
    t1, h1 = ref_sys(xmid)
    R.ref_atimes(h1, n, po._p(f), po._p(u))
    want = pos.reshape(-1) + dt * u
    R.ref_sys_free(h0)
    R.ref_sys_free(h1)
    disp = want - pos.reshape(-1)
    assert np.abs(xn - want).max() < 1e-12 * np.abs(disp).max()
    # z = M^{-1/2} y to the fit accuracy
    wv, V = np.linalg.eigh(0.5 * (M + M.T))
    zt = V @ ((V.T @ y) / np.sqrt(wv))
    assert np.linalg.norm(z - zt) < 1e-6 * np.linalg.norm(zt)


def _two_spheres(gap, L=10.0):
    pos = np.array([[2.0, 2.0, 2.0], [2.0 + gap, 2.0, 2.0]])
    return po.System(0, pos, box=(L, L, L), tr=1.0), pos


def test_bd_overlap_halving_keeps_draw():
    """a large first step that overlaps: dt is halved (fact recomputed) and
    the same draw is reused until the step is accepted.  Here the first two
    attempts overlap in the first stage, the third in the second, so the
    accepted attempt takes its first-stage product at the third attempt's
    intermediate configuration (REDO_scale sets no positions,
    brownian.c:2400, 2523)"""
    s, pos = _two_spheres(2.05)
    y = np.zeros(6)
    F = np.array([50.0, 0, 0, -50.0, 0, 0])    # pushes the pair together
    nch = 20
    a = po.cheb_coef(nch, 0, 0.3, 1.5)
    xn, dtt, used, nh, err = s.bd_mid_F(pos, y[None], 1.0, 0.01, nch, a,
                                        0.3, 1.5, F=F, dt_lim=1e-8)
    assert used == 1 and nh == 3
    assert dtt == pytest.approx(0.01 / 2**nh)
    assert s.check_overlap(xn, 1.0) == 0
    # the same control flow written out with the operator (no noise)
    def at(p, f):
        keep = s.pos.copy()
        s.pos[:] = p.reshape(s.pos.shape)
        u = s.atimes(f)
        s.pos[:] = keep
        return u

    x, dt, P, stages = pos.reshape(-1), 0.01, pos.reshape(-1), []
    for _ in range(8):
        x1 = x + 0.5 * dt * at(P, F)
        if s.check_overlap(x1, 1.0):
            stages.append(1)
            dt /= 2
            continue
        P = x1
        x2 = x + dt * at(x1, F)
        if s.check_overlap(x2, 1.0):
            stages.append(2)
            dt /= 2
            continue
        break
    assert stages == [1, 1, 2] and dt == dtt
    assert np.abs(xn - x2).max() < 1e-14 * np.abs(x2 - x).max()
    # the plain mid-point step at the final dt differs
    x3, dt3, _, nh3, _ = s.bd_mid_F(pos, y[None], 1.0, dtt, nch, a, 0.3, 1.5,
                                    F=F, dt_lim=1e-8)
    assert nh3 == 0 and not np.allclose(xn, x3, rtol=0, atol=1e-9)


def test_bd_overlap_redraw_below_dt_lim():
    """below dt_lim the draw is rejected and dt restored (REDO_FB); the
    reference keeps the halved-dt fact, fix_fact recomputes it"""
    s, pos = _two_spheres(2.05)
    F = np.array([50.0, 0, 0, -50.0, 0, 0])
    nch = 20
    a = po.cheb_coef(nch, 0, 0.3, 1.5)
    draws = np.zeros((3, 6))
    draws[1, 0] = -6.0        # second draw pulls the pair apart
    draws[1, 3] = 6.0
    draws[2] = draws[1]
    dt, dt_lim = 0.01, 0.003
    # dt 0.01 overlaps, 0.005 (>= dt_lim, same draw) overlaps, 0.0025 is
    # below dt_lim: redraw at dt = 0.01 with fact still that of dt = 0.005
    xn, dtt, used, nh, _ = s.bd_mid_F(pos, draws, 1.0, dt, nch, a, 0.3, 1.5,
                                      F=F, dt_lim=dt_lim)
    assert used == 2 and nh == 2 and dtt == dt
    fact_stale, fact_ok = math.sqrt(2 / 0.005), math.sqrt(2 / 0.01)
    xf, dtf, usedf, nhf, _ = s.bd_mid_F(pos, draws, 1.0, dt, nch, a, 0.3, 1.5,
                                        F=F, dt_lim=dt_lim, fix_fact=True)
    assert usedf == 2 and nhf == 2 and dtf == dt
    # both equal a plain step with the second draw, fact scaled accordingly:
    # Pe' = Pe fact_ok^2 / fact^2 gives the same fact at dt
    for xx, fact in ((xn, fact_stale), (xf, fact_ok)):
        pe = 2 / (fact**2 * dt)
        xw, dtw, uw, nw, _ = s.bd_mid_F(pos, draws[1:2], pe, dt, nch, a, 0.3,
                                        1.5, F=F, dt_lim=dt_lim)
        assert nw == 0 and uw == 1
        assert np.allclose(xx, xw, rtol=0, atol=1e-14)
    assert np.abs(xn - xf).max() > 1e-3
    # running out of draws reports -1
    _, dtn, usedn, _, _ = s.bd_mid_F(pos, draws[:1], 1.0, dt, nch, a, 0.3, 1.5,
                                     F=F, dt_lim=dt_lim)
    assert dtn == -1.0 and usedn == 1


@pytest.mark.parametrize("version,nm", [(2, 8), (2, 5), (1, 6), (0, 5)])
def test_dense_mixed_solve_residual(version, nm):
    """the restated solve (ewald-3fts-matrix.c:570-652, 1192-1290) returns
    generalised forces g and velocities v with v = M g on every component:
    the known ones reproduced, the unknown ones consistent"""
    L = 7.0
    rng = np.random.default_rng(version * 10 + nm)
    pos = np.array([[1, 1, 1], [4, 1.2, 0.8], [1.1, 4.2, 1], [4, 4, 1.3],
                    [1, 1.1, 4], [4.2, 1, 4], [1, 4, 4.1], [4, 4.1, 4.2]],
                   dtype=np.float64) + 0.2
    s = po.System(version, pos, box=(L, L, L), tr=1.0)
    d = po.dim(version)
    b = rng.standard_normal(8 * d)
    x = po.solve_mix_dense(s, nm, b)
    isH = np.array([i >= nm or q >= 6 for i in range(8) for q in range(d)])
    g = np.where(isH, x, b)          # forces / torques / stresslets
    v = np.where(isH, b, x)          # velocities / angular velocities / E
    assert np.abs(s.atimes(g) - v).max() < 1e-12 * np.abs(v).max()


def _quat_W(q):
    """W(Q) of ode-quaternion.c:28-48 (quaternion_W), the matrix whose
    inverse quaternion_Winv writes out"""
    return np.array([[+q[3], +q[2], -q[1], -q[0]],
                     [-q[2], +q[3], +q[0], -q[1]],
                     [+q[1], -q[0], +q[3], -q[2]],
                     [+q[0], +q[1], +q[2], +q[3]]])


def test_quaternion_rate():
    """dQ/dt = 1/2 W^{-1} (O, 0) (ode-quaternion.c:89-105): the columns the
    oracle applies equal numpy's inverse of W (what quaternion_Winv_lapack
    computes); |Q| is conserved to first order; a rotation about z at rate
    w turns the Euler-angle quaternion of quaternion_by_Euler
    (ode-quaternion.c:107-121, theta = 0: Q = (0, 0, sin(a/2), cos(a/2)))
    at da/dt = w"""
    rng = np.random.default_rng(4)
    for _ in range(5):
        q = rng.standard_normal(4)
        Winv = np.linalg.inv(_quat_W(q))
        for j in range(3):
            e = np.zeros(3)
            e[j] = 1.0
            assert np.allclose(2 * po.quat_dqdt(q, e), Winv[:, j],
                               rtol=0, atol=1e-14 * np.abs(Winv).max())
        o = rng.standard_normal(3)
        assert abs(q @ po.quat_dqdt(q, o)) < 1e-14 * (q @ q) * np.abs(o).max()
    a, w = 0.7, 1.3
    q = np.array([0.0, 0.0, math.sin(a / 2), math.cos(a / 2)])
    dq = po.quat_dqdt(q, np.array([0.0, 0.0, w]))
    want = 0.5 * w * np.array([0.0, 0.0, math.cos(a / 2), -math.sin(a / 2)])
    assert np.allclose(dq, want, rtol=0, atol=1e-15)


def _eight(L=8.0, shift=0.3):
    return np.array([[1, 1, 1], [4, 1.2, 0.8], [1.1, 4.2, 1], [4, 4, 1.3],
                     [1, 1.1, 4], [4.2, 1, 4], [1, 4, 4.1], [4, 4.1, 4.2]],
                    dtype=np.float64) + shift


def _ft_of_fts_dense(s):
    """(F, T) -> (U, O) of the FTS mobility problem with E = 0 from the
    restated dense solve (ewald-3fts-matrix.c:570-652)"""
    return po.ft_mobility_dense(s)


def test_fts_mobility_of_forces_and_torques():
    """the oracle's E = 0 mobility problem (BD_atimes_mob_FU,
    brownian.c:597-646) equals the restated dense solve and is a symmetric
    positive-definite map"""
    L = 8.0
    s = po.System(2, _eight(L), box=(L, L, L), tr=1.0)
    G = _ft_of_fts_dense(s)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(48)
    y = s.bd_op(x)
    assert np.abs(y - G @ x).max() < 1e-12 * np.abs(y).max()
    assert np.abs(G - G.T).max() < 1e-12 * np.abs(G).max()
    assert np.linalg.eigvalsh(0.5 * (G + G.T)).min() > 0


@pytest.mark.parametrize("version", [1, 2])
def test_bd_step_ft_fts_composition(version):
    """brownian.c:2352-2560 for FT / FTS: x and q after one step against
    the composition of the steps with the compiled reference's own
    chebyshev_eval_atimes and atimes_3all (FT), or with the restated dense
    solve and the Chebyshev restatement (FTS)"""
    import ctypes as C
    L = 8.0
    pos = _eight(L)
    s = po.System(version, pos, box=(L, L, L), tr=1.0)
    n6 = 48
    rng = np.random.default_rng(20 + version)
    y = rng.standard_normal(n6)
    F = rng.standard_normal(24)
    T = rng.standard_normal(24)
    q = rng.standard_normal(32)
    pe, dt, nch = 1.0, 1e-3, 30
    G0 = s.matrix() if version == 1 else _ft_of_fts_dense(s)
    w = np.linalg.eigvalsh(0.5 * (G0 + G0.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    a = po.cheb_coef(nch, 0, e0, e1)
    xn, qn, dtt, used, nh, err = s.bd_mid(pos, y[None], pe, dt, nch, a, e0,
                                          e1, q=q, F=F, T=T)
    assert used == 1 and nh == 0 and dtt == dt
    fact = math.sqrt(2 / (pe * dt))

    def ft(z):
        out = np.empty(n6)
        out.reshape(8, 6)[:, :3] = F.reshape(8, 3) + fact * z.reshape(8, 6)[:, :3]
        out.reshape(8, 6)[:, 3:] = T.reshape(8, 3) + fact * z.reshape(8, 6)[:, 3:]
        return out
    if version == 1:
        R = po.ref()

        def ref_sys(p):
            t = po.System(1, p, box=(L, L, L), tr=1.0)
            return t, R.ref_sys_new(C.byref(t.s))
        t0, h0 = ref_sys(pos)
        z = np.zeros(n6)
        R.ref_cheb_eval_sys(nch, po._p(a), h0, po._p(y), po._p(z), e0, e1)
        f = ft(z)
        u = np.zeros(n6)
        R.ref_atimes(h0, n6, po._p(f), po._p(u))
        xmid = pos.reshape(-1) + 0.5 * dt * u.reshape(8, 6)[:, :3].reshape(-1)
        t1, h1 = ref_sys(xmid)
        R.ref_atimes(h1, n6, po._p(f), po._p(u))
        R.ref_sys_free(h0)
        R.ref_sys_free(h1)
    else:
        z = po.cheb_eval_dense(a, G0, y, e0, e1)
        f = ft(z)
        u = G0 @ f
        xmid = pos.reshape(-1) + 0.5 * dt * u.reshape(8, 6)[:, :3].reshape(-1)
        s1 = po.System(2, xmid.reshape(-1, 3), box=(L, L, L), tr=1.0)
        u = _ft_of_fts_dense(s1) @ f
    U = u.reshape(8, 6)[:, :3].reshape(-1)
    O = u.reshape(8, 6)[:, 3:]
    want = pos.reshape(-1) + dt * U
    disp = np.abs(want - pos.reshape(-1)).max()
    assert np.abs(xn - want).max() < 1e-11 * disp
    qw = q.reshape(8, 4) + dt * np.stack([po.quat_dqdt(q.reshape(8, 4)[i], O[i])
                                          for i in range(8)])
    assert np.abs(qn - qw.reshape(-1)).max() < 1e-13 * np.abs(q).max()
    assert np.abs(qn - q).max() > 1e-6


@pytest.mark.parametrize("scheme,bb_n", [(1, 2.0), (1, 4.0), (2, 1.0)])
def test_bd_bb03_bm97_vs_reference_composition(scheme, bb_n):
    """BD_evolve_BB03 (brownian.c:2579-2920) and BD_evolve_BM97
    (2970-3190), F version, composed from the compiled reference's own
    chebyshev_eval_atimes and atimes_3all (no overlap in this step); BB03
    takes its deterministic product back at x: BD_add_FP (BD, x, f) moves
    the system there before it (2209-2212, 2800-2816)"""
    import ctypes as C
    L = 8.0
    pos = _eight(L)
    s = po.System(0, pos, box=(L, L, L), tr=1.0)
    n = pos.size
    rng = np.random.default_rng(30 + scheme)
    y = rng.standard_normal(n)
    F = rng.standard_normal(n)
    pe, dt, nch = 1.0, 1e-3, 30
    M = s.matrix()
    w = np.linalg.eigvalsh(0.5 * (M + M.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    a = po.cheb_coef(nch, 0, e0, e1)
    xn, _, dtt, used, nh, err = s.bd_mid(pos, y[None], pe, dt, nch, a, e0, e1,
                                         F=F, scheme=scheme, bb_n=bb_n)
    assert used == 1 and nh == 0 and dtt == dt
    R = po.ref()

    def ref_sys(p):
        t = po.System(0, p, box=(L, L, L), tr=1.0)
        return t, R.ref_sys_new(C.byref(t.s))

    def ref_op(h, f):
        u = np.zeros(n)
        R.ref_atimes(h, n, po._p(np.ascontiguousarray(f)), po._p(u))
        return u
    t0, h0 = ref_sys(pos)
    z = np.zeros(n)
    R.ref_cheb_eval_sys(nch, po._p(a), h0, po._p(y), po._p(z), e0, e1)
    fact = math.sqrt(2 / (pe * dt))
    x0 = pos.reshape(-1)
    if scheme == 2:
        u0 = ref_op(h0, F + fact * z)
        t1, h1 = ref_sys(x0 + dt * u0)
        u = 0.5 * (u0 + ref_op(h1, F + fact * z))
    else:
        u0 = ref_op(h0, fact * z)
        t1, h1 = ref_sys(x0 + dt / bb_n * u0)
        ub = u0 + 0.5 * bb_n * (ref_op(h1, fact * z) - u0)
        u = ref_op(h0, F) + ub
    R.ref_sys_free(h0)
    R.ref_sys_free(h1)
    want = x0 + dt * u
    assert np.abs(xn - want).max() < 1e-12 * np.abs(want - x0).max()


def test_fts_mobility_with_fixed_particles():
    """BD_atimes_mob_FU with particles nm..np-1 fixed (brownian.c:650-729):
    the oracle's (F, T) -> (U, O) equals the restated dense mixed solve
    (ewald-3fts-matrix.c:1192-1290) with the fixed particles at rest"""
    L = 8.0
    s = po.System(2, _eight(L), box=(L, L, L), tr=1.0)
    nm = 5
    rng = np.random.default_rng(8)
    ft = rng.standard_normal(6 * nm)
    uo = s.fts_mix_ft(nm, ft)
    b = np.zeros(88)
    b.reshape(8, 11)[:nm, :6] = ft.reshape(nm, 6)
    x = po.solve_mix_dense(s, nm, b)
    want = x.reshape(8, 11)[:nm, :6].reshape(-1)
    assert np.abs(uo - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("version", [0, 1])
def test_bd_step_fixed_particles_composition(version):
    """a mid-point step with particles 5..7 fixed: the noise uses the
    zero-padded product of the compiled reference's atimes_3all on the
    mobile rows (brownian.c:521-547, 562-588); the drift is the mixed
    problem of solve_mix_3f / 3ft (ewald-3f.c:269-330), the fixed particles
    at rest with unknown forces, here by the Schur complement of the
    reference operator's columns; the fixed particles stay"""
    import ctypes as C
    L = 8.0
    pos = _eight(L)
    s = po.System(version, pos, box=(L, L, L), tr=1.0)
    nm, d = 5, (3 if version == 0 else 6)
    n = d * nm
    rng = np.random.default_rng(40 + version)
    y = rng.standard_normal(n)
    F = rng.standard_normal(3 * nm)
    T = rng.standard_normal(3 * nm) if version else None
    pe, dt, nch = 1.0, 1e-3, 30
    G = s.matrix()[:n, :n]
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    a = po.cheb_coef(nch, 0, e0, e1)
    xn, _, dtt, used, nh, err = s.bd_mid(pos, y[None], pe, dt, nch, a, e0, e1,
                                         F=F, T=T, nm=nm)
    assert used == 1 and nh == 0 and dtt == dt and xn.size == 3 * nm
    R = po.ref()

    def ref_op(p, f):
        t = po.System(version, p, box=(L, L, L), tr=1.0)
        h = R.ref_sys_new(C.byref(t.s))
        M = np.zeros((d * 8, d * 8))
        for k in range(d * 8):
            e = np.zeros(d * 8)
            e[k] = 1.0
            u = np.zeros(d * 8)
            R.ref_atimes(h, d * 8, po._p(e), po._p(u))
            M[:, k] = u
        R.ref_sys_free(h)
        # U_f = 0: M_fm f + M_ff F_f = 0, U_m = M_mm f + M_mf F_f
        Ff = np.linalg.solve(M[n:, n:], -M[n:, :n] @ f)
        return M[:n, :n] @ f + M[:n, n:] @ Ff
    z = po.cheb_eval_dense(a, G, y, e0, e1)
    fact = math.sqrt(2 / (pe * dt))
    f = F + fact * z if version == 0 else np.concatenate(
        [F.reshape(nm, 3), T.reshape(nm, 3)], axis=1).reshape(-1) + fact * z
    x0 = pos.reshape(-1)
    u0 = ref_op(pos, f)
    U0 = u0 if version == 0 else u0.reshape(nm, 6)[:, :3].reshape(-1)
    pm = pos.copy()
    pm[:nm] = (x0[:3 * nm] + 0.5 * dt * U0).reshape(nm, 3)
    u1 = ref_op(pm, f)
    U1 = u1 if version == 0 else u1.reshape(nm, 6)[:, :3].reshape(-1)
    want = x0[:3 * nm] + dt * U1
    assert np.abs(xn - want).max() < 1e-12 * np.abs(want - x0[:3 * nm]).max()


def _approaching_pair(L, gap):
    """_eight with particles 0 and 1 at centre distance 2 + gap along x"""
    pos = _eight(L)
    pos[1] = pos[0] + np.array([2.0 + gap, 0.0, 0.0])
    return pos


def test_bd_retry_after_second_stage_overlap_keeps_positions():
    """a mid-point step whose second stage overlaps while dt / 2 stays above
    dt_lim: the retry (REDO_scale, brownian.c:2400, 2523) keeps the noise
    and takes its first-stage product at the intermediate configuration of
    the failed attempt, which the system still holds; composed here from the
    compiled reference's atimes_3all, and unlike the composition that takes
    it at x"""
    import ctypes as C
    L, gap = 8.0, 0.05
    pos = _approaching_pair(L, gap)
    s = po.System(0, pos, box=(L, L, L), tr=1.0)
    n = 24
    y = np.random.default_rng(44).standard_normal(n)
    F = np.zeros(n)
    F[0], F[3] = 100.0, -100.0
    pe, dt, nch = 1e6, 1e-3, 30
    G = s.matrix()
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    a = po.cheb_coef(nch, 0, e0, e1)
    xn, _, dtt, used, nh, err = s.bd_mid(pos, y[None], pe, dt, nch, a, e0,
                                         e1, F=F)
    assert used == 1 and nh == 1 and dtt == dt / 2
    R = po.ref()

    def M(p, f):
        t = po.System(0, p.reshape(-1, 3), box=(L, L, L), tr=1.0)
        h = R.ref_sys_new(C.byref(t.s))
        u = np.zeros(n)
        R.ref_atimes(h, n, po._p(np.ascontiguousarray(f)), po._p(u))
        R.ref_sys_free(h)
        return u

    def contact(p):
        d = p.reshape(-1, 3)
        return np.linalg.norm(d[1] - d[0]) <= 2.0

    x = pos.reshape(-1)
    z = po.cheb_eval_dense(a, G, y, e0, e1)
    f = F + math.sqrt(2 / (pe * dt)) * z
    x1 = x + 0.5 * dt * M(x, f)
    assert not contact(x1)
    assert contact(x + dt * M(x1, f))            # the second stage overlaps
    h = dt / 2
    f = F + math.sqrt(2 / (pe * h)) * z
    xm = x + 0.5 * h * M(x1, f)                   # first stage at x1
    want = x + h * M(xm, f)
    xm_x = x + 0.5 * h * M(x, f)                  # the same taken at x
    other = x + h * M(xm_x, f)
    disp = np.abs(want - x).max()
    assert np.abs(xn - want).max() < 1e-12 * disp
    assert np.abs(other - want).max() > 1e-6 * disp


def _uinf(p, Ui, Oi, Ei):
    E = np.array([[Ei[0], Ei[1], Ei[2]], [Ei[1], Ei[4], Ei[3]],
                  [Ei[2], Ei[3], -Ei[0] - Ei[4]]])
    p = p.reshape(-1, 3)
    return (Ui + np.cross(Oi, p) + p @ E.T).reshape(-1)


@pytest.mark.parametrize("scheme,nm", [(0, 8), (1, 8), (0, 5)])
def test_bd_step_in_imposed_flow_composition(scheme, nm):
    """F steps in an imposed flow (Ui, Oi, Ei, no shear): every drift is
    solve_mix_3f in the labo frame, M f plus the flow at the positions of
    the product (shift_rest_to_labo_U, f.c:621-660), so BB03's combination
    of three drift products carries the flow three times, as the reference
    computes it (brownian.c:2818-2829); with fixed particles their
    rest-frame velocities are minus the flow at the positions of particles
    0..nf-1 (ewald-3f.c:301); the noise stays in the rest frame.  Composed
    from the compiled reference's atimes_3all"""
    import ctypes as C
    L = 8.0
    pos = _eight(L)
    s = po.System(0, pos, box=(L, L, L), tr=1.0)
    Ui, Oi = np.array([0.2, -0.1, 0.3]), np.array([0.1, 0.2, -0.15])
    Ei = np.array([0.05, 0.1, -0.07, 0.02, -0.04])
    s.s.Ui[:], s.s.Oi[:], s.s.Ei[:] = Ui, Oi, Ei
    n = 3 * nm
    rng = np.random.default_rng(50 + scheme + nm)
    y = rng.standard_normal(n)
    F = rng.standard_normal(n)
    pe, dt, nch, bbn = 1.0, 1e-3, 30, 2.0
    G = s.matrix()[:n, :n]
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    a = po.cheb_coef(nch, 0, e0, e1)
    xn, _, dtt, used, nh, err = s.bd_mid(pos, y[None], pe, dt, nch, a, e0, e1,
                                         F=F, scheme=scheme, bb_n=bbn, nm=nm)
    assert used == 1 and nh == 0
    R = po.ref()

    def drift(p, f):
        full = pos.copy()
        full[:nm] = p.reshape(-1, 3)
        t = po.System(0, full, box=(L, L, L), tr=1.0)
        h = R.ref_sys_new(C.byref(t.s))
        M = np.zeros((24, 24))
        for k in range(24):
            e = np.zeros(24)
            e[k] = 1.0
            u = np.zeros(24)
            R.ref_atimes(h, 24, po._p(e), po._p(u))
            M[:, k] = u
        R.ref_sys_free(h)
        if nm == 8:
            um = M @ f
        else:
            uf0 = -_uinf(full[:8 - nm], Ui, Oi, Ei)
            Ff = np.linalg.solve(M[n:, n:], uf0 - M[n:, :n] @ f)
            um = M[:n, :n] @ f + M[:n, n:] @ Ff
        return um + _uinf(full[:nm], Ui, Oi, Ei)

    z = po.cheb_eval_dense(a, G, y, e0, e1)
    fz = math.sqrt(2 / (pe * dt)) * z
    x = pos[:nm].reshape(-1)
    if scheme == 0:
        x1 = x + 0.5 * dt * drift(x, F + fz)
        want = x + dt * drift(x1, F + fz)
    else:
        u0 = drift(x, fz)
        x1 = x + dt / bbn * u0
        uB = u0 + 0.5 * bbn * (drift(x1, fz) - u0)
        want = x + dt * (drift(x, F) + uB)
    disp = np.abs(want - x).max()
    assert np.abs(xn - want).max() < 1e-12 * disp


@pytest.mark.parametrize("scheme", [0, 1])
def test_bd_sheared_step_composition(scheme):
    """an F step of a sheared box (shear_mode 1, rate 0.8, reference shift
    s0 at t0): the noise and first stage at the cell shift of t, the second
    at t + h (mid-point h = dt/2), BB03's deterministic product back at t
    and at x (brownian.c:2393, 2468, 2741, 2805, 2808);
    the drift carries the shear flow of stokes_set_shear (Oi_z = -rate/2,
    Ei_xy = rate/2).  Composed from the compiled reference's atimes_3all at
    those shifts"""
    import ctypes as C
    L, rate, t, t0, s0, bbn = 8.0, 0.8, 3.7, 1.0, 0.4, 2.0
    pos = _eight(L)
    s = po.System(0, pos, box=(L, L, L), tr=1.0, shear_mode=1)
    Ui, Oi = np.zeros(3), np.array([0.0, 0.0, -0.5 * rate])
    Ei = np.array([0.0, 0.5 * rate, 0.0, 0.0, 0.0])
    s.s.Oi[:], s.s.Ei[:] = Oi, Ei
    s.s.shear_rate, s.s.bd_t, s.s.bd_t0, s.s.bd_s0 = rate, t, t0, s0

    def shift(tt):
        v = s0 + L * rate * (tt - t0)
        while v < -L / 2:
            v += L
        while v >= L / 2:
            v -= L
        return v

    n = 24
    rng = np.random.default_rng(60 + scheme)
    y = rng.standard_normal(n)
    F = rng.standard_normal(n)
    pe, dt, nch = 1.0, 1e-3, 30
    s.s.shear_shift = shift(t)
    G = s.matrix()
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    a = po.cheb_coef(nch, 0, e0, e1)
    xn, _, dtt, used, nh, err = s.bd_mid(pos, y[None], pe, dt, nch, a, e0, e1,
                                         F=F, scheme=scheme, bb_n=bbn)
    assert used == 1 and nh == 0
    R = po.ref()

    def drift(p, f, sh):
        tt = po.System(0, p.reshape(-1, 3), box=(L, L, L), tr=1.0,
                       shear_mode=1, shear_shift=sh)
        h = R.ref_sys_new(C.byref(tt.s))
        u = np.zeros(n)
        R.ref_atimes(h, n, po._p(np.ascontiguousarray(f)), po._p(u))
        R.ref_sys_free(h)
        return u + _uinf(p, Ui, Oi, Ei)

    z = po.cheb_eval_dense(a, G, y, e0, e1)
    fz = math.sqrt(2 / (pe * dt)) * z
    x = pos.reshape(-1)
    if scheme == 0:
        x1 = x + 0.5 * dt * drift(x, F + fz, shift(t))
        want = x + dt * drift(x1, F + fz, shift(t + 0.5 * dt))
    else:
        u0 = drift(x, fz, shift(t))
        x1 = x + dt / bbn * u0
        uB = u0 + 0.5 * bbn * (drift(x1, fz, shift(t + dt / bbn)) - u0)
        want = x + dt * (drift(x, F, shift(t)) + uB)
    disp = np.abs(want - x).max()
    assert np.abs(xn - want).max() < 1e-12 * disp
    # the shifts matter: the same without the change at t + h differs
    if scheme == 0:
        other = x + dt * drift(x1, F + fz, shift(t))
        assert np.abs(other - want).max() > 1e-9 * disp


def _compose_ref_step(scheme, x, F, ys, pe, dt, dt_lim, bbn, t, shift, L,
                      a, e0, e1, flow, shear_mode):
    """the reference's F step, BD_evolve_mid (brownian.c:2352-2560) or
    BD_evolve_BB03 (2579-2920), with its REDO / REDO_FB / REDO_scale control
    flow: each product is the compiled reference's atimes_3all at the
    positions and cell shift the reference's system holds at that point
    (stokes_set_pos_mobile and stokes_set_shear_shift at 2390-2393, 2466-2468;
    2627-2630, 2738-2741, 2802-2808 where BD_add_FP moves BB03 back to x),
    the noise chebyshev_eval_atimes of the dense operator where it is drawn;
    returns (x_new, dt_taken, path of overlap stages)"""
    import ctypes as C
    R = po.ref()
    n = x.size
    Ui, Oi, Ei = flow

    def M(p, sh, f):
        s = po.System(0, p.reshape(-1, 3), box=(L, L, L), tr=1.0,
                      shear_mode=shear_mode, shear_shift=sh)
        h = R.ref_sys_new(C.byref(s.s))
        u = np.zeros(n)
        R.ref_atimes(h, n, po._p(np.ascontiguousarray(f)), po._p(u))
        R.ref_sys_free(h)
        return u

    def drift(p, sh, f):
        return M(p, sh, f) + _uinf(p, Ui, Oi, Ei)

    def overlap(p, sh):
        s = po.System(0, p.reshape(-1, 3), box=(L, L, L), tr=1.0,
                      shear_mode=shear_mode, shear_shift=sh)
        return s.check_overlap(p, 1.0)

    def noise(p, sh, y):
        G = np.stack([M(p, sh, e) for e in np.eye(n)], axis=1)
        return po.cheb_eval_dense(a, G, y, e0, e1)

    path, k, dtl = [], 0, dt
    fact = math.sqrt(2 / (pe * dtl))
    label = "redo"
    while True:
        if label == "redo":
            P, S = x.copy(), shift(t)
        if label in ("redo", "redo_fb"):
            z = noise(P, S, ys[k])
            k += 1
        if scheme == 0:
            f = F + fact * z
            u0 = drift(P, S, f)
            x1 = x + 0.5 * dtl * u0
        else:
            u0 = drift(P, S, fact * z)
            x1 = x + dtl / bbn * u0
        if overlap(x1, S):
            stage = 1
        else:
            P, S = x1, shift(t + (0.5 * dtl if scheme == 0 else dtl / bbn))
            if scheme == 0:
                u = drift(P, S, f)
            else:
                uB = u0 + 0.5 * bbn * (drift(P, S, fact * z) - u0)
                P, S = x.copy(), shift(t)
                u = drift(P, S, F) + uB
            xn = x + dtl * u
            if not overlap(xn, S):
                return xn, dtl, path
            stage = 2
        path.append(stage)
        dtl *= 0.5
        if dtl < dt_lim:
            dtl = dt
            label = "redo_fb" if stage == 1 else "redo"
        else:
            fact = math.sqrt(2 / (pe * dtl))
            label = "scale"


@pytest.mark.parametrize("scheme,gap,F0,ys,pe,shear,want_path", [
    (0, 0.02, 100.0, 0.0, 1e6, 1, [1, 2]),
    (1, 0.06, 50.0, 3.0, 1.0, 1, [1, 2, 2]),
    (1, 0.02, 100.0, 0.0, 1e6, 0, [2, 2]),
])
def test_bd_retries_match_reference_control_flow(scheme, gap, F0, ys, pe,
                                                 shear, want_path):
    """steps that halve after first- and second-stage overlaps, sheared
    (shear_mode 1, rate 0.8) and not: the oracle's step equals the
    reference's control flow composed from its compiled atimes_3all.  The
    mid-point retry after a second-stage overlap starts from the failed
    attempt's intermediate configuration at the shift it left; BB03's starts
    from x at the shift of t (its deterministic product moved the system
    back); a first-stage retry keeps x and the shift of t"""
    L, rate, t, t0, s0, bbn = 8.0, 0.8, 3.7, 1.0, 0.4, 2.0
    pos = _eight(L)
    pos[1] = pos[0] + np.array([2.0 + gap, 0.0, 0.0])

    def shift(tt):
        if not shear:
            return 0.0
        v = s0 + L * rate * (tt - t0)
        while v < -L / 2:
            v += L
        while v >= L / 2:
            v -= L
        return v

    flow = ((np.zeros(3), np.array([0.0, 0.0, -0.5 * rate]),
             np.array([0.0, 0.5 * rate, 0.0, 0.0, 0.0])) if shear
            else (np.zeros(3), np.zeros(3), np.zeros(5)))
    s = po.System(0, pos, box=(L, L, L), tr=1.0, shear_mode=shear)
    s.s.Oi[:], s.s.Ei[:] = flow[1], flow[2]
    s.s.shear_rate, s.s.bd_t, s.s.bd_t0, s.s.bd_s0 = rate, t, t0, s0
    s.s.shear_shift = shift(t)
    n = 24
    y = np.random.default_rng(5).standard_normal(n) * 0.1
    y[0] += ys
    y[3] -= ys
    F = np.zeros(n)
    F[0], F[3] = F0, -F0
    G = s.matrix()
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    a = po.cheb_coef(30, 0, e0, e1)
    dt = 1e-3
    xn, _, dtt, used, nh, err = s.bd_mid(pos, np.stack([y] * 4), pe, dt, 30,
                                         a, e0, e1, F=F, scheme=scheme,
                                         bb_n=bbn)
    want, dtw, path = _compose_ref_step(scheme, pos.reshape(-1), F, [y] * 4,
                                        pe, dt, 1e-6, bbn, t, shift, L, a, e0,
                                        e1, flow, shear)
    assert path == want_path
    assert nh == len(want_path) and dtt == dtw and used == 1
    disp = np.abs(want - pos.reshape(-1)).max()
    assert np.abs(xn - want).max() < 1e-12 * disp
