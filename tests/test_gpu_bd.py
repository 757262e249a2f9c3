"""The device Brownian-dynamics step (stokes_hip_bd_*, F version, mid-point
scheme of brownian.c:2352-2560) against the oracle's restatement of the
same step, fed the Gaussian draw the device used.

Bars: FP64 operator -- positions after a step within 1e-10 of the oracle
relative to the displacement (the step is 31 FP64 operator products plus
the Chebyshev recursion); z = M^{-1/2} y within the fit accuracy of the
eigendecomposition closed form.  Each step is one replay of a captured
graph; replays alternate with eager operator calls and host syncs.
"""
import math

import numpy as np
import pytest

import libstokes_amd as ls
from libstokes_amd import workloads
import pyoracle as po
from conftest import relerr

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _setup(n, phi, seed=3, precision=ls.FP64):
    pos, L = workloads.configuration(n, phi, seed=seed)
    s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0, precision=precision)
    o = po.System(0, pos, box=(L, L, L), tr=1.0)
    return pos, L, s, o


def _spectrum(o):
    M = o.matrix()
    return np.linalg.eigh(0.5 * (M + M.T))


def test_bd_steps_match_oracle():
    n, phi = 27, 0.1
    _say(f"bd parity N={n} phi={phi}")
    pos, L, s, o = _setup(n, phi)
    w, V = _spectrum(o)
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    nch, pe, dt = 30, 1.0, 1e-3
    a = po.cheb_coef(nch, 0, e0, e1)
    rng = np.random.default_rng(2)
    F = rng.standard_normal(3 * n)
    Fd = ls.DeviceArray.from_host(F)
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1e-6, dt_lim=1e-6, rmin_ol=1.0,
               seed=7, force_dev=Fd, bounds=(e0, e1))
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    x = pos.reshape(-1).copy()
    ys = []
    for step in range(4):
        info = bd.step(xd, dt)
        y, z = bd.last_noise()
        ys.append(y)
        # overlaps halve dt with the same draw (brownian.c:2435-2459); a
        # refit changes the bounds: the oracle takes the bounds in use
        assert info.redraws == 0
        assert info.attempts == 1 + info.halvings + info.refits
        assert info.err_cheb < 1e-6
        o.pos[:] = x
        if step == 0:
            assert info.attempts == 1
            zt = V @ ((V.T @ y) / np.sqrt(w))
            assert relerr(z, zt) < 1e-7
        b0, b1 = info.eig
        a = po.cheb_coef(nch, 0, b0, b1)
        xo, dto, used, nh, err = o.bd_mid_F(x, y[None], pe, dt, nch, a, b0,
                                            b1, F=F)
        assert used == 1 and nh == info.halvings and dto == info.dt
        xg = xd.to_host()
        disp = np.abs(xo - x).max()
        _say(f"  step {step}: |x_gpu - x_oracle| / disp = "
             f"{np.abs(xg - xo).max() / disp:.2e}")
        assert np.abs(xg - xo).max() < 1e-10 * disp
        x = xo
        # eager operator work between replays: still the oracle's operator
        s.set_pos(x.reshape(-1, 3))
        v = rng.standard_normal(3 * n)
        o.pos[:] = x
        assert relerr(s.atimes(v), o.atimes(v)) < 1e-12
    # fresh draws each step, standard normal
    assert not np.array_equal(ys[0], ys[1])
    bd.close()


def test_bd_noise_statistics_and_seed():
    n = 2048
    pos, L = workloads.configuration(n, 0.2, seed=1)
    s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0, precision=ls.MIXED)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    x0 = pos.reshape(-1).copy()
    draws = []
    for rep in range(2):
        xd = ls.DeviceArray.from_host(x0)
        bd = ls.BD(s, peclet=1.0, n_cheb=16, eps=1.0, seed=99,
                   bounds=(0.2, 3.0))
        bd.step(xd, 1e-4)
        draws.append(bd.last_noise()[0])
        bd.close()
    y = draws[0]
    assert np.array_equal(draws[0], draws[1])          # same seed, same draw
    m, v = y.mean(), y.var()
    _say(f"  N(0,1) check on {y.size}: mean {m:.4f} var {v:.4f}")
    assert abs(m) < 5 / math.sqrt(y.size)
    assert abs(v - 1) < 6 * math.sqrt(2 / y.size)
    # fourth moment of a normal is 3
    assert abs((y**4).mean() - 3) < 0.3


def _pair(gap=2.05, L=10.0):
    pos = np.array([[2.0, 2.0, 2.0], [2.0 + gap, 2.0, 2.0]])
    s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0)
    o = po.System(0, pos, box=(L, L, L), tr=1.0)
    return pos, s, o


def test_bd_overlap_halving_matches_oracle():
    pos, s, o = _pair()
    F = np.array([50.0, 0, 0, -50.0, 0, 0])
    nch, e0, e1 = 20, 0.3, 1.5
    a = po.cheb_coef(nch, 0, e0, e1)
    bd = ls.BD(s, peclet=1e6, n_cheb=nch, eps=1.0, dt_lim=1e-8, seed=1,
               force_dev=ls.DeviceArray.from_host(F), bounds=(e0, e1))
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    info = bd.step(xd, 0.01)
    y, _ = bd.last_noise()
    xo, dto, used, nh, _ = o.bd_mid_F(pos.reshape(-1), y[None], 1e6, 0.01, nch,
                                      a, e0, e1, F=F, dt_lim=1e-8)
    _say(f"  halvings gpu {info.halvings} oracle {nh}, dt {info.dt} {dto}")
    assert info.halvings == nh >= 1 and info.redraws == 0
    assert info.dt == dto
    assert info.attempts == nh + 1
    xg = xd.to_host()
    assert np.abs(xg - xo).max() < 1e-10 * np.abs(xo - pos.reshape(-1)).max()
    bd.close()


def test_bd_redraw_below_dt_lim():
    """F = 10 closes the gap 0.05 within dt = 0.01 but not within 0.005,
    and dt_lim = 0.006: an overlap halves dt below dt_lim, so the draw is
    rejected and dt restored (brownian.c:2501-2525)"""
    F = np.array([10.0, 0, 0, -10.0, 0, 0])
    # without noise every attempt overlaps: the step gives up
    pos, s, o = _pair()
    bd = ls.BD(s, peclet=1e12, n_cheb=20, eps=1.0, dt_lim=0.006, seed=5,
               force_dev=ls.DeviceArray.from_host(F), bounds=(0.3, 1.5))
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    with pytest.raises(ls.StokesError, match="attempts"):
        bd.step(xd, 0.01)
    bd.close()
    # with noise some draw separates the pair; the accepted step is the
    # plain step of that draw at the restored dt
    nch, e0, e1 = 20, 0.3, 1.5
    a = po.cheb_coef(nch, 0, e0, e1)
    seen = 0
    for seed in range(1, 21):
        pos, s, o = _pair()
        bd = ls.BD(s, peclet=1.0, n_cheb=nch, eps=1.0, dt_lim=0.006,
                   seed=seed, force_dev=ls.DeviceArray.from_host(F),
                   bounds=(e0, e1))
        xd = ls.DeviceArray.from_host(pos.reshape(-1))
        info = bd.step(xd, 0.01)
        y, _ = bd.last_noise()
        bd.close()
        assert info.dt == 0.01 and info.redraws == info.halvings
        xg = xd.to_host()
        assert o.check_overlap(xg, 1.0) == 0
        xo, dto, used, nh, _ = o.bd_mid_F(pos.reshape(-1), y[None], 1.0, 0.01,
                                          nch, a, e0, e1, F=F, dt_lim=0.006)
        assert used == 1 and nh == 0 and dto == 0.01
        assert np.abs(xg - xo).max() < 1e-10 * np.abs(xo - pos.reshape(-1)).max()
        if info.redraws:
            seen += 1
            _say(f"  seed {seed}: redraws {info.redraws} attempts "
                 f"{info.attempts}")
        s.close()
    assert seen >= 1


def test_bd_lanczos_bounds_and_refit():
    n, phi = 27, 0.1
    pos, L, s, o = _setup(n, phi, seed=4)
    w, V = _spectrum(o)
    # no bounds: the first step estimates them by Lanczos
    bd = ls.BD(s, peclet=1.0, n_cheb=30, eps=1e-6, seed=3)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    info = bd.step(xd, 1e-4)
    _say(f"  lanczos bounds {info.eig[0]:.6f} {info.eig[1]:.6f} "
         f"exact {w[0]:.6f} {w[-1]:.6f}, err {info.err_cheb:.2e}")
    assert info.eig[0] <= w[0] * (1 + 1e-6) and info.eig[1] >= w[-1] * (1 - 1e-6)
    assert info.eig[0] > 0.5 * w[0] and info.eig[1] < 1.5 * w[-1]
    assert info.refits == 0 and info.err_cheb < 1e-6
    bd.close()
    # bounds far off: the accuracy check refits once and re-evaluates
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    bd = ls.BD(s, peclet=1.0, n_cheb=30, eps=1e-6, seed=3,
               bounds=(0.01 * w[0], 0.02 * w[0]))
    info = bd.step(xd, 1e-4)
    y, z = bd.last_noise()
    _say(f"  refit: refits {info.refits} err {info.err_cheb:.2e}")
    assert info.refits == 1 and info.err_cheb < 1e-6
    zt = V @ ((V.T @ y) / np.sqrt(w))
    assert relerr(z, zt) < 1e-6
    bd.close()


def test_bd_mixed_precision_step():
    n, phi = 64, 0.2
    pos, L, s, o = _setup(n, phi, seed=6, precision=ls.MIXED)
    w, V = _spectrum(o)
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    nch = 30
    a = po.cheb_coef(nch, 0, e0, e1)
    bd = ls.BD(s, peclet=1.0, n_cheb=nch, eps=1e-5, seed=8, bounds=(e0, e1))
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    info = bd.step(xd, 1e-3)
    y, z = bd.last_noise()
    xo, dto, used, nh, _ = o.bd_mid_F(pos.reshape(-1), y[None], 1.0, 1e-3,
                                      nch, a, e0, e1)
    disp = xo - pos.reshape(-1)
    e = relerr(xd.to_host() - pos.reshape(-1), disp)
    _say(f"  mixed step displacement rel err {e:.2e}")
    assert info.attempts == 1
    assert e < 1e-5
    bd.close()


def test_bd_bounds_enclose_spectrum_n512():
    """N = 512 (n = 1536): Ritz values lie inside the spectrum; the
    integrator's bounds (Ritz values widened) must enclose it, and steps
    then run with an accurate Chebyshev noise"""
    n, phi = 512, 0.2
    pos, L = workloads.configuration(n, phi, seed=7)
    s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0)
    M = s.matrix()
    w = np.linalg.eigvalsh(0.5 * (M + M.T))
    t0, t1 = s.lanczos_bounds(niter=160, seed=1, margin=0.0)
    _say(f"  spectrum [{w[0]:.5f}, {w[-1]:.5f}]  Ritz [{t0:.5f}, {t1:.5f}]")
    assert w[0] * (1 - 1e-9) <= t0 and t1 <= w[-1] * (1 + 1e-9)
    # the spectrum is wide (condition ~70 here): an order-30 fit of
    # 1/sqrt(x) reaches ~1e-4 in z.Mz, so eps is set to match
    bd = ls.BD(s, peclet=1.0, n_cheb=30, eps=1e-3, seed=2)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    for step in range(3):
        info = bd.step(xd, 1e-3)
        _say(f"  step {step}: {info.summary()}")
        assert info.err_cheb < 1e-3 and info.refits == 0
    assert info.eig[0] <= w[0] and info.eig[1] >= w[-1] * (1 - 1e-9)
    bd.close()


@pytest.mark.parametrize("version", [0, 1, 2])
def test_real_cache_products_equal_matrix_free(version):
    """stokes_hip_set_real_cache: the dense real-space matrix gives the
    matrix-free product (FP64 1e-13, MIXED 1e-6 against FP64)"""
    n = 96
    pos, L = workloads.configuration(n, 0.15, seed=12)
    x = np.random.default_rng(13).standard_normal(ls.dim(version) * n)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0)
    s.set_split(30.0)
    y0 = s.atimes(x)
    s.set_real_cache(True)
    y1 = s.atimes(x)
    s.set_real_cache(False)
    e64 = relerr(y1, y0)
    s.set_precision(ls.MIXED)
    s.set_real_cache(True)
    y2 = s.atimes(x)
    s.set_real_cache(False)
    em = relerr(y2, y0)
    print(f"  version {version}: cached FP64 {e64:.2e}, cached MIXED {em:.2e}",
          flush=True)
    assert e64 < 1e-13
    assert em < 1e-6
    s.close()


@pytest.mark.parametrize("version,n,nrhs", [(0, 97, 5), (2, 96, 7), (1, 97, 32),
                                             (2, 97, 1)])
def test_real_cache_block_products(version, n, nrhs):
    """products of several columns (and one column with a row length off the
    16-byte width) with the dense real-space matrix equal the matrix-free
    products (FP64 1e-13, MIXED 1e-6 against FP64)"""
    pos, L = workloads.configuration(n, 0.15, seed=14)
    d = ls.dim(version) * n
    x = np.random.default_rng(15).standard_normal(d * nrhs)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0)
    s.set_split(30.0)
    xd = ls.DeviceArray.from_host(x)
    yd = ls.DeviceArray(d * nrhs)

    def run(prec, cache):
        s.set_precision(prec)
        s.set_real_cache(cache)
        s.atimes_dev(xd, yd, nrhs=nrhs)
        s.sync()
        s.set_real_cache(False)
        return yd.to_host().reshape(nrhs, d)
    y0 = run(ls.FP64, False)
    y1 = run(ls.FP64, True)
    y2 = run(ls.MIXED, True)
    y3 = run(ls.MIXED, True)
    e64 = max(relerr(y1[b], y0[b]) for b in range(nrhs))
    em = max(relerr(y2[b], y0[b]) for b in range(nrhs))
    print(f"  version {version} N={n} nrhs={nrhs}: cached FP64 {e64:.2e}, "
          f"cached MIXED {em:.2e}", flush=True)
    assert e64 < 1e-13
    assert em < 1e-6
    assert np.array_equal(y2, y3)
    s.close()


def test_bd_steps_with_and_without_real_cache():
    """the BD step with the dense real-space matrix for its products at x
    follows the matrix-free step (FP64: 1e-12 of the displacement)"""
    n = 128
    pos, L = workloads.configuration(n, 0.15, seed=14)
    out = []
    for cache in (False, True):
        s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0, rmin=1.0)
        s.set_split(100.0)
        xd = ls.DeviceArray.from_host(pos.reshape(-1))
        bd = ls.BD(s, peclet=1.0, n_cheb=20, eps=1e-2, dt_lim=1e-6,
                   rmin_ol=1.0, seed=5)
        bd.set_real_cache(cache)
        for _ in range(3):
            bd.step(xd, 1e-3)
        out.append(xd.to_host())
        bd.close()
        s.close()
    disp = np.linalg.norm(out[0] - pos.reshape(-1))
    d = np.linalg.norm(out[1] - out[0]) / disp
    print(f"  BD 3 steps, cached vs matrix-free: {d:.2e} of the displacement",
          flush=True)
    assert d < 1e-12


@pytest.mark.parametrize("version,n", [(1, 27), (2, 12)])
def test_bd_ft_fts_steps_match_oracle(version, n):
    """FT / FTS steps (positions and quaternions) against the oracle's step
    fed with the device's draw; FTS with tight GMRES for the noise and the
    step's solves"""
    phi = 0.1
    _say(f"bd parity version={version} N={n} phi={phi}")
    pos, L = workloads.configuration(n, phi, seed=5)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0)
    o = po.System(version, pos, box=(L, L, L), tr=1.0)
    G = o.matrix() if version == 1 else po.ft_mobility_dense(o)
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    nch, pe, dt = 30, 1.0, 1e-3
    a = po.cheb_coef(nch, 0, e0, e1)
    rng = np.random.default_rng(3 + version)
    F = rng.standard_normal(3 * n)
    T = rng.standard_normal(3 * n)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    q = q.reshape(-1)
    Fd = ls.DeviceArray.from_host(F)
    Td = ls.DeviceArray.from_host(T)
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1e-6, dt_lim=1e-6, rmin_ol=1.0,
               seed=9, force_dev=Fd, bounds=(e0, e1))
    bd.set_torque(Td)
    if version == 2:
        s.set_iter(eps=1e-13)
        bd.set_fts_iter(30, 2000, 1e-13)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    qd = ls.DeviceArray.from_host(q)
    x = pos.reshape(-1).copy()
    for step in range(3):
        info = bd.step(xd, dt, qd)
        assert info.attempts == 1 and info.err_cheb < 1e-6
        y, z = bd.last_noise()
        assert y.size == 6 * n
        o.pos[:] = x
        xo, qo, dto, used, nh, err = o.bd_mid(x, y[None], pe, dt, nch, a, e0,
                                              e1, q=q, F=F, T=T)
        assert used == 1 and nh == 0 and dto == info.dt
        xg, qg = xd.to_host(), qd.to_host()
        disp = np.abs(xo - x).max()
        ex = np.abs(xg - xo).max() / disp
        eq = np.abs(qg - qo).max() / np.abs(qo - q).max()
        _say(f"  step {step}: x err / disp {ex:.2e}, q err / dq {eq:.2e}")
        assert ex < (1e-10 if version == 1 else 1e-8)
        assert eq < (1e-10 if version == 1 else 1e-8)
        x, q = xo, qo
    bd.close()


def test_bd_fts_reference_iteration_parameters():
    """FTS step with the reference's GMRES parameters for the noise
    (brownian.c:597-600: restart 20, eps 1e-6) and the Lanczos bounds of
    the integrator: close to the exact-solve oracle step"""
    n, phi = 12, 0.1
    pos, L = workloads.configuration(n, phi, seed=6)
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0)
    s.set_iter(eps=1e-12)
    o = po.System(2, pos, box=(L, L, L), tr=1.0)
    nch, pe, dt = 30, 1.0, 1e-3
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1e-2, dt_lim=1e-6, rmin_ol=1.0,
               seed=4)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    info = bd.step(xd, dt)
    _say(f"fts bd default iteration: {info.summary()}")
    y, _ = bd.last_noise()
    b0, b1 = info.eig
    G = po.ft_mobility_dense(o)
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    assert b0 < w[0] and b1 > w[-1] and b0 > 0.5 * w[0]
    a = po.cheb_coef(nch, 0, b0, b1)
    xo, _, dto, used, nh, err = o.bd_mid(pos, y[None], pe, dt, nch, a, b0, b1)
    disp = np.abs(xo - pos.reshape(-1)).max()
    e = np.abs(xd.to_host() - xo).max() / disp
    _say(f"  |x_gpu - x_oracle| / disp = {e:.2e}")
    assert e < 1e-4
    bd.close()


@pytest.mark.parametrize("version,n,scheme,bb_n", [(0, 27, 1, 2.0), (0, 27, 2, 1.0),
                                                   (1, 27, 1, 4.0), (1, 27, 2, 1.0),
                                                   (2, 12, 1, 2.0)])
def test_bd_schemes_match_oracle(version, n, scheme, bb_n):
    """Banchio-Brady 2003 and Ball-Melrose 1997 steps (positions and, for
    FT / FTS, quaternions) against the oracle's step fed with the device's
    draw"""
    phi = 0.1
    _say(f"bd scheme {scheme} (n = {bb_n}) version={version} N={n}")
    pos, L = workloads.configuration(n, phi, seed=7)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0)
    o = po.System(version, pos, box=(L, L, L), tr=1.0)
    G = (o.matrix() if version < 2 else po.ft_mobility_dense(o))
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    nch, pe, dt = 30, 1.0, 1e-3
    a = po.cheb_coef(nch, 0, e0, e1)
    rng = np.random.default_rng(11 + scheme)
    d3 = 3 * n
    F = rng.standard_normal(d3)
    T = rng.standard_normal(d3) if version > 0 else None
    q = None
    if version > 0:
        q = rng.standard_normal((n, 4))
        q = (q / np.linalg.norm(q, axis=1)[:, None]).reshape(-1)
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1e-6, dt_lim=1e-6, rmin_ol=1.0,
               seed=5, force_dev=ls.DeviceArray.from_host(F), bounds=(e0, e1))
    bd.set_scheme(scheme, bb_n)
    if version > 0:
        bd.set_torque(ls.DeviceArray.from_host(T))
    if version == 2:
        s.set_iter(eps=1e-13)
        bd.set_fts_iter(30, 2000, 1e-13)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    qd = ls.DeviceArray.from_host(q) if q is not None else None
    x = pos.reshape(-1).copy()
    for step in range(2):
        info = bd.step(xd, dt, qd)
        assert info.attempts == 1
        y, _ = bd.last_noise()
        o.pos[:] = x
        xo, qo, dto, used, nh, err = o.bd_mid(x, y[None], pe, dt, nch, a, e0,
                                              e1, q=q, F=F, T=T, scheme=scheme,
                                              bb_n=bb_n)
        assert used == 1 and nh == 0 and dto == info.dt
        disp = np.abs(xo - x).max()
        ex = np.abs(xd.to_host() - xo).max() / disp
        msg = f"  step {step}: x err / disp {ex:.2e}"
        assert ex < (1e-8 if version == 2 else 1e-10)
        if q is not None:
            eq = np.abs(qd.to_host() - qo).max() / np.abs(qo - q).max()
            msg += f", q err / dq {eq:.2e}"
            assert eq < (1e-8 if version == 2 else 1e-10)
        _say(msg)
        x, q = xo, qo
    bd.close()


@pytest.mark.parametrize("version,n,nm", [(0, 27, 20), (1, 27, 20), (2, 12, 8)])
def test_bd_fixed_particles_match_oracle(version, n, nm):
    """mid-point steps with particles nm..n-1 fixed (the noise by
    BD_atimes_mob_FU's zero padding for F / FT, the mixed problem for FTS;
    the drift by the mixed problem of solve_mix_3all, GMRES on the device;
    overlap check over the mobile particles) against the oracle's step fed
    with the device's draw; the fixed particles do not move"""
    phi = 0.1
    _say(f"bd fixed particles version={version} N={n} nm={nm}")
    pos, L = workloads.configuration(n, phi, seed=9)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0, nm=nm)
    o = po.System(version, pos, box=(L, L, L), tr=1.0)
    d = 3 if version == 0 else 6
    if version < 2:
        G = o.matrix()[:d * nm, :d * nm]
    else:
        G = np.stack([o.fts_mix_ft(nm, np.eye(6 * nm)[c])
                      for c in range(6 * nm)], axis=1)
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    nch, pe, dt = 30, 1.0, 1e-3
    a = po.cheb_coef(nch, 0, e0, e1)
    rng = np.random.default_rng(21 + version)
    F = rng.standard_normal(3 * nm)
    T = rng.standard_normal(3 * nm) if version > 0 else None
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1e-6, dt_lim=1e-6, rmin_ol=1.0,
               seed=3, force_dev=ls.DeviceArray.from_host(F), bounds=(e0, e1))
    if version > 0:
        bd.set_torque(ls.DeviceArray.from_host(T))
    s.set_iter(eps=1e-13)
    if version == 2:
        bd.set_fts_iter(30, 2000, 1e-13)
    x = pos.reshape(-1)[:3 * nm].copy()
    xd = ls.DeviceArray.from_host(x)
    for step in range(2):
        info = bd.step(xd, dt)
        assert info.attempts == 1
        y, _ = bd.last_noise()
        assert y.size == d * nm
        full = pos.reshape(-1).copy()
        full[:3 * nm] = x
        o.pos[:] = full
        xo, _, dto, used, nh, err = o.bd_mid(full, y[None], pe, dt, nch, a, e0,
                                             e1, F=F, T=T, nm=nm)
        assert used == 1 and nh == 0 and dto == info.dt
        disp = np.abs(xo - x).max()
        ex = np.abs(xd.to_host() - xo).max() / disp
        _say(f"  step {step}: x err / disp {ex:.2e}")
        assert ex < (1e-8 if version == 2 else 1e-10)
        x = xo
    bd.close()


@pytest.mark.parametrize("version,n", [(0, 27), (1, 27), (2, 12)])
def test_bd_unequal_radii_match_oracle(version, n):
    """mid-point steps of spheres with radii log-uniform on [0.5, 2]: the
    noise operator carries BD_atimes_mob_FU's Global scaling (U / a,
    O / a^3, brownian.c:430-480, 736), the drift does not, the overlap
    check uses a_i + a_j; against the oracle's step fed with the device's
    draw.  F: the Lanczos bounds of the integrator enclose the spectrum of
    the scaled operator"""
    phi = 0.1
    _say(f"bd unequal radii version={version} N={n}")
    pos, a, L = workloads.configuration_poly(n, phi, seed=31)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0, a=a)
    o = po.System(version, pos, box=(L, L, L), tr=1.0, a=a)
    d = 3 if version == 0 else 6
    if version < 2:
        G = o.matrix()
    else:
        G = po.ft_mobility_dense(o)
    sv = np.array([1.0 / a[i] if q < 3 else a[i] ** -3.0
                   for i in range(n) for q in range(d)])
    Gs = np.sqrt(sv)[:, None] * G * np.sqrt(sv)[None, :]
    w = np.linalg.eigvalsh(0.5 * (Gs + Gs.T))
    nch, pe, dt = 30, 1.0, 1e-3
    rng = np.random.default_rng(41 + version)
    F = rng.standard_normal(3 * n)
    T = rng.standard_normal(3 * n) if version > 0 else None
    kw = {} if version == 0 else {"bounds": (0.95 * w[0], 1.05 * w[-1])}
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1e-2, dt_lim=1e-6, rmin_ol=1.0,
               seed=6, force_dev=ls.DeviceArray.from_host(F), **kw)
    if version > 0:
        bd.set_torque(ls.DeviceArray.from_host(T))
    if version == 2:
        s.set_iter(eps=1e-13)
        bd.set_fts_iter(30, 2000, 1e-13)
    x = pos.reshape(-1).copy()
    xd = ls.DeviceArray.from_host(x)
    for step in range(2):
        info = bd.step(xd, dt)
        e0, e1 = info.eig
        if version == 0 and step == 0:
            _say(f"  bounds [{e0:.4g}, {e1:.4g}], spectrum [{w[0]:.4g}, {w[-1]:.4g}]")
            assert e0 < w[0] and e1 > w[-1]
        y, _ = bd.last_noise()
        o.pos[:] = x
        ac = po.cheb_coef(nch, 0, e0, e1)
        xo, _, dto, used, nh, err = o.bd_mid(x, y[None], pe, dt, nch, ac, e0,
                                             e1, F=F, T=T)
        assert used == 1 and nh == 0 and dto == info.dt
        disp = np.abs(xo - x).max()
        ex = np.abs(xd.to_host() - xo).max() / disp
        _say(f"  step {step}: x err / disp {ex:.2e}")
        assert ex < (1e-8 if version == 2 else 1e-10)
        x = xo
    bd.close()


def test_bd_reference_random_stream():
    """with stokes_hip_bd_set_rng_mt19937 the draws are the reference's own
    stream (MT19937 + polar method, KIrand.c:108-157), bitwise equal to the
    oracle's generator for the same seed across steps, and the steps equal
    the oracle's fed with that stream; a dt halving replays the draw"""
    n, phi = 27, 0.1
    pos, L, s, o = _setup(n, phi, seed=8)
    w, V = _spectrum(o)
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    nch, pe, dt = 30, 1.0, 1e-3
    a = po.cheb_coef(nch, 0, e0, e1)
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1e-6, dt_lim=1e-6, rmin_ol=1.0,
               seed=1, bounds=(e0, e1))
    bd.set_rng_mt19937(4357)
    rng = po.Rng(4357)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    x = pos.reshape(-1).copy()
    for step in range(3):
        info = bd.step(xd, dt)
        assert info.attempts == 1
        y, _ = bd.last_noise()
        yo = rng.gaussian(3 * n)
        assert np.array_equal(y, yo)
        o.pos[:] = x
        xo, _, _, used, nh, _ = o.bd_mid(x, yo[None], pe, dt, nch, a, e0, e1)
        disp = np.abs(xo - x).max()
        e = np.abs(xd.to_host() - xo).max() / disp
        _say(f"  step {step}: draws equal the MT19937 stream, x err / disp {e:.2e}")
        assert e < 1e-10
        x = xo
    bd.close()


@pytest.mark.parametrize("version", [0, 1, 2])
def test_bd_retry_after_second_stage_overlap(version):
    """eight spheres, two of them pushed together so that the second stage
    of the first attempt overlaps while dt / 2 stays above dt_lim: the retry
    keeps the noise and takes its first-stage product at the intermediate
    configuration of the failed attempt (REDO_scale, brownian.c:2400,
    2523), an eager attempt on the device; against the oracle's step fed
    with the device's draw"""
    L, gap = 8.0, 0.05
    pos = np.array([[1, 1, 1], [4, 1.2, 0.8], [1.1, 4.2, 1], [4, 4, 1.3],
                    [1, 1.1, 4], [4.2, 1, 4], [1, 4, 4.1], [4, 4.1, 4.2]],
                   dtype=np.float64) + 0.3
    pos[1] = pos[0] + np.array([2.0 + gap, 0.0, 0.0])
    _say(f"bd second-stage retry version={version}")
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0)
    o = po.System(version, pos, box=(L, L, L), tr=1.0)
    d = 3 if version == 0 else 6
    n = 8
    if version < 2:
        G = o.matrix()
    else:
        G = np.stack([o.fts_mix_ft(n, np.eye(6 * n)[c]) for c in range(6 * n)],
                     axis=1)
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    nch, pe, dt = 30, 1e6, 1e-3
    a = po.cheb_coef(nch, 0, e0, e1)
    F = np.zeros(3 * n)
    F[0], F[3] = 100.0, -100.0
    T = np.zeros(3 * n) if version > 0 else None
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1.0, dt_lim=1e-6, rmin_ol=1.0,
               seed=5, force_dev=ls.DeviceArray.from_host(F), bounds=(e0, e1))
    if version > 0:
        bd.set_torque(ls.DeviceArray.from_host(T))
    if version == 2:
        s.set_iter(eps=1e-13)
        bd.set_fts_iter(30, 2000, 1e-13)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    info = bd.step(xd, dt)
    y, _ = bd.last_noise()
    assert y.size == d * n
    xo, _, dto, used, nh, _ = o.bd_mid(pos.reshape(-1), y[None], pe, dt, nch,
                                       a, e0, e1, F=F, T=T)
    _say(f"  halvings gpu {info.halvings} oracle {nh}, dt {info.dt} {dto}, "
         f"overlaps of the last attempt {info.overlaps[0]} {info.overlaps[1]}")
    assert info.halvings == nh == 1 and info.redraws == 0 and used == 1
    assert info.dt == dto == dt / 2 and info.attempts == 2
    xg = xd.to_host()
    disp = np.abs(xo - pos.reshape(-1)).max()
    ex = np.abs(xg - xo).max() / disp
    _say(f"  x err / disp {ex:.2e}")
    assert ex < (1e-8 if version == 2 else 1e-10)
    bd.close()


@pytest.mark.parametrize("version,n,nm,scheme", [
    (0, 27, 27, 0), (0, 27, 27, 1), (1, 27, 20, 0), (2, 12, 12, 0),
    (2, 12, 8, 1)])
def test_bd_imposed_flow_matches_oracle(version, n, nm, scheme):
    """steps in an imposed flow (Ui, Oi, Ei; no shear): every drift product
    is solve_mix_3all in the labo frame (the flow added at the positions of
    the product; fixed particles at rest in the labo frame; FTS with
    E = -Ei), the noise in the rest frame; against the oracle's step fed
    with the device's draw"""
    phi = 0.1
    _say(f"bd imposed flow version={version} N={n} nm={nm} scheme={scheme}")
    pos, L = workloads.configuration(n, phi, seed=12)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0, nm=nm)
    o = po.System(version, pos, box=(L, L, L), tr=1.0)
    Ui, Oi = np.array([0.2, -0.1, 0.3]), np.array([0.1, 0.2, -0.15])
    Ei = np.array([0.05, 0.1, -0.07, 0.02, -0.04])
    s.L.stokes_set_Ui(s.sys, *Ui)
    s.L.stokes_set_Oi(s.sys, *Oi)
    s.L.stokes_set_Ei(s.sys, *Ei)
    o.s.Ui[:], o.s.Oi[:], o.s.Ei[:] = Ui, Oi, Ei
    d = 3 if version == 0 else 6
    if version < 2:
        G = o.matrix()[:d * nm, :d * nm]
    else:
        G = np.stack([o.fts_mix_ft(nm, np.eye(6 * nm)[c])
                      for c in range(6 * nm)], axis=1)
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    nch, pe, dt = 30, 1.0, 1e-3
    a = po.cheb_coef(nch, 0, e0, e1)
    rng = np.random.default_rng(31 + version)
    F = rng.standard_normal(3 * nm)
    T = rng.standard_normal(3 * nm) if version > 0 else None
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1e-6, dt_lim=1e-6, rmin_ol=1.0,
               seed=4, force_dev=ls.DeviceArray.from_host(F), bounds=(e0, e1))
    if version > 0:
        bd.set_torque(ls.DeviceArray.from_host(T))
    if scheme:
        bd.set_scheme(1, 2.0)
    s.set_iter(eps=1e-13)
    if version == 2:
        bd.set_fts_iter(30, 2000, 1e-13)
    x = pos.reshape(-1)[:3 * nm].copy()
    xd = ls.DeviceArray.from_host(x)
    for step in range(2):
        info = bd.step(xd, dt)
        assert info.attempts == 1
        y, _ = bd.last_noise()
        full = pos.reshape(-1).copy()
        full[:3 * nm] = x
        o.pos[:] = full
        xo, _, dto, used, nh, err = o.bd_mid(full, y[None], pe, dt, nch, a, e0,
                                             e1, F=F, T=T, nm=nm,
                                             scheme=scheme, bb_n=2.0)
        assert used == 1 and nh == 0 and dto == info.dt
        disp = np.abs(xo - x).max()
        ex = np.abs(xd.to_host() - xo).max() / disp
        _say(f"  step {step}: x err / disp {ex:.2e}")
        assert ex < (1e-8 if version == 2 else 1e-10)
        x = xo
    bd.close()


@pytest.mark.parametrize("version,n,scheme", [(0, 27, 0), (0, 27, 1),
                                              (1, 27, 2), (2, 12, 0)])
def test_bd_sheared_steps_match_oracle(version, n, scheme):
    """steps of a sheared box (shear_mode 1, stokes_set_shear's imposed
    flow, the cell shift from the step's time: t for the noise and the
    first stage, t + h for the second, BB03's deterministic product at t,
    brownian.c:2393-2468, 2630-2805, 3017-3092) against the oracle's step
    fed with the device's draw"""
    L0, phi, rate, t0, s0 = None, 0.1, 0.8, 1.0, 0.4
    _say(f"bd sheared version={version} N={n} scheme={scheme}")
    pos, L = workloads.configuration(n, phi, seed=14)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0)
    s.L.stokes_set_shear(s.sys, 1, rate)
    o = po.System(version, pos, box=(L, L, L), tr=1.0, shear_mode=1)
    o.s.Oi[:] = [0.0, 0.0, -0.5 * rate]
    o.s.Ei[:] = [0.0, 0.5 * rate, 0.0, 0.0, 0.0]
    o.s.shear_rate, o.s.bd_t0, o.s.bd_s0 = rate, t0, s0

    def shift(tt):
        v = s0 + L * rate * (tt - t0)
        while v < -L / 2:
            v += L
        while v >= L / 2:
            v -= L
        return v

    t = 2.3
    o.s.shear_shift = shift(t)
    d = 3 if version == 0 else 6
    if version < 2:
        G = o.matrix()
    else:
        G = np.stack([o.fts_mix_ft(n, np.eye(6 * n)[c]) for c in range(6 * n)],
                     axis=1)
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.9 * w[0], 1.1 * w[-1]
    nch, pe, dt = 30, 1.0, 1e-3
    a = po.cheb_coef(nch, 0, e0, e1)
    rng = np.random.default_rng(41 + version)
    F = rng.standard_normal(3 * n)
    T = rng.standard_normal(3 * n) if version > 0 else None
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1.0, dt_lim=1e-6, rmin_ol=1.0,
               seed=6, force_dev=ls.DeviceArray.from_host(F), bounds=(e0, e1))
    bd.set_shear_ref(t0, s0)
    if version > 0:
        bd.set_torque(ls.DeviceArray.from_host(T))
    if scheme:
        bd.set_scheme(scheme, 2.0)
    s.set_iter(eps=1e-13)
    if version == 2:
        bd.set_fts_iter(30, 2000, 1e-13)
    x = pos.reshape(-1).copy()
    xd = ls.DeviceArray.from_host(x)
    for step in range(2):
        info = bd.step(xd, dt, t=t)
        assert info.attempts == 1
        y, _ = bd.last_noise()
        o.pos[:] = x
        o.s.bd_t = t
        xo, _, dto, used, nh, err = o.bd_mid(x, y[None], pe, dt, nch, a, e0,
                                             e1, F=F, T=T, scheme=scheme,
                                             bb_n=2.0)
        assert used == 1 and nh == 0 and dto == info.dt
        disp = np.abs(xo - x).max()
        ex = np.abs(xd.to_host() - xo).max() / disp
        _say(f"  step {step} (t = {t}): x err / disp {ex:.2e}")
        assert ex < (1e-8 if version == 2 else 1e-10)
        x = xo
        t += dt
    bd.close()


@pytest.mark.parametrize("scheme,halvings", [(0, 3), (1, 3)])
def test_bd_sheared_retries_match_oracle(scheme, halvings):
    """a sheared F step (shear_mode 1, rate 0.8) that overlaps in its first
    stage and then in its second before it is accepted (the reference's
    random stream, seed 18): the mid-point retry after the second-stage
    overlap starts from the failed attempt's intermediate configuration at
    the shift it left, BB03's from x at the shift of t (brownian.c:2400,
    2523, 2637, 2808); against the oracle's step, itself pinned to the
    reference's control flow composed from its compiled atimes_3all"""
    L, gap, rate, t, t0, s0 = 8.0, 0.03, 0.8, 3.7, 1.0, 0.4
    pos = np.array([[1, 1, 1], [4, 1.2, 0.8], [1.1, 4.2, 1], [4, 4, 1.3],
                    [1, 1.1, 4], [4.2, 1, 4], [1, 4, 4.1], [4, 4.1, 4.2]],
                   dtype=np.float64) + 0.3
    pos[1] = pos[0] + np.array([2.0 + gap, 0.0, 0.0])
    _say(f"bd sheared retries scheme={scheme}")
    s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0)
    s.L.stokes_set_shear(s.sys, 1, rate)
    o = po.System(0, pos, box=(L, L, L), tr=1.0, shear_mode=1)
    o.s.Oi[:] = [0.0, 0.0, -0.5 * rate]
    o.s.Ei[:] = [0.0, 0.5 * rate, 0.0, 0.0, 0.0]
    o.s.shear_rate, o.s.bd_t0, o.s.bd_s0, o.s.bd_t = rate, t0, s0, t

    def shift(tt):
        v = s0 + L * rate * (tt - t0)
        while v < -L / 2:
            v += L
        while v >= L / 2:
            v -= L
        return v

    o.s.shear_shift = shift(t)
    G = o.matrix()
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    nch, pe, dt = 30, 1.0, 1e-3
    a = po.cheb_coef(nch, 0, e0, e1)
    F = np.zeros(24)
    F[0], F[3] = 30.0, -30.0
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1.0, dt_lim=1e-6, rmin_ol=1.0,
               seed=5, force_dev=ls.DeviceArray.from_host(F), bounds=(e0, e1))
    bd.set_rng_mt19937(18)
    bd.set_shear_ref(t0, s0)
    if scheme:
        bd.set_scheme(1, 2.0)
    s.set_iter(eps=1e-13)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    info = bd.step(xd, dt, t=t)
    y = po.Rng(18).gaussian(24)
    xo, _, dto, used, nh, _ = o.bd_mid(pos.reshape(-1), y[None], pe, dt, nch,
                                       a, e0, e1, F=F, scheme=scheme, bb_n=2.0)
    _say(f"  halvings gpu {info.halvings} oracle {nh}, dt {info.dt} {dto}")
    assert info.halvings == nh == halvings and info.redraws == 0 and used == 1
    assert info.dt == dto
    disp = np.abs(xo - pos.reshape(-1)).max()
    ex = np.abs(xd.to_host() - xo).max() / disp
    _say(f"  x err / disp {ex:.2e}")
    assert ex < 1e-10
    bd.close()


@pytest.mark.parametrize("version,n,nm,scheme", [
    (0, 16, 16, 0), (1, 16, 16, 0), (2, 12, 12, 0), (0, 16, 12, 0), (2, 12, 9, 0),
    (0, 16, 16, 2), (1, 16, 16, 2), (0, 16, 16, 1), (1, 16, 12, 1)])
def test_bd_step_with_lubrication(version, n, nm, scheme):
    """flag_lub (stokes_hip_bd_set_lub): the drift of both stages of the
    mid-point step is the lubricated mixed problem of solve_mix_3all
    (solve_mix_lub_3f, ewald-3f.c:646-705 -> solve_mob_lub_3f when nothing
    is fixed; the FT form, ewald-3ft.c:527-805; solve_mix_lub_3fts,
    ewald-3fts.c:1255-1306, with E = 0), particles nm..n-1 fixed at rest.
    Composed in numpy from the device's noise z and the oracle's maps M, L
    at x and at the mid-point (brownian.c:2416-2510): f = (F, T) + fact z
    on the mobile particles, v from (I + M L) v = M g with the mobile
    particles' (F, T) and everything else's velocity / strain known (zero)
    and the rest unknown, x_mid = x + dt/2 U1, U2 the same at x_mid, x_new
    = x + dt U2 (Ball-Melrose 1997, scheme 2: x1 = x + dt U1, x_new = x +
    dt (U1 + U(x1)) / 2; Banchio-Brady 2003, scheme 1 with n = 4: u0 the
    drift of the noise force alone at x, x1 = x + dt u0 / n, x_new = x + dt
    (U_F(x) + u0 + n/2 (U_z(x1) - u0)), brownian.c:2579-2920); within 1e-9
    of the displacement, and apart from the step without lubrication"""
    pe, dt, nch = 1.0, 1e-3, 30
    D = (3, 6, 11)[version]
    dn = 3 if version == 0 else 6
    qs = 6 if version == 2 else D
    pos, L = workloads.configuration(n, 0.2, seed=3, gap=2.02)
    s = ls.Stokes(version, pos, box=(L, L, L), tr=1.0, nm=nm)
    s.set_lub(4.0 + 1e-12, 3.0, 100, 1)
    s.set_iter(max_it=2000, restart=40, eps=1e-13)
    o = po.System(version, pos, box=(L, L, L), tr=1.0)
    if version < 2:
        G = o.matrix()[:dn * nm, :dn * nm]
    elif nm == n:
        G = po.ft_mobility_dense(o)
    else:
        G = np.stack([o.fts_mix_ft(nm, np.eye(6 * nm)[c])
                      for c in range(6 * nm)], axis=1)
    w = np.linalg.eigvalsh(0.5 * (G + G.T))
    e0, e1 = 0.95 * w[0], 1.05 * w[-1]
    rng = np.random.default_rng(5 + version)
    F = rng.standard_normal(3 * nm)
    T = rng.standard_normal(3 * nm)
    Fd = ls.DeviceArray.from_host(F)
    Td = ls.DeviceArray.from_host(T)
    bd = ls.BD(s, peclet=pe, n_cheb=nch, eps=1e-6, dt_lim=1e-6, rmin_ol=1.0,
               seed=7, force_dev=Fd, bounds=(e0, e1))
    if version > 0:
        bd.set_torque(Td)
    if version == 2:
        bd.set_fts_iter(30, 2000, 1e-13)
    bd.set_lub(True)
    if scheme:
        bd.set_scheme(scheme, 4.0)
    full = pos.reshape(-1).copy()
    x = full[:3 * nm].copy()
    xd = ls.DeviceArray.from_host(x)
    info = bd.step(xd, dt)
    assert info.attempts == 1 and info.dt == dt
    y, z = bd.last_noise()
    assert z.size == dn * nm
    fdet = np.zeros((nm, dn))
    fdet[:, :3] = F.reshape(nm, 3)
    if version > 0:
        fdet[:, 3:] = T.reshape(nm, 3)
    fdet = fdet.reshape(-1)
    fz = math.sqrt(2.0 / (pe * dt)) * z

    def drift(xp, lub=True, f=fdet + fz):
        p = full.copy()
        p[:3 * nm] = xp
        o.pos[:] = p
        Mx = o.matrix()
        M = po.trans_mat_ext2ext(n, Mx) if version == 2 else Mx
        Lm = o.lub_columns()
        assert np.linalg.norm(Lm) > 0
        if not lub:
            Lm = np.zeros_like(Lm)
        b = np.zeros((n, D))
        b[:nm, :dn] = f.reshape(nm, dn)
        b = b.reshape(-1)
        P = np.eye(D * n) + M @ Lm
        q, i = np.arange(D * n) % D, np.arange(D * n) // D
        K = np.flatnonzero((i < nm) & (q < qs))
        H = np.flatnonzero(~((i < nm) & (q < qs)))
        A = np.concatenate([P[:, K], -M[:, H]], axis=1)
        sol = np.linalg.solve(A, M[:, K] @ b[K] - P[:, H] @ b[H])
        v = np.empty(D * n)
        v[K] = sol[:K.size]
        v[H] = sol[K.size:]
        return v.reshape(n, D)[:nm, :3].reshape(-1)

    def step(lub):
        if scheme == 1:
            u0 = drift(x, lub, fz)
            u = drift(x + 0.25 * dt * u0, lub, fz)
            return x + dt * (drift(x, lub, fdet) + u0 + 2.0 * (u - u0))
        u1 = drift(x, lub)
        if scheme == 0:
            return x + dt * drift(x + 0.5 * dt * u1, lub)
        return x + 0.5 * dt * (u1 + drift(x + dt * u1, lub))

    xn = step(True)
    xg = xd.to_host()
    disp = np.abs(xn - x).max()
    e = np.abs(xg - xn).max() / disp
    # the same noise without lubrication moves the particles elsewhere
    x0 = step(False)
    d0 = np.abs(x0 - xn).max() / disp
    _say(f"  lubricated step version={version} N={n} nm={nm} scheme={scheme}: "
         f"|x_gpu - x_numpy| "
         f"/ disp = {e:.2e}, without lubrication {d0:.2e}")
    assert e < 1e-9
    assert d0 > 1e-3
    bd.close()
    Fd.free()
    Td.free()
    xd.free()
    s.close()
