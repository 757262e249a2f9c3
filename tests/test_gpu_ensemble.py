"""Ensembles of independent systems in one graph launch of batched kernels
(stokes_hip_ensemble_*): every box against the oracle on its own.

Boxes of uneven N with radii log-uniform on [0.5, 2] at volume fraction
0.25 (the batched-ensemble workload, at sizes the oracle finishes in
seconds), FTS.  Bars: FP64 1e-12, MIXED 1e-6 per box (normwise); repeated
launches bitwise equal; moved positions are picked up by the next launch;
every tile shape and per-box splitting gives the same products to the
bars; the batched FP64 lubrication correction equals each box's own
stokes_hip_lub_dev to 1e-14 (same pairs in the same order and the same
formulas, compiled into another kernel) and the oracle to 1e-12, also after
the near pairs outgrow the pair list.
"""
import numpy as np
import pytest

import libstokes_amd as ls
from libstokes_amd import workloads
import pyoracle as po
from conftest import relerr

pytestmark = pytest.mark.gpu

SIZES = (20, 35, 50, 64, 80, 100)


def _say(msg):
    print(msg, flush=True)


def _boxes(precision):
    boxes = []
    for k, n in enumerate(SIZES):
        pos, a, L = workloads.configuration_poly(n, 0.25, seed=100 + k)
        boxes.append((pos, a, L))
    systems = [ls.Stokes(2, pos, box=(L, L, L), tr=1.0, a=a,
                         precision=precision) for pos, a, L in boxes]
    return boxes, systems


@pytest.mark.parametrize("precision,tol", [(ls.FP64, 1e-12), (ls.MIXED, 1e-6)])
def test_ensemble_matches_oracle_per_box(precision, tol):
    boxes, systems = _boxes(precision)
    ens = ls.Ensemble(systems)
    off = ens.offsets
    pos_all = np.concatenate([b[0].reshape(-1) for b in boxes])
    rng = np.random.default_rng(3)
    x_all = rng.standard_normal(11 * off[-1])
    posd = ls.DeviceArray.from_host(pos_all)
    xd = ls.DeviceArray.from_host(x_all)
    yd = ls.DeviceArray(11 * off[-1])
    ens.atimes_dev(posd, xd, yd)
    y1 = yd.to_host()
    ens.atimes_dev(posd, xd, yd)
    y2 = yd.to_host()
    assert np.array_equal(y1, y2)
    for b, (pos, a, L) in enumerate(boxes):
        o = po.System(2, pos, box=(L, L, L), tr=1.0, a=a)
        sl = slice(11 * off[b], 11 * off[b + 1])
        e = relerr(y1[sl], o.atimes(x_all[sl]))
        _say(f"  box {b} N={SIZES[b]}: rel err {e:.2e}")
        assert e < tol
    # move the particles of box 2: the next launch sees the new positions
    pos2 = boxes[2][0] + 0.01 * rng.standard_normal(boxes[2][0].shape)
    pos_all[3 * off[2]:3 * off[3]] = pos2.reshape(-1)
    ls.check(ls.lib().stokes_hip_memcpy_h2d(
        posd.ptr, pos_all.ctypes.data, pos_all.nbytes), "h2d")
    ens.atimes_dev(posd, xd, yd)          # a replay of the same graph
    y3 = yd.to_host()
    o = po.System(2, pos2, box=(boxes[2][2],) * 3, tr=1.0, a=boxes[2][1])
    sl = slice(11 * off[2], 11 * off[3])
    assert relerr(y3[sl], o.atimes(x_all[sl])) < tol
    ens.close()
    for s in systems:
        s.close()


def test_ensemble_autotune_keeps_results():
    """the per-box splitting autotuner (stokes_hip_ensemble_autotune)
    changes only the splitting: every box still equals the oracle at the
    reference's tr = 1 (FP64 1e-12)"""
    boxes, systems = _boxes(ls.FP64)
    ens = ls.Ensemble(systems)
    off = ens.offsets
    pos_all = np.concatenate([b[0].reshape(-1) for b in boxes])
    x_all = np.random.default_rng(4).standard_normal(11 * off[-1])
    posd = ls.DeviceArray.from_host(pos_all)
    xd = ls.DeviceArray.from_host(x_all)
    yd = ls.DeviceArray(11 * off[-1])
    trs = ens.autotune(trs=[10.0, 100.0, 1000.0], reps=2)
    _say(f"  tuned splittings {trs}")
    assert set(trs) <= {10.0, 100.0, 1000.0}
    ens.atimes_dev(posd, xd, yd)
    y = yd.to_host()
    for b, (pos, a, L) in enumerate(boxes):
        o = po.System(2, pos, box=(L, L, L), tr=1.0, a=a)
        sl = slice(11 * off[b], 11 * off[b + 1])
        assert relerr(y[sl], o.atimes(x_all[sl])) < 1e-12
    ens.close()
    for s in systems:
        s.close()


@pytest.mark.parametrize("precision,tol", [(ls.FP64, 1e-12), (ls.MIXED, 1e-6)])
def test_ensemble_tiles_and_mixed_splittings(precision, tol):
    """tile shapes (chunks per real-space wave, particles per structure-
    factor split, k per back-projection split) and a different splitting
    per box change only the summation order"""
    boxes, systems = _boxes(precision)
    ens = ls.Ensemble(systems)
    off = ens.offsets
    pos_all = np.concatenate([b[0].reshape(-1) for b in boxes])
    x_all = np.random.default_rng(8).standard_normal(11 * off[-1])
    posd = ls.DeviceArray.from_host(pos_all)
    xd = ls.DeviceArray.from_host(x_all)
    yd = ls.DeviceArray(11 * off[-1])
    want = []
    for b, (pos, a, L) in enumerate(boxes):
        o = po.System(2, pos, box=(L, L, L), tr=1.0, a=a)
        want.append(o.atimes(x_all[11 * off[b]:11 * off[b + 1]]))
    trs = [0.0, 10.0, 100.0, 1000.0, 3.0, 300.0]
    for b, t in enumerate(trs):
        ens.set_split(t, box=b)
    for jw, jper, kper in [(1, 16, 16), (2, 256, 64), (4, 1024, 256)]:
        ens.set_tiles(jw, jper, kper)
        assert ens.tiles() == (jw, jper, kper)
        ens.atimes_dev(posd, xd, yd)
        y = yd.to_host()
        for b in range(len(boxes)):
            e = relerr(y[11 * off[b]:11 * off[b + 1]], want[b])
            assert e < tol, (jw, jper, kper, b, e)
        _say(f"  tiles ({jw}, {jper}, {kper}): worst box "
             f"{max(relerr(y[11 * off[b]:11 * off[b + 1]], want[b]) for b in range(len(boxes))):.2e}")
    ms = ens.phase_ms(posd, xd, yd)
    assert np.all(ms >= 0) and ms[0] > 0
    ens.close()
    for s in systems:
        s.close()


def _lub_boxes(sizes, seed0):
    boxes, systems = [], []
    for k, n in enumerate(sizes):
        pos, a, L = workloads.configuration_poly(n, 0.25, seed=seed0 + k)
        boxes.append((pos, a, L))
        s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, a=a, precision=ls.MIXED)
        s.set_lub(4.0 + 1e-12, 3.0, 100, 1)
        systems.append(s)
    return boxes, systems


def test_ensemble_lubrication_matches_boxes():
    """the batched lubrication correction (lubmax 3: gaps below one radius,
    series to order 100) against each box's own device correction (1e-14)
    and the oracle (1e-12); with the mobility product in the same launch,
    both outputs equal the separate launches bit for bit"""
    boxes, systems = _lub_boxes((30, 64, 20, 100), 300)
    ens = ls.Ensemble(systems)
    off = ens.offsets
    pos_all = np.concatenate([b[0].reshape(-1) for b in boxes])
    x_all = np.random.default_rng(9).standard_normal(11 * off[-1])
    posd = ls.DeviceArray.from_host(pos_all)
    xd = ls.DeviceArray.from_host(x_all)
    yl = ls.DeviceArray(11 * off[-1])
    ens.lub_dev(posd, xd, yl)
    ylub = yl.to_host()
    for b, (pos, a, L) in enumerate(boxes):
        sl = slice(11 * off[b], 11 * off[b + 1])
        own = systems[b].lub(x_all[sl])
        assert relerr(ylub[sl], own) <= 1e-14, b
        if b < 2:
            o = po.System(2, pos, box=(L, L, L), tr=1.0, a=a)
            want = o.lub(x_all[sl])
            e = relerr(ylub[sl], want)
            _say(f"  box {b} N={len(pos)}: lubrication vs oracle {e:.2e}")
            assert np.linalg.norm(want) > 0 and e <= 1e-12
    ym = ls.DeviceArray(11 * off[-1])
    ens.atimes_dev(posd, xd, ym)
    y_alone = ym.to_host()
    y2 = ls.DeviceArray(11 * off[-1])
    yl2 = ls.DeviceArray(11 * off[-1])
    ens.atimes_lub_dev(posd, xd, y2, yl2)
    assert np.array_equal(y2.to_host(), y_alone)
    assert np.array_equal(yl2.to_host(), ylub)
    ens.close()
    for s in systems:
        s.close()


def test_ensemble_lubrication_list_grows():
    """after a first launch sized the near-pair list, box 1 is squeezed into
    a cluster (every pair within lubmax: 2450 directed pairs against a list
    sized for 1639): the launch sees more pairs than
    the list holds, grows it, runs again and still equals the box's own
    correction (1e-14)"""
    boxes, systems = _lub_boxes((40, 50), 400)
    ens = ls.Ensemble(systems)
    off = ens.offsets
    pos_all = np.concatenate([b[0].reshape(-1) for b in boxes])
    x_all = np.random.default_rng(10).standard_normal(11 * off[-1])
    posd = ls.DeviceArray.from_host(pos_all)
    xd = ls.DeviceArray.from_host(x_all)
    yl = ls.DeviceArray(11 * off[-1])
    ens.lub_dev(posd, xd, yl)
    pos1, a1, L1 = boxes[1]
    rng = np.random.default_rng(11)
    squeezed = pos1[0] + rng.uniform(-0.4, 0.4, pos1.shape)
    pos_all[3 * off[1]:3 * off[2]] = squeezed.reshape(-1)
    ls.check(ls.lib().stokes_hip_memcpy_h2d(
        posd.ptr, pos_all.ctypes.data, pos_all.nbytes), "h2d")
    ens.lub_dev(posd, xd, yl)
    y = yl.to_host()
    systems[1].set_pos(squeezed)
    own = systems[1].lub(x_all[11 * off[1]:11 * off[2]])
    assert relerr(y[11 * off[1]:11 * off[2]], own) <= 1e-14
    own0 = systems[0].lub(x_all[:11 * off[1]])
    assert relerr(y[:11 * off[1]], own0) <= 1e-14
    ens.close()
    for s in systems:
        s.close()
