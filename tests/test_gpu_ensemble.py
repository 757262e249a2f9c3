"""Ensembles of independent systems in one graph launch
(stokes_hip_ensemble_*): every box against the oracle on its own.

Boxes of uneven N with radii log-uniform on [0.5, 2] at volume fraction
0.25 (the batched-ensemble workload, at sizes the oracle finishes in
seconds), FTS.  Bars: FP64 1e-12, MIXED 1e-6 per box (normwise); repeated
launches bitwise equal; moved positions are picked up by the next launch.
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
