"""Write tests/golden/golden.npz and golden_layout.json from the compiled
reference hot path (oracle/_ref/libref.so, built by `make -C oracle ref` from
the reference's ewald.c, f.c, ft.c, fts.c, non-ewald.c, chebyshev.c and
KIrand.c).  Run in the build container, where the reference tree exists:

    python tests/golden/make_golden.py

Each case stores its inputs (positions, box, xi, eps, radii, shear, x) and
the reference outputs (atimes_3all y, and the dense matrix for small N).
The Ewald parameters the reference ran with come from the oracle's
restatement of stokes_set_xi (the reference's stokes.c needs the absent
libiter header); they are stored so the product's set-up can be compared.
"""
import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

REF_SRC = "/root/reference/src"

# 9-particle FTS configuration of test/check-mob-fts.c:59-67 (L = 5.733683)
FIX9 = np.array([
    4.841648, 4.345871, 2.411424, 1.484547, 2.931487, 2.321764,
    4.494053, 1.739099, 2.732656, 1.436324, 5.216197, 5.634981,
    2.707117, 0.577389, 2.489404, 3.502632, 5.234916, 5.542215,
    4.615775, 3.146068, 0.080511, 5.510249, 1.060571, 0.710376,
    1.143043, 2.055843, 4.194753]).reshape(9, 3)
FIX9_L = 5.733683


def sc_l(phi):
    """box of the SC tests: l = (pi / 0.75 / phi)^(1/3) (check-ewald.c:69)"""
    return (math.pi / 0.75 / phi) ** (1.0 / 3.0)


def rsa(n, L, gap, rng):
    """random sequential addition with minimum-image centre distance >= gap"""
    pos = []
    while len(pos) < n:
        p = rng.uniform(0, L, 3)
        ok = True
        for q in pos:
            d = p - q
            d -= L * np.round(d / L)
            if d @ d < gap * gap:
                ok = False
                break
        if ok:
            pos.append(p)
    return np.array(pos)


def cases():
    rng = np.random.default_rng(20261203)
    out = []
    vn = {0: "F", 1: "FT", 2: "FTS"}
    # N=1 simple cubic (check-ewald-poly.c:64-85), x = 1, tr 1 / 4 / 10
    for v in (0, 1, 2):
        for tr in (1.0, 4.0, 10.0):
            l = sc_l(0.2)
            out.append(dict(name=f"sc1_{vn[v]}_tr{int(tr)}", version=v,
                            pos=np.zeros((1, 3)), box=(l, l, l), tr=tr,
                            x=np.ones(po.dim(v)), matrix=True))
    # N=2 in an elongated box (check-ewald-poly.c:195-231)
    for v in (0, 1, 2):
        for d in (0, 1, 2):
            l = sc_l(0.2)
            box = [l, l, l]
            box[d] = 2 * l
            pos = np.zeros((2, 3))
            pos[1, d] = l
            out.append(dict(name=f"sc2_dir{d}_{vn[v]}", version=v, pos=pos,
                            box=tuple(box), tr=1.0,
                            x=np.ones(2 * po.dim(v)), matrix=True))
    # N=8 simple cubic in a 2l box (check-ewald.c:93-127)
    for v in (0, 1, 2):
        l = sc_l(0.2)
        pos = np.array([[i, j, k] for (i, j, k) in
                        [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                         (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]],
                       dtype=float) * l
        out.append(dict(name=f"sc8_{vn[v]}", version=v, pos=pos,
                        box=(2 * l, 2 * l, 2 * l), tr=1.0,
                        x=np.ones(8 * po.dim(v)), matrix=True))
    # 9-particle fixture, random forces
    for v in (0, 1, 2):
        out.append(dict(name=f"fix9_{vn[v]}", version=v, pos=FIX9,
                        box=(FIX9_L,) * 3, tr=1.0,
                        x=rng.standard_normal(9 * po.dim(v)), matrix=True))
        out.append(dict(name=f"fix9_tr100_{vn[v]}", version=v, pos=FIX9,
                        box=(FIX9_L,) * 3, tr=100.0,
                        x=rng.standard_normal(9 * po.dim(v)), matrix=False))
    # polydisperse, shear, rmin clamp, open boundary
    radii = rng.uniform(0.6, 1.4, 9)
    for v in (0, 1, 2):
        out.append(dict(name=f"fix9_poly_{vn[v]}", version=v, pos=FIX9,
                        box=(FIX9_L,) * 3, tr=1.0, a=radii,
                        x=rng.standard_normal(9 * po.dim(v)), matrix=True))
        for sm in (1, 2):
            out.append(dict(name=f"fix9_shear{sm}_{vn[v]}", version=v,
                            pos=FIX9, box=(FIX9_L,) * 3, tr=1.0,
                            shear_mode=sm, shear_shift=0.37 * FIX9_L,
                            x=rng.standard_normal(9 * po.dim(v)),
                            matrix=(v == 2)))
        out.append(dict(name=f"fix9_rmin_{vn[v]}", version=v, pos=FIX9,
                        box=(FIX9_L,) * 3, tr=1.0, rmin=1.3,
                        x=rng.standard_normal(9 * po.dim(v)), matrix=False))
        out.append(dict(name=f"rpy_fix9_{vn[v]}", version=v, pos=FIX9,
                        box=None, x=rng.standard_normal(9 * po.dim(v)),
                        matrix=True))
        out.append(dict(name=f"rpy_fix9_poly_{vn[v]}", version=v, pos=FIX9,
                        box=None, a=radii,
                        x=rng.standard_normal(9 * po.dim(v)), matrix=True))
    # the N = 64 workload: phi = 0.1, RSA with gap 2.1, F/T/S ~ N(0,1)
    L = (64 * 4.0 * math.pi / 3.0 / 0.1) ** (1.0 / 3.0)
    pos64 = rsa(64, L, 2.1, rng)
    for v in (0, 2):
        out.append(dict(name=f"c1_{vn[v]}", version=v, pos=pos64,
                        box=(L, L, L), tr=1.0,
                        x=rng.standard_normal(64 * po.dim(v)), matrix=False))
    return out


def layout_offsets():
    """offsetof() of every struct stokes field in the reference header"""
    fields = [
        "version", "np", "nm", "pos", "a", "rmin", "twobody_nmax",
        "twobody_lub", "poly_table", "twobody_f_list", "slip", "slip_a",
        "slip_G32", "slip_G30", "slip_G52", "slip_table",
        "twobody_slip_f_list", "Ui", "Oi", "Ei", "shear_mode", "shear_rate",
        "shear_shift", "periodic", "ewald_eps", "rmax2", "rmaxx", "rmaxy",
        "rmaxz", "kmax", "kmaxx", "kmaxy", "kmaxz", "xi", "xi2", "xiaspi",
        "xia2", "pivol", "lx", "ly", "lz", "ilx", "ily", "ilz", "self_a",
        "self_c", "self_m", "flag_table", "nr", "rlx", "rly", "rlz", "rmx",
        "rmy", "rmz", "nk", "ex", "ey", "ez", "k", "k1", "k2", "k3", "ya",
        "yb", "yc", "yg", "yh", "ym", "lubmin2", "lubmax", "cpu1", "cpu2",
        "cpu3", "it"]
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "stokes.h"',
           "int main(void){",
           'printf("{\\"sizeof\\": %zu", sizeof(struct stokes));']
    for f in fields:
        src.append(f'printf(", \\"{f}\\": %zu", offsetof(struct stokes, {f}));')
    src += ['printf("}\\n");', "return 0;}"]
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "probe.c")
        exe = os.path.join(td, "probe")
        open(c, "w").write("\n".join(src))
        subprocess.check_call(["gcc", "-I", REF_SRC, c, "-o", exe])
        return json.loads(subprocess.check_output([exe]).decode())


def main():
    if not po.ref_available():
        sys.exit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    store = {}
    meta = []
    for c in cases():
        v = c["version"]
        s = po.System(v, c["pos"], box=c.get("box"), tr=c.get("tr"),
                      a=c.get("a"), rmin=c.get("rmin", 0.0),
                      shear_mode=c.get("shear_mode", 0),
                      shear_shift=c.get("shear_shift", 0.0))
        y = s.ref_atimes(c["x"])
        key = c["name"]
        store[f"{key}/pos"] = np.asarray(c["pos"], float)
        store[f"{key}/x"] = np.asarray(c["x"], float)
        store[f"{key}/y"] = y
        if c.get("a") is not None:
            store[f"{key}/a"] = np.asarray(c["a"], float)
        if c["matrix"]:
            store[f"{key}/mat"] = s.ref_matrix()
        m = dict(name=key, version=v, box=c.get("box"), tr=c.get("tr"),
                 rmin=c.get("rmin", 0.0), shear_mode=c.get("shear_mode", 0),
                 shear_shift=c.get("shear_shift", 0.0),
                 poly=c.get("a") is not None, matrix=c["matrix"])
        if c.get("box") is not None:
            m["params"] = s.params()
        meta.append(m)
    # Chebyshev on the matrix of test/check-brownian.c:101-110 (n = 10)
    R = po.ref()
    n = 10
    A = np.eye(n)
    for i in range(n):
        for j in range(i + 1, n):
            A[i, j] = A[j, i] = 1.0 / (i + j + 1)
    ev = np.linalg.eigvalsh(A)
    e0, e1 = 0.9 * ev[0], 1.1 * ev[-1]
    import ctypes as C
    for fid, nch in ((0, 50), (1, 30)):
        a = np.zeros(nch)
        R.ref_cheb_coef(C.c_int(nch), C.c_int(fid), C.c_double(e0),
                        C.c_double(e1), a.ctypes.data_as(C.POINTER(C.c_double)))
        y = np.ones(n)
        z = np.zeros(n)
        R.ref_cheb_eval_dense(
            C.c_int(nch), a.ctypes.data_as(C.POINTER(C.c_double)), C.c_int(n),
            A.ctypes.data_as(C.POINTER(C.c_double)),
            y.ctypes.data_as(C.POINTER(C.c_double)),
            z.ctypes.data_as(C.POINTER(C.c_double)), C.c_double(e0),
            C.c_double(e1))
        store[f"cheb{fid}/A"] = A
        store[f"cheb{fid}/a"] = a
        store[f"cheb{fid}/z"] = z
        store[f"cheb{fid}/bounds"] = np.array([e0, e1])
    # chebyshev_error_minvsqrt / _Rsqrt (chebyshev.c:241-310) on the same
    # matrix: low-order fits (errors of order 1e-3) and unrelated vectors
    rng = np.random.default_rng(77)
    for fid, nch in ((0, 5), (1, 4)):
        a = np.zeros(nch)
        R.ref_cheb_coef(C.c_int(nch), C.c_int(fid), C.c_double(e0),
                        C.c_double(e1), a.ctypes.data_as(C.POINTER(C.c_double)))
        y = rng.standard_normal(n)
        z = np.zeros(n)
        R.ref_cheb_eval_dense(
            C.c_int(nch), a.ctypes.data_as(C.POINTER(C.c_double)), C.c_int(n),
            A.ctypes.data_as(C.POINTER(C.c_double)),
            y.ctypes.data_as(C.POINTER(C.c_double)),
            z.ctypes.data_as(C.POINTER(C.c_double)), C.c_double(e0),
            C.c_double(e1))
        zr = rng.standard_normal(n)
        store[f"cheberr{fid}/y"] = y
        store[f"cheberr{fid}/z"] = np.stack([z, zr])
        store[f"cheberr{fid}/err"] = np.array(
            [po.ref_cheb_error_dense(fid, A, y, zz) for zz in (z, zr)])
    # MT19937 + polar Gaussian (KIrand.c), seeds 0 and 5489
    for seed in (0, 5489):
        h = R.ref_rng_new(seed)
        store[f"rng{seed}/int32"] = np.array(
            [R.ref_rng_int32(h) for _ in range(16)], dtype=np.uint64)
        R.ref_rng_free(h)
        h = R.ref_rng_new(seed)
        store[f"rng{seed}/gauss"] = np.array(
            [R.ref_rng_gaussian(h) for _ in range(16)])
        R.ref_rng_free(h)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **store)
    with open(os.path.join(HERE, "golden_cases.json"), "w") as f:
        json.dump(meta, f, indent=1)
    with open(os.path.join(HERE, "golden_layout.json"), "w") as f:
        json.dump(layout_offsets(), f, indent=1)
    print(f"wrote {len(meta)} operator cases")


if __name__ == "__main__":
    main()
