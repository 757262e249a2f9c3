"""The C-ABI library (CPU only: no compute calls).

* libstokes_amd/libstokes_hip.so loads and exports every symbol that
  include/*.h declares;
* struct stokes of include/stokes.h has the byte layout of the reference's
  struct stokes (offsets recorded from the reference header in
  tests/golden/golden_layout.json) and of the ctypes mirror;
* host-side set-up (stokes_set_l / stokes_set_xi / xi_by_tratio) gives the
  same parameters and image table as the oracle, bit for bit.
"""
import ctypes as C
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import libstokes_amd as ls
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src):
            name = m.group(1)
            if name in ("if", "for", "while", "sizeof", "return", "void",
                        "double", "int", "extern", "func", "atimes"):
                continue
            names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    lib = ls.lib()
    declared = _declared_functions()
    assert len(declared) >= 50
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(ls.exported_symbols()) <= declared


def _offsets_of(header_dir):
    fields = [k for k in json.load(open(os.path.join(
        ROOT, "tests", "golden", "golden_layout.json"))) if k != "sizeof"]
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "stokes.h"',
           "int main(void){",
           'printf("{\\"sizeof\\": %zu", sizeof(struct stokes));']
    for f in fields:
        src.append(f'printf(", \\"{f}\\": %zu", offsetof(struct stokes, {f}));')
    src += ['printf("}\\n");', "return 0;}"]
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "p.c")
        exe = os.path.join(td, "p")
        open(c, "w").write("\n".join(src))
        subprocess.check_call(["gcc", "-I", header_dir, c, "-o", exe])
        return json.loads(subprocess.check_output([exe]).decode())


def test_struct_stokes_layout_matches_reference():
    ref = json.load(open(os.path.join(ROOT, "tests", "golden",
                                      "golden_layout.json")))
    ours = _offsets_of(os.path.join(ROOT, "include"))
    assert ours == ref
    # the ctypes mirror used by tests and bench
    assert C.sizeof(ls.StokesStruct) == ref["sizeof"]
    for name, off in ref.items():
        if name != "sizeof":
            assert getattr(ls.StokesStruct, name).offset == off, name


@pytest.mark.parametrize("box,tr,eps", [
    ((13.89, 13.89, 13.89), 1.0, 1e-12), ((24.27, 24.27, 24.27), 1e3, 1e-12),
    ((8.0, 9.0, 10.0), 30.0, 1e-8), ((5.733683,) * 3, 1.0, 1e-12)])
def test_setup_matches_oracle_bitwise(box, tr, eps):
    pos = np.zeros((2, 3))
    s = ls.Stokes(0, pos, box=box, tr=tr, eps=eps)
    o = po.System(0, pos, box=box, tr=tr, eps=eps)
    assert s.params() == o.params()
    rl, rm = s.image_table()
    orl, orm = o.image_table()
    assert np.array_equal(rl, orl) and np.array_equal(rm, orm)
    s.close()


def test_shear_shift_window():
    s = ls.Stokes(0, np.zeros((1, 3)), box=(10.0, 8.0, 6.0), tr=1.0)
    L = s.L
    L.stokes_set_shear.argtypes = [C.POINTER(ls.StokesStruct), C.c_int,
                                   C.c_double]
    L.stokes_get_shear_shift.argtypes = [C.POINTER(ls.StokesStruct), C.c_double,
                                         C.c_double, C.c_double]
    L.stokes_set_shear(s.sys, 1, 0.7)
    assert s.s.Oi[2] == -0.35 and s.s.Ei[1] == 0.35
    for t in (0.0, 1.0, 3.7, 11.0):
        sh = L.stokes_get_shear_shift(s.sys, t, 0.0, 0.2)
        want = 0.2 + 8.0 * 0.7 * t
        want = (want + 5.0) % 10.0 - 5.0
        assert sh == pytest.approx(want, abs=1e-12)
        assert -5.0 <= sh < 5.0
    s.close()


def test_scalars_exported_match_oracle():
    lib = ls.lib()
    out = [C.c_double() for _ in range(11)]
    for r in (2.0, 3.3, 7.0):
        for xi in (0.2, 0.7):
            lib.scalars_ewald_real(C.c_int(2), C.c_double(xi), C.c_double(r),
                                   *[C.byref(v) for v in out])
            got = np.array([v.value for v in out])
            want = po.scalars_ewald_real(2, xi, r)
            assert np.allclose(got, want, rtol=1e-13, atol=1e-16)
            lib.scalars_ewald_real_poly(C.c_int(2), C.c_double(xi),
                                        C.c_double(r), C.c_double(1.3),
                                        C.c_double(0.6),
                                        *[C.byref(v) for v in out])
            got = np.array([v.value for v in out])
            want = po.scalars_ewald_real_poly(2, xi, r, 1.3, 0.6)
            assert np.allclose(got, want, rtol=1e-13, atol=1e-16)
    sc = np.zeros(11)
    lib.scalars_nonewald(2, C.c_double(2.7), sc.ctypes.data_as(C.POINTER(C.c_double)))
    want = np.zeros(11)
    po.lib().orc_scalars_nonewald(C.c_int(2), C.c_double(2.7),
                                  want.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.allclose(sc, want, rtol=1e-15, atol=0)


def test_scalar_derivatives_of_the_adjoint():
    """the radial derivatives of xa, ya used by the adjoint kernels equal
    central differences (h = 1e-5) of the oracle's scalars_ewald_real,
    itself bitwise equal to the compiled reference's"""
    lib = ls.lib()
    out = np.zeros(4)
    h = 1e-5
    for xi in (0.1, 0.4, 1.0):
        for r in (0.7, 2.0, 3.3, 6.0):
            lib.stokes_hip_scalars_real_f_deriv(
                C.c_double(xi), C.c_double(r),
                out.ctypes.data_as(C.POINTER(C.c_double)))
            s0 = po.scalars_ewald_real(0, xi, r)[:2]
            fd = (po.scalars_ewald_real(0, xi, r + h)[:2]
                  - po.scalars_ewald_real(0, xi, r - h)[:2]) / (2 * h)
            assert np.allclose(out[:2], s0, rtol=1e-13, atol=1e-16)
            assert np.allclose(out[2:], fd, rtol=1e-7, atol=1e-10)


def test_chebyshev_host_functions_match_oracle():
    lib = ls.lib()
    FUNC = C.CFUNCTYPE(C.c_double, C.c_double)
    f = FUNC(lambda x: 1.0 / np.sqrt(x))
    a = np.zeros(40)
    lib.chebyshev_coef(40, f, C.c_double(0.1), C.c_double(3.0),
                       a.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.allclose(a, po.cheb_coef(40, 0, 0.1, 3.0), rtol=0, atol=1e-15)
    for x in (0.2, 1.0, 2.9):
        assert lib.chebyshev_eval(40, a.ctypes.data_as(C.POINTER(C.c_double)),
                                  0.1, 3.0, x) == pytest.approx(
            po.cheb_eval(a, 0.1, 3.0, x), rel=1e-14)


def test_bd_refuses_shear_without_a_box():
    """shear needs a periodic system: an open system with a shear mode is
    refused before any device work"""
    pos = np.array([[1.0, 1.0, 1.0], [4.0, 1.0, 1.0]])
    s = ls.Stokes(0, pos, tr=1.0)
    out = C.c_void_p()
    s.L.stokes_set_shear(s.sys, 1, 0.5)
    rc = s.L.stokes_hip_bd_create(s.sys, None, 1.0, 20, 1e-3, 1e-6, 1.0, 1,
                                  C.byref(out))
    assert rc == -4 and not out.value                # STOKES_HIP_EUNSUP
    assert b"periodic" in s.L.stokes_hip_last_error()
