"""Trajectory files (stokes-nc.h): written by the library in the netCDF
classic format with 64-bit offsets and read back with scipy's independent
netCDF reader.

* the variable set of stokes_nc_init for every version, with and without
  fixed particles, polydisperse or not, with quaternions, time-changing
  imposed flow, shear and the random-generator state, is the flag table of
  the reference's stokes_nc_init (stokes-nc.c:1445-1598), with each
  variable's dimensions and long_name (stokes-nc.c:716-1366);
* values written by the setters read back bit for bit; unwritten entries
  and records skipped over hold the classic format's fill value;
* the index variables hold 0..n-1 and l is (0, 0, 0) until set;
* an existing file is not overwritten (NC_NOCLOBBER).
CPU only.
"""
import ctypes as C
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

import libstokes_amd as ls

FILL = 9.9692099683868690e+36

FIXED = ["l", "Ui0", "Oi0", "Ei0", "x0", "U0", "O0", "E0", "F0", "T0", "S0",
         "xf0", "Uf0", "Of0", "Ef0", "Ff0", "Tf0", "Sf0", "a", "af"]
RECORD = ["Ui", "Oi", "Ei", "x", "q", "U", "O", "E", "F", "T", "S", "xf", "Uf",
          "Of", "Ef", "Ff", "Tf", "Sf"]


def _lib():
    L = ls.lib()
    P = C.POINTER(C.c_double)
    L.stokes_nc_init.restype = C.c_void_p
    L.stokes_nc_init.argtypes = [C.c_char_p] + [C.c_int] * 8
    L.stokes_nc_x_init.restype = C.c_void_p
    L.stokes_nc_x_init.argtypes = [C.c_char_p, C.c_int]
    L.stokes_nc_free.argtypes = [C.c_void_p]
    L.stokes_nc_set_shear_rate.argtypes = [C.c_void_p, C.c_double]
    for v in FIXED:
        getattr(L, "stokes_nc_set_" + v.lower()).argtypes = [C.c_void_p, P]
    for v in RECORD:
        getattr(L, "stokes_nc_set_" + v.lower()).argtypes = [C.c_void_p,
                                                             C.c_int, P]
    L.stokes_nc_set_time.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.stokes_nc_set_shear_shift.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.stokes_nc_set_rng_state.argtypes = [C.c_void_p, C.c_int,
                                          C.POINTER(C.c_uint32), C.c_int,
                                          C.c_int, C.c_double]
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _expected(version, nf, poly, Q, it, shear, rng):
    """the flag table of stokes_nc_init (stokes-nc.c:1445-1598) as the set
    of variables it defines, with their dimensions"""
    v = {"p": ("p",), "vec": ("vec",), "stt": ("stt",), "quat": ("quat",),
         "time": ("time",), "l": ("vec",), "shear_mode": ()}
    if nf:
        v["pf"] = ("pf",)
    if rng:
        v.update({"rng": ("rng",), "mt": ("time", "rng"), "mti": ("time",),
                  "mt_Ghs": ("time",), "mt_Gs": ("time",)})
    if shear:
        v.update({"shear_rate": (), "shear_shift": ("time",)})
    if it:
        v.update({"Ui": ("time", "vec"), "Oi": ("time", "vec"),
                  "Ei": ("time", "stt")})
    else:
        v.update({"Ui0": ("vec",), "Oi0": ("vec",), "Ei0": ("stt",)})
    v.update({"F0": ("p", "vec"), "x": ("time", "p", "vec"),
              "U": ("time", "p", "vec")})
    if Q:
        v["q"] = ("time", "p", "quat")
    if poly:
        v["a"] = ("p",)
    if version >= 1:
        v.update({"T0": ("p", "vec"), "O": ("time", "p", "vec")})
    if version == 2:
        v.update({"E0": ("p", "stt"), "S": ("time", "p", "stt")})
    if nf:
        v.update({"xf0": ("pf", "vec"), "Uf0": ("pf", "vec"),
                  "Ff": ("time", "pf", "vec")})
        if poly:
            v["af"] = ("pf",)
        if version >= 1:
            v.update({"Of0": ("pf", "vec"), "Tf": ("time", "pf", "vec")})
        if version == 2:
            v.update({"Ef0": ("pf", "stt"), "Sf": ("time", "pf", "stt")})
    return v


LONG = {"x": "positions of particle center", "U": "translational velocities",
        "F0": "forces at t0", "Ff": "forces for fixed particles",
        "shear_shift": "cell shift for shear boundary condition",
        "mt": "random number generator table", "l": "lattice vector",
        "q": "quaternion of particle direction", "af": "radii of particle"}


@pytest.mark.parametrize("version", [0, 1, 2])
@pytest.mark.parametrize("nf", [0, 3])
@pytest.mark.parametrize("extras", [(0, 0, 0, 0, 0), (1, 1, 1, 2, 1)])
def test_variable_set_and_dimensions(tmp_path, version, nf, extras):
    poly, Q, it, shear, rng = extras
    L = _lib()
    fn = str(tmp_path / "t.nc").encode()
    nc = L.stokes_nc_init(fn, 5, nf, version, poly, Q, it, shear, rng)
    assert nc
    L.stokes_nc_free(nc)
    with netcdf_file(fn.decode(), "r", mmap=False) as f:
        want = _expected(version, nf, poly, Q, it, shear, rng)
        assert set(f.variables) == set(want)
        for name, dims in want.items():
            assert f.variables[name].dimensions == dims, name
        for name, text in LONG.items():
            if name in f.variables:
                assert f.variables[name].long_name == text.encode(), name
        assert f.dimensions["p"] == 5 and f.dimensions["vec"] == 3
        assert f.dimensions["stt"] == 5 and f.dimensions["quat"] == 4
        assert f.dimensions["time"] is None
        assert np.array_equal(f.variables["p"][:], np.arange(5))
        assert np.array_equal(f.variables["stt"][:], np.arange(5))
        if nf:
            assert f.dimensions["pf"] == nf
            assert np.array_equal(f.variables["pf"][:], np.arange(nf))
        if rng:
            assert np.array_equal(f.variables["rng"][:], np.arange(624))
        assert f.variables["shear_mode"].getValue() == shear
        assert np.array_equal(f.variables["l"][:], np.zeros(3))


def test_values_round_trip_and_fill(tmp_path):
    L = _lib()
    np_, nf, nt = 6, 2, 4
    fn = str(tmp_path / "traj.nc")
    nc = L.stokes_nc_init(fn.encode(), np_, nf, 2, 1, 1, 0, 1, 1)
    rng = np.random.default_rng(3)
    l = np.array([10.0, 11.0, 12.5])
    L.stokes_nc_set_l(nc, _p(l))
    L.stokes_nc_set_shear_rate(nc, 0.25)
    fixed = {"Ui0": 3, "Oi0": 3, "Ei0": 5, "F0": 3 * np_, "T0": 3 * np_,
             "E0": 5 * np_, "a": np_, "af": nf, "xf0": 3 * nf, "Uf0": 3 * nf,
             "Ef0": 5 * nf}             # Of0 left unwritten
    vals = {k: rng.standard_normal(n) for k, n in fixed.items()}
    for k, v in vals.items():
        getattr(L, "stokes_nc_set_" + k.lower())(nc, _p(v))
    rec = {"x": 3 * np_, "q": 4 * np_, "U": 3 * np_, "O": 3 * np_,
           "S": 5 * np_, "Ff": 3 * nf, "Tf": 3 * nf, "Sf": 5 * nf}
    steps = [0, 1, 3]                  # record 2 skipped: fill values
    rv = {(k, s): rng.standard_normal(n) for k, n in rec.items() for s in steps}
    mt = rng.integers(0, 2**32, 624, dtype=np.uint32)
    for s in steps:
        L.stokes_nc_set_time(nc, s, 0.1 * s)
        L.stokes_nc_set_shear_shift(nc, s, -0.5 * s)
        for k in rec:
            getattr(L, "stokes_nc_set_" + k.lower())(nc, s, _p(rv[(k, s)]))
        L.stokes_nc_set_rng_state(
            nc, s, mt.ctypes.data_as(C.POINTER(C.c_uint32)), 100 + s, s % 2,
            0.5 * s)
    L.stokes_nc_free(nc)
    with netcdf_file(fn, "r", mmap=False) as f:
        assert f.variables["time"].shape == (nt,)
        assert np.array_equal(f.variables["l"][:], l)
        assert f.variables["shear_rate"].getValue() == 0.25
        for k, v in vals.items():
            got = f.variables[k][:].reshape(-1)
            assert np.array_equal(got, v), k
        for s in steps:
            assert f.variables["time"][s] == 0.1 * s
            assert f.variables["shear_shift"][s] == -0.5 * s
            for k in rec:
                got = f.variables[k][s].reshape(-1)
                assert np.array_equal(got, rv[(k, s)]), (k, s)
            assert np.array_equal(
                f.variables["mt"][s].astype(np.int64) & 0xFFFFFFFF,
                mt.astype(np.int64))
            assert f.variables["mti"][s] == 100 + s
            assert f.variables["mt_Ghs"][s] == s % 2
            assert f.variables["mt_Gs"][s] == 0.5 * s
        assert f.variables["time"][2] == FILL
        assert np.all(f.variables["x"][2] == FILL)
        assert f.variables["mti"][2] == -2147483647
        # never written: fill values
        assert f.variables["Of0"].shape == (nf, 3)
        assert np.all(f.variables["Of0"][:] == FILL)


def test_x_only_file_and_no_clobber(tmp_path):
    L = _lib()
    fn = str(tmp_path / "x.nc")
    nc = L.stokes_nc_x_init(fn.encode(), 3)
    x = np.arange(9.0)
    for s in range(3):
        L.stokes_nc_set_time(nc, s, float(s))
        L.stokes_nc_set_x(nc, s, _p(x + s))
    L.stokes_nc_free(nc)
    with netcdf_file(fn, "r", mmap=False) as f:
        assert set(f.variables) == {"p", "vec", "stt", "quat", "time", "l",
                                    "shear_mode", "x"}
        assert np.array_equal(f.variables["x"][:].reshape(3, 9),
                              np.stack([x + s for s in range(3)]))
    # an existing file: the reference's NC_NOCLOBBER error exits
    import subprocess
    import sys
    code = ("import ctypes as C, libstokes_amd as ls; L = ls.lib(); "
            "L.stokes_nc_x_init.restype = C.c_void_p; "
            f"L.stokes_nc_x_init({fn!r}.encode(), 3)")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True,
                       cwd=os.path.dirname(os.path.dirname(
                           os.path.abspath(__file__))))
    assert r.returncode == 1 and b"cannot create" in r.stderr
    assert os.path.getsize(fn) > 0


class _NC(C.Structure):
    """the head of struct stokes_nc: the reference's first int fields"""
    _fields_ = [(n, C.c_int) for n in (
        "id", "np", "p_dim", "p_id", "npf", "pf_dim", "pf_id", "nvec",
        "vec_dim", "vec_id", "nstt", "stt_dim", "stt_id", "nquat",
        "quat_dim", "quat_id", "nrng", "rng_dim", "rng_id", "flag_rng",
        "ntime")]


def _rlib():
    L = _lib()
    P = C.POINTER(C.c_double)
    for f in ("stokes_nc_open", "stokes_nc_reopen"):
        getattr(L, f).restype = C.POINTER(_NC)
        getattr(L, f).argtypes = [C.c_char_p]
    for f in ("stokes_nc_get_array1d", "stokes_nc_get_data0"):
        getattr(L, f).argtypes = [C.c_void_p, C.c_char_p, P]
    L.stokes_nc_get_data.argtypes = [C.c_void_p, C.c_char_p, C.c_int, P]
    L.stokes_nc_get_time.argtypes = [C.c_void_p, P]
    L.stokes_nc_get_time_step.argtypes = [C.c_void_p, C.c_int]
    L.stokes_nc_get_time_step.restype = C.c_double
    L.stokes_nc_get_rng_state.argtypes = [
        C.c_void_p, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_int),
        C.POINTER(C.c_int), P]
    return L


def test_reader_and_append(tmp_path):
    """stokes_nc_open reads what the writer wrote and a file scipy wrote;
    stokes_nc_reopen appends records that scipy then reads"""
    L = _rlib()
    fn = str(tmp_path / "r.nc")
    np_, nf = 4, 2
    nc = L.stokes_nc_init(fn.encode(), np_, nf, 1, 1, 1, 0, 0, 1)
    rng = np.random.default_rng(5)
    a, uf0, f0 = (rng.standard_normal(k) for k in (np_, 3 * nf, 3 * np_))
    L.stokes_nc_set_a(nc, _p(a))
    L.stokes_nc_set_uf0(nc, _p(uf0))
    L.stokes_nc_set_f0(nc, _p(f0))
    xs = [rng.standard_normal(3 * np_) for _ in range(3)]
    qs = [rng.standard_normal(4 * np_) for _ in range(3)]
    mt = rng.integers(0, 2**32, 624, dtype=np.uint32)
    for s in range(3):
        L.stokes_nc_set_time(nc, s, 0.5 * s)
        L.stokes_nc_set_x(nc, s, _p(xs[s]))
        L.stokes_nc_set_q(nc, s, _p(qs[s]))
    L.stokes_nc_set_rng_state(nc, 1, mt.ctypes.data_as(C.POINTER(C.c_uint32)),
                              7, 1, -0.25)
    L.stokes_nc_free(nc)

    r = L.stokes_nc_open(fn.encode())
    h = r.contents
    assert (h.np, h.npf, h.nvec, h.nstt, h.nquat, h.ntime) == (np_, nf, 3, 5,
                                                             4, 3)
    assert h.flag_rng == 1 and h.nrng == 624
    out = np.zeros(3 * np_)
    L.stokes_nc_get_array1d(r, b"a", _p(out))
    assert np.array_equal(out[:np_], a)
    L.stokes_nc_get_data0(r, b"Uf0", _p(out))
    assert np.array_equal(out[:3 * nf], uf0)
    L.stokes_nc_get_data0(r, b"F0", _p(out))
    assert np.array_equal(out, f0)
    q = np.zeros(4 * np_)
    for s in range(3):
        L.stokes_nc_get_data(r, b"x", s, _p(out))
        assert np.array_equal(out, xs[s])
        L.stokes_nc_get_data(r, b"q", s, _p(q))
        assert np.array_equal(q, qs[s])
    t = np.zeros(3)
    L.stokes_nc_get_time(r, _p(t))
    assert np.array_equal(t, [0.0, 0.5, 1.0])
    assert L.stokes_nc_get_time_step(r, 2) == 1.0
    mt2 = np.zeros(624, np.uint32)
    mti, ghs, gs = C.c_int(), C.c_int(), C.c_double()
    L.stokes_nc_get_rng_state(r, 1, mt2.ctypes.data_as(C.POINTER(C.c_uint32)),
                              C.byref(mti), C.byref(ghs), C.byref(gs))
    assert np.array_equal(mt2, mt) and (mti.value, ghs.value, gs.value) == (
        7, 1, -0.25)
    L.stokes_nc_free(r)

    # append a record to the same file, read it back with scipy
    r = L.stokes_nc_reopen(fn.encode())
    x3 = rng.standard_normal(3 * np_)
    L.stokes_nc_set_time(r, 3, 1.5)
    L.stokes_nc_set_x(r, 3, _p(x3))
    assert r.contents.ntime == 4
    L.stokes_nc_free(r)
    with netcdf_file(fn, "r", mmap=False) as f:
        assert f.variables["time"].shape == (4,)
        assert np.array_equal(f.variables["x"][3].reshape(-1), x3)
        assert np.array_equal(f.variables["x"][1].reshape(-1), xs[1])

    # a CDF-1 file written by scipy (32-bit offsets, other attributes)
    fn2 = str(tmp_path / "s.nc")
    with netcdf_file(fn2, "w", version=1) as f:
        f.history = "written by scipy"
        f.createDimension("time", None)        # scipy: the record one first
        for d, n in (("p", 2), ("vec", 3), ("stt", 5), ("quat", 4)):
            f.createDimension(d, n)
        v = f.createVariable("time", "d", ("time",))
        x = f.createVariable("x", "d", ("time", "p", "vec"))
        x.units = "a"
        for s in range(2):
            v[s] = 2.0 * s
            x[s] = np.arange(6.0).reshape(2, 3) + 10 * s
    r = L.stokes_nc_open(fn2.encode())
    assert r.contents.np == 2 and r.contents.ntime == 2
    out = np.zeros(6)
    L.stokes_nc_get_data(r, b"x", 1, _p(out))
    assert np.array_equal(out, np.arange(6.0) + 10)
    assert L.stokes_nc_get_time_step(r, 1) == 2.0
    L.stokes_nc_free(r)
