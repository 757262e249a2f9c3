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


# ---- BD parameter files (stokes_nc_set_by_params and its readers) ----

class _Params:
    """the arguments of stokes_nc_set_by_params for one system"""

    def __init__(self, version, np_, nm, poly, periodic, shear_mode, seed):
        r = np.random.default_rng(seed)
        nf = np_ - nm
        self.version, self.np, self.nm, self.nf = version, np_, nm, nf
        self.a = r.uniform(0.5, 2.0, np_) if poly else None
        self.periodic = periodic
        self.Ui, self.Oi, self.Ei = r.normal(size=3), r.normal(size=3), \
            r.normal(size=5)
        self.F, self.T, self.E = r.normal(size=3 * nm), \
            r.normal(size=3 * nm), r.normal(size=5 * nm)
        self.uf, self.of, self.ef = r.normal(size=3 * max(nf, 1)), \
            r.normal(size=3 * max(nf, 1)), r.normal(size=5 * max(nf, 1))
        self.xf = r.normal(size=3 * max(nf, 1))
        self.lat = np.array([10.0, 11.0, 12.0]) if periodic else np.zeros(3)
        self.shear_mode = shear_mode
        self.shear_rate = 0.75 if shear_mode else 0.0

    def system(self):
        s = _new_sys(self.version, self.np, self.nm)
        st = s.contents
        if self.a is not None:
            ls.lib().stokes_set_radius(s, _p(self.a))
        if self.periodic:
            ls.lib().stokes_set_l(s, *self.lat)
            st.periodic = 1
        st.shear_mode = self.shear_mode
        st.shear_rate = self.shear_rate
        return s

    def args(self):
        return [_p(v) for v in (self.Ui, self.Oi, self.Ei, self.F, self.T,
                                self.E, self.uf, self.of, self.ef, self.xf,
                                self.lat)]


def _new_sys(version, np_, nm):
    L = ls.lib()
    s = L.stokes_init()
    s.contents.version = version
    L.stokes_set_np(s, np_, nm)
    return s


def _param_lib():
    L = _lib()
    P = C.POINTER(C.c_double)
    L.stokes_nc_set_by_params.restype = C.c_void_p
    L.stokes_nc_set_by_params.argtypes = ([C.c_char_p, C.c_void_p, C.c_int]
                                          + [P] * 11
                                          + [C.c_int, C.c_double, C.c_int])
    L.stokes_nc_check_params.restype = C.c_int
    L.stokes_nc_check_params.argtypes = ([C.c_void_p, C.c_void_p, C.c_int]
                                         + [P] * 11 + [C.c_double])
    L.stokes_nc_get_version.restype = C.c_int
    L.stokes_nc_get_version.argtypes = [C.c_void_p]
    L.stokes_nc_get_params.argtypes = ([C.c_void_p, C.c_void_p,
                                        C.POINTER(C.c_int)] + [P] * 11
                                       + [C.POINTER(C.c_int), P,
                                          C.POINTER(C.c_int)])
    L.stokes_nc_open.restype = C.c_void_p
    L.stokes_nc_open.argtypes = [C.c_char_p]
    return L


CASES = [(0, 6, 6, False, True, 0), (1, 7, 5, True, True, 0),
         (2, 5, 5, True, False, 0), (2, 8, 6, False, True, 1),
         (1, 4, 4, False, True, 2), (0, 9, 3, True, False, 0)]


@pytest.mark.parametrize("version,np_,nm,poly,periodic,shear", CASES)
def test_params_file_round_trip(tmp_path, version, np_, nm, poly, periodic,
                                shear):
    """stokes_nc_set_by_params writes the BD parameters with the constant
    imposed flow and the flag table of stokes_nc_init (stokes-nc.c:
    2731-2841); scipy reads the entries back bit for bit; the library's
    stokes_nc_check_params (stokes-nc.c:2869-3188) finds no mismatch, and
    one for each entry changed; stokes_nc_get_version and
    stokes_nc_get_params (stokes-nc.c:3191-3540) return the version, sizes,
    radii, lattice, shear and every vector written"""
    L = _param_lib()
    prm = _Params(version, np_, nm, poly, periodic, shear, seed=np_ + nm)
    sys = prm.system()
    fn = str(tmp_path / "params.nc").encode()
    nc = L.stokes_nc_set_by_params(fn, C.cast(sys, C.c_void_p), 1, *prm.args(),
                                   shear, prm.shear_rate, 0)
    assert nc
    L.stokes_nc_free(nc)
    nf = np_ - nm
    with netcdf_file(fn.decode(), "r", mmap=False) as f:
        v = f.variables
        assert np.array_equal(v["Ui0"][:], prm.Ui)
        assert np.array_equal(v["Oi0"][:], prm.Oi)
        assert np.array_equal(v["Ei0"][:], prm.Ei)
        assert np.array_equal(v["F0"][:].reshape(-1), prm.F)
        assert ("T0" in v) == (version >= 1) and ("E0" in v) == (version == 2)
        if version >= 1:
            assert np.array_equal(v["T0"][:].reshape(-1), prm.T)
        if version == 2:
            assert np.array_equal(v["E0"][:].reshape(-1), prm.E)
        assert np.array_equal(v["l"][:], prm.lat)
        assert "q" in v and "Ui" not in v
        if poly:
            assert np.array_equal(v["a"][:], prm.a[:nm])
        if nf:
            assert np.array_equal(v["xf0"][:].reshape(-1), prm.xf[:3 * nf])
            assert np.array_equal(v["Uf0"][:].reshape(-1), prm.uf[:3 * nf])
            if poly:
                assert np.array_equal(v["af"][:], prm.a[nm:])
        assert int(v["shear_mode"].getValue()) == shear
        if shear:
            assert float(v["shear_rate"].getValue()) == prm.shear_rate
    nc = L.stokes_nc_open(fn)
    sp = C.cast(sys, C.c_void_p)
    assert L.stokes_nc_check_params(nc, sp, 1, *prm.args(), 1e-12) == 0
    assert L.stokes_nc_get_version(nc) == version
    f0, e4 = prm.F[0], prm.Ei[4]
    prm.F[0] += 1.0
    assert L.stokes_nc_check_params(nc, sp, 1, *prm.args(), 1e-12) == 1
    prm.F[0] = f0
    prm.Ei[4] += 1e-3
    assert L.stokes_nc_check_params(nc, sp, 1, *prm.args(), 1e-12) == 1
    assert L.stokes_nc_check_params(nc, sp, 1, *prm.args(), 1e-2) == 0
    prm.Ei[4] = e4
    # read everything back into a fresh system and fresh arrays
    got = _new_sys(0, 1, 1)
    q, sm, rng = C.c_int(), C.c_int(), C.c_int()
    sr = C.c_double()
    out = _Params(version, np_, nm, poly, periodic, shear, seed=99)
    for name in ("Ui", "Oi", "Ei", "F", "T", "E", "uf", "of", "ef", "xf",
                 "lat"):
        setattr(out, name, np.zeros_like(getattr(out, name)))
    L.stokes_nc_get_params(nc, C.cast(got, C.c_void_p), C.byref(q),
                           *out.args(), C.byref(sm), C.byref(sr),
                           C.byref(rng))
    g = got.contents
    assert (g.version, g.np, g.nm, g.periodic) == (version, np_, nm,
                                                  int(periodic))
    assert q.value == 1 and rng.value == 0 and sm.value == shear
    assert g.shear_mode == shear
    if shear:
        assert sr.value == prm.shear_rate and g.shear_rate == prm.shear_rate
    if periodic:
        assert (g.lx, g.ly, g.lz) == tuple(prm.lat)
    if poly:
        assert np.array_equal(np.ctypeslib.as_array(g.a, (np_,)), prm.a)
    else:
        assert not g.a
    for name, n in (("Ui", 3), ("Oi", 3), ("Ei", 5), ("F", 3 * nm),
                    ("xf", 3 * nf), ("uf", 3 * nf)):
        assert np.array_equal(getattr(out, name)[:n], getattr(prm, name)[:n])
    if version >= 1:
        assert np.array_equal(out.T, prm.T)
    if version == 2:
        assert np.array_equal(out.E, prm.E)
    L.stokes_nc_free(nc)
    ls.lib().stokes_free(got)
    ls.lib().stokes_free(sys)


def test_print_actives(tmp_path):
    """stokes_nc_print_actives (stokes-nc.c:52-115): the table of the defined
    entries, the shear mode and the random-generator flag"""
    L = _lib()
    L.stokes_nc_print_actives.argtypes = [C.c_void_p, C.c_void_p]
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fn = str(tmp_path / "act.nc").encode()
    txt = str(tmp_path / "act.txt").encode()
    nc = L.stokes_nc_init(fn, 4, 2, 1, 1, 1, 0, 2, 1)
    out = libc.fopen(txt, b"w")
    L.stokes_nc_print_actives(nc, out)
    libc.fclose(out)
    L.stokes_nc_free(nc)
    text = open(txt.decode()).read().splitlines()
    assert text == [
        "    | x | u o e | f t s | ui oi ei",
        "----+---+-------+-------+----------",
        "x0  | 0 | 0 0 0 | 1 1 0 | 1  1  1",
        "xf0 | 1 | 1 1 0 | 0 0 0 |",
        "----+---+-------+-------+----------",
        "x   | 1 | 1 1 0 | 0 0 0 | 0  0  0",
        "xf  | 0 | 0 0 0 | 1 1 0 |",
        "----+---+-------+-------+----------",
        "q   | 1 |",
        "a   | 1 |",
        "af  | 1 |",
        "----+---+-------+------------------",
        "shear_mode      | 2",
        "----------------+------------------",
        "states of RNG   | 1"]
