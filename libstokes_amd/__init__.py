"""libstokes_amd -- Python view of the MI355X libstokes operator.

The product is the C-ABI library ``libstokes_amd/libstokes_hip.so`` (C host
code + HIP kernels for gfx950) with the reference's ``struct stokes`` /
``atimes_3all`` interface (include/*.h).  This module binds it with ctypes
so that tests and the benchmark drive exactly the entry points a C caller
uses.  There is no fallback: if the library is missing, or it cannot reach a
GPU, every call raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstokes_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

FP64 = 0
MIXED = 1

# symbols declared in include/*.h (checked by the CPU tests)
EXPORTS = [
    # stokes.h
    "stokes_init", "stokes_free", "stokes_set_np", "stokes_set_Ui",
    "stokes_set_Oi", "stokes_set_Ei", "stokes_set_shear",
    "stokes_get_shear_shift", "stokes_set_shear_shift", "stokes_set_l",
    "stokes_set_xi", "xi_by_tratio", "stokes_set_iter", "stokes_set_pos",
    "stokes_set_pos_mobile", "stokes_set_pos_fixed", "stokes_set_radius",
    "stokes_unset_radius",
    # ewald.h
    "scalars_ewald_real", "scalars_ewald_real_poly",
    "scalars_mob_poly_scale_SD_", "atimes_3all", "make_matrix_mob_3all",
    "atimes_3all_matrix", "atimes_ewald_3all", "make_matrix_mob_ewald_3all",
    # non-ewald.h
    "scalars_nonewald", "atimes_nonewald_3all",
    "make_matrix_mob_nonewald_3all",
    # matrix.h
    "dot_prod_matrix", "multiply_extmat_with_extvec_3fts",
    # chebyshev.h
    "chebyshev_coef", "chebyshev_eval", "chebyshev_eval_atimes",
    "chebyshev_error_minvsqrt", "chebyshev_error_Rsqrt",
    # stokes_hip.h
    "stokes_hip_last_error", "stokes_hip_set_precision",
    "stokes_hip_set_recip_g_fix", "stokes_hip_set_split",
    "stokes_hip_autotune",
    "stokes_hip_set_timing", "stokes_hip_malloc", "stokes_hip_free",
    "stokes_hip_memcpy_h2d", "stokes_hip_memcpy_d2h", "stokes_hip_sync",
    "stokes_hip_copy_calibrate",
    "stokes_hip_set_pos_dev", "stokes_hip_use_host_pos",
    "stokes_hip_atimes_dev", "stokes_hip_atimes_dev_noalloc", "stokes_hip_atimes",
    "stokes_hip_make_matrix_dev",
    "stokes_hip_make_matrix", "stokes_hip_extmat_dev",
    "stokes_hip_cheb_eval_dev", "stokes_hip_lanczos_bounds",
    "stokes_hip_adjoint_dev", "stokes_hip_scalars_real_f_deriv", "stokes_hip_solve_mix_dev", "solve_mob_3fts_matrix_0",
    "solve_mix_3fts_matrix_0", "stokes_hip_solve_mob_3fts_dev",
    "solve_mob_3fts_0", "stokes_hip_solve_mix_3fts_dev",
    "stokes_hip_solve_res_3fts_dev", "solve_res_3fts_0", "solve_res_3fts",
    "solve_mob_3fts", "solve_mix_3fts_0", "solve_mix_3fts",
    "solve_res_3f", "solve_res_3f_0", "solve_mob_3f", "solve_mix_3f",
    "solve_res_3ft", "solve_res_3ft_0", "solve_mob_3ft", "solve_mix_3ft",
    "shift_labo_to_rest_U", "shift_rest_to_labo_U", "shift_labo_to_rest_O",
    "shift_rest_to_labo_O", "shift_labo_to_rest_E",
    "stokes_hip_ensemble_create", "stokes_hip_ensemble_atimes",
    "stokes_hip_ensemble_lub", "stokes_hip_ensemble_atimes_lub",
    "stokes_hip_ensemble_set_precision", "stokes_hip_ensemble_set_split",
    "stokes_hip_ensemble_set_tiles", "stokes_hip_ensemble_get_tiles",
    "stokes_hip_ensemble_phase_ms",
    "stokes_hip_ensemble_refresh", "stokes_hip_ensemble_offset",
    "stokes_hip_ensemble_destroy",
    "stokes_hip_bd_create", "stokes_hip_bd_set_bounds",
    "stokes_hip_bd_set_fact_fix", "stokes_hip_bd_step",
    "stokes_hip_bd_set_real_cache", "stokes_hip_set_real_cache",
    "stokes_hip_bd_last_noise", "stokes_hip_bd_destroy",
    "stokes_hip_bd_step_q", "stokes_hip_bd_set_torque",
    "stokes_hip_bd_step_t", "stokes_hip_bd_set_shear_ref",
    "stokes_hip_bd_set_fts_iter", "stokes_hip_ensemble_autotune",
    "stokes_hip_bd_set_scheme", "stokes_hip_bd_set_lub", "stokes_hip_set_dense_blocks",
    "stokes_hip_set_overlap",
    "stokes_hip_bd_set_rng_mt19937",
    # lub.h, twobody.h
    "calc_lub_3f", "calc_lub_3ft", "calc_lub_3fts", "make_matrix_lub_3f",
    "make_matrix_lub_3ft", "make_matrix_lub_3fts", "stokes_hip_lub_dev",
    "stokes_hip_make_matrix_lub_dev", "stokes_hip_solve_res_lub_3fts_dev",
    "stokes_hip_solve_mob_lub_3fts_dev",
    "stokes_hip_solve_lub_3fts_matrix_dev", "solve_res_lub_3fts",
    "solve_res_lub_3fts_0", "solve_mob_lub_3fts", "solve_mob_lub_3fts_0",
    "solve_res_lub_3fts_matrix", "solve_res_lub_3fts_matrix_0",
    "solve_mob_lub_3fts_matrix", "solve_mob_lub_3fts_matrix_0",
    "stokes_hip_solve_res_lub_dev", "stokes_hip_solve_mix_lub_dev",
    "solve_res_lub_3f", "solve_res_lub_3f_0", "atimes_mob_lub_3f",
    "solve_mob_lub_3f", "solve_mix_lub_3f", "solve_res_lub_3ft",
    "solve_res_lub_3ft_0", "solve_mob_lub_3ft", "solve_mix_lub_3ft",
    "solve_mix_lub_3fts", "solve_mix_lub_3fts_0",
    "cond_lub", "cond_lub_poly", "twobody_f_init", "twobody_f_free",
    "twobody_far", "twobody_lub", "scalar_two_body_res",
]


class StokesStruct(C.Structure):
    """struct stokes of include/stokes.h (same layout as the reference)."""
    _fields_ = [
        ("version", C.c_int), ("np", C.c_int), ("nm", C.c_int),
        ("pos", _dp), ("a", _dp), ("rmin", C.c_double),
        ("twobody_nmax", C.c_int), ("twobody_lub", C.c_int),
        ("poly_table", _ip), ("twobody_f_list", _vp),
        ("slip", _dp), ("slip_a", _dp), ("slip_G32", _dp),
        ("slip_G30", _dp), ("slip_G52", _dp), ("slip_table", _ip),
        ("twobody_slip_f_list", _vp),
        ("Ui", C.c_double * 3), ("Oi", C.c_double * 3), ("Ei", C.c_double * 5),
        ("shear_mode", C.c_int), ("shear_rate", C.c_double),
        ("shear_shift", C.c_double),
        ("periodic", C.c_int),
        ("ewald_eps", C.c_double), ("rmax2", C.c_double),
        ("rmaxx", C.c_int), ("rmaxy", C.c_int), ("rmaxz", C.c_int),
        ("kmax", C.c_double),
        ("kmaxx", C.c_int), ("kmaxy", C.c_int), ("kmaxz", C.c_int),
        ("xi", C.c_double), ("xi2", C.c_double), ("xiaspi", C.c_double),
        ("xia2", C.c_double), ("pivol", C.c_double),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("ilx", C.c_int * 27), ("ily", C.c_int * 27), ("ilz", C.c_int * 27),
        ("self_a", C.c_double), ("self_c", C.c_double), ("self_m", C.c_double),
        ("flag_table", C.c_int), ("nr", C.c_int),
        ("rlx", _dp), ("rly", _dp), ("rlz", _dp),
        ("rmx", _ip), ("rmy", _ip), ("rmz", _ip),
        ("nk", C.c_int),
        ("ex", _dp), ("ey", _dp), ("ez", _dp), ("k", _dp), ("k1", _dp),
        ("k2", _dp), ("k3", _dp), ("ya", _dp), ("yb", _dp), ("yc", _dp),
        ("yg", _dp), ("yh", _dp), ("ym", _dp),
        ("lubmin2", C.c_double), ("lubmax", C.c_double),
        ("cpu1", C.c_double), ("cpu2", C.c_double), ("cpu3", C.c_double),
        ("it", _vp),
    ]


class BDInfo(C.Structure):
    """stokes_hip_bd_info of include/stokes_hip.h"""
    _fields_ = [("dt", C.c_double), ("err_cheb", C.c_double),
                ("eig", C.c_double * 2), ("attempts", C.c_int),
                ("halvings", C.c_int), ("redraws", C.c_int),
                ("refits", C.c_int), ("overlaps", C.c_ulonglong * 2)]

    def summary(self):
        return (f"dt {self.dt:g} err {self.err_cheb:.2e} bounds "
                f"[{self.eig[0]:.4g}, {self.eig[1]:.4g}] attempts "
                f"{self.attempts} halvings {self.halvings} redraws "
                f"{self.redraws} refits {self.refits} last overlaps "
                f"{self.overlaps[0]}/{self.overlaps[1]}")


class StokesError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libstokes_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise StokesError(
            f"{LIB_PATH} is not built; run `make -C libstokes_amd` "
            "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    sp = C.POINTER(StokesStruct)
    L.stokes_init.restype = sp
    L.stokes_free.argtypes = [sp]
    L.stokes_set_np.argtypes = [sp, C.c_int, C.c_int]
    L.stokes_set_l.argtypes = [sp, C.c_double, C.c_double, C.c_double]
    L.stokes_set_xi.argtypes = [sp, C.c_double, C.c_double]
    L.stokes_set_Ui.argtypes = [sp] + [C.c_double] * 3
    L.stokes_set_shear.argtypes = [sp, C.c_int, C.c_double]
    L.stokes_set_Oi.argtypes = [sp] + [C.c_double] * 3
    L.stokes_set_Ei.argtypes = [sp] + [C.c_double] * 5
    L.xi_by_tratio.restype = C.c_double
    L.xi_by_tratio.argtypes = [sp, C.c_double]
    L.stokes_set_pos.argtypes = [sp, _dp]
    L.stokes_set_radius.argtypes = [sp, _dp]
    L.stokes_unset_radius.argtypes = [sp]
    L.stokes_get_shear_shift.restype = C.c_double
    L.stokes_hip_last_error.restype = C.c_char_p
    for f in ("stokes_hip_set_precision", "stokes_hip_set_timing",
              "stokes_hip_set_recip_g_fix"):
        getattr(L, f).argtypes = [sp, C.c_int]
    L.stokes_hip_set_split.argtypes = [sp, C.c_double]
    L.stokes_hip_set_real_cache.argtypes = [sp, C.c_int]
    L.stokes_hip_set_dense_blocks.argtypes = [sp, C.c_int]
    L.stokes_hip_set_overlap.argtypes = [sp, C.c_int]
    L.stokes_hip_solve_mix_dev.argtypes = [sp, _vp, _vp]
    L.stokes_hip_solve_mob_3fts_dev.argtypes = [sp, _vp, _vp,
                                                C.POINTER(C.c_int), _dp]
    L.solve_mob_3fts_0.argtypes = [sp, _vp] + [_dp] * 6
    L.solve_mob_3fts_0.restype = None
    for f in ("stokes_hip_solve_mix_3fts_dev", "stokes_hip_solve_res_3fts_dev"):
        getattr(L, f).argtypes = [sp, _vp, _vp, C.POINTER(C.c_int), _dp]
    for f, a in (("solve_res_3fts_0", [sp] + [_dp] * 6),
                 ("solve_res_3fts", [sp] + [_dp] * 6),
                 ("solve_mob_3fts", [sp] + [_dp] * 6),
                 ("solve_mix_3fts_0", [sp, _vp] + [_dp] * 12),
                 ("solve_mix_3fts", [sp] + [_dp] * 12),
                 ("solve_res_3f", [sp] + [_dp] * 2),
                 ("solve_res_3f_0", [sp] + [_dp] * 2),
                 ("solve_mob_3f", [sp] + [_dp] * 2),
                 ("solve_mix_3f", [sp] + [_dp] * 4),
                 ("solve_res_3ft", [sp] + [_dp] * 4),
                 ("solve_res_3ft_0", [sp] + [_dp] * 4),
                 ("solve_mob_3ft", [sp] + [_dp] * 4),
                 ("solve_mix_3ft", [sp] + [_dp] * 8)):
        getattr(L, f).argtypes = a
        getattr(L, f).restype = None
    for f in ("solve_res_lub_3fts", "solve_res_lub_3fts_0",
              "solve_mob_lub_3fts", "solve_mob_lub_3fts_0",
              "solve_res_lub_3fts_matrix", "solve_res_lub_3fts_matrix_0",
              "solve_mob_lub_3fts_matrix", "solve_mob_lub_3fts_matrix_0"):
        getattr(L, f).argtypes = [sp] + [_dp] * 6
        getattr(L, f).restype = None
    for f in ("calc_lub_3f", "calc_lub_3ft", "calc_lub_3fts"):
        getattr(L, f).argtypes = [sp, _dp, _dp]
        getattr(L, f).restype = None
    for f in ("make_matrix_lub_3f", "make_matrix_lub_3ft",
              "make_matrix_lub_3fts"):
        getattr(L, f).argtypes = [sp, _dp]
        getattr(L, f).restype = None
    L.stokes_hip_lub_dev.argtypes = [sp, _vp, _vp]
    L.stokes_hip_make_matrix_lub_dev.argtypes = [sp, _vp]
    for f in ("stokes_hip_solve_res_lub_3fts_dev",
              "stokes_hip_solve_mob_lub_3fts_dev"):
        getattr(L, f).argtypes = [sp, _vp, _vp, C.POINTER(C.c_int), _dp]
    L.stokes_hip_solve_lub_3fts_matrix_dev.argtypes = [sp, C.c_int, _vp,
                                                       _vp]
    for f in ("stokes_hip_solve_res_lub_dev", "stokes_hip_solve_mix_lub_dev"):
        getattr(L, f).argtypes = [sp, _vp, _vp, C.POINTER(C.c_int), _dp]
    for f, a in (("solve_res_lub_3f", [sp] + [_dp] * 2),
                 ("solve_res_lub_3f_0", [sp] + [_dp] * 2),
                 ("solve_mob_lub_3f", [sp] + [_dp] * 2),
                 ("solve_mix_lub_3f", [sp] + [_dp] * 4),
                 ("solve_res_lub_3ft", [sp] + [_dp] * 4),
                 ("solve_res_lub_3ft_0", [sp] + [_dp] * 4),
                 ("solve_mob_lub_3ft", [sp] + [_dp] * 4),
                 ("solve_mix_lub_3ft", [sp] + [_dp] * 8),
                 ("solve_mix_lub_3fts", [sp] + [_dp] * 12),
                 ("solve_mix_lub_3fts_0", [sp] + [_dp] * 12)):
        getattr(L, f).argtypes = a
        getattr(L, f).restype = None
    L.atimes_mob_lub_3f.argtypes = [C.c_int, _dp, _dp, sp]
    L.atimes_mob_lub_3f.restype = None
    L.stokes_set_iter.argtypes = [sp, C.c_char_p, C.c_int, C.c_int,
                                  C.c_double, C.c_int, _vp]
    L.stokes_hip_adjoint_dev.argtypes = [sp, _vp, _vp, _vp, _vp]
    L.stokes_hip_scalars_real_f_deriv.argtypes = [C.c_double, C.c_double,
                                                  _dp]
    L.stokes_hip_scalars_real_f_deriv.restype = None
    L.solve_mob_3fts_matrix_0.argtypes = [sp] + [_dp] * 6
    L.solve_mix_3fts_matrix_0.argtypes = [sp] + [_dp] * 12
    L.stokes_hip_autotune.argtypes = [sp, C.c_int, _dp, C.c_int, C.c_int,
                                      _dp, _dp]
    L.stokes_hip_atimes.argtypes = [sp, C.c_int, _dp, _dp]
    L.stokes_hip_atimes_dev.argtypes = [sp, C.c_int, _vp, _vp]
    L.stokes_hip_make_matrix.argtypes = [sp, _dp]
    L.stokes_hip_make_matrix_dev.argtypes = [sp, _vp]
    L.stokes_hip_extmat_dev.argtypes = [sp, _vp, C.c_int, _vp, _vp]
    L.stokes_hip_cheb_eval_dev.argtypes = [sp, C.c_int, _dp, C.c_int, _vp,
                                           _vp, C.c_double, C.c_double]
    L.stokes_hip_lanczos_bounds.argtypes = [sp, C.c_int, C.c_ulong,
                                            C.c_double, _dp, _dp]
    L.stokes_hip_malloc.argtypes = [C.POINTER(_vp), C.c_size_t]
    L.stokes_hip_free.argtypes = [_vp]
    L.stokes_hip_memcpy_h2d.argtypes = [_vp, _vp, C.c_size_t]
    L.stokes_hip_memcpy_d2h.argtypes = [_vp, _vp, C.c_size_t]
    L.stokes_hip_sync.argtypes = [sp]
    L.stokes_hip_copy_calibrate.argtypes = [_vp, _vp, C.c_size_t]
    L.stokes_hip_set_pos_dev.argtypes = [sp, _vp]
    L.stokes_hip_use_host_pos.argtypes = [sp]
    L.chebyshev_eval.restype = C.c_double
    L.chebyshev_eval.argtypes = [C.c_int, _dp, C.c_double, C.c_double,
                                 C.c_double]
    L.scalars_nonewald.argtypes = [C.c_int, C.c_double, _dp]
    L.stokes_hip_ensemble_create.argtypes = [C.POINTER(sp), C.c_int,
                                             C.POINTER(_vp)]
    L.stokes_hip_ensemble_atimes.argtypes = [_vp, _vp, _vp, _vp]
    L.stokes_hip_ensemble_lub.argtypes = [_vp, _vp, _vp, _vp]
    L.stokes_hip_ensemble_atimes_lub.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.stokes_hip_ensemble_set_precision.argtypes = [_vp, C.c_int]
    L.stokes_hip_ensemble_set_split.argtypes = [_vp, C.c_int, C.c_double]
    L.stokes_hip_ensemble_set_tiles.argtypes = [_vp, C.c_int, C.c_int,
                                                C.c_int]
    L.stokes_hip_ensemble_get_tiles.argtypes = [_vp, _ip, _ip, _ip]
    L.stokes_hip_ensemble_phase_ms.argtypes = [_vp, _vp, _vp, _vp, _vp, _dp]
    L.stokes_hip_ensemble_refresh.argtypes = [_vp]
    L.stokes_hip_ensemble_autotune.argtypes = [_vp, _dp, C.c_int, C.c_int,
                                               _dp]
    L.stokes_hip_ensemble_offset.argtypes = [_vp, C.c_int]
    L.stokes_hip_ensemble_offset.restype = C.c_size_t
    L.stokes_hip_ensemble_destroy.argtypes = [_vp]
    L.stokes_hip_ensemble_destroy.restype = None
    L.stokes_hip_bd_create.argtypes = [sp, _vp, C.c_double, C.c_int,
                                       C.c_double, C.c_double, C.c_double,
                                       C.c_ulonglong, C.POINTER(_vp)]
    L.stokes_hip_bd_set_bounds.argtypes = [_vp, C.c_double, C.c_double]
    L.stokes_hip_bd_set_fact_fix.argtypes = [_vp, C.c_int]
    L.stokes_hip_bd_set_real_cache.argtypes = [_vp, C.c_int]
    L.stokes_hip_bd_step.argtypes = [_vp, _vp, C.c_double,
                                     C.POINTER(BDInfo)]
    L.stokes_hip_bd_step_q.argtypes = [_vp, _vp, _vp, C.c_double,
                                       C.POINTER(BDInfo)]
    L.stokes_hip_bd_step_t.argtypes = [_vp, _vp, _vp, C.c_double, C.c_double,
                                       C.POINTER(BDInfo)]
    L.stokes_hip_bd_set_shear_ref.argtypes = [_vp, C.c_double, C.c_double]
    L.stokes_hip_bd_set_torque.argtypes = [_vp, _vp]
    L.stokes_hip_bd_set_scheme.argtypes = [_vp, C.c_int, C.c_double]
    L.stokes_hip_bd_set_lub.argtypes = [_vp, C.c_int]
    L.stokes_hip_bd_set_rng_mt19937.argtypes = [_vp, C.c_ulong]
    L.stokes_hip_bd_set_fts_iter.argtypes = [_vp, C.c_int, C.c_int,
                                             C.c_double]
    L.stokes_hip_bd_last_noise.argtypes = [_vp, _vp, _vp]
    L.stokes_hip_bd_destroy.argtypes = [_vp]
    L.stokes_hip_bd_destroy.restype = None
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


def check(rc, what):
    if rc != 0:
        msg = lib().stokes_hip_last_error()
        raise StokesError(f"{what} failed ({rc}): "
                          f"{msg.decode() if msg else ''}")


def dim(version):
    return (3, 6, 11)[version]


class DeviceArray:
    """A float64 buffer in GPU memory owned by the library's allocator."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = _vp()
        check(lib().stokes_hip_malloc(C.byref(self.ptr), 8 * max(self.n, 1)),
              "stokes_hip_malloc")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        d = cls(a.size)
        check(lib().stokes_hip_memcpy_h2d(d.ptr, a.ctypes.data_as(_vp),
                                          8 * a.size), "h2d")
        return d

    def to_host(self):
        out = np.empty(self.n)
        check(lib().stokes_hip_memcpy_d2h(out.ctypes.data_as(_vp), self.ptr,
                                          8 * self.n), "d2h")
        return out

    def free(self):
        if self.ptr:
            check(lib().stokes_hip_free(self.ptr), "stokes_hip_free")
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stokes:
    """A struct stokes created and configured through the C API."""

    def __init__(self, version, pos, box=None, tr=None, xi=None, eps=1e-12,
                 a=None, rmin=0.0, shear_mode=0, shear_shift=0.0,
                 precision=FP64, nm=None):
        L = lib()
        self.L = L
        self.sys = L.stokes_init()
        s = self.sys.contents
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        self.np = pos.size // 3
        self.nm = self.np if nm is None else int(nm)
        s.version = version
        self.version = version
        L.stokes_set_np(self.sys, self.np, self.nm)
        L.stokes_set_pos(self.sys, _p(pos))
        if a is not None:
            a = np.ascontiguousarray(a, dtype=np.float64)
            L.stokes_set_radius(self.sys, _p(a))
        s.rmin = rmin
        if box is not None:
            s.periodic = 1
            L.stokes_set_l(self.sys, *[float(b) for b in box])
            if xi is None:
                xi = L.xi_by_tratio(self.sys, float(tr if tr else 1.0))
            L.stokes_set_xi(self.sys, float(xi), float(eps))
        s.shear_mode = shear_mode
        s.shear_shift = shear_shift
        self.n = dim(version) * self.np
        if precision != FP64:
            self.set_precision(precision)

    @property
    def s(self):
        return self.sys.contents

    def set_pos(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        self.L.stokes_set_pos(self.sys, _p(pos))

    def set_precision(self, mode):
        check(self.L.stokes_hip_set_precision(self.sys, mode),
              "set_precision")

    def set_lub(self, lubmin2=None, lubmax=None, nmax=None, lub_form=None):
        """lubrication parameters of struct stokes (lubmin2, lubmax,
        twobody_nmax, twobody_lub)"""
        s = self.s
        if lubmin2 is not None:
            s.lubmin2 = float(lubmin2)
        if lubmax is not None:
            s.lubmax = float(lubmax)
        if nmax is not None:
            s.twobody_nmax = int(nmax)
        if lub_form is not None:
            s.twobody_lub = int(lub_form)

    def lub(self, x):
        """y = L x, the lubrication correction of the system's version
        (stokes_hip_lub_dev through host staging)"""
        x = np.ascontiguousarray(x, dtype=np.float64)
        xd = DeviceArray.from_host(x)
        yd = DeviceArray(x.size)
        check(self.L.stokes_hip_lub_dev(self.sys, xd.ptr, yd.ptr), "lub_dev")
        return yd.to_host()

    def lub_dev(self, x_dev, y_dev):
        check(self.L.stokes_hip_lub_dev(self.sys, x_dev.ptr, y_dev.ptr),
              "lub_dev")

    def lub_matrix(self):
        """the dense extracted-form lubrication matrix of the system's
        version (stokes_hip_make_matrix_lub_dev), (n, n) row-major"""
        md = DeviceArray(self.n * self.n)
        check(self.L.stokes_hip_make_matrix_lub_dev(self.sys, md.ptr),
              "make_matrix_lub_dev")
        return md.to_host().reshape(self.n, self.n)

    def solve_lub(self, b, kind="res", matrix=False):
        """lubricated FTS problem on the device: kind "res" (U O E -> F T
        S) or "mob" (F T E -> U O S), per particle; GMRES or (matrix) the
        dense solve; returns (x, iterations, relative residual), the last
        two None for the dense solve"""
        bd = DeviceArray.from_host(np.ascontiguousarray(b, np.float64))
        xd = DeviceArray(bd.n)
        if matrix:
            check(self.L.stokes_hip_solve_lub_3fts_matrix_dev(
                self.sys, 0 if kind == "res" else 1, bd.ptr, xd.ptr),
                "solve_lub_3fts_matrix")
            return xd.to_host(), None, None
        it, rr = C.c_int(), C.c_double()
        f = (self.L.stokes_hip_solve_res_lub_3fts_dev if kind == "res"
             else self.L.stokes_hip_solve_mob_lub_3fts_dev)
        check(f(self.sys, bd.ptr, xd.ptr, C.byref(it), C.byref(rr)),
              "solve_" + kind + "_lub")
        return xd.to_host(), it.value, rr.value

    def solve_lub_any(self, b, kind="mix"):
        """lubricated problem of the system's version on the device, in
        the fluid-rest frame: kind "res" (velocities of every particle ->
        forces, M f = (I + M L) v) or "mix" (known [(F, T, E)_mobile, (U, O,
        E)_fixed] -> unknowns [(U, O, S)_mobile, (F, T, S)_fixed], the
        system's nm mobile particles); per particle as the reference packs
        them; returns (x, iterations, relative residual)"""
        bd = DeviceArray.from_host(np.ascontiguousarray(b, np.float64))
        xd = DeviceArray(bd.n)
        it, rr = C.c_int(), C.c_double()
        f = (self.L.stokes_hip_solve_res_lub_dev if kind == "res"
             else self.L.stokes_hip_solve_mix_lub_dev)
        check(f(self.sys, bd.ptr, xd.ptr, C.byref(it), C.byref(rr)),
              "solve_" + kind + "_lub")
        x = xd.to_host()
        bd.free()
        xd.free()
        return x, it.value, rr.value

    def set_recip_g_fix(self, on=True):
        check(self.L.stokes_hip_set_recip_g_fix(self.sys, 1 if on else 0),
              "set_recip_g_fix")

    def set_split(self, tratio):
        """compute at the splitting of tratio (0: sys->xi); results stay
        those of sys->xi to the Ewald tolerance"""
        check(self.L.stokes_hip_set_split(self.sys, float(tratio)),
              "set_split")

    def autotune(self, nrhs=1, trs=None, reps=3):
        """time candidate splittings, keep the fastest; returns (tr, ms)"""
        bt, bm = C.c_double(), C.c_double()
        if trs is None:
            tp, nt = None, 0
        else:
            arr = np.ascontiguousarray(trs, dtype=np.float64)
            tp, nt = _p(arr), arr.size
        check(self.L.stokes_hip_autotune(self.sys, nrhs, tp, nt, reps,
                                         C.byref(bt), C.byref(bm)),
              "autotune")
        return bt.value, bm.value

    def solve_mix(self, b, nm=None):
        """dense FP64 mixed problem (stokes_hip_solve_mix_dev): b holds the
        known values per particle (mobile: F T E, fixed: U O E), the result
        the unknowns (mobile: U O S, fixed: F T S)"""
        if nm is not None:
            self.s.nm = nm
        bd = DeviceArray.from_host(np.ascontiguousarray(b, np.float64))
        xd = DeviceArray(self.n)
        check(self.L.stokes_hip_solve_mix_dev(self.sys, bd.ptr, xd.ptr),
              "solve_mix")
        return xd.to_host()

    def adjoint_dev(self, g_dev, f_dev, dx_dev=None, df_dev=None):
        """d(g . M(x) f)/dx and /df into device arrays (F version)"""
        check(self.L.stokes_hip_adjoint_dev(
            self.sys, g_dev.ptr, f_dev.ptr if f_dev is not None else None,
            dx_dev.ptr if dx_dev is not None else None,
            df_dev.ptr if df_dev is not None else None), "adjoint")

    def solve_mob_gmres(self, fte):
        """FTS mobility problem by device GMRES: fte per particle (F T E)
        -> (U O S); returns (x, iterations, relative residual)"""
        bd = DeviceArray.from_host(np.ascontiguousarray(fte, np.float64))
        xd = DeviceArray(bd.n)
        it, rr = C.c_int(), C.c_double()
        check(self.L.stokes_hip_solve_mob_3fts_dev(self.sys, bd.ptr, xd.ptr,
                                                   C.byref(it),
                                                   C.byref(rr)),
              "solve_mob_3fts")
        return xd.to_host(), it.value, rr.value

    def solve_gmres(self, b, kind="mix", nm=None):
        """device GMRES: kind "mix" (b per particle: mobile F T E, fixed
        U O E -> mobile U O S, fixed F T S; nm mobile particles) or "res"
        (U O E -> F T S); returns (x, iterations, relative residual)"""
        if nm is not None:
            self.s.nm = nm
        bd = DeviceArray.from_host(np.ascontiguousarray(b, np.float64))
        xd = DeviceArray(bd.n)
        it, rr = C.c_int(), C.c_double()
        f = (self.L.stokes_hip_solve_res_3fts_dev if kind == "res"
             else self.L.stokes_hip_solve_mix_3fts_dev)
        check(f(self.sys, bd.ptr, xd.ptr, C.byref(it), C.byref(rr)),
              "solve_" + kind)
        return xd.to_host(), it.value, rr.value

    def set_real_cache(self, on=True):
        """dense real-space matrix of the current positions for the
        following products (keep the positions fixed)"""
        check(self.L.stokes_hip_set_real_cache(self.sys, int(bool(on))),
              "set_real_cache")

    def set_overlap(self, on=True):
        """FP64 block products: reciprocal part on a second stream (default on)"""
        check(self.L.stokes_hip_set_overlap(self.sys, int(bool(on))),
              "set_overlap")

    def set_dense_blocks(self, on=True):
        """per-call dense real-space matrix for wide MIXED blocks"""
        check(self.L.stokes_hip_set_dense_blocks(self.sys, int(bool(on))),
              "set_dense_blocks")

    def set_iter(self, solver="gmres", max_it=2000, restart=30, eps=1e-12):
        self.L.stokes_set_iter(self.sys, solver.encode(), max_it, restart,
                               eps, 0, None)

    def set_timing(self, on=True):
        check(self.L.stokes_hip_set_timing(self.sys, 1 if on else 0),
              "set_timing")

    def params(self):
        s = self.s
        return dict(xi=s.xi, rmax2=s.rmax2, rmaxx=s.rmaxx, rmaxy=s.rmaxy,
                    rmaxz=s.rmaxz, kmax=s.kmax, kmaxx=s.kmaxx, kmaxy=s.kmaxy,
                    kmaxz=s.kmaxz, nr=s.nr, self_a=s.self_a, self_c=s.self_c,
                    self_m=s.self_m, pivol=s.pivol)

    def image_table(self):
        s = self.s
        rl = np.array([[s.rlx[m], s.rly[m], s.rlz[m]] for m in range(s.nr)])
        rm = np.array([[s.rmx[m], s.rmy[m], s.rmz[m]] for m in range(s.nr)],
                      dtype=np.int32)
        return rl, rm

    def atimes(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        check(self.L.stokes_hip_atimes(self.sys, self.n, _p(x), _p(y)),
              "atimes")
        return y

    def atimes_dev(self, x_dev, y_dev, nrhs=1):
        check(self.L.stokes_hip_atimes_dev(self.sys, nrhs, x_dev.ptr,
                                           y_dev.ptr), "atimes_dev")

    def matrix(self):
        m = np.empty((self.n, self.n))
        check(self.L.stokes_hip_make_matrix(self.sys, _p(m)), "make_matrix")
        return m

    def matrix_dev(self, mat_dev):
        check(self.L.stokes_hip_make_matrix_dev(self.sys, mat_dev.ptr),
              "make_matrix_dev")

    def extmat_dev(self, mat_dev, x_dev, y_dev, nrhs=1):
        check(self.L.stokes_hip_extmat_dev(self.sys, mat_dev.ptr, nrhs,
                                           x_dev.ptr, y_dev.ptr), "extmat")

    def cheb_dev(self, acoef, y_dev, z_dev, e0, e1, nrhs=1):
        a = np.ascontiguousarray(acoef, dtype=np.float64)
        check(self.L.stokes_hip_cheb_eval_dev(self.sys, a.size, _p(a), nrhs,
                                              y_dev.ptr, z_dev.ptr, e0, e1),
              "cheb_eval_dev")

    def lanczos_bounds(self, niter=40, seed=1, margin=0.0):
        lo, hi = C.c_double(), C.c_double()
        check(self.L.stokes_hip_lanczos_bounds(self.sys, niter, seed, margin,
                                               C.byref(lo), C.byref(hi)),
              "lanczos")
        return lo.value, hi.value

    def sync(self):
        check(self.L.stokes_hip_sync(self.sys), "sync")

    def times(self):
        s = self.s
        return s.cpu1, s.cpu2, s.cpu3

    def close(self):
        if self.sys:
            self.L.stokes_free(self.sys)
            self.sys = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """The operator of several independent systems as one graph launch
    (stokes_hip_ensemble_* of include/stokes_hip.h)."""

    def __init__(self, systems):
        self.L = lib()
        self.systems = list(systems)
        arr = (C.POINTER(StokesStruct) * len(self.systems))(
            *[s.sys for s in self.systems])
        h = _vp()
        check(self.L.stokes_hip_ensemble_create(arr, len(self.systems),
                                                C.byref(h)), "ensemble_create")
        self.h = h
        self.offsets = [self.L.stokes_hip_ensemble_offset(h, b)
                        for b in range(len(self.systems) + 1)]

    def atimes_dev(self, pos_dev, x_dev, y_dev):
        check(self.L.stokes_hip_ensemble_atimes(self.h, pos_dev.ptr, x_dev.ptr,
                                                y_dev.ptr), "ensemble_atimes")

    def lub_dev(self, pos_dev, x_dev, ylub_dev):
        """the FP64 lubrication correction of every box"""
        check(self.L.stokes_hip_ensemble_lub(self.h, pos_dev.ptr, x_dev.ptr,
                                             ylub_dev.ptr), "ensemble_lub")

    def atimes_lub_dev(self, pos_dev, x_dev, y_dev, ylub_dev):
        """M x and the lubrication correction L x in one launch"""
        check(self.L.stokes_hip_ensemble_atimes_lub(
            self.h, pos_dev.ptr, x_dev.ptr, y_dev.ptr, ylub_dev.ptr),
            "ensemble_atimes_lub")

    def set_precision(self, mode):
        check(self.L.stokes_hip_ensemble_set_precision(self.h, mode),
              "ensemble_set_precision")

    def set_split(self, tr, box=-1):
        check(self.L.stokes_hip_ensemble_set_split(self.h, box, float(tr)),
              "ensemble_set_split")

    def set_tiles(self, jw, jper, kper):
        check(self.L.stokes_hip_ensemble_set_tiles(self.h, jw, jper, kper),
              "ensemble_set_tiles")

    def tiles(self):
        v = [C.c_int(), C.c_int(), C.c_int()]
        check(self.L.stokes_hip_ensemble_get_tiles(self.h, *[C.byref(t)
                                                             for t in v]),
              "ensemble_get_tiles")
        return tuple(t.value for t in v)

    def phase_ms(self, pos_dev, x_dev, y_dev, ylub_dev=None):
        """(real, reciprocal, combine, lubrication) ms of one eager launch
        (y_dev or ylub_dev None: that product skipped)"""
        ms = np.zeros(4)
        check(self.L.stokes_hip_ensemble_phase_ms(
            self.h, pos_dev.ptr, x_dev.ptr, y_dev.ptr if y_dev else None,
            ylub_dev.ptr if ylub_dev else None, _p(ms)), "ensemble_phase_ms")
        return ms

    def refresh(self):
        check(self.L.stokes_hip_ensemble_refresh(self.h), "ensemble_refresh")

    def autotune(self, trs=None, reps=3):
        """splittings per box and tile shape; returns the splittings kept"""
        nb = len(self.systems)
        best = np.zeros(nb)
        if trs is None:
            tp, nt = None, 0
        else:
            arr = np.ascontiguousarray(trs, dtype=np.float64)
            tp, nt = _p(arr), arr.size
        check(self.L.stokes_hip_ensemble_autotune(self.h, tp, nt, reps,
                                                  _p(best)),
              "ensemble_autotune")
        return best

    def close(self):
        if self.h:
            self.L.stokes_hip_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BD:
    """Mid-point Brownian-dynamics integrator of one F, FT or FTS system
    (stokes_hip_bd_* of include/stokes_hip.h; brownian.c:2352-2560)."""

    def __init__(self, stokes, peclet=1.0, n_cheb=30, eps=1e-6, dt_lim=1e-6,
                 rmin_ol=1.0, seed=1, force_dev=None, bounds=None,
                 fact_fix=False):
        self.L = lib()
        self.stokes = stokes
        # the captured step reads the force buffer on every replay: keep it
        self.force_dev = force_dev
        h = _vp()
        check(self.L.stokes_hip_bd_create(
            stokes.sys, force_dev.ptr if force_dev is not None else None,
            float(peclet), int(n_cheb), float(eps), float(dt_lim),
            float(rmin_ol), int(seed), C.byref(h)), "bd_create")
        self.h = h
        if bounds is not None:
            check(self.L.stokes_hip_bd_set_bounds(self.h, *map(float, bounds)),
                  "bd_set_bounds")
        if fact_fix:
            check(self.L.stokes_hip_bd_set_fact_fix(self.h, 1),
                  "bd_set_fact_fix")

    def set_real_cache(self, on=True):
        """dense real-space matrix for the products at x of each step"""
        check(self.L.stokes_hip_bd_set_real_cache(self.h, int(bool(on))),
              "bd_set_real_cache")

    def set_torque(self, torque_dev):
        """constant torques [3 np] (FT, FTS); read on every step: kept"""
        self.torque_dev = torque_dev
        check(self.L.stokes_hip_bd_set_torque(
            self.h, torque_dev.ptr if torque_dev is not None else None),
            "bd_set_torque")

    def set_rng_mt19937(self, seed):
        """draw the noise from the reference's MT19937 + polar stream"""
        check(self.L.stokes_hip_bd_set_rng_mt19937(self.h, int(seed)),
              "bd_set_rng_mt19937")

    def set_lub(self, on=True):
        """the drift of each stage with lubrication (BD_params_init's
        flag_lub): the lubricated mixed problem, attempts run eagerly"""
        check(self.L.stokes_hip_bd_set_lub(self.h, 1 if on else 0),
              "bd_set_lub")

    def set_scheme(self, scheme, bb_n=1.0):
        """0 mid-point, 1 Banchio-Brady 2003 (step parameter bb_n), 2
        Ball-Melrose 1997"""
        check(self.L.stokes_hip_bd_set_scheme(self.h, int(scheme),
                                              float(bb_n)), "bd_set_scheme")

    def set_fts_iter(self, restart=20, max_iter=2000, eps=1e-6):
        """FTS: GMRES parameters of the noise products"""
        check(self.L.stokes_hip_bd_set_fts_iter(self.h, int(restart),
                                                int(max_iter), float(eps)),
              "bd_set_fts_iter")

    def step(self, x_dev, dt, q_dev=None, t=None):
        """one step; t: the time of the step (needed for a sheared system)"""
        info = BDInfo()
        qp = q_dev.ptr if q_dev is not None else None
        if t is None:
            rc = self.L.stokes_hip_bd_step_q(self.h, x_dev.ptr, qp, float(dt),
                                             C.byref(info))
        else:
            rc = self.L.stokes_hip_bd_step_t(self.h, x_dev.ptr, qp, float(t),
                                             float(dt), C.byref(info))
        if rc != 0:
            msg = self.L.stokes_hip_last_error()
            raise StokesError(f"bd_step failed ({rc}): "
                              f"{msg.decode() if msg else ''} "
                              f"[{info.summary()}]")
        return info

    def set_shear_ref(self, t0, s0):
        check(self.L.stokes_hip_bd_set_shear_ref(self.h, float(t0), float(s0)),
              "bd_set_shear_ref")

    def last_noise(self):
        n = (3 if self.stokes.version == 0 else 6) * self.stokes.s.nm
        y = DeviceArray(n)
        z = DeviceArray(n)
        check(self.L.stokes_hip_bd_last_noise(self.h, y.ptr, z.ptr),
              "bd_last_noise")
        return y.to_host(), z.to_host()

    def close(self):
        if self.h:
            self.L.stokes_hip_bd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cheb_coef_inv_sqrt(n_cheb, e0, e1):
    """Chebyshev coefficients of 1/sqrt(x) on [e0, e1] (chebyshev_coef of
    the library, chebyshev.c:33-61)"""
    L = lib()
    a = np.zeros(n_cheb)
    f = _INV_SQRT_CB
    L.chebyshev_coef(int(n_cheb), f, C.c_double(e0), C.c_double(e1), _p(a))
    return a


def cheb_coef_sqrt(n_cheb, e0, e1):
    """Chebyshev coefficients of sqrt(x) on [e0, e1]"""
    L = lib()
    a = np.zeros(n_cheb)
    L.chebyshev_coef(int(n_cheb), _SQRT_CB, C.c_double(e0), C.c_double(e1),
                     _p(a))
    return a


_FUNC_CB = C.CFUNCTYPE(C.c_double, C.c_double)
_INV_SQRT_CB = _FUNC_CB(lambda x: 1.0 / np.sqrt(x))
_SQRT_CB = _FUNC_CB(lambda x: float(np.sqrt(x)))


def exported_symbols():
    """Names declared in include/*.h that the library must export."""
    return list(EXPORTS)
