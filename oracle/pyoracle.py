"""ctypes bindings to the test oracle (oracle/liboracle.so) and to the
compiled reference hot path (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (libstokes_amd)
never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcSys(C.Structure):
    """Mirror of orc_sys (oracle.h)."""
    _fields_ = [
        ("version", C.c_int), ("np", C.c_int),
        ("pos", _dp), ("a", _dp), ("rmin", C.c_double), ("periodic", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("pivol", C.c_double),
        ("xi", C.c_double), ("xi2", C.c_double), ("xiaspi", C.c_double),
        ("xia2", C.c_double), ("ewald_eps", C.c_double),
        ("self_a", C.c_double), ("self_c", C.c_double), ("self_m", C.c_double),
        ("rmax2", C.c_double), ("rmaxx", C.c_int), ("rmaxy", C.c_int),
        ("rmaxz", C.c_int),
        ("kmax", C.c_double), ("kmaxx", C.c_int), ("kmaxy", C.c_int),
        ("kmaxz", C.c_int),
        ("nr", C.c_int), ("rlx", _dp), ("rly", _dp), ("rlz", _dp),
        ("rmx", _ip), ("rmy", _ip), ("rmz", _ip),
        ("shear_mode", C.c_int), ("shear_shift", C.c_double),
        ("fix_recip_g", C.c_int),
        ("Ui", C.c_double * 3), ("Oi", C.c_double * 3), ("Ei", C.c_double * 5),
        ("shear_rate", C.c_double), ("bd_t", C.c_double),
        ("bd_t0", C.c_double), ("bd_s0", C.c_double),
    ]


def _p(a):
    return a.ctypes.data_as(_dp)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is not built (run `make -C oracle` / `make -C oracle ref`)")
    return C.CDLL(path)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load(ORACLE_SO)
        L = _lib
        L.orc_xi_by_tratio.restype = C.c_double
        L.orc_ewald_real_bound.restype = C.c_double
        L.orc_ewald_recip_bound.restype = C.c_double
        L.orc_cheb_eval.restype = C.c_double
        L.orc_cheb_error_minvsqrt.restype = C.c_double
        L.orc_cheb_error_Rsqrt.restype = C.c_double
        L.orc_twobody_table.argtypes = [C.c_int, C.c_int, C.c_double, _dp]
        L.orc_twobody_res.argtypes = [C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_int,
                                      _dp]
        L.orc_two_body_res.argtypes = [C.c_double, _dp]
        L.orc_minv_poly.argtypes = [C.c_int, C.c_double, C.c_double,
                                    C.c_double, _dp]
        L.orc_minv_fts.argtypes = [C.c_double, _dp]
        L.orc_block_atimes_sign.argtypes = [C.c_int, _dp, _dp, C.c_double,
                                            C.c_double, C.c_double, _dp,
                                            C.c_int]
        L.orc_lub_3fts.argtypes = [C.c_void_p, C.c_double, C.c_double,
                                   C.c_int, C.c_int, _dp, _dp]
        L.orc_func.restype = C.c_double
        L.orc_rng_int32.restype = C.c_ulong
        L.orc_rng_real2.restype = C.c_double
        L.orc_rng_gaussian.restype = C.c_double
        L.orc_bd_mid_F.restype = C.c_double
        L.orc_bd_mid.restype = C.c_double
        L.orc_bd_step.restype = C.c_double
        L.orc_bd_step_nm.restype = C.c_double
        L.orc_check_overlap.restype = C.c_int
        L.orc_xi_by_tratio.argtypes = [C.POINTER(OrcSys), C.c_double]
        L.orc_set_xi.argtypes = [C.POINTER(OrcSys), C.c_double, C.c_double]
        L.orc_set_l.argtypes = [C.POINTER(OrcSys), C.c_double, C.c_double,
                                C.c_double]
        L.orc_ewald_real_bound.argtypes = [C.c_double, C.c_double]
        L.orc_ewald_recip_bound.argtypes = [C.c_double, C.c_double]
    return _lib


_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = _load(REF_SO)
        R = _ref
        R.ref_sys_new.restype = C.c_void_p
        R.ref_sys_new.argtypes = [C.POINTER(OrcSys)]
        R.ref_sys_free.argtypes = [C.c_void_p]
        R.ref_atimes.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        R.ref_matrix.argtypes = [C.c_void_p, _dp]
        R.ref_cpu_split.restype = C.c_double
        R.ref_cpu_split.argtypes = [C.c_void_p, C.c_int]
        R.ref_cheb_eval.restype = C.c_double
        R.ref_cheb_error_minvsqrt_dense.restype = C.c_double
        R.ref_cheb_error_Rsqrt_dense.restype = C.c_double
        R.ref_rng_new.restype = C.c_void_p
        R.ref_rng_new.argtypes = [C.c_ulong]
        R.ref_rng_free.argtypes = [C.c_void_p]
        R.ref_rng_int32.restype = C.c_ulong
        R.ref_rng_int32.argtypes = [C.c_void_p]
        R.ref_rng_gaussian.restype = C.c_double
        R.ref_rng_gaussian.argtypes = [C.c_void_p]
        R.ref_cheb_eval_sys.argtypes = [C.c_int, _dp, C.c_void_p, _dp, _dp,
                                        C.c_double, C.c_double]
        R.ref_twobody_table.argtypes = [C.c_int, C.c_int, C.c_double, _dp]
        R.ref_twobody_res.argtypes = [C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_int,
                                      _dp]
        R.ref_two_body_res.argtypes = [C.c_double, _dp]
        R.ref_minv_fts.argtypes = [C.c_double, _dp]
        R.ref_fts_block.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        R.ref_lub_3fts_mono.argtypes = [C.c_void_p, C.c_double, C.c_double,
                                        _dp, _dp]
        R.ref_lub_matrix_mono.argtypes = [C.c_void_p, C.c_double,
                                          C.c_double, _dp]
        R.scalars_nonewald_poly_full.argtypes = [C.c_int, C.c_double,
                                                 C.c_double, C.c_double, _dp]
    return _ref


def dim(version):
    return (3, 6, 11)[version]


class System:
    """Oracle-side system: holds numpy buffers alive for the orc_sys."""

    def __init__(self, version, pos, box=None, xi=None, tr=None, eps=1e-12,
                 a=None, rmin=0.0, shear_mode=0, shear_shift=0.0,
                 fix_recip_g=False):
        L = lib()
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        self.np = self.pos.size // 3
        self.a = None if a is None else np.ascontiguousarray(a, np.float64)
        s = OrcSys()
        L.orc_sys_init(C.byref(s))
        s.version = version
        s.np = self.np
        s.pos = _p(self.pos)
        s.a = _p(self.a) if self.a is not None else None
        s.rmin = rmin
        s.shear_mode = shear_mode
        s.shear_shift = shear_shift
        s.fix_recip_g = 1 if fix_recip_g else 0
        if box is not None:
            s.periodic = 1
            L.orc_set_l(C.byref(s), *[float(b) for b in box])
            if xi is None:
                xi = L.orc_xi_by_tratio(C.byref(s), float(tr if tr else 1.0))
            L.orc_set_xi(C.byref(s), float(xi), float(eps))
        self.s = s
        self.version = version
        self.n = dim(version) * self.np

    def __del__(self):
        try:
            lib().orc_sys_free_tables(C.byref(self.s))
        except Exception:
            pass

    def params(self):
        s = self.s
        return dict(xi=s.xi, rmax2=s.rmax2, rmaxx=s.rmaxx, rmaxy=s.rmaxy,
                    rmaxz=s.rmaxz, kmax=s.kmax, kmaxx=s.kmaxx, kmaxy=s.kmaxy,
                    kmaxz=s.kmaxz, nr=s.nr, self_a=s.self_a, self_c=s.self_c,
                    self_m=s.self_m, pivol=s.pivol)

    def image_table(self):
        s = self.s
        nr = s.nr
        rl = np.array([[s.rlx[m], s.rly[m], s.rlz[m]] for m in range(nr)])
        rm = np.array([[s.rmx[m], s.rmy[m], s.rmz[m]] for m in range(nr)],
                      dtype=np.int32)
        return rl, rm

    def atimes(self, x, rows=None):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n)
        if rows is None:
            lib().orc_atimes_3all(C.byref(self.s), _p(x), _p(y))
        else:
            lib().orc_atimes_ewald_rows(C.byref(self.s), _p(x), _p(y),
                                        int(rows[0]), int(rows[1]))
        return y

    def matrix(self):
        m = np.zeros((self.n, self.n))
        if self.s.periodic:
            lib().orc_matrix_ewald(C.byref(self.s), _p(m))
        else:
            lib().orc_matrix_nonewald(C.byref(self.s), _p(m))
        return m

    def check_overlap(self, pos, rmin=1.0):
        """overlapping (pair, image) count at positions pos (brownian.c:299-375)"""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1)
        return lib().orc_check_overlap(C.byref(self.s), _p(pos),
                                       C.c_double(rmin))

    def bd_mid_F(self, x, draws, peclet, dt, ncheb, acoef, e0, e1, F=None,
                 dt_lim=1e-6, rmin_ol=1.0, fix_fact=False):
        """one mid-point BD step (brownian.c:2352-2560), F version; returns
        (x_new, dt_taken, draws_used, halvings, err_cheb)"""
        x = np.array(x, dtype=np.float64).reshape(-1)
        draws = np.ascontiguousarray(draws, np.float64).reshape(-1)
        nd = draws.size // x.size
        a = np.ascontiguousarray(acoef, np.float64)
        used, nh, err = C.c_int(), C.c_int(), C.c_double()
        Fp = None if F is None else _p(np.ascontiguousarray(F, np.float64))
        L = lib()
        dtt = L.orc_bd_mid_F(C.byref(self.s), _p(x), Fp, C.c_double(peclet),
                             C.c_double(dt), C.c_double(dt_lim),
                             C.c_double(rmin_ol), C.c_int(ncheb), _p(a),
                             C.c_double(e0), C.c_double(e1), C.c_int(nd),
                             _p(draws), C.c_int(1 if fix_fact else 0),
                             C.byref(used), C.byref(nh), C.byref(err))
        return x, dtt, used.value, nh.value, err.value

    def bd_mid(self, x, draws, peclet, dt, ncheb, acoef, e0, e1, q=None,
               F=None, T=None, dt_lim=1e-6, rmin_ol=1.0, fix_fact=False,
               scheme=0, bb_n=1.0, nm=None):
        """one BD step (scheme 0 mid-point brownian.c:2352-2560, 1 BB03,
        2 BM97) of the system's version (F, FT, FTS); draws[k] has 3 np (F)
        or 6 np values; returns (x_new, q_new, dt_taken, draws_used,
        halvings, err_cheb)"""
        x = np.array(x, dtype=np.float64).reshape(-1)
        nm = self.np if nm is None else int(nm)
        x = x[:3 * nm].copy()
        nop = x.size * (1 if self.version == 0 else 2)
        draws = np.ascontiguousarray(draws, np.float64).reshape(-1)
        nd = draws.size // nop
        a = np.ascontiguousarray(acoef, np.float64)
        qq = None if q is None else np.array(q, dtype=np.float64).reshape(-1)
        used, nh, err = C.c_int(), C.c_int(), C.c_double()

        def opt(v):
            return None if v is None else _p(np.ascontiguousarray(v, np.float64))
        Fa = None if F is None else np.ascontiguousarray(F, np.float64)
        Ta = None if T is None else np.ascontiguousarray(T, np.float64)
        dtt = lib().orc_bd_step_nm(
            C.byref(self.s), C.c_int(nm), C.c_int(scheme), C.c_double(bb_n),
            _p(x),
            None if qq is None else _p(qq), opt(Fa),
            opt(Ta), C.c_double(peclet), C.c_double(dt), C.c_double(dt_lim),
            C.c_double(rmin_ol), C.c_int(ncheb), _p(a), C.c_double(e0),
            C.c_double(e1), C.c_int(nd), _p(draws),
            C.c_int(1 if fix_fact else 0), C.byref(used), C.byref(nh),
            C.byref(err))
        return x, qq, dtt, used.value, nh.value, err.value

    def fts_mix_ft(self, nm, ft):
        """(F, T) -> (U, O) of the nm mobile particles, the rest at rest,
        FTS (BD_atimes_mob_FU with fixed particles)"""
        ft = np.ascontiguousarray(ft, np.float64).reshape(-1)
        uo = np.zeros_like(ft)
        lib().orc_fts_mix_ft(C.byref(self.s), C.c_int(nm), _p(ft), _p(uo))
        return uo

    def bd_op(self, x):
        """the BD mobility of the system's version (BD_atimes_mob_FU): F ->
        U, (F, T) -> (U, O), or for FTS the E = 0 mobility problem"""
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        y = np.zeros_like(x)
        lib().orc_bd_op_cb(C.c_int(x.size), _p(x), _p(y), C.byref(self.s))
        return y

    # ---- compiled reference on the same system --------------------------
    def ref_atimes(self, x):
        R = ref()
        h = R.ref_sys_new(C.byref(self.s))
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n)
        R.ref_atimes(h, self.n, _p(x), _p(y))
        R.ref_sys_free(h)
        return y

    def lub(self, uoe, lubmin2=4.0 + 1e-12, lubmax=3.0, nmax=100,
            flag_lub=1):
        """calc_lub_3fts (lub.c:576-768) of the oracle: fts = L(uoe);
        lubmin2 is r^2 (equal spheres) or s^2 (unequal), lubmax r or s"""
        uoe = np.ascontiguousarray(uoe, np.float64)
        fts = np.zeros_like(uoe)
        lib().orc_lub_3fts(C.byref(self.s), lubmin2, lubmax, nmax, flag_lub,
                           _p(uoe), _p(fts))
        return fts

    def ref_lub_mono(self, uoe, lubmin2=4.0 + 1e-12, lubmax=3.0):
        """the compiled reference's calc_lub_fts_2b over lub.c's pair loop
        (equal spheres)"""
        R = ref()
        h = R.ref_sys_new(C.byref(self.s))
        uoe = np.ascontiguousarray(uoe, np.float64)
        fts = np.zeros_like(uoe)
        R.ref_lub_3fts_mono(h, lubmin2, lubmax, _p(uoe), _p(fts))
        R.ref_sys_free(h)
        return fts

    def ref_lub_matrix_mono(self, lubmin2=4.0 + 1e-12, lubmax=3.0):
        """the compiled reference's matrix_lub_{f,ft,fts}_2b over
        lub-matrix.c's pair loop (equal spheres): the extracted-form
        lubrication matrix, (n, n)"""
        R = ref()
        h = R.ref_sys_new(C.byref(self.s))
        m = np.zeros((self.n, self.n))
        R.ref_lub_matrix_mono(h, lubmin2, lubmax, _p(m))
        R.ref_sys_free(h)
        return m

    def lub_columns(self, **kw):
        """the linear map of the oracle's lubrication correction, column by
        column: L[:, k] = lub(e_k)"""
        eye = np.eye(self.n)
        return np.stack([self.lub(eye[k], **kw) for k in range(self.n)], 1)

    def ref_matrix(self):
        R = ref()
        h = R.ref_sys_new(C.byref(self.s))
        m = np.zeros((self.n, self.n))
        R.ref_matrix(h, _p(m))
        R.ref_sys_free(h)
        return m


def ext_to_map(version, np_, mat):
    """the linear map of an extracted-form matrix: mat . T for FTS
    (trans_mat_ext2ext), mat itself for F and FT"""
    return trans_mat_ext2ext(np_, mat) if version == 2 else mat


def trans_mat_ext2ext(np_, mat):
    """mat . T on every 11-column particle block, T(F, T, S) = extracted
    form (ewald-3fts-matrix.c:383-437)"""
    t = np.eye(11)
    t[6, 6] = t[10, 10] = 2.0
    t[6, 10] = t[10, 6] = 1.0
    t[7, 7] = t[8, 8] = t[9, 9] = 2.0
    return mat @ np.kron(np.eye(np_), t)


def solve_mix_dense(system, nm, b):
    """Dense mixed problem of ewald-3fts-matrix.c:570-652 / 1192-1290 (and
    the same partition for F, FT): the true map A = M T split into known
    force components (mobile F, T) and known velocity components (mobile E,
    every component of the fixed particles), solve_linear of
    matrix.c:266-312 with numpy's LAPACK inverse for lapack_inv_.
    b: known values per particle (mobile F T E, fixed U O E); returns the
    unknowns (mobile U O S, fixed F T S).  Numpy restatement: LAPACK is not
    part of the reference tree."""
    npart, v = system.np, system.version
    d = dim(v)
    mat = system.matrix()
    if v == 2:
        mat = trans_mat_ext2ext(npart, mat)
    isH = np.zeros(npart * d, dtype=bool)
    for i in range(npart):
        for q in range(d):
            isH[i * d + q] = i >= nm or q >= 6
    h = np.flatnonzero(isH)
    l = np.flatnonzero(~isH)
    b = np.asarray(b, np.float64)
    x = np.empty_like(b)
    if h.size == 0:
        return mat @ b
    A = mat[np.ix_(h, h)]
    B = mat[np.ix_(h, l)]
    Cm = mat[np.ix_(l, h)]
    D = mat[np.ix_(l, l)]
    Ai = np.linalg.inv(A)
    I_, J_ = Ai, -Ai @ B
    K_, L_ = Cm @ Ai, D - Cm @ Ai @ B
    x[l] = L_ @ b[l] + K_ @ b[h]
    x[h] = J_ @ b[l] + I_ @ b[h]
    return x


def ft_mobility_dense(system):
    """(F, T) -> (U, O) particle-wise of the FTS mobility problem with E =
    0 (BD_atimes_mob_FU, brownian.c:597-646) as a dense (6 np)^2 matrix,
    from the restated dense solve (solve_mix_dense)"""
    n = system.np
    B = np.zeros((11 * n, 6 * n))
    for c in range(6 * n):
        B[11 * (c // 6) + c % 6, c] = 1.0
    X = solve_mix_dense(system, n, B)
    return X.reshape(n, 11, 6 * n)[:, :6, :].reshape(6 * n, 6 * n)


def extmat_fts(np_, m, x):
    y = np.zeros(11 * np_)
    m = np.ascontiguousarray(m, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    lib().orc_extmat_fts(int(np_), _p(m), _p(x), _p(y))
    return y


def scalars_ewald_real(version, xi, r):
    o = np.zeros(11)
    lib().orc_scalars_ewald_real(C.c_int(version), C.c_double(xi),
                                 C.c_double(r), _p(o))
    return o


def scalars_ewald_real_poly(version, xi, r, aa, ab, scale_sd=False):
    o = np.zeros(11)
    lib().orc_scalars_ewald_real_poly(C.c_int(version), C.c_double(xi),
                                      C.c_double(r), C.c_double(aa),
                                      C.c_double(ab), _p(o))
    if scale_sd:
        lib().orc_scalars_scale_SD(C.c_int(version), C.c_double(aa), _p(o))
    return o


def ref_scalars_ewald_real(version, xi, r, poly=None):
    R = ref()
    o = [C.c_double() for _ in range(11)]
    if poly is None:
        R.scalars_ewald_real(C.c_int(version), C.c_double(xi), C.c_double(r),
                             *[C.byref(v) for v in o])
    else:
        R.scalars_ewald_real_poly(C.c_int(version), C.c_double(xi),
                                  C.c_double(r), C.c_double(poly[0]),
                                  C.c_double(poly[1]),
                                  *[C.byref(v) for v in o])
    return np.array([v.value for v in o])


def cheb_coef(ncheb, func_id, x0, x1):
    a = np.zeros(ncheb)
    lib().orc_cheb_coef(C.c_int(ncheb), C.c_int(func_id), C.c_double(x0),
                        C.c_double(x1), _p(a))
    return a


def cheb_eval(a, x0, x1, x):
    return lib().orc_cheb_eval(C.c_int(len(a)), _p(np.ascontiguousarray(a)),
                               C.c_double(x0), C.c_double(x1), C.c_double(x))


class _Dense(C.Structure):
    _fields_ = [("n", C.c_int), ("m", _dp)]


def cheb_eval_dense(a, mat, y, e0, e1):
    mat = np.ascontiguousarray(mat, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    n = y.size
    z = np.zeros(n)
    d = _Dense(n, _p(mat))
    L = lib()
    cb = C.cast(L.orc_atimes_dense_cb, C.c_void_p)
    L.orc_cheb_eval_atimes(C.c_int(len(a)), _p(np.ascontiguousarray(a)),
                           C.c_int(n), _p(y), _p(z), C.c_double(e0),
                           C.c_double(e1), cb, C.byref(d))
    return z


def cheb_error_dense(kind, mat, y, z):
    """the oracle's chebyshev_error_minvsqrt (kind 0, chebyshev.c:241-269)
    or chebyshev_error_Rsqrt (kind 1, 282-310) with the dense operator mat"""
    mat = np.ascontiguousarray(mat, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    d = _Dense(y.size, _p(mat))
    L = lib()
    cb = C.cast(L.orc_atimes_dense_cb, C.c_void_p)
    f = L.orc_cheb_error_minvsqrt if kind == 0 else L.orc_cheb_error_Rsqrt
    return f(C.c_int(y.size), _p(y), _p(z), cb, C.byref(d))


def ref_cheb_error_dense(kind, mat, y, z):
    """the compiled reference's chebyshev_error_minvsqrt / _Rsqrt"""
    mat = np.ascontiguousarray(mat, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    R = ref()
    f = (R.ref_cheb_error_minvsqrt_dense if kind == 0
         else R.ref_cheb_error_Rsqrt_dense)
    return f(C.c_int(y.size), _p(mat), _p(y), _p(z))


def cheb_eval_sys(a, system, y, e0, e1):
    y = np.ascontiguousarray(y, np.float64)
    z = np.zeros(y.size)
    L = lib()
    cb = C.cast(L.orc_atimes_sys_cb, C.c_void_p)
    L.orc_cheb_eval_atimes(C.c_int(len(a)), _p(np.ascontiguousarray(a)),
                           C.c_int(y.size), _p(y), _p(z), C.c_double(e0),
                           C.c_double(e1), cb, C.byref(system.s))
    return z


class Rng(C.Structure):
    _fields_ = [("mt", C.c_ulong * 624), ("mti", C.c_int),
                ("has_saved", C.c_int), ("saved", C.c_double)]

    def __init__(self, seed):
        super().__init__()
        lib().orc_rng_seed(C.byref(self), C.c_ulong(seed))

    def int32(self):
        return lib().orc_rng_int32(C.byref(self))

    def gaussian(self, n=None):
        if n is None:
            return lib().orc_rng_gaussian(C.byref(self))
        return np.array([lib().orc_rng_gaussian(C.byref(self))
                         for _ in range(n)])


def quat_dqdt(q, o):
    """quaternion_dQdt (ode-quaternion.c:89-105) of the oracle"""
    q = np.ascontiguousarray(q, np.float64)
    o = np.ascontiguousarray(o, np.float64)
    d = np.zeros(4)
    lib().orc_quat_dqdt(_p(q), _p(o), _p(d))
    return d


# ---- lubrication (orc_lub.c) ---------------------------------------------
TWOBODY_FAMILIES = ("XA", "YA", "YB", "XC", "YC", "XG", "YG", "YH", "XM",
                    "YM", "ZM")


def twobody_table(fam, n, lam, reference=False):
    """f_0..f_n of family fam (index into TWOBODY_FAMILIES) at lambda"""
    f = np.zeros(n + 1)
    (ref().ref_twobody_table if reference else lib().orc_twobody_table)(
        fam, n, lam, _p(f))
    return f


def twobody_res(version, r, aa, ab, n=100, flag_lub=1, flag_scale=1,
                reference=False):
    """twobody_scalars_res (twobody.c:2786-2814): 22 scalars"""
    out = np.zeros(22)
    (ref().ref_twobody_res if reference else lib().orc_twobody_res)(
        version, r, aa, ab, n, flag_lub, flag_scale, _p(out))
    return out


def two_body_res(s, reference=False):
    """scalar_two_body_res (two-body-res.c:31): 22 scalars"""
    out = np.zeros(22)
    (ref().ref_two_body_res if reference else lib().orc_two_body_res)(
        s, _p(out))
    return out


def minv_fts(s, reference=False):
    """scalar_minv_fts (fts.c:1490): 22 scalars"""
    out = np.zeros(22)
    (ref().ref_minv_fts if reference else lib().orc_minv_fts)(s, _p(out))
    return out


def minv_poly(version, r, aa, ab):
    """scalars_minv_*_poly (minv-poly.c): 44 scalars (11, 12, 21, 22)"""
    out = np.zeros(44)
    lib().orc_minv_poly(version, r, aa, ab, _p(out))
    return out


def nonewald_poly_full(version, r, aa, ab):
    """the compiled reference's scalars_nonewald_poly_full (non-ewald.c:
    230-318): 44 scalars"""
    out = np.zeros(44)
    ref().scalars_nonewald_poly_full(version, r, aa, ab, _p(out))
    return out


def fts_block(x, e, sc, self_, reference=False):
    """y = B(e; sc) x, matrix_fts_atimes / matrix_fts_self_atimes"""
    x = np.ascontiguousarray(x, np.float64)
    e = np.ascontiguousarray(e, np.float64)
    sc = np.ascontiguousarray(sc, np.float64)
    y = np.zeros(11)
    if reference:
        ref().ref_fts_block(int(self_), _p(x), _p(y), _p(e), _p(sc))
    else:
        lib().orc_block_atimes_sign(2, _p(x), _p(y), e[0], e[1], e[2], _p(sc),
                                    int(self_))
    return y
