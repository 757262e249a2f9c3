"""Benchmark: FTS Ewald mobility matvecs per second at N = 1024.

Workload (the 11N x 11N FTS configuration the headline metric is quoted
on): N = 1024 monodisperse spheres (a = 1) at volume fraction 0.3 in a
cubic box, Ewald tolerance 1e-12, F/T/S ~ N(0,1), seeded synthetic
positions (non-overlapping, gap >= 2.1a).  One step = one call of the
batched matvec Y = M(x) . X on B = 128 right-hand sides (11N x 128), from
device-resident positions and forces (positions re-read every call).

value   = matvecs/s (128 per step) of the MIXED batched path (FP32 real
          space and reciprocal sums on the matrix cores, FP64 k-blocks,
          partial sums and self term), at the Ewald splitting chosen by
          timing the batched call, with the largest per-column relative
          error against the FP64 GPU path reported (must be <= 1e-6).
fp64    = matvecs/s of the FP64 GPU path: batched at the same splitting,
          single right-hand side at its own tuned splitting and at the
          reference's default splitting (tr = 1); mixed_over_fp64 divides
          value by the best of these.
single  = the same figures for one right-hand side per call.
roofline: the real-space part of the batched MIXED call (for equal
          spheres k_real_moments + k_real_mfma_mom: FP32 image moments per
          pair, then the pair blocks times the columns on the FP32 matrix
          cores); algorithmic flops = pair-images inside r_max x 300 +
          pairs with an image inside r_max x 2 x 121 x B; time from HIP
          events around that part on the library stream.
cpu_baseline: the oracle (FP64 C restatement of the reference loops, one
          core) timed on a bounded row sample of the same workload, with
          the oracle's time over the compiled reference's on the
          reference's own CPU case measured beside it.
bd, ensemble: the BD step (N = 2048 F) and the 256-box ensemble figures.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import libstokes_amd as ls  # noqa: E402
from libstokes_amd import workloads  # noqa: E402

METRIC = ("FTS Ewald M·f/sec at N=1024 (mixed vs FP64 GPU, rel err"
          "≤1e-6); BD steps/s")
PEAK_FP32 = 157.3   # TFLOP/s, MI355X (vector and dense matrix FP32)
PEAK_FP64 = 78.6    # TFLOP/s
FLOP_PER_IMAGE_FTS = 300.0


def pair_images_in_cutoff(pos, s):
    """exact count of (i, j, image) with 0 < r^2 <= r_max^2 (ewald.c:688-690)
    and of the pairs (i, j) with at least one such image"""
    rl, _ = s.image_table()
    rmax2 = s.s.rmax2
    d0 = pos[None, :, :] - pos[:, None, :]
    d0 = d0.reshape(-1, 3)
    count = 0
    hit = np.zeros(len(d0), bool)
    for m in range(len(rl)):
        d = d0 + rl[m]
        r2 = np.einsum("ij,ij->i", d, d)
        inside = (r2 <= rmax2) & (r2 > 0.0)
        count += int(np.count_nonzero(inside))
        hit |= inside
    return count, int(np.count_nonzero(hit))


def measured_traffic(kernel, n, tr, nrhs=1):
    """HBM bytes per launch of `kernel` at (n, tr) from the committed counter
    summary profiles/traffic.txt (FETCH_SIZE and WRITE_SIZE per dispatch,
    each scaled by the calibration copy), or None"""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 4 and not line.startswith("#"):
            kv = dict(x.split("=", 1) for x in f[1:] if "=" in x)
            if (f[0] == kernel.replace(" ", "") and int(kv.get("n", -1)) == n
                    and float(kv.get("tr", -1)) == float(tr)
                    and int(kv.get("nrhs", 1)) == nrhs):
                return float(kv["bytes"])
    return None


def time_matvec(s, xd, yd, steps, warmup, nrhs=1):
    for _ in range(warmup):
        s.atimes_dev(xd, yd, nrhs)
    s.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.atimes_dev(xd, yd, nrhs)
    s.sync()
    return (time.perf_counter() - t0) / steps


def kernel_times(s, xd, yd, reps, nrhs=1):
    """mean (real-space kernel, rest) ms from the library's HIP events"""
    s.set_timing(True)
    real, recip = [], []
    for _ in range(reps):
        s.atimes_dev(xd, yd, nrhs)
        _, r, k = s.times()
        real.append(r)
        recip.append(k)
    s.set_timing(False)
    return float(np.mean(real)), float(np.mean(recip))


def cpu_baseline(pos, L, x, tr, budget_s):
    """oracle rows [0, r) of the same workload, one core"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    o = po.System(2, pos, box=(L, L, L), tr=tr)
    n = len(pos)
    rows = 1
    t = 0.0
    while True:
        t0 = time.perf_counter()
        o.atimes(x, rows=(0, rows))
        t = time.perf_counter() - t0
        if t * 4 > budget_s or rows * 4 > n:
            break
        rows *= 4
    rate = (rows / n) / t
    return dict(value=rate, unit="matvecs/s", cores=1, kind="port",
                sample=f"oracle FP64 rows 0..{rows - 1} of the N={n} FTS "
                       f"matvec (tr={tr:g}), {t:.1f} s; rate = rows/N/time",
                host=host_cpu())


def oracle_over_reference():
    """the oracle's time over the compiled reference's (oracle/_ref) for one
    FP64 FTS matvec of the reference's CPU-runnable case (N = 64, phi = 0.1,
    tr = 1), both on one core of this host; None without oracle/_ref"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    if not po.ref_available():
        return None
    pos, L = workloads.configuration(64, 0.1, seed=21)
    x = np.random.default_rng(3).standard_normal(11 * 64)
    o = po.System(2, pos, box=(L, L, L), tr=1.0)
    o.atimes(x)
    o.ref_atimes(x)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        y1 = o.atimes(x)
    ta = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        y2 = o.ref_atimes(x)
    tb = (time.perf_counter() - t0) / reps
    return {"case": "FTS N=64 phi=0.1 tr=1, one core", "oracle_s": ta,
            "reference_s": tb, "oracle_over_reference": ta / tb,
            "max_rel_diff": float(np.abs(y1 - y2).max() / np.abs(y2).max())}


def host_cpu():
    model = "?"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible"


def cpu_baseline_threads(pos, L, x, tr, budget_s, threads):
    """oracle rows spread over host threads (row ranges are independent;
    the oracle call releases the GIL)"""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    o = po.System(2, pos, box=(L, L, L), tr=tr)
    n = len(pos)
    per = 1
    with ThreadPoolExecutor(threads) as ex:
        while True:
            rows = min(n, per * threads)
            bounds = [(i * rows // threads, (i + 1) * rows // threads)
                      for i in range(threads)]
            t0 = time.perf_counter()
            list(ex.map(lambda b: o.atimes(x, rows=b), bounds))
            t = time.perf_counter() - t0
            if t * 4 > budget_s or rows * 4 > n:
                break
            per *= 4
    return dict(value=(rows / n) / t, unit="matvecs/s", cores=threads,
                kind="port",
                sample=f"oracle FP64 rows 0..{rows - 1} of the N={n} FTS "
                       f"matvec (tr={tr:g}) over {threads} threads, "
                       f"{t:.1f} s; rate = rows/N/time", host=host_cpu())


def bench_bd(args, log):
    """BD steps/s: mid-point steps (brownian.c:2352-2560), F version,
    N = 2048 spheres at phi = 0.2, Chebyshev order 30, dt = 1e-3, Pe = 1;
    spectral bounds by Lanczos on the first (untimed) step"""
    n, phi = 2048, 0.2
    pos, L = workloads.configuration(n, phi, seed=2)
    out = {"workload": f"BD mid-point F N={n} phi={phi} n_cheb=30 dt=1e-3 "
                       f"Pe=1 rmin=1 (31 operator products per step)"}
    for name, prec in (("mixed", ls.MIXED), ("fp64", ls.FP64)):
        # r clamped at contact (sys->rmin = 1): without an excluded-volume
        # force this is how the reference runs dense BD; it also skips the
        # overlap-driven dt control (brownian.c:2436, 2502)
        # splitting: timed on whole steps (32 of the 33 products of a step
        # use the dense real-space matrix at x, so the balance differs from
        # a lone product); the result stays that of tr = 1
        best = None
        for tr in (30.0, 100.0, 300.0, 1000.0, 3000.0):
            s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0, precision=prec,
                          rmin=1.0)
            s.set_split(tr)
            xt = ls.DeviceArray.from_host(pos.reshape(-1))
            bdt = ls.BD(s, peclet=1.0, n_cheb=30, eps=5e-2, dt_lim=1e-6,
                        rmin_ol=1.0, seed=11)
            try:
                bdt.step(xt, 1e-3)
                s.sync()
                t0 = time.perf_counter()
                for _ in range(3):
                    bdt.step(xt, 1e-3)
                s.sync()
                t = (time.perf_counter() - t0) / 3
            except ls.StokesError:
                t = float("inf")
            log(f"BD {name} tr={tr:g}: {t * 1e3:.3f} ms/step")
            if best is None or t < best[1]:
                best = (tr, t)
            bdt.close()
            xt.free()
            s.close()
        tr_bd = best[0]
        s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0, precision=prec,
                      rmin=1.0)
        s.set_split(tr_bd)
        xd = ls.DeviceArray.from_host(pos.reshape(-1))
        # order 30 on this spectrum (condition ~10^2) is accurate to ~1e-2
        # in z.Mz: eps is set so the fit is not redone every step
        bd = ls.BD(s, peclet=1.0, n_cheb=30, eps=5e-2, dt_lim=1e-6,
                   rmin_ol=1.0, seed=11)
        try:
            for _ in range(2):
                info = bd.step(xd, 1e-3)
                log(f"BD {name} warm-up: {info.summary()}")
        except ls.StokesError as e:
            out[name] = {"error": str(e)}
            log(f"BD {name}: {e}")
            bd.close()
            s.close()
            continue
        s.sync()
        att = halv = redr = 0
        t0 = time.perf_counter()
        for _ in range(args.bd_steps):
            info = bd.step(xd, 1e-3)
            att += info.attempts
            halv += info.halvings
            redr += info.redraws
        s.sync()
        t = time.perf_counter() - t0
        out[name] = {"steps_per_s": args.bd_steps / t,
                     "ms_per_step": t / args.bd_steps * 1e3,
                     "attempts_per_step": att / args.bd_steps,
                     "halvings": halv, "redraws": redr,
                     "err_cheb_last": info.err_cheb, "tr": tr_bd,
                     "eig": [info.eig[0], info.eig[1]]}
        log(f"BD {name}: {out[name]['ms_per_step']:.3f} ms/step")
        bd.close()
        # 64 noise vectors as one 3N x 64 block per Chebyshev recursion
        # (z = M^{-1/2} y, order 30, the bounds of the run above)
        e0, e1 = info.eig
        from libstokes_amd import cheb_coef_inv_sqrt
        acoef = cheb_coef_inv_sqrt(30, e0, e1)
        nb = 64
        yb = ls.DeviceArray.from_host(
            np.random.default_rng(4).standard_normal(nb * 3 * n))
        zb = ls.DeviceArray(nb * 3 * n)
        s.cheb_dev(acoef, yb, zb, e0, e1, nrhs=nb)
        s.sync()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            s.cheb_dev(acoef, yb, zb, e0, e1, nrhs=nb)
        s.sync()
        tc = (time.perf_counter() - t0) / reps
        # the same block with the real-space part as a dense matrix built at
        # this configuration (the build is inside the timed region): each of
        # the 30 products is one GEMM plus the matrix-free reciprocal part
        # the GEMM cost does not depend on the splitting, the reciprocal
        # part does: the splitting is chosen for this variant on its own
        def cached_block():
            s.set_real_cache(True)
            s.cheb_dev(acoef, yb, zb, e0, e1, nrhs=nb)
            s.set_real_cache(False)
        tcc, trc = None, None
        for trx in (10.0, 30.0, 100.0, 300.0, 1000.0):
            s.set_split(trx)
            cached_block()
            s.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                cached_block()
            s.sync()
            tx = (time.perf_counter() - t0) / reps
            log(f"cheb block 64 {name} cached tr={trx:g}: {tx * 1e3:.3f} ms")
            if tcc is None or tx < tcc:
                tcc, trc = tx, trx
        s.set_split(trc)
        cached_block()
        s.sync()
        zc = zb.to_host()
        s.set_split(tr_bd)
        s.cheb_dev(acoef, yb, zb, e0, e1, nrhs=nb)
        s.sync()
        out[name]["cheb_block64_cached_vs_matrix_free"] = float(
            np.linalg.norm(zc - zb.to_host()) / np.linalg.norm(zb.to_host()))
        out[name]["cheb_block64_matrix_free_ms"] = tc * 1e3
        out[name]["cheb_block64_cached_ms"] = tcc * 1e3
        out[name]["cheb_block64_cached_tr"] = trc
        tc = min(tc, tcc)
        out[name]["cheb_block64_ms"] = tc * 1e3
        out[name]["noise_vectors_per_s_block64"] = nb / tc
        # accuracy of column 0: |z.Mz - y.y| / z.Mz (chebyshev.c:241-269)
        y0 = yb.to_host()[:3 * n]
        z0 = zb.to_host()[:3 * n]
        mz = s.atimes(z0)
        out[name]["cheb_block64_err_col0"] = abs(z0 @ mz - y0 @ y0) / (z0 @ mz)
        # the same block for z = M^{1/2} y (the other semantics of the noise:
        # the drift-free Brownian displacement), same number of products;
        # accuracy |y.My - z.z| / y.My (chebyshev.c:282-310)
        from libstokes_amd import cheb_coef_sqrt
        asq = cheb_coef_sqrt(30, e0, e1)
        use_cache = tcc <= out[name]["cheb_block64_matrix_free_ms"] * 1e-3
        s.set_split(trc if use_cache else tr_bd)

        def sqrt_block():
            if use_cache:
                s.set_real_cache(True)
            s.cheb_dev(asq, yb, zb, e0, e1, nrhs=nb)
            if use_cache:
                s.set_real_cache(False)
        sqrt_block()
        s.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            sqrt_block()
        s.sync()
        ts = (time.perf_counter() - t0) / reps
        s.set_split(tr_bd)
        z0 = zb.to_host()[:3 * n]
        my = s.atimes(y0)
        out[name]["cheb_block64_sqrt_ms"] = ts * 1e3
        out[name]["cheb_block64_sqrt_err_col0"] = abs(y0 @ my - z0 @ z0) / (y0 @ my)
        log(f"cheb block 64 {name}: {tc * 1e3:.3f} ms (1/sqrt), "
            f"{ts * 1e3:.3f} ms (sqrt)")
        s.close()
    # the same steps with the overlap check on (sys->rmin = 0: no contact
    # clamp; an overlap after a step halves dt or redraws the noise,
    # brownian.c:2436-2560), MIXED, at the splitting of the run above;
    # dt_lim = dt / 8, so that an attempt chain reaches a redraw after three
    # halvings (without an excluded-volume force a pair near contact is
    # separated by a new draw, not by a shorter step)
    tr_ol = out.get("mixed", {}).get("tr", 1000.0)
    s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0, precision=ls.MIXED, rmin=0.0)
    s.set_split(tr_ol)
    xd = ls.DeviceArray.from_host(pos.reshape(-1))
    bd = ls.BD(s, peclet=1.0, n_cheb=30, eps=5e-2, dt_lim=1.25e-4,
               rmin_ol=1.0, seed=11)
    # without an excluded-volume force, pairs collect near contact and each
    # attempt's chance of a clean step falls; the library stops after 64
    # attempts (the reference retries without limit), so the figure is the
    # rate over the steps done before that, with the count
    att = halv = redr = done = 0
    err = None
    s.sync()
    t0 = time.perf_counter()
    for _ in range(args.bd_steps):
        try:
            info = bd.step(xd, 1e-3)
        except ls.StokesError as e:
            err = str(e)
            break
        att += info.attempts
        halv += info.halvings
        redr += info.redraws
        done += 1
    s.sync()
    t = time.perf_counter() - t0
    out["mixed_overlap_check"] = {
        "rmin": 0.0, "dt_lim": 1.25e-4, "tr": tr_ol, "steps_done": done,
        "steps_per_s": done / t if done else 0.0,
        "attempts_per_step": att / done if done else None,
        "halvings": halv, "redraws": redr, "stopped": err}
    log(f"BD mixed, overlap check on: {done} steps, "
        f"{att / max(done, 1):.2f} attempts/step" + (f"; {err}" if err else ""))
    bd.close()
    xd.free()
    s.close()
    return out


def _poly_box(args):
    n, seed = args
    return workloads.configuration_poly(n, 0.25, seed=seed)


def ensemble_boxes(nbox, log, workers=8):
    """the ensemble's configurations: N log-uniform on [50, 2000], radii
    log-uniform on [0.5, 2], phi = 0.25 (generated by worker processes
    before this process first touches the GPU; workers 0: in this process)"""
    import multiprocessing as mp
    rng = np.random.default_rng(0)
    ns = np.exp(rng.uniform(math.log(50), math.log(2000), nbox)).astype(int)
    t0 = time.perf_counter()
    work = [(int(n), 1000 + k) for k, n in enumerate(ns)]
    if workers > 0:
        with mp.get_context("spawn").Pool(workers) as pool:
            boxes = pool.map(_poly_box, work)
    else:
        boxes = [_poly_box(w) for w in work]
    log(f"ensemble: {nbox} configurations in {time.perf_counter() - t0:.1f} s")
    return boxes


def bench_dense_build(args, pos, L, posd, log):
    """the dense FP64 extracted-form mobility matrix of the headline system
    (make_matrix_mob_ewald_3all, N = 1024 FTS: (11 N)^2 doubles = 1.015 GB),
    timed at the reference's splitting (tr = 1) and at the fastest of a few
    other splittings (results those of tr = 1); checked against the
    matrix-free FP64 product at that splitting (mat . T(x) = M x)"""
    n = len(pos)
    nd = 11 * n
    mat = ls.DeviceArray(nd * nd)
    x = np.random.default_rng(7).standard_normal(nd)
    xd = ls.DeviceArray.from_host(x)
    yd = ls.DeviceArray(nd)
    res = {}
    best = None
    for tr in ([1.0] + ([args.tr] if args.tr else [3.0, 10.0, 30.0])):
        s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, precision=ls.FP64)
        s.set_split(0.0 if tr == 1.0 else tr)
        ls.check(s.L.stokes_hip_set_pos_dev(s.sys, posd.ptr), "set_pos_dev")
        s.matrix_dev(mat)                      # first call: set-up, buffers
        s.sync()
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):
            s.matrix_dev(mat)
        s.sync()
        t = (time.perf_counter() - t0) / reps
        res[f"{tr:g}"] = t * 1e3
        log(f"dense build tr={tr:g}: {t * 1e3:.2f} ms")
        if best is None or t < best[1]:
            best = (tr, t)
        s.close()
    tr, t = best
    s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, precision=ls.FP64)
    s.set_split(0.0 if tr == 1.0 else tr)
    ls.check(s.L.stokes_hip_set_pos_dev(s.sys, posd.ptr), "set_pos_dev")
    s.matrix_dev(mat)
    s.extmat_dev(mat, xd, yd)
    ym = yd.to_host()
    s.atimes_dev(xd, yd)
    yf = yd.to_host()
    s.close()
    err = float(np.linalg.norm(ym - yf) / np.linalg.norm(yf))
    nbytes = 8.0 * nd * nd
    # algorithmic work of the symmetric build at the splitting kept: each
    # unordered pair image inside r_max once (300 flop; the images of (j, i)
    # are those of (i, j) reflected) and, per unordered pair (i <= j) and k,
    # cos and sin of k.(x_j - x_i) from the phase table (6 flop) and the 68
    # moments accumulated (136 flop); the 11 x 11 blocks are a fixed linear
    # map of the moments, written twice
    sh = ls.Stokes(2, pos, box=(L, L, L), tr=tr)
    pin, _ = pair_images_in_cutoff(pos, sh)
    rl, _ = sh.image_table()
    r2 = np.einsum("ij,ij->i", rl, rl)
    pself = n * int(np.count_nonzero((r2 > 0.0) & (r2 <= sh.s.rmax2)))
    nk = half_space_k(sh)
    sh.close()
    upairs = (pin - pself) / 2 + pself
    flop = upairs * FLOP_PER_IMAGE_FTS + n * (n + 1) / 2 * nk * 142.0
    achieved = flop / t / 1e12
    mat.free()
    return {"workload": f"dense FP64 FTS mobility matrix N={n} phi={args.phi} "
                        f"((11N)^2 doubles, {nbytes / 1e9:.3f} GB)",
            "ms_by_tr": res, "tr": tr, "ms": t * 1e3,
            "builds_per_s": 1.0 / t,
            "write_gb_per_s": nbytes / t / 1e9,
            "rel_err_matrix_vs_matrix_free": err,
            "roofline": {"bound": "fp64", "achieved": achieved,
                         "peak": PEAK_FP64, "unit": "TFLOP/s",
                         "frac": achieved / PEAK_FP64, "traffic": None,
                         "kernel": "k_matrix_sym<2> (FP64 vector, pairs i <= j "
                                   "once: unordered pair images x 300 + "
                                   "N (N + 1) / 2 nk x 142 flop; the 1.015 GB "
                                   "write alone would take 0.13 ms at 8 TB/s)",
                         "unordered_pair_images": upairs, "nk": nk}}


def half_space_k(s):
    """reciprocal vectors of the half space with |k| <= k_max (the k table of
    the operator at the system's own splitting, cubic box)"""
    st = s.s
    kx = st.kmaxx
    m = np.arange(-kx, kx + 1)
    m1, m2, m3 = np.meshgrid(np.arange(0, kx + 1), m, m, indexing="ij")
    kk = (2.0 * math.pi / st.lx) ** 2 * (m1 ** 2 + m2 ** 2 + m3 ** 2)
    half = (m1 > 0) | ((m1 == 0) & (m2 > 0)) | ((m1 == 0) & (m2 == 0) & (m3 > 0))
    return int(np.count_nonzero(half & (np.sqrt(kk) <= st.kmax)))


def pair_images_poly(pos, L, rmax):
    """exact count of (i, j, image) with 0 < r <= rmax in a cubic periodic
    box of side L (every lattice image: the reference's table holds all
    cells within reach for an unsheared cubic box)"""
    from scipy.spatial import cKDTree
    n = int(np.ceil(rmax / L))
    shifts = np.array([(a, b, c) for a in range(-n, n + 1)
                       for b in range(-n, n + 1)
                       for c in range(-n, n + 1)], float) * L
    cloud = (pos[None, :, :] + shifts[:, None, :]).reshape(-1, 3)
    lo, hi = pos.min(0) - rmax, pos.max(0) + rmax
    keep = np.all((cloud >= lo) & (cloud <= hi), axis=1)
    t_pos, t_cloud = cKDTree(pos), cKDTree(cloud[keep])
    # pairs at distance 0 are each particle with itself (primary image)
    return int(t_pos.count_neighbors(t_cloud, rmax)) - len(pos)


def bench_ensemble(args, boxes, log):
    """Batched ensemble (C4): nbox independent FTS boxes (ensemble_boxes),
    MIXED mobility product and the FP64 lubrication correction (lubmax 3 in
    units of the mean radius: gaps below one radius, series to order 100)
    in one graph launch of batched kernels, every box at the splitting and
    the kernels at the tile shape the ensemble autotuner chose (results
    those of each box's own tr = 1 splitting)"""
    nbox = len(boxes)
    systems = []
    for pos, a, L in boxes:
        s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, a=a, precision=ls.MIXED)
        s.set_lub(4.0 + 1e-12, 3.0, 100, 1)
        systems.append(s)
    ens = ls.Ensemble(systems)
    t0 = time.perf_counter()
    trs = ens.autotune(trs=[10.0, 30.0, 100.0, 300.0, 1000.0, 3000.0])
    vals, cnt = np.unique(trs, return_counts=True)
    tuned = {f"{v:g}": int(c) for v, c in zip(vals, cnt)}
    tiles = ens.tiles()
    log(f"ensemble: splittings and tiles tuned in "
        f"{time.perf_counter() - t0:.1f} s: {tuned}, tiles {tiles}")
    tot = ens.offsets[-1]
    posd = ls.DeviceArray.from_host(
        np.concatenate([b[0].reshape(-1) for b in boxes]))
    xd = ls.DeviceArray.from_host(
        np.random.default_rng(1).standard_normal(11 * tot))
    yd = ls.DeviceArray(11 * tot)
    yl = ls.DeviceArray(11 * tot)
    reps = max(3, args.steps // 4)

    def timed(fn):
        fn()                               # capture + first launch
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps

    t_mob = timed(lambda: ens.atimes_dev(posd, xd, yd))
    t_all = timed(lambda: ens.atimes_lub_dev(posd, xd, yd, yl))
    log(f"ensemble: {t_all * 1e3:.2f} ms per launch with lubrication, "
        f"{t_mob * 1e3:.2f} ms mobility only")
    ms = np.min([ens.phase_ms(posd, xd, yd, yl) for _ in range(3)], axis=0)
    # algorithmic work of the real-space phase: every pair image inside
    # each box's own r_max at its tuned splitting
    t0 = time.perf_counter()
    pin = 0
    for (pos, a, L), tr in zip(boxes, trs):
        sh = ls.Stokes(2, pos, box=(L, L, L), tr=float(tr) if tr > 0 else 1.0,
                       a=a)
        pin += pair_images_poly(pos, L, math.sqrt(sh.s.rmax2))
        sh.close()
    log(f"ensemble: {pin} pair images counted in "
        f"{time.perf_counter() - t0:.1f} s")
    peak_valu = PEAK_FP64          # FP32 FMA without packed math, per lane
    achieved = pin * FLOP_PER_IMAGE_FTS / (ms[0] * 1e-3) / 1e12
    out = {"workload": f"{nbox} FTS boxes, N log-uniform [50, 2000] "
                       f"(sum {tot}), radii log-uniform [0.5, 2], phi 0.25, "
                       f"MIXED mobility + FP64 lubrication (lubmax 3), one "
                       f"graph launch of batched kernels",
           "nbox": nbox, "sum_n": int(tot), "ms_per_launch": t_all * 1e3,
           "ms_per_launch_mobility_only": t_mob * 1e3,
           "phase_ms_eager": {"real": ms[0], "reciprocal": ms[1],
                              "combine": ms[2], "lubrication": ms[3]},
           "tuned_tr_counts": tuned,
           "tiles": {"chunks_per_wave": tiles[0],
                     "particles_per_sf_split": tiles[1],
                     "k_per_back_split": tiles[2]},
           "box_products_per_s": nbox / t_all,
           "particle_products_per_s": tot / t_all,
           "roofline": {"bound": "valu", "achieved": achieved,
                        "peak": peak_valu, "unit": "TFLOP/s",
                        "frac": achieved / peak_valu, "traffic": None,
                        "kernel": "k_ens_real<2, float> (FP32 pair scalars "
                                  "and contraction, all boxes in one launch;"
                                  " peak: FP32 FMA without packed math)",
                        "kernel_ms": ms[0], "pair_images": pin,
                        "flop_per_pair_image": FLOP_PER_IMAGE_FTS}}
    ens.close()
    for s in systems:
        s.close()
    return out


def profile_run(args):
    """kernels for the counter runs: k_copy_calibrate on 256 MiB (a known
    byte count), then --steps matvecs of both precisions at the splitting
    of --tr (system at the reference's tr = 1, as in the timed run)"""
    n = args.n
    pos, L = workloads.configuration(n, args.phi, seed=1)
    x = np.random.default_rng(2).standard_normal(11 * n)
    nb = 256 << 20
    a = ls.DeviceArray(nb // 8)
    b = ls.DeviceArray(nb // 8)
    ls.check(ls.lib().stokes_hip_copy_calibrate(b.ptr, a.ptr, nb), "copy")
    posd = ls.DeviceArray.from_host(pos)
    xd = ls.DeviceArray.from_host(x)
    yd = ls.DeviceArray(11 * n)
    tr = args.tr if args.tr else 3000.0
    nrhs = args.profile_batch
    if nrhs > 1:
        xd = ls.DeviceArray.from_host(
            np.random.default_rng(3).standard_normal(nrhs * 11 * n))
        yd = ls.DeviceArray(nrhs * 11 * n)
    for prec in (ls.MIXED, ls.FP64):
        s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, precision=prec)
        s.set_split(0.0 if tr == 1.0 else tr)
        ls.check(s.L.stokes_hip_set_pos_dev(s.sys, posd.ptr), "set_pos_dev")
        t = time_matvec(s, xd, yd, args.steps, args.warmup, nrhs=nrhs)
        print(json.dumps({"precision": "mixed" if prec else "fp64", "tr": tr,
                          "nrhs": nrhs, "ms_per_call": t * 1e3,
                          "copy_calibration_bytes": nb}), flush=True)
        s.close()


def profile_bd_run(args):
    """short run for rocprofv3: --steps BD steps (F, N=2048, phi=0.2, order
    30) per precision at the splitting --tr (default 1000), then one
    64-column Chebyshev block per precision with the cached real-space
    operator at tr = 10"""
    n = 2048
    pos, L = workloads.configuration(n, 0.2, seed=2)
    tr = args.tr if args.tr else 1000.0
    for name, prec in (("mixed", ls.MIXED), ("fp64", ls.FP64)):
        s = ls.Stokes(0, pos, box=(L, L, L), tr=1.0, precision=prec, rmin=1.0)
        s.set_split(tr)
        xd = ls.DeviceArray.from_host(pos.reshape(-1))
        bd = ls.BD(s, peclet=1.0, n_cheb=30, eps=5e-2, dt_lim=1e-6,
                   rmin_ol=1.0, seed=11)
        for _ in range(2):
            bd.step(xd, 1e-3)
        s.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            info = bd.step(xd, 1e-3)
        s.sync()
        t = (time.perf_counter() - t0) / args.steps
        e0, e1 = info.eig
        bd.close()
        from libstokes_amd import cheb_coef_inv_sqrt
        a = cheb_coef_inv_sqrt(30, e0, e1)
        yb = ls.DeviceArray.from_host(
            np.random.default_rng(4).standard_normal(64 * 3 * n))
        zb = ls.DeviceArray(64 * 3 * n)
        s.set_split(10.0)
        s.set_real_cache(True)
        s.cheb_dev(a, yb, zb, e0, e1, nrhs=64)
        s.set_real_cache(False)
        s.sync()
        print(json.dumps({"precision": name, "bd_tr": tr,
                          "bd_ms_per_step": t * 1e3, "bd_steps": args.steps,
                          "block_tr": 10.0}), flush=True)
        s.close()


def profile_ens_run(args):
    """short run for rocprofv3: the ensemble workload (--ens-boxes boxes,
    MIXED mobility + FP64 lubrication) at one splitting for every box (--tr,
    default 300) and the tile shape (4, 1024, 256): one graph launch, then
    --steps eager launches on one stream (kernels one after another)"""
    log = lambda m: print(m, file=sys.stderr, flush=True)
    boxes = ensemble_boxes(args.ens_boxes, log, workers=0)
    systems = []
    for pos, a, L in boxes:
        s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, a=a, precision=ls.MIXED)
        s.set_lub(4.0 + 1e-12, 3.0, 100, 1)
        systems.append(s)
    ens = ls.Ensemble(systems)
    tr = args.tr if args.tr else 300.0
    ens.set_split(tr)
    ens.set_tiles(4, 1024, 256)
    tot = ens.offsets[-1]
    posd = ls.DeviceArray.from_host(
        np.concatenate([b[0].reshape(-1) for b in boxes]))
    xd = ls.DeviceArray.from_host(
        np.random.default_rng(1).standard_normal(11 * tot))
    yd = ls.DeviceArray(11 * tot)
    yl = ls.DeviceArray(11 * tot)
    ens.atimes_lub_dev(posd, xd, yd, yl)
    # eager launches on one stream (each kernel's duration its own)
    ms = np.min([ens.phase_ms(posd, xd, yd, yl) for _ in range(args.steps)],
                axis=0)
    print(json.dumps({"ensemble_boxes": len(boxes), "sum_n": int(tot),
                      "tr": tr, "tiles": [4, 1024, 256],
                      "phase_ms": {"real": ms[0], "reciprocal": ms[1],
                                   "combine": ms[2], "lubrication": ms[3]},
                      "steps": args.steps}), flush=True)
    ens.close()
    for s in systems:
        s.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--phi", type=float, default=0.3)
    ap.add_argument("--tr", type=float, default=None,
                    help="Ewald splitting for the tuned runs (default: tune)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the dense-matrix build figure")
    ap.add_argument("--only-dense", action="store_true",
                    help="only the dense-matrix build figure")
    ap.add_argument("--batch", type=int, default=128,
                    help="right-hand sides per launch for the batched figures")
    ap.add_argument("--ens-boxes", type=int, default=256,
                    help="boxes of the batched-ensemble figure (0: skip)")
    ap.add_argument("--bd-steps", type=int, default=100,
                    help="timed BD steps per precision (0: skip BD)")
    ap.add_argument("--profile-batch", type=int, default=1,
                    help="right-hand sides per call in --profile mode")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: a 256 MiB calibration "
                         "copy, then --steps mixed and FP64 matvecs at --tr")
    ap.add_argument("--profile-bd", action="store_true",
                    help="short run for rocprofv3: --steps BD steps of each "
                         "precision at --tr, one cached 64-column Chebyshev "
                         "block of each")
    ap.add_argument("--profile-ens", action="store_true",
                    help="short run for rocprofv3: --steps launches of the "
                         "ensemble (mobility + lubrication) at --tr")
    args = ap.parse_args()
    if args.profile_ens:
        return profile_ens_run(args)
    if args.profile:
        return profile_run(args)
    if args.profile_bd:
        return profile_bd_run(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; "
                          "--gpus must be 1"}))
        sys.exit(2)

    log = lambda m: print(m, file=sys.stderr, flush=True)
    # CPU-only set-up first: the worker processes start before this process
    # touches the GPU
    boxes = ensemble_boxes(args.ens_boxes, log) if args.ens_boxes > 0 else None
    n = args.n
    pos, L = workloads.configuration(n, args.phi, seed=1)
    x = np.random.default_rng(2).standard_normal(11 * n)
    posd = ls.DeviceArray.from_host(pos)
    xd = ls.DeviceArray.from_host(x)
    yd = ls.DeviceArray(11 * n)
    if args.only_dense:
        print(json.dumps({"dense_build": bench_dense_build(args, pos, L, posd,
                                                           log)}))
        return

    def make(tr, prec):
        # the system keeps the reference's default splitting (tr = 1); the
        # operator is computed at the splitting of tr, which leaves the
        # result that of tr = 1 to the Ewald tolerance
        s = ls.Stokes(2, pos, box=(L, L, L), tr=1.0, precision=prec)
        s.set_split(0.0 if tr == 1.0 else tr)
        ls.check(s.L.stokes_hip_set_pos_dev(s.sys, posd.ptr), "set_pos_dev")
        return s

    # FP64 at the reference's default splitting
    s64_ref = make(1.0, ls.FP64)
    t64_ref = time_matvec(s64_ref, xd, yd, max(3, args.steps // 4), 1)
    y64_ref = yd.to_host()
    log(f"fp64 tr=1: {t64_ref * 1e3:.3f} ms")
    trs = [args.tr] if args.tr else [30.0, 50.0, 100.0, 300.0, 1000.0, 3000.0, 1e4]

    # ---- one right-hand side per call: splitting tuned on the mixed path
    best = None
    for tr in trs:
        s = make(tr, ls.MIXED)
        t = time_matvec(s, xd, yd, 3, 1)
        log(f"single mixed tr={tr:g}: {t * 1e3:.3f} ms")
        if best is None or t < best[1]:
            best = (tr, t)
        s.close()
    tr1 = best[0]
    s64 = make(tr1, ls.FP64)
    t64 = time_matvec(s64, xd, yd, args.steps, args.warmup)
    y64 = yd.to_host()
    s64.close()
    smix = make(tr1, ls.MIXED)
    tmix = time_matvec(smix, xd, yd, args.steps, args.warmup)
    ymix = yd.to_host()
    smix.close()
    err1 = float(np.linalg.norm(ymix - y64) / np.linalg.norm(y64))
    err_ref_split = float(np.linalg.norm(y64 - y64_ref) / np.linalg.norm(y64_ref))
    err_mix_tr1 = float(np.linalg.norm(ymix - y64_ref) / np.linalg.norm(y64_ref))

    # ---- batched: B right-hand sides per call, splitting tuned on it
    nb = args.batch
    X = np.random.default_rng(3).standard_normal(nb * 11 * n)
    xb = ls.DeviceArray.from_host(X)
    yb = ls.DeviceArray(nb * 11 * n)
    bestb = None
    for tr in trs:
        s = make(tr, ls.MIXED)
        t = time_matvec(s, xb, yb, 3, 1, nrhs=nb)
        log(f"batch {nb} mixed tr={tr:g}: {t * 1e3:.3f} ms")
        if bestb is None or t < bestb[1]:
            bestb = (tr, t)
        s.close()
    trb = bestb[0]
    smixb = make(trb, ls.MIXED)
    tb_mix = time_matvec(smixb, xb, yb, args.steps, args.warmup, nrhs=nb)
    Yb_mix = yb.to_host().reshape(nb, -1)
    # FP64 batched at its own best splitting (the balance of its real-space
    # and reciprocal kernels differs from the mixed one)
    best64 = None
    for tr in trs:
        s = make(tr, ls.FP64)
        t = time_matvec(s, xb, yb, 2, 1, nrhs=nb)
        log(f"batch {nb} fp64 tr={tr:g}: {t * 1e3:.3f} ms")
        if best64 is None or t < best64[1]:
            best64 = (tr, t)
        s.close()
    tr64b = best64[0]
    s64b = make(tr64b, ls.FP64)
    tb_64 = time_matvec(s64b, xb, yb, max(3, args.steps // 4), 1, nrhs=nb)
    Yb_64 = yb.to_host().reshape(nb, -1)
    s64b.close()
    errs = [float(np.linalg.norm(Yb_mix[c] - Yb_64[c]) / np.linalg.norm(Yb_64[c]))
            for c in range(nb)]
    # column 0 against the FP64 single-rhs path at the reference splitting
    x0d = ls.DeviceArray.from_host(X[:11 * n])
    s64_ref.atimes_dev(x0d, yd)
    y0 = yd.to_host()
    err_b_tr1 = float(np.linalg.norm(Yb_mix[0] - y0) / np.linalg.norm(y0))
    log(f"batch {nb}: mixed {tb_mix * 1e3:.3f} ms (tr={trb:g}), fp64 "
        f"{tb_64 * 1e3:.3f} ms (tr={tr64b:g}), max col err {max(errs):.2e}")
    fp64_best = max(1.0 / t64, 1.0 / t64_ref, nb / tb_64)

    # roofline of the dominant kernel of the batched MIXED call: the
    # real-space kernel at the splitting it ran with
    real_ms, rest_ms = kernel_times(smixb, xb, yb, 5, nrhs=nb)
    smixb.close()
    shadow = ls.Stokes(2, pos, box=(L, L, L), tr=trb)
    pin, npairs = pair_images_in_cutoff(pos, shadow)
    shadow.close()
    flops = pin * FLOP_PER_IMAGE_FTS + npairs * 2.0 * 121.0 * nb
    achieved = flops / (real_ms * 1e-3) / 1e12
    kname = "k_real_moments<2>+k_real_mfma_mom<2, 4>"
    roof = dict(bound="mfma", achieved=achieved, peak=PEAK_FP32,
                unit="TFLOP/s", frac=achieved / PEAK_FP32,
                traffic=measured_traffic(kname, n, trb, nb),
                kernel=f"{kname} (real-space part of the call: FP32 image "
                       f"moments of each pair i <= j once, then the pair "
                       f"blocks x {nb} columns on the FP32 matrix cores)",
                kernel_ms=real_ms, rest_of_call_ms=rest_ms,
                pair_images=pin, pairs=npairs,
                flop_per_pair_image=FLOP_PER_IMAGE_FTS,
                flop_per_pair_per_column=242.0)

    # ---- the same batched MIXED product with the real-space blocks written
    # as a dense FP32 matrix once per call (stokes_hip_set_dense_blocks;
    # rebuilt from the positions on every call) times the block on the
    # matrix cores: not the matrix-free headline, reported beside it
    nd_cols = 2 * nb
    X2 = np.concatenate([X, np.random.default_rng(5).standard_normal(
        nb * 11 * n)])
    xb2 = ls.DeviceArray.from_host(X2)
    yb2 = ls.DeviceArray(nd_cols * 11 * n)
    bestd = None
    for tr in ([args.tr] if args.tr else [30.0, 100.0, 300.0]):
        s = make(tr, ls.MIXED)
        s.set_dense_blocks(True)
        t = time_matvec(s, xb2, yb2, 3, 1, nrhs=nd_cols)
        log(f"dense-block {nd_cols} mixed tr={tr:g}: {t * 1e3:.3f} ms")
        if bestd is None or t < bestd[1]:
            bestd = (tr, t)
        s.close()
    sd = make(bestd[0], ls.MIXED)
    sd.set_dense_blocks(True)
    td = time_matvec(sd, xb2, yb2, args.steps, args.warmup, nrhs=nd_cols)
    Yd = yb2.to_host().reshape(nd_cols, -1)
    errs_d = [float(np.linalg.norm(Yd[c] - Yb_64[c]) / np.linalg.norm(Yb_64[c]))
              for c in range(nb)]
    # the GEMM of one 128-column pass (HIP events around the real-space
    # part: the block's FP32 rows and the split-K GEMM)
    gemm_ms, _ = kernel_times(sd, xb2, yb2, 5, nrhs=nd_cols)
    sd.close()
    ndim = 11 * n
    gflop = 2.0 * ndim * ndim * min(128, nd_cols)
    dense = {"nrhs": nd_cols, "tr": bestd[0], "ms_per_call": td * 1e3,
             "matvecs_per_s": nd_cols / td,
             "rel_err_vs_fp64_max_column": max(errs_d),
             "roofline": dict(bound="mfma",
                              achieved=gflop / (gemm_ms * 1e-3) / 1e12,
                              peak=PEAK_FP32, unit="TFLOP/s",
                              frac=gflop / (gemm_ms * 1e-3) / 1e12 / PEAK_FP32,
                              traffic=None,
                              kernel="k_cache_gemm<8> (one 128-column pass: "
                                     "2 (11N)^2 128 flop)",
                              kernel_ms=gemm_ms)}
    log(f"dense-block {nd_cols}: {td * 1e3:.3f} ms (tr={bestd[0]:g}), "
        f"{nd_cols / td:.0f} matvecs/s, max col err {max(errs_d):.2e}")

    out = {
        "metric": METRIC,
        "value": nb / tb_mix,
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tb_mix * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"FTS Ewald matvec N={n} phi={args.phi} a=1 "
                               f"eps=1e-12, {nb} right-hand sides per call",
                   "n": n, "phi": args.phi, "tr": trb, "nrhs": nb},
        "rel_err_mixed_vs_fp64_max_column": max(errs),
        "rel_err_mixed_vs_fp64_tr1_column0": err_b_tr1,
        "fp64_matvecs_per_s_batched": nb / tb_64,
        "fp64_batched_tr": tr64b,
        "fp64_matvecs_per_s_single": 1.0 / t64,
        "fp64_matvecs_per_s_single_tr1": 1.0 / t64_ref,
        "mixed_over_fp64": (nb / tb_mix) / fp64_best,
        "single": {"tr": tr1, "mixed_matvecs_per_s": 1.0 / tmix,
                   "fp64_matvecs_per_s": 1.0 / t64,
                   "mixed_over_fp64": t64 / tmix,
                   "rel_err_mixed_vs_fp64": err1,
                   "rel_err_fp64_tuned_vs_tr1": err_ref_split,
                   "rel_err_mixed_vs_fp64_tr1": err_mix_tr1},
        "roofline": roof,
        "dense_real_blocks": dense,
    }
    if not args.no_dense:
        out["dense_build"] = bench_dense_build(args, pos, L, posd, log)
    if args.bd_steps > 0:
        out["bd"] = bench_bd(args, log)
    if args.ens_boxes > 0:
        out["ensemble"] = bench_ensemble(args, boxes, log)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pos, L, x, 1.0, args.cpu_budget)
        ratio = oracle_over_reference()
        if ratio is not None:
            out["cpu_baseline"]["oracle_vs_reference"] = ratio
        thr = max(1, min(16, os.cpu_count() or 1))
        out["cpu_baseline_threads"] = cpu_baseline_threads(
            pos, L, x, 1.0, args.cpu_budget, thr)
    print(json.dumps(out))
    s64_ref.close()


if __name__ == "__main__":
    main()
