"""Synthetic particle configurations of the benchmark workloads.

Monodisperse spheres (a = 1) in a cubic box of volume fraction phi:
L = (N * 4 pi / 3 / phi)^(1/3).  Positions are non-overlapping with a
minimum centre distance `gap` (2.1 a: the reference's lubrication cut-off
scale) under the minimum-image convention.  Dilute systems use random
sequential addition; dense ones start from a random subset of a simple
cubic lattice and are randomised by hard-sphere Monte Carlo sweeps.
Everything is seeded (numpy PCG64) and deterministic.
"""
import math

import numpy as np


def box_length(n, phi, a=1.0):
    return (n * 4.0 * math.pi * a ** 3 / 3.0 / phi) ** (1.0 / 3.0)


def _min_image(d, L):
    return d - L * np.round(d / L)


def rsa(n, L, gap, rng, max_tries=200000):
    pos = np.empty((n, 3))
    k = 0
    tries = 0
    g2 = gap * gap
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("random sequential addition did not converge")
        p = rng.uniform(0.0, L, 3)
        if k:
            d = _min_image(pos[:k] - p, L)
            if np.min(np.einsum("ij,ij->i", d, d)) < g2:
                continue
        pos[k] = p
        k += 1
    return pos


def lattice_mc(n, L, gap, rng, sweeps=20, step=0.4):
    m = int(math.ceil(n ** (1.0 / 3.0)))
    while (L / m) < gap:
        raise ValueError("box too small for a lattice start at this gap")
    sites = (np.indices((m, m, m)).reshape(3, -1).T + 0.5) * (L / m)
    pick = rng.choice(len(sites), size=n, replace=False)
    pos = sites[pick].copy()
    g2 = gap * gap
    for _ in range(sweeps):
        for i in rng.permutation(n):
            trial = pos[i] + rng.uniform(-step, step, 3)
            d = _min_image(pos - trial, L)
            r2 = np.einsum("ij,ij->i", d, d)
            r2[i] = np.inf
            if r2.min() >= g2:
                pos[i] = trial % L
    return pos


def configuration(n, phi, seed=0, gap=2.1):
    """(positions [n,3], box length) for n unit spheres at volume fraction
    phi, deterministic in seed"""
    rng = np.random.default_rng(seed)
    L = box_length(n, phi)
    phi_eff = phi * (gap / 2.0) ** 3
    if phi_eff < 0.25:
        pos = rsa(n, L, gap, rng)
    else:
        pos = lattice_mc(n, L, gap, rng)
    return pos, L


def configuration_poly(n, phi, seed=0, a_lo=0.5, a_hi=2.0, gap=1.05,
                       max_tries=2000000):
    """(positions [n,3], radii [n], box length) for n spheres with radii
    log-uniform on [a_lo, a_hi] at volume fraction phi; centres at least
    gap (a_i + a_j) apart (minimum image), placed largest first by random
    sequential addition; deterministic in seed"""
    rng = np.random.default_rng(seed)
    a = np.exp(rng.uniform(math.log(a_lo), math.log(a_hi), n))
    L = (np.sum(4.0 * math.pi / 3.0 * a ** 3) / phi) ** (1.0 / 3.0)
    order = np.argsort(-a)
    pos = np.empty((n, 3))
    # the placed spheres in placement order
    pp = np.empty((n, 3))
    pa = np.empty(n)
    k = 0
    tries = 0
    for idx in order:
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("random sequential addition did not converge")
            p = rng.uniform(0.0, L, 3)
            if k:
                d = _min_image(pp[:k] - p, L)
                lim = gap * (pa[:k] + a[idx])
                if np.any(np.einsum("ij,ij->i", d, d) < lim * lim):
                    continue
            pos[idx] = p
            pp[k] = p
            pa[k] = a[idx]
            k += 1
            break
    return pos, a, L


def min_distance(pos, L):
    best = np.inf
    for i in range(len(pos) - 1):
        d = _min_image(pos[i + 1:] - pos[i], L)
        best = min(best, float(np.sqrt(np.min(np.einsum("ij,ij->i", d, d)))))
    return best
