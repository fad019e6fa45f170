"""Observables and the seed-spread comparison for statistical tests of CPM
dynamics.

The reference's colour phase is concurrent (kernels2d.cu:292-302 read
cellArea / cellPerimeter while other threads' atomicAdds at :552-569 land),
so its outcome depends on the interleaving.  Two executions bound what it
can produce: every table read before any write of the phase (the engine's
phase-start snapshot, oracle.sweep) and every read after the writes that
precede it in one sequential order (oracle.sweep_live).  The tests compare
the engine with the live-read execution and with the serial random-sequential
CPM (oracle.serial_sweep), with tolerances taken from the spread over seeds
alone: a difference passes when Welch's t statistic is within the two-sided
critical value at ALPHA per comparison.
"""
from __future__ import annotations

import numpy as np
from scipy import stats as sps

import oracle as O

ALPHA = 1e-4  # per-comparison false-alarm rate of compare()


def sorting_lattice(L=64, seed=0):
    """The reference's examples/sorting.py set-up scaled to L: 3x3 cells on
    a 6-site grid inside a disc of radius L/3, types 1/2 at random;
    lam_A = 1, A0 = 40, J_11 = 14, J_22 = 2, J_12 = 11, J_m = 16, T = 10."""
    rng = np.random.default_rng(seed)
    lat = np.zeros((L, L), np.uint32)
    n = 0
    for y in range(0, L - 1, 6):
        for x in range(0, L - 1, 6):
            if (x - L / 2) ** 2 + (y - L / 2) ** 2 < (L * L) / 9:
                n += 1
                t = rng.integers(1, 3)
                lat[y:y + 3, x:x + 3] = (np.uint32(t) << np.uint32(24)) | np.uint32(n)
    J = [[0, 16, 16], [16, 14, 11], [16, 11, 2]]
    p = O.Params.make(10, J=J, lam_area=[0, 1, 1], target_area=[0, 40, 40])
    return lat, n, p, J


NAMES = ["adhesion/site", "mean area", "area std", "sorting index"]
AREA_STD = 2


def observables(lat, n, p, area, J):
    """(adhesion energy per site, mean cell area, std of cell areas, sorting
    index), the sorting index being the heterotypic share of cell-cell
    contacts.  The adhesion energy is sum_{a<=b} J_ab N_ab over the unordered
    unlike-cell neighbour pairs N_ab (the area term of H is a function of the
    area moments, which the second and third entries carry)."""
    pr = O.energy_stats(lat, 2, n, p)["pairs"]  # upper triangle, a <= b
    adh = float(np.sum(np.triu(np.asarray(J, float)[:3, :3] * pr[:3, :3]))) / lat.size
    a = area[1:][area[1:] > 0]
    sort = pr[1, 2] / max(1, pr[1, 1] + pr[1, 2] + pr[2, 2])
    return np.array([adh, a.mean(), a.std(), sort])


def compare(a, b, names=None, skip=()):
    """a, b: (seeds, times, observables).  Welch's t for every time and
    observable (not in skip); ok when every |t| <= the two-sided critical
    value at ALPHA with Welch-Satterthwaite degrees of freedom.  Returns
    (ok, rows (time, name, mean_a, mean_b, t, critical))."""
    names = names or NAMES
    na, nb = a.shape[0], b.shape[0]
    ma, mb = a.mean(0), b.mean(0)
    va, vb = a.var(0, ddof=1) / na, b.var(0, ddof=1) / nb
    se = np.sqrt(va + vb)
    t = (ma - mb) / np.maximum(se, 1e-300)
    df = (va + vb) ** 2 / np.maximum(va ** 2 / (na - 1) + vb ** 2 / (nb - 1), 1e-300)
    crit = sps.t.ppf(1 - ALPHA / 2, df)
    rows, ok = [], True
    for i in range(a.shape[1]):
        for k in range(a.shape[2]):
            if k in skip:
                continue
            rows.append((i, names[k], ma[i, k], mb[i, k], t[i, k], crit[i, k]))
            ok = ok and abs(t[i, k]) <= crit[i, k]
    return ok, rows


def welch_t(a, b):
    """Welch's t of mean(a) - mean(b) and its critical value at ALPHA."""
    na, nb = len(a), len(b)
    va, vb = np.var(a, ddof=1) / na, np.var(b, ddof=1) / nb
    df = (va + vb) ** 2 / (va ** 2 / (na - 1) + vb ** 2 / (nb - 1))
    return (np.mean(a) - np.mean(b)) / np.sqrt(va + vb), sps.t.ppf(1 - ALPHA / 2, df)
