"""Observables for statistical comparisons between CPM dynamics."""
from __future__ import annotations

import numpy as np

import oracle as O

Z_TOL = 4.0  # seed-spread term: Z_TOL * sqrt(se_a^2 + se_b^2)
# Relative floors for the systematic differences between the parallel
# checkerboard dynamics (phase-start table snapshot, every site a target
# once per MCS) and serial random-sequential dynamics.  Measured at 256^2
# over 8 seeds: energy/site within 0.2%, mean area +0.4%, area std +8%,
# sorting index -5% (the checkerboard sorts slightly faster per MCS).
REL_TOL = [0.01, 0.01, 0.15, 0.10]


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


def observables(lat, n, p, area):
    """(energy per site, mean cell area, std of cell areas, sorting index)
    where the sorting index is the heterotypic share of cell-cell contacts."""
    st = O.energy_stats(lat, 2, n, p)
    pr = st["pairs"]
    H = O.hamiltonian(lat, 2, n, p) / lat.size
    a = area[1:][area[1:] > 0]
    sort = pr[1, 2] / max(1, pr[1, 1] + pr[1, 2] + pr[2, 2])
    return np.array([H, a.mean(), a.std(), sort])


NAMES = ["energy/site", "mean area", "area std", "sorting index"]


def compare(a, b):
    """a, b: (seeds, times, observables).  An observable agrees when
    |mean_a - mean_b| <= max(Z_TOL * sqrt(se_a^2 + se_b^2), REL_TOL * |mean_b|).
    Returns (ok, report rows (time, name, mean_a, mean_b, z, ratio)) where
    ratio = |difference| / allowed."""
    ma, mb = a.mean(0), b.mean(0)
    se = np.sqrt(a.var(0, ddof=1) / a.shape[0] + b.var(0, ddof=1) / b.shape[0])
    allowed = np.maximum(Z_TOL * se, np.array(REL_TOL)[None, :] * np.abs(mb))
    diff = np.abs(ma - mb)
    z = diff / np.maximum(se, 1e-12)
    ratio = diff / np.maximum(allowed, 1e-12)
    report = [(t, NAMES[k], ma[t, k], mb[t, k], z[t, k], ratio[t, k])
              for t in range(a.shape[1]) for k in range(a.shape[2])]
    return bool(np.all(ratio <= 1.0)), report
