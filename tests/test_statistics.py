"""Stochastic observables of the deterministic checkerboard dynamics vs a
serial random-sequential CPM (the dynamics the reference's parallel sweep
approximates), over seeds.

Tolerance (stats_common.compare): for every observable at every checkpoint,
|mean_checkerboard - mean_serial| <= max(4 * sqrt(se_cb^2 + se_serial^2),
rel * |mean_serial|) with rel = 1% (energy/site, mean area), 15% (area std),
10% (sorting index); 16 seeds, standard errors from the seed spread.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from stats_common import compare, observables, sorting_lattice

SEEDS = 16
CHECKPOINTS = 4
MCS_PER_CHECKPOINT = 50


def _trajectories(mode):
    out = []
    for seed in range(SEEDS):
        lat, n, p, _ = sorting_lattice(L=64, seed=seed)
        tab = O.tally(lat, 2, n)
        traj = []
        for k in range(CHECKPOINTS):
            m0 = k * MCS_PER_CHECKPOINT
            if mode == "serial":
                O.serial_sweep(lat, 2, tab, p, 1000 + seed, m0, MCS_PER_CHECKPOINT)
            else:
                c, u = mode
                O.sweep(lat, 2, tab, p, c, u, 1000 + seed, m0, MCS_PER_CHECKPOINT, 4)
            traj.append(observables(lat, n, p, tab.area))
        out.append(traj)
    return np.array(out)


@pytest.fixture(scope="module")
def serial():
    return _trajectories("serial")


@pytest.mark.parametrize("c,u", [(1, 1), (2, 4)])
def test_checkerboard_matches_serial_cpm(serial, c, u):
    cb = _trajectories((c, u))
    ok, report = compare(cb, serial)
    bad = [r for r in report if r[5] > 1.0]
    assert ok, bad
    # the dynamics did something: cells grew from 9 sites towards 40
    assert cb[:, -1, 1].mean() > 30


def test_tolerance_detects_a_different_temperature(serial):
    """The same comparison separates T = 10 from T = 30 (the test has power)."""
    out = []
    for seed in range(SEEDS):
        lat, n, p, _ = sorting_lattice(L=64, seed=seed)
        p.temperature = 30.0
        tab = O.tally(lat, 2, n)
        traj = []
        for k in range(CHECKPOINTS):
            O.sweep(lat, 2, tab, p, 1, 1, 1000 + seed, k * MCS_PER_CHECKPOINT, MCS_PER_CHECKPOINT, 4)
            traj.append(observables(lat, n, p, tab.area))
        out.append(traj)
    ok, _ = compare(np.array(out), serial)
    assert not ok
