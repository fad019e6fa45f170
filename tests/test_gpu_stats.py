"""GPU dynamics vs the serial random-sequential CPM at the reference's
sorting.py size (256^2): energy, area histogram moments and sorting index
over 8 seeds agree within the tolerance of stats_common.compare."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import run_kwargs
from stats_common import compare, observables, sorting_lattice

pytestmark = pytest.mark.gpu

SEEDS = 8
CHECKPOINTS = 3
STEP = 50


def test_gpu_sorting_statistics_match_serial_cpm():
    import gpucpm
    print("CASE gpu-stats 256", flush=True)
    gpu, ser = [], []
    for seed in range(SEEDS):
        lat0, n, p, J = sorting_lattice(L=256, seed=seed)
        sim = gpucpm.Cpm(256, 2, n, 10, False)
        sim.set_constraints(cell_type=1, lambda_area=1, target_area=40)
        sim.set_constraints(cell_type=2, lambda_area=1, target_area=40)
        for a, b in ((0, 1), (0, 2), (1, 1), (1, 2), (2, 2)):
            sim.set_constraints(cell_type=a, other_cell_type=b, adhesion=J[a][b])
        sim.get_state()[...] = lat0.view(np.int32)
        sim.push_to_gpu()
        sim.set_seed(77 + seed)
        traj = []
        for _ in range(CHECKPOINTS):
            sim.run(**run_kwargs(1, 1, STEP))
            sim.synchronize()
            sim.pull_from_gpu()
            lat = sim.get_state().view(np.uint32).copy()
            traj.append(observables(lat, n, p, sim.get_cell_table()["area"]))
        gpu.append(traj)
        lat = lat0.copy()
        tab = O.tally(lat, 2, n)
        traj = []
        for k in range(CHECKPOINTS):
            O.serial_sweep(lat, 2, tab, p, 500 + seed, k * STEP, STEP)
            traj.append(observables(lat, n, p, tab.area))
        ser.append(traj)
    ok, report = compare(np.array(gpu), np.array(ser))
    for r in report:
        print("t=%d %-14s gpu %.5g serial %.5g z=%.2f used=%.2f" % ((r[0] + 1) * STEP, *r[1:]))
    assert ok, [r for r in report if r[5] > 1.0]
