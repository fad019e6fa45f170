"""The online autotuner (Cpm.set_autotune / Batch.set_autotune): it times
candidate kernel forms on a run's own launches, so the runs it samples must
stay bit-exact against the oracle, and once settled every later run uses one
form.  Every case prints its name before touching the GPU."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal, run_kwargs, run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("d", [2, 3])
def test_autotuned_runs_are_exact_and_settle(d):
    from gpucpm import workloads as W
    wl = W.sorting2d(L=128, grid=9, edge=10, seed=71) if d == 2 else W.plain3d(L=32, ncells=60, seed=72)
    print("CASE autotune single", d, wl.L, flush=True)
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(17)
    sim.set_autotune(True)
    kernels = []
    per_run = 8 if d == 2 else 2
    # 8 samples each of the launch forms (per phase, single workgroup; the
    # per-phase runs also sample their kernel forms), then settled runs
    nruns = 24
    for _ in range(nruns):
        sim.run(**run_kwargs(1, 1, per_run))
        kernels.append(sim.last_launch()["kernel"])
    sim.synchronize()
    sim.pull_from_gpu()
    lat_o, tab_o = run_oracle(wl, 1, 1, 17, per_run * nruns)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, d)
    # settled: the last runs all use one kernel
    assert len(set(kernels[-3:])) == 1, kernels
    assert "k_sweep_phase_compact" in kernels[-1] or "k_small_lds" in kernels[-1]
    assert any("k_small_lds" in k_ for k_ in kernels) and any("k_sweep_phase" in k_ for k_ in kernels)


def test_autotuned_interactive_runs_are_exact():
    """One-MCS runs (the notebook pattern) accumulate samples across calls."""
    from gpucpm import workloads as W
    wl = W.sorting2d(L=64, grid=5, edge=10, seed=73)
    print("CASE autotune interactive", wl.L, flush=True)
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(3)
    sim.set_autotune(True)
    for _ in range(40):
        sim.run(**run_kwargs(2, 1, 1, shared=0, positions_per_thread=4))
        sim.pull_from_gpu()
    lat_o, tab_o = run_oracle(wl, 2, 1, 3, 40, shared=False, ppt=4)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)


def test_autotuned_batch_is_exact():
    import gpucpm
    from gpucpm import workloads as W
    wls = [W.walled(w) if i % 2 else w
           for i, w in enumerate(W.batched_sorting(nrep=5, seed=74, sizes=(32, 64, 128)))]
    print("CASE autotune batch", [w.L for w in wls], flush=True)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(50 + i)
        sims.append(s)
    batch = gpucpm.Batch(sims)
    batch.set_autotune(True)
    for _ in range(4):
        batch.run(**run_kwargs(1, 1, 16))
    batch.synchronize()
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.pull_from_gpu()
        lat_o, tab_o = run_oracle(wl, 1, 1, 50 + i, 64)
        np.testing.assert_array_equal(s.get_state().view(np.uint32), lat_o)
        assert_tables_equal(s.get_cell_table(), tab_o, 2)
