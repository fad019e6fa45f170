"""Seeded combinations of the run's options, GPU against the oracle
bit-exactly: dimension, size, schedule (c, u, shared roles, positions per
thread), cell_sync, the Hamiltonian terms (fixed wall, chemotaxis, act,
persistence), the random stream and the site width.  The list is drawn
once from a fixed seed so every run tests the same cases; each case prints
its options before touching the GPU."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal, run_kwargs, run_oracle, make_cpm

pytestmark = pytest.mark.gpu


def _draw_cases(n=32, seed=2026):
    rng = np.random.default_rng(seed)
    cases = []
    while len(cases) < n:
        d = int(rng.choice([2, 3], p=[0.6, 0.4]))
        c = int(rng.choice([1, 2]))
        shared = int(rng.integers(0, 2))
        ppt = 2 * c if shared else int(rng.choice([c * k for k in (1, 2, 4) if c * k >= 2]))
        u = int(rng.integers(1, c ** d + 1)) if c > 1 else 1
        cell_sync = int(rng.random() < 0.25)
        terms = [t for t, pr in (("wall", 0.3), ("chem", 0.3), ("act", 0.3), ("pers", 0.5))
                 if rng.random() < pr]
        if d == 2:
            terms = [t for t in terms if t != "pers"]
        rng_mode = "minstd" if rng.random() < 0.3 else "philox"
        narrow = str(int(rng.random() < 0.7))
        # refused combinations stay out (their refusal is tested elsewhere)
        if "act" in terms and shared and c == 1 and cell_sync:
            continue
        cases.append(dict(d=d, c=c, u=u, shared=shared, ppt=ppt, cell_sync=cell_sync, terms=terms,
                          rng=rng_mode, narrow=narrow, seed=int(rng.integers(1, 1 << 30))))
    return cases


CASES = _draw_cases()


def _workload(case):
    from gpucpm import workloads as W
    d, terms = case["d"], case["terms"]
    if d == 2:
        wl = W.sorting2d(L=64, grid=5, edge=10, seed=case["seed"] % 1000)
    else:
        wl = W.plain3d(L=32, ncells=50, seed=case["seed"] % 1000)
    if "wall" in terms:
        wl = W.walled(wl)
    if "chem" in terms:
        wl = W.chemotactic(wl, lam=(0.0, 20.0, 5.0, 0.0) if "wall" in terms else (0.0, 20.0, 5.0))
    if "act" in terms:
        wl = W.active(wl) if d == 2 else W.active(wl, lam=(0.0, 200.0), act_max=(0, 30))
    if "pers" in terms:
        wl = W.persistent(wl) if d == 2 else W.persistent(wl, lam=(0.0, 200.0), diffusion=(0.0, 0.8))
    return wl


@pytest.mark.parametrize("idx", range(len(CASES)))
def test_combination(idx, engine_options):
    case = CASES[idx]
    engine_options(narrow=case["narrow"])
    wl = _workload(case)
    d, c, u = case["d"], case["c"], case["u"]
    n = 8 if d == 2 else 3
    inner = n // 2 if (case["rng"] == "minstd" or "pers" in case["terms"]) else n
    print("CASE combination", idx, case, flush=True)
    sim = make_cpm(wl)
    sim.push_to_gpu()
    sim.set_seed(case["seed"])
    sim.set_rng(case["rng"])
    sim.run(**run_kwargs(c, u, n, inner, shared=case["shared"], positions_per_thread=case["ppt"],
                         cell_sync=case["cell_sync"]))
    sim.synchronize()
    sim.pull_from_gpu()
    lat_o, tab_o = run_oracle(wl, c, u, case["seed"], n, inner=inner, shared=bool(case["shared"]),
                              ppt=case["ppt"], rng=case["rng"], cell_sync=bool(case["cell_sync"]))
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, d)
    if getattr(wl, "act", None) is not None:
        np.testing.assert_array_equal(sim.get_act_state().view(np.uint32), tab_o.act)
    if hasattr(tab_o, "pref"):
        np.testing.assert_array_equal(sim.get_pref_dirs().reshape(-1), np.asarray(tab_o.pref).reshape(-1))
