"""The benchmark configurations at their stated scale (the GPU path through
the C ABI against the CPU oracle on all host cores):

- 2D sorting, 1024^2, 3969 cells: the full 10,000 MCS, lattice and tables
  bit-exact;
- 3D, 256^3, 20,000 cells: the full 1,000 MCS, lattice and tables bit-exact;
- 64 batched sorting replicas, L in {128, 256, 512, 1024}: 2,000 MCS in one
  batch, every replica bit-identical to its own single run, its tables equal
  to a recount, and all 64 bit-exact against the oracle over a 20-MCS prefix
  (the run loop each replica reproduces is cpm.cu:405-481);
- 32 replicas of 512^2 sorting, 2,000 MCS: the statistics kernel equals a
  numpy recount for all 32, and dH/dtheta is within 1e-5 of FP64 central
  differences of the oracle's Hamiltonian.
"""
from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from helpers import assert_tables_equal, run_kwargs, run_oracle

pytestmark = pytest.mark.gpu

NCPU = len(os.sched_getaffinity(0))


def _announce(*a):
    print("CASE", *a, flush=True)


def _single(wl, seed, nmcs, c=1, u=1):
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(seed)
    sim.run(**run_kwargs(c, u, nmcs))
    sim.synchronize()
    sim.pull_from_gpu()
    return sim, sim.get_state().view(np.uint32).copy()


def test_sorting2d_full_run_bit_exact():
    from gpucpm import workloads as W
    wl = W.sorting2d()
    _announce("sorting2d full", wl.L, wl.ncells, 10000)
    sim, lat = _single(wl, 20240601, 10000)
    lat_o, tab_o = run_oracle(wl, 1, 1, 20240601, 10000, nthreads=NCPU)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)
    # the tightly packed squares under a strong perimeter pressure barely
    # move (a few hundred sites differ at the end); the run is not idle
    assert np.count_nonzero(lat != wl.lattice) > 100


def test_plain3d_full_run_bit_exact():
    from gpucpm import workloads as W
    wl = W.plain3d()
    _announce("plain3d full", wl.L, wl.ncells, 1000)
    sim, lat = _single(wl, 33, 1000)
    lat_o, tab_o = run_oracle(wl, 1, 1, 33, 1000, nthreads=NCPU)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 3)


def test_batched_replicas_full_config():
    import gpucpm
    import oracle as O
    from gpucpm import workloads as W
    wls = W.batched_sorting(64, seed=4)
    _announce("batched", len(wls), sorted({w.L for w in wls}), 2000)

    def fresh():
        sims = []
        for i, wl in enumerate(wls):
            s = wl.make_cpm()
            s.push_to_gpu()
            s.set_seed(9000 + i)
            sims.append(s)
        return sims

    sims = fresh()
    batch = gpucpm.Batch(sims)
    batch.run(**run_kwargs(1, 1, 20))
    batch.synchronize()
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.pull_from_gpu()
        lat_o, tab_o = run_oracle(wl, 1, 1, 9000 + i, 20, nthreads=NCPU)
        np.testing.assert_array_equal(s.get_state().view(np.uint32), lat_o, err_msg=f"replica {i}")
        assert_tables_equal(s.get_cell_table(), tab_o, 2)
    batch.run(**run_kwargs(1, 1, 1980))
    batch.synchronize()
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.pull_from_gpu()
        lat = s.get_state().view(np.uint32).copy()
        tab = O.tally(lat, 2, wl.ncells)
        table = s.get_cell_table()
        np.testing.assert_array_equal(table["area"], tab.area)
        np.testing.assert_array_equal(table["perimeter"], tab.perim)
        _, lat1 = _single(wl, 9000 + i, 2000)
        np.testing.assert_array_equal(lat, lat1, err_msg=f"replica {i} batch vs single run")
        assert s.mcs == 2000


def test_gradient_full_config():
    import gpucpm
    import oracle as O
    from gpucpm import workloads as W
    from gpucpm.autograd import ThetaLayout, energy
    from test_gradient_cpu import numpy_stats
    layout = ThetaLayout(3)
    wls = [W.sorting2d(L=512, grid=40, edge=10, seed=700 + i) for i in range(32)]
    _announce("gradient", len(wls), 512, 2000)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(7000 + i)
        sims.append(s)
    gpucpm.Batch(sims).run(**run_kwargs(1, 1, 2000))
    p = wls[0].params()
    theta = layout.pack(p["J"], p["lam_area"], p["target_area"], p["lam_perim"], p["target_perim"],
                        dtype=torch.float32).requires_grad_(True)
    H = energy(theta, sims, layout)
    H.backward()
    g = theta.grad.double().numpy()
    pairs, mom = gpucpm.energy_stats_many(sims)
    stats = []
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.pull_from_gpu()
        lat = s.get_state().view(np.uint32).copy()
        p_np, m_np = numpy_stats(lat, 2)
        np.testing.assert_array_equal(np.triu(pairs[i]), np.triu(p_np), err_msg=f"replica {i}")
        np.testing.assert_array_equal(mom[i], m_np, err_msg=f"replica {i}")
        st = O.energy_stats(lat, 2, wl.ncells, O.Params.make(1.0))
        spin, typ = lat & 0xFFFFFF, lat >> 24
        ncell = [len(np.unique(spin[(typ == t) & (spin > 0)])) for t in range(3)]
        stats.append((st, ncell))

    def H64(th):
        nt, nj = layout.nt, layout.nj
        J = np.zeros((20, 20))
        for k, (a, b) in enumerate(layout.pairs):
            J[a, b] = th[k]
        tot = 0.0
        for st, ncell in stats:
            tot += float(np.sum(J * st["pairs"]))
            for t in range(nt):
                la, a0 = th[nj + t], th[nj + nt + t]
                lp, p0 = th[nj + 2 * nt + t], th[nj + 3 * nt + t]
                tot += la * (st["sa2"][t] - 2 * a0 * st["sa1"][t] + ncell[t] * a0 * a0)
                tot += lp * (st["sp2"][t] - 2 * p0 * st["sp1"][t] + ncell[t] * p0 * p0)
        return tot / len(stats)

    th = theta.detach().double().numpy()
    assert abs(float(H.detach()) - H64(th)) / abs(H64(th)) < 1e-6
    h = 1e-3
    fd = np.array([(H64(th + h * e) - H64(th - h * e)) / (2 * h) for e in np.eye(th.size)])
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1.0)
    assert rel.max() < 1e-5, (rel.max(), layout.names()[int(rel.argmax())])
