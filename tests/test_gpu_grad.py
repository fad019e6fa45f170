"""dH/dtheta on the GPU: the statistics kernel vs an independent numpy
reduction (exact), and the gradient of sampled configurations vs the FP64
oracle Hamiltonian by central differences (relative error < 1e-5)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import run_kwargs
from test_gradient_cpu import numpy_stats

pytestmark = pytest.mark.gpu


def test_stats_kernel_matches_numpy_2d_and_3d():
    from gpucpm import workloads as W
    for wl, c, u, n in ((W.sorting2d(L=128, grid=9, edge=10, seed=50), 1, 1, 40),
                        (W.plain3d(L=32, ncells=60, seed=51), 1, 1, 10)):
        print("CASE stats", wl.name, wl.L, flush=True)
        sim = wl.make_cpm()
        sim.push_to_gpu()
        sim.run(**run_kwargs(c, u, n), seed=9)
        pairs, mom = sim.energy_stats()
        sim.pull_from_gpu()
        lat = sim.get_state().view(np.uint32).copy()
        p_np, m_np = numpy_stats(lat, wl.d)
        np.testing.assert_array_equal(np.triu(pairs), np.triu(p_np))
        np.testing.assert_array_equal(mom, m_np)


def test_sampled_gradient_vs_fp64_oracle_fd():
    """Sorting replicas (C2 parameters, 512^2), forward run, then dH/dtheta
    (FP32) for J, lam_A, A0, lam_P, P0 against central differences (step
    1e-3) of the oracle's FP64 Hamiltonian on the same lattices."""
    import gpucpm
    import oracle as O
    from gpucpm import workloads as W
    from gpucpm.autograd import ThetaLayout, energy
    layout = ThetaLayout(3)
    wls = [W.sorting2d(L=512, grid=40, edge=10, seed=60 + i) for i in range(4)]
    print("CASE grad", len(wls), flush=True)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(500 + i)
        sims.append(s)
    gpucpm.Batch(sims).run(**run_kwargs(1, 1, 100))
    p = wls[0].params()
    theta = layout.pack(p["J"], p["lam_area"], p["target_area"], p["lam_perim"], p["target_perim"],
                        dtype=torch.float32).requires_grad_(True)
    H = energy(theta, sims, layout)
    H.backward()
    g = theta.grad.double().numpy()

    lats = []
    for s in sims:
        s.synchronize()
        s.pull_from_gpu()
        lats.append(s.get_state().view(np.uint32).copy())

    def H64(th):
        nt, nj = layout.nt, layout.nj
        J = np.zeros((20, 20))
        for k, (a, b) in enumerate(layout.pairs):
            J[a, b] = J[b, a] = th[k]
        tot = 0.0
        for lat, wl in zip(lats, wls):
            st = O.energy_stats(lat, 2, wl.ncells, O.Params.make(1.0))
            # FP64 energy at real theta from the oracle's per-site pair counts
            # and its per-cell recount (moments rebuilt from sa1/sa2 at A0 = 0)
            tot += float(np.sum(np.triu(J) * st["pairs"]))
            area_terms = st["sa2"], st["sa1"]
            for t in range(nt):
                la, a0 = th[nj + t], th[nj + nt + t]
                lp, p0 = th[nj + 2 * nt + t], th[nj + 3 * nt + t]
                nA = _count_cells(lat, t)
                tot += la * (st["sa2"][t] - 2 * a0 * st["sa1"][t] + nA * a0 * a0)
                tot += lp * (st["sp2"][t] - 2 * p0 * st["sp1"][t] + nA * p0 * p0)
        return tot / len(lats)

    th = theta.detach().double().numpy()
    Hgpu = float(H.detach())
    assert abs(Hgpu - H64(th)) / abs(H64(th)) < 1e-6
    h = 1e-3
    fd = np.array([(H64(th + h * e) - H64(th - h * e)) / (2 * h) for e in np.eye(th.size)])
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1.0)
    assert rel.max() < 1e-5, (rel.max(), layout.names()[int(rel.argmax())])


def _count_cells(lat, t):
    spin = lat & 0xFFFFFF
    typ = lat >> 24
    return len(np.unique(spin[(typ == t) & (spin > 0)]))


def test_contrastive_gradient_is_difference_of_means():
    import gpucpm
    from gpucpm import workloads as W
    from gpucpm.autograd import ThetaLayout, contrastive_gradient
    layout = ThetaLayout(3)
    data = [W.sorting2d(L=256, grid=20, edge=10, seed=70 + i, J_AB=8) for i in range(2)]
    model = [W.sorting2d(L=256, grid=20, edge=10, seed=80 + i, J_AB=11) for i in range(2)]
    print("CASE contrastive", flush=True)
    sims_d, sims_m = [], []
    for wl, out in [(w, sims_d) for w in data] + [(w, sims_m) for w in model]:
        s = wl.make_cpm()
        s.push_to_gpu()
        out.append(s)
    gpucpm.Batch(sims_d + sims_m).run(**run_kwargs(1, 1, 50))
    p = model[0].params()
    theta = layout.pack(p["J"], p["lam_area"], p["target_area"], p["lam_perim"], p["target_perim"])
    g = contrastive_gradient(theta, sims_d, sims_m, layout).numpy()
    k = layout.pairs.index((1, 2))
    nd = np.mean([s.energy_stats()[0][1, 2] for s in sims_d])
    nm = np.mean([s.energy_stats()[0][1, 2] for s in sims_m])
    assert g[k] == pytest.approx(nd - nm, rel=1e-6)


def test_stats_many_equal_single_and_numpy():
    """energy_stats_many (every engine's statistics queued, one wait) equals
    energy_stats one by one and the numpy reduction, also read straight from
    the 16-bit lattice a run leaves current."""
    import gpucpm
    from gpucpm import workloads as W
    wls = [W.sorting2d(L=64, grid=5, edge=10, seed=60 + i) for i in range(3)] + \
          [W.walled(W.sorting2d(L=128, grid=9, edge=10, seed=63))]
    print("CASE stats many", [w.L for w in wls], flush=True)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.run(**run_kwargs(1, 1, 10), seed=30 + i)
        sims.append(s)
    pairs, mom = gpucpm.energy_stats_many(sims)  # before any pull: 16-bit sites current
    for i, s in enumerate(sims):
        p1, m1 = s.energy_stats()
        np.testing.assert_array_equal(pairs[i], p1)
        np.testing.assert_array_equal(mom[i], m1)
        s.pull_from_gpu()
        p_np, m_np = numpy_stats(s.get_state().view(np.uint32).copy(), 2)
        np.testing.assert_array_equal(np.triu(pairs[i]), np.triu(p_np))
        np.testing.assert_array_equal(mom[i], m_np)


def test_energy_grad_entry_equals_autograd_formula():
    """cpm_energy_grad (H and dH/dtheta in one C call) equals the autograd
    layer's closed forms on the same statistics, and its derivatives equal
    central differences of its own H (H is quadratic in theta: exact up to
    rounding)."""
    from gpucpm import workloads as W
    from gpucpm.autograd import ThetaLayout, energy_and_grad, stats_to_arrays
    wl = W.sorting2d(L=128, grid=9, edge=10, seed=64)
    print("CASE energy_grad", wl.L, flush=True)
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.run(**run_kwargs(1, 1, 20), seed=5)
    layout = ThetaLayout(3)
    p = wl.params()
    th = layout.pack(p["J"], p["lam_area"], p["target_area"], p["lam_perim"], p["target_perim"],
                     dtype=torch.float64).numpy()
    H, g = sim.energy_grad(th, 3)
    N, m = stats_to_arrays(*sim.energy_stats(), layout)
    H2, g2 = energy_and_grad(th, N, m, layout)
    np.testing.assert_allclose(H, H2, rtol=1e-14)
    np.testing.assert_allclose(g, g2, rtol=1e-14)
    h = 0.5
    for i in range(th.size):
        e = np.zeros_like(th)
        e[i] = h
        fd = (sim.energy_grad(th + e, 3)[0] - sim.energy_grad(th - e, 3)[0]) / (2 * h)
        np.testing.assert_allclose(g[i], fd, rtol=1e-9, atol=1e-6)
