"""dH/dtheta closed forms (gpucpm.autograd) pinned on the CPU.

The statistics here come from an independent numpy reduction of the
lattice; H is checked against the oracle's FP64 Hamiltonian and dH/dtheta
against central finite differences of that Hamiltonian at step 1e-3.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from gpucpm.autograd import CpmEnergy, ThetaLayout, energy_and_grad, stats_to_arrays


def numpy_stats(lat, d):
    """pairs (20x20, a<=b) and moments (20x5) by numpy, independently of the GPU kernel."""
    spin = (lat & 0xFFFFFF).astype(np.int64)
    typ = (lat >> 24).astype(np.int64)
    pairs = np.zeros((20, 20), np.int64)
    unlike = np.zeros(lat.shape, np.int64)
    for o in np.ndindex(*(3,) * d):
        if all(x == 1 for x in o):
            continue
        sh = [x - 1 for x in o]
        nb_full = np.roll(lat, shift=[-s for s in sh], axis=tuple(range(d)))
        nb_spin = nb_full & 0xFFFFFF
        diff = nb_full != lat
        unlike += nb_spin != spin
        a = np.minimum(typ, nb_full >> 24)
        b = np.maximum(typ, nb_full >> 24)
        np.add.at(pairs, (a[diff], b[diff]), 1)
    pairs //= 2
    n = spin.max()
    area = np.bincount(spin.ravel(), minlength=n + 1)
    per = np.bincount(spin.ravel(), weights=unlike.ravel(), minlength=n + 1).astype(np.int64)
    ctype = np.zeros(n + 1, np.int64)
    ctype[spin.ravel()] = typ.ravel()
    mom = np.zeros((20, 5), np.int64)
    for c in range(1, n + 1):
        if area[c]:
            t = ctype[c]
            mom[t] += [1, area[c], area[c] ** 2, per[c], per[c] ** 2]
    return pairs, mom


def _lattice(seed, L=32, d=2):
    rng = np.random.default_rng(seed)
    lat = np.zeros((L,) * d, np.uint32)
    for k in range(1, 12):
        c = rng.integers(0, L - 6, size=d)
        sl = tuple(slice(int(x), int(x) + int(rng.integers(2, 6))) for x in c)
        lat[sl] = (np.uint32(rng.integers(1, 3)) << np.uint32(24)) | np.uint32(k)
    return lat


THETA = dict(J=[[0, 16, 16], [16, 2, 11], [16, 11, 14]], lam_area=[0, 2.5, 1.25],
             target_area=[0, 20, 12], lam_perim=[0, 0.5, 0.75], target_perim=[0, 30, 25])


@pytest.mark.parametrize("d,L", [(2, 32), (3, 16)])
def test_closed_form_energy_matches_oracle_hamiltonian(d, L):
    lat = _lattice(d, L, d)
    layout = ThetaLayout(3)
    ncells = int((lat & 0xFFFFFF).max())
    pairs, mom = numpy_stats(lat, d)
    N, m = stats_to_arrays(pairs, mom, layout)
    th = layout.pack(**THETA, dtype=torch.float64).numpy()
    H, _ = energy_and_grad(th, N, m, layout)
    p = O.Params.make(1.0, J=THETA["J"], lam_area=THETA["lam_area"], target_area=THETA["target_area"],
                      lam_perim=THETA["lam_perim"], target_perim=THETA["target_perim"])
    assert H == pytest.approx(O.hamiltonian(lat, d, ncells, p), rel=1e-12)
    st = O.energy_stats(lat, d, ncells, p)
    np.testing.assert_array_equal(np.triu(st["pairs"]), np.triu(pairs))


def test_gradient_matches_central_differences():
    """dH/dtheta vs central differences of the FP64 Hamiltonian (step 1e-3).
    H is linear in J and lambda and quadratic in the targets, so the central
    difference is exact up to rounding; the FP32 gradient must agree to 1e-5."""
    d = 2
    lat = _lattice(7)
    layout = ThetaLayout(3)
    pairs, mom = numpy_stats(lat, d)
    N, m = stats_to_arrays(pairs, mom, layout)
    th = layout.pack(**THETA, dtype=torch.float64).numpy()
    _, g64 = energy_and_grad(th, N, m, layout)
    h = 1e-3
    fd = np.zeros_like(th)
    for i in range(th.size):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (energy_and_grad(tp, N, m, layout)[0] - energy_and_grad(tm, N, m, layout)[0]) / (2 * h)
    g32 = g64.astype(np.float32).astype(np.float64)
    scale = np.maximum(np.abs(fd), 1.0)
    assert np.max(np.abs(g32 - fd) / scale) < 1e-5
    # the J block equals the integer pair counts
    assert np.array_equal(g64[: layout.nj], N)


def test_autograd_function_backward():
    class FakeSim:
        def __init__(self, lat):
            self.lat = lat

        def energy_stats(self):
            return numpy_stats(self.lat, 2)

    layout = ThetaLayout(3)
    sims = [FakeSim(_lattice(s)) for s in (1, 2, 3)]
    theta = layout.pack(**THETA, dtype=torch.float32).requires_grad_(True)
    H = CpmEnergy.apply(theta, sims, layout)
    H.backward()
    th = theta.detach().double().numpy()
    want = np.mean([energy_and_grad(th, *stats_to_arrays(*s.energy_stats(), layout), layout)[1]
                    for s in sims], axis=0)
    np.testing.assert_allclose(theta.grad.double().numpy(), want, rtol=1e-6)
