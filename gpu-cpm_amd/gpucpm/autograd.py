"""Differentiable CPM energy: H(sigma; theta) and dH/dtheta through torch.autograd.

The Hamiltonian of the sweep (adhesion, area and perimeter terms;
stickiness enters dH asymmetrically and is not part of H) is

    H = sum_{a<=b} J_ab N_ab
      + sum_t lam_A[t] sum_{c: type t} (A_c - A0[t])^2
      + sum_t lam_P[t] sum_{c: type t} (P_c - P0[t])^2

with N_ab the number of unlike-cell neighbour pairs of types {a, b}.  The
HIP statistics kernel (``Cpm.energy_stats``) reduces the lattice and cell
tables to theta-independent integers: N_ab and, per type, the count, sum A,
sum A^2, sum P and sum P^2.  From those, H and every partial derivative are
closed-form polynomials in theta.  The sums are exact integers and they are
combined in float64; the results are delivered in the dtype of theta.

theta layout for ``ntypes`` types (``ThetaLayout``):
    J upper triangle (a <= b, row-major) | lam_A[t] | A0[t] | lam_P[t] | P0[t]
"""
from __future__ import annotations

import numpy as np
import torch


class ThetaLayout:
    def __init__(self, ntypes: int):
        self.nt = ntypes
        self.pairs = [(a, b) for a in range(ntypes) for b in range(a, ntypes)]
        self.nj = len(self.pairs)
        self.size = self.nj + 4 * ntypes

    def pack(self, J, lam_area, target_area, lam_perim, target_perim, dtype=torch.float32):
        J = np.asarray(J, np.float64)
        v = [J[a, b] for a, b in self.pairs]
        for arr in (lam_area, target_area, lam_perim, target_perim):
            v.extend(float(x) for x in list(arr)[: self.nt])
        return torch.tensor(v, dtype=dtype)

    def names(self):
        n = [f"J[{a},{b}]" for a, b in self.pairs]
        for key in ("lam_A", "A0", "lam_P", "P0"):
            n.extend(f"{key}[{t}]" for t in range(self.nt))
        return n


def stats_to_arrays(pairs: np.ndarray, mom: np.ndarray, layout: ThetaLayout):
    """Float64 arrays of the statistics used by the closed forms."""
    N = np.array([pairs[a, b] for a, b in layout.pairs], np.float64)
    m = mom[: layout.nt].astype(np.float64)  # (nt, 5): n, SA, SA2, SP, SP2
    return N, m


def energy_and_grad(theta: np.ndarray, N: np.ndarray, m: np.ndarray, layout: ThetaLayout):
    """H(theta) and dH/dtheta in float64 from the statistics."""
    nt, nj = layout.nt, layout.nj
    J = theta[:nj]
    la, a0, lp, p0 = (theta[nj + k * nt: nj + (k + 1) * nt] for k in range(4))
    n, SA, SA2, SP, SP2 = (m[:, k] for k in range(5))
    qa = SA2 - 2.0 * a0 * SA + n * a0 * a0  # sum (A - A0)^2
    qp = SP2 - 2.0 * p0 * SP + n * p0 * p0
    H = float(np.dot(J, N) + np.dot(la, qa) + np.dot(lp, qp))
    g = np.concatenate([N, qa, la * (2.0 * n * a0 - 2.0 * SA), qp, lp * (2.0 * n * p0 - 2.0 * SP)])
    return H, g


class CpmEnergy(torch.autograd.Function):
    """H(sigma; theta) averaged over the given simulations' current lattices.

    forward(theta, sims, layout) runs the statistics kernel on every sim;
    backward returns grad_output * mean_s dH(sigma_s)/dtheta.
    """

    @staticmethod
    def forward(ctx, theta, sims, layout):
        from . import Cpm, energy_stats_many
        th = theta.detach().to(torch.float64).cpu().numpy()
        Hs, gs = [], []
        if all(isinstance(s, Cpm) for s in sims):  # engines: one batched call
            stats = zip(*energy_stats_many(sims))
        else:  # any object with energy_stats()
            stats = (s.energy_stats() for s in sims)
        for pairs, mom in stats:
            N, m = stats_to_arrays(pairs, mom, layout)
            H, g = energy_and_grad(th, N, m, layout)
            Hs.append(H)
            gs.append(g)
        grad = torch.tensor(np.mean(gs, axis=0), dtype=theta.dtype, device=theta.device)
        ctx.save_for_backward(grad)
        return torch.tensor(float(np.mean(Hs)), dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None


def energy(theta, sims, layout):
    """Mean H over ``sims`` at parameters ``theta`` (differentiable in theta)."""
    return CpmEnergy.apply(theta, list(sims), layout)


def contrastive_gradient(theta, data_sims, model_sims, layout):
    """E_data[dH/dtheta] - E_model[dH/dtheta]: the gradient of the negative
    log-likelihood of the data under the Boltzmann model, times T (the
    descent direction moves the model's statistics toward the data's)."""
    t1 = theta.detach().clone().requires_grad_(True)
    energy(t1, data_sims, layout).backward()
    t2 = theta.detach().clone().requires_grad_(True)
    energy(t2, model_sims, layout).backward()
    return t1.grad - t2.grad
