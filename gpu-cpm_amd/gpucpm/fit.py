"""Contrastive parameter fitting driven by the dH/dtheta backward pass.

The gradient of the negative log-likelihood of data under the Boltzmann
model exp(-H(sigma; theta) / T) is (E_data[dH/dtheta] - E_model[dH/dtheta]) / T
(``autograd.contrastive_gradient``).  For one adhesion entry J_ab the
statistic is N_ab, the number of unlike-cell neighbour pairs of types a, b.

``fit_adhesion`` fits J_ab so that the model's N_ab after ``mcs`` Monte
Carlo steps from fresh initial lattices matches the data's: every
evaluation builds the model replicas at the current J (common initial
lattices and seeds across evaluations), runs them as one batch for ``mcs``
MCS and takes the contrastive gradient g(J) = E_data[N_ab] - E_model[N_ab]
through the autograd layer; E_model falls with J, so g rises with J and its
zero is the fit.  Until g changes sign J moves by Newton-like steps, g
divided by the slope estimated from the model at J0 and J0 - 1 (damped,
clipped); once two evaluations bracket the zero, by regula falsi inside the
bracket (with the Illinois halving, so the bracket keeps shrinking from
both sides).  J is kept on the engine's 2^-8 grid.  Sampling both sides
after the same number of MCS compares like with like when the dynamics
have not reached equilibrium (cell sorting coarsens for a long time).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .autograd import ThetaLayout, contrastive_gradient


@dataclass
class FitResult:
    J: float                              # the fitted J_ab
    history: list = field(default_factory=list)   # (J, gradient) per step
    slope: float = 0.0                    # d E_model[N_ab] / dJ used for the steps


def _quantize(J):
    return round(float(J) * 256.0) / 256.0


def fit_adhesion(data_sims, make_model, pair, J0, mcs, run_kwargs, steps=8, damping=0.7,
                 clip=1.0, ntypes=3, theta=None):
    """Fit J_pair from J0 toward the data.

    data_sims: pushed and already run Cpm engines (the data samples).
    make_model(J): fresh, pushed Cpm engines with J_pair = J (the same
        initial lattices and seeds on every call).
    run_kwargs: Cpm.run keyword arguments without iterations /
        inner_iterations; the model runs ``mcs`` MCS per evaluation.
    theta: the model's parameter vector in ThetaLayout(ntypes) (only the
        pair's own entry changes; the others do not enter its gradient).
    """
    from . import Batch
    layout = ThetaLayout(ntypes)
    a, b = sorted(pair)
    k = layout.pairs.index((a, b))
    th = torch.zeros(layout.size, dtype=torch.float64) if theta is None else theta.clone().double()

    def grad_at(J):
        sims = make_model(J)
        Batch(sims).run(**run_kwargs, iterations=mcs, inner_iterations=mcs)
        for s in sims:
            s.synchronize()
        th[k] = J
        g = contrastive_gradient(th, data_sims, sims, layout)
        return float(g[k])

    J = _quantize(J0)
    g = grad_at(J)
    Jp = _quantize(J - 1.0)
    gp = grad_at(Jp)
    slope = g - gp
    if not slope > 0:
        raise RuntimeError(f"fit_adhesion: the model's N_{a}{b} does not respond to J near {J0} "
                           f"(gradient {g:.4g} at J0, {gp:.4g} at J0 - 1)")
    res = FitResult(J=J, slope=slope)
    res.history += [(J, g), (Jp, gp)]
    lo = hi = None  # evaluations with g < 0 and g > 0: [J, g, times kept]
    for Jx, gx in ((J, g), (Jp, gp)):
        if gx < 0 and (lo is None or Jx > lo[0]):
            lo = [Jx, gx, 0]
        if gx > 0 and (hi is None or Jx < hi[0]):
            hi = [Jx, gx, 0]
    cur_J, cur_g = Jp, gp
    for _ in range(steps):
        if lo is not None and hi is not None:
            Jn = lo[0] - lo[1] * (hi[0] - lo[0]) / (hi[1] - lo[1])
        else:
            Jn = cur_J - float(np.clip(damping * cur_g / slope, -clip, clip))
        Jn = _quantize(Jn)
        if any(Jn == Jh for Jh, _ in res.history):
            break  # the zero is resolved to the 2^-8 grid
        gn = grad_at(Jn)
        res.history.append((Jn, gn))
        cur_J, cur_g = Jn, gn
        if gn == 0:
            lo = hi = None
            res.J = Jn
            return res
        side, other = (lo, hi) if gn < 0 else (hi, lo)
        if side is None:
            side = [Jn, gn, 0]
        else:
            side[:] = [Jn, gn, side[2] + 1]
            if other is not None:
                other[2] = 0
            if other is not None and side[2] >= 2:
                other[1] /= 2.0  # Illinois: the end kept twice counts half
                side[2] = 0
        if gn < 0:
            lo = side
        else:
            hi = side
    if lo is not None and hi is not None:
        res.J = _quantize(lo[0] - lo[1] * (hi[0] - lo[0]) / (hi[1] - lo[1]))
    else:
        res.J = cur_J
    return res
