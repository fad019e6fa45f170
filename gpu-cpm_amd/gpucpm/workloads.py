"""Synthetic workloads: the benchmark configurations as lattices + constraints.

Each builder returns a ``Workload``: the initial lattice (uint32, ids
``cell | type << 24``), the number of cells, the temperature and the list of
``set_constraints`` keyword sets that define the Hamiltonian.  Placement uses
``numpy.random.default_rng(seed)``.

- ``plain2d``    256^2, 100 cells (5x5 squares) of one type, area constraint.
- ``sorting2d``  1024^2, 63x63 squares of 10x10 tiling the centre, two types,
                 area + perimeter constraints, differential adhesion.
- ``plain3d``    256^3, 20000 cubes of 5^3 at random non-overlapping positions,
                 volume + surface constraints, 26-neighbourhood.
- ``sorting_small_cells``  the reference sorting example's Hamiltonian
                 (3x3 cells, area term only), which sorts visibly.
- ``batched_sorting``  64 sorting replicas of mixed sizes and parameters.
- ``walled``     any of these inside a one-site frame of a fixed cell type.
- ``chemotactic`` any of these with a chemotaxis term and a smooth field.
- ``active``     any of these with the act term and a random act lattice.
- ``persistent`` any 3D one with the persistence term.
The shapes may be scaled down (``L`` / ``ncells`` arguments) for tests.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    L: int
    d: int
    ncells: int
    temperature: int
    lattice: np.ndarray
    constraints: list = field(default_factory=list)
    field: np.ndarray = None  # chemotaxis field [d][L]^d (float64), or None
    act: np.ndarray = None    # initial act lattice (uint32), or None
    history: int = 1          # centroid history length of the persistence term

    def make_cpm(self, cls=None):
        """Build a Cpm with this workload's lattice and constraints (not pushed)."""
        if cls is None:
            from . import Cpm as cls
        sim = cls(self.L, self.d, self.ncells, self.temperature, self.field is not None,
                  self.history)
        for kw in self.constraints:
            kw = dict(kw)
            sim.set_constraints(kw.pop("cell_type"), **kw)
        sim.get_state()[...] = self.lattice.view(np.int32)
        if self.field is not None:
            sim.get_field()[...] = self.field
        if self.act is not None:
            sim.get_act_state()[...] = self.act.view(np.int32)
        return sim

    def params(self):
        """Dense parameter arrays (J, lam_area, target_area, lam_perim, target_perim)
        as the constraints define them (adhesion written symmetrically)."""
        J = np.zeros((20, 20), np.int64)
        la = np.zeros(20)
        a0 = np.zeros(20, np.int64)
        lp = np.zeros(20)
        p0 = np.zeros(20, np.int64)
        S = np.zeros((20, 20), np.int64)
        fixed = np.zeros(20, np.int64)
        chem = np.zeros(20)
        lam_act = np.zeros(20)
        act_max = np.zeros(20, np.int64)
        lam_pers = np.zeros(20)
        pers_diff = np.zeros(20)
        for kw in self.constraints:
            t = kw["cell_type"]
            if "adhesion" in kw:
                o = kw["other_cell_type"]
                J[t, o] = J[o, t] = kw["adhesion"]
            if "stickiness" in kw:
                S[t, kw["other_cell_type"]] = kw["stickiness"]
            if "lambda_area" in kw:
                la[t], a0[t] = kw["lambda_area"], kw["target_area"]
            if "lambda_perimeter" in kw:
                lp[t], p0[t] = kw["lambda_perimeter"], kw["target_perimeter"]
            if kw.get("fixed", -1) >= 0:
                fixed[t] = kw["fixed"]
            if kw.get("lambda_chemotaxis", -1) != -1:
                chem[t] = kw["lambda_chemotaxis"]
            if kw.get("lambda_act", -1) != -1 and kw.get("max_act", -1) != -1:
                lam_act[t], act_max[t] = kw["lambda_act"], kw["max_act"]
            if all(kw.get(k, -1) != -1 for k in ("lambda_persistence", "persistence_diffusion",
                                                  "persistence_time")):
                lam_pers[t], pers_diff[t] = kw["lambda_persistence"], kw["persistence_diffusion"]
        return dict(J=J, S=S, lam_area=la, target_area=a0, lam_perim=lp, target_perim=p0,
                    temperature=float(self.temperature), fixed=fixed, lam_chem=chem,
                    field=self.field, lam_act=lam_act, act_max=act_max, lam_pers=lam_pers,
                    pers_diff=pers_diff)


def _place_boxes(L, d, n, edge, rng, types):
    lat = np.zeros((L,) * d, np.uint32)
    placed = 0
    tries = 0
    while placed < n:
        tries += 1
        if tries > 1000 * n + 1000:
            raise RuntimeError("could not place all boxes")
        corner = rng.integers(0, L - edge + 1, size=d)
        sl = tuple(slice(int(c), int(c) + edge) for c in corner[::-1])  # [z][y][x]
        if lat[sl].any():
            continue
        placed += 1
        lat[sl] = (np.uint32(types[placed - 1]) << np.uint32(24)) | np.uint32(placed)
    return lat


def plain2d(L=256, ncells=100, seed=1):
    """256^2, 100 type-1 cells as 5x5 squares; A0=50, lam_A=50, J(1,0)=20,
    J(1,1)=40, T=20."""
    rng = np.random.default_rng(seed)
    lat = _place_boxes(L, 2, ncells, 5, rng, [1] * ncells)
    cons = [dict(cell_type=1, lambda_area=50, target_area=50),
            dict(cell_type=1, other_cell_type=0, adhesion=20),
            dict(cell_type=1, other_cell_type=1, adhesion=40)]
    return Workload("plain2d", L, 2, ncells, 20, lat, cons)


def sorting2d(L=1024, grid=63, edge=10, seed=2, J_AB=11, temperature=10, ncells=None):
    """Two-type cell sorting: grid x grid squares of edge x edge tiling the
    centre (the first ``ncells`` slots in row-major order when given); types
    1/2 Bernoulli(1/2); A0=100, lam_A=25, P0=40, lam_P=2; J_m1=J_m2=16,
    J_11=2, J_22=14, J_12=J_AB, T=10."""
    rng = np.random.default_rng(seed)
    n = grid * grid if ncells is None else int(ncells)
    types = rng.integers(1, 3, size=n)
    lat = np.zeros((L, L), np.uint32)
    off = (L - grid * edge) // 2
    for k in range(n):
        gy, gx = divmod(k, grid)
        y0, x0 = off + gy * edge, off + gx * edge
        lat[y0:y0 + edge, x0:x0 + edge] = (np.uint32(types[k]) << np.uint32(24)) | np.uint32(k + 1)
    cons = []
    for t in (1, 2):
        cons.append(dict(cell_type=t, lambda_area=25, target_area=100))
        cons.append(dict(cell_type=t, lambda_perimeter=2, target_perimeter=40))
        cons.append(dict(cell_type=t, other_cell_type=0, adhesion=16))
    cons.append(dict(cell_type=1, other_cell_type=1, adhesion=2))
    cons.append(dict(cell_type=2, other_cell_type=2, adhesion=14))
    cons.append(dict(cell_type=1, other_cell_type=2, adhesion=int(J_AB)))
    return Workload("sorting2d", L, 2, n, int(temperature), lat, cons)


def sorting_small_cells(L=256, seed=5, J_AB=11, temperature=10):
    """The Hamiltonian of the reference's own sorting example (examples/
    sorting.py), which sorts visibly: 3x3 cells on a 6-site grid inside a
    disc of radius L/3, types 1/2 at random; lam_A = 1, A0 = 40, no
    perimeter term; J_m1 = J_m2 = 16, J_11 = 14, J_22 = 2, J_12 = J_AB,
    T = 10.  (The 10x10 squares of ``sorting2d`` under its perimeter
    pressure keep their interfaces in place, so J_12 leaves no trace there.)"""
    rng = np.random.default_rng(seed)
    lat = np.zeros((L, L), np.uint32)
    n = 0
    for y in range(0, L - 1, 6):
        for x in range(0, L - 1, 6):
            if (x - L / 2) ** 2 + (y - L / 2) ** 2 < (L * L) / 9:
                n += 1
                t = rng.integers(1, 3)
                lat[y:y + 3, x:x + 3] = (np.uint32(t) << np.uint32(24)) | np.uint32(n)
    cons = [dict(cell_type=t, lambda_area=1, target_area=40) for t in (1, 2)]
    cons += [dict(cell_type=t, other_cell_type=0, adhesion=16) for t in (1, 2)]
    cons += [dict(cell_type=1, other_cell_type=1, adhesion=14),
             dict(cell_type=2, other_cell_type=2, adhesion=2),
             dict(cell_type=1, other_cell_type=2, adhesion=int(round(J_AB)))]
    return Workload("sorting_small_cells", L, 2, n, int(temperature), lat, cons)


def batched_sorting(nrep=64, seed=4, sizes=(128, 256, 512, 1024)):
    """Batched replicas: nrep sorting lattices with L drawn uniformly from
    ``sizes``, L^2/250 cells (10x10 squares tiling the centre), integer
    J_AB uniform in [6, 14] and integer T uniform in [5, 20]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nrep):
        L = int(rng.choice(sizes))
        n = L * L // 250
        grid = int(np.ceil(np.sqrt(n)))
        J_AB = int(rng.integers(6, 15))
        T = int(rng.integers(5, 21))
        wl = sorting2d(L=L, grid=grid, edge=10, seed=int(rng.integers(0, 2**31)), J_AB=J_AB,
                       temperature=T, ncells=n)
        wl.name = f"replica{i}"
        out.append(wl)
    return out


def walled(wl: Workload, wall_type=3, J_wall=(8, 4, 12)) -> Workload:
    """``wl`` inside a fixed wall: the lattice sites with any coordinate 0
    become one extra cell (id ``ncells + 1``) of ``wall_type``, which
    ``set_constraints(fixed=1)`` makes unmovable.  ``J_wall`` holds the
    adhesion of the wall with medium, type 1 and type 2."""
    lat = wl.lattice.copy()
    wid = wl.ncells + 1
    edge = np.zeros(lat.shape, bool)
    for ax in range(wl.d):
        sl = [slice(None)] * wl.d
        sl[ax] = 0
        edge[tuple(sl)] = True
    lat[edge] = (np.uint32(wall_type) << np.uint32(24)) | np.uint32(wid)
    cons = list(wl.constraints) + [dict(cell_type=wall_type, fixed=1)]
    for other, j in zip((0, 1, 2), J_wall):
        cons.append(dict(cell_type=wall_type, other_cell_type=other, adhesion=int(j)))
    return Workload(wl.name + "_walled", wl.L, wl.d, wid, wl.temperature, lat, cons, wl.field,
                    wl.act, wl.history)


def chemotactic(wl: Workload, lam=(0.0, 20.0, 5.0), seed=7, amplitude=2.0) -> Workload:
    """``wl`` with a chemotaxis term: lambda_chemotaxis ``lam[t]`` for types
    0, 1, 2 and a smooth random field (a sum of a few periodic modes, values
    within +-amplitude)."""
    rng = np.random.default_rng(seed)
    L, d = wl.L, wl.d
    grids = np.meshgrid(*([np.arange(L) * 2 * np.pi / L] * d), indexing="ij")[::-1]  # x, y(, z)
    f = np.zeros((d,) + (L,) * d)
    for a in range(d):
        for _ in range(3):
            k = rng.integers(-2, 3, size=d)
            ph = rng.uniform(0, 2 * np.pi)
            f[a] += np.cos(sum(int(k[i]) * grids[i] for i in range(d)) + ph)
    f *= amplitude / max(1e-12, np.abs(f).max())
    cons = list(wl.constraints) + [dict(cell_type=t, lambda_chemotaxis=float(v))
                                   for t, v in enumerate(lam) if v != 0.0]
    return Workload(wl.name + "_chemo", L, d, wl.ncells, wl.temperature, wl.lattice.copy(), cons, f,
                    wl.act, wl.history)


def active(wl: Workload, lam=(0.0, 300.0, 150.0), act_max=(0, 40, 30), seed=9) -> Workload:
    """``wl`` with the act term: lambda_act ``lam[t]`` and max_act
    ``act_max[t]`` for types 0, 1, 2, and an act lattice of recent random
    stamps (MCS in [0, 40))."""
    rng = np.random.default_rng(seed)
    cons = list(wl.constraints) + [dict(cell_type=t, lambda_act=float(l), max_act=int(m))
                                   for t, (l, m) in enumerate(zip(lam, act_max)) if l != 0.0]
    act = rng.integers(0, 40, size=wl.lattice.shape).astype(np.uint32)
    return Workload(wl.name + "_act", wl.L, wl.d, wl.ncells, wl.temperature, wl.lattice.copy(),
                    cons, wl.field, act, wl.history)


def persistent(wl: Workload, lam=(0.0, 200.0, 100.0), diffusion=(0.0, 0.8, 0.5), history=3,
               time=10) -> Workload:
    """``wl`` with the persistence term (3D): lambda_persistence ``lam[t]``
    and persistence_diffusion ``diffusion[t]`` for types 0, 1, 2, and a
    centroid history of ``history`` chunks."""
    cons = list(wl.constraints) + [dict(cell_type=t, lambda_persistence=float(l),
                                        persistence_diffusion=float(D), persistence_time=int(time))
                                   for t, (l, D) in enumerate(zip(lam, diffusion)) if l != 0.0]
    return Workload(wl.name + "_pers", wl.L, wl.d, wl.ncells, wl.temperature, wl.lattice.copy(),
                    cons, wl.field, wl.act, history)


def plain3d(L=256, ncells=20000, seed=3):
    """256^3, 20000 type-1 cubes of 5^3; V0=150, lam_V=25, S0=150, lam_S=1,
    J(1,0)=10, J(1,1)=20, T=20."""
    rng = np.random.default_rng(seed)
    lat = _place_boxes(L, 3, ncells, 5, rng, [1] * ncells)
    cons = [dict(cell_type=1, lambda_area=25, target_area=150),
            dict(cell_type=1, lambda_perimeter=1, target_perimeter=150),
            dict(cell_type=1, other_cell_type=0, adhesion=10),
            dict(cell_type=1, other_cell_type=1, adhesion=20)]
    return Workload("plain3d", L, 3, ncells, 20, lat, cons)
