"""The reference-stream mode of the oracle (rng = "minstd"): one MINSTD
state per reference thread, consumed in the reference kernel's own order.

The pin is a restatement written the way the reference kernel is written --
thread-major: each thread seeds its state once per launch and walks its
MCS, colour-round, checkerboard and update loops in turn
(kernels2d.cu:72-106,170-260,276-290,522; kernels3d.cu:86-215;
random.cu:5-28) -- whereas the oracle is phase-major (every region of one
colour phase, then the next phase).  Lattices whose copies are never
accepted keep each thread's consumption independent of the others, so the
per-thread sequences of (target, source, acceptance draw) must agree
exactly, including the reference's 3D rejection test that compares z with
y (kernels3d.cu:208).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as O

M = 2147483647.0


def next_random(seed: int) -> int:
    """random.cu:5-12: 16807 x mod (2^31 - 1), in double."""
    temp = float(seed) * 16807.0
    return int(temp - M * math.floor(temp * (1.0 / M)))


def int32(v: int) -> int:
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >= 1 << 31 else v


def reference_thread_streams(lat, d, c, u, ppt, shared, seed32, nmcs, fixed, tpb=4):
    """{thread index: [(target, source, acceptance draw or -1), ...]} of one
    launch of nmcs MCS, for a lattice that no attempt changes.  Threads are
    laid out as the reference lays them out: L / ppt per axis in blocks of
    tpb (threads_per_block, cpm.cu:363); with shared roles each thread's ppt
    sites are G = ppt / 2c regions of 2c sites, visited one block slice at a
    time (kernels2d.cu:72-86,92-106,136-146; kernels3d.cu:65-108)."""
    L = lat.shape[0]
    flat = lat.reshape(-1)
    R = 2 * c if shared else ppt
    K = R // c
    G = ppt // R  # global blocks (1 without shared roles)
    mult = c ** d // u
    nt = L // ppt

    def index(x, y, z):  # kernels.cu:20-42
        return (z * L + y) * L + x if d == 3 else y * L + x

    def neighbour(x, seed):  # random.cu:22-28
        seed = next_random(seed)
        return (x + seed % 3 - 1) & (L - 1), seed

    def region_base(t, g):  # kernels2d.cu:136-146
        return (t // tpb) * G * tpb * R + (t % tpb) * R + g * tpb * R

    out = {}
    for t in range(nt ** d):
        tx, ty, tz = t % nt, (t // nt) % nt, (t // (nt * nt) if d == 3 else 0)
        seed = next_random(int32(index(tx * ppt, ty * ppt, tz * ppt) + seed32))
        ev = []
        for _ in range(nmcs):
            for _ in range(mult):
                for gx in range(G):
                    for gy in range(G):
                        for gz in (range(G) if d == 3 else (0,)):
                            bx, by = region_base(tx, gx), region_base(ty, gy)
                            bz = region_base(tz, gz) if d == 3 else 0
                            for cx in range(K):
                                for cy in range(K):
                                    for cz in (range(K) if d == 3 else (0,)):
                                        for _ in range(u):
                                            seed = next_random(seed)
                                            a = seed % c
                                            seed = next_random(seed)
                                            b = seed % c
                                            cc = 0
                                            if d == 3:
                                                seed = next_random(seed)
                                                cc = seed % c
                                            px, py, pz = bx + cx * c + a, by + cy * c + b, bz + cz * c + cc
                                            while True:
                                                qx, seed = neighbour(px, seed)
                                                qy, seed = neighbour(py, seed)
                                                qz = 0
                                                if d == 3:
                                                    qz, seed = neighbour(pz, seed)
                                                if d == 2:
                                                    rej = qx == px and qy == py
                                                elif shared:
                                                    rej = qx == px and qy == py and qz == pz
                                                else:  # kernels3d.cu:208 compares z with y
                                                    rej = qx == px and qy == py and qz == py
                                                if not rej:
                                                    break
                                            if shared:  # drawn site = target
                                                tgt, src = index(px, py, pz), index(qx, qy, qz)
                                            else:       # drawn site = source
                                                tgt, src = index(qx, qy, qz), index(px, py, pz)
                                            draw = -1
                                            if flat[src] != flat[tgt] and not fixed[int(flat[tgt]) >> 24]:
                                                seed = next_random(seed)
                                                draw = seed
                                            ev.append((tgt, src, draw))
        out[t] = ev
    return out


def _frozen_params(fixed_type=None):
    """Copies are never accepted: every cell has area 1 and target area 1,
    so any copy costs 2 lambda_A = 200 >> the table (T = 1)."""
    kw = dict(J=np.zeros((20, 20)), lam_area=np.full(20, 100.0), target_area=np.ones(20))
    fixed = np.zeros(20, np.int64)
    if fixed_type is not None:
        fixed[fixed_type] = 1
    return O.Params.make(1.0, fixed=fixed, **kw), fixed


def _distinct_lattice(L, d, types=(1, 2)):
    n = L ** d
    ids = np.arange(1, n + 1, dtype=np.uint32)
    t = np.array(types, np.uint32)[np.arange(n) % len(types)]
    return ((t << np.uint32(24)) | ids).reshape((L,) * d), n


CASES = [  # d, L, c, u, shared, ppt, threads_per_block
    (2, 16, 2, 1, False, 4, 4),
    (2, 16, 2, 3, False, 8, 4),
    (2, 16, 1, 1, True, 2, 4),
    (2, 16, 2, 4, True, 4, 4),
    (3, 8, 1, 1, False, 2, 4),
    (3, 8, 2, 3, False, 4, 4),
    (3, 8, 1, 1, True, 2, 4),
    # shared roles with ppt > 2c: each thread loops over ppt / 2c regions per
    # axis, placed by threads_per_block
    (2, 16, 1, 1, True, 4, 1),
    (2, 16, 1, 1, True, 4, 2),
    (2, 32, 2, 2, True, 8, 2),
    (2, 32, 2, 4, True, 8, 4),
    (3, 8, 1, 1, True, 4, 2),
    (3, 16, 1, 1, True, 8, 1),
]


@pytest.mark.parametrize("d,L,c,u,shared,ppt,tpb", CASES)
@pytest.mark.parametrize("lattice", ["uniform", "distinct", "distinct_fixed"])
def test_oracle_refstream_equals_thread_major_restatement(d, L, c, u, shared, ppt, tpb, lattice):
    seed = 0x7FFF1234 if lattice == "uniform" else -19
    if lattice == "uniform":
        lat = np.full((L,) * d, (1 << 24) | 1, np.uint32)
        ncells = 1
        p, fixed = _frozen_params()
    else:
        lat, ncells = _distinct_lattice(L, d)
        p, fixed = _frozen_params(2 if lattice == "distinct_fixed" else None)
    tab = O.tally(lat, d, ncells)
    nmcs = 2
    ref = reference_thread_streams(lat, d, c, u, ppt, shared, int32(seed), nmcs, fixed, tpb)
    nev = sum(len(v) for v in ref.values())
    trace = O.sweep(lat, d, tab, p, c, u, seed & 0xFFFFFFFF, 0, nmcs, shared=shared, ppt=ppt,
                    rng="minstd", trace_cap=nev + 8, tpb=tpb)
    assert len(trace) == nev
    got = {}
    for rid, t, s, w in trace.tolist():
        got.setdefault(rid, []).append((t, s, w))
    assert got == ref
    ev = [e for v in ref.values() for e in v]
    if lattice == "uniform":
        assert all(w == -1 for (_, _, w) in ev)
    elif lattice == "distinct":
        # every copy between distinct sites takes a draw; a target equal to
        # its source comes only from the 3D z-versus-y rejection test
        assert all((w >= 1) == (t != s) for (t, s, w) in ev)
        assert any(t == s for (t, s, _) in ev) == (d == 3 and not shared)
    else:
        assert any(w == -1 for (t, s, w) in ev if t != s) and any(w >= 1 for (_, _, w) in ev)


def test_refstream_reseeds_every_call():
    """The reference seeds its threads once per launch from the same host
    seed (cpm.cu:347-360): two calls replay the same draws."""
    lat, n = _distinct_lattice(16, 2)
    p, _ = _frozen_params()
    tab = O.tally(lat, 2, n)
    a = O.sweep(lat, 2, tab, p, 2, 1, 5, 0, 1, shared=False, ppt=4, rng="minstd", trace_cap=1 << 12)
    b = O.sweep(lat, 2, tab, p, 2, 1, 5, 1, 1, shared=False, ppt=4, rng="minstd", trace_cap=1 << 12)
    np.testing.assert_array_equal(a, b)


def test_ref_threshold_table_is_exact_comparison():
    """thr31[k] = ceil(exp(-k / (scale T)) (2^31 - 1)): u < thr31[k] exactly
    when u / (2^31 - 1) < exp(-k / (scale T)) (checked in exact rationals at
    the table's edges)."""
    from fractions import Fraction
    T, scale = 7.0, 4
    thr = O.ref_threshold_table(T, scale)
    assert thr[0] == 0xFFFFFFFF
    for k in range(1, len(thr)):
        p = Fraction(math.exp(-k / (scale * T)))
        for uu in (int(thr[k]) - 1, int(thr[k])):
            assert (uu < thr[k]) == (Fraction(uu, 2147483647) < p)
    k = len(thr)
    assert math.ceil(math.exp(-k / (scale * T)) * M) <= 1


def test_refstream_sweep_keeps_tables_exact():
    """A live lattice under the reference stream: areas and perimeters equal
    a recount and the run is thread-count invariant."""
    from gpucpm import workloads as W
    wl = W.sorting2d(L=64, grid=5, edge=10, seed=3)
    import helpers
    p = helpers.oracle_params(wl)
    res = []
    for nth in (1, 4):
        lat = wl.lattice.copy()
        tab = O.tally(lat, 2, wl.ncells)
        for ch in range(3):
            O.sweep(lat, 2, tab, p, 2, 1, 77, 10 * ch, 10, nthreads=nth, shared=False, ppt=4,
                    rng="minstd")
        rec = O.tally(lat, 2, wl.ncells)
        np.testing.assert_array_equal(tab.area, rec.area)
        np.testing.assert_array_equal(tab.perim, rec.perim)
        res.append(lat)
    np.testing.assert_array_equal(res[0], res[1])
    assert np.count_nonzero(res[0] != wl.lattice) > 0
