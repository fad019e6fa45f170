"""Pins for the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible mistake in the oracle fails it:
published known-answer vectors for the generators, H(after) - H(before)
for dH, an independent numpy recount for the tallies, the arithmetic mean
for centroids, exp() for the acceptance table, and thread-count invariance
for the phase independence of the checkerboard schedule.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


# ------------------------------------------------------------- generators --

def test_philox_random123_kat():
    kat = _json("philox4x32_10_kat.json")
    for v in kat["vectors"]:
        ctr = [int(x, 16) for x in v["ctr"]]
        key = [int(x, 16) for x in v["key"]]
        out = [int(x) for x in O.philox(ctr, key)]
        assert out == [int(x, 16) for x in v["out"]]


def test_minstd_kat_and_double_formula():
    kat = _json("minstd_kat.json")
    x = kat["seed"]
    for _ in range(kat["steps"]):
        x = O.minstd_int(x)
    assert x == kat["value"]
    # random.cu:5-12 evaluates the step in double; it equals the integer map.
    rng = np.random.default_rng(0)
    for s in rng.integers(1, 2**31 - 1, size=2000):
        assert O.minstd_f64(int(s)) == O.minstd_int(int(s))


# ------------------------------------------------------------- parameters --

def test_energy_scale_is_smallest_dyadic():
    p = O.Params.make(10, lam_area=[0, 25, 0.5], lam_perim=[0, 0.25])
    assert O.energy_scale(p) == 4
    p = O.Params.make(10, lam_area=[0, 3])
    assert O.energy_scale(p) == 1
    p = O.Params.make(10, lam_area=[0, 0.1])
    assert O.energy_scale(p) == -1


@pytest.mark.parametrize("T,scale", [(10, 1), (20, 1), (7.5, 4), (1, 2)])
def test_threshold_table_matches_exp(T, scale):
    thr = O.threshold_table(T, scale)
    k = np.arange(1, len(thr))
    p = np.exp(-k / (scale * T))
    # floor(p * 2^32): error below one unit, monotone, ends where it hits zero
    assert np.all(thr[1:] <= p * 2.0**32)
    assert np.all(thr[1:] > p * 2.0**32 - 1)
    assert np.all(np.diff(thr[1:].astype(np.int64)) <= 0)
    assert math.floor(math.exp(-len(thr) / (scale * T)) * 2.0**32) == 0


def test_zero_temperature_table():
    thr = O.threshold_table(0.0, 1)
    assert len(thr) == 1


# ------------------------------------------------------------- energy -----

def _random_lattice(rng, L, d, ncells, ntypes=2, fill=0.7):
    shape = (L,) * d
    ids = rng.integers(1, ncells + 1, size=shape)
    ids[rng.random(shape) > fill] = 0
    types = np.concatenate([[0], rng.integers(1, ntypes + 1, size=ncells)])
    return ((types[ids].astype(np.uint32) << np.uint32(24)) | ids.astype(np.uint32)).astype(np.uint32)


def _sym(rng, n, lo, hi):
    a = rng.integers(lo, hi, size=(n, n))
    return np.triu(a) + np.triu(a, 1).T


@pytest.mark.parametrize("d,L", [(2, 4), (2, 8), (3, 4)])
def test_delta_h_equals_energy_difference(d, L):
    """dH for every (target, neighbour source) pair equals H(after) - H(before),
    with dyadic lambdas (scale 4) and a symmetric J.  H sums over live cells,
    so moves that delete a cell's last site are left out: there the
    incremental dH (like kernels2d.cu:515-518) still charges the dead cell's
    (0 - A0)^2 and (0 - P0)^2 terms."""
    rng = np.random.default_rng(L * 10 + d)
    ncells = 5
    lat = _random_lattice(rng, L, d, ncells)
    p = O.Params.make(3, J=_sym(rng, 3, 0, 30), lam_area=[0, 2.25, 3.5], target_area=[0, 4, 3],
                      lam_perim=[0, 0.75, 1.5], target_perim=[0, 10, 6])
    scale = O.energy_scale(p)
    assert scale == 4
    tab = O.tally(lat, d, ncells)
    H0 = O.hamiltonian(lat, d, ncells, p)
    flat = lat.reshape(-1)
    offs = [o for o in np.ndindex(*(3,) * d) if any(x != 1 for x in o)]
    checked = 0
    for t in range(flat.size):
        tc = np.unravel_index(t, lat.shape)  # (z, y, x) order
        for o in offs:
            sc = tuple((tc[i] + o[i] - 1) % L for i in range(d))
            s = int(np.ravel_multi_index(sc, lat.shape))
            if flat[s] == flat[t]:
                continue
            if (flat[t] & 0xFFFFFF) and tab.area[flat[t] & 0xFFFFFF] == 1:
                continue
            dh = O.delta_h(lat, d, tab, p, scale, t, s)
            after = lat.copy()
            after.reshape(-1)[t] = flat[s]
            H1 = O.hamiltonian(after, d, ncells, p)
            assert dh == round((H1 - H0) * scale), (t, s)
            checked += 1
    assert checked > 20


def test_stickiness_closed_form():
    """Stickiness adds -S[t_t][t_n] for every neighbour n unlike the target
    (kernels2d.cu:400)."""
    rng = np.random.default_rng(3)
    L, d, ncells = 8, 2, 4
    lat = _random_lattice(rng, L, d, ncells)
    J = _sym(rng, 3, 0, 20)
    S = rng.integers(0, 9, size=(3, 3))
    p0 = O.Params.make(5, J=J)
    p1 = O.Params.make(5, J=J, S=S)
    tab = O.tally(lat, d, ncells)
    flat = lat.reshape(-1)
    for t in range(flat.size):
        y, x = divmod(t, L)
        s = ((y + 1) % L) * L + x
        if flat[s] == flat[t]:
            continue
        nb = [lat[(y + dy) % L, (x + dx) % L] for dy in (-1, 0, 1) for dx in (-1, 0, 1) if dx or dy]
        extra = sum(int(S[flat[t] >> 24, n >> 24]) for n in nb if n != flat[t])
        assert O.delta_h(lat, d, tab, p1, 1, t, s) == O.delta_h(lat, d, tab, p0, 1, t, s) - extra


# ------------------------------------------------------------- tallies ----

def _numpy_recount(lat, d, ncells):
    """Independent recount with np.roll: area and Moore/26 unlike-neighbour count."""
    spin = (lat & 0xFFFFFF).astype(np.int64)
    area = np.bincount(spin.reshape(-1), minlength=ncells + 1)
    unlike = np.zeros(lat.shape, np.int64)
    for o in np.ndindex(*(3,) * d):
        if all(x == 1 for x in o):
            continue
        unlike += np.roll(spin, shift=[1 - x for x in o], axis=tuple(range(d))) != spin
    perim = np.bincount(spin.reshape(-1), weights=unlike.reshape(-1), minlength=ncells + 1)
    area[0] = perim[0] = 0
    return area, perim.astype(np.int64)


@pytest.mark.parametrize("d,L", [(2, 32), (3, 16)])
def test_tally_matches_numpy_recount(d, L):
    rng = np.random.default_rng(d)
    lat = _random_lattice(rng, L, d, 30)
    tab = O.tally(lat, d, 30)
    area, perim = _numpy_recount(lat, d, 30)
    np.testing.assert_array_equal(tab.area, area)
    np.testing.assert_array_equal(tab.perim, perim)


def test_tally_rejects_bad_ids():
    with pytest.raises(ValueError):
        O.tally(np.zeros((6, 6), np.uint32), 2, 1)  # L must be a power of two
    lat = np.zeros((8, 8), np.uint32)
    lat[0, 0] = 9
    with pytest.raises(ValueError):
        O.tally(lat, 2, 8)
    lat[0, 0] = (20 << 24) | 1
    with pytest.raises(ValueError):
        O.tally(lat, 2, 8)


def test_centroid_is_arithmetic_mean_for_compact_cells():
    L = 32
    lat = np.zeros((L, L), np.uint32)
    lat[3:8, 10:14] = (1 << 24) | 1
    lat[20:23, 2:9] = (1 << 24) | 2
    lat[0:2, 30:32] = (1 << 24) | 3
    lat[30:32, 30:32] = (1 << 24) | 3  # wraps across the y boundary
    tab = O.tally(lat, 2, 3)
    cen = O.centroids(L, 2, tab)
    for c in (1, 2):
        ys, xs = np.nonzero((lat & 0xFFFFFF) == c)
        assert cen[c, 0] == pytest.approx(xs.mean())
        assert cen[c, 1] == pytest.approx(ys.mean())
    # the wrapped cell covers rows 30, 31, 0, 1: unwrapped mean 31.5
    assert cen[3, 0] == pytest.approx(30.5)
    assert cen[3, 1] == pytest.approx(31.5)
    assert np.all(np.isnan(cen[0]))


# ------------------------------------------------------------- sweep ------

def _sorting_small(L=64, seed=0):
    rng = np.random.default_rng(seed)
    lat = np.zeros((L, L), np.uint32)
    n = 0
    for y0 in range(8, L - 8, 6):
        for x0 in range(8, L - 8, 6):
            n += 1
            lat[y0:y0 + 6, x0:x0 + 6] = (np.uint32(rng.integers(1, 3)) << np.uint32(24)) | np.uint32(n)
    J = [[0, 16, 16], [16, 2, 11], [16, 11, 14]]
    p = O.Params.make(10, J=J, lam_area=[0, 2, 2], target_area=[0, 36, 36],
                      lam_perim=[0, 0.5, 0.5], target_perim=[0, 60, 60])
    return lat, n, p


@pytest.mark.parametrize("c,u", [(1, 1), (2, 4), (4, 5), (8, 64)])
def test_sweep_keeps_tables_exact_and_is_thread_invariant(c, u):
    lat, n, p = _sorting_small()
    tab = O.tally(lat, 2, n)
    lat1, tab1 = lat.copy(), tab.copy()
    O.sweep(lat, 2, tab, p, c, u, seed=11, mcs0=0, nmcs=25, nthreads=1)
    O.sweep(lat1, 2, tab1, p, c, u, seed=11, mcs0=0, nmcs=25, nthreads=8)
    np.testing.assert_array_equal(lat, lat1)
    np.testing.assert_array_equal(tab.csum, tab1.csum)
    area, perim = _numpy_recount(lat, 2, n)
    np.testing.assert_array_equal(tab.area, area)
    np.testing.assert_array_equal(tab.perim, perim)
    assert np.count_nonzero(lat != _sorting_small()[0]) > 50


def test_sweep_3d_tables_exact():
    rng = np.random.default_rng(5)
    L = 16
    lat = np.zeros((L, L, L), np.uint32)
    for k in range(1, 9):
        z, y, x = rng.integers(0, L - 4, size=3)
        lat[z:z + 4, y:y + 4, x:x + 4] = (1 << 24) | k
    p = O.Params.make(20, J=[[0, 10], [10, 20]], lam_area=[0, 5], target_area=[0, 80],
                      lam_perim=[0, 0.5], target_perim=[0, 400])
    tab = O.tally(lat, 3, 8)
    O.sweep(lat, 3, tab, p, 2, 8, seed=3, mcs0=0, nmcs=10, nthreads=4)
    area, perim = _numpy_recount(lat, 3, 8)
    np.testing.assert_array_equal(tab.area, area)
    np.testing.assert_array_equal(tab.perim, perim)


def test_sweep_split_equals_single_run():
    lat, n, p = _sorting_small(seed=2)
    a, ta = lat.copy(), O.tally(lat, 2, n)
    b, tb = lat.copy(), O.tally(lat, 2, n)
    O.sweep(a, 2, ta, p, 2, 4, 9, 0, 30)
    O.sweep(b, 2, tb, p, 2, 4, 9, 0, 12)
    O.sweep(b, 2, tb, p, 2, 4, 9, 12, 18)
    np.testing.assert_array_equal(a, b)


def test_sweep_copies_stay_local():
    """With c = 1, u = 1 one MCS is four colour phases and every site is a
    target exactly once.  A target copies a neighbour of another colour,
    which itself may have copied in an earlier phase, so after one MCS every
    site holds a value found within distance 4 of it in the initial lattice,
    and at distance 1 for sites of the first colour phase (even x, even y)."""
    lat, n, p = _sorting_small(seed=4)
    p.temperature = 1000.0  # accept almost everything
    before = lat.copy()
    tab = O.tally(lat, 2, n)
    O.sweep(lat, 2, tab, p, 1, 1, 5, 0, 1)
    L = lat.shape[0]
    changed = np.argwhere(lat != before)
    assert len(changed) > 100
    for y, x in changed:
        r = 1 if (x % 2 == 0 and y % 2 == 0) else 4
        win = before[np.ix_([(y + k) % L for k in range(-r, r + 1)],
                            [(x + k) % L for k in range(-r, r + 1)])]
        assert lat[y, x] in win


def test_sweep_rejects_bad_schedule():
    lat, n, p = _sorting_small()
    tab = O.tally(lat, 2, n)
    for c, u in [(3, 1), (128, 1), (2, 0), (2, 5)]:
        with pytest.raises(ValueError):
            O.sweep(lat, 2, tab, p, c, u, 1, 0, 1)


def test_cell_sync_is_deterministic_and_stricter():
    """cell_sync (kernels.cu:123-141, restated as: at most one committed flip
    per cell per colour phase, lowest sub-block index wins) is thread-count
    invariant, keeps exact tables and commits no more flips than the plain
    sweep from the same state."""
    lat, n, p = _sorting_small(seed=3)
    p.temperature = 1000.0
    res = []
    for cs in (False, True):
        for nt in (1, 8):
            l = lat.copy()
            t = O.tally(l, 2, n)
            O.sweep(l, 2, t, p, 1, 1, 5, 0, 1, nt, cell_sync=cs)
            area, perim = _numpy_recount(l, 2, n)
            np.testing.assert_array_equal(t.area, area)
            np.testing.assert_array_equal(t.perim, perim)
            res.append(l)
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[2], res[3])
    plain, sync = res[0], res[2]
    assert 0 < np.count_nonzero(sync != lat) < np.count_nonzero(plain != lat)
    # at most one flip per cell per colour phase: over one MCS (four
    # phases) no cell's area moves by more than 4
    t0 = O.tally(lat, 2, n)
    t1 = O.tally(sync, 2, n)
    assert np.max(np.abs(t1.area.astype(int) - t0.area)) <= 4
    # u > 1: one claim/commit round per attempt index, so per MCS (4 colour
    # phases x u indices, one round) a cell changes at most 4 u times
    res = []
    for nth in (1, 8):
        a = lat.copy()
        t = O.tally(a, 2, n)
        O.sweep(a, 2, t, p, 2, 4, 5, 0, 1, nthreads=nth, cell_sync=True)
        res.append((a, t))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    area, perim = _numpy_recount(res[0][0], 2, n)
    np.testing.assert_array_equal(res[0][1].area, area)
    np.testing.assert_array_equal(res[0][1].perim, perim)
    t0 = O.tally(lat, 2, n)
    assert np.max(np.abs(res[0][1].area.astype(int) - t0.area)) <= 16
    plain4 = lat.copy()
    O.sweep(plain4, 2, O.tally(plain4, 2, n), p, 2, 4, 5, 0, 1)
    assert 0 < np.count_nonzero(res[0][0] != lat) <= np.count_nonzero(plain4 != lat)


# ------------------------------------------------------- fixed types ------

def _walled_blocks(L=32, fixed=1):
    """No medium at all: a one-site frame (cell 5, type 3) and four type-1
    cells filling the inside.  The wall adheres badly to cells and well to
    medium, so copies of medium from the wall are favoured."""
    lat = np.zeros((L, L), np.uint32)
    h = L // 2
    for k, (y0, x0) in enumerate([(0, 0), (0, h), (h, 0), (h, h)]):
        lat[y0:y0 + h, x0:x0 + h] = (1 << 24) | (k + 1)
    lat[0, :] = lat[:, 0] = (3 << 24) | 5
    J = np.zeros((20, 20))
    J[1, 1] = 6
    J[0, 1] = J[1, 0] = 2
    J[1, 3] = J[3, 1] = 30
    fx = np.zeros(20)
    fx[3] = fixed
    p = O.Params.make(10, J=J, lam_area=[0, 1, 0, 0], target_area=[0, 225, 0, 0],
                      fixed=fx)
    return lat, 5, p


@pytest.mark.parametrize("serial", [False, True])
def test_fixed_type_never_moves_and_copies_medium(serial):
    """settings.h:24 / kernels2d.cu:281-290: a fixed target keeps its value
    and a copy from a fixed source writes medium.  The lattice starts with
    no medium, so every medium site after the run was created by the second
    rule; with the wall not fixed no medium can ever appear and the wall
    moves."""
    for fixed in (1, 0):
        lat, n, p = _walled_blocks(fixed=fixed)
        before = lat.copy()
        tab = O.tally(lat, 2, n)
        if serial:
            O.serial_sweep(lat, 2, tab, p, seed=3, mcs0=0, nmcs=20)
        else:
            O.sweep(lat, 2, tab, p, 2, 4, seed=3, mcs0=0, nmcs=20)
        wall_before = (before >> 24) == 3
        wall_after = (lat >> 24) == 3
        area, perim = _numpy_recount(lat, 2, n)
        np.testing.assert_array_equal(tab.area, area)
        np.testing.assert_array_equal(tab.perim, perim)
        if fixed:
            np.testing.assert_array_equal(wall_after, wall_before)
            assert np.all(lat[wall_before] == before[wall_before])
            nmed = int(np.count_nonzero(lat == 0))
            assert nmed > 10
            # medium first appears next to the wall; after 20 MCS it is still
            # within a few sites of it (x or y small, or wrapped to L-1)
            ys, xs = np.nonzero(lat == 0)
            d = np.minimum(np.minimum(xs, 32 - xs), np.minimum(ys, 32 - ys))
            assert d.max() <= 10
        else:
            assert np.count_nonzero(lat == 0) == 0
            assert np.count_nonzero(wall_after != wall_before) > 0


@pytest.mark.parametrize("stick,expect", [(0, "most"), (100, "most"), (-100, "none")])
def test_fixed_source_onto_medium_stamps_act(stick, expect):
    """kernels2d.cu:281-290,532-534: a copy from a fixed source onto a
    medium target is a medium -> medium attempt that the reference still
    evaluates and, when accepted, stamps act[target] = tick.  A fixed frame in
    pure medium, no cells: that is the only possible move, the lattice never
    changes, and only medium sites next to the frame get stamps.  Its energy
    is the adhesion difference, -S(0, wall) per wall neighbour (stickiness of
    medium to the wall): 0 or negative is always accepted, and with c = 1
    every site is a target once per MCS with a uniform source, so a site with
    w wall neighbours is stamped within k MCS with probability
    1 - (1 - w/8)^k; +100 per neighbour at T = 10 is never accepted."""
    L, k, mcs0 = 32, 3, 5
    lat = np.zeros((L, L), np.uint32)
    lat[0, :] = lat[:, 0] = (3 << 24) | 1
    act = np.zeros((L, L), np.uint32)
    fx = np.zeros(20)
    fx[3] = 1
    S = np.zeros((20, 20), np.int64)
    S[0, 3] = stick
    p = O.Params.make(10, S=S, fixed=fx, lam_act=[0, 5.0], act_max=[0, 40], act=act)
    before = lat.copy()
    tab = O.tally(lat, 2, 1)
    O.sweep(lat, 2, tab, p, 1, 1, seed=17, mcs0=mcs0, nmcs=k)
    np.testing.assert_array_equal(lat, before)
    wall = (lat >> 24) == 3
    near = np.zeros_like(wall)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            near |= np.roll(np.roll(wall, dy, 0), dx, 1)
    near &= ~wall
    stamped = act != 0
    assert not np.any(stamped & ~near)
    assert np.all((act[stamped] >= mcs0) & (act[stamped] < mcs0 + k))
    if expect == "none":
        assert not np.any(stamped)
        return
    # straight runs of the frame: w = 3 wall neighbours
    w = np.zeros((L, L), int)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            w += np.roll(np.roll(wall, dy, 0), dx, 1)
    sel = near & (w == 3)
    frac = stamped[sel].mean()
    pexp = 1 - (5 / 8) ** k
    sd = np.sqrt(pexp * (1 - pexp) / sel.sum())
    assert abs(frac - pexp) < 4 * sd, (frac, pexp, sd)


# --------------------------------------------------------- chemotaxis ------

def _one_cell(L=16, d=2, t=1):
    lat = np.zeros((L,) * d, np.uint32)
    sl = tuple(slice(L // 2 - 2, L // 2 + 2) for _ in range(d))
    lat[sl] = (np.uint32(t) << np.uint32(24)) | np.uint32(1)
    return lat


@pytest.mark.parametrize("d", [2, 3])
def test_chemotaxis_energy_closed_form(d):
    """kernels2d.cu:325-334 / kernels3d.cu:314-325: a copy from source s to
    target t adds -lambda (d . f(s)) with d the unit vector from s to t;
    lambda of the source type, or of the target type when the source is
    medium.  The expected values come from the continuous formula; the
    oracle's fixed point (scale 2^8) must agree to within its rounding."""
    L = 16
    lat = _one_cell(L, d)
    n = lat.size
    rng = np.random.default_rng(3)
    field = rng.uniform(-3, 3, size=(d,) + lat.shape)
    lam = np.zeros(20)
    lam[1] = 2.5
    p0 = O.Params.make(10, J=[[0, 4], [4, 8]], lam_area=[0, 1], target_area=[0, 16])
    p1 = O.Params.make(10, J=[[0, 4], [4, 8]], lam_area=[0, 1], target_area=[0, 16],
                       lam_chem=lam, field=field)
    scale = O.energy_scale(p1)
    assert scale == 256
    tab = O.tally(lat, d, 1)
    flat = lat.reshape(-1)
    coords = np.stack(np.unravel_index(np.arange(n), lat.shape)[::-1], 1)  # x, y(, z)
    fq = field.reshape(d, -1)
    checked = 0
    for t in range(n):
        for off in np.ndindex(*(3,) * d):
            o = np.array(off) - 1
            if not o.any():
                continue
            sc = (coords[t] + o) % L
            s = int(np.ravel_multi_index(tuple(sc[::-1]), lat.shape))
            if flat[s] == flat[t]:
                continue
            e = O.delta_h(lat, d, tab, p1, scale, t, s) - O.delta_h(lat, d, tab, p0, scale, t, s)
            lam_used = 2.5 if (flat[s] >> 24) == 1 or (flat[t] >> 24) == 1 else 0.0
            dvec = -o / np.linalg.norm(o)
            expect = -lam_used * float(dvec @ fq[:, s]) * scale
            # final rounding (0.5) plus the field's 2^-16 and the
            # multiplier's 2^-8 quantization (below 0.05 at these sizes)
            assert abs(e - expect) <= 0.55, (t, s, e, expect)
            checked += 1
    assert checked > 50


def test_chemotaxis_moves_cell_up_the_field():
    """A cell in a uniform field f = (f, 0) with lambda > 0 extends along +x
    (the copy direction aligned with f lowers the energy) and its centroid
    drifts in +x; with lambda = 0 it does not.  Over 300 MCS the free cell's
    x displacement has a spread of about 4.4 sites per seed, so the means
    over six seeds are compared (standard error about 1.8)."""
    L = 64
    lat0 = np.zeros((L, L), np.uint32)
    lat0[28:36, 28:36] = (1 << 24) | 1
    field = np.zeros((2, L, L))
    field[0] = 1.0
    drift = {}
    for lam in (0.0, 4.0):
        dx = []
        for seed in range(1, 7):
            lat = lat0.copy()
            p = O.Params.make(10, J=[[0, 8], [8, 16]], lam_area=[0, 2], target_area=[0, 64],
                              lam_chem=[0, lam], field=field)
            tab = O.tally(lat, 2, 1)
            x0 = O.centroids(L, 2, tab)[1, 0]
            O.sweep(lat, 2, tab, p, 1, 1, seed=seed, mcs0=0, nmcs=300)
            area, perim = _numpy_recount(lat, 2, 1)
            np.testing.assert_array_equal(tab.area, area)
            np.testing.assert_array_equal(tab.perim, perim)
            c = O.centroids(L, 2, tab)[1]
            dx.append((c[0] - x0 + L / 2) % L - L / 2)
        drift[lam] = float(np.mean(dx))
    assert drift[4.0] > 10.0, drift
    assert abs(drift[0.0]) < 5.5, drift


# ---------------------------------------------------------------- act ------

def test_act_geometric_mean_fixed_point():
    """The fixed-point geometric mean (log2 table, 2^x table + second-order
    term) against numpy in double, over random value sets of 1..27 values."""
    rng = np.random.default_rng(8)
    for _ in range(400):
        n = int(rng.integers(1, 28))
        v = rng.integers(1, 200, size=n)
        g = O.act_geo_mean(sum(O.act_log2(int(x)) for x in v), n) / 65536.0
        ref = float(np.exp(np.mean(np.log(v.astype(np.float64)))))
        assert abs(g - ref) <= 2e-6 * ref + 2.0 / 65536, (v, g, ref)
    assert O.act_geo_mean(O.act_log2(1), 1) == 65536
    assert O.act_geo_mean(3 * O.act_log2(8), 3) == 8 * 65536


def _act_setup(L=16, d=2, lam=40.0, amax=20, seed=1):
    rng = np.random.default_rng(seed)
    lat = np.zeros((L,) * d, np.uint32)
    sl = tuple(slice(4, 10) for _ in range(d))
    lat[sl] = (1 << 24) | 1
    sl2 = tuple(slice(9, 13) for _ in range(d))
    lat[sl2] = (2 << 24) | 2
    act = rng.integers(80, 101, size=lat.shape).astype(np.uint32)
    la = np.zeros(20)
    la[1], la[2] = lam, lam / 2
    am = np.zeros(20, np.int64)
    am[1], am[2] = amax, amax + 10
    p = O.Params.make(10, J=[[0, 4, 4], [4, 8, 6], [4, 6, 8]], lam_area=[0, 1, 1],
                      target_area=[0, 30, 30], lam_act=la, act_max=am, act=act)
    return lat, act, p, la, am


@pytest.mark.parametrize("d", [2, 3])
def test_act_energy_closed_form(d):
    """kernels2d.cu:312-343,386-397,417-473: -lambda/actMax (G_s - G_t), G the
    geometric mean of max(actMax + act - tick, 0) over the site and its
    same-cell neighbours (0 for medium), lambda/actMax of the target type or
    of the source type when the target is medium."""
    L = 16
    lat, act, p, la, am = _act_setup(L, d)
    scale = O.energy_scale(p)
    assert scale == 256
    tick = 100
    flat, aflat = lat.reshape(-1), act.reshape(-1)
    coords = np.stack(np.unravel_index(np.arange(lat.size), lat.shape)[::-1], 1)

    def gmean(si):
        v = flat[si]
        if v & 0xFFFFFF == 0:
            return 0.0
        amax = am[v >> 24]
        vals = [min(max(amax + int(aflat[si]) - tick, 0), amax)]
        for off in np.ndindex(*(3,) * d):
            o = np.array(off) - 1
            if not o.any():
                continue
            ni = int(np.ravel_multi_index(tuple(((coords[si] + o) % L)[::-1]), lat.shape))
            if flat[ni] == v:
                vals.append(min(max(amax + int(aflat[ni]) - tick, 0), amax))
        vals = np.array(vals, np.float64)
        return 0.0 if (vals == 0).any() else float(np.exp(np.log(vals).mean()))

    checked = 0
    for t in range(0, lat.size, 3):
        for off in np.ndindex(*(3,) * d):
            o = np.array(off) - 1
            if not o.any():
                continue
            s = int(np.ravel_multi_index(tuple(((coords[t] + o) % L)[::-1]), lat.shape))
            if flat[s] == flat[t]:
                continue
            tl = flat[t] >> 24 if flat[t] >> 24 else flat[s] >> 24
            if la[tl] == 0:
                continue
            e = O.act_energy(lat, d, p, scale, t, s, tick)
            expect = -la[tl] / am[tl] * (gmean(s) - gmean(t)) * scale
            assert abs(e - expect) <= 0.5 + 1e-3 * abs(expect) + 0.01, (t, s, e, expect)
            checked += 1
    assert checked > 20


def test_act_sweep_keeps_tables_exact_and_stamps_copies():
    """An accepted copy sets act[target] = tick (kernels2d.cu:532-534): after
    one MCS at tick m every site that changed holds act m; the tables stay
    equal to a recount; with c = 1 (decided then applied per phase) likewise,
    also with cell_sync, and the run is thread-count invariant."""
    lat, act, p, _, _ = _act_setup(32, 2)
    before, act0 = lat.copy(), act.copy()
    tab = O.tally(lat, 2, 2)
    O.sweep(lat, 2, tab, p, 2, 1, seed=4, mcs0=123, nmcs=1, nthreads=4)
    changed = lat != before
    assert changed.sum() > 5
    assert (act[changed] == 123).all()
    assert (act[~changed & (act != act0)] == 123).all()
    area, perim = _numpy_recount(lat, 2, 2)
    np.testing.assert_array_equal(tab.area, area)
    np.testing.assert_array_equal(tab.perim, perim)
    for d, cs in ((2, False), (3, False), (2, True), (3, True)):
        outs = []
        for nth in (1, 8):
            l1, a1, p1, _, _ = _act_setup(16, d, seed=3)
            before = l1.copy()
            t1 = O.tally(l1, d, 2)
            O.sweep(l1, d, t1, p1, 1, 1, seed=6, mcs0=40, nmcs=6, nthreads=nth, cell_sync=cs)
            area, perim = _numpy_recount(l1, d, 2)
            np.testing.assert_array_equal(t1.area, area)
            np.testing.assert_array_equal(t1.perim, perim)
            assert np.count_nonzero(l1 != before) > 0
            outs.append((l1, a1))
        np.testing.assert_array_equal(outs[0][0], outs[1][0])
        np.testing.assert_array_equal(outs[0][1], outs[1][1])
    # thread-count invariance with act (phase independence at c >= 2)
    la3, ac3, p3, _, _ = _act_setup(32, 2, seed=2)
    lb3, ab3 = la3.copy(), ac3.copy()
    p3b = O.Params.make(10, J=[[0, 4, 4], [4, 8, 6], [4, 6, 8]], lam_area=[0, 1, 1],
                        target_area=[0, 30, 30], lam_act=np.array(p3.lam_act),
                        act_max=np.array(p3.act_max), act=ab3)
    ta, tb = O.tally(la3, 2, 2), O.tally(lb3, 2, 2)
    O.sweep(la3, 2, ta, p3, 2, 3, seed=9, mcs0=50, nmcs=20, nthreads=1)
    O.sweep(lb3, 2, tb, p3b, 2, 3, seed=9, mcs0=50, nmcs=20, nthreads=8)
    np.testing.assert_array_equal(la3, lb3)
    np.testing.assert_array_equal(ac3, ab3)


@pytest.mark.parametrize("d,cell_sync", [(2, False), (3, False), (2, True), (3, True)])
def test_act_single_site_checkerboard_decides_on_phase_start(d, cell_sync):
    """Shared roles, c = 1, with an act term whose energy is exactly zero
    (lambda_act = 2^-10, actMax = 1: |m dG| < 2^31, so E = 0): the two-pass
    phases (decide on the phase-start lattice, then apply) accept exactly what
    the act-free sweep accepts, and stamp act[t] = tick on every copy.  With
    cell_sync the decide pass also claims and the apply pass checks the
    claims: the result is the act-free cell_sync sweep's."""
    lat, act, _, _, _ = _act_setup(16, d, seed=5)
    lat0 = lat.copy()
    la = np.zeros(20)
    la[1] = la[2] = 2.0 ** -10
    am = np.zeros(20, np.int64)
    am[1] = am[2] = 1
    J = [[0, 4, 4], [4, 8, 6], [4, 6, 8]]
    p_act = O.Params.make(10, J=J, lam_area=[0, 1, 1], target_area=[0, 30, 30], lam_act=la,
                          act_max=am, act=act)
    p_off = O.Params.make(10, J=J, lam_area=[0, 1, 1], target_area=[0, 30, 30])
    ta, tb = O.tally(lat, d, 2), O.tally(lat0, d, 2)
    act0 = act.copy()
    for m in range(4):
        before = lat.copy()
        O.sweep(lat, d, ta, p_act, 1, 1, seed=21, mcs0=100 + m, nmcs=1, cell_sync=cell_sync)
        O.sweep(lat0, d, tb, p_off, 1, 1, seed=21, mcs0=100 + m, nmcs=1, cell_sync=cell_sync)
        np.testing.assert_array_equal(lat, lat0)
        assert (act[lat != before] == 100 + m).all()
    np.testing.assert_array_equal(ta.area, tb.area)
    np.testing.assert_array_equal(ta.perim, tb.perim)
    assert np.count_nonzero(act != act0) > 0


def test_act_cell_migrates():
    """Act cells (lambda_act > 0) move persistently: over 400 MCS the
    centroid of a single act cell travels farther than without the act
    term."""
    L = 64
    dist = {}
    for lam in (0.0, 300.0):
        lat = np.zeros((L, L), np.uint32)
        lat[28:36, 28:36] = (1 << 24) | 1
        act = np.zeros((L, L), np.uint32)
        p = O.Params.make(20, J=[[0, 10], [10, 0]], lam_area=[0, 5], target_area=[0, 64],
                          lam_act=[0, lam], act_max=[0, 40], act=act)
        tab = O.tally(lat, 2, 1)
        c0 = O.centroids(L, 2, tab)[1].copy()
        path = 0.0
        prev = c0
        for k in range(8):
            O.sweep(lat, 2, tab, p, 2, 1, seed=11, mcs0=1 + 50 * k, nmcs=50, nthreads=4)
            c = O.centroids(L, 2, tab)[1].copy()
            dv = (c - prev + L / 2) % L - L / 2
            path += float(np.hypot(*dv))
            prev = c
        dist[lam] = path
    assert dist[300.0] > 4.0 * dist[0.0], dist


# -------------------------------------------------------- persistence ------

def _pref_tables(shift):
    """One 4^3 cell, then the same cell translated by shift (periodic)."""
    L = 16
    a = np.zeros((L, L, L), np.uint32)
    a[6:10, 6:10, 6:10] = (1 << 24) | 1
    b = np.roll(a, shift=tuple(shift[::-1]), axis=(0, 1, 2))
    return L, O.tally(a, 3, 1), O.tally(b, 3, 1)


def test_pref_dir_update_closed_forms():
    """kernels.cu:158-212 / cpm.cu:487-514: with diffusion 0 the preferred
    direction becomes the unit displacement of the centroid over the history
    window; with diffusion 1 it stays; with a history of 1 there is no
    displacement and it stays; displacements wrap across the boundary."""
    for shift, expect in (((2, 0, 0), (1, 0, 0)), ((0, -3, 4), (0, -0.6, 0.8)),
                          ((9, 0, 0), (-1, 0, 0))):  # 9 on L = 16 wraps to -7
        L, ta, tb = _pref_tables(shift)
        for D, H in ((0.0, 2), (1.0, 2), (0.0, 1)):
            ps = O.PrefState(L, 3, ta, H, seed=5)
            p0 = ps.pref.copy()
            ps.update(tb, [0, 1], np.array([0.0, D] + [0.0] * 18))
            p1 = ps.pref.reshape(-1, 3)[1]
            if D == 0.0 and H == 2:
                np.testing.assert_allclose(p1, expect, atol=1e-12)
            else:
                np.testing.assert_array_equal(ps.pref, p0)
            np.testing.assert_allclose(np.linalg.norm(ps.pref.reshape(-1, 3)[1:], axis=1), 1.0, atol=1e-12)
            np.testing.assert_array_equal(ps.pref_q, np.round(ps.pref * 65536).astype(np.int32))
    # diffusion 0.5 blends the old direction and the displacement
    L, ta, tb = _pref_tables((3, 0, 0))
    ps = O.PrefState(L, 3, ta, 2, seed=5)
    old = ps.pref.reshape(-1, 3)[1].copy()
    ps.update(tb, [0, 1], np.array([0.0, 0.5] + [0.0] * 18))
    mix = 0.5 * np.array([1.0, 0, 0]) + 0.5 * old
    np.testing.assert_allclose(ps.pref.reshape(-1, 3)[1], mix / np.linalg.norm(mix), atol=1e-12)


def test_persistence_energy_closed_form():
    """kernels3d.cu:327-344: -lambda (d . p), d the unit vector from source to
    target, p the source cell's preferred direction (the target cell's when the
    source is medium), lambda of the source type (target type likewise)."""
    L = 16
    lat = _one_cell(L, 3)
    n = lat.size
    tab = O.tally(lat, 3, 1)
    ps = O.PrefState(L, 3, tab, 2, seed=9)
    p0 = O.Params.make(10, J=[[0, 4], [4, 8]], lam_area=[0, 1], target_area=[0, 64])
    p1 = O.Params.make(10, J=[[0, 4], [4, 8]], lam_area=[0, 1], target_area=[0, 64],
                       lam_pers=[0, 3.0])
    p1.set_pref(ps.pref_q)
    scale = O.energy_scale(p1)
    assert scale == 256
    flat = lat.reshape(-1)
    coords = np.stack(np.unravel_index(np.arange(n), lat.shape)[::-1], 1)
    pref = ps.pref.reshape(-1, 3)
    checked = 0
    for t in range(0, n, 5):
        for off in np.ndindex(3, 3, 3):
            o = np.array(off) - 1
            if not o.any():
                continue
            s = int(np.ravel_multi_index(tuple(((coords[t] + o) % L)[::-1]), lat.shape))
            if flat[s] == flat[t]:
                continue
            cell = (flat[s] & 0xFFFFFF) or (flat[t] & 0xFFFFFF)
            e = O.delta_h(lat, 3, tab, p1, scale, t, s) - O.delta_h(lat, 3, tab, p0, scale, t, s)
            expect = -3.0 * float((-o / np.linalg.norm(o)) @ pref[cell]) * scale
            assert abs(e - expect) <= 0.55, (t, s, e, expect)
            checked += 1
    assert checked > 30


def test_persistence_makes_cells_travel():
    """A 3D cell with a persistence term, its preferred direction updated
    from the centroid history after every chunk, travels far more than one
    without."""
    L = 32
    path = {}
    for lam in (0.0, 200.0):
        lat = np.zeros((L, L, L), np.uint32)
        lat[13:19, 13:19, 13:19] = (1 << 24) | 1
        tab = O.tally(lat, 3, 1)
        ps = O.PrefState(L, 3, tab, 3, seed=3)
        p = O.Params.make(20, J=[[0, 10], [10, 0]], lam_area=[0, 5], target_area=[0, 216],
                          lam_pers=[0, lam])
        p.set_pref(ps.pref_q)
        prev, total = O.pref_centroids(L, 3, tab).reshape(-1, 3)[1].copy(), 0.0
        for k in range(20):
            O.sweep(lat, 3, tab, p, 2, 1, seed=3, mcs0=10 * k, nmcs=10, nthreads=8)
            ps.update(tab, [0, 1], np.array([0.0, 0.8] + [0.0] * 18))
            c = O.pref_centroids(L, 3, tab).reshape(-1, 3)[1].copy()
            total += float(np.linalg.norm((c - prev + L / 2) % L - L / 2))
            prev = c
        path[lam] = total
    assert path[200.0] > 5 * path[0.0], path


# ------------------------------------------------- run() roles/schedule ------

def test_schedule_colours_and_groups():
    """Shared roles: K = 2 colours, one group.  Non-shared (the reference's
    default): K = ppt / c colours; groups S with S ppt - c >= 3."""
    assert O.schedule(64, 2, 1, True, 2) == (2, 1)
    assert O.schedule(64, 2, 4, True, 99) == (2, 1)
    assert O.schedule(64, 2, 4, False, 8) == (2, 1)   # example2d.py
    assert O.schedule(64, 2, 2, False, 4) == (2, 2)   # sorting.py: ppt - c = 2
    assert O.schedule(64, 2, 1, False, 2) == (2, 2)
    assert O.schedule(64, 2, 1, False, 8) == (8, 1)
    assert O.schedule(64, 2, 1, False, 1) == (1, 4)
    assert O.schedule(4, 2, 1, False, 2) == (2, 2)
    with pytest.raises(ValueError):
        O.schedule(64, 2, 2, False, 3)


def _python_sweep(lat, d, ncells, p, c, u, seed, mcs0, nmcs, shared, ppt):
    """Plain-Python restatement of the run() phase loop (kernels2d.cu:72-243
    in the deterministic form of cpm_oracle.h): the schedule, the roles and
    the Philox counters; the energy comes from the oracle's pinned delta_h."""
    L = lat.shape[0]
    K, S = O.schedule(L, d, c, shared, ppt)
    R = K * c
    N = 3 ** d - 1
    offs = [o for o in np.ndindex(*(3,) * d) if any(x != 1 for x in o)]
    offs = [tuple(np.array(o[::-1]) - 1) for o in offs]  # (dx, dy[, dz]), dx fastest
    scale = O.energy_scale(p)
    thr = O.threshold_table(p.temperature, scale)
    flat = lat.reshape(-1)
    tab = O.tally(lat, d, ncells)
    rounds = c ** d // u
    h = L // R // S
    key = [seed & 0xFFFFFFFF, seed >> 32]
    nsb = L // c
    lc = int(np.log2(c))

    def idx(x):
        return int(sum(int(x[a]) % L * L ** a for a in range(d)))

    for m in range(nmcs):
        mcs = mcs0 + m
        for r in range(rounds):
            for pc in range(K ** d * S ** d):
                col, grp = divmod(pc, S ** d)
                cc = [(col // K ** (d - 1 - a)) % K for a in range(d)]   # cbX outermost
                gg = [(grp // S ** (d - 1 - a)) % S for a in range(d)]
                phase = r * K ** d * S ** d + pc
                snap = tab.copy()
                for qi in range(h ** d):
                    qa = [(qi // h ** a) % h for a in range(d)]           # x fastest
                    b = [(qa[a] * S + gg[a]) * K + cc[a] for a in range(d)]
                    sb = sum(b[a] * nsb ** a for a in range(d))
                    # c = 1: x-neighbouring same-colour sub-blocks share one
                    # block (the odd one takes words 3 and 0)
                    odd = qa[0] & 1 if c == 1 else 0
                    for i in range(u):
                        w = [int(x) for x in O.philox([sb - odd * S * K, mcs, phase, i], key)]
                        if odd:
                            w = [w[0], w[3], w[0], w[3]]
                        pos = [b[a] * c + ((int(w[0]) >> (a * lc)) & (c - 1)) for a in range(d)]
                        dn = (int(w[1]) * N) >> 32
                        nbr = [(pos[a] + offs[dn][a]) % L for a in range(d)]
                        t, s_ = (pos, nbr) if shared else (nbr, pos)
                        ti, si = idx(t), idx(s_)
                        vs, vt = int(flat[si]), int(flat[ti])
                        if vs == vt:
                            continue
                        dh = O.delta_h(lat, d, snap, p, scale, ti, si)
                        if not (dh <= 0 or (dh < len(thr) and int(w[2]) < int(thr[dh]))):
                            continue
                        flat[ti] = vs
                        tab = O.tally(lat, d, ncells)  # exact tables (live)
    return tab


@pytest.mark.parametrize("d,L,c,u,shared,ppt", [(2, 8, 1, 1, False, 2), (2, 8, 2, 2, False, 4),
                                                (2, 16, 1, 1, False, 8), (2, 8, 2, 3, True, 4),
                                                (3, 8, 1, 1, False, 2), (2, 16, 1, 1, True, 2),
                                                (3, 8, 1, 1, True, 2)])
def test_sweep_matches_python_restatement(d, L, c, u, shared, ppt):
    """The C oracle's schedule and roles equal a plain-Python loop over the
    phases (shared and non-shared roles)."""
    rng = np.random.default_rng(L + c + u)
    lat = np.zeros((L,) * d, np.uint32)
    for k in range(1, 4):
        sl = tuple(slice(int(x), int(x) + 3) for x in rng.integers(0, L - 3, size=d))
        lat[sl] = (np.uint32(1 + k % 2) << np.uint32(24)) | np.uint32(k)
    p = O.Params.make(20, J=[[0, 6, 6], [6, 8, 4], [6, 4, 8]], lam_area=[0, 1, 1],
                      target_area=[0, 9, 9])
    a = lat.copy()
    tab_py = _python_sweep(a, d, 3, p, c, u, 77, 5, 2, shared, ppt)
    b = lat.copy()
    tb = O.tally(b, d, 3)
    O.sweep(b, d, tb, p, c, u, 77, 5, 2, nthreads=4, shared=shared, ppt=ppt)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(tab_py.area, tb.area)
    np.testing.assert_array_equal(tab_py.perim, tb.perim)
    assert np.count_nonzero(a != lat) > 0


def test_nonshared_sweep_is_thread_invariant_and_exact():
    """Non-shared roles with groups (ppt - c < 3): parallel = serial, exact
    tables (the groups keep concurrent attempts apart)."""
    lat, n, p = _sorting_small(seed=7)
    for c, ppt, u in ((1, 2, 1), (2, 4, 1), (2, 4, 4), (4, 8, 2)):
        a, b = lat.copy(), lat.copy()
        ta, tb = O.tally(a, 2, n), O.tally(b, 2, n)
        O.sweep(a, 2, ta, p, c, u, 3, 0, 10, nthreads=1, shared=False, ppt=ppt)
        O.sweep(b, 2, tb, p, c, u, 3, 0, 10, nthreads=8, shared=False, ppt=ppt)
        np.testing.assert_array_equal(a, b)
        area, perim = _numpy_recount(a, 2, n)
        np.testing.assert_array_equal(ta.area, area)
        np.testing.assert_array_equal(ta.perim, perim)


def _counted_sweep(lat, d, n, p, c, u, seed, nmcs, nthreads=4, **kw):
    tab = O.tally(lat, d, n)
    counts = np.zeros(4, np.int64)
    O.sweep(lat, d, tab, p, c, u, seed, 0, nmcs, nthreads, counts=counts, **kw)
    return counts


def test_attempt_counters_closed_forms():
    """The oracle's attempt counters: with c = u = 1 every site is a target
    exactly once per MCS and a committed flip changes its target, so after
    one MCS the committed count is the number of changed sites; the counts
    are nested (committed <= accepted <= evaluated <= differing <= L^d per
    MCS), equal where nothing can refuse a step (committed = accepted without
    cell_sync; evaluated = differing without fixed types), and independent of
    the thread count."""
    from stats_common import sorting_lattice
    lat0, n, p, _ = sorting_lattice(L=64, seed=3)
    lat = lat0.copy()
    c1 = _counted_sweep(lat, 2, n, p, 1, 1, 11, 1)
    assert c1[3] == np.count_nonzero(lat != lat0) > 0
    assert c1[0] == c1[1] and c1[2] == c1[3]
    for nthreads in (1, 8):
        lat = lat0.copy()
        c10 = _counted_sweep(lat, 2, n, p, 2, 4, 12, 10, nthreads=nthreads)
        if nthreads == 1:
            ref = c10
        np.testing.assert_array_equal(c10, ref)
    assert 64 * 64 * 10 >= ref[0] >= ref[1] >= ref[2] >= ref[3] > 0


def test_attempt_counters_with_cell_sync_and_fixed_types():
    """cell_sync refuses accepted proposals whose cells another sub-block
    claimed (committed < accepted on a busy lattice); a fixed wall makes
    attempts that never reach the energy (evaluated < differing)."""
    from stats_common import sorting_lattice
    lat0, n, p, _ = sorting_lattice(L=64, seed=4)
    lat = lat0.copy()
    cs = _counted_sweep(lat, 2, n, p, 2, 1, 13, 10, cell_sync=True)
    assert cs[3] < cs[2] <= cs[1] == cs[0]
    wall = lat0.copy()
    wall[0, :] = wall[:, 0] = (3 << 24) | (n + 1)
    pw = O.Params.make(10, J=[[0, 16, 16, 5], [16, 14, 11, 5], [16, 11, 2, 5], [5, 5, 5, 0]],
                       lam_area=[0, 1, 1, 0], target_area=[0, 40, 40, 0], fixed=[0, 0, 0, 1])
    cw = _counted_sweep(wall, 2, n + 1, pw, 1, 1, 14, 10)
    assert cw[1] < cw[0] and cw[2] == cw[3]
    assert np.all(wall[0, :] == (3 << 24) | (n + 1))
