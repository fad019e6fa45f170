"""The statistics kernel's pair-counting forms on lattices that load them
hard: cell ids drawn per site (most neighbour pairs unlike, up to every pair
of a thread's sites in one type-pair field), L from 8 to 256, 16-bit sites
(current after one MCS).  Three types: the three-type chunk kernel (every
engine of the launch eligible); four: byte counters; six: per-pair LDS
histogram adds; a launch mixing three- and four-type engines uses the
four-type form for all.  Every form equals the numpy reduction exactly."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import run_kwargs
from test_gradient_cpu import numpy_stats

pytestmark = pytest.mark.gpu


def _random_engine(L, ntypes, seed, medium=0.3, distinct=False):
    import gpucpm
    rng = np.random.default_rng(seed)
    if distinct:  # every site its own cell of type 1
        ncells = L * L
        ctype = np.ones(ncells + 1, np.int64)
        ids = np.arange(1, ncells + 1).reshape(L, L)
    else:
        ncells = min(L * L // 3, 1500)
        ctype = rng.integers(1, ntypes, size=ncells + 1)
        ids = rng.integers(1, ncells + 1, size=(L, L))
        ids[rng.random((L, L)) < medium] = 0
    lat = np.where(ids > 0, (ctype[ids].astype(np.uint32) << np.uint32(24)) | ids.astype(np.uint32), 0)
    sim = gpucpm.Cpm(L, 2, ncells, 10, False)
    for a in range(ntypes):
        for b in range(a, ntypes):
            sim.set_constraints(cell_type=a, other_cell_type=b, adhesion=4 + a + b)
    sim.get_state()[...] = lat.astype(np.uint32).view(np.int32)
    sim.push_to_gpu()
    sim.set_seed(seed)
    sim.run(**run_kwargs(1, 1, 1))  # the 16-bit copy is now current
    return sim


def _check(sims):
    import gpucpm
    pairs, mom = gpucpm.energy_stats_many(sims)
    for i, s in enumerate(sims):
        s.pull_from_gpu()
        p_np, m_np = numpy_stats(s.get_state().view(np.uint32).copy(), 2)
        np.testing.assert_array_equal(np.triu(pairs[i]), np.triu(p_np), err_msg=f"engine {i}")
        np.testing.assert_array_equal(mom[i], m_np, err_msg=f"engine {i}")


@pytest.mark.parametrize("ntypes", [3, 4, 6])
def test_stats_forms_on_random_lattices(ntypes):
    print("CASE stats forms", ntypes, flush=True)
    sims = [_random_engine(L, ntypes, 100 * ntypes + i) for i, L in enumerate((8, 16, 32, 64, 128, 256))]
    _check(sims)


def test_stats_three_type_fields_at_their_limit():
    """Every site its own cell of type 1 and no medium: every positive-half
    pair is a (1, 1) unlike pair, the largest count one field can take."""
    print("CASE stats field limit", flush=True)
    sims = [_random_engine(L, 3, 7 + i, distinct=True) for i, L in enumerate((8, 32, 64))]
    _check(sims)


def test_stats_mixed_launch():
    print("CASE stats mixed", flush=True)
    sims = [_random_engine(64, 3, 31), _random_engine(128, 4, 32), _random_engine(32, 3, 33)]
    _check(sims)
