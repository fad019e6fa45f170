"""Stochastic observables of the engine's checkerboard dynamics (phase-start
snapshot of the cell tables, oracle.sweep, which the GPU reproduces bit for
bit) against the two other executions they must be read against, over
seeds, at 128^2 (the reference's examples/sorting.py set-up scaled down):

- the live-read checkerboard (oracle.sweep_live): the same schedule, stream
  and decisions, every attempt reading the tables as the earlier attempts of
  its phase left them -- the other end of what the reference's concurrent
  phase can produce.  Adhesion energy, mean area and sorting index agree with
  it to seed precision; the area spread is wider under the snapshot (a
  cell's attempts in one phase all see its phase-start area, so the area
  term's feedback lags by up to a phase) and never narrower;
- the serial random-sequential CPM (oracle.serial_sweep): the live-read
  checkerboard reproduces its area spread, so the engine's wider spread is
  the snapshot's alone.  (Both checkerboard executions sort faster per MCS
  than the serial CPM, about -5% in the sorting index at 1024^2: an effect
  of the reference's colour schedule itself, reported, not asserted.)

Tolerances come from the seed spread alone (stats_common.compare: Welch's t
against its two-sided critical value at alpha = 1e-4 per comparison); the
last tests plant changes in the oracle's parameters to show the comparison
has power.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from stats_common import AREA_STD, compare, observables, sorting_lattice, welch_t

SEEDS = 32
CHECKPOINTS = 4
MCS_PER_CHECKPOINT = 50
L = 128


def _one(mode, seed, c=1, u=1, temperature=None, J12=None):
    lat, n, p, J = sorting_lattice(L=L, seed=seed)
    if J12 is not None:  # (the observables keep the unchanged J)
        Jm = [row[:] for row in J]
        Jm[1][2] = Jm[2][1] = J12
        p = O.Params.make(10, J=Jm, lam_area=[0, 1, 1], target_area=[0, 40, 40])
    if temperature is not None:
        p.temperature = temperature
    tab = O.tally(lat, 2, n)
    traj = []
    for k in range(CHECKPOINTS):
        m0 = k * MCS_PER_CHECKPOINT
        if mode == "serial":
            O.serial_sweep(lat, 2, tab, p, 1000 + seed, m0, MCS_PER_CHECKPOINT)
        elif mode == "live":
            O.sweep_live(lat, 2, tab, p, c, u, 1000 + seed, m0, MCS_PER_CHECKPOINT)
        else:
            O.sweep(lat, 2, tab, p, c, u, 1000 + seed, m0, MCS_PER_CHECKPOINT, 1)
        traj.append(observables(lat, n, p, tab.area, J))
    return traj


def _trajectories(mode, **kw):
    # (the oracle's ctypes calls release the GIL: one thread per seed)
    with ThreadPoolExecutor(8) as ex:
        return np.array(list(ex.map(lambda s: _one(mode, s, **kw), range(SEEDS))))


@pytest.fixture(scope="module")
def live():
    return _trajectories("live")


@pytest.fixture(scope="module")
def snapshot():
    return _trajectories("snapshot")


def _report(rows):
    for r in rows:
        print("t=%d %-14s %.5g vs %.5g t=%.2f (critical %.2f)" % ((r[0] + 1) * MCS_PER_CHECKPOINT, *r[1:]))


@pytest.mark.parametrize("c,u", [(1, 1), (2, 4)])
def test_snapshot_matches_live_reads(live, snapshot, c, u):
    snap = snapshot if (c, u) == (1, 1) else _trajectories("snapshot", c=c, u=u)
    lv = live if (c, u) == (1, 1) else _trajectories("live", c=c, u=u)
    ok, rows = compare(snap, lv, skip=(AREA_STD,))
    _report(rows)
    assert ok, [r for r in rows if abs(r[4]) > r[5]]
    # the area spread: never narrower than under live reads
    for k in range(CHECKPOINTS):
        t, crit = welch_t(snap[:, k, AREA_STD], lv[:, k, AREA_STD])
        assert t > -crit, (k, t)
    # the dynamics did something: cells grew from 9 sites towards 40
    assert snap[:, -1, 1].mean() > 30


def test_snapshot_widens_the_area_spread(live, snapshot):
    """The one systematic effect of the snapshot at c = 1, u = 1: the area
    spread (each seed's mean over the checkpoints) is wider than under live
    reads (about +9% at 64^2 and 128^2, +8% at 256^2 and 1024^2 with this
    set-up)."""
    s, lv = snapshot[:, :, AREA_STD].mean(1), live[:, :, AREA_STD].mean(1)
    t, crit = welch_t(s, lv)
    print("area std snapshot %.4f live %.4f (ratio %.3f) t=%.2f" % (s.mean(), lv.mean(), s.mean() / lv.mean(), t))
    assert t > crit


def test_live_reads_match_serial_area_spread(live):
    """The live-read checkerboard reproduces the serial CPM's area spread and
    adhesion energy: the snapshot alone makes the engine's spread wider."""
    serial = _trajectories("serial")
    ok, rows = compare(live, serial, skip=(1, 3))
    _report(rows)
    assert ok, [r for r in rows if abs(r[4]) > r[5]]
    for k in range(CHECKPOINTS):
        print("t=%d sorting index live %.4f serial %.4f (the schedule's own effect)" % (
            (k + 1) * MCS_PER_CHECKPOINT, live[:, k, 3].mean(), serial[:, k, 3].mean()))


def test_comparison_detects_a_different_temperature(live):
    """The comparison separates T = 10 from T = 30."""
    ok, _ = compare(_trajectories("snapshot", temperature=30.0), live, skip=(AREA_STD,))
    assert not ok


def test_comparison_detects_a_changed_adhesion(live):
    """A planted change of one adhesion entry (J_12 = 11 -> 12, integer as
    the reference's) in the oracle's snapshot run is detected."""
    ok, rows = compare(_trajectories("snapshot", J12=12), live, skip=(AREA_STD,))
    assert not ok
