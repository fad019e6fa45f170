"""GPU parity: the HIP sweep through the C ABI vs the CPU oracle, bit-exact.

Every case prints its name, shape and schedule before touching the GPU.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal, run_kwargs, run_oracle, make_cpm

pytestmark = pytest.mark.gpu


def _gpu():
    import gpucpm
    from gpucpm import workloads
    return gpucpm, workloads


def _run_gpu(wl, c, u, seed, nmcs, inner=None, splits=None):
    gpucpm, _ = _gpu()
    sim = make_cpm(wl)
    sim.push_to_gpu()
    sim.set_seed(seed)
    for n in (splits or [nmcs]):
        sim.run(**run_kwargs(c, u, n, inner or n))
    sim.synchronize()
    sim.pull_from_gpu()
    return sim, sim.get_state().view(np.uint32).copy(), sim.get_cell_table()


def _announce(*a):
    print("CASE", *a, flush=True)


def test_tally_after_push_matches_oracle():
    _, W = _gpu()
    for wl in (W.plain2d(L=64, ncells=20, seed=5), W.sorting2d(L=128, grid=9, edge=10, seed=6),
               W.plain3d(L=32, ncells=40, seed=7)):
        _announce("tally", wl.name, wl.L, wl.d)
        sim = make_cpm(wl)
        sim.push_to_gpu()
        _, tab = run_oracle(wl, 1, 1, 0, 0)
        assert_tables_equal(sim.get_cell_table(), tab, wl.d)


@pytest.mark.parametrize("mode", ["phase", "small"])
@pytest.mark.parametrize("c,u", [(1, 1), (2, 1), (2, 4), (4, 3), (4, 16)])
def test_sorting2d_small_bit_exact(c, u, mode, engine_options):
    _, W = _gpu()
    engine_options(mode=mode)
    wl = W.sorting2d(L=128, grid=9, edge=10, seed=11)
    _announce("sorting2d", wl.L, c, u, mode)
    nmcs = 30
    _, lat, table = _run_gpu(wl, c, u, seed=1234, nmcs=nmcs)
    lat_o, tab_o = run_oracle(wl, c, u, 1234, nmcs)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 2)


@pytest.mark.parametrize("c,u", [(1, 1), (2, 8), (2, 3)])
def test_plain3d_small_bit_exact(c, u):
    _, W = _gpu()
    wl = W.plain3d(L=32, ncells=60, seed=12)
    _announce("plain3d", wl.L, c, u)
    nmcs = 10
    _, lat, table = _run_gpu(wl, c, u, seed=99, nmcs=nmcs)
    lat_o, tab_o = run_oracle(wl, c, u, 99, nmcs)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 3)


def test_plain2d_full_config_bit_exact():
    """256^2, 100 cells, 500 MCS -- the full plain 2D configuration."""
    _, W = _gpu()
    wl = W.plain2d()
    _announce("plain2d-full", wl.L, 2, 4)
    _, lat, table = _run_gpu(wl, 2, 4, seed=1, nmcs=500)
    lat_o, tab_o = run_oracle(wl, 2, 4, 1, 500)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 2)


def test_sorting2d_full_size_prefix_bit_exact():
    """1024^2 sorting lattice, first 20 MCS."""
    _, W = _gpu()
    wl = W.sorting2d()
    _announce("sorting2d-full", wl.L, 1, 1)
    _, lat, table = _run_gpu(wl, 1, 1, seed=2, nmcs=20)
    lat_o, tab_o = run_oracle(wl, 1, 1, 2, 20)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 2)


def test_plain3d_full_size_prefix_bit_exact():
    """256^3 lattice with 20000 cubes, first 2 MCS."""
    _, W = _gpu()
    wl = W.plain3d()
    _announce("plain3d-full", wl.L, 1, 1)
    _, lat, table = _run_gpu(wl, 1, 1, seed=3, nmcs=2)
    lat_o, tab_o = run_oracle(wl, 1, 1, 3, 2)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 3)


def test_split_runs_equal_one_run():
    """run(a) + run(b) == run(a + b): the stream is keyed by the MCS counter."""
    _, W = _gpu()
    wl = W.sorting2d(L=128, grid=9, edge=10, seed=13)
    _announce("split", wl.L)
    _, lat1, t1 = _run_gpu(wl, 2, 4, seed=5, nmcs=40)
    _, lat2, t2 = _run_gpu(wl, 2, 4, seed=5, nmcs=40, splits=[15, 25])
    _, lat3, t3 = _run_gpu(wl, 2, 4, seed=5, nmcs=40, inner=8)
    np.testing.assert_array_equal(lat1, lat2)
    np.testing.assert_array_equal(lat1, lat3)
    for k in ("area", "perimeter", "csum"):
        np.testing.assert_array_equal(t1[k], t2[k])
        np.testing.assert_array_equal(t1[k], t3[k])


def test_full_size_tables_equal_recount():
    """Size-independent property at the 1024^2 shape: after 200 MCS the
    incrementally kept tables equal a full recount of the final lattice."""
    import oracle as O
    _, W = _gpu()
    wl = W.sorting2d()
    _announce("recount", wl.L)
    _, lat, table = _run_gpu(wl, 2, 4, seed=77, nmcs=200)
    tab = O.tally(lat, 2, wl.ncells)
    np.testing.assert_array_equal(table["area"], tab.area)
    np.testing.assert_array_equal(table["perimeter"], tab.perim)
    assert table["area"].sum() == np.count_nonzero(lat & 0xFFFFFF)


@pytest.mark.parametrize("mode", ["phase", "global", "lds"])
def test_launch_forms_agree_with_oracle(mode, engine_options):
    """The per-phase, persistent-global and persistent-LDS launch forms all
    reproduce the oracle in 2D; 3D runs per phase."""
    _, W = _gpu()
    engine_options(mode=mode)
    cases = [(W.plain2d(L=64, ncells=20, seed=21), 2, 4, 40), (W.sorting2d(L=128, grid=9, edge=10, seed=22), 1, 1, 30)]
    if mode == "phase":
        # the persistent global form is switched off in 3D (a 256^3 run
        # exceeded its time limit); 3D always launches per phase
        cases.append((W.plain3d(L=32, ncells=60, seed=23), 1, 1, 8))
    for wl, c, u, n in cases:
        _announce("mode", mode, wl.name, wl.L, c, u)
        _, lat, table = _run_gpu(wl, c, u, seed=31, nmcs=n, inner=max(1, n // 3))
        lat_o, tab_o = run_oracle(wl, c, u, 31, (n // max(1, n // 3)) * max(1, n // 3))
        np.testing.assert_array_equal(lat, lat_o)
        assert_tables_equal(table, tab_o, wl.d)


@pytest.mark.skip(reason="the 3D persistent global form is switched off (CPM_ENOTIMPL): a 256^3 run "
                         "of k_persist_global<3> (2048 blocks, the occupancy limit) stopped at a grid "
                         "barrier in an event-timed pass and was killed at its time limit.  The barrier "
                         "protocol is deadlock-free whenever every block is resident "
                         "(tests/test_barrier_model.py), so a block of that grid was not resident; what "
                         "kept it out is not established, so the case stays off the GPU")
def test_persistent_global_3d_matches_oracle(engine_options):
    _, W = _gpu()
    engine_options(mode="global")
    wl = W.plain3d(L=32, ncells=60, seed=23)
    _announce("mode global 3D", wl.L)
    _, lat, table = _run_gpu(wl, 1, 1, seed=31, nmcs=8, inner=2)
    lat_o, tab_o = run_oracle(wl, 1, 1, 31, 8)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 3)


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_lds_tile_size_does_not_change_results(tile, engine_options):
    _, W = _gpu()
    engine_options(mode="lds")
    engine_options(tile=str(tile))
    wl = W.plain2d(L=128, ncells=40, seed=24)
    _announce("tile", tile)
    _, lat, table = _run_gpu(wl, 2, 3, seed=8, nmcs=25)
    lat_o, tab_o = run_oracle(wl, 2, 3, 8, 25)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 2)


@pytest.mark.parametrize("box", ["0", "1"])
def test_3d_box_staging_matches_oracle(box, engine_options):
    """3D with c = 1 reads global memory (box3 = 0, default) or stages
    16x16x8 boxes in LDS (1); both reproduce the oracle."""
    _, W = _gpu()
    engine_options(box3=box)
    wl = W.plain3d(L=64, ncells=300, seed=25)
    _announce("box3", box, wl.L)
    _, lat, table = _run_gpu(wl, 1, 1, seed=17, nmcs=6)
    lat_o, tab_o = run_oracle(wl, 1, 1, 17, 6)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 3)


def test_cell_sync_matches_oracle():
    """cell_sync = 1 (at most one flip per cell per colour phase, lowest
    sub-block wins) in 2D and 3D, bit-exact against the oracle."""
    _, W = _gpu()
    for wl, n in ((W.plain2d(L=64, ncells=20, seed=26), 30), (W.plain3d(L=32, ncells=60, seed=27), 6)):
        _announce("cell_sync", wl.name, wl.L)
        gpucpm, _ = _gpu()
        sim = make_cpm(wl)
        sim.push_to_gpu()
        sim.set_seed(41)
        sim.run(**run_kwargs(1, 1, n, cell_sync=1))
        sim.synchronize()
        sim.pull_from_gpu()
        import oracle as O
        from helpers import oracle_params
        lat_o = wl.lattice.copy()
        tab_o = O.tally(lat_o, wl.d, wl.ncells)
        O.sweep(lat_o, wl.d, tab_o, oracle_params(wl), 1, 1, 41, 0, n, 8, cell_sync=True)
        np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
        assert_tables_equal(sim.get_cell_table(), tab_o, wl.d)


@pytest.mark.parametrize("mode,d,c,u", [("phase", 2, 1, 1), ("phase", 2, 2, 4), ("global", 2, 2, 3),
                                        ("lds", 2, 1, 1), ("phase", 3, 1, 1), ("phase", 3, 2, 8)])
def test_fixed_wall_bit_exact(mode, d, c, u, engine_options):
    """A fixed cell type (Cpm::setFixed, kernels2d.cu:281-290) as a frame
    around the cells: bit-exact against the oracle in every launch form, and
    the wall is unchanged."""
    _, W = _gpu()
    engine_options(mode=mode)
    base = W.sorting2d(L=128, grid=9, edge=10, seed=28) if d == 2 else W.plain3d(L=32, ncells=60, seed=29)
    wl = W.walled(base)
    _announce("fixed", mode, wl.name, wl.L, c, u)
    n = 30 if d == 2 else 8
    _, lat, table = _run_gpu(wl, c, u, seed=51, nmcs=n)
    lat_o, tab_o = run_oracle(wl, c, u, 51, n)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, d)
    wall = (wl.lattice >> 24) == 3
    np.testing.assert_array_equal(lat[wall], wl.lattice[wall])
    assert np.count_nonzero(lat != wl.lattice) > 0


@pytest.mark.parametrize("small,compact,full,att", [
    ("0", "0", "0", "1"), ("1", "0", "0", "1"), ("1", "1", "0", "1"), ("1", "1", "1", "1"),
    ("0", "1", "0", "1"), ("0", "1", "1", "1"), ("0", "1", "0", "2"), ("0", "1", "0", "4"),
    ("1", "1", "0", "2"), ("1", "1", "0", "4"), ("0", "1", "0", "8"), ("1", "1", "0", "8")])
@pytest.mark.parametrize("narrow", ["1", "0"])
def test_phase_kernel_forms_agree_with_oracle(small, compact, full, att, narrow, engine_options):
    """The per-phase kernel forms (parameters staged in LDS or read
    directly; attempts finished in place or queued and finished densely;
    16-bit or 32-bit sites) reproduce the oracle in 2D and 3D, with and
    without a fixed wall."""
    _, W = _gpu()
    engine_options(narrow=narrow)
    engine_options(small=small)
    engine_options(compact=compact)
    engine_options(full=full)
    engine_options(attempts=att)
    cases = [(W.sorting2d(L=128, grid=9, edge=10, seed=61), 1, 1, 30),
             (W.walled(W.sorting2d(L=64, grid=5, edge=10, seed=62)), 1, 1, 30),
             (W.sorting2d(L=128, grid=9, edge=10, seed=63), 2, 4, 20),
             (W.plain3d(L=32, ncells=60, seed=64), 1, 1, 8),
             (W.walled(W.plain3d(L=32, ncells=60, seed=65)), 1, 1, 8)]
    for wl, c, u, n in cases:
        _announce("forms", small, compact, full, att, narrow, wl.name, wl.L, c, u)
        _, lat, table = _run_gpu(wl, c, u, seed=71, nmcs=n)
        lat_o, tab_o = run_oracle(wl, c, u, 71, n)
        np.testing.assert_array_equal(lat, lat_o)
        assert_tables_equal(table, tab_o, wl.d)


@pytest.mark.parametrize("small,compact,att", [("0", "1", "1"), ("1", "1", "4"), ("0", "0", "1"),
                                               ("1", "0", "1")])
def test_chemotaxis_bit_exact(small, compact, att, engine_options):
    """The chemotaxis term (Cpm::setChemotaxis, kernels2d.cu:325-334,
    kernels3d.cu:314-325) with a smooth field, in 2D and 3D, every kernel
    form, bit-exact against the oracle."""
    _, W = _gpu()
    engine_options(small=small)
    engine_options(compact=compact)
    engine_options(attempts=att)
    cases = [(W.chemotactic(W.sorting2d(L=128, grid=9, edge=10, seed=81)), 1, 1, 30),
             (W.chemotactic(W.sorting2d(L=64, grid=5, edge=10, seed=82)), 2, 4, 20),
             (W.chemotactic(W.walled(W.sorting2d(L=64, grid=5, edge=10, seed=83))), 1, 1, 20),
             (W.chemotactic(W.plain3d(L=32, ncells=60, seed=84), lam=(0.0, 15.0)), 1, 1, 8)]
    for wl, c, u, n in cases:
        _announce("chemo", small, compact, att, wl.name, wl.L, c, u)
        sim, lat, table = _run_gpu(wl, c, u, seed=91, nmcs=n)
        lat_o, tab_o = run_oracle(wl, c, u, 91, n)
        np.testing.assert_array_equal(lat, lat_o)
        assert_tables_equal(table, tab_o, wl.d)
        # pull restores the pushed field
        np.testing.assert_array_equal(sim.get_field(), wl.field)


def test_chemotaxis_needs_a_field():
    """A chemotaxis lambda on an engine made with has_field = 0 is refused at
    run (the reference would read an unallocated field)."""
    gpucpm, W = _gpu()
    wl = W.sorting2d(L=64, grid=5, edge=10, seed=85)
    sim = make_cpm(wl)
    sim.set_constraints(1, lambda_chemotaxis=2.0)
    sim.push_to_gpu()
    with pytest.raises(ValueError):
        sim.run(**run_kwargs(1, 1, 1))


@pytest.mark.parametrize("small,compact,att,mode", [("0", "1", "1", "phase"), ("1", "1", "4", "phase"),
                                                    ("0", "0", "1", "phase"), ("1", "0", "1", "phase"),
                                                    ("0", "1", "1", "global"), ("0", "1", "1", "lds")])
def test_act_bit_exact(small, compact, att, mode, engine_options):
    """The act term (Cpm::setAct, kernels2d.cu:312-343,386-397,417-473,
    532-534) in 2D and 3D, every kernel and launch form, bit-exact against
    the oracle, act lattice included; combined with chemotaxis and a fixed
    wall."""
    _, W = _gpu()
    engine_options(small=small)
    engine_options(compact=compact)
    engine_options(attempts=att)
    engine_options(mode=mode)
    cases = [(W.active(W.sorting2d(L=128, grid=9, edge=10, seed=101)), 2, 1, 30),
             (W.active(W.sorting2d(L=64, grid=5, edge=10, seed=102)), 2, 4, 20),
             (W.active(W.chemotactic(W.walled(W.sorting2d(L=64, grid=5, edge=10, seed=103)))), 4, 3, 20)]
    if mode == "phase":
        cases.append((W.active(W.plain3d(L=32, ncells=60, seed=104), lam=(0.0, 200.0),
                               act_max=(0, 30)), 2, 1, 6))
    for wl, c, u, n in cases:
        _announce("act", small, compact, att, mode, wl.name, wl.L, c, u)
        sim, lat, table = _run_gpu(wl, c, u, seed=111, nmcs=n)
        lat_o, tab_o = run_oracle(wl, c, u, 111, n)
        np.testing.assert_array_equal(lat, lat_o)
        assert_tables_equal(table, tab_o, wl.d)
        np.testing.assert_array_equal(sim.get_act_state().view(np.uint32), tab_o.act)


@pytest.mark.parametrize("narrow", ["1", "0"])
def test_act_single_site_checkerboard_bit_exact(narrow, engine_options):
    """With shared roles and positions_per_checkerboard = 1 the act term's
    source neighbourhood reaches same-phase targets: each phase is decided
    against its start lattice and then applied (two launches), bit-exact
    against the oracle's two passes in 2D and 3D (also with the reference
    stream, and with cell_sync: the decide launch also claims, the apply
    launch checks the claims).  Batches: test_gpu_batch.py."""
    _, W = _gpu()
    engine_options(narrow=narrow)
    cases = [(W.active(W.sorting2d(L=128, grid=9, edge=10, seed=105)), 30, "philox"),
             (W.active(W.chemotactic(W.walled(W.sorting2d(L=64, grid=5, edge=10, seed=106)))), 20,
              "philox"),
             (W.active(W.plain3d(L=32, ncells=60, seed=107), lam=(0.0, 200.0), act_max=(0, 30)), 6,
              "philox"),
             (W.active(W.sorting2d(L=64, grid=5, edge=10, seed=108)), 12, "minstd")]
    cases = [(wl, n, rng, False) for wl, n, rng in cases] + [
        (W.active(W.sorting2d(L=128, grid=9, edge=10, seed=110)), 20, "philox", True),
        (W.active(W.plain3d(L=32, ncells=60, seed=111), lam=(0.0, 200.0), act_max=(0, 30)), 5,
         "philox", True),
        (W.active(W.sorting2d(L=64, grid=5, edge=10, seed=112)), 12, "minstd", True)]
    for wl, n, rng, cs in cases:
        _announce("act c=1", narrow, rng, "cell_sync" if cs else "", wl.name, wl.L, wl.d)
        sim = make_cpm(wl)
        sim.push_to_gpu()
        sim.set_seed(113)
        sim.set_rng(rng)
        sim.run(**run_kwargs(1, 1, n, n // 2 if rng == "minstd" else n, positions_per_thread=2,
                             cell_sync=int(cs)))
        sim.synchronize()
        sim.pull_from_gpu()
        lat = sim.get_state().view(np.uint32).copy()
        lat_o, tab_o = run_oracle(wl, 1, 1, 113, n, inner=n // 2 if rng == "minstd" else None,
                                  rng=rng, cell_sync=cs)
        np.testing.assert_array_equal(lat, lat_o)
        assert_tables_equal(sim.get_cell_table(), tab_o, wl.d)
        np.testing.assert_array_equal(sim.get_act_state().view(np.uint32), tab_o.act)
        assert np.count_nonzero(lat != wl.lattice) > 0


@pytest.mark.parametrize("small,compact,att", [("1", "1", "4"), ("0", "1", "1"), ("1", "0", "1")])
def test_persistence_bit_exact(small, compact, att, engine_options):
    """The 3D persistence term (Cpm::setPersistence, kernels3d.cu:327-344)
    with preferred directions updated from the centroid history after every
    inner_iterations chunk (kernels.cu:158-212, cpm.cu:487-514): lattice,
    tables and preferred directions bit-exact against the oracle."""
    _, W = _gpu()
    engine_options(small=small)
    engine_options(compact=compact)
    engine_options(attempts=att)
    cases = [(W.persistent(W.plain3d(L=32, ncells=60, seed=121), lam=(0.0, 300.0), diffusion=(0.0, 0.7)),
              1, 1, 12, 2),
             (W.persistent(W.plain3d(L=32, ncells=40, seed=122), lam=(0.0, 150.0), diffusion=(0.0, 0.0),
                           history=1), 2, 4, 6, 3),
             (W.persistent(W.chemotactic(W.plain3d(L=32, ncells=40, seed=123), lam=(0.0, 10.0)),
                           lam=(0.0, 200.0), diffusion=(0.0, 0.5), history=4), 1, 1, 12, 3)]
    for wl, c, u, n, inner in cases:
        _announce("pers", small, compact, att, wl.name, wl.L, c, u, inner)
        sim, lat, table = _run_gpu(wl, c, u, seed=131, nmcs=n, inner=inner)
        lat_o, tab_o = run_oracle(wl, c, u, 131, n, inner=inner)
        np.testing.assert_array_equal(lat, lat_o)
        assert_tables_equal(table, tab_o, wl.d)
        np.testing.assert_array_equal(sim.get_pref_dirs().reshape(-1), tab_o.pref)


def test_persistence_has_no_effect_in_2d():
    """The reference has no 2D persistence term (kernels2d.cu has none and
    the prefDir update runs for dimensionality 3 only, cpm.cu:487)."""
    _, W = _gpu()
    base = W.sorting2d(L=64, grid=5, edge=10, seed=124)
    wl = W.persistent(base)
    _announce("pers2d")
    _, lat_a, _ = _run_gpu(base, 1, 1, seed=5, nmcs=10)
    _, lat_b, _ = _run_gpu(wl, 1, 1, seed=5, nmcs=10)
    np.testing.assert_array_equal(lat_a, lat_b)


@pytest.mark.parametrize("compact,att", [("1", "1"), ("1", "4"), ("0", "1")])
def test_nonshared_roles_bit_exact(compact, att, engine_options):
    """run(shared=0), the reference's default (kernels2d.cu:227-243): the
    source drawn in the sub-block, the target its neighbour, K = ppt / c
    colours and groups that keep concurrent attempts apart; bit-exact against
    the oracle in 2D and 3D, with every Hamiltonian term."""
    _, W = _gpu()
    engine_options(compact=compact)
    engine_options(attempts=att)
    cases = [(W.sorting2d(L=128, grid=9, edge=10, seed=141), 1, 2, 1, 20),
             (W.sorting2d(L=128, grid=9, edge=10, seed=142), 2, 4, 1, 10),
             (W.sorting2d(L=64, grid=5, edge=10, seed=143), 2, 4, 4, 10),
             (W.sorting2d(L=128, grid=9, edge=10, seed=144), 4, 8, 1, 6),
             (W.sorting2d(L=64, grid=5, edge=10, seed=145), 1, 8, 1, 10),
             (W.active(W.chemotactic(W.walled(W.sorting2d(L=64, grid=5, edge=10, seed=146)))), 1, 2, 1, 10),
             (W.plain3d(L=32, ncells=60, seed=147), 1, 2, 1, 4),
             (W.active(W.plain3d(L=32, ncells=40, seed=148), lam=(0.0, 200.0), act_max=(0, 30)), 2, 4, 1, 3)]
    for wl, c, ppt, u, n in cases:
        _announce("nonshared", compact, att, wl.name, wl.L, c, ppt, u)
        gpucpm, _ = _gpu()
        sim = make_cpm(wl)
        sim.push_to_gpu()
        sim.set_seed(151)
        sim.run(**run_kwargs(c, u, n, shared=0, positions_per_thread=ppt))
        sim.synchronize()
        sim.pull_from_gpu()
        lat_o, tab_o = run_oracle(wl, c, u, 151, n, shared=False, ppt=ppt)
        np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
        assert_tables_equal(sim.get_cell_table(), tab_o, wl.d)


def test_nonshared_cell_sync_and_batch():
    """Non-shared roles with cell_sync, and in a batch, against the oracle."""
    gpucpm, W = _gpu()
    import oracle as O
    from helpers import oracle_params
    wl = W.sorting2d(L=64, grid=5, edge=10, seed=152)
    _announce("nonshared cell_sync")
    sim = make_cpm(wl)
    sim.push_to_gpu()
    sim.set_seed(3)
    sim.run(**run_kwargs(1, 1, 10, cell_sync=1, shared=0, positions_per_thread=2))
    sim.synchronize()
    sim.pull_from_gpu()
    lat_o = wl.lattice.copy()
    tab_o = O.tally(lat_o, 2, wl.ncells)
    O.sweep(lat_o, 2, tab_o, oracle_params(wl), 1, 1, 3, 0, 10, 8, cell_sync=True, shared=False, ppt=2)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    _announce("nonshared batch")
    wls = W.batched_sorting(nrep=4, seed=153, sizes=(64, 128))
    sims = []
    for i, w in enumerate(wls):
        s = w.make_cpm()
        s.push_to_gpu()
        s.set_seed(20 + i)
        sims.append(s)
    gpucpm.Batch(sims).run(**run_kwargs(2, 1, 8, shared=0, positions_per_thread=4))
    for i, (w, s) in enumerate(zip(wls, sims)):
        s.synchronize()
        s.pull_from_gpu()
        lat_o, tab_o = run_oracle(w, 2, 1, 20 + i, 8, shared=False, ppt=4)
        np.testing.assert_array_equal(s.get_state().view(np.uint32), lat_o)
        assert_tables_equal(s.get_cell_table(), tab_o, 2)


@pytest.mark.parametrize("d,c,u,shared", [(2, 2, 4, 1), (2, 2, 3, 0), (3, 2, 8, 1)])
def test_cell_sync_several_updates_matches_oracle(d, c, u, shared):
    """cell_sync with u > 1: a claim and a commit launch per attempt index
    (at most one flip per cell per index), bit-exact against the oracle."""
    gpucpm, W = _gpu()
    import oracle as O
    from helpers import oracle_params
    wl = W.sorting2d(L=64, grid=5, edge=10, seed=161) if d == 2 else W.plain3d(L=32, ncells=60, seed=162)
    _announce("cell_sync u", d, c, u, shared)
    sim = make_cpm(wl)
    sim.push_to_gpu()
    sim.set_seed(17)
    n = 6 if d == 2 else 2
    sim.run(**run_kwargs(c, u, n, cell_sync=1, shared=shared, positions_per_thread=2 * c))
    sim.synchronize()
    sim.pull_from_gpu()
    lat_o = wl.lattice.copy()
    tab_o = O.tally(lat_o, d, wl.ncells)
    O.sweep(lat_o, d, tab_o, oracle_params(wl), c, u, 17, 0, n, 8, cell_sync=True, shared=bool(shared),
            ppt=2 * c)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, d)
