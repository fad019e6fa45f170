"""Shared test helpers: oracle runs of a workload and GPU-vs-oracle comparisons."""
from __future__ import annotations

import numpy as np

import oracle as O

# Kernel options and launch form applied to every engine the GPU tests make
# through make_cpm (Cpm.set_kernel_options, Cpm.set_launch); set per test by
# the engine_options fixture (tests/conftest.py).  None of them changes
# results.
ENGINE_OPTIONS = {}
_FORMS = {"auto": -1, "phase": 0, "global": 1, "lds": 2, "small": 3, "resident": 4}


def make_cpm(wl):
    """wl.make_cpm() with the current test's kernel options and launch form.
    Kernel options select among the per-phase kernels, so a test that sets
    them without a launch form runs one launch per phase."""
    sim = wl.make_cpm()
    opts = dict(ENGINE_OPTIONS)
    mode, tile = opts.pop("mode", None), int(opts.pop("tile", 0))
    opts = {k: int(v) for k, v in opts.items()}
    if opts:
        sim.set_kernel_options(**opts)
        if mode is None:
            mode = "phase"
    if mode is not None:
        sim.set_launch(_FORMS[mode], tile)
    return sim


def oracle_params(wl) -> "O.Params":
    p = wl.params()
    return O.Params.make(p["temperature"], J=p["J"], S=p["S"], lam_area=p["lam_area"],
                         target_area=p["target_area"], lam_perim=p["lam_perim"],
                         target_perim=p["target_perim"], fixed=p["fixed"], lam_chem=p["lam_chem"],
                         field=p["field"], lam_act=p["lam_act"], act_max=p["act_max"],
                         lam_pers=p["lam_pers"])


def run_oracle(wl, c, u, seed, nmcs, mcs0=0, nthreads=8, lattice=None, inner=None, shared=True,
               ppt=None, rng="philox", cell_sync=False, counts=None, tpb=4):
    """Tally + sweep in the oracle; returns (lattice, Tables).  With a 3D
    persistence term the run is cut into chunks of ``inner`` MCS with a
    preferred-direction update after each (tables.pref holds the result);
    with rng="minstd" each chunk is one reference launch (its streams are
    seeded at the chunk's start; tpb is threads_per_block, which places the
    regions with shared roles and ppt > 2c).  counts (int64[4]) collects the
    attempt counters (counter stream only)."""
    lat = np.ascontiguousarray(wl.lattice if lattice is None else lattice).astype(np.uint32).copy()
    tab = O.tally(lat, wl.d, wl.ncells)
    p = oracle_params(wl)
    act = None
    if getattr(wl, "act", None) is not None:
        act = np.ascontiguousarray(wl.act, np.uint32).copy()
        p.set_act(act)
    pw = wl.params()
    pers = wl.d == 3 and bool(np.any(pw["lam_pers"] != 0))
    if pers:
        ps = O.PrefState(wl.L, wl.d, tab, wl.history, seed)
        p.set_pref(ps.pref_q)
        ctype = np.zeros(wl.ncells + 1, np.int32)
        v = lat.reshape(-1)
        ctype[v & 0xFFFFFF] = (v >> 24).astype(np.int32)
        ctype[0] = 0
        inner = inner or nmcs
        for m in range(0, nmcs, inner):
            O.sweep(lat, wl.d, tab, p, c, u, seed, mcs0 + m, inner, nthreads, shared=shared, ppt=ppt,
                    rng=rng, cell_sync=cell_sync, counts=counts, tpb=tpb)
            ps.update(tab, ctype, pw["pers_diff"])
        tab.pref = ps.pref
    elif nmcs and rng == "minstd":
        inner = inner or nmcs
        for m in range(0, nmcs, inner):
            O.sweep(lat, wl.d, tab, p, c, u, seed, mcs0 + m, inner, nthreads, shared=shared, ppt=ppt,
                    rng=rng, cell_sync=cell_sync, tpb=tpb)
    elif nmcs:
        O.sweep(lat, wl.d, tab, p, c, u, seed, mcs0, nmcs, nthreads, shared=shared, ppt=ppt,
                cell_sync=cell_sync, counts=counts)
    if act is not None:
        tab.act = act
    return lat, tab


def run_kwargs(c, u, iterations, inner=None, **extra):
    """run() kwargs; shared roles unless shared=0 is passed."""
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=2 * c, positions_per_checkerboard=c,
              updates_per_checkerboard_switch=u, updates_per_barrier=1,
              iterations=iterations, inner_iterations=inner or iterations, shared=1)
    kw.update(extra)
    return kw


def assert_tables_equal(gpu_table: dict, tab: "O.Tables", d: int):
    n = tab.ncells + 1
    np.testing.assert_array_equal(gpu_table["area"], tab.area)
    np.testing.assert_array_equal(gpu_table["perimeter"], tab.perim)
    np.testing.assert_array_equal(gpu_table["anchor"].reshape(-1)[: 3 * n], tab.anchor)
    np.testing.assert_array_equal(gpu_table["csum"].reshape(-1)[: 3 * n], tab.csum)


def lattice_digest(lat: np.ndarray) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(lat).view(np.uint32).tobytes()).hexdigest()
