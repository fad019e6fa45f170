"""gpucpm -- MI355X Cellular Potts engine, drop-in for the reference module.

``gpucpm.Cpm`` keeps the reference's Python surface (shabaz/gpu-cpm
src/python_wrapper.cpp:27-380): constructor, ``set_constraints``,
``add_cell``, ``push_to_gpu``, ``run``, ``synchronize``, ``pull_from_gpu``,
``get_state``, ``get_centroids`` -- same names, positional/keyword rules,
defaults and dtypes.  Every call goes through the C ABI of libgpucpm.so
(include/gpucpm.h), whose sweep runs as HIP kernels on the GPU.

Extensions: ``run(..., seed=...)``, ``set_seed``, ``seed``, ``mcs``,
``get_cell_table``.
"""
from __future__ import annotations

import ctypes
import operator

import numpy as np

from . import _lib
from ._lib import RunArgs, check, lib

__all__ = ["Cpm", "Batch", "RunArgs"]

_UNSET = -1


def _int(v, name):
    """PyArg 'i' conversion: integers (anything with __index__) only."""
    try:
        return operator.index(v)
    except TypeError:
        raise TypeError(f"{name}: an integer is required (got {type(v).__name__})") from None


def _float(v, name):
    """PyArg 'd' conversion."""
    try:
        return float(v)
    except (TypeError, ValueError):
        raise TypeError(f"{name}: must be real number, not {type(v).__name__}") from None


def _run_args(cell_sync, block_sync, global_sync, threads_per_block, positions_per_thread,
              positions_per_checkerboard, updates_per_checkerboard_switch, updates_per_barrier,
              iterations, inner_iterations, shared=0, partial_dispatch=0, centroid_tracking=False):
    a = RunArgs()
    vals = dict(cell_sync=cell_sync, block_sync=block_sync, global_sync=global_sync,
                threads_per_block=threads_per_block, positions_per_thread=positions_per_thread,
                positions_per_checkerboard=positions_per_checkerboard,
                updates_per_checkerboard_switch=updates_per_checkerboard_switch,
                updates_per_barrier=updates_per_barrier, iterations=iterations,
                inner_iterations=inner_iterations, shared=shared, partial_dispatch=partial_dispatch)
    for k, v in vals.items():
        setattr(a, k, _int(v, k))
    a.centroid_tracking = int(bool(centroid_tracking))
    return a


class Cpm:
    """One Cellular Potts simulation (reference: class Cpm, src/cpm.h:78-112).

    ``Cpm(dimension, dimensionality, number_of_cells, temperature, has_field
    [, history=1, has_act=True])`` -- positional only, as PyCpm_init parses
    "iiiip|ip" (python_wrapper.cpp:35).
    """

    def __init__(self, dimension, dimensionality, number_of_cells, temperature, has_field,
                 history=1, has_act=True, /):
        self._h = None
        self.dimension = _int(dimension, "dimension")
        self.dimensionality = _int(dimensionality, "dimensionality")
        self.number_of_cells = _int(number_of_cells, "number_of_cells")
        temperature = _int(temperature, "temperature")
        self._has_field = bool(has_field)
        history = _int(history, "history")
        self._has_act = bool(has_act)
        h = ctypes.c_void_p()
        check(lib.cpm_create(self.dimension, self.dimensionality, self.number_of_cells,
                             temperature, int(self._has_field), history, int(self._has_act),
                             ctypes.byref(h)))
        self._h = h
        self._field = None
        if self._has_field:
            # engine-owned host field (cpm.cu:189-190), viewed in place
            p = ctypes.POINTER(ctypes.c_double)()
            check(lib.cpm_host_field(self._h, ctypes.byref(p)))
            shape = (self.dimensionality,) + (self.dimension,) * self.dimensionality
            n = int(np.prod(shape))
            buf = (ctypes.c_double * n).from_address(ctypes.addressof(p.contents))
            buf._owner = self
            self._field = np.ctypeslib.as_array(buf).reshape(shape)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and h.value and lib is not None:
            lib.cpm_destroy(h)

    # -------------------------------------------------------------- setup --
    def set_constraints(self, cell_type, *, other_cell_type=_UNSET, lambda_perimeter=-1.0,
                        target_perimeter=_UNSET, lambda_area=-1.0, target_area=_UNSET,
                        lambda_act=-1.0, max_act=_UNSET, adhesion=_UNSET, fixed=_UNSET,
                        lambda_chemotaxis=-1.0, lambda_persistence=-1.0,
                        persistence_diffusion=-1.0, persistence_time=_UNSET,
                        set_adhesion_symmetric=1, stickiness=_UNSET):
        """Set Hamiltonian constraints of one cell type (python_wrapper.cpp:55-129).

        -1 means "not given"; area, perimeter, act, adhesion and stickiness
        are applied only when both halves of their pair are given.  Adhesion
        is always written symmetrically: the reference parses
        ``set_adhesion_symmetric`` as an int into a double that is never 0
        (python_wrapper.cpp:81-82), so J[a][b] and J[b][a] are both set.
        """
        ct = _int(cell_type, "cell_type")
        oct_ = _int(other_cell_type, "other_cell_type")
        lp, tp = _float(lambda_perimeter, "lambda_perimeter"), _int(target_perimeter, "target_perimeter")
        la, ta = _float(lambda_area, "lambda_area"), _int(target_area, "target_area")
        lact, mact = _float(lambda_act, "lambda_act"), _int(max_act, "max_act")
        adh, fx = _int(adhesion, "adhesion"), _int(fixed, "fixed")
        lchem = _float(lambda_chemotaxis, "lambda_chemotaxis")
        lpers = _float(lambda_persistence, "lambda_persistence")
        pdiff = _float(persistence_diffusion, "persistence_diffusion")
        ptime = _int(persistence_time, "persistence_time")
        _int(set_adhesion_symmetric, "set_adhesion_symmetric")
        stick = _int(stickiness, "stickiness")

        # Terms outside the implemented Hamiltonian are refused, not ignored.

        if fx >= 0:
            check(lib.cpm_set_fixed(self._h, ct, fx))
        if lpers != -1 and pdiff != -1 and ptime != -1:
            check(lib.cpm_set_persistence(self._h, ct, ptime, pdiff, lpers))
        if lchem != -1:
            check(lib.cpm_set_chemotaxis(self._h, ct, lchem))
        if lp != -1 and tp != -1:
            check(lib.cpm_set_perimeter(self._h, ct, tp, lp))
        if la != -1 and ta != -1:
            check(lib.cpm_set_area(self._h, ct, ta, la))
        if lact != -1 and mact != -1:
            check(lib.cpm_set_act(self._h, ct, mact, lact))
        if oct_ != -1 and adh != -1:
            check(lib.cpm_set_adhesion(self._h, ct, oct_, adh, 1))
        if oct_ != -1 and stick != -1:
            check(lib.cpm_set_stickiness(self._h, ct, oct_, stick))
        return None

    def add_cell(self, type, x, y, z=-1, /):
        """Place a one-site cell of ``type`` at (x, y[, z]) in the host lattice
        (python_wrapper.cpp:131-148, cpm.cu:160-174)."""
        t, x, y, z = _int(type, "type"), _int(x, "x"), _int(y, "y"), _int(z, "z")
        if self.dimensionality == 3 and z == -1:
            z = 0  # the 2D form indexes y*L + x, i.e. the z = 0 plane
        if self.dimensionality == 2:
            if z not in (-1, 0):
                raise ValueError("add_cell: z given for a 2D lattice")
            z = -1
        check(lib.cpm_add_cell(self._h, t, x, y, z))
        return None

    # ------------------------------------------------------------ transfer --
    def push_to_gpu(self):
        """Copy the host lattice to the GPU and rebuild the cell tables
        (cpm.cu:541-597)."""
        check(lib.cpm_push(self._h))

    def pull_from_gpu(self):
        """Copy the GPU lattice into the host lattice (cpm.cu:531-540)."""
        check(lib.cpm_pull(self._h))

    def synchronize(self):
        """Wait for queued GPU work (cpm.cu:653-655)."""
        check(lib.cpm_sync(self._h))

    # ----------------------------------------------------------------- run --
    def run(self, cell_sync, block_sync, global_sync, threads_per_block, positions_per_thread,
            positions_per_checkerboard, updates_per_checkerboard_switch, updates_per_barrier,
            iterations, inner_iterations, shared=0, partial_dispatch=0, centroid_tracking=False,
            *, seed=None):
        """Run (iterations // inner_iterations) * inner_iterations MCS
        (python_wrapper.cpp:182-227, cpm.cu:291-517).

        The checkerboard sub-block edge is ``positions_per_checkerboard`` and
        each sub-block makes ``updates_per_checkerboard_switch`` attempts per
        colour phase.  The other schedule arguments choose a thread mapping
        in the reference and do not change results here.  ``seed`` (an
        extension) sets the RNG key before running.
        """
        a = _run_args(cell_sync, block_sync, global_sync, threads_per_block, positions_per_thread,
                      positions_per_checkerboard, updates_per_checkerboard_switch,
                      updates_per_barrier, iterations, inner_iterations, shared, partial_dispatch,
                      centroid_tracking)
        if seed is not None:
            self.set_seed(seed)
        check(lib.cpm_run(self._h, ctypes.byref(a)))
        return None

    # ------------------------------------------------------------ readback --
    def get_state(self):
        """Writable int32 view of the host lattice, shape (L,)*d, indexed
        [y][x] / [z][y][x] (python_wrapper.cpp:292-308).  The view keeps this
        Cpm alive."""
        p = ctypes.POINTER(ctypes.c_uint32)()
        check(lib.cpm_host_lattice(self._h, ctypes.byref(p)))
        n = self.dimension ** self.dimensionality
        buf = (ctypes.c_int32 * n).from_address(ctypes.addressof(p.contents))
        buf._owner = self
        return np.ctypeslib.as_array(buf).reshape((self.dimension,) * self.dimensionality)

    def get_centroids(self):
        """float64 (active_cells, d) centroids [x, y(, z)] of the cells added
        with add_cell (python_wrapper.cpp:151-177, cpm.cu:68-110)."""
        n = ctypes.c_int()
        check(lib.cpm_active_cells(self._h, ctypes.byref(n)))
        out = np.zeros((n.value, self.dimensionality), np.float64)
        rows = ctypes.c_int()
        check(lib.cpm_get_centroids(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    n.value, ctypes.byref(rows)))
        return out

    def get_field(self):
        """float64 (d, L, ...) view of the host field (python_wrapper.cpp:274-290):
        fill it, then push_to_gpu; the chemotaxis term reads it at the source
        site of each copy.  The view keeps this Cpm alive."""
        if self._field is None:
            raise ValueError("get_field: this Cpm was created with has_field=False")
        return self._field

    def get_pref_dirs(self):
        """Extension: float64 (number_of_cells + 1, 3) preferred directions of
        the persistence term (row 0 unused); drawn from the seed, then
        updated after every inner_iterations chunk (3D)."""
        out = np.zeros((self.number_of_cells + 1, 3), np.float64)
        check(lib.cpm_get_pref_dirs(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set_pref_dirs(self, dirs):
        """Extension: set the preferred directions ((number_of_cells + 1, 3))."""
        a = np.ascontiguousarray(dirs, np.float64).reshape(self.number_of_cells + 1, 3)
        check(lib.cpm_set_pref_dirs(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def get_act_state(self):
        """int32 view of the host act lattice (python_wrapper.cpp:312-328):
        the MCS of the last copy into each site, pushed and pulled with the
        lattice.  The view keeps this Cpm alive."""
        p = ctypes.POINTER(ctypes.c_uint32)()
        check(lib.cpm_host_act(self._h, ctypes.byref(p)))
        n = self.dimension ** self.dimensionality
        buf = (ctypes.c_int32 * n).from_address(ctypes.addressof(p.contents))
        buf._owner = self
        return np.ctypeslib.as_array(buf).reshape((self.dimension,) * self.dimensionality)

    # ---------------------------------------------------------- extensions --
    def get_cell_table(self):
        """All n+1 rows of the cell tables: area, perimeter (Moore/26 count
        of unlike neighbours), anchor and csum (integer centroid state) and
        centroid (float64 [x, y(, z)], NaN for empty rows)."""
        rows = self.number_of_cells + 1
        area = np.zeros(rows, np.int32)
        perim = np.zeros(rows, np.int32)
        anchor = np.zeros(rows * 3, np.int32)
        csum = np.zeros(rows * 3, np.int64)
        P = ctypes.POINTER
        check(lib.cpm_get_cell_table(self._h, area.ctypes.data_as(P(ctypes.c_int32)),
                                     perim.ctypes.data_as(P(ctypes.c_int32)),
                                     anchor.ctypes.data_as(P(ctypes.c_int32)),
                                     csum.ctypes.data_as(P(ctypes.c_int64))))
        d, L = self.dimensionality, self.dimension
        cen = np.full((rows, d), np.nan)
        alive = area > 0
        for ax in range(d):
            v = anchor[ax::3][alive] + csum[ax::3][alive] / area[alive]
            v = np.where(v < 0, v + L, v)
            v = np.where(v >= L, v - L, v)
            cen[alive, ax] = v
        return {"area": area, "perimeter": perim, "anchor": anchor.reshape(rows, 3),
                "csum": csum.reshape(rows, 3), "centroid": cen}

    def energy_stats(self):
        """Sufficient statistics of H on the device lattice (exact integers):
        pairs (20, 20) int64, upper triangle: unlike-cell neighbour pairs by
        type pair; moments (20, 5) int64: per type count, sum A, sum A^2,
        sum P, sum P^2 over live cells."""
        pairs = np.zeros((20, 20), np.int64)
        mom = np.zeros((20, 5), np.int64)
        P = ctypes.POINTER(ctypes.c_int64)
        check(lib.cpm_energy_stats(self._h, pairs.ctypes.data_as(P), mom.ctypes.data_as(P)))
        return pairs, mom

    def energy_grad(self, theta, ntypes):
        """(H, dH/dtheta) in float64 on the current lattice, theta laid out as
        gpucpm.autograd.ThetaLayout(ntypes) (the C entry cpm_energy_grad)."""
        th = np.ascontiguousarray(np.asarray(theta, np.float64).reshape(-1))
        nt = int(ntypes)
        if th.size != nt * (nt + 1) // 2 + 4 * nt:
            raise ValueError("theta has the wrong length for ntypes")
        g = np.zeros_like(th)
        H = ctypes.c_double()
        D = ctypes.POINTER(ctypes.c_double)
        check(lib.cpm_energy_grad(self._h, nt, th.ctypes.data_as(D), ctypes.byref(H), g.ctypes.data_as(D)))
        return H.value, g

    def set_adhesion_real(self, cell_type, other_cell_type, adhesion, symmetric=True):
        """Adhesion as a multiple of 2^-8 (extension used by parameter fitting)."""
        check(lib.cpm_set_adhesion_real(self._h, _int(cell_type, "cell_type"),
                                        _int(other_cell_type, "other_cell_type"),
                                        float(adhesion), int(bool(symmetric))))

    def timing_begin(self):
        """Start bracketing every sweep launch with HIP events."""
        check(lib.cpm_timing_begin(self._h))

    def timing_end(self):
        """Synchronize; return (kernel_ms, launches, site_updates) since timing_begin."""
        ms, n, upd = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        check(lib.cpm_timing_end(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(upd)))
        return ms.value, n.value, upd.value

    def set_launch(self, form=-1, tile=0):
        """Launch form (-1 auto, 0 per phase, 1 persistent global, 2 persistent
        LDS tiles, 3 one workgroup with the lattice in LDS, 4 resident 2D
        tiles with neighbour-only waits) and LDS tile edge (0 auto).  Results
        never depend on them."""
        check(lib.cpm_set_launch(self._h, _int(form, "form"), _int(tile, "tile")))

    def set_kernel_options(self, small=None, compact=None, full=None, attempts=None,
                           narrow=None, box3=None):
        """Options of the per-phase kernels (None keeps the default): small
        (scalar parameter loads, at most four types), compact (queue only
        attempts whose sites differ, u = 1), full (queue whole
        neighbourhoods), attempts (drawn per thread: 1, 2, 4, 8), narrow
        (16-bit sites when they fit), box3 (3D, c = 1: LDS boxes).  Results
        never depend on them."""
        v = [-1 if x is None else int(x) for x in (small, compact, full, attempts, narrow, box3)]
        check(lib.cpm_set_kernel_options(self._h, *v))

    @property
    def resident_fallbacks(self):
        """(guard, waits): chunks of resident-form runs (launch form 4) redone
        with one launch per phase because a cell spread over more than two
        tiles per axis, or because a neighbour wait gave up."""
        g, w = ctypes.c_int64(), ctypes.c_int64()
        check(lib.cpm_resident_fallbacks(self._h, ctypes.byref(g), ctypes.byref(w)))
        return g.value, w.value

    def last_launch(self):
        """Launch form of the last run: dict(form, kernel, blocks, tile)."""
        f, b, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib.cpm_last_launch(self._h, ctypes.byref(f), ctypes.byref(b), ctypes.byref(t)))
        kernel = lib.cpm_last_kernel(self._h).decode()
        return {"form": f.value, "kernel": kernel, "blocks": b.value, "tile": t.value}

    def set_autotune(self, on=True):
        """Time each candidate kernel form on this workload's own launches and
        keep the fastest (all forms compute the same results)."""
        check(lib.cpm_set_autotune(self._h, int(bool(on))))

    def set_counters(self, on=True):
        """Count attempts in the sweep (the reference's analytics block,
        cpm.cu:215, kernels2d.cu:292,526,531); switching on zeroes the counts."""
        check(lib.cpm_set_counters(self._h, int(bool(on))))

    def counters(self, reset=False):
        """dict of attempts drawn, differing (source != target), evaluated
        (after the fixed-type rules), accepted and committed since the
        counters were switched on or last reset."""
        out = (ctypes.c_int64 * 5)()
        check(lib.cpm_get_counters(self._h, out, int(bool(reset))))
        return dict(zip(("attempts", "differing", "evaluated", "accepted", "committed"), out))

    _RNG = {"philox": 0, "minstd": 1}

    def set_rng(self, name):
        """Random stream: "philox" (default; counter-based, keyed by sub-block,
        MCS, phase and attempt) or "minstd", the reference's per-thread
        stream (random.cu, kernels2d.cu:99): one MINSTD state per reference
        thread of positions_per_thread sites, reseeded every inner_iterations
        chunk from the thread's first site index + the low 32 bits of the
        seed, for comparisons with the reference.  With shared=1 and
        positions_per_thread > 2 * positions_per_checkerboard the threads sit
        in blocks of threads_per_block per axis and visit their regions one
        block slice at a time, as the reference's do; a Batch's engines must
        all use the same stream."""
        if name not in self._RNG:
            raise ValueError("rng must be 'philox' or 'minstd'")
        check(lib.cpm_set_rng(self._h, self._RNG[name]))

    @property
    def rng(self):
        v = ctypes.c_int()
        check(lib.cpm_get_rng(self._h, ctypes.byref(v)))
        return {0: "philox", 1: "minstd"}[v.value]

    def set_seed(self, seed):
        check(lib.cpm_set_seed(self._h, _int(seed, "seed") & (2**64 - 1)))

    @property
    def seed(self):
        v = ctypes.c_uint64()
        check(lib.cpm_get_seed(self._h, ctypes.byref(v)))
        return v.value

    @property
    def mcs(self):
        v = ctypes.c_uint64()
        check(lib.cpm_get_mcs(self._h, ctypes.byref(v)))
        return v.value

    @mcs.setter
    def mcs(self, value):
        check(lib.cpm_set_mcs(self._h, _int(value, "mcs")))


def energy_stats_many(sims):
    """Cpm.energy_stats of several engines at once (one statistics launch
    for all of them): pairs (n, 20, 20) and moments (n, 20, 5), int64."""
    sims = list(sims)
    n = len(sims)
    pairs = np.zeros((n, 20, 20), np.int64)
    mom = np.zeros((n, 20, 5), np.int64)
    hs = (ctypes.c_void_p * max(n, 1))(*[s._h for s in sims])
    P = ctypes.POINTER(ctypes.c_int64)
    check(lib.cpm_energy_stats_many(ctypes.cast(hs, ctypes.POINTER(ctypes.c_void_p)), n,
                                    pairs.ctypes.data_as(P), mom.ctypes.data_as(P)))
    return pairs, mom


def energy_stats_many_noalloc(sims):
    """energy_stats_many without any device allocation (for callers that
    cannot allocate): at most 64 engines; overlapping calls on one GPU race."""
    sims = list(sims)
    n = len(sims)
    pairs = np.zeros((n, 20, 20), np.int64)
    mom = np.zeros((n, 20, 5), np.int64)
    hs = (ctypes.c_void_p * max(n, 1))(*[s._h for s in sims])
    P = ctypes.POINTER(ctypes.c_int64)
    check(lib.cpm_energy_stats_many_noalloc(ctypes.cast(hs, ctypes.POINTER(ctypes.c_void_p)), n,
                                            pairs.ctypes.data_as(P), mom.ctypes.data_as(P)))
    return pairs, mom


def energy_stats_kernel_ms():
    """Kernel time (ms) of the last energy_stats / energy_stats_many call:
    one statistics launch covers every engine of a call."""
    v = ctypes.c_double()
    check(lib.cpm_energy_stats_timing(ctypes.byref(v)))
    return v.value


class Batch:
    """Batched replicas: several pushed Cpm engines (same dimensionality; any
    sizes, parameters, temperatures and seeds) advanced together, by default
    with one launch per colour phase covering every replica.  Each replica
    ends bit-identical to calling its own ``run`` with the same arguments."""

    def __init__(self, sims):
        self.sims = list(sims)
        arr = (ctypes.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        h = ctypes.c_void_p()
        check(lib.cpm_batch_create(arr, len(self.sims), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and h.value and lib is not None:
            lib.cpm_batch_destroy(h)

    def run(self, cell_sync, block_sync, global_sync, threads_per_block, positions_per_thread,
            positions_per_checkerboard, updates_per_checkerboard_switch, updates_per_barrier,
            iterations, inner_iterations, shared=0, partial_dispatch=0, centroid_tracking=False):
        """Same arguments as Cpm.run, applied to every replica."""
        a = _run_args(cell_sync, block_sync, global_sync, threads_per_block, positions_per_thread,
                      positions_per_checkerboard, updates_per_checkerboard_switch,
                      updates_per_barrier, iterations, inner_iterations, shared, partial_dispatch,
                      centroid_tracking)
        check(lib.cpm_batch_run(self._h, ctypes.byref(a)))

    def synchronize(self):
        check(lib.cpm_batch_sync(self._h))

    def set_autotune(self, on=True):
        """Time the candidate kernel forms (attempts per thread, staged or
        scalar parameters) on this batch's own launches and keep the fastest."""
        check(lib.cpm_batch_set_autotune(self._h, int(bool(on))))

    def set_form(self, form):
        """0: one launch per colour phase for all replicas (default);
        1: one persistent cooperative launch per chunk (2D);
        2: the per-phase launches of up to 64 MCS captured as a hip graph and
        replayed; -1: choose (form 0, or with set_autotune the fastest of 0,
        1 and 2 by whole-run timings)."""
        check(lib.cpm_batch_set_form(self._h, _int(form, "form")))

    def graph_info(self):
        """(form the last run used, graph captures, graph replays)."""
        f, c, r = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        check(lib.cpm_batch_graph_info(self._h, ctypes.byref(f), ctypes.byref(c), ctypes.byref(r)))
        return f.value, c.value, r.value

    def set_kernel_options(self, small=None, compact=None, attempts=None):
        """Cpm.set_kernel_options for batched runs (attempts: 1, 4 or 8)."""
        v = [-1 if x is None else int(x) for x in (small, compact, attempts)]
        check(lib.cpm_batch_set_kernel_options(self._h, *v))

    def set_blocks(self, blocks_per_work):
        """Blocks per unit of work in each replica's group (default 1.0)."""
        check(lib.cpm_batch_set_blocks(self._h, float(blocks_per_work)))

    def timing_begin(self):
        """Start bracketing every batched launch with HIP events."""
        check(lib.cpm_batch_timing_begin(self._h))

    def timing_end(self):
        """Synchronize; return (kernel_ms, launches, site_updates) since timing_begin."""
        ms, n, upd = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        check(lib.cpm_batch_timing_end(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(upd)))
        return ms.value, n.value, upd.value

    @property
    def last_kernel(self):
        """Name of the per-phase batch kernel of the last run."""
        return lib.cpm_batch_last_kernel(self._h).decode()

    @property
    def blocks(self):
        n, b = ctypes.c_int(), ctypes.c_int()
        check(lib.cpm_batch_info(self._h, ctypes.byref(n), ctypes.byref(b)))
        return b.value
