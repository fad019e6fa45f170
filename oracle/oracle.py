"""ctypes wrapper over the CPU oracle (libcpm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (the ``gpucpm``
package) never imports it.

The oracle restates the reference sweep (shabaz/gpu-cpm src/kernels2d.cu,
src/kernels3d.cu, src/kernels.cu, src/random.cu); see cpm_oracle.h for the
file:line map and the deterministic phase rules.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

NTYPES = 20
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libcpm_oracle.so")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIBPATH):
        subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE], check=True)
    return _LIBPATH


class Params(ctypes.Structure):
    """Mirror of orc_params (settings.h:4-31, hot-path fields only)."""

    _fields_ = [
        ("J", (ctypes.c_int32 * NTYPES) * NTYPES),
        ("S", (ctypes.c_int32 * NTYPES) * NTYPES),
        ("lam_area", ctypes.c_double * NTYPES),
        ("target_area", ctypes.c_int32 * NTYPES),
        ("lam_perim", ctypes.c_double * NTYPES),
        ("target_perim", ctypes.c_int32 * NTYPES),
        ("temperature", ctypes.c_double),
        ("fixed", ctypes.c_int32 * NTYPES),
        ("lam_chem", ctypes.c_double * NTYPES),
        ("field_q", ctypes.POINTER(ctypes.c_int32)),
        ("lam_act", ctypes.c_double * NTYPES),
        ("act_max", ctypes.c_int32 * NTYPES),
        ("act", ctypes.POINTER(ctypes.c_uint32)),
        ("lam_pers", ctypes.c_double * NTYPES),
        ("pref_q", ctypes.POINTER(ctypes.c_int32)),
    ]

    def set_pref(self, pref_q):
        """Attach quantized preferred directions (int32 [(ncells+1)*3])."""
        assert pref_q.dtype == np.int32 and pref_q.flags["C_CONTIGUOUS"]
        self._pref_q = pref_q
        self.pref_q = _ptr(pref_q, ctypes.c_int32)

    def set_act(self, act):
        """Attach the act lattice (uint32, C-contiguous, the lattice's shape);
        sweeps update it in place."""
        assert act.dtype == np.uint32 and act.flags["C_CONTIGUOUS"]
        self._act = act
        self.act = _ptr(act, ctypes.c_uint32)

    def set_field(self, field):
        """Attach a field [d][L]^d (float64): quantized to 2^-16 as the engine
        does at push.  The quantized copy is kept alive on this object."""
        self._field_q = quantize_field(field)
        self.field_q = _ptr(self._field_q, ctypes.c_int32)

    @classmethod
    def make(cls, temperature, J=None, S=None, lam_area=None, target_area=None,
             lam_perim=None, target_perim=None, fixed=None, lam_chem=None, field=None,
             lam_act=None, act_max=None, act=None, lam_pers=None):
        p = cls()
        p.temperature = float(temperature)
        for name, val in (("J", J), ("S", S)):
            if val is not None:
                arr = np.zeros((NTYPES, NTYPES), np.int32)
                v = np.asarray(val, np.int32)
                arr[: v.shape[0], : v.shape[1]] = v
                for a in range(NTYPES):
                    for b in range(NTYPES):
                        getattr(p, name)[a][b] = int(arr[a, b])
        for name, val in (("lam_area", lam_area), ("target_area", target_area),
                          ("lam_perim", lam_perim), ("target_perim", target_perim),
                          ("fixed", fixed), ("lam_chem", lam_chem), ("lam_act", lam_act),
                          ("act_max", act_max), ("lam_pers", lam_pers)):
            if val is not None:
                for t, x in enumerate(val):
                    getattr(p, name)[t] = float(x) if name.startswith("lam") else int(x)
        if field is not None:
            p.set_field(field)
        if act is not None:
            p.set_act(act)
        return p

    def J_array(self):
        return np.array([[self.J[a][b] for b in range(NTYPES)] for a in range(NTYPES)])


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIBPATH)
        P = ctypes.POINTER
        u32p, i32p, i64p, f64p = P(ctypes.c_uint32), P(ctypes.c_int32), P(ctypes.c_int64), P(ctypes.c_double)
        L.orc_minstd_next_f64.argtypes = [ctypes.c_int32]
        L.orc_minstd_next_f64.restype = ctypes.c_int32
        L.orc_minstd_next_int.argtypes = [ctypes.c_uint32]
        L.orc_minstd_next_int.restype = ctypes.c_uint32
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_philox4x32_10.restype = None
        L.orc_energy_scale.argtypes = [P(Params)]
        L.orc_energy_scale.restype = ctypes.c_int
        L.orc_threshold_table.argtypes = [ctypes.c_double, ctypes.c_int, u32p, ctypes.c_int]
        L.orc_threshold_table.restype = ctypes.c_int
        L.orc_tally.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p, i32p, i32p, i32p, i64p]
        L.orc_tally.restype = ctypes.c_int
        L.orc_sweep.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p, i32p, i32p, i32p, i64p,
                                P(Params), ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
        L.orc_sweep.restype = ctypes.c_int
        L.orc_sweep_sync.argtypes = L.orc_sweep.argtypes + [ctypes.c_int]
        L.orc_sweep_sync.restype = ctypes.c_int
        L.orc_sweep_ex.argtypes = L.orc_sweep.argtypes + [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_sweep_ex.restype = ctypes.c_int
        L.orc_sweep_counts.argtypes = L.orc_sweep_ex.argtypes + [i64p]
        L.orc_sweep_counts.restype = ctypes.c_int
        L.orc_sweep_live.argtypes = L.orc_sweep.argtypes[:-1] + [ctypes.c_int, ctypes.c_int]
        L.orc_sweep_live.restype = ctypes.c_int
        L.orc_sweep_refstream.argtypes = L.orc_sweep_ex.argtypes + [
            P(ctypes.c_int64), ctypes.c_int64, P(ctypes.c_int64), ctypes.c_int]
        L.orc_sweep_refstream.restype = ctypes.c_int
        L.orc_ref_threshold_table.argtypes = [ctypes.c_double, ctypes.c_int, u32p, ctypes.c_int]
        L.orc_ref_threshold_table.restype = ctypes.c_int
        L.orc_schedule.argtypes = [ctypes.c_int] * 5 + [P(ctypes.c_int), P(ctypes.c_int)]
        L.orc_schedule.restype = ctypes.c_int
        L.orc_serial_sweep.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p, i32p, i32p, i32p,
                                       i64p, P(Params), ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int]
        L.orc_serial_sweep.restype = ctypes.c_int
        L.orc_delta_h.argtypes = [ctypes.c_int, ctypes.c_int, u32p, i32p, i32p, P(Params), ctypes.c_int,
                                  ctypes.c_int64, ctypes.c_int64]
        L.orc_delta_h.restype = ctypes.c_int64
        L.orc_hamiltonian.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p, P(Params)]
        L.orc_hamiltonian.restype = ctypes.c_double
        L.orc_energy_stats.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p, P(Params),
                                       i64p, i64p, i64p, i64p, i64p]
        L.orc_energy_stats.restype = ctypes.c_int
        L.orc_centroids.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, i64p, f64p]
        L.orc_centroids.restype = None
        L.orc_quantize_field.argtypes = [ctypes.c_int64, f64p, i32p]
        L.orc_quantize_field.restype = ctypes.c_int
        L.orc_chem_multiplier.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_chem_multiplier.restype = ctypes.c_int64
        L.orc_act_log2.argtypes = [ctypes.c_int]
        L.orc_act_log2.restype = ctypes.c_int64
        L.orc_act_geo_mean.argtypes = [ctypes.c_int64, ctypes.c_int]
        L.orc_act_geo_mean.restype = ctypes.c_int64
        L.orc_act_multiplier.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_act_multiplier.restype = ctypes.c_int64
        L.orc_act_energy.argtypes = [ctypes.c_int, ctypes.c_int, u32p, P(Params), ctypes.c_int,
                                     ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32]
        L.orc_act_energy.restype = ctypes.c_int64
        L.orc_pref_centroids.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, i64p, f64p]
        L.orc_pref_centroids.restype = None
        L.orc_update_pref_dirs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, i64p,
                                           i32p, f64p, f64p, f64p, ctypes.c_int, P(ctypes.c_int), i32p]
        L.orc_update_pref_dirs.restype = None
        L.orc_init_pref_dirs.argtypes = [ctypes.c_int, ctypes.c_uint64, f64p]
        L.orc_init_pref_dirs.restype = None
        _lib = L
    return _lib


def _ptr(a, ctype):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def minstd_f64(x: int) -> int:
    return lib().orc_minstd_next_f64(x)


def minstd_int(x: int) -> int:
    return lib().orc_minstd_next_int(x)


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32).copy()
    k = np.asarray(key, np.uint32).copy()
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_ptr(c, ctypes.c_uint32), _ptr(k, ctypes.c_uint32), _ptr(out, ctypes.c_uint32))
    return out


def quantize_field(field) -> np.ndarray:
    """float64 field -> int32 round(f * 2^16) (orc_quantize_field)."""
    f = np.ascontiguousarray(field, np.float64)
    out = np.zeros(f.shape, np.int32)
    if lib().orc_quantize_field(f.size, _ptr(f, ctypes.c_double), _ptr(out, ctypes.c_int32)):
        raise ValueError("field values must satisfy |f| < 2^15")
    return out


def act_log2(v: int) -> int:
    return lib().orc_act_log2(int(v))


def act_geo_mean(sum_lg: int, count: int) -> int:
    """Geometric mean in 2^-16 units from a sum of act_log2 values."""
    return lib().orc_act_geo_mean(int(sum_lg), int(count))


def act_energy(lat, d, p: Params, scale: int, t: int, s: int, tick: int) -> int:
    lat32 = np.ascontiguousarray(lat).view(np.uint32)
    return lib().orc_act_energy(lat32.shape[0], d, _ptr(lat32, ctypes.c_uint32), ctypes.byref(p),
                                scale, t, s, tick & 0xFFFFFFFF)


def init_pref_dirs(ncells: int, seed: int) -> np.ndarray:
    out = np.zeros((ncells + 1) * 3, np.float64)
    lib().orc_init_pref_dirs(int(ncells), int(seed) & (2**64 - 1), _ptr(out, ctypes.c_double))
    return out


def pref_centroids(L, d, tables: Tables) -> np.ndarray:
    out = np.zeros((tables.ncells + 1) * 3, np.float64)
    lib().orc_pref_centroids(L, d, tables.ncells, _ptr(tables.area, ctypes.c_int32),
                             _ptr(tables.anchor, ctypes.c_int32), _ptr(tables.csum, ctypes.c_int64),
                             _ptr(out, ctypes.c_double))
    return out


class PrefState:
    """Preferred directions and the centroid history ring (H entries, all
    holding the centroids when the state is made)."""

    def __init__(self, L, d, tables: Tables, H: int, seed: int):
        self.L, self.d, self.H = L, d, H
        self.pref = init_pref_dirs(tables.ncells, seed)
        self.ring = np.tile(pref_centroids(L, d, tables), H).copy()
        self.active = 0
        self.pref_q = quantize_field(self.pref)  # llround(p 2^16), as after each update

    def update(self, tables: Tables, ctype, diffusion):
        act = ctypes.c_int(self.active)
        ct = np.ascontiguousarray(ctype, np.int32)
        df = np.ascontiguousarray(diffusion, np.float64)
        lib().orc_update_pref_dirs(self.L, self.d, tables.ncells, _ptr(tables.area, ctypes.c_int32),
                                   _ptr(tables.anchor, ctypes.c_int32), _ptr(tables.csum, ctypes.c_int64),
                                   _ptr(ct, ctypes.c_int32), _ptr(df, ctypes.c_double),
                                   _ptr(self.pref, ctypes.c_double), _ptr(self.ring, ctypes.c_double),
                                   self.H, ctypes.byref(act), _ptr(self.pref_q, ctypes.c_int32))
        self.active = act.value


def act_multiplier(lam: float, act_max: int, scale: int) -> int:
    return lib().orc_act_multiplier(lam, int(act_max), int(scale))


def chem_multiplier(lam: float, scale: int, k: int) -> int:
    return lib().orc_chem_multiplier(lam, scale, k)


def energy_scale(p: Params) -> int:
    return lib().orc_energy_scale(ctypes.byref(p))


def threshold_table(temperature: float, scale: int, cap: int = 1 << 22):
    out = np.zeros(cap, np.uint32)
    n = lib().orc_threshold_table(float(temperature), int(scale), _ptr(out, ctypes.c_uint32), cap)
    if n < 0:
        raise ValueError("threshold table does not fit")
    return out[:n].copy()


class Tables:
    """Cell tables in the oracle's layout (area, perimeter, anchor, csum)."""

    def __init__(self, ncells):
        self.ncells = ncells
        self.area = np.zeros(ncells + 1, np.int32)
        self.perim = np.zeros(ncells + 1, np.int32)
        self.anchor = np.zeros((ncells + 1) * 3, np.int32)
        self.csum = np.zeros((ncells + 1) * 3, np.int64)

    def copy(self):
        t = Tables(self.ncells)
        for k in ("area", "perim", "anchor", "csum"):
            setattr(t, k, getattr(self, k).copy())
        return t


def tally(lat: np.ndarray, d: int, ncells: int) -> Tables:
    lat = np.ascontiguousarray(lat).view(np.uint32)
    L = lat.shape[0]
    t = Tables(ncells)
    rc = lib().orc_tally(L, d, ncells, _ptr(lat, ctypes.c_uint32), _ptr(t.area, ctypes.c_int32),
                         _ptr(t.perim, ctypes.c_int32), _ptr(t.anchor, ctypes.c_int32),
                         _ptr(t.csum, ctypes.c_int64))
    if rc:
        raise ValueError("bad lattice: spin above ncells, type above 19, or L not a power of two")
    return t


def schedule(L: int, d: int, c: int, shared: bool, ppt: int):
    """(K colours per axis, S groups per axis) of a run() schedule."""
    K, S = ctypes.c_int(), ctypes.c_int()
    if lib().orc_schedule(L, d, c, int(bool(shared)), int(ppt), ctypes.byref(K), ctypes.byref(S)):
        raise ValueError("bad positions_per_checkerboard / positions_per_thread")
    return K.value, S.value


def sweep(lat: np.ndarray, d: int, tables: Tables, p: Params, c: int, u: int, seed: int,
          mcs0: int, nmcs: int, nthreads: int = 1, cell_sync: bool = False, shared: bool = True,
          ppt: int = None, rng: str = "philox", trace_cap: int = 0, counts: np.ndarray = None,
          tpb: int = 4):
    """Run nmcs MCS in place on lat (uint32 view) and tables.  shared selects
    the roles of the run() kwarg of that name; ppt is positions_per_thread
    (used without shared roles).  rng = "minstd" uses the reference's
    per-thread MINSTD stream (one call = one reference launch: the streams
    are seeded at entry); with trace_cap > 0 the attempts' (region, target,
    source, acceptance draw or -1) rows are returned (tpb: threads_per_block,
    which places the regions with shared roles and ppt > 2c).  counts (int64[4],
    added to): attempts with differing source and target, evaluated,
    accepted, committed (counter stream only)."""
    assert lat.dtype in (np.uint32, np.int32) and lat.flags["C_CONTIGUOUS"]
    lat32 = lat.view(np.uint32)
    L = lat.shape[0]
    args = (L, d, tables.ncells, _ptr(lat32, ctypes.c_uint32),
            _ptr(tables.area, ctypes.c_int32), _ptr(tables.perim, ctypes.c_int32),
            _ptr(tables.anchor, ctypes.c_int32), _ptr(tables.csum, ctypes.c_int64),
            ctypes.byref(p), int(c), int(u), int(seed) & (2**64 - 1),
            int(mcs0) & 0xFFFFFFFF, int(nmcs), int(nthreads), int(bool(cell_sync)),
            int(bool(shared)), int(ppt if ppt is not None else 2 * c))
    if rng == "philox" and counts is not None:
        assert counts.dtype == np.int64 and counts.size >= 4
        rc = lib().orc_sweep_counts(*args, _ptr(counts, ctypes.c_int64))
        trace = None
    elif rng == "philox":
        rc = lib().orc_sweep_ex(*args)
        trace = None
    elif rng == "minstd":
        trace = np.zeros((max(trace_cap, 1), 4), np.int64)
        n = ctypes.c_int64(0)
        rc = lib().orc_sweep_refstream(*args, _ptr(trace, ctypes.c_int64) if trace_cap else None,
                                       int(trace_cap), ctypes.byref(n), int(tpb))
        trace = trace[:n.value] if trace_cap else None
    else:
        raise ValueError("rng must be 'philox' or 'minstd'")
    if rc:
        raise ValueError(f"orc_sweep failed with code {rc}")
    return trace


def sweep_live(lat: np.ndarray, d: int, tables: Tables, p: Params, c: int, u: int, seed: int,
               mcs0: int, nmcs: int, shared: bool = True, ppt: int = None):
    """The live-table checkerboard (orc_sweep_live): sweep()'s schedule and
    stream, each phase's attempts one at a time in a per-phase random order
    against the live area / perimeter tables."""
    assert lat.dtype in (np.uint32, np.int32) and lat.flags["C_CONTIGUOUS"]
    lat32 = lat.view(np.uint32)
    rc = lib().orc_sweep_live(lat.shape[0], d, tables.ncells, _ptr(lat32, ctypes.c_uint32),
                              _ptr(tables.area, ctypes.c_int32), _ptr(tables.perim, ctypes.c_int32),
                              _ptr(tables.anchor, ctypes.c_int32), _ptr(tables.csum, ctypes.c_int64),
                              ctypes.byref(p), int(c), int(u), int(seed) & (2**64 - 1),
                              int(mcs0) & 0xFFFFFFFF, int(nmcs), int(bool(shared)),
                              int(ppt if ppt is not None else 2 * c))
    if rc:
        raise ValueError(f"orc_sweep_live failed with code {rc}")


def ref_threshold_table(temperature: float, scale: int, cap: int = 1 << 16) -> np.ndarray:
    out = np.zeros(cap, np.uint32)
    n = lib().orc_ref_threshold_table(float(temperature), int(scale), _ptr(out, ctypes.c_uint32), cap)
    if n < 0:
        raise ValueError("table too long")
    return out[:n]


def serial_sweep(lat: np.ndarray, d: int, tables: Tables, p: Params, seed: int, mcs0: int,
                 nmcs: int) -> None:
    """Serial random-sequential CPM (statistical comparisons only)."""
    lat32 = lat.view(np.uint32)
    rc = lib().orc_serial_sweep(lat.shape[0], d, tables.ncells, _ptr(lat32, ctypes.c_uint32),
                                _ptr(tables.area, ctypes.c_int32), _ptr(tables.perim, ctypes.c_int32),
                                _ptr(tables.anchor, ctypes.c_int32), _ptr(tables.csum, ctypes.c_int64),
                                ctypes.byref(p), int(seed) & (2**64 - 1), int(mcs0) & 0xFFFFFFFF,
                                int(nmcs))
    if rc:
        raise ValueError(f"orc_serial_sweep failed with code {rc}")


def delta_h(lat, d, tables: Tables, p: Params, scale: int, t: int, s: int) -> int:
    lat32 = np.ascontiguousarray(lat).view(np.uint32)
    return lib().orc_delta_h(lat32.shape[0], d, _ptr(lat32, ctypes.c_uint32),
                             _ptr(tables.area, ctypes.c_int32), _ptr(tables.perim, ctypes.c_int32),
                             ctypes.byref(p), scale, t, s)


def hamiltonian(lat, d, ncells, p: Params) -> float:
    lat32 = np.ascontiguousarray(lat).view(np.uint32)
    return lib().orc_hamiltonian(lat32.shape[0], d, ncells, _ptr(lat32, ctypes.c_uint32), ctypes.byref(p))


def energy_stats(lat, d, ncells, p: Params):
    lat32 = np.ascontiguousarray(lat).view(np.uint32)
    pairs = np.zeros((NTYPES, NTYPES), np.int64)
    sa2, sa1, sp2, sp1 = (np.zeros(NTYPES, np.int64) for _ in range(4))
    rc = lib().orc_energy_stats(lat32.shape[0], d, ncells, _ptr(lat32, ctypes.c_uint32), ctypes.byref(p),
                                _ptr(pairs, ctypes.c_int64), _ptr(sa2, ctypes.c_int64),
                                _ptr(sa1, ctypes.c_int64), _ptr(sp2, ctypes.c_int64),
                                _ptr(sp1, ctypes.c_int64))
    if rc:
        raise ValueError("bad lattice")
    return {"pairs": pairs, "sa2": sa2, "sa1": sa1, "sp2": sp2, "sp1": sp1}


def centroids(L, d, tables: Tables):
    out = np.zeros((tables.ncells + 1) * d, np.float64)
    lib().orc_centroids(L, d, tables.ncells, _ptr(tables.area, ctypes.c_int32),
                        _ptr(tables.anchor, ctypes.c_int32), _ptr(tables.csum, ctypes.c_int64),
                        _ptr(out, ctypes.c_double))
    return out.reshape(tables.ncells + 1, d)
