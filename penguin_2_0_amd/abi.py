"""ctypes mirror of include/penguin.h and a thin Python wrapper over the C ABI.

This module only marshals arguments; every computation happens in libpenguin.so.  Loading
fails loudly when the library has not been built -- there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libpenguin.so"
NPAD = 6

PN_OK, PN_ERR_ARG, PN_ERR_HIP, PN_ERR_LAYOUT, PN_ERR_STATE, PN_ERR_UNSUPPORTED, PN_ERR_NUMERIC, PN_ERR_IO = range(8)
PN_FP64, PN_FP32 = 0, 1
ERR_NAMES = {0: "PN_OK", 1: "PN_ERR_ARG", 2: "PN_ERR_HIP", 3: "PN_ERR_LAYOUT", 4: "PN_ERR_STATE",
             5: "PN_ERR_UNSUPPORTED", 6: "PN_ERR_NUMERIC", 7: "PN_ERR_IO"}


class PnBody(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("m", "x", "y", "vx", "vy", "fx", "fy", "rs")]


class PnSymdisk(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("r", "p", "u", "v", "w")]


class PnGrid(C.Structure):
    _fields_ = [("imax", C.c_int32), ("jmax", C.c_int32), ("kmax", C.c_int32)]


class PnParams(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "ndim", "eos", "flat_flag", "plnt_flag", "frrot_flag", "visc_flag", "fargo_flag",
        "ngeomx", "ngeomy", "ngeomz", "grspx", "grspy", "grspz", "init_p_iso")] + \
        [(n, C.c_double) for n in (
            "p_alpha", "p_beta", "ss_alpha", "sc_h", "vis_nu", "M_p", "R_p", "t_growth", "rs_fac",
            "gam", "courant", "endtime", "tmovie", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax")] + \
        [(n, C.c_int32) for n in ("nlft", "nrgh", "kill_flag", "bary_flag")]

    def replace(self, **kw) -> "PnParams":
        out = PnParams()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(PnParams))
        for k, v in kw.items():
            if not hasattr(out, k):
                raise AttributeError(k)
            setattr(out, k, v)
        return out

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class PnRunState(C.Structure):
    _fields_ = [("dt", C.c_double), ("simtime", C.c_double), ("FrRot", C.c_double),
                ("planet", PnBody), ("ncycle", C.c_int64), ("timem", C.c_double), ("land", C.c_int32),
                ("event", C.c_int32)]


class PnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


def ring_dtype(jmax: int) -> np.dtype:
    """numpy view of hydr_ring (variable_types.h:67-97) for a runtime jmax."""
    f8, i4 = np.float64, np.int32
    dt = np.dtype([("r", f8, (jmax,)), ("p", f8, (jmax,)), ("u", f8, (jmax,)), ("v", f8, (jmax,)),
                   ("w", f8, (jmax,)), ("x", f8), ("dx", f8), ("xvol", f8), ("xc", f8),
                   ("y", f8, (jmax,)), ("dy", f8, (jmax,)), ("yvol", f8, (jmax,)), ("yc", f8, (jmax,)),
                   ("z", f8), ("dz", f8), ("zvol", f8), ("zc", f8), ("i", i4), ("k", i4),
                   ("rot_v", f8), ("res_v", f8), ("rot_j", i4), ("inc_j", i4)])
    assert dt.itemsize == 72 * jmax + 96
    return dt


_lib = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libpenguin.so; raises if it is missing (build with penguin_2_0_amd.build)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run `python -m penguin_2_0_amd.build` first")
    lib = C.CDLL(str(p))
    vp, sz, d, i = C.c_void_p, C.c_size_t, C.c_double, C.c_int
    P = C.POINTER
    sig = {
        "pn_params_default": (None, [P(PnParams)]),
        "pn_ring_bytes": (sz, [C.c_int32]),
        "pn_init_disk": (i, [P(PnGrid), P(PnParams), C.c_uint64, d, vp, vp, vp, vp, vp, P(PnSymdisk)]),
        "pn_create": (i, [P(PnGrid), P(PnParams), i, P(vp)]),
        "pn_destroy": (i, [vp]),
        "pn_last_error": (C.c_char_p, [vp]),
        "pn_load_rings": (i, [vp, vp, sz, vp, vp, sz, vp, vp, sz, P(PnSymdisk), sz]),
        "pn_store_rings": (i, [vp, vp, sz]),
        "pn_get_dt": (i, [vp, d, d, P(d)]),
        "pn_index_shift": (i, [vp, d, d]),
        "pn_sweep2d": (i, [vp, d, P(PnBody), d]),
        "pn_reduce": (i, [vp, P(PnBody), d, i, P(d)]),
        "pn_sweep_axis": (i, [vp, i, d, P(PnBody), d]),
        "pn_time_kernels": (i, [vp, i, P(d)]),
        "pn_tune": (i, [vp, i, P(i), P(i)]),
        "pn_set_variants": (i, [vp, i, i]),
        "pn_set_packing": (i, [vp, i]),
        "pn_set_variant_z": (i, [vp, i]),
        "pn_get_variant_z": (i, [vp, P(i)]),
        "pn_get_variants": (i, [vp, P(i), P(i)]),
        "pn_init_planet": (None, [P(PnParams), P(PnBody), P(d)]),
        "pn_planet_forces": (i, [vp, P(PnBody), d, d]),
        "pn_kick_drift_kick": (None, [P(PnParams), P(PnBody), P(d), d, i]),
        "pn_run_begin": (i, [vp, P(PnRunState)]),
        "pn_run_restart": (i, [vp, i, P(PnRunState)]),
        "pn_run": (i, [vp, i, i, P(PnRunState)]),
        "pn_totals": (i, [vp, P(d), P(d)]),
        "pn_totals_noalloc": (i, [vp, P(d), P(d)]),
        "pn_ring_means": (i, [vp, i, P(d), sz]),
        "pn_copy_bandwidth": (i, [sz, i, P(d)]),
        "pn_lab_fields": (i, [vp, sz, i, i, i, vp, vp, vp, vp, vp]),
        "pn_damp_edges": (i, [vp, d]),
        "pn_save_snapshot": (i, [vp, C.c_char_p]),
        "pn_load_snapshot": (i, [vp, C.c_char_p]),
        "pn_label": (i, [P(PnGrid), P(PnParams), C.c_char_p, sz]),
        "pn_frame_name": (i, [i, C.c_char_p, sz]),
        "pn_write_para": (i, [P(PnGrid), P(PnParams), C.c_char_p, C.c_char_p, C.c_char_p]),
        "pn_write_initial": (i, [P(PnGrid), P(PnParams), C.c_char_p]),
        "pn_batch_create": (i, [i, P(PnGrid), P(PnParams), i, P(vp)]),
        "pn_batch_destroy": (i, [vp]),
        "pn_batch_last_error": (C.c_char_p, [vp]),
        "pn_batch_size": (i, [vp]),
        "pn_batch_disk": (vp, [vp, i]),
        "pn_batch_run_begin": (i, [vp, P(PnRunState)]),
        "pn_batch_run_restart": (i, [vp, i, P(PnRunState)]),
        "pn_batch_run": (i, [vp, i, i, P(PnRunState)]),
        "pn_batch_tune": (i, [vp, i, P(i), P(i)]),
        "pn_batch_set_variants": (i, [vp, i, i]),
        "pn_batch_get_variants": (i, [vp, P(i), P(i)]),
        "pn_batch_time_kernels": (i, [vp, i, P(d)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """GB/s of a plain device copy of nbytes (read + write), timed on the device."""
    out = C.c_double()
    rc = load_library().pn_copy_bandwidth(nbytes, reps, C.byref(out))
    if rc != PN_OK:
        raise PnError(rc, "pn_copy_bandwidth failed")
    return out.value


def params_default() -> PnParams:
    p = PnParams()
    load_library().pn_params_default(C.byref(p))
    return p


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Disk:
    """Host buffers of one disk in the reference's byte layout."""

    def __init__(self, grid: PnGrid, params: PnParams):
        self.grid, self.params = grid, params
        self.dtype = ring_dtype(grid.jmax)
        n = grid.imax * grid.kmax
        self.rings = np.zeros(n, self.dtype)
        self.lft = np.zeros(NPAD * grid.kmax, self.dtype)
        self.rgh = np.zeros(NPAD * grid.kmax, self.dtype)
        three_d = params.ndim == 3
        self.udr = np.zeros(NPAD * grid.imax, self.dtype) if three_d else None
        self.top = np.zeros(NPAD * grid.imax, self.dtype) if three_d else None
        self.val = (PnSymdisk * n)()

    def copy(self) -> "Disk":
        d = Disk.__new__(Disk)
        d.grid, d.params, d.dtype = self.grid, self.params, self.dtype
        d.rings, d.lft, d.rgh = self.rings.copy(), self.lft.copy(), self.rgh.copy()
        d.udr = None if self.udr is None else self.udr.copy()
        d.top = None if self.top is None else self.top.copy()
        n = len(self.val)
        d.val = (PnSymdisk * n)()
        C.memmove(d.val, self.val, C.sizeof(PnSymdisk) * n)
        return d

    def val_array(self) -> np.ndarray:
        return np.ctypeslib.as_array(C.cast(self.val, C.POINTER(C.c_double)), shape=(len(self.val), 5)).copy()


def init_disk(grid: PnGrid, params: PnParams, seed: int, noise: float) -> Disk:
    d = Disk(grid, params)
    rc = load_library().pn_init_disk(C.byref(grid), C.byref(params), seed, noise, _ptr(d.rings), _ptr(d.lft),
                                     _ptr(d.rgh), _ptr(d.udr), _ptr(d.top), d.val)
    if rc != PN_OK:
        raise PnError(rc, "pn_init_disk")
    return d


def lab_fields(rings: np.ndarray, grid: PnGrid) -> dict:
    """The reader's lab-frame fields (pn_lab_fields): r, p, u, v, w as [i][j][k] arrays."""
    out = {f: np.empty((grid.imax, grid.jmax, grid.kmax)) for f in ("r", "p", "u", "v", "w")}
    rings = np.ascontiguousarray(rings)
    rc = load_library().pn_lab_fields(_ptr(rings), rings.nbytes, grid.imax, grid.jmax, grid.kmax,
                                      *(_ptr(out[f]) for f in ("r", "p", "u", "v", "w")))
    if rc != PN_OK:
        raise PnError(rc, "pn_lab_fields")
    return out


def label(grid: PnGrid, params: PnParams) -> str:
    """The run label of every output file name (pn_label, create_label)."""
    buf = C.create_string_buffer(256)
    rc = load_library().pn_label(C.byref(grid), C.byref(params), buf, len(buf))
    if rc != PN_OK:
        raise PnError(rc, "pn_label")
    return buf.value.decode()


def frame_name(num: int) -> str:
    """The five-digit snapshot suffix (pn_frame_name, frame_num)."""
    buf = C.create_string_buffer(8)
    rc = load_library().pn_frame_name(num, buf, len(buf))
    if rc != PN_OK:
        raise PnError(rc, "pn_frame_name")
    return buf.value.decode()


def write_para(grid: PnGrid, params: PnParams, mainp: str, label_: str, path) -> None:
    rc = load_library().pn_write_para(C.byref(grid), C.byref(params), mainp.encode(), label_.encode(),
                                      str(path).encode())
    if rc != PN_OK:
        raise PnError(rc, "pn_write_para")


def write_initial(grid: PnGrid, params: PnParams, path) -> None:
    rc = load_library().pn_write_initial(C.byref(grid), C.byref(params), str(path).encode())
    if rc != PN_OK:
        raise PnError(rc, "pn_write_initial")


class Context:
    """One device context (pn_create / pn_destroy)."""

    def __init__(self, grid: PnGrid, params: PnParams, precision: int = PN_FP64):
        self.lib = load_library()
        self.grid, self.params, self.precision = grid, params, precision
        h = C.c_void_p()
        rc = self.lib.pn_create(C.byref(grid), C.byref(params), precision, C.byref(h))
        self.h = h
        if rc != PN_OK:
            msg = self.lib.pn_last_error(h).decode() if h.value else ""
            self.close()
            raise PnError(rc, msg)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            if getattr(self, "_owned", True):
                self.lib.pn_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int, what: str):
        if rc != PN_OK:
            raise PnError(rc, f"{what}: {self.lib.pn_last_error(self.h).decode()}")

    def load(self, disk: Disk):
        ut = 0 if disk.udr is None else disk.udr.nbytes
        self._chk(self.lib.pn_load_rings(self.h, _ptr(disk.rings), disk.rings.nbytes, _ptr(disk.lft), _ptr(disk.rgh),
                                         disk.lft.nbytes, _ptr(disk.udr), _ptr(disk.top), ut, disk.val,
                                         len(disk.val)), "pn_load_rings")

    def store(self) -> np.ndarray:
        out = np.zeros(self.grid.imax * self.grid.kmax, ring_dtype(self.grid.jmax))
        self._chk(self.lib.pn_store_rings(self.h, _ptr(out), out.nbytes), "pn_store_rings")
        return out

    def get_dt(self, dt_prev: float, FrRot: float) -> float:
        d = C.c_double()
        self._chk(self.lib.pn_get_dt(self.h, dt_prev, FrRot, C.byref(d)), "pn_get_dt")
        return d.value

    def index_shift(self, dt: float, FrRot: float):
        self._chk(self.lib.pn_index_shift(self.h, dt, FrRot), "pn_index_shift")

    def sweep2d(self, dt: float, planet: PnBody, FrRot: float):
        self._chk(self.lib.pn_sweep2d(self.h, dt, C.byref(planet), FrRot), "pn_sweep2d")

    def sweep_axis(self, axis: int, dt: float, planet: PnBody, FrRot: float):
        self._chk(self.lib.pn_sweep_axis(self.h, axis, dt, C.byref(planet), FrRot), "pn_sweep_axis")

    def time_kernels(self, nrep: int) -> dict:
        ms = (C.c_double * 6)()
        self._chk(self.lib.pn_time_kernels(self.h, nrep, ms), "pn_time_kernels")
        return dict(zip(("cfl", "index_shift", "sweep_x", "sweep_y", "sweep_z", "rotate"), list(ms)))

    def tune(self, nrep: int = 3) -> tuple[int, int]:
        vx, vy = C.c_int(), C.c_int()
        self._chk(self.lib.pn_tune(self.h, nrep, C.byref(vx), C.byref(vy)), "pn_tune")
        return vx.value, vy.value

    def set_variants(self, var_x: int, var_y: int):
        self._chk(self.lib.pn_set_variants(self.h, var_x, var_y), "pn_set_variants")

    def damp_edges(self, dt: float):
        self._chk(self.lib.pn_damp_edges(self.h, dt), "pn_damp_edges")

    def variants(self) -> tuple[int, int]:
        vx, vy = C.c_int(), C.c_int()
        self._chk(self.lib.pn_get_variants(self.h, C.byref(vx), C.byref(vy)), "pn_get_variants")
        return vx.value, vy.value

    def reduce(self, planet: PnBody, exclude: float, mode: int) -> float:
        d = C.c_double()
        self._chk(self.lib.pn_reduce(self.h, C.byref(planet), exclude, mode, C.byref(d)), "pn_reduce")
        return d.value

    def totals(self) -> tuple[float, float]:
        m, l = C.c_double(), C.c_double()
        self._chk(self.lib.pn_totals(self.h, C.byref(m), C.byref(l)), "pn_totals")
        return m.value, l.value

    def set_variant_z(self, var_z: int):
        """The 3-D FP32 theta sweep's own shape (pn_set_variant_z)."""
        self._chk(self.lib.pn_set_variant_z(self.h, var_z), "pn_set_variant_z")

    def variant_z(self) -> int:
        v = C.c_int()
        self._chk(self.lib.pn_get_variant_z(self.h, C.byref(v)), "pn_get_variant_z")
        return v.value

    def set_packing(self, on: bool):
        """Two short last radial / theta windows per wavefront (pn_set_packing)."""
        self._chk(self.lib.pn_set_packing(self.h, 1 if on else 0), "pn_set_packing")

    def totals_noalloc(self) -> tuple[float, float]:
        """pn_totals on the library's file-scope scratch (pn_totals_noalloc)."""
        m, l = C.c_double(), C.c_double()
        self._chk(self.lib.pn_totals_noalloc(self.h, C.byref(m), C.byref(l)), "pn_totals_noalloc")
        return m.value, l.value

    def ring_means(self, field: int | str) -> np.ndarray:
        """Azimuthal mean of a field per interior ring (pn_ring_means); field 0..4 or one of
        'r', 'p', 'u', 'v', 'w'.  Returns imax * kmax values in ring order i + imax k."""
        if isinstance(field, str):
            field = "rpuvw".index(field)
        n = self.grid.imax * self.grid.kmax
        out = np.zeros(n)
        self._chk(self.lib.pn_ring_means(self.h, field, out.ctypes.data_as(C.POINTER(C.c_double)), n),
                  "pn_ring_means")
        return out

    def run_begin(self) -> PnRunState:
        st = PnRunState()
        self._chk(self.lib.pn_run_begin(self.h, C.byref(st)), "pn_run_begin")
        return st

    def run_restart(self, startat: int) -> PnRunState:
        """Prime a run resumed from picture startat (simtime = tmovie * startat)."""
        st = PnRunState()
        self._chk(self.lib.pn_run_restart(self.h, startat, C.byref(st)), "pn_run_restart")
        return st

    def run(self, nsteps: int, st: PnRunState, graph: bool = True) -> PnRunState:
        self._chk(self.lib.pn_run(self.h, nsteps, 1 if graph else 0, C.byref(st)), "pn_run")
        return st

    def save_snapshot(self, path: str):
        self._chk(self.lib.pn_save_snapshot(self.h, str(path).encode()), "pn_save_snapshot")

    def load_snapshot(self, path: str):
        self._chk(self.lib.pn_load_snapshot(self.h, str(path).encode()), "pn_load_snapshot")

    @classmethod
    def borrowed(cls, h: int, grid: PnGrid, params: PnParams, precision: int) -> "Context":
        """A wrapper over a context another object owns (a batch's disk): close() leaves it."""
        self = cls.__new__(cls)
        self.lib = load_library()
        self.grid, self.params, self.precision = grid, params, precision
        self.h = C.c_void_p(h)
        self._owned = False
        return self


class Batch:
    """Many independent FP32 disks advanced together (pn_batch_*)."""

    def __init__(self, grids: list[PnGrid], params: list[PnParams], precision: int = PN_FP32):
        self.lib = load_library()
        n = len(grids)
        self.grids, self.params = list(grids), list(params)
        ga = (PnGrid * n)(*grids)
        pa = (PnParams * n)(*params)
        h = C.c_void_p()
        rc = self.lib.pn_batch_create(n, ga, pa, precision, C.byref(h))
        self.h = h
        if rc != PN_OK:
            msg = self.lib.pn_batch_last_error(h).decode() if h.value else ""
            self.close()
            raise PnError(rc, msg)
        self.disks = [Context.borrowed(self.lib.pn_batch_disk(h, d), grids[d], params[d], precision)
                      for d in range(n)]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.pn_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int, what: str):
        if rc != PN_OK:
            raise PnError(rc, f"{what}: {self.lib.pn_batch_last_error(self.h).decode()}")

    def __len__(self) -> int:
        return len(self.disks)

    def run_begin(self):
        st = (PnRunState * len(self))()
        self._chk(self.lib.pn_batch_run_begin(self.h, st), "pn_batch_run_begin")
        return st

    def run_restart(self, startat: int):
        """Prime every disk for a run resumed from picture startat."""
        st = (PnRunState * len(self))()
        self._chk(self.lib.pn_batch_run_restart(self.h, startat, st), "pn_batch_run_restart")
        return st

    def run(self, nsteps: int, states, graph: bool = True):
        self._chk(self.lib.pn_batch_run(self.h, nsteps, 1 if graph else 0, states), "pn_batch_run")
        return states

    def tune(self, nrep: int = 3) -> tuple[int, int]:
        vx, vy = C.c_int(), C.c_int()
        self._chk(self.lib.pn_batch_tune(self.h, nrep, C.byref(vx), C.byref(vy)), "pn_batch_tune")
        return vx.value, vy.value

    def set_variants(self, var_x: int, var_y: int):
        self._chk(self.lib.pn_batch_set_variants(self.h, var_x, var_y), "pn_batch_set_variants")

    def variants(self) -> tuple[int, int]:
        vx, vy = C.c_int(), C.c_int()
        self._chk(self.lib.pn_batch_get_variants(self.h, C.byref(vx), C.byref(vy)), "pn_batch_get_variants")
        return vx.value, vy.value

    def time_kernels(self, nrep: int) -> dict:
        ms = (C.c_double * 4)()
        self._chk(self.lib.pn_batch_time_kernels(self.h, nrep, ms), "pn_batch_time_kernels")
        return dict(zip(("pre_step", "sweep_x", "sweep_y", "post_step"), list(ms)))
