"""ctypes binding of the FP64 CPU oracle (libpn_oracle.so).

TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module, and only as what results are compared with, or as the timed CPU
baseline.  The product library (penguin_2_0_amd/libpenguin.so) never loads it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from penguin_2_0_amd.abi import NPAD, PnBody, PnGrid, PnParams, PnRunState, PnSymdisk, ring_dtype

ORACLE_PATH = Path(__file__).resolve().parent / "libpn_oracle.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_PATH.exists():
            raise RuntimeError(f"{ORACLE_PATH} is not built; run `python -m penguin_2_0_amd.build --oracle`")
        L = C.CDLL(str(ORACLE_PATH))
        vp, d, i = C.c_void_p, C.c_double, C.c_int
        P = C.POINTER
        G, Pp, B = P(PnGrid), P(PnParams), P(PnBody)
        dp = P(C.c_double)
        sig = {
            "orc_init_disk": (i, [G, Pp, C.c_uint64, d, vp, vp, vp, vp, vp, P(PnSymdisk)]),
            "orc_get_dt": (d, [G, Pp, vp, d, d]),
            "orc_index_shift": (None, [G, Pp, vp, P(PnSymdisk), d, d]),
            "orc_sweepx": (None, [G, Pp, vp, vp, vp, d, B, d]),
            "orc_sweepy": (None, [G, Pp, vp, d, B, d]),
            "orc_sweepz": (None, [G, Pp, vp, vp, vp, d, B, d]),
            "orc_rotate": (None, [G, vp]),
            "orc_damp": (None, [G, Pp, vp, P(PnSymdisk), d]),
            "orc_reduce": (d, [G, Pp, vp, B, d, i]),
            "orc_totals": (None, [G, Pp, vp, dp, dp]),
            "orc_ring_means": (None, [G, vp, i, dp]),
            "orc_set_threads": (None, [i]),
            "orc_get_threads": (i, []),
            "orc_init_planet": (None, [Pp, B, dp]),
            "orc_planet_forces": (None, [G, Pp, vp, B, d, d]),
            "orc_kick_drift_kick": (None, [Pp, B, dp, d, i]),
            "orc_run_begin": (None, [G, Pp, vp, P(PnSymdisk), P(PnRunState)]),
            "orc_run": (None, [G, Pp, vp, vp, vp, vp, vp, P(PnSymdisk), i, P(PnRunState)]),
            "orc_parabola_window": (None, [i, i, i, dp, dp, dp, dp, dp, dp]),
            "orc_paraset_coeffs": (None, [i, i, i, dp, dp]),
            "orc_flatten_window": (None, [i, dp, dp, dp]),
            "orc_riemann": (None, [i, d, d, d, d, d, d, d, dp, dp]),
            "orc_derivs": (None, [dp, dp, i, dp, dp]),
            "orc_force_rad_cyl": (d, [B, d, d, d, d]),
            "orc_force_azi_cyl": (d, [B, d, d, d, d]),
            "orc_potential": (d, [Pp, B, d, d, d]),
            "orc_gravity": (d, [i, i, i, B, d, d, d, d]),
            "orc_grid": (None, [i, d, d, i, dp, dp, dp]),
            "orc_ppmlr_window": (None, [Pp, i, i, d, d, B, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]),
            "orc_splitmix64": (C.c_uint64, [C.c_uint64]),
            "orc_noise_u": (d, [C.c_uint64, C.c_uint64]),
        }
        for n, (res, args) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def vptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleDisk:
    """A disk state held in hydr_ring bytes, advanced by the oracle."""

    def __init__(self, grid: PnGrid, params: PnParams, seed: int = 1, noise: float = 1e-4):
        self.grid, self.params = grid, params
        dt = ring_dtype(grid.jmax)
        n = grid.imax * grid.kmax
        self.rings = np.zeros(n, dt)
        self.lft = np.zeros(NPAD * grid.kmax, dt)
        self.rgh = np.zeros(NPAD * grid.kmax, dt)
        self.udr = np.zeros(NPAD * grid.imax, dt) if params.ndim == 3 else None
        self.top = np.zeros(NPAD * grid.imax, dt) if params.ndim == 3 else None
        self.val = (PnSymdisk * n)()
        rc = lib().orc_init_disk(C.byref(grid), C.byref(params), seed, noise, vptr(self.rings), vptr(self.lft),
                                 vptr(self.rgh), vptr(self.udr), vptr(self.top), self.val)
        if rc != 0:
            raise RuntimeError(f"orc_init_disk failed: {rc}")

    def run_begin(self) -> PnRunState:
        st = PnRunState()
        lib().orc_run_begin(C.byref(self.grid), C.byref(self.params), vptr(self.rings), self.val, C.byref(st))
        return st

    def run(self, nsteps: int, st: PnRunState) -> PnRunState:
        lib().orc_run(C.byref(self.grid), C.byref(self.params), vptr(self.rings), vptr(self.lft), vptr(self.rgh),
                      vptr(self.udr), vptr(self.top), self.val, nsteps, C.byref(st))
        return st

    def get_dt(self, dt_prev: float, FrRot: float) -> float:
        return lib().orc_get_dt(C.byref(self.grid), C.byref(self.params), vptr(self.rings), dt_prev, FrRot)

    def index_shift(self, dt: float, FrRot: float):
        lib().orc_index_shift(C.byref(self.grid), C.byref(self.params), vptr(self.rings), self.val, dt, FrRot)

    def sweepx(self, dt: float, planet: PnBody, FrRot: float):
        lib().orc_sweepx(C.byref(self.grid), C.byref(self.params), vptr(self.rings), vptr(self.lft),
                         vptr(self.rgh), dt, C.byref(planet), FrRot)

    def sweepy(self, dt: float, planet: PnBody, FrRot: float):
        lib().orc_sweepy(C.byref(self.grid), C.byref(self.params), vptr(self.rings), dt, C.byref(planet), FrRot)

    def sweepz(self, dt: float, planet: PnBody, FrRot: float):
        lib().orc_sweepz(C.byref(self.grid), C.byref(self.params), vptr(self.rings), vptr(self.udr),
                         vptr(self.top), dt, C.byref(planet), FrRot)

    def rotate(self):
        lib().orc_rotate(C.byref(self.grid), vptr(self.rings))

    def damp(self, dt: float):
        lib().orc_damp(C.byref(self.grid), C.byref(self.params), vptr(self.rings), self.val, dt)

    def reduce(self, planet: PnBody, exclude: float, mode: int) -> float:
        return lib().orc_reduce(C.byref(self.grid), C.byref(self.params), vptr(self.rings), C.byref(planet),
                                exclude, mode)

    def ring_means(self, field: int | str) -> np.ndarray:
        if isinstance(field, str):
            field = "rpuvw".index(field)
        out = np.zeros(self.grid.imax * self.grid.kmax)
        lib().orc_ring_means(C.byref(self.grid), vptr(self.rings), field, dptr(out))
        return out

    def totals(self) -> tuple[float, float]:
        m, l = C.c_double(), C.c_double()
        lib().orc_totals(C.byref(self.grid), C.byref(self.params), vptr(self.rings), C.byref(m), C.byref(l))
        return m.value, l.value
