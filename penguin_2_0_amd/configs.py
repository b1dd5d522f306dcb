"""The five disk workloads as (grid, parameters, seed, steps).

Parameter mapping onto the reference's global.h constants:
  r in [0.4, 2.5] log-spaced (grspx 1), phi in [0, 2 pi) uniform; Sigma = r^-p_alpha,
  c_s^2 = sc_h^2 r^-p_beta, v_phi pressure-corrected Keplerian, v_r = -1.5 nu / r when
  viscosity is on (else 0); a planet of mass ratio q at r = 1 with Plummer softening
  rs_fac * h(1), fixed in the co-rotating frame (plnt_flag 1) and at full mass from t = 0
  (t_growth 0); FARGO 1 and the co-rotating frame on; CFL 0.5; fixed r ghost rings (the only
  r boundary the reference implements).  Seeded relative noise 1e-4 on Sigma.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .abi import PnGrid, PnParams, params_default

HPI = 1.570796326794896619231321691639751442


@dataclass
class Workload:
    name: str
    grid: PnGrid
    params: PnParams
    seed: int
    steps: int
    noise: float = 1.0e-4

    @property
    def cells(self) -> int:
        return self.grid.imax * self.grid.jmax * self.grid.kmax


def disk2d(nr: int, nphi: int, q: float = 1.0e-4, **over) -> PnParams:
    p = params_default().replace(
        ndim=2, eos=0, flat_flag=1, plnt_flag=1, frrot_flag=1, visc_flag=0, fargo_flag=1,
        ngeomx=1, ngeomy=3, ngeomz=0, grspx=1, grspy=0, grspz=0, init_p_iso=0,
        p_alpha=1.0, p_beta=1.0, sc_h=0.05, ss_alpha=0.0, vis_nu=0.0,
        M_p=q, R_p=1.0, t_growth=0.0, rs_fac=0.6, gam=1.0, courant=0.5,
        xmin=0.4, xmax=2.5, ymin=0.0, ymax=2.0 * math.pi)
    return p.replace(**over) if over else p


def c1() -> Workload:
    """Plain FP64 2-D r-phi locally isothermal disk, 256 x 768, 100 steps."""
    return Workload("c1_disk2d_256x768", PnGrid(256, 768, 1), disk2d(256, 768), seed=1, steps=100)


def c2() -> Workload:
    """The same disk at 1024 x 3072, 1000 steps."""
    return Workload("c2_disk2d_1024x3072", PnGrid(1024, 3072, 1), disk2d(1024, 3072), seed=2, steps=1000)


def c3() -> Workload:
    """3-D spherical r-phi-theta isothermal disk, 256 x 768 x 64, theta = pi/2 +- 3h."""
    p = disk2d(256, 768, q=3.0e-5).replace(
        ndim=3, ngeomx=2, ngeomy=4, ngeomz=5, p_alpha=1.5, p_beta=1.0,
        zmin=HPI - 0.15, zmax=HPI + 0.15)
    return Workload("c3_disk3d_256x768x64", PnGrid(256, 768, 64), p, seed=3, steps=500)


def c4_members(seed: int = 4, n: int = 48) -> list[Workload]:
    """48 independent 2-D disks with uneven grids: N_r in {128..512}, N_phi = 3 N_r, planet q
    log-uniform on [1e-5, 1e-3], alpha log-uniform on [1e-4, 1e-2] (nu = alpha h^2)."""
    rng = np.random.default_rng(seed)
    sizes = [128, 192, 256, 384, 512]
    out = []
    for d in range(n):
        nr = int(sizes[rng.integers(len(sizes))])
        q = float(10.0 ** rng.uniform(-5.0, -3.0))
        alpha = float(10.0 ** rng.uniform(-4.0, -2.0))
        p = disk2d(nr, 3 * nr, q=q, visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05)
        out.append(Workload(f"c4_disk{d:02d}_{nr}x{3 * nr}", PnGrid(nr, 3 * nr, 1), p, seed=1000 + d, steps=2000))
    return out


def c5() -> Workload:
    """Large adiabatic 2-D disk, 4096 x 12288, gamma 1.4, initial P = c_s^2 Sigma, q = 1e-3."""
    p = disk2d(4096, 12288, q=1.0e-3, eos=2, gam=1.4, init_p_iso=1)
    return Workload("c5_disk2d_adiabatic_4096x12288", PnGrid(4096, 12288, 1), p, seed=5, steps=5000)


def c5e3() -> Workload:
    """The C5 disk with the energy-consistent adiabatic step (EOS 3) in place of EOS 2."""
    w = c5()
    return Workload("c5_disk2d_adiabatic_eos3_4096x12288", w.grid, w.params.replace(eos=3), seed=5, steps=5000)


def small(nr: int = 48, nphi: int = 144, **over) -> Workload:
    """A small disk of the same physics for fast parity tests."""
    return Workload(f"disk2d_{nr}x{nphi}", PnGrid(nr, nphi, 1), disk2d(nr, nphi, **over), seed=7, steps=10)
