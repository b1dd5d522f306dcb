"""Pins of the oracle's planet integrator (orc_kick_drift_kick, orc_planet_forces; planet.cu:28-93)
and of the disk-to-planet force reductions (orc_reduce modes 8 and 9; device_func.cu:83-108),
each against what the mathematics fixes:

- a circular orbit stays circular, and the co-rotating frame reproduces the inertial orbit;
- on an eccentric Kepler orbit the kick-drift-kick conserves the angular momentum to rounding
  and the energy to second order in dt, and the orbit closes after one Kepler period;
- modes 8 / 9 equal a brute-force sum over the lab-frame cells of minus the gradient of the
  softened planet potential (taken here by finite differences of the potential, not from the
  oracle's force functions) times (Sigma - r^-p_alpha) times the cell volume, outside the
  excluded circle; a density field mirror-symmetric about the planet's azimuth has no
  azimuthal force.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle
from penguin_2_0_amd import configs
from penguin_2_0_amd.abi import PnBody, PnGrid

TWO_PI = 2.0 * math.pi


def kdk_orbit(P, b: PnBody, FrRot: float, dt: float, nsteps: int):
    """The planet part of the step loop (main.cu:385-421 as orc_run orders it): half kick,
    drift, forces at the new position, half kick.  Returns per-step (x, y, vx, vy + FrRot,
    FrRot) after each step."""
    L = oracle.lib()
    g = PnGrid(1, 12, 1)
    fr = C.c_double(FrRot)
    L.orc_planet_forces(C.byref(g), C.byref(P), None, C.byref(b), fr.value, 0.0)
    out = []
    for _ in range(nsteps):
        L.orc_kick_drift_kick(C.byref(P), C.byref(b), C.byref(fr), 0.5 * dt, 1)
        frame = fr.value
        L.orc_kick_drift_kick(C.byref(P), C.byref(b), C.byref(fr), dt, 0)
        L.orc_planet_forces(C.byref(g), C.byref(P), None, C.byref(b), fr.value, 0.0)
        L.orc_kick_drift_kick(C.byref(P), C.byref(b), C.byref(fr), 0.5 * dt, 1)
        out.append((b.x, b.y, b.vx, b.vy + fr.value, fr.value, frame))
    return np.array(out)


def star_only(frrot: int):
    # plnt_flag 1: the planet feels the star only (no disk reduction), M_p tiny
    return configs.disk2d(8, 24).replace(plnt_flag=1, frrot_flag=frrot, M_p=1e-6, t_growth=0.0)


def start(x, y, vx, omega):
    return PnBody(m=1e-6, x=x, y=y, vx=vx, vy=omega, fx=0.0, fy=0.0, rs=0.0)


def test_circular_orbit_stays_circular_in_both_frames():
    dt, n = 0.01, 700   # a bit more than one orbit at r = 1
    a = kdk_orbit(star_only(0), start(1.0, math.pi, 0.0, 1.0), 0.0, dt, n)
    assert np.abs(a[:, 0] - 1.0).max() < 1e-13
    assert np.abs(a[:, 2]).max() < 1e-13
    assert np.abs(a[:, 3] - 1.0).max() < 1e-13
    # the azimuth advances at the orbital rate, wrapped into [0, 2 pi)
    want = np.mod(math.pi + dt * np.arange(1, n + 1), TWO_PI)
    assert np.abs(a[:, 1] - want).max() < 1e-11
    # co-rotating: the frame turns with the planet, which stays at pi
    b = kdk_orbit(star_only(1), start(1.0, math.pi, 0.0, 0.0), 1.0, dt, n)
    assert np.abs(b[:, 1] - math.pi).max() < 1e-13
    assert np.abs(b[:, 4] - 1.0).max() < 1e-13


def kepler(a):
    x, vx, om = a[:, 0], a[:, 2], a[:, 3]
    E = 0.5 * (vx * vx + (x * om) ** 2) - 1.0 / x
    Lz = x * x * om
    return E, Lz


def test_eccentric_orbit_conserves_angular_momentum_and_energy_to_second_order():
    x0, vx0, om0 = 1.0, 0.15, 1.1
    E0 = 0.5 * (vx0 ** 2 + (x0 * om0) ** 2) - 1.0 / x0
    L0 = x0 * x0 * om0
    sma = -0.5 / E0
    period = TWO_PI * sma ** 1.5
    errs, closes = [], []
    for n in (400, 800, 1600):
        dt = period / n
        a = kdk_orbit(star_only(0), start(x0, 1.0, vx0, om0), 0.0, dt, n)
        E, Lz = kepler(a)
        # the drift carries L = x^2 Omega across the radius change and the kicks add no torque
        assert np.abs(Lz - L0).max() < 1e-13 * L0
        errs.append(np.abs(E - E0).max())
        # after one Kepler period the orbit closes: back at x0 with vx0, one turn further
        closes.append(max(abs(a[-1, 0] - x0), abs(a[-1, 2] - vx0),
                          abs(math.remainder(a[-1, 1] - 1.0, TWO_PI))))
    assert errs[0] < 1e-3
    assert errs[0] / errs[1] > 3.6 and errs[1] / errs[2] > 3.6
    assert closes[0] < 2e-2
    assert closes[0] / closes[1] > 3.6 and closes[1] / closes[2] > 3.6


def test_corotating_frame_reproduces_the_inertial_orbit():
    """With the frame turning at the planet's rate, x, vx and the inertial angular speed are
    the inertial run's, the planet stays at its azimuth, and the frame angle sum(FrRot dt)
    is the inertial run's azimuth advance."""
    dt, n = 0.005, 1500
    a = kdk_orbit(star_only(0), start(1.0, math.pi, 0.15, 1.1), 0.0, dt, n)
    b = kdk_orbit(star_only(1), start(1.0, math.pi, 0.15, 0.0), 1.1, dt, n)
    assert np.abs(b[:, 0] - a[:, 0]).max() < 1e-12
    assert np.abs(b[:, 2] - a[:, 2]).max() < 1e-12
    assert np.abs(b[:, 3] - a[:, 3]).max() < 1e-12
    assert np.abs(b[:, 1] - math.pi).max() < 1e-13
    turned = math.pi + np.cumsum(b[:, 5] * dt)
    assert np.abs(np.mod(turned, TWO_PI) - a[:, 1]).max() < 1e-10


# ---------------------------------------------------------------------------------------
# disk forces on the planet, modes 8 and 9

def potential(q, xp, yp, rs, bary, r, phi):
    """The softened planet potential the sweeps use (forces.cu:4-21): with the indirect term
    of the star's reflex motion, or in the barycentric frame with the star's own pull."""
    c = np.cos(phi - yp)
    if bary:
        stm = 1.0 / (1.0 + q)
        plm, plx, stx = q * stm, xp * stm, xp * q * stm
        Rp = np.sqrt(r * r + plx * plx - 2.0 * plx * r * c + rs * rs)
        Rs = np.sqrt(r * r + stx * stx + 2.0 * stx * r * c)
        return -plm / Rp - stm / Rs
    return -q / np.sqrt(r * r + xp * xp - 2.0 * xp * r * c + rs * rs) + q * r * c / (xp * xp)


def brute_force(o, b: PnBody, exclude: float, bary: int):
    P = o.params
    R = o.rings
    J = o.grid.jmax
    h = 1e-6
    m8 = m9 = 0.0
    for n in range(len(R)):
        r = R["xc"][n]
        j0 = (np.arange(J) + R["rot_j"][n]) % J
        phi = R["yc"][n][j0]
        dsig = R["r"][n] - r ** -P.p_alpha
        vol = R["xvol"][n] * R["yvol"][n][j0]
        keep = (r - b.x) ** 2 + r * r * (phi - b.y) ** 2 > (exclude * P.sc_h) ** 2
        fr = -(potential(b.m, b.x, b.y, b.rs, bary, r + h, phi) - potential(b.m, b.x, b.y, b.rs, bary, r - h, phi)) / (2 * h)
        fp = -(potential(b.m, b.x, b.y, b.rs, bary, r, phi + h) - potential(b.m, b.x, b.y, b.rs, bary, r, phi - h)) / (2 * h) / r
        m8 += np.sum(np.where(keep, -fr * dsig * vol, 0.0))
        m9 += np.sum(np.where(keep, -fp * r * dsig * vol, 0.0))
    return m8, m9


@pytest.mark.parametrize("bary", [0, 1])
@pytest.mark.parametrize("exclude", [0.0, 4.0])
def test_disk_forces_on_the_planet_brute_force(bary, exclude):
    w = configs.small(20, 60, plnt_flag=2, bary_flag=bary, M_p=3e-4)
    o = oracle.OracleDisk(w.grid, w.params, w.seed, 0.0)
    rng = np.random.default_rng(5)
    o.rings["r"] *= 1.0 + 0.05 * rng.uniform(-1, 1, o.rings["r"].shape)
    o.rings["rot_j"] = rng.integers(0, w.grid.jmax, len(o.rings))
    b = PnBody(m=3e-4, x=1.03, y=2.2, vx=0.0, vy=0.0, fx=0.0, fy=0.0, rs=0.02)
    got8 = o.reduce(b, exclude, 8)
    got9 = o.reduce(b, exclude, 9)
    want8, want9 = brute_force(o, b, exclude, bary)
    assert abs(got8 - want8) <= 1e-6 * abs(want8)
    assert abs(got9 - want9) <= 1e-6 * abs(want9)
    if exclude == 0.0:
        # the unperturbed disk (Sigma = r^-p_alpha) exerts no force in either mode
        o0 = oracle.OracleDisk(w.grid, w.params, w.seed, 0.0)
        assert o0.reduce(b, 0.0, 8) == 0.0 and o0.reduce(b, 0.0, 9) == 0.0


def test_mirror_symmetric_density_has_no_azimuthal_force():
    w = configs.small(24, 72, plnt_flag=2, M_p=1e-3)
    o = oracle.OracleDisk(w.grid, w.params, w.seed, 0.0)
    J = w.grid.jmax
    rng = np.random.default_rng(11)
    # lab columns j and J-1-j lie at pi -+ the same angle on the uniform grid
    half = 1.0 + 0.1 * rng.uniform(-1, 1, (len(o.rings), J // 2))
    lab = np.concatenate([half, half[:, ::-1]], axis=1)
    o.rings["r"] *= lab   # rot_j = 0: data index = lab column
    b = PnBody(m=1e-3, x=1.0, y=math.pi, vx=0.0, vy=0.0, fx=0.0, fy=0.0, rs=0.03)
    m8, m9 = o.reduce(b, 0.0, 8), o.reduce(b, 0.0, 9)
    w9 = brute_force(o, b, 0.0, 0)[1]
    # the sum of |terms| sets the rounding scale of mode 9
    scale = abs(brute_force(o, b, 0.0, 0)[0])
    assert abs(m8) > 1e-6
    assert abs(m9) < 1e-13 * scale and abs(w9) < 1e-9 * scale
    # the same field with the planet moved off the symmetry axis feels a torque
    b.y = math.pi + 0.3
    assert abs(o.reduce(b, 0.0, 9)) > 1e-3 * scale

