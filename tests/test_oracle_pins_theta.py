"""Pins of the oracle's theta sweep (orc_sweepz), which the reference holds only as pieces
(states.cu:29-39, evolve.cu:26-29,63-67,110-124, remap.cu:50-54, forces.cu:298-372,
init.cpp:415-541) with no driver (main.cu:376 calls an undefined bundle_sweep3D), against
what the mathematics fixes:

- spherical theta (ngeomz 5): a uniform state at rest stays at rest; the mass sum of
  rho * dvol over a pencil changes only by what crosses its ends; a rotating atmosphere with
  the same specific angular momentum l on every theta ring, rho = exp(-l^2 cot^2(theta) /
  (2 c^2 r^2)), is an exact equilibrium of the theta momentum equation (pressure against the
  centrifugal term l^2 cos(theta) / (r sin(theta))^3), and the v_theta one sweep creates
  falls at second order with the theta resolution;
- cylindrical z (ngeomz 0) far from the star: a shock tube along z converges to the exact
  isothermal solution, which pins the pencil gather over theta rings, the ghost rings below
  and above, and the component order of the sweep.
"""
from __future__ import annotations

import math

import numpy as np

import oracle
from penguin_2_0_amd import configs
from penguin_2_0_amd.abi import NPAD, PnBody, PnGrid

from test_oracle_pins import exact_isothermal

HPI = 0.5 * math.pi


def theta_disk(nth, half=0.3, nr=2, nphi=12, c=0.1, r0=2.0, th0=HPI, **over):
    """Thin shell r in [r0, 1.001 r0], theta in th0 -+ half, no planet, no frame rotation."""
    P = configs.disk2d(nr, nphi).replace(ndim=3, ngeomx=2, ngeomy=4, ngeomz=5, plnt_flag=0, fargo_flag=0,
                                         visc_flag=0, frrot_flag=0, p_beta=0.0, sc_h=c, xmin=r0,
                                         xmax=1.001 * r0, zmin=th0 - half, zmax=th0 + half).replace(**over)
    o = oracle.OracleDisk(PnGrid(nr, nphi, nth), P, 1, 0.0)
    for A in (o.rings, o.lft, o.rgh, o.udr, o.top):
        A["rot_j"] = 0
        A["res_v"] = 0.0
    return P, o


def all_rings(o):
    return (o.rings, o.udr, o.top)


def test_theta_sweep_keeps_a_uniform_state_at_rest():
    P, o = theta_disk(40)
    for A in all_rings(o):
        A["r"] = 1.3
        A["p"] = 1.3 * 0.01
        A["u"] = 0.0
        A["v"] = 0.0
        A["w"] = 0.0
    for _ in range(5):
        o.sweepz(0.01, PnBody(), 0.0)
    assert np.abs(o.rings["r"] - 1.3).max() < 1e-14
    assert np.abs(o.rings["w"]).max() < 1e-15


def test_theta_sweep_conserves_mass_away_from_the_ends():
    nth = 64
    P, o = theta_disk(nth)
    rng = np.random.default_rng(2)
    R = o.rings
    for A in all_rings(o):
        A["r"] = 1.0
        A["p"] = 0.01
        A["u"] = 0.0
        A["v"] = 0.0
        A["w"] = 0.0
    th = R["zc"]
    bump = np.exp(-((th - HPI) / 0.05) ** 2)
    R["r"] = 1.0 + bump[:, None] * (0.5 + 0.01 * rng.uniform(-1, 1, R["r"].shape))
    R["p"] = R["r"] * 0.01
    R["w"] = (0.05 * bump)[:, None]
    vol = R["xvol"] * R["zvol"]

    def mass():
        return ((R["r"] * R["yvol"]).sum(axis=1) * vol).sum()

    m0 = mass()
    for _ in range(10):
        o.sweepz(0.02, PnBody(), 0.0)
    assert abs(mass() - m0) < 1e-14 * m0
    assert np.abs(R["w"]).max() > 1e-3   # it moved


def _equilibrium_residual(nth, L=2.0, c=0.1):
    P, o = theta_disk(nth, c=c)
    for A in all_rings(o):
        rad, th = A["xc"], A["zc"]
        rho = np.exp(-(L * L / (2.0 * c * c * rad * rad)) / np.tan(th) ** 2)
        A["r"] = rho[:, None]
        A["p"] = (c * c * rho)[:, None]
        A["u"] = 0.0
        A["v"] = (L / (rad * np.sin(th)))[:, None]
        A["w"] = 0.0
    dth = o.rings["dz"][0]
    dt = 0.2 * dth / c
    o.sweepz(dt, PnBody(), 0.0)
    a = np.abs(o.rings["w"][:, 0]) / dt
    off = np.abs(o.rings["zc"] - HPI)
    band = (off > 0.08) & (off < 0.22)
    return np.median(a), a[band].max(), a[off < 0.05].max()


def test_theta_sweep_rotating_equilibrium_residual_converges():
    """Median and monotone-band residuals fall at second order.  At the density maximum on
    the equator the parabolae are flattened (the monotonicity constraint of parabola.cu at a
    local extremum), which leaves a first-order residual there."""
    (m1, b1, e1), (m2, b2, e2), (m3, b3, e3) = (_equilibrium_residual(n) for n in (32, 64, 128))
    # the centrifugal acceleration itself is ~ l^2 cos / (r sin)^3 ~ 0.15 at the pencil ends
    assert b1 < 1e-4 and e1 < 1e-2
    assert m1 / m2 > 3.5 and m2 / m3 > 3.5
    assert b1 / b2 > 3.5 and b2 / b3 > 3.5
    assert e1 / e2 > 1.8 and e2 / e3 > 1.8


def test_theta_sweep_without_rotation_leaves_the_stratified_state_unbalanced():
    """The same density with l = 0 is not in equilibrium: the pressure gradient accelerates
    the gas by about c^2 d ln(rho) / (r d theta) -- the residual stays O(1) as theta is
    refined, so the balance above is the centrifugal term's work."""
    P, o = theta_disk(64)
    c, L = 0.1, 2.0
    for A in all_rings(o):
        rad, th = A["xc"], A["zc"]
        rho = np.exp(-(L * L / (2.0 * c * c * rad * rad)) / np.tan(th) ** 2)
        A["r"] = rho[:, None]
        A["p"] = (c * c * rho)[:, None]
        A["u"] = 0.0
        A["v"] = 0.0
        A["w"] = 0.0
    dt = 0.2 * o.rings["dz"][0] / c
    o.sweepz(dt, PnBody(), 0.0)
    assert np.abs(o.rings["w"][:, 0]).max() / dt > 0.1


def _z_shock_tube(n):
    """Isothermal shock tube along z in cylindrical geometry (ngeomz 0) at R = 1e6, where
    the star's vertical pull -z / (R^2 + z^2)^1.5 is 1e-18."""
    P = configs.disk2d(2, 12).replace(ndim=3, ngeomx=1, ngeomy=3, ngeomz=0, plnt_flag=0, fargo_flag=0,
                                      visc_flag=0, frrot_flag=0, grspx=0, grspz=0, p_beta=0.0, sc_h=1.0,
                                      xmin=1.0e6, xmax=1.0e6 + 1.0e-3, zmin=0.0, zmax=1.0)
    o = oracle.OracleDisk(PnGrid(2, 12, n), P, 1, 0.0)
    rl, rr = 1.0, 0.25
    for A in all_rings(o):
        left = A["zc"] < 0.5
        A["r"] = np.where(left, rl, rr)[:, None]
        A["p"] = A["r"] * 1.0
        A["u"] = 0.0
        A["v"] = 0.0
        A["w"] = 0.0
        A["rot_j"] = 0
        A["res_v"] = 0.0
    t_end = 0.15
    nstep = int(t_end / (0.3 / n)) + 1
    for _ in range(nstep):
        o.sweepz(t_end / nstep, PnBody(), 0.0)
    zc = o.rings["zc"].reshape(n, 2)[:, 0]
    rho = o.rings["r"].reshape(n, 2, 12)[:, 0, 0]
    ue, pe = exact_isothermal(1.0, rl, 0.0, rr, 0.0)
    S = pe * ue / (pe - rr)
    xi = (zc - 0.5) / t_end
    rex = np.where(xi < -1.0, rl, np.where(xi < ue - 1.0, rl * np.exp(-xi - 1.0), np.where(xi < S, pe, rr)))
    vz = o.rings["w"].reshape(n, 2, 12)[:, 0, 0]
    return np.abs(rho - rex).mean(), vz


def test_z_sweep_shock_tube_converges_to_the_exact_solution():
    e1, _ = _z_shock_tube(200)
    e2, vz = _z_shock_tube(400)
    assert e1 < 5e-3 and e2 < 0.75 * e1
    assert vz.min() > -1e-12 and vz.max() > 0.2   # the gas moves towards +z only


def steady_outflow(th, c, th0=0.75, u0=0.03):
    """Steady isothermal flow along theta on a sphere with no forces: rho u sin(theta) = F
    (continuity) and u^2 / 2 + c^2 ln(rho) = B (Bernoulli), subsonic branch through
    rho = 1, u = u0 at th0."""
    F = u0 * math.sin(th0)
    B = 0.5 * u0 * u0
    out_u = np.empty_like(th)
    for n, t in enumerate(th):
        lo, hi = 1e-12, c   # the subsonic root: g(u) decreases from +inf to g(c) < 0
        for _ in range(200):
            u = 0.5 * (lo + hi)
            g = 0.5 * u * u + c * c * math.log(F / (u * math.sin(t))) - B
            if g > 0.0:
                lo = u
            else:
                hi = u
        out_u[n] = 0.5 * (lo + hi)
    return F / (out_u * np.sin(th)), out_u


def _outflow_residual(nth, c=0.1):
    P, o = theta_disk(nth, c=c, th0=0.75, half=0.25)
    for A in all_rings(o):
        rho, u = steady_outflow(A["zc"], c)
        A["r"] = rho[:, None]
        A["p"] = (c * c * rho)[:, None]
        A["u"] = 0.0
        A["v"] = 0.0
        A["w"] = u[:, None]
    rad = o.rings["xc"][0]
    dt = 0.3 * o.rings["dz"][0] * rad / c
    w0, r0 = o.rings["w"][:, 0].copy(), o.rings["r"][:, 0].copy()
    o.sweepz(dt, PnBody(), 0.0)
    off = np.abs(o.rings["zc"] - 0.75)
    band = off < 0.2
    return (np.abs(o.rings["w"][:, 0] - w0)[band].max() / dt,
            np.abs(o.rings["r"][:, 0] / r0 - 1.0)[band].max() / dt)


def test_theta_sweep_steady_outflow_residual_converges():
    """A steady subsonic flow along theta off the equator (cot(theta) ~ 1): at a fixed
    Courant number the density residual falls at second order; the velocity residual falls
    at first order, being proportional to dt (the radial sweep leaves the same O(dt) residual
    on a steady wind, so it is the Lagrange-plus-remap split's, not the theta geometry's)."""
    (u1, r1), (u2, r2), (u3, r3) = (_outflow_residual(n) for n in (32, 64, 128))
    assert u1 < 5e-5 and r1 < 1e-5
    assert u1 / u2 > 1.8 and u2 / u3 > 1.8
    assert r1 / r2 > 3.5 and r2 / r3 > 3.5
