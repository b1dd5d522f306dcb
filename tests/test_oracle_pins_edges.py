"""Pins of the oracle's radial boundary codes (set_boundx, boundary.cu:86-206; the oracle's
ghost_values) by what each code implies for the flow through a planar radial sweep
(ngeomx 0: no gravity, no curvature):

- 0 reflect on both sides: the mass between the two mirror planes (the edge cells' centres,
  about which the ghost cells mirror the interior) is conserved to rounding, and a state
  mirror-symmetric about the middle of the domain stays symmetric;
- 1 outflow: a uniform flow stays uniform whatever the fixed ghost rings hold, and a
  pulse leaves the domain with little reflected;
- 2 periodic in radius: mass is conserved to rounding and a pulse carried once across the
  domain comes back to its start, the error falling at second order;
- 4 outflow of r, p, u on the left but of r, p only on the right: a uniform flow stays
  uniform at the left edge and not at the right, and matter flowing in from the left
  carries the fixed ghost ring's transverse velocity.
"""
from __future__ import annotations

import numpy as np

import oracle
from penguin_2_0_amd import configs
from penguin_2_0_amd.abi import PnBody, PnGrid


def tube(n, nlft, nrgh, eos=0, x0=0.0, nphi=12):
    """A planar tube of n cells on [x0, x0 + 1] with the given boundary codes; isothermal
    with c = 1 (eos 0) or adiabatic with gamma 1.4 (eos 3)."""
    P = configs.disk2d(n, nphi).replace(ngeomx=0, plnt_flag=0, fargo_flag=0, visc_flag=0, frrot_flag=0,
                                        grspx=0, xmin=x0, xmax=x0 + 1.0, sc_h=1.0, p_beta=0.0,
                                        nlft=nlft, nrgh=nrgh)
    if eos:
        P = P.replace(eos=eos, gam=1.4, init_p_iso=1)
    o = oracle.OracleDisk(PnGrid(n, nphi, 1), P, 1, 0.0)
    for A in (o.rings, o.lft, o.rgh):
        A["u"] = 0.0
        A["v"] = 0.0
        A["w"] = 0.0
        A["rot_j"] = 0
        A["res_v"] = 0.0
    return P, o


def set_fields(o, rho, u, p=None):
    R = o.rings
    R["r"] = rho[:, None]
    R["u"] = u[:, None]
    R["p"] = (rho if p is None else p)[:, None]


def sweeps(o, dt, n):
    for _ in range(n):
        o.sweepx(dt, PnBody(), 0.0)


def test_reflecting_edges_conserve_mass_between_the_mirror_planes():
    n = 160
    P, o = tube(n, 0, 0)
    x = o.rings["xc"]
    rho = 1.0 + 0.4 * np.exp(-((x - 0.2) / 0.06) ** 2) + 0.2 * np.exp(-((x - 0.75) / 0.1) ** 2)
    # the mirror image of the edge cell is itself, so its u must vanish (and then stays 0)
    s = (x - x[0]) / (x[-1] - x[0])
    u = 0.3 * np.sin(3.0 * np.pi * s)
    u[0] = u[-1] = 0.0
    set_fields(o, rho, u)
    wt = np.ones(n)
    wt[0] = wt[-1] = 0.5

    def mass():
        return (o.rings["r"][:, 0] * o.rings["xvol"] * wt).sum()

    def plain():
        return (o.rings["r"][:, 0] * o.rings["xvol"]).sum()

    m0, p0 = mass(), plain()
    sweeps(o, 0.4 / n, 200)   # waves reach and leave both walls
    assert abs(mass() - m0) < 1e-14 * m0
    # the plain sum over the cells is not conserved: mass crossed the edge cells' outer faces
    assert abs(plain() - p0) > 1e-6 * p0


def test_reflecting_edges_keep_mirror_symmetry():
    n = 128
    P, o = tube(n, 0, 0)
    x = o.rings["xc"]
    rho = 1.0 + 0.5 * np.exp(-((x - 0.3) / 0.05) ** 2) + 0.5 * np.exp(-((x - 0.7) / 0.05) ** 2)
    u = 0.2 * np.sin(2.0 * np.pi * x) * np.exp(-((x - 0.5) / 0.3) ** 2)   # odd about x = 1/2
    set_fields(o, rho, u)
    sweeps(o, 0.4 / n, 150)
    r, v = o.rings["r"][:, 0], o.rings["u"][:, 0]
    assert np.abs(r - r[::-1]).max() < 1e-12
    assert np.abs(v + v[::-1]).max() < 1e-12
    assert np.abs(r - rho).max() > 0.05   # the flow did move


def test_outflow_keeps_a_uniform_flow_against_different_ghost_rings():
    P, o = tube(96, 1, 1)
    for A in (o.lft, o.rgh):
        A["r"] = 3.0
        A["p"] = 3.0
        A["u"] = -0.4
    set_fields(o, np.full(96, 1.5), np.full(96, 0.35))
    sweeps(o, 0.3 / 96, 40)
    assert np.abs(o.rings["r"] - 1.5).max() < 1e-13
    assert np.abs(o.rings["u"] - 0.35).max() < 1e-13


def test_outflow_lets_a_pulse_leave():
    n = 128
    P, o = tube(n, 1, 1, eos=3, x0=0.0)
    x = o.rings["xc"]
    amp = 0.3
    rho = 1.0 + amp * np.exp(-((x - 0.5) / 0.05) ** 2)
    set_fields(o, rho, np.full(n, 1.0), np.full(n, 1.0))   # a contact moving right at u = 1
    sweeps(o, 0.2 / n, int(0.8 / (0.2 / n)))
    left = np.abs(o.rings["r"][:, 0] - 1.0).max()
    assert left < 0.02 * amp
    # the same pulse between reflecting walls is still there
    P, o2 = tube(n, 0, 0, eos=3, x0=0.0)
    set_fields(o2, rho, np.full(n, 1.0), np.full(n, 1.0))
    sweeps(o2, 0.2 / n, int(0.8 / (0.2 / n)))
    assert np.abs(o2.rings["r"][:, 0] - 1.0).max() > 0.3 * amp


def _periodic_pulse_error(n):
    P, o = tube(n, 2, 2, eos=3)
    x = o.rings["xc"]
    rho = 1.0 + 0.2 * np.exp(-((x - 0.5) / 0.08) ** 2)
    set_fields(o, rho, np.full(n, 1.0), np.full(n, 1.0))
    m0 = (o.rings["r"][:, 0] * o.rings["xvol"]).sum()
    steps = 2 * n   # once across at u = 1: dt = 1 / (2 n)
    sweeps(o, 1.0 / steps, steps)
    m1 = (o.rings["r"][:, 0] * o.rings["xvol"]).sum()
    return np.abs(o.rings["r"][:, 0] - rho).mean(), abs(m1 - m0) / m0


def test_periodic_radius_carries_a_pulse_back_to_its_start():
    e1, dm1 = _periodic_pulse_error(64)
    e2, dm2 = _periodic_pulse_error(128)
    assert dm1 < 1e-14 and dm2 < 1e-14
    assert e1 < 5e-3
    assert e1 / e2 > 3.5


def test_code4_copies_u_on_the_left_only():
    n = 96
    P, o = tube(n, 4, 4)
    for A in (o.lft, o.rgh):
        A["r"] = 1.0
        A["p"] = 1.0
        A["u"] = 0.0
    set_fields(o, np.full(n, 2.0), np.full(n, 0.3))
    sweeps(o, 0.3 / n, 20)
    r, u = o.rings["r"][:, 0], o.rings["u"][:, 0]
    # left: r, p and u copied from the edge cell, so no jump there
    assert np.abs(r[: n // 3] - 2.0).max() < 1e-13 and np.abs(u[: n // 3] - 0.3).max() < 1e-13
    # right: u from the fixed ghost ring (0) against 0.3 inside
    assert np.abs(u[-3:] - 0.3).max() > 1e-3


def test_code4_inflow_carries_the_ghost_rings_transverse_velocity():
    n = 96
    P, o = tube(n, 4, 4)
    for A in (o.lft, o.rgh):
        A["v"] = 0.7
        A["w"] = -0.2
    set_fields(o, np.full(n, 1.0), np.full(n, 0.5))
    steps = 60   # the inflow fills x < 0.5 * 60 * 0.3 / 96 ~ 0.09
    sweeps(o, 0.3 / n, steps)
    v, w = o.rings["v"][:, 0], o.rings["w"][:, 0]
    assert np.abs(v[:4] - 0.7).max() < 1e-3 and np.abs(w[:4] + 0.2).max() < 1e-3
    # (the last cells see the right side's fixed ghost v through the reconstruction)
    assert np.abs(v[n // 2:n - 10]).max() < 1e-13
    # outflow (code 1) copies the edge cell's v = 0 instead
    P, o1 = tube(n, 1, 1)
    for A in (o1.lft, o1.rgh):
        A["v"] = 0.7
    set_fields(o1, np.full(n, 1.0), np.full(n, 0.5))
    sweeps(o1, 0.3 / n, steps)
    assert np.abs(o1.rings["v"]).max() < 1e-13
