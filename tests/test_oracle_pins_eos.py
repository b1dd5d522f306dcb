"""Pins of the oracle's isentropic (EOS 1) and adiabatic (EOS 2) sweeps (riemann.cu:55-108,
evolve.cu:129-160, remap.cu:144-176, the EOS closure of ppm.cu:138-144) against exact
solutions:

- the EOS 1 Riemann iteration solves the two-shock relations of a gamma-law gas, which are
  exact when both waves are shocks: for colliding streams its p* and u* are the exact
  gamma-law solution's;
- an EOS 1 shock tube (p = K rho^gamma after every sweep: mass and momentum only) converges
  to the exact barotropic Riemann solution, which differs from the gamma-law one in the
  shock's jump conditions;
- an EOS 2 shock tube (Sod's problem), run far from the star so that the potential the
  reference's energy update subtracts is negligible and with no transverse motion, converges
  to the exact gamma-law solution.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from penguin_2_0_amd import configs
from penguin_2_0_amd.abi import PnBody, PnGrid

from test_oracle_pins import exact_gamma_law, riem


@pytest.mark.parametrize("case", [
    (1.0, 1.0, 1.0, 1.0, -1.0, 1.0),
    (1.0, 0.5, 1.0, 0.25, -2.0, 0.3),
    (0.4, 3.0, 2.0, 1.5, 0.0, 0.5),
    (2.0, 2.0, 5.0, 1.0, -1.5, 1.0),
])
def test_isentropic_riemann_equals_the_exact_two_shock_solution(case):
    rl, ul, pl, rr, ur, pr = case
    gam = 1.4
    um, pm = riem(1, gam, pl, rl, ul, pr, rr, ur)
    # the exact gamma-law solution at the contact (both sides must be shocks for the case)
    lo = exact_gamma_law(rl, ul, pl, rr, ur, pr, gam, np.array([um - 1e-9, um + 1e-9]))
    assert lo[2, 0] > max(pl, pr)
    assert abs(pm - lo[2, 0]) < 1e-7 * pm
    assert abs(um - lo[1, 0]) < 1e-7 * max(1.0, abs(um))


def exact_barotropic(rl, ul, rr, ur, gam, K, s):
    """Exact Riemann solution of the isentropic gas p = K rho^gam (mass and momentum
    conserved, no energy equation): left rarefaction or shock, right rarefaction or shock;
    sampled at s = x / t.  Returns rho, u."""
    c = lambda r: math.sqrt(gam * K * r ** (gam - 1.0))
    p = lambda r: K * r ** gam

    def f(rs, rk):   # velocity change across the wave facing state k
        if rs > rk:
            return math.sqrt((p(rs) - p(rk)) * (rs - rk) / (rs * rk))
        return 2.0 / (gam - 1.0) * (c(rs) - c(rk))

    lo, hi = 1e-9, 50.0 * max(rl, rr)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid, rl) + f(mid, rr) + ur - ul > 0.0:
            hi = mid
        else:
            lo = mid
    rs = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (f(rs, rr) - f(rs, rl))
    out = np.zeros((2, len(s)))
    for n, x in enumerate(s):
        side = -1.0 if x < us else 1.0
        rk, uk = (rl, ul) if side < 0 else (rr, ur)
        if rs > rk:   # shock: mass flux across it
            S = (rs * us - rk * uk) / (rs - rk)
            inside = (x > S) if side < 0 else (x < S)
            out[:, n] = (rs, us) if inside else (rk, uk)
        else:         # rarefaction
            ck, cs = c(rk), c(rs)
            head, tail = uk + side * ck, us + side * cs
            if (x < head) if side < 0 else (x > head):
                out[:, n] = (rk, uk)
            elif (x > tail) if side < 0 else (x < tail):
                out[:, n] = (rs, us)
            else:
                cc = 2.0 / (gam + 1.0) * (ck - side * 0.5 * (gam - 1.0) * (uk - x))
                u = x - side * cc   # u - c = s on the left fan, u + c = s on the right
                out[:, n] = (rk * (cc / ck) ** (2.0 / (gam - 1.0)), u)
    return out


def test_exact_barotropic_solution_satisfies_its_relations():
    """The test's own solver: across the shock rho u and rho u^2 + p are conserved in the
    shock's frame, and the rarefaction keeps u + 2c/(gam - 1)."""
    gam, K = 1.4, 1.0 / 1.4
    s = np.array([-5.0, 0.0, 0.5, 5.0])
    ex = exact_barotropic(1.0, 0.0, 0.2, 0.0, gam, K, s)
    r2, u2 = ex[:, 2]
    S = (r2 * u2) / (r2 - 0.2)
    assert abs((r2 * (u2 - S)) - (0.2 * (0.0 - S))) < 1e-12
    assert abs((r2 * (u2 - S) ** 2 + K * r2 ** gam) - (0.2 * S * S + K * 0.2 ** gam)) < 1e-9
    c = lambda r: math.sqrt(gam * K * r ** (gam - 1.0))
    assert abs(u2 + 2 * c(r2) / (gam - 1) - 2 * c(1.0) / (gam - 1)) < 1e-9


def _isentropic_tube(n, gam=1.4):
    P = configs.disk2d(n, 12).replace(ngeomx=0, plnt_flag=0, fargo_flag=0, visc_flag=0, frrot_flag=0,
                                      grspx=0, xmin=0.0, xmax=1.0, sc_h=1.0, p_beta=0.0, eos=1, gam=gam)
    o = oracle.OracleDisk(PnGrid(n, 12, 1), P, 1, 0.0)
    K = 1.0 / gam   # the closure p = c_s^2 rho^gam / gam with c_s = 1
    rl, rr = 1.0, 0.2
    for A in (o.rings, o.lft, o.rgh):
        rho = np.where(A["xc"] < 0.5, rl, rr)
        A["r"] = rho[:, None]
        A["p"] = (K * rho ** gam)[:, None]
        A["u"] = 0.0
        A["v"] = 0.0
        A["rot_j"] = 0
        A["res_v"] = 0.0
    t_end = 0.2
    nstep = int(t_end / (0.3 / n)) + 1
    for _ in range(nstep):
        o.sweepx(t_end / nstep, PnBody(), 0.0)
    s = (o.rings["xc"] - 0.5) / t_end
    ex = exact_barotropic(rl, 0.0, rr, 0.0, gam, K, s)
    gl = exact_gamma_law(rl, 0.0, K * rl ** gam, rr, 0.0, K * rr ** gam, gam, s)
    rho, u = o.rings["r"][:, 0], o.rings["u"][:, 0]
    assert np.allclose(o.rings["p"][:, 0], K * rho ** gam, rtol=1e-14)
    return (np.abs(rho - ex[0]).mean(), np.abs(u - ex[1]).mean(),
            np.abs(rho - gl[0]).mean(), np.abs(u - gl[1]).mean())


def test_isentropic_shock_tube_converges_to_the_barotropic_solution():
    (r1, u1, _, _), (r2, u2, gr2, gu2) = _isentropic_tube(200), _isentropic_tube(400)
    assert r2 < 3e-3 and u2 < 3e-3
    assert r2 < 0.75 * r1 and u2 < 0.75 * u1
    # the gamma-law solution (energy conserved across the shock) is measurably further away
    assert gr2 > 2.5 * r2 and gu2 > 2.0 * u2


def _sod_eos2(n, gam=1.4, x0=1.0e6):
    """Sod's problem in a planar tube at x0 = 1e6 from the star: the potential the EOS 2
    energy update subtracts (evolve.cu:149-157) is -1/x0 = -1e-6 there, uniform to 1e-12."""
    P = configs.disk2d(n, 12).replace(ngeomx=0, plnt_flag=0, fargo_flag=0, visc_flag=0, frrot_flag=0,
                                      grspx=0, xmin=x0, xmax=x0 + 1.0, eos=2, gam=gam, init_p_iso=1)
    o = oracle.OracleDisk(PnGrid(n, 12, 1), P, 1, 0.0)
    for A in (o.rings, o.lft, o.rgh):
        left = A["xc"] < x0 + 0.5
        A["r"] = np.where(left, 1.0, 0.125)[:, None] * np.ones(12)
        A["p"] = np.where(left, 1.0, 0.1)[:, None] * np.ones(12)
        A["u"] = 0.0
        A["v"] = 0.0
        A["rot_j"] = 0
        A["res_v"] = 0.0
    t_end = 0.15
    nstep = int(t_end / (0.3 / n / 1.8)) + 1
    for _ in range(nstep):
        o.sweepx(t_end / nstep, PnBody(), 0.0)
    ex = exact_gamma_law(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, gam, (o.rings["xc"] - x0 - 0.5) / t_end)
    return (np.abs(o.rings["r"][:, 0] - ex[0]).mean(), np.abs(o.rings["p"][:, 0] - ex[2]).mean(),
            np.abs(o.rings["u"][:, 0] - ex[1]).mean())


def test_adiabatic_eos2_shock_tube_converges_to_the_exact_solution():
    e1, e2 = _sod_eos2(100), _sod_eos2(200)
    assert e2[0] < 4e-3 and e2[1] < 4e-3 and e2[2] < 8e-3
    assert all(b < 0.75 * a for a, b in zip(e1, e2))


def toro_f(p, rk, pk, gam):
    """Velocity change across the wave facing state k at pressure p (Toro, eq. 4.6-4.7):
    shock branch (p - pk) sqrt(A / (p + B)), A = 2 / ((gam + 1) rk), B = (gam - 1) pk / (gam + 1);
    rarefaction branch 2 c_k / (gam - 1) ((p / pk)^((gam - 1) / (2 gam)) - 1)."""
    if p > pk:
        return (p - pk) * math.sqrt(2.0 / ((gam + 1.0) * rk) / (p + (gam - 1.0) / (gam + 1.0) * pk))
    ck = math.sqrt(gam * pk / rk)
    return 2.0 * ck / (gam - 1.0) * ((p / pk) ** ((gam - 1.0) / (2.0 * gam)) - 1.0)


@pytest.mark.parametrize("case", [
    (1.0, 0.0, 1.0, 0.125, 0.0, 0.1),     # Sod: a shock to the right
    (0.125, 0.0, 0.1, 1.0, 0.0, 1.0),     # mirrored: a shock to the left
    (1.0, 1.0, 1.0, 1.0, -1.0, 1.0),      # colliding: two shocks
    (1.0, 0.3, 2.0, 0.5, 0.1, 0.4),
])
def test_adiabatic_riemann_velocity_follows_the_wave_relations_of_its_pressure(case):
    """EOS 2 / 3 take p* from the two-rarefaction closed form and then u* from the exact
    wave relation of each side at that p* (shock or rarefaction): u* = (ul + ur)/2 +
    (f_R(p*) - f_L(p*))/2."""
    rl, ul, pl, rr, ur, pr = case
    gam = 1.4
    for eos in (2, 3):
        um, pm = riem(eos, gam, pl, rl, ul, pr, rr, ur)
        want = 0.5 * (ul + ur) + 0.5 * (toro_f(pm, rr, pr, gam) - toro_f(pm, rl, pl, gam))
        assert abs(um - want) < 1e-13 * max(1.0, abs(um))
    assert pm > min(pl, pr)
