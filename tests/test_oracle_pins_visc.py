"""Pins of the oracle's composed viscous update (visc_r / visc_p after the remap of each
sweep; viscosity.cu:370-472, with the quadratic-fit derivatives of viscosity.cu:2-19 taken on
the window's left cell edges, ppm.cu:130, 277-278), against what the mathematics fixes.  Each
check compares a sweep with viscosity on against the same sweep with it off, so the
difference is the viscous increment alone:

- a state uniform along the ring gets no azimuthal viscous increment;
- a transverse shear v_r = A sin(k phi) on a ring at rest diffuses at the rate nu (k/R)^2, the
  discrete increment converging at second order;
- in the radial sweep of a disk in radial force balance, rigid rotation gets no viscous
  torque, and a linear shear Omega(r) gets dv_phi/dt = (1/(rho R^2)) d/dR(nu rho R^3 dOmega/dR),
  the discrete increment converging at second order on a uniform radial grid (where the
  left-edge stencil's derivatives sit at the cell centres to second order).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.integrate import quad

import oracle
from penguin_2_0_amd import configs
from penguin_2_0_amd.abi import PnBody, PnGrid

TWO_PI = 2.0 * math.pi


def twin(grid, P):
    """The same state in two oracle disks, with viscosity on and off."""
    a = oracle.OracleDisk(grid, P.replace(visc_flag=1), 1, 0.0)
    b = oracle.OracleDisk(grid, P.replace(visc_flag=0), 1, 0.0)
    return a, b


def copy_state(a, b):
    for f in ("rings", "lft", "rgh"):
        getattr(b, f)[:] = getattr(a, f)


def ring(nphi, nu, rad0=50.0):
    P = configs.disk2d(2, nphi).replace(plnt_flag=0, fargo_flag=0, frrot_flag=0, xmin=rad0, xmax=rad0 * 1.0001,
                                        p_beta=0.0, p_alpha=0.0, vis_nu=nu)
    a, b = twin(PnGrid(2, nphi, 1), P)
    for o in (a, b):
        o.rings["res_v"] = 0.0
        o.rings["rot_v"] = 0.0
        o.rings["rot_j"] = 0
    return a, b


def test_phi_viscosity_leaves_a_ring_uniform_state_alone():
    a, b = ring(96, 1e-2)
    R = a.rings
    R["r"] = np.array([1.2, 0.7])[:, None]
    R["p"] = R["r"] * 0.0025
    R["v"] = 0.013
    R["u"] = np.array([0.002, -0.001])[:, None]
    copy_state(a, b)
    a.sweepy(0.3, PnBody(), 0.0)
    b.sweepy(0.3, PnBody(), 0.0)
    for f in ("r", "u", "v"):
        assert np.array_equal(a.rings[f], b.rings[f])


def _shear_error(nphi, nu=2.0, k=3, amp=1e-3):
    a, b = ring(nphi, nu)
    R = a.rings
    R["r"] = 1.0
    R["p"] = 0.0025
    R["v"] = 0.0
    phi = R["yc"][0]
    R["u"] = amp * np.sin(k * phi)[None, :]
    u0 = R["u"].copy()
    copy_state(a, b)
    rad = R["xc"][:, None]
    dt = 0.2 * (TWO_PI / nphi * rad.min()) ** 2 / nu   # well inside the explicit limit
    a.sweepy(dt, PnBody(), 0.0)
    b.sweepy(dt, PnBody(), 0.0)
    # the ring at rest with uniform pressure: the hydro part leaves v_r as it is (to the
    # rounding of the remap's mass ratio)
    assert np.abs(b.rings["u"] - u0).max() < 1e-15 * amp * 10
    rate = (a.rings["u"] - b.rings["u"]) / dt
    want = -nu * (k / rad) ** 2 * u0
    return np.abs(rate - want).max() / np.abs(want).max()


def test_phi_viscosity_diffuses_a_transverse_shear_at_the_analytic_rate():
    e1, e2, e3 = (_shear_error(n) for n in (48, 96, 192))
    assert e1 < 2e-2
    assert e1 / e2 > 3.5 and e2 / e3 > 3.5


def balanced_disk(n, omega, nu=1e-3, c=0.05, p_alpha=0.5, lo=0.95, hi=1.05):
    """A cylindrical disk on a uniform radial grid r in [lo, hi], no planet, rotating at
    Omega(r), with the density that balances gravity, rotation and isothermal pressure:
    c^2 d ln(rho)/dr = Omega^2 r - 1/r^2."""
    P = configs.disk2d(n, 12).replace(plnt_flag=0, fargo_flag=0, frrot_flag=0, grspx=0, xmin=lo, xmax=hi,
                                      p_beta=0.0, sc_h=c, p_alpha=p_alpha, vis_nu=nu)
    a, b = twin(PnGrid(n, 12, 1), P)
    lnrho = lambda r: quad(lambda s: (omega(s) ** 2 * s - 1.0 / s ** 2) / (c * c), 0.5 * (lo + hi), r,
                           epsabs=1e-13, epsrel=1e-13)[0]
    for A in (a.rings, a.lft, a.rgh):
        xc = A["xc"]
        rho = np.exp([lnrho(x) for x in xc])
        A["r"] = rho[:, None]
        A["p"] = (c * c * rho)[:, None]
        A["u"] = 0.0
        A["v"] = (omega(xc) * xc)[:, None]
        A["w"] = 0.0
        A["rot_j"] = 0
        A["res_v"] = 0.0
    copy_state(a, b)
    return a, b, lnrho


def _rigid_increment(n):
    # sub-Keplerian (Omega 0.9): the balancing density falls monotonically outwards, so the
    # parabolae are nowhere flattened at an extremum
    a, b, _ = balanced_disk(n, lambda r: 0.9 + 0.0 * r)
    dt = 1e-3
    a.sweepx(dt, PnBody(), 0.0)
    b.sweepx(dt, PnBody(), 0.0)
    inner = slice(8, n - 8)
    return np.abs(a.rings["v"] - b.rings["v"])[inner].max() / dt


def test_radial_viscosity_exerts_no_torque_on_rigid_rotation():
    """What is left is the viscous response to the small flow the hydro residual starts
    (~1e-7 here); the linear shear below gets ~0.1, and a viscous term taken on v_phi rather
    than on Omega = v_phi / R would give ~1e-3."""
    for n in (64, 256):
        assert _rigid_increment(n) < 1e-6


def _radial_shear_error(n, s=-1.5, nu=1e-3, p_alpha=0.5):
    omega = lambda r: 0.9 + s * (r - 1.0)
    a, b, lnrho = balanced_disk(n, omega, nu=nu, p_alpha=p_alpha)
    dt = 1e-3
    a.sweepx(dt, PnBody(), 0.0)
    b.sweepx(dt, PnBody(), 0.0)
    R = a.rings["xc"]
    rate = (a.rings["v"] - b.rings["v"])[:, 0] / dt
    # (1/(rho R^2)) d/dR(nu rho R^3 s) with nu = vis_nu R^p_alpha, d ln rho / dR from the balance
    c = a.params.sc_h
    dlnrho = (omega(R) ** 2 * R - 1.0 / R ** 2) / (c * c)
    nuR = nu * R ** p_alpha
    want = s * nuR * (3.0 + p_alpha + R * dlnrho)
    inner = slice(8, n - 8)
    return np.abs(rate - want)[inner].max() / np.abs(want)[inner].max()


def test_radial_viscosity_drives_a_shear_at_the_analytic_rate():
    e1, e2, e3 = (_radial_shear_error(n) for n in (64, 128, 256))
    assert e1 < 5e-2
    assert e1 / e2 > 3.5 and e2 / e3 > 3.5


def _radial_vr_error(n, nu=1e-3, p_alpha=0.5, amp=1e-4, c=0.2):
    """v_r = amp sin(1.5 pi (r - 0.3) / 0.2) on a balanced disk at r in [0.3, 0.5] (close in,
    so that the curvature terms are a few percent of the operator): the viscous increment of
    a short step against the divergence of the cylindrical stress, 2 d/dR(nu rho u') +
    2 nu rho u'/R - 2 nu rho u / R^2, over rho."""
    omega = lambda r: 0.95 * r ** -1.5   # sub-Keplerian: the density falls monotonically
    a, b, _ = balanced_disk(n, omega, nu=nu, p_alpha=p_alpha, c=c, lo=0.3, hi=0.5)
    k = 1.5 * math.pi / 0.2
    for o in (a, b):
        for A in (o.rings, o.lft, o.rgh):
            A["u"] = (amp * np.sin(k * (A["xc"] - 0.3)))[:, None]
    dt = 1e-8
    a.sweepx(dt, PnBody(), 0.0)
    b.sweepx(dt, PnBody(), 0.0)
    R = a.rings["xc"]
    rate = (a.rings["u"] - b.rings["u"])[:, 0] / dt
    dlnrho = (omega(R) ** 2 * R - 1.0 / R ** 2) / (c * c)
    x = k * (R - 0.3)
    u, u1, u2 = amp * np.sin(x), amp * k * np.cos(x), -amp * k * k * np.sin(x)
    nuR = nu * R ** p_alpha
    want = 2.0 * nuR * (u2 + (p_alpha / R + dlnrho) * u1 + u1 / R - u / R ** 2)
    inner = slice(n // 8, n - n // 8)
    return np.abs(rate - want)[inner].max() / np.abs(want)[inner].max()


def test_radial_viscosity_diffuses_the_radial_velocity_at_the_analytic_rate():
    e1, e2, e3 = (_radial_vr_error(n) for n in (64, 128, 256))
    assert e1 < 2e-2
    assert e1 / e2 > 3.5 and e2 / e3 > 3.5


def _phi_vphi_error(nphi, nu=2.0, k=2, amp=1e-6):
    """v_phi = amp sin(k phi) on a ring at rest: the viscous increment of a short step
    against (1/rho) d/ds(2 nu rho dv/ds) = 2 nu v'' (the reference drops the curvature terms,
    viscosity.cu:445,457)."""
    a, b = ring(nphi, nu)
    for o in (a, b):
        R = o.rings
        R["r"] = 1.0
        R["p"] = 0.0025
        R["u"] = 0.0
        R["v"] = amp * np.sin(k * R["yc"][0])[None, :]
    rad = a.rings["xc"][:, None]
    dt = 1e-4
    v0 = a.rings["v"].copy()
    a.sweepy(dt, PnBody(), 0.0)
    b.sweepy(dt, PnBody(), 0.0)
    rate = (a.rings["v"] - b.rings["v"]) / dt
    want = -2.0 * nu * (k / rad) ** 2 * v0
    return np.abs(rate - want).max() / np.abs(want).max()


def test_phi_viscosity_diffuses_the_azimuthal_velocity_at_the_analytic_rate():
    e1, e2, e3 = (_phi_vphi_error(n) for n in (48, 96, 192))
    assert e1 < 2e-2
    assert e1 / e2 > 3.5 and e2 / e3 > 3.5


# ---------------------------------------------------------------------------------------
# theta viscosity (device_viscosity_z, viscosity.cu:474-518), in the theta sweep

def theta_twin(nth, nu, r0=2.0, half=0.3, c=0.1):
    from test_oracle_pins_theta import theta_disk
    P, a = theta_disk(nth, half=half, c=c, r0=r0, vis_nu=nu, visc_flag=1)
    _, b = theta_disk(nth, half=half, c=c, r0=r0, vis_nu=nu, visc_flag=0)
    return a, b


def _theta_rings(o):
    return (o.rings, o.udr, o.top)


def test_theta_viscosity_leaves_a_theta_uniform_state_alone():
    a, b = theta_twin(40, 1e-2)
    for o in (a, b):
        for A in _theta_rings(o):
            A["r"] = 1.3
            A["p"] = 1.3 * 0.01
            A["u"] = 0.02
            A["v"] = 0.0   # (a uniform v_phi is not a uniform l = r sin(theta) v_phi)
            A["w"] = 0.0
    a.sweepz(0.01, PnBody(), 0.0)
    b.sweepz(0.01, PnBody(), 0.0)
    for f in ("r", "u", "v", "w"):
        assert np.array_equal(a.rings[f], b.rings[f])


def _theta_shear_error(nth, field, factor, nu=1e-3, k=4.0, amp=1e-6, r0=2.0):
    """field = amp sin(k (theta - pi/2)) on a shell at rest: the viscous increment of a short
    step against factor * nu * d2/ds2 of it along the arc s = r theta."""
    a, b = theta_twin(nth, nu, r0=r0)
    for o in (a, b):
        for A in _theta_rings(o):
            A["r"] = 1.0
            A["p"] = 0.01
            A["u"] = 0.0
            A["v"] = 0.0
            A["w"] = 0.0
            A[field] = (amp * np.sin(k * (A["zc"] - 0.5 * math.pi)))[:, None]
    f0 = a.rings[field].copy()
    dt = 1e-5
    a.sweepz(dt, PnBody(), 0.0)
    b.sweepz(dt, PnBody(), 0.0)
    rate = (a.rings[field] - b.rings[field]) / dt
    rad = a.rings["xc"][:, None]
    want = -factor * nu * (k / rad) ** 2 * f0
    return np.abs(rate - want).max() / np.abs(want).max()


def test_theta_viscosity_diffuses_each_velocity_at_the_analytic_rate():
    """v_r and v_phi across theta at nu k^2 / r^2, v_theta along it at twice that."""
    for field, factor in (("u", 1.0), ("v", 1.0), ("w", 2.0)):
        e1, e2, e3 = (_theta_shear_error(n, field, factor) for n in (32, 64, 128))
        assert e1 < 3e-2, field
        assert e1 / e2 > 3.5 and e2 / e3 > 3.5, field
