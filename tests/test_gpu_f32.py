"""FP32-bulk path against the plain FP64 path, through the C ABI, on the same seeded inputs.

The FP32 path carries Sigma, v_r and v_phi as FP32 deviations from per-ring FP64
backgrounds and is not bit-compatible with the FP64 path; the bars are relative L1 / Linf
errors in the lab frame (each ring's data rolled by its rot_j) with the tolerances written
next to each check, while integer work (rot_j) and graph replay are exact.
"""
from __future__ import annotations

import numpy as np
import pytest

from penguin_2_0_amd import abi, configs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def lab(state: np.ndarray, f: str) -> np.ndarray:
    a = state[f]
    return np.stack([np.roll(a[i], int(state["rot_j"][i])) for i in range(a.shape[0])])


def rel(a, b):
    d = np.abs(a - b)
    return d.sum() / np.abs(b).sum(), d.max() / np.abs(b).max()


def run(w, prec, n, disk=None, variants=None, graph=True):
    d = disk if disk is not None else abi.init_disk(w.grid, w.params, w.seed, w.noise)
    with abi.Context(w.grid, w.params, prec) as ctx:
        ctx.load(d)
        if variants is not None:
            ctx.set_variants(*variants)
        st = ctx.run_begin()
        ctx.run(n, st, graph=graph)
        return ctx.store(), st


def test_f32_load_store_roundtrip():
    """Upload rounds each deviation from the background once; nothing else changes."""
    w = configs.c1()
    say("case: fp32 roundtrip", w.name)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
        ctx.load(d)
        out = ctx.store()
    val = d.val_array()
    for f, bg in (("r", val[:, 0:1]), ("v", val[:, 3:4]), ("u", 0.0)):
        dev = np.abs(d.rings[f] - bg)
        err = np.abs(out[f] - d.rings[f])
        # one FP32 rounding of the deviation, plus the FP64 re-addition
        assert np.all(err <= 2.0 ** -24 * dev + 4e-16 * np.abs(d.rings[f]) + 1e-300), f
    assert np.array_equal(out["rot_j"], d.rings["rot_j"])
    # the isothermal closure: P = c_s^2 Sigma with the ring's c_s^2
    cs2 = d.rings["p"] / d.rings["r"]
    assert np.allclose(out["p"], out["r"] * cs2, rtol=1e-14, atol=0)


def test_f32_eos3_load_store_roundtrip():
    """EOS 3 carries q = P / Sigma as its deviation from q_b = P_b / Sigma_b: the round trip
    rounds that deviation once, so P comes back within one FP32 ulp of Sigma (q - q_b)."""
    w = configs.small(96, 288, eos=3, gam=1.4, M_p=1e-3)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
        ctx.load(d)
        out = ctx.store()
    val = d.val_array()
    q, qb = d.rings["p"] / d.rings["r"], (val[:, 1] / val[:, 0])[:, None]
    dr = np.abs(d.rings["r"] - val[:, 0:1])
    assert np.all(np.abs(out["r"] - d.rings["r"]) <= 2.0 ** -24 * dr + 4e-16 * d.rings["r"]), "r"
    # P' = Sigma' (q_b + dq'): the Sigma rounding enters through q, the dq rounding through Sigma
    bound = out["r"] * 2.0 ** -24 * np.abs(q - qb) + q * 2.0 ** -24 * dr + 1e-15 * d.rings["p"]
    assert np.all(np.abs(out["p"] - d.rings["p"]) <= bound), "p"


@pytest.mark.parametrize("w,n,tol", [(configs.small(130, 250), 50, 1e-6), (configs.c1(), 100, 1e-6)],
                         ids=["disk2d_130x250", "c1"])
def test_f32_vs_fp64(w, n, tol):
    say("case: fp32 vs fp64", w.name, n, "steps")
    a, sa = run(w, abi.PN_FP32, n)
    b, sb = run(w, abi.PN_FP64, n)
    assert np.array_equal(a["rot_j"], b["rot_j"])
    assert abs(sa.dt - sb.dt) <= 1e-6 * sb.dt
    assert abs(sa.simtime - sb.simtime) <= 1e-6 * sb.simtime
    for f, t1, tinf in (("r", tol, 1e-4), ("v", tol, 1e-5)):
        l1, li = rel(lab(a, f), lab(b, f))
        say(f"  {f}: L1 {l1:.3e} Linf {li:.3e}")
        assert l1 <= t1 and li <= tinf, (f, l1, li)
    # v_r starts at zero: its error is judged against the largest |v_r|
    l1, li = rel(lab(a, "u"), lab(b, "u"))
    say(f"  u: L1 {l1:.3e} Linf {li:.3e}")
    assert l1 <= 1e-4 and li <= 1e-2


def visc_disk(nr, nphi, alpha):
    w = configs.small(nr, nphi, visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05)
    w.name += f"_visc{alpha:g}"
    return w


@pytest.mark.parametrize("w", [visc_disk(130, 250, 1e-2), visc_disk(256, 768, 1e-3)], ids=lambda w: w.name)
def test_f32_viscous_vs_fp64(w):
    """Explicit viscosity in perturbation form tracks the FP64 viscous update."""
    n = 100
    say("case: fp32 viscous vs fp64", w.name, n, "steps")
    a, sa = run(w, abi.PN_FP32, n)
    b, sb = run(w, abi.PN_FP64, n)
    assert np.array_equal(a["rot_j"], b["rot_j"])
    # dt follows the largest signal speed, i.e. the Linf of the state
    assert abs(sa.dt - sb.dt) <= 1e-5 * sb.dt
    for f, t1, tinf in (("r", 1e-6, 1e-4), ("v", 1e-6, 1e-5), ("u", 1e-4, 1e-2)):
        l1, li = rel(lab(a, f), lab(b, f))
        say(f"  {f}: L1 {l1:.3e} Linf {li:.3e}")
        assert l1 <= t1 and li <= tinf, (f, l1, li)


def test_f32_viscous_keeps_reference_azimuthal_windows():
    """The azimuthal viscosity reads halo cells at the window edges, so a viscous disk must run
    the reference's 128-cell azimuthal windows: the tuner picks among those only and a longer
    window is refused.  A strongly forced viscous disk then tracks FP64 closely (with 256-cell
    windows it drifted to L1 4e-6 by 100 steps)."""
    w = configs.small(64, 192, M_p=3e-4, visc_flag=1, ss_alpha=1e-2, vis_nu=1e-2 * 0.05 * 0.05)
    w.seed = 101
    say("case: fp32 viscous windows", w.name)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
        ctx.load(d)
        for vy in (2, 3):
            with pytest.raises(abi.PnError) as e:
                ctx.set_variants(0, vy)
            assert e.value.code == abi.PN_ERR_ARG
        st = ctx.run_begin()
        ctx.run(100, st)
        assert ctx.variants()[1] in (0, 1)
        a = ctx.store()
    b, _ = run(w, abi.PN_FP64, 100, disk=d)
    l1, li = rel(lab(a, "r"), lab(b, "r"))
    say(f"  rho L1 {l1:.3e} Linf {li:.3e}")
    assert l1 <= 1e-6 and li <= 1e-4


def test_f32_every_variant():
    """Every (cells per lane, waves, columns) variant of both sweeps is a valid step."""
    w = configs.small(130, 250)
    say("case: fp32 variants", w.name)
    b, _ = run(w, abi.PN_FP64, 20)
    for vx in range(4):
        for vy in range(4):
            a, _ = run(w, abi.PN_FP32, 20, variants=(vx, vy))
            l1, li = rel(lab(a, "r"), lab(b, "r"))
            say(f"  variants ({vx}, {vy}): rho L1 {l1:.3e} Linf {li:.3e}")
            assert np.array_equal(a["rot_j"], b["rot_j"])
            assert l1 <= 1e-6 and li <= 1e-4


def test_f32_graph_equals_eager_bitwise():
    w = configs.small(130, 250)
    say("case: fp32 graph vs eager", w.name)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    outs = []
    for graph in (False, True):
        with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
            ctx.load(d)
            ctx.set_variants(0, 3)
            st = ctx.run_begin()
            for _ in range(3):
                ctx.run(2, st, graph=graph)
                ctx.totals()
            outs.append((ctx.store(), st.dt, st.simtime))
    assert np.array_equal(outs[0][0].view(np.uint8), outs[1][0].view(np.uint8))
    assert outs[0][1] == outs[1][1] and outs[0][2] == outs[1][2]


def test_f32_roll_invariance_bitwise():
    """The same physical state stored with every ring rolled by m and rot_j lowered by m steps
    to the same lab-frame bytes."""
    w = configs.small(48, 144)
    say("case: fp32 roll invariance", w.name)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    rng = np.random.default_rng(5)
    d2 = d.copy()
    J = w.grid.jmax
    for n, m in enumerate(rng.integers(0, J, size=len(d2.rings))):
        for f in ("r", "p", "u", "v", "w"):
            d2.rings[f][n] = np.roll(d.rings[f][n], -m)
        d2.rings["rot_j"][n] = (d.rings["rot_j"][n] + m) % J
    a, _ = run(w, abi.PN_FP32, 3, disk=d, variants=(0, 3))
    b, _ = run(w, abi.PN_FP32, 3, disk=d2, variants=(0, 3))
    for f in ("r", "u", "v"):
        assert np.array_equal(lab(a, f), lab(b, f)), f


def test_f32_totals_track_fp64():
    """Mass and angular momentum of the FP32 state (summed in FP64) follow the FP64 path's."""
    w = configs.c1()
    say("case: fp32 totals", w.name)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    tot = {}
    for prec in (abi.PN_FP32, abi.PN_FP64):
        with abi.Context(w.grid, w.params, prec) as ctx:
            ctx.load(d)
            m0, l0 = ctx.totals()
            st = ctx.run_begin()
            ctx.run(100, st)
            m1, l1 = ctx.totals()
            tot[prec] = (m0, l0, m1, l1)
    a, b = tot[abi.PN_FP32], tot[abi.PN_FP64]
    say(f"  fp32 {a}\n  fp64 {b}")
    assert abs(a[0] - b[0]) <= 1e-12 * b[0] and abs(a[1] - b[1]) <= 1e-12 * abs(b[1])
    assert abs(a[2] - b[2]) <= 1e-9 * b[2] and abs(a[3] - b[3]) <= 1e-9 * abs(b[3])


HPI = 1.570796326794896619231321691639751442


def disk3d(nr, nphi, nth):
    p = configs.disk2d(nr, nphi, q=3.0e-5).replace(
        ndim=3, ngeomx=2, ngeomy=4, ngeomz=5, p_alpha=1.5, p_beta=1.0, zmin=HPI - 0.15, zmax=HPI + 0.15)
    return configs.Workload(f"disk3d_{nr}x{nphi}x{nth}", abi.PnGrid(nr, nphi, nth), p, seed=11, steps=10)


def test_f32_3d_load_store_roundtrip():
    w = disk3d(24, 72, 12)
    say("case: fp32 3-D roundtrip", w.name)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
        ctx.load(d)
        out = ctx.store()
    val = d.val_array()
    for f, bg in (("r", val[:, 0:1]), ("v", val[:, 3:4]), ("u", 0.0), ("w", 0.0)):
        dev = np.abs(d.rings[f] - bg)
        err = np.abs(out[f] - d.rings[f])
        assert np.all(err <= 2.0 ** -24 * dev + 4e-16 * np.abs(d.rings[f]) + 1e-300), f


# 3-D bars: Sigma and v_phi as in 2-D; v_r and v_theta start at zero and are judged
# against their own largest values
def damped(w):
    w.params.kill_flag = 1
    w.name += "_damping"
    return w


def bary3d(w):
    w.params = w.params.replace(bary_flag=1, M_p=1e-3)
    w.name += "_bary"
    return w


def bounded(w, nlft, nrgh):
    w.params = w.params.replace(nlft=nlft, nrgh=nrgh)
    w.name += f"_bound{nlft}{nrgh}"
    return w


@pytest.mark.parametrize("w,n", [(disk3d(24, 72, 12), 20), (disk3d(64, 192, 32), 50), (damped(disk3d(64, 192, 32)), 50),
                                 (bounded(disk3d(24, 72, 12), 1, 0), 30), (bounded(disk3d(64, 192, 32), 4, 2), 30),
                                 (bary3d(disk3d(64, 192, 32)), 30)],
                         ids=["disk3d_24x72x12", "disk3d_64x192x32", "disk3d_64x192x32_damping",
                              "disk3d_24x72x12_outflow_reflect", "disk3d_64x192x32_mixed_periodic",
                              "disk3d_64x192x32_bary"])
def test_f32_3d_vs_fp64(w, n):
    say("case: fp32 3-D vs fp64", w.name, n, "steps")
    a, sa = run(w, abi.PN_FP32, n)
    b, sb = run(w, abi.PN_FP64, n)
    assert np.array_equal(a["rot_j"], b["rot_j"])
    assert abs(sa.dt - sb.dt) <= 1e-5 * sb.dt
    for f, t1, tinf in (("r", 1e-6, 1e-4), ("v", 1e-6, 1e-5), ("u", 1e-3, 1e-2), ("w", 1e-3, 1e-2)):
        l1, li = rel(lab(a, f), lab(b, f))
        say(f"  {f}: L1 {l1:.3e} Linf {li:.3e}")
        assert l1 <= t1 and li <= tinf, (f, l1, li)


def visc3d(nr, nphi, nth, alpha=3e-3):
    w = disk3d(nr, nphi, nth)
    w.params = w.params.replace(visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05, M_p=1e-4)
    w.name += f"_visc{alpha:g}"
    return w


@pytest.mark.parametrize("w,n", [(visc3d(24, 72, 12), 30), (visc3d(64, 192, 32, 1e-2), 50)],
                         ids=["disk3d_24x72x12_visc", "disk3d_64x192x32_visc"])
def test_f32_3d_viscous_vs_fp64(w, n):
    """The viscous 3-D disk on the FP32 path: viscosity in perturbation form in the radial
    (device_viscosity_r with its v_theta term), azimuthal (device_viscosity_p with v_theta)
    and theta sweeps (device_viscosity_z, n = vis_nu rho), against the FP64 path with the 3-D
    bars, in two theta shapes.  The theta viscosity reads halo cells at window edges, so the
    192-cell theta windows are refused for a viscous disk."""
    say("case: fp32 3-D viscous vs fp64", w.name, n, "steps")
    b, sb = run(w, abi.PN_FP64, n)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    err = {}
    for vz in (0, 3):
        with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
            ctx.load(d)
            ctx.set_variants(0, 1)
            ctx.set_variant_z(vz)
            st = ctx.run_begin()
            ctx.run(n, st, graph=True)
            a = ctx.store()
        assert np.array_equal(a["rot_j"], b["rot_j"])
        assert abs(st.dt - sb.dt) <= 1e-5 * sb.dt
        for f, t1, tinf in (("r", 1e-6, 1e-4), ("v", 1e-6, 1e-5), ("u", 1e-3, 1e-2), ("w", 1e-3, 1e-2)):
            l1, li = rel(lab(a, f), lab(b, f))
            err[f] = max(err.get(f, 0.0), l1)
            say(f"  theta shape {vz}: {f} L1 {l1:.3e} Linf {li:.3e}")
            assert l1 <= t1 and li <= tinf, (vz, f, l1, li)
    # the viscosity moves the state by far more than the FP32 path's distance from FP64
    p0 = w.params.replace(visc_flag=0, ss_alpha=0.0, vis_nu=0.0)
    c, _ = run(configs.Workload(w.name + "_inviscid", w.grid, p0, seed=w.seed, steps=n), abi.PN_FP64, n, disk=d)
    for f in ("r", "u", "w"):
        l1, _ = rel(lab(c, f), lab(b, f))
        say(f"  inviscid vs viscous FP64: {f} L1 {l1:.3e} (FP32 distance {err[f]:.3e})")
        assert l1 > 10 * err[f], (f, l1, err[f])
    with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
        with pytest.raises(abi.PnError) as e:
            ctx.set_variant_z(4)
        assert e.value.code == abi.PN_ERR_ARG
        ctx.load(d)
        ctx.tune()
        assert ctx.variant_z() != 4


def test_f32_3d_moving_planet_vs_fp64():
    """A planet that moves in 3-D (under the star alone: the reference's disk-on-planet
    reductions are 2-D only), in the inertial frame so that it crosses the grid: the FP32 path
    with its CFL pass before each step and the per-step planet tables against FP64, fields
    with the 3-D bars and the planet's position."""
    w = disk3d(24, 72, 12)
    w.params = w.params.replace(plnt_flag=2, M_p=1e-3, frrot_flag=0)
    w.name += "_moving_planet_inertial"
    n = 30
    say("case: fp32 3-D moving planet vs fp64", w.name, n, "steps")
    a, sa = run(w, abi.PN_FP32, n)
    b, sb = run(w, abi.PN_FP64, n)
    assert np.array_equal(a["rot_j"], b["rot_j"])
    assert abs(sa.dt - sb.dt) <= 1e-5 * sb.dt
    for f, t1, tinf in (("r", 1e-6, 1e-4), ("v", 1e-6, 1e-5), ("u", 1e-3, 1e-2), ("w", 1e-3, 1e-2)):
        l1, li = rel(lab(a, f), lab(b, f))
        say(f"  {f}: L1 {l1:.3e} Linf {li:.3e}")
        assert l1 <= t1 and li <= tinf, (f, l1, li)
    # the orbit is the star's alone, so the two planets sit on one circle at their own times:
    # the azimuths differ by the angular speed times the difference of the summed dt
    lag = (sa.planet.y - sb.planet.y) - sb.planet.vy * (sa.simtime - sb.simtime)
    say(f"  planet ({sa.planet.x!r}, {sa.planet.y!r}) vs ({sb.planet.x!r}, {sb.planet.y!r}); "
        f"simtime {sa.simtime!r} vs {sb.simtime!r}; azimuth beyond the time difference {lag:.3e}")
    assert abs(sb.planet.y - 3.141592653589793) > 1.0   # it moved
    assert abs(sa.planet.x - sb.planet.x) <= 1e-12 and abs(lag) <= 1e-11


def test_f32_3d_every_variant_and_graph():
    """Every radial / theta variant is a valid 3-D step, and graph replay equals eager."""
    w = disk3d(24, 72, 12)
    say("case: fp32 3-D variants", w.name)
    b, _ = run(w, abi.PN_FP64, 10)
    for vx in range(4):
        a, _ = run(w, abi.PN_FP32, 10, variants=(vx, 3))
        l1, li = rel(lab(a, "r"), lab(b, "r"))
        say(f"  variant {vx}: rho L1 {l1:.3e} Linf {li:.3e}")
        assert l1 <= 1e-6 and li <= 1e-4
    e, _ = run(w, abi.PN_FP32, 6, variants=(0, 3), graph=False)
    g, _ = run(w, abi.PN_FP32, 6, variants=(0, 3), graph=True)
    assert np.array_equal(e.view(np.uint8), g.view(np.uint8))


@pytest.mark.parametrize("nth", [64, 200])
def test_f32_theta_variants(nth):
    """Every theta shape (pn_set_variant_z) is a valid 3-D step: the four radial shapes and the
    192-cell one (K = 3) that runs two theta pencils per wavefront when a window holds at most
    96 cells with its halos (64 theta cells: 76), and one per wavefront on the full windows of
    200 theta cells.  The packed and unpacked K = 3 runs give the same bytes."""
    w = disk3d(16, 48, nth)
    say("case: fp32 theta variants", w.name)
    b, _ = run(w, abi.PN_FP64, 10)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    outs = {}
    for vz, pack in [(0, True), (1, True), (2, True), (3, True), (4, True), (4, False)]:
        with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
            ctx.load(d)
            ctx.set_variants(0, 3)
            ctx.set_variant_z(vz)
            ctx.set_packing(pack)
            assert ctx.variant_z() == vz
            st = ctx.run_begin()
            ctx.run(10, st, graph=True)
            a = ctx.store()
        outs[(vz, pack)] = a
        assert np.array_equal(a["rot_j"], b["rot_j"])
        for f, t1, tinf in (("r", 1e-6, 1e-4), ("v", 1e-6, 1e-5), ("u", 1e-3, 1e-2), ("w", 1e-3, 1e-2)):
            l1, li = rel(lab(a, f), lab(b, f))
            say(f"  theta variant {vz} packing {pack}: {f} L1 {l1:.3e} Linf {li:.3e}")
            assert l1 <= t1 and li <= tinf, (vz, f, l1, li)
    assert np.array_equal(outs[(4, True)].view(np.uint8), outs[(4, False)].view(np.uint8))
    same = np.array_equal(outs[(4, True)].view(np.uint8), outs[(0, True)].view(np.uint8))
    say(f"  K = 3 and K = 2 theta windows give the same bytes: {same}")
    with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
        with pytest.raises(abi.PnError) as e:
            ctx.set_variant_z(5)
        assert e.value.code == abi.PN_ERR_ARG


def flags_case(suffix, **over):
    w = configs.small(96, 288, **over)
    w.name += suffix
    return w


# every flag combination the FP32 path accepts, against the FP64 path on the same state
FLAG_CASES = [flags_case("_fargo2", fargo_flag=2), flags_case("_inertial", frrot_flag=0),
              flags_case("_noflat", flat_flag=0), flags_case("_noplanet", plnt_flag=0),
              flags_case("_nofargo", fargo_flag=0), flags_case("_growing_planet", t_growth=5.0),
              flags_case("_strong_planet", M_p=1e-3),
              # a planet that moves under the disk's forces, reduced from the FP32 state
              flags_case("_moving_planet", plnt_flag=2, M_p=1e-3),
              # damping zones at both radial edges (the separate CFL kernel then gives dt)
              flags_case("_damping", kill_flag=1),
              # more than 2048 slots: Index_Shift, the table rows and the roll over the whole grid
              configs.small(2048, 116, M_p=1e-3),
              # radial boundary codes (set_boundx, boundary.cu:86-206): ghost rings copied from the
              # interior before each radial sweep, carrying their source rings' backgrounds (draws
              # where no ring's FARGO shift per step lies within the two paths' dt difference of a
              # half cell, so that rot_j agrees; with the reflecting inner edge and M_p = 1e-4,
              # ring 4 rounds its shift the other way near step 20)
              flags_case("_reflect", nlft=0, nrgh=0, M_p=1e-3), flags_case("_outflow", nlft=1, nrgh=1),
              flags_case("_periodic_r", nlft=2, nrgh=2), flags_case("_outflow_reflect", nlft=1, nrgh=0),
              flags_case("_mixed_viscous", nlft=4, nrgh=4, M_p=1e-3, visc_flag=1, ss_alpha=1e-2, vis_nu=2.5e-5),
              flags_case("_outflow_damping_moving_planet", nlft=1, nrgh=1, kill_flag=1, plnt_flag=2, M_p=1e-3),
              # the barycentric frame (bary_flag, forces.cu:28-40, 116-126): the star's offset pull
              # as a deviation force beside the planet's, no indirect term
              flags_case("_bary", bary_flag=1, M_p=1e-3), flags_case("_bary_weak_planet", bary_flag=1),
              flags_case("_bary_moving_planet", bary_flag=1, plnt_flag=2, M_p=1e-3),
              flags_case("_bary_inertial", bary_flag=1, frrot_flag=0, M_p=1e-3)]


@pytest.mark.parametrize("w", FLAG_CASES, ids=lambda w: w.name)
def test_f32_flags_vs_fp64(w):
    n = 50
    say("case: fp32 flags", w.name, n, "steps")
    a, sa = run(w, abi.PN_FP32, n)
    b, sb = run(w, abi.PN_FP64, n)
    assert np.array_equal(a["rot_j"], b["rot_j"])
    assert abs(sa.dt - sb.dt) <= 1e-5 * sb.dt
    for f, t1, tinf in (("r", 1e-6, 1e-4), ("v", 1e-6, 1e-5)):
        l1, li = rel(lab(a, f), lab(b, f))
        say(f"  {f}: L1 {l1:.3e} Linf {li:.3e}")
        assert l1 <= t1 and li <= tinf, (f, l1, li)
    if w.params.plnt_flag == 2:
        dp = max(abs(sa.planet.x - sb.planet.x), abs(sa.planet.y - sb.planet.y),
                 abs(sa.planet.vx - sb.planet.vx), abs(sa.planet.vy - sb.planet.vy))
        say(f"  planet: x {sb.planet.x!r} y {sb.planet.y!r}, largest difference {dp:.3e}")
        assert sb.planet.vx != 0.0 and dp <= 5e-9   # the disk has moved the planet; both paths alike


# EOS 3 (the energy-consistent adiabatic step): Sigma, v_r, v_phi and q = p / Sigma in FP32
# against the FP64 path; p is judged like Sigma, v_r (noise-driven, small) against its own
# largest value
EOS3_CASES = [flags_case("_eos3", eos=3, gam=1.4, init_p_iso=1),
              flags_case("_eos3_adiabatic_ic", eos=3, gam=1.4),
              flags_case("_eos3_strong_planet", eos=3, gam=1.4, init_p_iso=1, M_p=1e-3),
              flags_case("_eos3_fargo2", eos=3, gam=1.4, init_p_iso=1, fargo_flag=2),
              flags_case("_eos3_inertial", eos=3, gam=1.4, init_p_iso=1, frrot_flag=0),
              flags_case("_eos3_noflat", eos=3, gam=1.4, init_p_iso=1, flat_flag=0),
              flags_case("_eos3_moving_planet", eos=3, gam=1.4, init_p_iso=1, plnt_flag=2, M_p=1e-3),
              flags_case("_eos3_damping", eos=3, gam=1.4, init_p_iso=1, kill_flag=1),
              flags_case("_eos3_reflect", eos=3, gam=1.4, init_p_iso=1, nlft=0, nrgh=0),
              flags_case("_eos3_mixed", eos=3, gam=1.4, init_p_iso=1, nlft=4, nrgh=4),
              flags_case("_eos3_bary", eos=3, gam=1.4, init_p_iso=1, bary_flag=1, M_p=1e-3),
              # viscosity (velocities only, as on the FP64 path) with either initial pressure
              flags_case("_eos3_viscous", eos=3, gam=1.4, init_p_iso=1, visc_flag=1, ss_alpha=1e-2, vis_nu=2.5e-5,
                         M_p=1e-3),
              flags_case("_eos3_viscous_adiabatic_ic", eos=3, gam=1.4, visc_flag=1, ss_alpha=3e-3, vis_nu=7.5e-6)]


@pytest.mark.parametrize("w", EOS3_CASES, ids=lambda w: w.name)
def test_f32_eos3_vs_fp64(w):
    n = 50
    say("case: fp32 eos3", w.name, n, "steps")
    a, sa = run(w, abi.PN_FP32, n)
    b, sb = run(w, abi.PN_FP64, n)
    assert np.array_equal(a["rot_j"], b["rot_j"])
    assert abs(sa.dt - sb.dt) <= 1e-5 * sb.dt
    for f, t1, tinf in (("r", 1e-6, 1e-4), ("p", 1e-6, 1e-4), ("v", 1e-6, 1e-5), ("u", 1e-3, 1e-3)):
        l1, li = rel(lab(a, f), lab(b, f))
        say(f"  {f}: L1 {l1:.3e} Linf {li:.3e}")
        assert l1 <= t1 and li <= tinf, (f, l1, li)


def test_f32_eos3_every_variant_and_graph():
    w = EOS3_CASES[2]
    say("case: fp32 eos3 variants", w.name)
    b, _ = run(w, abi.PN_FP64, 30)
    for vx in range(4):
        for vy in range(4):
            a, _ = run(w, abi.PN_FP32, 30, variants=(vx, vy))
            l1, _ = rel(lab(a, "p"), lab(b, "p"))
            say(f"  variants ({vx}, {vy}): p L1 {l1:.3e}")
            assert l1 <= 1e-7
    e, _ = run(w, abi.PN_FP32, 6, graph=False)
    g, _ = run(w, abi.PN_FP32, 6, graph=True)
    assert np.array_equal(e.view(np.uint8), g.view(np.uint8))


def test_f32_landing_at_the_fp64_steps():
    """The landing on picture times takes the same steps on the FP32 path as on the FP64 one."""
    w = configs.small(96, 288)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    with abi.Context(w.grid, w.params, abi.PN_FP64) as c:
        c.load(d)
        s = c.run_begin()
        dt0 = c.get_dt(s.dt, s.FrRot)
    w.params.tmovie = 7.5 * dt0
    out = {}
    for prec in (abi.PN_FP64, abi.PN_FP32):
        with abi.Context(w.grid, w.params, prec) as c:
            c.load(d)
            s = c.run_begin()
            s.land = 1
            c.run(100, s)
            out[prec] = (s.event, s.ncycle, s.simtime, c.store())
    say("case: fp32 landing", out[abi.PN_FP64][:3], out[abi.PN_FP32][:3])
    assert out[abi.PN_FP32][:2] == out[abi.PN_FP64][:2] == (1, 8)
    assert abs(out[abi.PN_FP32][2] - w.params.tmovie) <= 1e-12 * w.params.tmovie
    l1, _ = rel(lab(out[abi.PN_FP32][3], "r"), lab(out[abi.PN_FP64][3], "r"))
    assert l1 <= 1e-6


@pytest.mark.parametrize("prec", [abi.PN_FP64, abi.PN_FP32], ids=["fp64", "fp32"])
def test_3d_graph_after_an_odd_number_of_steps(prec):
    """A 3-D step flips the state buffer and a captured graph holds two steps: after a call
    that ended on an odd step, the next graph replay must start from the other buffer."""
    w = disk3d(24, 72, 12)
    a, _ = run(w, prec, 7, graph=False)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    with abi.Context(w.grid, w.params, prec) as ctx:
        ctx.load(d)
        st = ctx.run_begin()
        ctx.run(3, st, graph=True)
        ctx.run(4, st, graph=True)
        b = ctx.store()
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("over", [dict(), dict(eos=3, gam=1.4, init_p_iso=1)], ids=["isothermal", "eos3"])
def test_f32_damp_edges_matches_fp64(over):
    """pn_damp_edges on an FP32 context (kill_wave.cu:3-43 on the deviations) against the FP64
    context on the same state: the rings inside the zones relax towards their background
    record, and the FP32 result is the FP64 one to the rounding of the stored deviations."""
    w = configs.small(96, 288, kill_flag=1, **over)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    # the disk starts with v_r = 0 and v_phi on its background: give every field a deviation
    rng = np.random.default_rng(5)
    shape = d.rings["r"].shape
    d.rings["u"] += 1e-3 * rng.uniform(-1.0, 1.0, shape)
    d.rings["v"] *= 1.0 + 1e-4 * rng.uniform(-1.0, 1.0, shape)
    if over:
        d.rings["p"] *= 1.0 + 1e-4 * rng.uniform(-1.0, 1.0, shape)
    say("case: fp32 damp_edges", w.name, over)
    out = {}
    for prec in (abi.PN_FP32, abi.PN_FP64):
        with abi.Context(w.grid, w.params, prec) as ctx:
            ctx.load(d)
            ctx.damp_edges(0.1)
            out[prec] = ctx.store()
    a, b = out[abi.PN_FP32], out[abi.PN_FP64]
    val = d.val_array()
    fields = [("r", 0), ("u", 2), ("v", 3)] + ([("p", 1)] if over else [])
    for f, col in fields:
        bg = val[:, col:col + 1]
        dev0, dev = np.abs(d.rings[f] - bg), np.abs(b[f] - bg)
        err = np.abs(a[f] - b[f]).sum() / dev.sum()
        say(f"  {f}: innermost ring deviation {dev0[0].sum():.3e} -> {dev[0].sum():.3e}, FP32 vs FP64 {err:.2e}")
        assert dev[0].sum() < dev0[0].sum() and err <= 1e-6, f
