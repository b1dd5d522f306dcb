"""Batched FP32 disks (pn_batch_*) against single-disk runs, through the C ABI.

A batch runs the same per-disk arithmetic as a single FP32 context, with the work of all
disks in one launch per sweep; with the sweep shapes pinned to the same variants every
disk must come out as its own single-context run, and close to its own FP64 run (relative
L1 in the lab frame, tolerances written next to each check).
"""
from __future__ import annotations

import numpy as np
import pytest

from penguin_2_0_amd import abi, configs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def lab(state, f):
    a = state[f]
    return np.stack([np.roll(a[i], int(state["rot_j"][i])) for i in range(a.shape[0])])


def rel(a, b):
    d = np.abs(a - b)
    return d.sum() / np.abs(b).sum(), d.max() / np.abs(b).max()


def members():
    """Uneven disks like the parameter sweep: different grids, planet masses, viscosities."""
    out = []
    for d, (nr, q, alpha) in enumerate([(48, 1e-5, 1e-4), (64, 3e-4, 1e-2), (96, 1e-3, 3e-3), (130, 1e-4, 1e-3),
                                        (48, 1e-4, 0.0)]):
        over = dict(visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05) if alpha > 0 else {}
        w = configs.small(nr, 3 * nr, M_p=q, **over)
        w.seed = 100 + d
        out.append(w)
    return out


def run_batch(ws, n, variants=(0, 1), graph=True):
    disks = [abi.init_disk(w.grid, w.params, w.seed, w.noise) for w in ws]
    with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
        for ctx, d in zip(b.disks, disks):
            ctx.load(d)
        b.set_variants(*variants)
        st = b.run_begin()
        b.run(n, st, graph=graph)
        return [ctx.store() for ctx in b.disks], [st[d] for d in range(len(ws))]


def run_single(w, prec, n, variants=(0, 1), disk=None):
    d = disk if disk is not None else abi.init_disk(w.grid, w.params, w.seed, w.noise)
    with abi.Context(w.grid, w.params, prec) as ctx:
        ctx.load(d)
        if prec == abi.PN_FP32:
            ctx.set_variants(*variants)
        st = ctx.run_begin()
        ctx.run(n, st)
        return ctx.store(), st


def test_batch_equals_single_disks():
    ws = members()
    n = 30
    say("case: batch of", len(ws), "disks,", n, "steps")
    outs, sts = run_batch(ws, n)
    for w, o, s in zip(ws, outs, sts):
        ref, rs = run_single(w, abi.PN_FP32, n)
        say(f"  {w.name}: dt {s.dt!r} vs {rs.dt!r}")
        assert s.dt == rs.dt and s.simtime == rs.simtime and s.FrRot == rs.FrRot
        assert np.array_equal(o["rot_j"], ref["rot_j"])
        for f in ("r", "u", "v"):
            l1, li = rel(o[f], ref[f])
            assert li <= 1e-12, (w.name, f, l1, li)


def test_batch_vs_fp64():
    """Each batched disk tracks its own FP64 run at 30 and 100 steps (the strongly forced disks
    have thermal planet masses up to 8); the flow's own sensitivity -- the drift of an FP64 run
    from the FP32-rounded initial state -- is printed beside it."""
    ws = members()
    disks = [abi.init_disk(w.grid, w.params, w.seed, w.noise) for w in ws]
    say("case: batch vs fp64, 30 and 100 steps")
    with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
        for ctx, d in zip(b.disks, disks):
            ctx.load(d)
        b.set_variants(0, 1)
        st = b.run_begin()
        b.run(30, st)
        o30 = [(ctx.store(), st[k].dt) for k, ctx in enumerate(b.disks)]
        b.run(70, st)
        o100 = [ctx.store() for ctx in b.disks]
    for k, (w, d) in enumerate(zip(ws, disks)):
        ref, rs = run_single(w, abi.PN_FP64, 30)
        o, dt = o30[k]
        assert np.array_equal(o["rot_j"], ref["rot_j"])
        l1, li = rel(lab(o, "r"), lab(ref, "r"))
        say(f"  {w.name} 30 steps: rho L1 {l1:.3e} Linf {li:.3e}; dt {dt:.9e} vs {rs.dt:.9e}")
        assert l1 <= 1e-6 and li <= 1e-4
        assert abs(dt - rs.dt) <= 1e-5 * rs.dt
        ref, _ = run_single(w, abi.PN_FP64, 100)
        dr = d.copy()
        for f in ("r", "p", "u", "v"):
            dr.rings[f] = dr.rings[f].astype(np.float32).astype(np.float64)
        flo, _ = run_single(w, abi.PN_FP64, 100, disk=dr)
        l1, _ = rel(lab(o100[k], "r"), lab(ref, "r"))
        floor, _ = rel(lab(flo, "r"), lab(ref, "r"))
        say(f"  {w.name} 100 steps: rho L1 {l1:.3e}, sensitivity floor {floor:.3e}")
        assert l1 <= 1e-6


def test_batch_of_large_c4_members_within_bar_after_1000_steps():
    """The 20 members of the 48-disk sweep with 384 or 512 rings, as one batch for 1000
    steps with the tuner's shapes: every disk within max(1e-5, its floor) of its own FP64 run
    in relative L1(Sigma), the floor being the FP64 run from the FP32-rounded initial state
    (measured: all within 1.2e-6)."""
    ws = [w for w in configs.c4_members() if w.grid.imax >= 384]
    n = 1000
    disks = [abi.init_disk(w.grid, w.params, w.seed, w.noise) for w in ws]
    say("case: batch of the", len(ws), "C4 members with 384 or 512 rings,", n, "steps")
    with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
        for ctx, d in zip(b.disks, disks):
            ctx.load(d)
        st = b.run_begin()
        b.run(n, st, graph=True)
        outs = [ctx.store() for ctx in b.disks]
    for w, d, o in zip(ws, disks, outs):
        ref, _ = run_single(w, abi.PN_FP64, n, disk=d)
        dr = d.copy()
        for f in ("r", "p", "u", "v"):
            dr.rings[f] = dr.rings[f].astype(np.float32).astype(np.float64)
        flo, _ = run_single(w, abi.PN_FP64, n, disk=dr)
        l1, li = rel(lab(o, "r"), lab(ref, "r"))
        floor, _ = rel(lab(flo, "r"), lab(ref, "r"))
        say(f"  {w.name}: rho L1 {l1:.3e} Linf {li:.3e}, floor {floor:.3e}")
        assert l1 <= max(1e-5, floor), w.name


def test_batch_graph_equals_eager_bitwise():
    ws = members()[:3]
    say("case: batch graph vs eager")
    a, sa = run_batch(ws, 6, graph=False)
    b, sb = run_batch(ws, 6, graph=True)
    for x, y, s, t in zip(a, b, sa, sb):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
        assert s.dt == t.dt and s.simtime == t.simtime


@pytest.mark.parametrize("variants", [(1, 0), (2, 1), (3, 0)])
def test_batch_variants(variants):
    ws = members()[:3]
    say("case: batch variants", variants)
    outs, _ = run_batch(ws, 10, variants=variants)
    for w, o in zip(ws, outs):
        ref, _ = run_single(w, abi.PN_FP32, 10, variants=variants)
        for f in ("r", "u", "v"):
            assert rel(o[f], ref[f])[1] <= 1e-12, (w.name, f)


def test_batch_restart_from_pictures(tmp_path):
    """pn_batch_run_restart: every disk of a fresh batch loaded from its picture and resumed
    reproduces the unbroken batch, with simtime counted from tmovie * startat."""
    ws = members()
    for w in ws:
        w.params.tmovie = 0.25
    disks = [abi.init_disk(w.grid, w.params, w.seed, w.noise) for w in ws]
    paths = [str(tmp_path / f"binary_disk{d}_00002") for d in range(len(ws))]
    say("case: batch restart", len(ws), "disks, startat 2")
    with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
        for ctx, d in zip(b.disks, disks):
            ctx.load(d)
        b.set_variants(0, 1)
        st = b.run_begin()
        b.run(6, st)
        for ctx, path in zip(b.disks, paths):
            ctx.save_snapshot(path)
        t6 = [st[d].simtime for d in range(len(ws))]
        b.run(5, st)
        want = [ctx.store() for ctx in b.disks]
        t11 = [st[d].simtime for d in range(len(ws))]
    with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
        for ctx, d, path in zip(b.disks, disks, paths):
            ctx.load(d)
            ctx.load_snapshot(path)
        b.set_variants(0, 1)
        st = b.run_restart(2)
        assert all(st[d].simtime == 0.5 and st[d].timem == 0.0 for d in range(len(ws)))
        b.run(5, st)
        for d, ctx in enumerate(b.disks):
            got = ctx.store()
            say(f"  disk {d}: simtime {st[d].simtime!r} vs {0.5 + t11[d] - t6[d]!r}")
            assert np.array_equal(got.view(np.uint8), want[d].view(np.uint8)), d
            assert abs(st[d].simtime - (0.5 + t11[d] - t6[d])) <= 1e-12



def moving_members():
    """Disks whose planets move under the disks' forces (plnt_flag 2) beside fixed ones."""
    out = []
    for d, (nr, q, alpha, flag) in enumerate([(64, 1e-3, 3e-3, 2), (96, 3e-4, 0.0, 1), (130, 1e-3, 1e-3, 2),
                                              (48, 1e-4, 0.0, 2)]):
        over = dict(visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05) if alpha > 0 else {}
        w = configs.small(nr, 3 * nr, M_p=q, plnt_flag=flag, **over)
        w.seed = 300 + d
        out.append(w)
    return out


def test_batch_with_moving_planets_equals_single_disks():
    """A planet that moves takes its CFL pass before the step and its disk forces and closing
    kick after the sweeps, inside the batch's captured step: every disk equals its own
    single-context FP32 run, planet included."""
    ws = moving_members()
    n = 30
    say("case: batch with moving planets,", len(ws), "disks,", n, "steps")
    outs, sts = run_batch(ws, n)
    for w, o, s in zip(ws, outs, sts):
        ref, rs = run_single(w, abi.PN_FP32, n)
        say(f"  {w.name} plnt_flag {w.params.plnt_flag}: planet ({s.planet.x!r}, {s.planet.y!r}) vs "
            f"({rs.planet.x!r}, {rs.planet.y!r}); dt {s.dt!r} vs {rs.dt!r}")
        assert s.dt == rs.dt and s.simtime == rs.simtime and s.FrRot == rs.FrRot
        assert s.planet.x == rs.planet.x and s.planet.y == rs.planet.y and s.planet.vx == rs.planet.vx
        assert np.array_equal(o["rot_j"], ref["rot_j"])
        for f in ("r", "u", "v"):
            l1, li = rel(o[f], ref[f])
            assert li <= 1e-12, (w.name, f, l1, li)
    # the planets did move
    assert any(abs(s.planet.x - 1.0) > 0 for w, s in zip(ws, sts) if w.params.plnt_flag == 2)


def test_batch_with_damping_zones_equals_single_disks():
    """Damping zones (kill_wave.cu:3-43) change the state after the sweeps, so a damped disk
    takes its CFL pass before the step and its damping between the batched sweeps and the
    post-step: every disk, damped or not, with a fixed or a moving planet, equals its own
    single-context FP32 run."""
    ws = []
    for d, (nr, q, alpha, flag, kill) in enumerate([(48, 1e-4, 0.0, 1, 1), (64, 3e-4, 3e-3, 1, 1),
                                                    (96, 1e-3, 1e-3, 2, 1), (130, 1e-4, 0.0, 1, 0)]):
        over = dict(visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05) if alpha > 0 else {}
        w = configs.small(nr, 3 * nr, M_p=q, plnt_flag=flag, kill_flag=kill, **over)
        w.seed = 400 + d
        ws.append(w)
    n = 30
    say("case: batch with damping zones,", len(ws), "disks,", n, "steps")
    outs, sts = run_batch(ws, n)
    for w, o, s in zip(ws, outs, sts):
        ref, rs = run_single(w, abi.PN_FP32, n)
        say(f"  {w.name} kill_flag {w.params.kill_flag} plnt_flag {w.params.plnt_flag}: dt {s.dt!r} vs {rs.dt!r}")
        assert s.dt == rs.dt and s.simtime == rs.simtime and s.FrRot == rs.FrRot
        assert s.planet.x == rs.planet.x and s.planet.y == rs.planet.y
        assert np.array_equal(o["rot_j"], ref["rot_j"])
        for f in ("r", "u", "v"):
            l1, li = rel(o[f], ref[f])
            assert li <= 1e-12, (w.name, f, l1, li)


def test_batch_with_radial_boundary_codes_equals_single_disks():
    """Radial boundary codes (set_boundx, boundary.cu:86-206): each disk's ghost rings are
    copied from its interior between the batched pre-step and the radial sweep, so every disk
    -- each with its own codes, one of them damped with a moving planet -- equals its own
    single-context FP32 run."""
    ws = []
    for d, (nr, q, alpha, codes, extra) in enumerate([(48, 1e-4, 0.0, (0, 0), {}), (64, 3e-4, 3e-3, (4, 4), {}),
                                                      (96, 1e-3, 0.0, (1, 2), dict(plnt_flag=2, kill_flag=1)),
                                                      (130, 1e-4, 1e-3, (3, 3), {})]):
        over = dict(visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05) if alpha > 0 else {}
        w = configs.small(nr, 3 * nr, M_p=q, nlft=codes[0], nrgh=codes[1], **over, **extra)
        w.seed = 500 + d
        ws.append(w)
    n = 30
    say("case: batch with radial boundary codes,", len(ws), "disks,", n, "steps")
    outs, sts = run_batch(ws, n)
    for w, o, s in zip(ws, outs, sts):
        ref, rs = run_single(w, abi.PN_FP32, n)
        say(f"  {w.name} codes ({w.params.nlft}, {w.params.nrgh}): dt {s.dt!r} vs {rs.dt!r}")
        assert s.dt == rs.dt and s.simtime == rs.simtime and s.FrRot == rs.FrRot
        assert s.planet.x == rs.planet.x and s.planet.y == rs.planet.y
        assert np.array_equal(o["rot_j"], ref["rot_j"])
        for f in ("r", "u", "v"):
            l1, li = rel(o[f], ref[f])
            assert li <= 1e-12, (w.name, f, l1, li)


def test_batch_in_the_barycentric_frame_equals_single_disks():
    """The barycentric frame (bary_flag, forces.cu:28-40, 116-126) runs kernels of its own, so
    the disks of a batch share it: every disk of a barycentric batch -- fixed and moving planets,
    viscous or not -- equals its own single-context FP32 run."""
    ws = []
    for d, (nr, q, alpha, flag) in enumerate([(48, 1e-3, 0.0, 1), (64, 3e-4, 3e-3, 2), (96, 1e-4, 1e-3, 1)]):
        over = dict(visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05) if alpha > 0 else {}
        w = configs.small(nr, 3 * nr, M_p=q, plnt_flag=flag, bary_flag=1, **over)
        w.seed = 600 + d
        ws.append(w)
    n = 30
    say("case: batch in the barycentric frame,", len(ws), "disks,", n, "steps")
    outs, sts = run_batch(ws, n)
    for w, o, s in zip(ws, outs, sts):
        ref, rs = run_single(w, abi.PN_FP32, n)
        say(f"  {w.name} plnt_flag {w.params.plnt_flag}: dt {s.dt!r} vs {rs.dt!r}")
        assert s.dt == rs.dt and s.simtime == rs.simtime and s.FrRot == rs.FrRot
        assert s.planet.x == rs.planet.x and s.planet.y == rs.planet.y
        assert np.array_equal(o["rot_j"], ref["rot_j"])
        for f in ("r", "u", "v"):
            l1, li = rel(o[f], ref[f])
            assert li <= 1e-12, (w.name, f, l1, li)


def test_batch_of_adiabatic_disks_equals_single_disks():
    """Disks with the energy-consistent adiabatic step (EOS 3) run the batch kernels that carry
    q = P / Sigma as a fourth field; the disks of such a batch -- viscous or not, fixed or moving
    planets, radial boundary codes -- each equal their own single-context FP32 run, P included."""
    ws = []
    for d, (nr, q, alpha, extra) in enumerate([(48, 1e-3, 0.0, {}), (64, 3e-4, 3e-3, dict(plnt_flag=2)),
                                               (96, 1e-4, 1e-3, dict(nlft=1, nrgh=0, init_p_iso=0))]):
        over = dict(visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05) if alpha > 0 else {}
        w = configs.small(nr, 3 * nr, **{"M_p": q, "eos": 3, "gam": 1.4, "init_p_iso": 1, **over, **extra})
        w.seed = 700 + d
        ws.append(w)
    n = 30
    say("case: batch of EOS 3 disks,", len(ws), "disks,", n, "steps")
    outs, sts = run_batch(ws, n)
    for w, o, s in zip(ws, outs, sts):
        ref, rs = run_single(w, abi.PN_FP32, n)
        say(f"  {w.name} visc {w.params.visc_flag} plnt_flag {w.params.plnt_flag}: dt {s.dt!r} vs {rs.dt!r}")
        assert s.dt == rs.dt and s.simtime == rs.simtime and s.FrRot == rs.FrRot
        assert s.planet.x == rs.planet.x and s.planet.y == rs.planet.y
        assert np.array_equal(o["rot_j"], ref["rot_j"])
        for f in ("r", "p", "u", "v"):
            l1, li = rel(o[f], ref[f])
            assert li <= 1e-12, (w.name, f, l1, li)


def test_batch_landing_matches_single_disks():
    """Picture times and end times differ per disk: the batch stops after the first step in
    which any disk lands, and every disk -- landed or not -- equals its own single-context
    run with the landing over the same steps; a second call goes on to the next landing."""
    ws = members()[:3] + moving_members()[:1]
    for w, (tm, te) in zip(ws, [(0.11, 5.0), (0.07, 5.0), (0.5, 0.09), (0.2, 5.0)]):
        w.params.tmovie, w.params.endtime = tm, te
    disks = [abi.init_disk(w.grid, w.params, w.seed, w.noise) for w in ws]
    say("case: batch landing,", len(ws), "disks")
    with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
        for ctx, d in zip(b.disks, disks):
            ctx.load(d)
        b.set_variants(0, 1)
        st = b.run_begin()
        for k in range(len(ws)):
            st[k].land = 1
        singles = []
        for w, d in zip(ws, disks):
            c = abi.Context(w.grid, w.params, abi.PN_FP32)
            c.load(d)
            c.set_variants(0, 1)
            s1 = c.run_begin()
            s1.land = 1
            singles.append((c, s1))
        try:
            for call in range(2):
                n0 = st[0].ncycle
                b.run(1000, st)
                n = st[0].ncycle - n0
                events = [st[k].event for k in range(len(ws))]
                say(f"  call {call}: {n} steps, events {events}")
                assert 0 < n < 1000 and any(events)
                assert all(st[k].ncycle - n0 == n for k in range(len(ws)))
                for k, (w, (c, s1)) in enumerate(zip(ws, singles)):
                    c.run(n, s1)
                    assert s1.ncycle == st[k].ncycle and s1.event == st[k].event, (w.name, s1.ncycle, s1.event)
                    assert s1.simtime == st[k].simtime and s1.dt == st[k].dt and s1.timem == st[k].timem
                    o, ref = b.disks[k].store(), c.store()
                    assert np.array_equal(o["rot_j"], ref["rot_j"])
                    for f in ("r", "u", "v"):
                        l1, li = rel(o[f], ref[f])
                        assert li <= 1e-12, (w.name, f, l1, li)
                if any(st[k].event == 2 for k in range(len(ws))):
                    break
        finally:
            for c, _ in singles:
                c.close()
