"""Two short last windows per wavefront (pn_set_packing) change nothing but the schedule: the
FP32 radial and theta sweeps with packing on and off give the same bytes, step after step,
for every pencil geometry that packs -- the 2-D cylindrical radial sweep (inviscid, viscous,
EOS 3), the 3-D spherical radial and theta sweeps (inviscid and viscous), and a batch of uneven
disks -- in every radial variant.  Short last windows: 128 rings (24 cells with halos), 48 rings (one 60-cell window),
24 x 72 x 12 (one 36-cell window)."""
from __future__ import annotations

import numpy as np
import pytest

from penguin_2_0_amd import abi, configs

pytestmark = pytest.mark.gpu

HPI = 1.570796326794896619231321691639751442


def say(*a):
    print(*a, flush=True)


def disk3d(nr, nphi, nth):
    p = configs.disk2d(nr, nphi, q=3.0e-5).replace(
        ndim=3, ngeomx=2, ngeomy=4, ngeomz=5, p_alpha=1.5, p_beta=1.0, zmin=HPI - 0.15, zmax=HPI + 0.15)
    return configs.Workload(f"disk3d_{nr}x{nphi}x{nth}", abi.PnGrid(nr, nphi, nth), p, seed=11, steps=10)


def visc(nr, nphi, alpha):
    w = configs.small(nr, nphi, visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05)
    w.name += "_visc"
    return w


def eos3(nr, nphi):
    w = configs.small(nr, nphi, eos=3, gam=1.4, init_p_iso=1, M_p=1e-3)
    w.name += "_eos3"
    return w


def visc3d(nr, nphi, nth):
    w = disk3d(nr, nphi, nth)
    w.params = w.params.replace(visc_flag=1, ss_alpha=3e-3, vis_nu=3e-3 * 0.05 * 0.05)
    w.name += "_visc"
    return w


CASES = [configs.small(128, 384, M_p=1e-3), configs.small(48, 144), visc(128, 384, 3e-3), eos3(128, 384),
         disk3d(24, 72, 12), visc3d(24, 72, 12)]


def run(w, disk, pack, var_x, n):
    with abi.Context(w.grid, w.params, abi.PN_FP32) as ctx:
        ctx.load(disk)
        ctx.set_packing(pack)
        ctx.set_variants(var_x, 1 if w.params.visc_flag else 3)
        st = ctx.run_begin()
        ctx.run(n, st, graph=True)
        return ctx.store(), st


@pytest.mark.parametrize("w", CASES, ids=lambda w: w.name)
def test_packed_tails_give_the_same_bytes(w):
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    for vx in range(4):
        say("case: packing", w.name, "radial variant", vx)
        a, sa = run(w, d, True, vx, 12)
        b, sb = run(w, d, False, vx, 12)
        assert sa.dt == sb.dt and sa.simtime == sb.simtime
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (w.name, vx)


def test_packed_tails_in_a_batch():
    ws = []
    for d, (nr, q, alpha) in enumerate([(128, 1e-3, 3e-3), (48, 1e-4, 0.0), (256, 3e-4, 1e-3), (192, 1e-5, 1e-2)]):
        over = dict(visc_flag=1, ss_alpha=alpha, vis_nu=alpha * 0.05 * 0.05) if alpha > 0 else {}
        w = configs.small(nr, 3 * nr, M_p=q, **over)
        w.seed = 200 + d
        ws.append(w)
    disks = [abi.init_disk(w.grid, w.params, w.seed, w.noise) for w in ws]
    outs = []
    for pack in (True, False):
        say("case: packing in a batch,", "on" if pack else "off")
        with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
            for ctx, disk in zip(b.disks, disks):
                ctx.load(disk)
                ctx.set_packing(pack)
            b.set_variants(0, 1)
            st = b.run_begin()
            b.run(12, st, graph=True)
            outs.append([ctx.store() for ctx in b.disks])
    for w, a, c in zip(ws, *outs):
        assert np.array_equal(a.view(np.uint8), c.view(np.uint8)), w.name
