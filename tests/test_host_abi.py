"""Host side of the C ABI, no GPU needed: the library loads and exports every declared
entry point, parameter defaults equal global.h, the hydr_ring layout, synthetic disks equal
the oracle's byte for byte, the planet's host integrator, and argument errors."""
from __future__ import annotations

import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
from penguin_2_0_amd import abi, configs
from penguin_2_0_amd.abi import PnBody, PnGrid, PnParams

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "penguin.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pn_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = abi.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_defaults_are_global_h():
    p = abi.params_default()
    twopi = 6.283185307179586476925286766559
    hpi = 1.570796326794896619231321691639751442
    assert (p.ndim, p.eos, p.flat_flag, p.plnt_flag, p.frrot_flag, p.visc_flag, p.fargo_flag) == (2, 0, 1, 2, 1, 1, 1)
    assert (p.ngeomx, p.ngeomy, p.ngeomz, p.grspx, p.grspy, p.grspz) == (1, 3, 0, 1, 0, 0)
    assert (p.p_alpha, p.p_beta, p.ss_alpha, p.sc_h) == (1.5, 0.5, 0.0001, 0.05)
    assert p.vis_nu == 0.0001 * 0.05 * 0.05
    assert (p.M_p, p.R_p, p.t_growth, p.rs_fac) == (0.001, 1.0, 1000.0 * twopi, 0.5)
    assert (p.gam, p.courant, p.endtime, p.tmovie) == (1.0, 0.5, 10000.0 * twopi, 200.0 * twopi)
    assert (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax) == (0.2, 3.0, 0.0, twopi, hpi - 0.065, hpi)


def test_ring_layout():
    lib = abi.load_library()
    for J in (1, 12, 768, 3072, 12288):
        assert lib.pn_ring_bytes(J) == 72 * J + 96 == abi.ring_dtype(J).itemsize
    assert lib.pn_ring_bytes(768) == 55392 and lib.pn_ring_bytes(12288) == 884832
    dt = abi.ring_dtype(10)
    off = {n: dt.fields[n][1] for n in dt.names}
    # member order of variable_types.h:67-97
    assert off["r"] == 0 and off["w"] == 320 and off["x"] == 400 and off["xc"] == 424
    assert off["y"] == 432 and off["yc"] == 672 and off["z"] == 752 and off["zc"] == 776
    assert off["i"] == 784 and off["k"] == 788 and off["rot_v"] == 792 and off["res_v"] == 800
    assert off["rot_j"] == 808 and off["inc_j"] == 812


CASES = {
    "c1": lambda: configs.c1(),
    "viscous": lambda: configs.c4_members()[0],
    "adiabatic_small": lambda: configs.Workload("adi", PnGrid(40, 96, 1), configs.c5().params, 5, 1),
    "sph3d_small": lambda: configs.Workload("sph", PnGrid(16, 36, 8), configs.c3().params, 3, 1),
}


@pytest.mark.parametrize("case", list(CASES))
def test_init_disk_matches_oracle_bytes(case):
    w = CASES[case]()
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    o = oracle.OracleDisk(w.grid, w.params, w.seed, w.noise)
    assert np.array_equal(d.rings.view(np.uint8), o.rings.view(np.uint8))
    assert np.array_equal(d.lft.view(np.uint8), o.lft.view(np.uint8))
    assert np.array_equal(d.rgh.view(np.uint8), o.rgh.view(np.uint8))
    if w.params.ndim == 3:
        assert np.array_equal(d.udr.view(np.uint8), o.udr.view(np.uint8))
        assert np.array_equal(d.top.view(np.uint8), o.top.view(np.uint8))
    ov = np.ctypeslib.as_array(C.cast(o.val, C.POINTER(C.c_double)), shape=(len(o.val), 5))
    assert np.array_equal(d.val_array(), ov)
    # ring order i + imax k and the ring's own (i, k)
    n = np.arange(w.grid.imax * w.grid.kmax)
    assert np.array_equal(d.rings["i"], n % w.grid.imax) and np.array_equal(d.rings["k"], n // w.grid.imax)


def test_planet_host_integrator_matches_oracle():
    lib = abi.load_library()
    olib = oracle.lib()
    for P in (configs.c1().params, abi.params_default()):
        b1, b2 = PnBody(), PnBody()
        f1, f2 = C.c_double(), C.c_double()
        lib.pn_init_planet(C.byref(P), C.byref(b1), C.byref(f1))
        olib.orc_init_planet(C.byref(P), C.byref(b2), C.byref(f2))
        assert bytes(b1) == bytes(b2) and f1.value == f2.value
        # give it a radial kick so drift and kick do something
        for b in (b1, b2):
            b.fx, b.fy, b.vx = 1e-3, 2e-4, 1e-3
        for dt, kick in ((0.01, 1), (0.02, 0), (0.01, 1)):
            lib.pn_kick_drift_kick(C.byref(P), C.byref(b1), C.byref(f1), dt, kick)
            olib.orc_kick_drift_kick(C.byref(P), C.byref(b2), C.byref(f2), dt, kick)
            assert bytes(b1) == bytes(b2) and f1.value == f2.value
    # init_planet closed forms: Keplerian planet at R_p, co-rotating frame
    P = configs.c1().params
    b, f = PnBody(), C.c_double()
    lib.pn_init_planet(C.byref(P), C.byref(b), C.byref(f))
    assert b.x == 1.0 and b.y == math.pi and b.vy == 0.0 and f.value == 1.0
    assert abs(b.rs - 0.6 * 0.05) < 1e-16 and b.m == P.M_p   # t_growth 0: full mass


def test_argument_errors_before_any_device_call():
    lib = abi.load_library()
    g = PnGrid(64, 96, 1)
    for bad in (dict(ndim=3), dict(eos=7), dict(ngeomx=2), dict(fargo_flag=5), dict(plnt_flag=3)):
        P = configs.disk2d(64, 96).replace(**bad)
        h = C.c_void_p()
        rc = lib.pn_create(C.byref(g), C.byref(P), 0, C.byref(h))
        assert rc in (abi.PN_ERR_ARG, abi.PN_ERR_UNSUPPORTED), (bad, rc)
        assert lib.pn_last_error(h)
        lib.pn_destroy(h)
    with pytest.raises(abi.PnError):
        abi.Context(PnGrid(64, 8, 1), configs.disk2d(64, 8))   # fewer phi cells than two halos
    # the synthetic-disk builder rejects a grid that contradicts ndim
    d = abi.Disk(PnGrid(8, 24, 1), configs.disk2d(8, 24))
    rc = lib.pn_init_disk(C.byref(PnGrid(8, 24, 2)), C.byref(configs.disk2d(8, 24)), 1, 0.0,
                          d.rings.ctypes.data_as(C.c_void_p), d.lft.ctypes.data_as(C.c_void_p),
                          d.rgh.ctypes.data_as(C.c_void_p), None, None, d.val)
    assert rc == abi.PN_ERR_ARG
    # priming a run needs a context and a state record
    assert lib.pn_run_begin(None, None) == abi.PN_ERR_ARG
    assert lib.pn_run_restart(None, 2, None) == abi.PN_ERR_ARG


def _gpu_visible() -> bool:
    import os
    return any(Path(p).exists() for p in ("/dev/kfd",)) and os.access("/dev/kfd", os.R_OK | os.W_OK)


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-device error path")
def test_create_without_device_fails_loudly():
    with pytest.raises(abi.PnError) as e:
        abi.Context(PnGrid(32, 48, 1), configs.disk2d(32, 48))
    assert e.value.code == abi.PN_ERR_HIP


def test_batch_argument_errors_before_any_device_call():
    g = [PnGrid(64, 192, 1), PnGrid(48, 144, 1)]
    p = [configs.disk2d(64, 192), configs.disk2d(48, 144)]
    # the batched launch is built for the FP32 path only
    with pytest.raises(abi.PnError) as e:
        abi.Batch(g, p, precision=abi.PN_FP64)
    assert e.value.code == abi.PN_ERR_UNSUPPORTED
    with pytest.raises(abi.PnError) as e:
        abi.Batch([], [])
    assert e.value.code == abi.PN_ERR_ARG
    # the disks of a batch share the frame
    with pytest.raises(abi.PnError) as e:
        abi.Batch([PnGrid(32, 48, 1)] * 2, [configs.disk2d(32, 48), configs.disk2d(32, 48, bary_flag=1)])
    assert e.value.code == abi.PN_ERR_UNSUPPORTED
    # batches hold 2-D disks
    p3 = configs.disk2d(24, 72).replace(ndim=3, ngeomx=2, ngeomy=4, ngeomz=5, zmin=1.4, zmax=1.7)
    with pytest.raises(abi.PnError) as e:
        abi.Batch([PnGrid(24, 72, 12)], [p3])
    assert e.value.code == abi.PN_ERR_UNSUPPORTED


def test_fp32_parameter_errors_before_any_device_call():
    # the FP32 path is built for the isothermal closure and the 2-D EOS 3 disk
    for over in (dict(eos=1), dict(eos=2)):
        with pytest.raises(abi.PnError) as e:
            abi.Context(PnGrid(32, 48, 1), configs.disk2d(32, 48, gam=1.4, **over), abi.PN_FP32)
        assert e.value.code == abi.PN_ERR_UNSUPPORTED
    with pytest.raises(abi.PnError) as e:
        abi.Batch([PnGrid(32, 48, 1)], [configs.disk2d(32, 48, eos=1, gam=1.4)])
    assert e.value.code == abi.PN_ERR_UNSUPPORTED
    # the disks of a batch share the closure (EOS 3 runs kernels of its own)
    with pytest.raises(abi.PnError) as e:
        abi.Batch([PnGrid(32, 48, 1)] * 2, [configs.disk2d(32, 48), configs.disk2d(32, 48, eos=3, gam=1.4)])
    assert e.value.code == abi.PN_ERR_UNSUPPORTED
    # parameter values outside their ranges are argument errors on either path
    for over in (dict(bary_flag=2), dict(nlft=5), dict(nrgh=-1)):
        with pytest.raises(abi.PnError) as e:
            abi.Context(PnGrid(32, 48, 1), configs.disk2d(32, 48, **over), abi.PN_FP32)
        assert e.value.code == abi.PN_ERR_ARG


def test_eos3_is_two_dimensional():
    p3 = configs.c3().params.replace(eos=3, gam=1.4)
    with pytest.raises(abi.PnError) as e:
        abi.Context(PnGrid(24, 72, 12), p3)
    assert e.value.code == abi.PN_ERR_UNSUPPORTED
    with pytest.raises(abi.PnError) as e:
        abi.Context(PnGrid(32, 48, 1), configs.disk2d(32, 48, eos=4, gam=1.4))
    assert e.value.code == abi.PN_ERR_ARG


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-device error path")
def test_fp32_and_batch_without_device_fail_loudly():
    with pytest.raises(abi.PnError) as e:
        abi.Context(PnGrid(32, 48, 1), configs.disk2d(32, 48), abi.PN_FP32)
    assert e.value.code == abi.PN_ERR_HIP
    with pytest.raises(abi.PnError) as e:
        abi.Batch([PnGrid(32, 48, 1)], [configs.disk2d(32, 48)])
    assert e.value.code == abi.PN_ERR_HIP


def test_lab_fields_unrotates_like_the_reader():
    """pn_lab_fields against a brute-force roll: lab column j of ring (i, k) is data index
    (j - rot_j) mod jmax, output [i][j][k]; ring order does not matter."""
    for grid, p in ((PnGrid(6, 20, 1), configs.disk2d(6, 20)),
                    (PnGrid(5, 16, 3), configs.disk2d(5, 16).replace(ndim=3, ngeomx=2, ngeomy=4, ngeomz=5,
                                                                      zmin=1.4, zmax=1.7))):
        d = abi.init_disk(grid, p, 3, 1e-3)
        rng = np.random.default_rng(1)
        d.rings["rot_j"] = rng.integers(0, grid.jmax, size=len(d.rings))
        for f in ("u", "v", "w"):
            d.rings[f] = rng.standard_normal(d.rings[f].shape)
        out = abi.lab_fields(d.rings, grid)
        for f in ("r", "p", "u", "v", "w"):
            ref = np.empty((grid.imax, grid.jmax, grid.kmax))
            for ring in d.rings:
                for j in range(grid.jmax):
                    ref[ring["i"], j, ring["k"]] = ring[f][(j - ring["rot_j"]) % grid.jmax]
            assert np.array_equal(out[f], ref), f
        perm = rng.permutation(len(d.rings))
        out2 = abi.lab_fields(d.rings[perm], grid)
        assert all(np.array_equal(out[f], out2[f]) for f in out)
        # a repeated (i, k) and a short buffer are errors
        bad = d.rings.copy()
        bad[1]["i"], bad[1]["k"] = bad[0]["i"], bad[0]["k"]
        with pytest.raises(abi.PnError):
            abi.lab_fields(bad, grid)
        with pytest.raises(abi.PnError):
            abi.lab_fields(d.rings[:-1], grid)


def test_output_names_follow_the_reference_formulas():
    """create_label (output.cu:57-88) and frame_num (global.cu:27-38), evaluated by hand."""
    c1 = configs.c1()
    # 256x768x1; round(0.05/0.001) = 50; round(1e-4/3e-6) = 33; round(100*0.6) = 60;
    # int(0.4*100) = 40; int(2.5*100) = 250; EOS 0 with p_beta 1: int(10) -> _b10
    assert abi.label(c1.grid, c1.params) == "256x768x1_c50_p33_s60_i40_o250_b10_rot_fargo"
    # int_to_string takes an int: 0.29*100 = 28.999999999999996 is truncated to 28
    assert "_i28_" in abi.label(c1.grid, c1.params.replace(xmin=0.29))
    p = c1.params.replace(plnt_flag=2, visc_flag=1, ss_alpha=1e-3, eos=1, gam=1.4, frrot_flag=0, fargo_flag=2)
    assert abi.label(c1.grid, p) == "256x768x1_c50_pm33_s60_i40_o250_a10_g14"
    # p_alpha + p_beta/2 - 3/2 == 3/2 and p_beta == 0
    assert abi.label(c1.grid, c1.params.replace(p_alpha=3.0, p_beta=0.0, plnt_flag=0)) == \
        "256x768x1_c50_i40_o250_flatvor_iso_rot_fargo"
    assert abi.label(c1.grid, c1.params.replace(eos=2, gam=1.4)).endswith("_o250_ad_rot_fargo")
    assert abi.label(c1.grid, c1.params.replace(eos=3, gam=1.4)).endswith("_o250_ad3_rot_fargo")
    assert [abi.frame_name(n) for n in (0, 7, 42, 1234, 99999)] == ["00000", "00007", "00042", "01234", "99999"]
    for bad in (-1, 100000):
        with pytest.raises(abi.PnError):
            abi.frame_name(bad)


def test_para_and_initial_files_read_back_as_the_reader_reads_them(tmp_path):
    """write_para_file (output.cu:90-107) parsed in read_para's token order (read/post.cu:72-89);
    write_initial_file (output.cu:109-120) against the oracle's radial grid."""
    w = configs.c1()
    lab_ = abi.label(w.grid, w.params)
    abi.write_para(w.grid, w.params, "/work/run", lab_, tmp_path / "para.dat")
    t = (tmp_path / "para.dat").read_text().split()
    assert len(t) == 18
    assert [int(x) for x in t[:4]] == [w.grid.imax, w.grid.jmax, w.grid.kmax, 1]
    P = w.params
    for got, want in zip((t[4], t[5], t[6]), (P.gam, P.sc_h, P.tmovie)):
        assert abs(float(got) - want) <= 5e-6 * abs(want)   # 6 significant digits
    assert t[7] == "/work/run" and t[8] == lab_
    assert [int(x) for x in t[9:12]] == [P.grspx, P.grspy, P.grspz]
    for got, want in zip(t[12:], (P.xmin, P.xmax, P.ymin, P.ymax, P.zmin, P.zmax)):
        assert abs(float(got) - want) <= 5e-6 * abs(want)
    abi.write_initial(w.grid, w.params, tmp_path / "initial.dat")
    rows = np.loadtxt(tmp_path / "initial.dat")
    n = w.grid.imax
    xa, xc, dx = np.zeros(n), np.zeros(n), np.zeros(n)
    oracle.lib().orc_grid(n, P.xmin, P.xmax, P.grspx, oracle.dptr(xa), oracle.dptr(xc), oracle.dptr(dx))
    assert rows.shape == (n, 2)
    np.testing.assert_allclose(rows[:, 0], xa, rtol=5e-6)
    np.testing.assert_allclose(rows[:, 1], dx, rtol=5e-6)


def test_run_loop_driver_parses_its_arguments_without_a_device():
    """penguin_main (the reference's run loop over the C ABI) is built beside the library; its
    argument errors end before any device call."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent.parent / "penguin_2_0_amd" / "penguin_main"
    assert exe.exists()
    res = subprocess.run([str(exe), "--help"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0 and "usage: penguin_main" in res.stdout
    for bad in ([], ["--out", "x", "--set", "no_such_field=1"], ["--out", "x", "--set", "plnt_flag=1.5"],
                ["--out", "x", "--nprin", "0"], ["--bogus"]):
        res = subprocess.run([str(exe), *bad], capture_output=True, text=True, timeout=60)
        assert res.returncode == 2, (bad, res.stdout, res.stderr)


def test_run_loop_driver_sets_every_parameter(tmp_path):
    """Every pn_params field can be set from penguin_main's command line: each --set parses
    (a parse error exits 2 before any device call; without a device the run then stops at the
    context with exit 1)."""
    import subprocess
    from pathlib import Path
    from penguin_2_0_amd.abi import PnParams
    exe = Path(__file__).resolve().parent.parent / "penguin_2_0_amd" / "penguin_main"
    p = configs.disk2d(16, 48, endtime=0.2, tmovie=0.1)   # short, should a device be present
    args = [str(exe), "--out", str(tmp_path / "run"), "--imax", "16", "--jmax", "48"]
    for name, _ in PnParams._fields_:
        args += ["--set", f"{name}={getattr(p, name)!r}"]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode != 2, res.stderr
