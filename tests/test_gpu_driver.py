"""The C++ run loop (penguin_main: kernel(), main.cu:113-437, over the C ABI) against the same
calls made from Python: every picture it writes -- on a fresh run and on a run resumed from a
picture -- holds the bytes the Python-driven loop saves, and its time file ends at the end time
with the mass of the final state."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from penguin_2_0_amd import abi, configs

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent.parent / "penguin_2_0_amd" / "penguin_main"


def say(*a):
    print(*a, flush=True)


def workload():
    w = configs.small(48, 144)
    w.params.tmovie = 0.5
    w.params.endtime = 1.6
    return w


def driver(w, out, prec, *extra):
    args = [str(DRIVER), "--out", str(out), "--imax", str(w.grid.imax), "--jmax", str(w.grid.jmax), "--kmax",
            str(w.grid.kmax), "--seed", str(w.seed), "--noise", repr(w.noise), "--nprin", "5"]
    if prec == abi.PN_FP32:
        args.append("--fp32")
    for name, val in w.params.as_dict().items():
        args += ["--set", f"{name}={val!r}"]
    res = subprocess.run(args + list(extra), capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    return res.stdout


def python_pictures(w, prec, d, first, restart_from=None):
    """The same loop through the Python binding: pictures first, first + 1, ... as bytes."""
    pics = {}
    tmp = Path(first)
    with abi.Context(w.grid, w.params, prec) as ctx:
        ctx.load(d)
        if restart_from is None:
            npic = 0
            ctx.save_snapshot(str(tmp))
            pics[npic] = tmp.read_bytes()
            st = ctx.run_begin()
        else:
            npic, path = restart_from
            ctx.load_snapshot(str(path))
            st = ctx.run_restart(npic)
        st.land = 1
        while True:
            ctx.run(1 << 20, st, graph=True)
            if st.event >= 1:
                npic += 1
                ctx.save_snapshot(str(tmp))
                pics[npic] = tmp.read_bytes()
            if st.event == 2:
                break
        mass = ctx.reduce(st.planet, 0.0, 0)
    return pics, st, mass


@pytest.mark.parametrize("prec", [abi.PN_FP64, abi.PN_FP32], ids=["fp64", "fp32"])
def test_driver_pictures_equal_the_python_loop(tmp_path, prec):
    assert DRIVER.exists(), "build penguin_main first (python -m penguin_2_0_amd.build --lib)"
    w = workload()
    say("case: penguin_main", w.name, "precision", prec, "tmovie", w.params.tmovie, "endtime", w.params.endtime)
    out = tmp_path / "run"
    log = driver(w, out, prec)
    say("  " + " | ".join(line for line in log.splitlines() if "check point" in line))
    label = abi.label(w.grid, w.params)
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    pics, st, mass = python_pictures(w, prec, d, tmp_path / "py_snap")
    assert sorted(pics) == [0, 1, 2, 3, 4]   # pictures at 0.5, 1.0, 1.5 and the end time 1.6
    for npic, want in pics.items():
        got = (out / "binary" / f"binary_{label}_{abi.frame_name(npic)}").read_bytes()
        assert got == want, npic
    assert (out / "read" / f"para_{label}.dat").exists() and (out / "files" / f"initial_{label}.dat").exists()
    rows = np.loadtxt(out / "files" / f"time_{label}.dat", ndmin=2)
    say(f"  time file: {rows.shape[0]} rows, last simtime {rows[-1, 0]!r} mass {rows[-1, 1]!r}; python {st.simtime!r} {mass!r}")
    assert rows[-1, 0] == st.simtime == w.params.endtime and rows[-1, 1] == mass
    assert rows.shape[1] == 2 + 9   # simtime, mass, the torque and the planet's orbit
    # resumed from picture 2 (main.cu:244-330): the pictures after it are the Python loop's
    res = out / "binary" / f"binary_{label}_{abi.frame_name(2)}"
    keep = tmp_path / "picture2"
    keep.write_bytes(res.read_bytes())
    driver(w, out, prec, "--restart", "2")
    again, _, _ = python_pictures(w, prec, d, tmp_path / "py_snap2", restart_from=(2, keep))
    assert sorted(again) == [3, 4]
    for npic, want in again.items():
        got = (out / "binary" / f"binary_{label}_{abi.frame_name(npic)}").read_bytes()
        assert got == want, ("restart", npic)
