"""Throughput of the hydro step on one MI355X, as one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2] [--precision fp32|fp64]

A step is main.cu:345-387 -- CFL dt, FARGO split, planet half kick, r sweep, phi sweep,
integer phi roll, planet update -- advanced on the device with dt and the planet kept
there (a captured HIP graph per step).  value = zone-updates/s = cells * K / wall time of
the K timed steps of the measured path, with the state resident in HBM.

With the default FP32 path the plain FP64 path is run beside it on the same initial state
for the same W + K steps: its throughput gives the speedup, and its final state is the
reference for the relative L1 / Linf errors of Sigma, v_r and v_phi (and p for an adiabatic
disk, e.g. --config c5e3: the 4096 x 12288 disk with the energy-consistent EOS 3), compared
in the lab frame, and for the mass and angular-momentum drift.  The "batched" object times the
parameter sweep of 48 uneven viscous disks (N_r 128..512) advanced as one FP32 batch against
the same disks run one by one on the plain FP64 path, with each disk's error set beside its
sensitivity floor (the drift of the FP64 run from the FP32-rounded initial state).

The roofline entry prices the dominant sweep kernel of the measured path -- its algorithmic
bytes per launch (one read and one write of the minimal state fields of every cell: 3 for
the 2-D isothermal disk, one more in 3-D, one more for a non-isothermal EOS) over its
mean launch time, measured with HIP events on the context's stream -- against HBM
bandwidth (and against the bandwidth a plain 1 GiB device copy reaches on the same GPU,
copy_measured_gbs); traffic and the VALU instruction count per launch come from the
committed rocprofv3 counter runs (profiles/counters.json), the latter priced against the
VALU issue rate, which is what bounds these sweeps.  The CPU baseline times the FP64 oracle
on the host cores on a bounded prefix of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from penguin_2_0_amd import abi, configs  # noqa: E402

METRIC = "zone-updates/s per GPU: FP32-batched vs FP64 plain at L1(ρ)≤1e-5; % HBM peak"
HBM_PEAK_GBS = 8000.0
PROFILE_COUNTERS = ROOT / "profiles" / "counters.json"
# one wave64 FP32 VALU instruction per CU per clock (4 SIMDs x 16 lanes; packed FP32 not used):
# 256 CUs x 2.4 GHz
VALU_PEAK_INSTS = 256 * 2.4e9


def cpu_baseline(w: configs.Workload, budget_s: float = 12.0) -> dict:
    """FP64 oracle on the host cores, on the first steps of the same workload."""
    import oracle

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = oracle.OracleDisk(w.grid, w.params, w.seed, w.noise)
    st = o.run_begin()
    t0 = time.perf_counter()
    o.run(1, st)
    one = time.perf_counter() - t0
    n = max(1, min(40, int(budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.run(n, st)
    dt = time.perf_counter() - t0
    return {"value": w.cells * n / dt, "unit": "zone-updates/s", "cores": cores, "kind": "port",
            "sample": f"{w.name}: steps 2..{n + 1} of the FP64 oracle (OpenMP over pencils), "
                      f"{n} steps in {dt:.2f} s"}


def timed_run(w: configs.Workload, prec: int, steps: int, warmup: int, graph: bool = True) -> dict:
    disk = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    ctx = abi.Context(w.grid, w.params, prec)
    ctx.load(disk)
    st = ctx.run_begin()
    m0, l0 = ctx.totals()
    ctx.run(warmup, st, graph=graph)
    t0 = time.perf_counter()
    ctx.run(steps, st, graph=graph)   # returns after the stream has drained
    wall = time.perf_counter() - t0
    m1, l1 = ctx.totals()
    state = ctx.store()
    variants = ctx.variants()
    km = ctx.time_kernels(5)         # after the snapshot: timing advances the state
    ctx.close()
    return {"wall": wall, "value": w.cells * steps / wall, "state": state, "st": st, "km": km,
            "variants": variants, "dmass": (m1 - m0) / m0, "dangmom": (l1 - l0) / l0}


def batched_run(steps: int, warmup: int, graph: bool = True) -> dict:
    """The parameter sweep of 48 uneven viscous disks: the FP32 batch (one launch per sweep
    for all disks) against every disk's plain FP64 context run on its own, same steps."""
    ws = configs.c4_members()
    cells = sum(w.cells for w in ws)
    disks = [abi.init_disk(w.grid, w.params, w.seed, w.noise) for w in ws]
    with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
        for ctx, d in zip(b.disks, disks):
            ctx.load(d)
        st = b.run_begin()
        b.run(warmup, st, graph=graph)
        t0 = time.perf_counter()
        b.run(steps, st, graph=graph)
        wall = time.perf_counter() - t0
        out32 = [ctx.store() for ctx in b.disks]
        km = b.time_kernels(5)
    wall64, l1s, floors, worst = 0.0, [], [], None
    for k, (w, d, o) in enumerate(zip(ws, disks, out32)):
        with abi.Context(w.grid, w.params, abi.PN_FP64) as ctx:
            ctx.load(d)
            s = ctx.run_begin()
            ctx.run(warmup, s)
            t0 = time.perf_counter()
            ctx.run(steps, s)
            wall64 += time.perf_counter() - t0
            ref = ctx.store()
        # the disk's sensitivity floor: how far the same FP64 run drifts when only its
        # initial state is rounded to FP32
        dr = d.copy()
        for f in ("r", "p", "u", "v"):
            dr.rings[f] = dr.rings[f].astype(np.float32).astype(np.float64)
        with abi.Context(w.grid, w.params, abi.PN_FP64) as ctx:
            ctx.load(dr)
            s = ctx.run_begin()
            ctx.run(warmup + steps, s)
            floors.append(rel_err(lab(ctx.store(), "r"), lab(ref, "r"))[0])
        l1, li = rel_err(lab(o, "r"), lab(ref, "r"))
        l1s.append(l1)
        if worst is None or l1 >= worst["l1_rho"]:
            worst = {"disk": w.name, "l1_rho": l1, "linf_rho": li, "l1_vphi": rel_err(lab(o, "v"), lab(ref, "v"))[0],
                     "fp64_from_fp32_rounded_start_l1_rho": floors[-1]}
    value = cells * steps / wall
    v64 = cells * steps / wall64
    l1s, floors = np.array(l1s), np.array(floors)
    return {"workload": f"c4_batch_{len(ws)}_disks", "cells": cells, "steps": steps, "warmup": warmup,
            "value": value, "unit": "zone-updates/s", "ms_per_step": wall / steps * 1e3, "kernels_ms": km,
            "fp64_plain_value": v64, "speedup_vs_fp64_plain": value / v64,
            "accuracy_vs_fp64": {"after_steps": warmup + steps, "median_l1_rho": float(np.median(l1s)),
                                 "disks_l1_rho_le_1e-5": int((l1s <= 1e-5).sum()),
                                 "median_floor_l1_rho": float(np.median(floors)),
                                 "disks_floor_le_1e-5": int((floors <= 1e-5).sum()),
                                 "disks_l1_rho_le_max_1e-5_floor": int((l1s <= np.maximum(1e-5, floors)).sum()),
                                 "median_l1_over_floor": float(np.median(l1s / np.maximum(floors, 1e-300))),
                                 "worst": worst}}


def copy_probe() -> int:
    """The library's plain copy kernel on 256 MiB, 5 times after one warm-up: run under the
    HBM byte counters to scale them (the sweeps' traffic is quoted in the copy's units)."""
    gbs = abi.copy_bandwidth(256 << 20, 5)
    print(json.dumps({"copy_probe": {"bytes": 256 << 20, "launches": 6, "gbs": gbs}}), flush=True)
    return 0


def lab(state: np.ndarray, f: str) -> np.ndarray:
    """Field f with each ring's data rolled to lab columns (data j sits at j + rot_j)."""
    a = state[f]
    return np.stack([np.roll(a[i], int(state["rot_j"][i])) for i in range(a.shape[0])])


def rel_err(a: np.ndarray, b: np.ndarray) -> tuple[float, float]:
    d = np.abs(a - b)
    return float(d.sum() / np.abs(b).sum()), float(d.max() / np.abs(b).max())


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-ref", action="store_true", help="skip the FP64 leg of an FP32 run")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched parameter-sweep leg")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--floor", action="store_true",
                    help="also run the FP64 path from the FP32-rounded initial state: the flow's own "
                         "sensitivity, the error floor no FP32 arithmetic can go below")
    ap.add_argument("--copy-probe", action="store_true", help="only run a plain device copy of known size")
    ap.add_argument("--eager", action="store_true",
                    help="launch every step's kernels directly instead of replaying a captured HIP graph "
                         "(the same kernels; for tools that do not follow graph launches)")
    a = ap.parse_args(argv)
    if a.copy_probe:
        return copy_probe()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported (the disk is not "
              f"sharded across GPUs in this build)", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    w = getattr(configs, a.config)()
    prec = abi.PN_FP64 if a.precision == "fp64" else abi.PN_FP32
    run = timed_run(w, prec, a.steps, a.warmup, graph=not a.eager)
    value = run["value"]
    km = run["km"]

    # dominant kernel, timed on the context's stream
    dom = max(("sweep_x", "sweep_y", "sweep_z"), key=lambda k: km[k])
    wbytes = 8 if prec == abi.PN_FP64 else 4
    # the minimal state: Sigma (rho), v_r, v_phi; + v_theta in 3-D; + P unless isothermal
    fields = 3 + (1 if w.params.ndim == 3 else 0) + (1 if w.params.eos != 0 else 0)
    bytes_per_launch = 2 * fields * wbytes * w.cells
    achieved = bytes_per_launch / (km[dom] * 1e-3) / 1e9
    # HBM bytes and VALU instructions per launch of the dominant kernel (its variant) from the
    # committed rocprofv3 counter runs
    var = run["variants"][1] if dom == "sweep_y" else run["variants"][0]
    prof = {}
    if PROFILE_COUNTERS.exists():
        prof = json.loads(PROFILE_COUNTERS.read_text()).get(f"{w.name}:{a.precision}:{dom}:v{var}", {})
    traffic = prof.get("traffic")
    copy_gbs = abi.copy_bandwidth()
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": dom, "variant": var, "kernel_ms": km[dom], "bytes_per_launch": bytes_per_launch,
                "traffic_over_algorithmic": traffic / bytes_per_launch if traffic else None, "kernels_ms": km,
                "copy_measured_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs}
    if "sq_insts_valu" in prof:
        # what actually bounds the sweeps: VALU issue, against one instruction per CU per clock
        ips = prof["sq_insts_valu"] / (km[dom] * 1e-3)
        roofline["valu"] = {"insts_per_launch": prof["sq_insts_valu"], "achieved_insts_per_s": ips,
                            "peak_insts_per_s": VALU_PEAK_INSTS, "frac": ips / VALU_PEAK_INSTS}
    out = {
        "metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": run["wall"] / a.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == abi.PN_FP64 else "f32",
        "data": "synthetic (seeded disk initial condition of the workload)",
        "config": {"workload": w.name, "cells": w.cells,
                   "path": "plain FP64" if prec == abi.PN_FP64 else "FP32 bulk (single disk)", "graph": not a.eager},
        "roofline": roofline,
        "conservation": {"mass_drift": run["dmass"], "angmom_drift": run["dangmom"]},
    }
    if prec == abi.PN_FP32 and not a.no_ref:
        ref = timed_run(w, abi.PN_FP64, a.steps, a.warmup, graph=not a.eager)
        errs = {f: rel_err(lab(run["state"], f), lab(ref["state"], f)) for f in ("r", "u", "v", "p")}
        out["fp64_plain"] = {"value": ref["value"], "ms_per_step": ref["wall"] / a.steps * 1e3,
                             "speedup": value / ref["value"], "kernels_ms": ref["km"],
                             "mass_drift": ref["dmass"], "angmom_drift": ref["dangmom"]}
        out["accuracy_vs_fp64"] = {
            "after_steps": a.warmup + a.steps,
            "l1_rho": errs["r"][0], "linf_rho": errs["r"][1],
            "l1_vr": errs["u"][0], "linf_vr": errs["u"][1],
            "l1_vphi": errs["v"][0], "linf_vphi": errs["v"][1],
            **({"l1_p": errs["p"][0], "linf_p": errs["p"][1]} if w.params.eos != 0 else {}),
            "rot_j_equal": bool(np.array_equal(run["state"]["rot_j"], ref["state"]["rot_j"])),
            "dt_rel_diff": abs(run["st"].dt - ref["st"].dt) / ref["st"].dt,
        }
        if a.floor:
            d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
            for f in ("r", "p", "u", "v"):
                d.rings[f] = d.rings[f].astype(np.float32).astype(np.float64)
            with abi.Context(w.grid, w.params, abi.PN_FP64) as ctx:
                ctx.load(d)
                s = ctx.run_begin()
                ctx.run(a.warmup + a.steps, s, graph=not a.eager)
                fl = ctx.store()
            out["accuracy_vs_fp64"]["floor"] = {
                f"l1_{k}": rel_err(lab(fl, f), lab(ref["state"], f))[0]
                for k, f in (("rho", "r"), ("vr", "u"), ("vphi", "v"), ("p", "p"))}
    if prec == abi.PN_FP32 and not a.no_batch:
        out["batched"] = batched_run(a.batch_steps, a.warmup, graph=not a.eager)
    if rank == 0 and not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    if rank == 0:
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
