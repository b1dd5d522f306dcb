"""Throughput of the hydro step on one MI355X, as one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4|c2|c3|c5e3]

A step is main.cu:345-387 -- CFL dt, FARGO split, planet half kick, r sweep, phi sweep,
integer phi roll, planet update -- advanced on the device with dt and the planet kept
there (a captured HIP graph per step).  value = zone-updates/s = cells * K / wall time of
the K timed steps, with the state resident in HBM.

Default workload (--config c4): the parameter sweep of 48 independent viscous 2-D disks with
uneven grids (N_r 128..512, N_phi = 3 N_r; planet q and alpha drawn from a seed), advanced
as one FP32-bulk batch -- one launch per sweep for all disks.  Beside it, on the same disks:

  fp64_plain        every disk on the plain FP64 path, one after the other, same W + K steps:
                    its throughput and the batch's speedup over it;
  accuracy_vs_fp64  a fresh batch and every disk's FP64 run, 1000 steps each (whatever K is),
                    relative L1 / Linf of Sigma in the lab frame per disk, each set beside the
                    disk's sensitivity floor (the FP64 run started from the initial state
                    rounded to FP32: how far the flow itself carries a rounding of its start)
                    and the widest of three FP64 runs whose start is nudged by 2^-25 (every
                    value times 1 + 2^-25 u, u uniform in [-1, 1], seeds 1..3);
  single_disk       the 1024 x 3072 disk on the FP32-bulk and the FP64 path (K timed steps
                    each), with their errors after 1000 steps;
  roofline          the batch's dominant sweep kernel: algorithmic bytes per launch (one read
                    and one write of Sigma, v_r, v_phi of every cell, FP32) over its mean
                    duration, timed with HIP events on the batch's stream, against 8 TB/s;
                    traffic and VALU instructions per launch from the committed rocprofv3
                    counter runs (profiles/counters.json);
  cpu_baseline      the FP64 oracle on the host cores (OpenMP over pencils) and on one core,
                    on a bounded sample of the same 48 disks.

--config c2 / c3 / c5e3 print the same line for one disk (1024 x 3072; the 3-D 256 x 768 x 64
disk; the 4096 x 12288 adiabatic disk with EOS 3) on the FP32-bulk path with the FP64 path
beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
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
ACCURACY_STEPS = 1000


def progress(*a) -> None:
    """One line per finished leg on stderr (the JSON result alone goes to stdout)."""
    print("bench:", *a, file=sys.stderr, flush=True)


def lab(state: np.ndarray, f: str) -> np.ndarray:
    """Field f with each ring's data rolled to lab columns (data j sits at j + rot_j)."""
    a = state[f]
    return np.stack([np.roll(a[i], int(state["rot_j"][i])) for i in range(a.shape[0])])


def rel_err(a: np.ndarray, b: np.ndarray) -> tuple[float, float]:
    d = np.abs(a - b)
    return float(d.sum() / np.abs(b).sum()), float(d.max() / np.abs(b).max())


def fp32_rounded(d: abi.Disk) -> abi.Disk:
    """The disk with its initial fields rounded to FP32 (the start of the sensitivity floor)."""
    d = d.copy()
    for f in ("r", "p", "u", "v"):
        d.rings[f] = d.rings[f].astype(np.float32).astype(np.float64)
    return d


def nudged(d: abi.Disk, seed: int, rel: float = 2.0**-25) -> abi.Disk:
    """The disk with every initial value multiplied by 1 + rel u, u uniform in [-1, 1]."""
    d = d.copy()
    rng = np.random.default_rng(seed)
    for f in ("r", "p", "u", "v"):
        d.rings[f] = d.rings[f] * (1.0 + rel * rng.uniform(-1.0, 1.0, size=d.rings[f].shape))
    return d


def run_ctx(w: configs.Workload, disk: abi.Disk, prec: int, steps: int, warmup: int = 0,
            graph: bool = True, times: bool = False) -> dict:
    """Load disk, run warmup then steps (the latter timed); the final state and rates."""
    with abi.Context(w.grid, w.params, prec) as ctx:
        ctx.load(disk)
        st = ctx.run_begin()
        m0, l0 = ctx.totals()
        ctx.run(warmup, st, graph=graph)
        t0 = time.perf_counter()
        ctx.run(steps, st, graph=graph)   # returns after the stream has drained
        wall = time.perf_counter() - t0
        m1, l1 = ctx.totals()
        out = {"wall": wall, "value": w.cells * steps / wall if steps else 0.0, "state": ctx.store(), "st": st,
               "variants": (*ctx.variants(), ctx.variant_z()), "dmass": (m1 - m0) / m0, "dangmom": (l1 - l0) / l0}
        if times:
            out["km"] = ctx.time_kernels(5)   # after the snapshot: timing advances the state
    return out


def counters(key: str) -> dict:
    if PROFILE_COUNTERS.exists():
        return json.loads(PROFILE_COUNTERS.read_text()).get(key, {})
    return {}


def roofline_entry(kernel: str, variant: int, kernel_ms: float, bytes_per_launch: float, prof: dict,
                   km: dict, copy_gbs: float) -> dict:
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    traffic = prof.get("traffic")
    r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel, "variant": variant,
         "kernel_ms": kernel_ms, "bytes_per_launch": bytes_per_launch,
         "traffic_over_algorithmic": traffic / bytes_per_launch if traffic else None, "kernels_ms": km,
         "copy_measured_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs}
    if "sq_insts_valu" in prof:
        # what bounds the sweeps: VALU issue, against one instruction per CU per clock
        ips = prof["sq_insts_valu"] / (kernel_ms * 1e-3)
        r["valu"] = {"insts_per_launch": prof["sq_insts_valu"], "achieved_insts_per_s": ips,
                     "peak_insts_per_s": VALU_PEAK_INSTS, "frac": ips / VALU_PEAK_INSTS}
    return r


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ws: list[configs.Workload], budget_s: float = 10.0) -> dict:
    """The FP64 oracle on the host cores, on a bounded prefix of the same workload: every disk
    of ws for n steps on all the OpenMP threads (n sized to about budget_s), then every disk for
    one step on one thread.  Steps after the first are timed on all threads."""
    import oracle

    L = oracle.lib()
    threads = L.orc_get_threads()
    cells = sum(w.cells for w in ws)
    disks = [oracle.OracleDisk(w.grid, w.params, w.seed, w.noise) for w in ws]
    states = [d.run_begin() for d in disks]
    t0 = time.perf_counter()
    for d, st in zip(disks, states):
        d.run(1, st)
    one = time.perf_counter() - t0
    n = max(1, min(20, int(budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    for d, st in zip(disks, states):
        d.run(n, st)
    wall = time.perf_counter() - t0
    L.orc_set_threads(1)
    try:
        t0 = time.perf_counter()
        for d, st in zip(disks, states):
            d.run(1, st)
        wall1 = time.perf_counter() - t0
    finally:
        L.orc_set_threads(threads)
    name = ws[0].name if len(ws) == 1 else f"c4_batch_{len(ws)}_disks"
    return {"value": cells * n / wall, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": f"{name}: steps 2..{n + 1} of every disk on the FP64 oracle (OpenMP over pencils, "
                      f"{threads} threads), {n} steps in {wall:.2f} s; then one step of every disk on 1 thread "
                      f"in {wall1:.2f} s",
            "one_core": {"value": cells / wall1, "unit": "zone-updates/s", "cores": 1},
            "cpu_model": cpu_model()}


def batch_leg(ws: list[configs.Workload], disks: list[abi.Disk], steps: int, warmup: int, graph: bool) -> dict:
    """The batch timed for steps after warmup, then its kernels timed one by one."""
    with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
        for ctx, d in zip(b.disks, disks):
            ctx.load(d)
        st = b.run_begin()
        b.run(warmup, st, graph=graph)
        t0 = time.perf_counter()
        b.run(steps, st, graph=graph)
        wall = time.perf_counter() - t0
        km = b.time_kernels(10)
        variants = b.variants()
    return {"wall": wall, "km": km, "variants": variants}


def batch_accuracy(ws: list[configs.Workload], disks: list[abi.Disk], steps: int, graph: bool,
                   report: str | None) -> dict:
    """A fresh batch and every disk's FP64 run (its floor, three nudged starts) for `steps` steps."""
    with abi.Batch([w.grid for w in ws], [w.params for w in ws]) as b:
        for ctx, d in zip(b.disks, disks):
            ctx.load(d)
        st = b.run_begin()
        b.run(steps, st, graph=graph)
        out32 = [ctx.store() for ctx in b.disks]
    rows = []
    for w, d, o in zip(ws, disks, out32):
        ref = run_ctx(w, d, abi.PN_FP64, steps, graph=graph)["state"]
        flo = run_ctx(w, fp32_rounded(d), abi.PN_FP64, steps, graph=graph)["state"]
        nud = max(rel_err(lab(run_ctx(w, nudged(d, s), abi.PN_FP64, steps, graph=graph)["state"], "r"),
                          lab(ref, "r"))[0] for s in (1, 2, 3))
        l1, li = rel_err(lab(o, "r"), lab(ref, "r"))
        rows.append({"disk": w.name, "q": w.params.M_p, "alpha": w.params.ss_alpha, "l1_rho": l1, "linf_rho": li,
                     "l1_vphi": rel_err(lab(o, "v"), lab(ref, "v"))[0],
                     "floor_l1_rho": rel_err(lab(flo, "r"), lab(ref, "r"))[0], "nudge_l1_rho": nud})
    l1s = np.array([r["l1_rho"] for r in rows])
    floors = np.array([r["floor_l1_rho"] for r in rows])
    nudges = np.array([r["nudge_l1_rho"] for r in rows])
    ok = l1s <= np.maximum(1e-5, floors)
    worst = max(rows, key=lambda r: r["l1_rho"] / max(1e-5, r["floor_l1_rho"]))
    if report:
        Path(report).parent.mkdir(parents=True, exist_ok=True)
        Path(report).write_text(json.dumps(rows, indent=1))
    return {"after_steps": steps, "median_l1_rho": float(np.median(l1s)),
            "disks_l1_rho_le_1e-5": int((l1s <= 1e-5).sum()),
            "median_floor_l1_rho": float(np.median(floors)),
            "disks_floor_le_1e-5": int((floors <= 1e-5).sum()),
            "disks_l1_rho_le_max_1e-5_floor": int(ok.sum()),
            "disks_l1_rho_le_max_1e-5_floor_nudge": int((l1s <= np.maximum(1e-5, np.maximum(floors, nudges))).sum()),
            "median_l1_over_floor": float(np.median(l1s / np.maximum(floors, 1e-300))),
            "worst_over_bar": worst,
            "per_disk_l1_rho": [float(f"{x:.3g}") for x in l1s],
            "per_disk_floor_l1_rho": [float(f"{x:.3g}") for x in floors],
            "per_disk_nudge_l1_rho": [float(f"{x:.3g}") for x in nudges]}


def single_disk_accuracy(w: configs.Workload, steps: int, graph: bool, floor: bool) -> dict:
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    a = run_ctx(w, d, abi.PN_FP32, steps, graph=graph)
    b = run_ctx(w, d, abi.PN_FP64, steps, graph=graph)
    errs = {f: rel_err(lab(a["state"], f), lab(b["state"], f)) for f in ("r", "u", "v", "p")}
    out = {"after_steps": steps, "l1_rho": errs["r"][0], "linf_rho": errs["r"][1],
           "l1_vr": errs["u"][0], "linf_vr": errs["u"][1], "l1_vphi": errs["v"][0], "linf_vphi": errs["v"][1],
           **({"l1_p": errs["p"][0], "linf_p": errs["p"][1]} if w.params.eos != 0 else {}),
           "rot_j_equal": bool(np.array_equal(a["state"]["rot_j"], b["state"]["rot_j"])),
           "dt_rel_diff": abs(a["st"].dt - b["st"].dt) / b["st"].dt,
           "mass_drift_fp32": a["dmass"], "mass_drift_fp64": b["dmass"],
           "angmom_drift_fp32": a["dangmom"], "angmom_drift_fp64": b["dangmom"]}
    if floor:
        fl = run_ctx(w, fp32_rounded(d), abi.PN_FP64, steps, graph=graph)["state"]
        out["floor"] = {f"l1_{k}": rel_err(lab(fl, f), lab(b["state"], f))[0]
                        for k, f in (("rho", "r"), ("vr", "u"), ("vphi", "v"), ("p", "p"))}
    return out


def single_disk_rates(w: configs.Workload, steps: int, warmup: int, graph: bool) -> dict:
    d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    a = run_ctx(w, d, abi.PN_FP32, steps, warmup, graph=graph, times=True)
    b = run_ctx(w, d, abi.PN_FP64, steps, warmup, graph=graph, times=True)
    return {"workload": w.name, "cells": w.cells, "steps": steps, "warmup": warmup,
            "fp32": {"value": a["value"], "ms_per_step": a["wall"] / steps * 1e3, "kernels_ms": a["km"],
                     "variants": a["variants"]},
            "fp64_plain": {"value": b["value"], "ms_per_step": b["wall"] / steps * 1e3, "kernels_ms": b["km"]},
            "speedup": a["value"] / b["value"], "_km32": a["km"], "_var32": a["variants"]}


def copy_probe() -> int:
    """The library's plain copy kernel on 256 MiB, 5 times after one warm-up: run under the
    HBM byte counters to scale them (the sweeps' traffic is quoted in the copy's units)."""
    gbs = abi.copy_bandwidth(256 << 20, 5)
    print(json.dumps({"copy_probe": {"bytes": 256 << 20, "launches": 6, "gbs": gbs}}), flush=True)
    return 0


def main_batch(a) -> dict:
    ws = configs.c4_members()
    cells = sum(w.cells for w in ws)
    disks = [abi.init_disk(w.grid, w.params, w.seed, w.noise) for w in ws]
    graph = not a.eager
    bl = batch_leg(ws, disks, a.steps, a.warmup, graph)
    value = cells * a.steps / bl["wall"]
    km = bl["km"]
    progress(f"FP32 batch of {len(ws)} disks: {bl['wall'] / a.steps * 1e3:.4f} ms per step")
    if a.batch_only:
        return {"value": value, "ms_per_step": bl["wall"] / a.steps * 1e3, "km": km}
    # the plain FP64 path on the same disks, one after the other, same steps
    wall64 = 0.0
    for w, d in zip(ws, disks):
        wall64 += run_ctx(w, d, abi.PN_FP64, a.steps, a.warmup, graph=graph)["wall"]
    v64 = cells * a.steps / wall64
    progress(f"FP64 disks one by one: {wall64 / a.steps * 1e3:.4f} ms per step")
    dom = max(("sweep_x", "sweep_y"), key=lambda k: km[k])
    var = bl["variants"][0] if dom == "sweep_x" else bl["variants"][1]
    prof = counters(f"c4_batch_48_disks:fp32:{dom}:v{var}")
    roof = roofline_entry(dom, var, km[dom], 2 * 3 * 4 * cells, prof, km, abi.copy_bandwidth())
    out = {
        "metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": bl["wall"] / a.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded disk initial conditions of the workload)",
        "config": {"workload": f"c4_batch_{len(ws)}_disks", "disks": len(ws), "cells": cells,
                   "grids": "N_r in {128,192,256,384,512}, N_phi = 3 N_r", "path": "FP32 bulk, batched",
                   "graph": graph},
        "roofline": roof,
        "fp64_plain": {"value": v64, "ms_per_step": wall64 / a.steps * 1e3, "speedup": value / v64,
                       "how": "every disk on the plain FP64 path, one after the other"},
    }
    if not a.no_ref:
        out["accuracy_vs_fp64"] = batch_accuracy(ws, disks, ACCURACY_STEPS, graph, a.disk_report)
        progress(f"batch accuracy after {ACCURACY_STEPS} steps done")
        c2 = configs.c2()
        sd = single_disk_rates(c2, a.steps, a.warmup, graph)
        progress("1024 x 3072 disk rates done")
        sd.pop("_km32")
        sd.pop("_var32")
        sd["accuracy_vs_fp64"] = single_disk_accuracy(c2, ACCURACY_STEPS, graph, a.floor)
        out["single_disk"] = sd
        progress("1024 x 3072 disk accuracy done")
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ws, a.cpu_budget)
        progress("CPU baseline done")
    return out


def main_single(a) -> dict:
    w = getattr(configs, a.config)()
    graph = not a.eager
    sd = single_disk_rates(w, a.steps, a.warmup, graph)
    km, var32 = sd.pop("_km32"), sd.pop("_var32")
    dom = max(("sweep_x", "sweep_y", "sweep_z"), key=lambda k: km[k])
    var = {"sweep_x": var32[0], "sweep_y": var32[1], "sweep_z": var32[2]}[dom]
    fields = 3 + (1 if w.params.ndim == 3 else 0) + (1 if w.params.eos != 0 else 0)
    prof = counters(f"{w.name}:fp32:{dom}:v{var}")
    roof = roofline_entry(dom, var, km[dom], 2 * fields * 4 * w.cells, prof, km, abi.copy_bandwidth())
    out = {
        "metric": METRIC, "value": sd["fp32"]["value"], "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": sd["fp32"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded disk initial condition of the workload)",
        "config": {"workload": w.name, "cells": w.cells, "path": "FP32 bulk (single disk)", "graph": graph},
        "roofline": roof,
        "fp64_plain": {**sd["fp64_plain"], "speedup": sd["speedup"]},
    }
    if not a.no_ref:
        out["accuracy_vs_fp64"] = single_disk_accuracy(w, a.accuracy_steps or ACCURACY_STEPS, graph, a.floor)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline([w], a.cpu_budget)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c4", choices=["c4", "c2", "c3", "c5e3"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-ref", action="store_true", help="skip the FP64 accuracy legs")
    ap.add_argument("--batch-only", action="store_true",
                    help="c4: time only the FP32 batch (no FP64 runs, no CPU leg): for tracing the batch kernels")
    ap.add_argument("--floor", action="store_true",
                    help="single disks: also run the FP64 path from the FP32-rounded initial state")
    ap.add_argument("--accuracy-steps", type=int, default=0,
                    help="single disks: steps of the accuracy leg (default 1000)")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of the all-cores CPU leg")
    ap.add_argument("--disk-report", default=None, help="c4: write every disk's errors to this JSON file")
    ap.add_argument("--copy-probe", action="store_true", help="only run a plain device copy of known size")
    ap.add_argument("--eager", action="store_true",
                    help="launch every step's kernels directly instead of replaying a captured HIP graph "
                         "(the same kernels; for tools that do not follow graph launches)")
    a = ap.parse_args(argv)
    if a.copy_probe:
        return copy_probe()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported (a disk is not "
              f"sharded across GPUs in this build)", file=sys.stderr)
        return 2
    if a.steps < 1 or a.warmup < 0:
        print("--steps must be at least 1 and --warmup at least 0", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))
    out = main_batch(a) if a.config == "c4" else main_single(a)
    if rank == 0:
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
