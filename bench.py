"""Throughput of the hydro step on one MI355X, as one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2] [--precision fp64|fp32]

A step is main.cu:345-387 -- CFL dt, FARGO split, planet half kick, r sweep, phi sweep,
integer phi roll, planet update -- advanced on the device with dt and the planet kept
there (a captured HIP graph per step).  value = zone-updates/s = cells * K / wall time of
the K timed steps, with the state resident in HBM.  The roofline entry prices the dominant
sweep kernel, timed with HIP events on the context's stream, against HBM bandwidth.  The
CPU baseline times the FP64 oracle on the host cores on a bounded prefix of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from penguin_2_0_amd import abi, configs  # noqa: E402

METRIC = "zone-updates/s per GPU: FP32-batched vs FP64 plain at L1(ρ)≤1e-5; % HBM peak"
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = ROOT / "profiles" / "traffic.json"


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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported (the disk is not "
              f"sharded across GPUs in this build)", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    w = getattr(configs, a.config)()
    prec = abi.PN_FP64 if a.precision == "fp64" else abi.PN_FP32
    disk = abi.init_disk(w.grid, w.params, w.seed, w.noise)
    ctx = abi.Context(w.grid, w.params, prec)
    ctx.load(disk)
    st = ctx.run_begin()
    ctx.run(a.warmup, st, graph=True)
    t0 = time.perf_counter()
    ctx.run(a.steps, st, graph=True)   # returns after the stream has drained
    t1 = time.perf_counter()
    wall = t1 - t0
    value = w.cells * a.steps / wall

    # dominant kernel, timed on the context's stream
    km = ctx.time_kernels(5)
    dom = max(("sweep_x", "sweep_y"), key=lambda k: km[k])
    wbytes = 8 if prec == abi.PN_FP64 else 4
    fields = 3   # Sigma, v_r, v_phi: the minimal 2-D isothermal state
    bytes_per_launch = 2 * fields * wbytes * w.cells
    achieved = bytes_per_launch / (km[dom] * 1e-3) / 1e9
    traffic = None
    if PROFILE_TRAFFIC.exists():
        try:
            traffic = json.loads(PROFILE_TRAFFIC.read_text()).get(f"{w.name}:{a.precision}:{dom}")
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": dom, "kernel_ms": km[dom], "kernels_ms": km}
    out = {
        "metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == abi.PN_FP64 else "f32",
        "data": "synthetic (seeded disk initial condition of the workload)",
        "config": {"workload": w.name, "cells": w.cells, "path": "plain FP64" if prec == abi.PN_FP64 else "FP32 bulk",
                   "graph": True},
        "roofline": roofline,
    }
    ctx.close()
    if rank == 0 and not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    if rank == 0:
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
