"""Build the HIP library (gfx950) and the CPU oracle in-tree.

    python -m penguin_2_0_amd.build            # both
    python -m penguin_2_0_amd.build --lib      # only libpenguin.so

Outputs:
    penguin_2_0_amd/libpenguin.so   the product: C ABI of include/penguin.h
    penguin_2_0_amd/penguin_main    the reference's run loop as a C++ caller of that ABI
    oracle/libpn_oracle.so          the FP64 CPU restatement used by the tests as checker
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "penguin_2_0_amd"
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libpenguin.so"
DRIVER_SRC = PKG / "host" / "penguin_main.cpp"
DRIVER = PKG / "penguin_main"
ORACLE_SRC = ROOT / "oracle" / "pn_oracle.c"
ORACLE_LIB = ROOT / "oracle" / "libpn_oracle.so"

ARCH = "gfx950"
# packed FP32 (v_pk_*_f32) is switched off for all device code; FP64 kernels round every
# operation separately (no FMA contraction) so they reproduce the reference's arithmetic
HIP_COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
# the FP32 sweeps use native reciprocal / sqrt / exp / log (no reassociation: the order of
# every sum is the source order)
FAST_F32 = ["-ffp-contract=fast", "-fno-math-errno", "-fno-honor-infinities", "-fno-honor-nans",
            "-fno-signed-zeros", "-freciprocal-math", "-fapprox-func", "-fgpu-flush-denormals-to-zero"]
UNITS = {
    "pn_kernels_fp64.hip": ["-ffp-contract=off"],
    "pn_kernels_fp32.hip": FAST_F32,
    "pn_kernels_common.hip": ["-ffp-contract=off"],
    "pn_runtime.cpp": ["-ffp-contract=off"],
    "pn_batch.cpp": ["-ffp-contract=off"],
    "pn_host.cpp": ["-ffp-contract=off"],
}


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build_lib(jobs: int = 8) -> Path:
    hipcc = _hipcc()
    OBJ.mkdir(exist_ok=True)
    cmds = []
    objs = []
    for unit, extra in UNITS.items():
        src = CSRC / unit
        obj = OBJ / (src.stem + ".o")
        lang = ["-x", "hip"] if src.suffix == ".hip" else []
        cmds.append([hipcc, *HIP_COMMON, *extra, *lang, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> Path:
    """The C++ run loop over the C ABI, linked against the in-tree libpenguin.so."""
    cxx = shutil.which("g++") or "c++"
    tmp = DRIVER.with_suffix(".tmp")
    _run([cxx, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", "-o", str(tmp), str(DRIVER_SRC),
          f"-L{PKG}", "-lpenguin", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, DRIVER)
    return DRIVER


def build_oracle() -> Path:
    cc = shutil.which("gcc") or "cc"
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([cc, "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
          f"-I{ROOT / 'include'}", "-o", str(tmp), str(ORACLE_SRC), "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="build only the HIP library")
    ap.add_argument("--oracle", action="store_true", help="build only the CPU oracle")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    if not a.oracle:
        print("built", build_lib(a.j))
        print("built", build_driver())
    if not a.lib:
        print("built", build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main())
