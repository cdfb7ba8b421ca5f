"""Build the native libraries in-tree.

* ``gat_amd/lib/libgat_hip.so``  — HIP kernels + the C ABI of include/gat_hip.h,
  compiled by hipcc for gfx950 only.
* ``gat_amd/lib/libgat_host.so`` — the C++ host mirror of the reference's index
  and query classes (Grid, FSG, STIG, SystemTest), the synthetic-workload
  generator and the C ABI of include/gat_host.h.  Plain g++: it never includes a
  HIP header and reaches the device only through libgat_hip.so.

Floating point: every compile uses -ffp-contract=off and no fast-math flag, so
each float operation is individually rounded as in the reference's host code.
HIP code additionally disables packed-fp32 instruction selection.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "gat_amd"
LIB_DIR = PKG / "lib"
CSRC = PKG / "csrc"
HOST = PKG / "host"
INCLUDE = ROOT / "include"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")

HIP_SOURCES = ["runtime.hip", "util.hip", "range.hip", "range_batched.hip", "edr.hip", "edr_batched.hip", "build.hip"]
HOST_SOURCES = ["MBB.cpp", "Synth.cpp", "DeviceArena.cpp", "RangeEngine.cpp", "Grid.cpp", "GridSimilarity.cpp", "MortonGrid.cpp", "FSG.cpp",
                "STIG.cpp", "SystemTest.cpp", "Server.cpp", "Autotune.cpp", "MixedRange.cpp", "TrajFile.cpp",
                "host_capi.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-fvisibility=hidden",
]
HOST_FLAGS = ["-O2", "-m64", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-pthread",
              "-fvisibility=hidden"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(verbose: bool = True) -> Path:
    """Compile every HIP translation unit for gfx950 and link libgat_hip.so."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    obj_dir = ROOT / "_build" / "hip"
    obj_dir.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    out = LIB_DIR / "libgat_hip.so"
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", *objs, "-o", str(tmp),
          f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, out)
    return out


def build_host(verbose: bool = True) -> Path:
    """Compile the host mirror with g++ and link it against libgat_hip.so."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    obj_dir = ROOT / "_build" / "host"
    obj_dir.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        _run(["g++", *HOST_FLAGS, f"-I{INCLUDE}", f"-I{HOST}", "-c", str(HOST / src), "-o", str(obj)])
        objs.append(str(obj))
    out = LIB_DIR / "libgat_host.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-shared", "-pthread", "-Wl,--no-undefined", *objs, "-o", str(tmp), f"-L{LIB_DIR}", "-lgat_hip",
          "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_hip()
    build_host()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"hip": build_hip, "host": build_host, "all": build_all}[w]()
