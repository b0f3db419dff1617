"""Build the gfx950 hot-path library (liblsgpu.so) and the host tools.

Compiles every HIP source under ``leachingsim_amd/csrc`` for gfx950 with
hipcc and links them into ``leachingsim_amd/liblsgpu.so`` (in-tree, so the
library travels with the source tree).  The C++ host mirror of the
reference's ``Simulation`` API and the ``lsim`` driver are built against it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "liblsgpu.so"
LSIM = PKG / "lsim"
HOSTLIB = PKG / "liblshost.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: no fused multiply-add, so every expression rounds as the
# reference's scalar C++ does (bit-exact masks, reproducible fields).
# -packed-fp32-ops: no packed FP32 instructions (v_pk_*_f32 are kept out of
# every kernel).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", str(ROOT / "include"),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l
                    and l.strip() != "1 warning generated." and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")
    if out.strip():
        print(out, file=sys.stderr)


def build_library(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objs = []

    def compile_one(src: Path) -> Path:
        obj = BUILD / (src.stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp),
          *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    odir = ROOT / "oracle"
    _run(["make", "-s", "-C", str(odir), "-B", "liblsoracle.so"])
    return odir / "liblsoracle.so"


def build_host() -> Path | None:
    """C++ host side: liblshost.so (file formats, CPU-testable) and the lsim
    driver (Simulation over liblsgpu.so, the reference's command line)."""
    if not HOST.exists():
        return None
    cxx = ["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-I", str(ROOT / "include")]
    io = [str(HOST / "io.cpp"), str(HOST / "steplog.cpp")]
    tmp = HOSTLIB.with_suffix(".so.tmp")
    _run([*cxx, "-shared", *io, str(HOST / "capi.cpp"), "-o", str(tmp)])
    os.replace(tmp, HOSTLIB)
    _run([*cxx, *io, str(HOST / "simulation.cpp"), str(HOST / "lsim.cpp"), "-o", str(LSIM),
          "-L", str(PKG), "-llsgpu", "-Wl,-rpath,$ORIGIN"])
    return LSIM


def build_all(jobs: int = 8) -> None:
    build_library(jobs)
    build_oracle()
    build_host()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}")
