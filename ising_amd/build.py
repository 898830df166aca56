"""Build recipes for the native parts of ising_amd.

* ``libising_hip.so`` -- HIP kernels + C ABI, hipcc for gfx950, built in-tree
  under ``ising_amd/lib`` so that it travels with the source tree.
* the Fortran drop-in layer (``libising_fortran.so`` + ``pt_gpu``), built with
  the ROCm LLVM flang when it is present.

Each builder compiles every time it is called (it is the build check), and
returns the path of what it built.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB_DIR = PKG / "lib"
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = "gfx950"

# Packed FP32 instructions are kept off (wrong results have been observed on
# this hardware with some of their forms).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    # a*b + c is fused only inside one source expression, so an inlined
    # device function rounds the same way in every kernel that calls it (the
    # fused sweep and the half sweeps give the same chains bit for bit)
    "-ffp-contract=on",
    # matrix-core results land in ordinary VGPRs (no accumulator-register copies)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")


def _atomic_replace(tmp: Path, dst: Path) -> None:
    os.replace(tmp, dst)


def build_hip(verbose: bool = False) -> Path:
    """Compile ``csrc/ising_hip.hip`` into ``lib/libising_hip.so``."""
    LIB_DIR.mkdir(exist_ok=True)
    out = LIB_DIR / "libising_hip.so"
    tmp = LIB_DIR / "libising_hip.so.tmp"
    cmd = [HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", str(CSRC / "ising_hip.hip"), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    _atomic_replace(tmp, out)
    return out


def flang_available() -> bool:
    return Path(FLANG).exists() or shutil.which(FLANG) is not None


def build_fortran(verbose: bool = False) -> Path | None:
    """Compile the Fortran drop-in modules into ``lib/libising_fortran.so`` and
    link the PT driver ``lib/pt_gpu`` and the check program ``lib/test_dropin``
    against it (needs flang)."""
    if not flang_available() or not FSRC.exists():
        return None
    LIB_DIR.mkdir(exist_ok=True)
    mod_dir = LIB_DIR / "fmod"
    mod_dir.mkdir(exist_ok=True)
    srcs = [FSRC / n for n in ("ising_gpu_c.f90", "ising_chainmesh.f90", "energymin_gpu.f90", "pt_utils_gpu.f90")]
    lib = LIB_DIR / "libising_fortran.so"
    tmp = LIB_DIR / "libising_fortran.so.tmp"
    flags = ["-O2", "-fopenmp", f"-J{mod_dir}"]
    cmd = [FLANG, *flags, "-fPIC", "-shared", *map(str, srcs), f"-L{LIB_DIR}", "-lising_hip",
           "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    _atomic_replace(tmp, lib)
    for prog in ("pt_gpu", "test_dropin"):
        exe = LIB_DIR / prog
        tmp = LIB_DIR / f"{prog}.tmp"
        cmd = [FLANG, *flags, f"-I{mod_dir}", str(FSRC / f"{prog}.f90"), f"-L{LIB_DIR}",
               "-lising_fortran", "-lising_hip", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        _atomic_replace(tmp, exe)
    return lib


def build_all(verbose: bool = False) -> dict:
    out = {"hip": build_hip(verbose)}
    out["fortran"] = build_fortran(verbose)
    return out


if __name__ == "__main__":
    for k, v in build_all(verbose=True).items():
        print(k, v)
