"""Build the in-tree HIP library ``roundabout_amd/lib/librdb_colphys.so``.

hipcc cross-compiles for gfx950 (no GPU needed).  Device code is compiled
with ``-ffp-contract=off`` (no FMA contraction: the reference's CPU policy,
cmake/compiler_flags.cmake:335-362, which the bitwise parity relies on) and
without the packed-FP32 target feature.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "librdb_colphys.so"

SOURCES = ["colphys_kernels.hip", "colphys_api.hip", "mlp_kernels.hip", "mlp_net.hip", "mlp_api.hip", "surrogate.hip",
           "momentum.hip", "boundary.hip", "remap.hip", "contrib.hip"]
ARCH = os.environ.get("RDB_COLPHYS_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise FileNotFoundError("hipcc not found")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off", "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result",
    ]


def build(verbose: bool = False) -> Path:
    LIB_DIR.mkdir(exist_ok=True)
    cc = hipcc()
    cmds, objs = [], []
    for src in SOURCES:
        obj = LIB_DIR / (Path(src).stem + ".o")
        cmds.append([cc, *flags(), "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    # the translation units compile independently: one hipcc per source
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        def run(cmd):
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
        for f in [pool.submit(run, c) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        Path(o).unlink(missing_ok=True)
    return LIB


FORTRAN_DIR = PKG / "fortran"
FORTRAN_CHECK = FORTRAN_DIR / "colphys_fortran_check"


def flang() -> str | None:
    for cand in (os.environ.get("FC"), "/opt/rocm/llvm/bin/flang", shutil.which("flang")):
        if cand and Path(cand).exists():
            return cand
    return None


def build_fortran(verbose: bool = False) -> Path | None:
    """Compile the ISO_C_BINDING host module and its check program (AMD flang)."""
    fc = flang()
    if fc is None:
        return None
    mod_dir = LIB_DIR / "fortran_mod"
    mod_dir.mkdir(parents=True, exist_ok=True)
    obj = mod_dir / "rdb_colphys.o"
    cmds = [
        [fc, "-O2", f"-J{mod_dir}", "-c", str(FORTRAN_DIR / "rdb_colphys.F90"), "-o", str(obj)],
        [fc, "-O2", f"-J{mod_dir}", str(FORTRAN_DIR / "colphys_fortran_check.F90"), str(obj),
         f"-L{LIB_DIR}", "-lrdb_colphys", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../lib",
         "-o", str(FORTRAN_CHECK)],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return FORTRAN_CHECK


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_fortran(verbose="-v" in sys.argv)
    print(LIB)
