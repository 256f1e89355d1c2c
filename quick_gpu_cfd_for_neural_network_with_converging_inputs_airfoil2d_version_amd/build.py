"""Build libamr_hip.so in-tree: host C++ (g++) + HIP kernels for gfx950 (hipcc), one shared library.

The mesher sources are compiled with g++ -O3 -ffp-contract=off so that the mesh arrays round
exactly as the reference CPU build does; the .hip sources are compiled by hipcc for gfx950
with packed FP32 instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
OUT = PACKAGE / "libamr_hip.so"
DRIVER = PACKAGE / "quick_airfoil_hip"
DRIVER_SOURCE = "host/quick_airfoil_hip.cpp"
CALLER_SOURCE = ROOT / "tests" / "capi" / "api_caller.cpp"
CALLER = ROOT / "tests" / "capi" / "api_caller"
# and of the legacy-grid backend API (euler2d_gpu.h)
GRID_CALLER_SOURCE = ROOT / "tests" / "capi" / "ugrid_caller.cpp"
GRID_CALLER = ROOT / "tests" / "capi" / "ugrid_caller"
OBJ = ROOT / "_build" / "obj"

HOST_SOURCES = ["host/cutcell.cpp", "host/mesher.cpp", "host/viscous_prep.cpp", "host/mesh_capi.cpp",
                "host/cfd_api.cpp", "host/steady.cpp", "host/batch_steady.cpp", "host/ugrid_axis.cpp",
                "host/euler2d_api.cpp", "host/ugrid_cut.cpp"]
HIP_SOURCES = ["hip/amr_kernels_f64.hip", "hip/amr_kernels_f32.hip", "hip/amr_device.hip", "hip/amr_session.hip",
               "hip/amr_batch.hip", "hip/ugrid.hip"]
# Per-source extra flags: FP32 quotients and square roots of the FP32 kernels are not correctly
# rounded (about 1 ulp: reciprocal and multiply) instead of the scaled IEEE sequences, and FP32
# denormals flush to zero, so that a quotient is v_rcp + v_mul without the frexp / ldexp range
# scaling (8% fewer VALU instructions in the flux kernels).  Flow quantities never come near
# 1e-38; the FP64 state, control and reductions keep full IEEE behaviour.
# The tensor-product grid solver is compiled without FMA contraction: its arithmetic follows the
# reference's formulas operation for operation, so results are bitwise its CPU restatement's.
HIP_EXTRA = {"hip/amr_kernels_f32.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"],
             "hip/ugrid.hip": ["-ffp-contract=off"]}
INCLUDES = ["-I", str(ROOT / "include"), "-I", str(CSRC / "host"), "-I", str(CSRC / "hip")]
ARCH = os.environ.get("HIPAMR_ARCH", "gfx950")


def _hipcc() -> str:
    for candidate in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    done = subprocess.run(cmd, capture_output=True, text=True)
    if done.returncode != 0:
        raise RuntimeError(f"command failed ({done.returncode}): {' '.join(cmd)}\n{done.stdout}\n{done.stderr}")
    return done


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source and link PACKAGE/libamr_hip.so (always rebuilds)."""
    OBJ.mkdir(parents=True, exist_ok=True)
    hipcc = _hipcc()
    commands, objects = [], []
    for rel in HOST_SOURCES:
        obj = OBJ / (Path(rel).stem + ".o")
        commands.append(["g++", "-std=c++17", "-O3", "-DNDEBUG", "-fPIC", "-ffp-contract=off", "-Wall",
                         *INCLUDES, "-c", str(CSRC / rel), "-o", str(obj)])
        objects.append(obj)
    for rel in HIP_SOURCES:
        obj = OBJ / (Path(rel).stem + ".hip.o")
        commands.append([hipcc, "-x", "hip", f"--offload-arch={ARCH}", "-std=c++17", "-O3", "-DNDEBUG", "-fPIC",
                         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", *INCLUDES,
                         *HIP_EXTRA.get(rel, []), "-c", str(CSRC / rel), "-o", str(obj)])
        objects.append(obj)
    with ThreadPoolExecutor(max_workers=jobs or min(8, len(commands))) as pool:
        list(pool.map(lambda c: _run(c, verbose), commands))
    tmp = OUT.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)], verbose)
    tmp.replace(OUT)
    # The solver executable (the reference's quick_airfoil command line) over the library.
    driver_obj = OBJ / "quick_airfoil_hip.o"
    _run(["g++", "-std=c++17", "-O3", "-DNDEBUG", "-ffp-contract=off", "-Wall", *INCLUDES, "-c",
          str(CSRC / DRIVER_SOURCE), "-o", str(driver_obj)], verbose)
    tmp = DRIVER.with_suffix(".tmp")
    _run(["g++", str(driver_obj), "-L", str(PACKAGE), "-l:libamr_hip.so", "-L", "/opt/rocm/lib",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", str(tmp)], verbose)
    tmp.replace(DRIVER)
    # The test suite's callers of the C++ backend APIs (reference-style callers of
    # static_amr2d_gpu.h and euler2d_gpu.h linked against the library), when the tests are present.
    for source, target in ((CALLER_SOURCE, CALLER), (GRID_CALLER_SOURCE, GRID_CALLER)):
        if not source.exists():
            continue
        tmp = target.with_suffix(".tmp")
        _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", str(ROOT / "include"), str(source), "-L", str(PACKAGE),
              "-l:libamr_hip.so", "-L", "/opt/rocm/lib", f"-Wl,-rpath,$ORIGIN/../../{PACKAGE.name}",
              "-Wl,-rpath,/opt/rocm/lib", "-o", str(tmp)], verbose)
        tmp.replace(target)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
