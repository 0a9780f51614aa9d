"""In-tree build of the native libraries.

libfe_host.so : host C++ (meshes, numbering, coloring, coefficient fields)
libfe_hip.so  : gfx950 HIP kernels + C ABI (include/fe_hip.h)
libfeng.so    : C++ mirror of the reference feLinearSystem / feBilinearForm
                API on top of the two C ABIs (feng_amd/cpp)

Run `python -m feng_amd.build` (or __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "feng_amd"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
INC = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["fe_ctx.hip", "fe_csr.hip", "fe_assemble.hip", "fe_krylov.hip", "fe_precond.hip", "fe_gmres.hip", "fe_learned.hip", "fe_batch.hip", "fe_amg.hip", "fe_norm.hip", "fe_constraints.hip", "fe_gather.hip", "fe_amg_pcg.hip", "fe_fieldsplit.hip", "fe_direct.hip"]

# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{INC}", f"-I{CSRC}",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_host() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libfe_host.so"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wextra",
          f"-I{INC}", "-o", out, CSRC / "host" / "fe_host.cpp"])
    return out


def build_hip(jobs: int = 4) -> Path:
    LIB.mkdir(exist_ok=True)
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", CSRC / src, "-o", obj]
        print("+", " ".join(str(c) for c in cmd), flush=True)
        procs.append(subprocess.Popen([str(c) for c in cmd]))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError(f"hipcc failed on {src}")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    out = LIB / "libfe_hip.so"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs, "-lrocsolver", "-lrocblas"])
    return out


def build_cpp() -> Path | None:
    cpp = PKG / "cpp"
    if not (cpp / "feng_api.cpp").exists():
        return None
    out = LIB / "libfeng.so"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra",
          f"-I{INC}", f"-I{cpp}", "-o", out, cpp / "feng_api.cpp",
          f"-L{LIB}", "-lfe_hip", "-lfe_host", "-Wl,-rpath,$ORIGIN"])
    ex = cpp / "examples"
    if ex.exists():
        for src in sorted(ex.glob("*.cpp")):
            _run(["g++", "-std=c++17", "-O2", "-Wall", f"-I{INC}", f"-I{cpp}", "-o",
                  LIB / src.stem, src, f"-L{LIB}", "-lfeng", "-lfe_hip", "-lfe_host",
                  "-Wl,-rpath,$ORIGIN"])
    return out


def build_all() -> None:
    build_host()
    build_hip()
    build_cpp()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "hip": build_hip, "cpp": build_cpp, "all": build_all}[what]()
