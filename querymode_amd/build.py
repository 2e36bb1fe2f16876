"""Build recipe for the native libraries (no cmake, no JIT cache).

* ``libqmgpu.so``     the gfx950 HIP kernels + the C-ABI (hipcc, in-tree)
* ``libqmdatagen.so`` the deterministic synthetic-column generator (gcc)
* ``qm_napi.node``    the Node-API addon over the C-ABI (gcc), when the Node
  headers are installed

Everything lands in ``querymode_amd/lib`` so the built files travel with the
source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = LIB / "_obj"
INCLUDE = PKG.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QM_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["k_scan.hip", "k_part.hip", "k_result.hip", "k_topk.hip", "k_join.hip", "k_stats.hip", "k_sort.hip",
               "k_pruns.hip", "k_post.hip", "k_radix.hip",
               "k_batch.hip", "k_runs.hip", "qm_plan.cpp", "qm_api.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 ops are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def _up_to_date(src: Path, out: Path) -> bool:
    """QM_BUILD_INCREMENTAL=1: keep an object newer than its source and every header."""
    if os.environ.get("QM_BUILD_INCREMENTAL") != "1" or not out.exists():
        return False
    deps = [src, *CSRC.glob("*.h"), *INCLUDE.glob("*.h")]
    return all(d.stat().st_mtime < out.stat().st_mtime for d in deps)


def _compile_one(src: str) -> Path:
    out = OBJ / (src.rsplit(".", 1)[0] + ".o")
    if not _up_to_date(CSRC / src, out):
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(CSRC / src), "-o", str(out)])
    return out


def build_gpu_lib(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_one, HIP_SOURCES))
    target = LIB / "libqmgpu.so"
    tmp = LIB / "libqmgpu.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, target)
    return target


def build_datagen() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    target = LIB / "libqmdatagen.so"
    tmp = LIB / "libqmdatagen.so.tmp"
    _run(["gcc", "-O3", "-fPIC", "-shared", "-fopenmp", "-std=c11", str(CSRC / "datagen.c"), "-o", str(tmp), "-lm"])
    os.replace(tmp, target)
    return target


NODE_INCLUDE = Path(os.environ.get("QM_NODE_INCLUDE", "/usr/include/node"))


def build_napi() -> Path | None:
    """Node-API addon (C, node_api.h) linked against libqmgpu.so."""
    if not (NODE_INCLUDE / "node_api.h").exists():
        return None
    target = LIB / "qm_napi.node"
    tmp = LIB / "qm_napi.node.tmp"
    _run(["gcc", "-O2", "-Wall", "-shared", "-fPIC", "-DNODE_GYP_MODULE_NAME=qm_napi", f"-I{NODE_INCLUDE}",
          str(PKG / "napi" / "qm_napi.c"), "-o", str(tmp), f"-L{LIB}", "-lqmgpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, target)
    return target


def build_all() -> None:
    build_datagen()
    build_gpu_lib()
    build_napi()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
