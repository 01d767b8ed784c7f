"""Build libhps_mi355x.so for gfx950 with hipcc (in-tree, next to the package).

Usage: python -m jaxhps_amd.csrc.build   (or __graft_entry__.build()).
Each .hip file is compiled to an object in _build/ in parallel, then linked.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
REPO = PKG.parent
LIB = PKG / "libhps_mi355x.so"
BUILD = REPO / "_build"
SOURCES = ["gemm.hip", "hps_capi.hip", "refine.hip", "mlp.hip", "interp.hip", "scatter.hip", "adaptive.hip", "hps_prof.hip"]
# the LU kernels take most of the compile time: one object per element type,
# compiled in parallel (lu.hip's HPS_LU_TYPES selects the instantiations)
VARIANTS = [("lu.hip", f"lu_{t}", [f"-DHPS_LU_TYPES={bit}"]) for t, bit in (("f64", 1), ("f32", 2), ("c128", 4), ("c64", 8))]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # packed FP32 instructions are disabled: some of their forms have given
    # wrong results on this hardware pool
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{REPO / 'include'}",
]


def _compile(src: str, incremental: bool = False, stem: str = None, defines=()) -> Path:
    obj = BUILD / ((stem or Path(src).stem) + ".o")
    if incremental and obj.exists():
        deps = [HERE / src, *HERE.glob("*.h"), *(REPO / "include").glob("*.h")]
        if obj.stat().st_mtime > max(d.stat().st_mtime for d in deps):
            return obj
    cmd = [HIPCC, *FLAGS, *defines, "-c", str(HERE / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True, incremental: bool = False) -> Path:
    """Compile every source (incremental=True keeps objects newer than their
    source and every header) and link the library."""
    BUILD.mkdir(exist_ok=True)
    jobs = [(f, None, ()) for f in SOURCES] + VARIANTS
    with cf.ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 8)) as ex:
        objs = list(ex.map(lambda j: _compile(j[0], incremental, j[1], j[2]), jobs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
