"""Build libswp_core.so (gfx950) in-tree with hipcc.

``python -m sweep_amd.build`` compiles every ``sweep_amd/csrc/*.hip`` into
``sweep_amd/_lib/libswp_core.so``.  The library is loaded from there by
``sweep_amd._core``; nothing is installed elsewhere.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libswp_core.so"
ARCH = os.environ.get("SWP_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libswp_core)")


def sources():
    return sorted(CSRC.glob("*.hip"))


# The parity build: the same sources with floating-point contraction off (no
# a * b + c fused into one rounding) and SWP_EXACT_DIV (stencil sums divided
# by h^2 or h where the oracle divides, instead of multiplied by a reciprocal),
# for checking that a record's distance from the oracle, which does neither,
# comes from these two rounding choices alone.  Tests load it in a subprocess
# through SWP_CORE_LIBRARY; the product path loads LIB.
LIB_NOCONTRACT = OUT_DIR / "libswp_core_nocontract.so"


def build(verbose: bool = False, jobs: int | None = None, contract: bool = True) -> Path:
    """Compile every source and link the shared library; returns its path
    (contract=False: the parity build, LIB_NOCONTRACT)."""
    OUT_DIR.mkdir(exist_ok=True)
    obj_dir = OUT_DIR / ("obj" if contract else "obj_nocontract")
    obj_dir.mkdir(exist_ok=True)
    cc = hipcc()
    srcs = sources()
    flags = HIPCC_FLAGS if contract else HIPCC_FLAGS + ["-ffp-contract=off", "-DSWP_EXACT_DIV"]
    lib = LIB if contract else LIB_NOCONTRACT

    def compile_one(src: Path) -> Path:
        obj = obj_dir / (src.stem + ".o")
        cmd = [cc, *flags, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stdout}\n{r.stderr}")
        return obj

    jobs = jobs or min(len(srcs), 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = lib.with_suffix(".so.tmp")
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv, contract="--no-contract" not in sys.argv)
    print(path)
