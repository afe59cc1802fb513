"""Build ``libwbmhd.so`` for gfx950 in-tree with hipcc.

``python -m wellbalanced_amd.build`` compiles every ``csrc/*.hip`` translation
unit and links them into ``wellbalanced_amd/libwbmhd.so``.  The FP64 parity units
are compiled with ``-ffp-contract=off`` so that no multiply-add is fused; the
mixed-precision units (``*_mixed.hip``) fuse the multiply-adds within each source
expression (``-ffp-contract=on``): the same expression gives the same bits wherever it
is inlined, so a face flux recomputed by the refluxing equals the stage's.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
LIBRARY = HERE / "libwbmhd.so"
ARCH = "gfx950"


def _hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the ROCm toolchain is needed to build libwbmhd.so")
    return path


def _flags(source: Path) -> list[str]:
    flags = [
        f"--offload-arch={ARCH}",
        "-O3",
        "-fPIC",
        "-std=c++17",
        f"-I{INCLUDE}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics",
    ]
    if source.stem.endswith("_mixed"):
        flags.append("-ffp-contract=on")
    else:
        flags.append("-ffp-contract=off")
    return flags


def build(verbose: bool = False) -> Path:
    """Compile all HIP sources and link ``libwbmhd.so``; returns its path."""
    hipcc = _hipcc()
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)

    def compile_one(src: Path) -> Path:
        obj = objdir / (src.stem + ".o")
        cmd = [hipcc, *_flags(src), "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stderr}")
        return obj

    workers = min(len(sources), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        objects = list(pool.map(compile_one, sources))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
