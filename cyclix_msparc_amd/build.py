"""Build the gfx950 HIP library in-tree: cyclix_msparc_amd/_lib/libcyclix_msparc.so.

Each ``csrc/*.hip`` translation unit is compiled with hipcc for gfx950 in
parallel, then linked against rocBLAS/rocSOLVER (used only for the FP64 dense
generalized eigenproblem of Rayleigh-Ritz).  Usage: ``python -m
cyclix_msparc_amd.build`` (also called by ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libcyclix_msparc.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    h = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    if not Path(h).exists():
        raise RuntimeError("hipcc not found; the gfx950 library cannot be built")
    return h


# per-file flags: the stencil's plane loop is unrolled by the queue depth
# (13, or 13 pairs), above the compiler's default limit for a forced unroll;
# the queue must stay in registers, which needs the full unroll
def _file_flags(name: str):
    if name.startswith("cyx_stencil_"):
        return ["-mllvm", "-pragma-unroll-threshold=262144"]
    return []


def _compile(src: Path, obj: Path, extra):
    cmd = [_hipcc(), *HIPCC_FLAGS, *_file_flags(src.name), *extra, "-I", str(CSRC), "-c", str(src), "-o",
           str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return r.stderr


def build(verbose: bool = False, extra=()) -> Path:
    OUT_DIR.mkdir(exist_ok=True)
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [objdir / (s.stem + ".o") for s in srcs]
    workers = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        futs = [ex.submit(_compile, s, o, list(extra)) for s, o in zip(srcs, objs)]
        for f in futs:
            msg = f.result()
            if verbose and msg:
                print(msg, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp),
           f"-L{ROCM / 'lib'}", "-lrocsolver", "-lrocblas", "-lrocfft", f"-Wl,-rpath,{ROCM / 'lib'}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
