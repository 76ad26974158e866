"""Builds zxtalori_amd/_lib/libzxtalori.so for gfx950 with hipcc (no CMake, no JIT).

    python -m zxtalori_amd.build         # or python zxtalori_amd/build.py

Everything is compiled with -ffp-contract=off (the exact-order image/chord arithmetic
depends on it; the pair kernel writes its fmas explicitly); device code with correctly
rounded FP32 division and sqrt and without packed-FP32 instructions.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "_lib" / "libzxtalori.so"
ARCH = os.environ.get("ZXTALORI_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-ffp-contract=off"]
DEVICE = [f"--offload-arch={ARCH}", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["kernels.hip", "spacing_static.hip", "engine.cpp", "batch.cpp", "abi.cpp", "hopf.cpp", "transport.cpp"]


def hipcc():
    found = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(found).exists():
        raise RuntimeError("hipcc not found (ROCm is required to build libzxtalori.so)")
    return found


def compile_one(src, obj):
    args = [hipcc(), *COMMON]
    if src.suffix == ".hip":
        args += ["-x", "hip", *DEVICE]
    args += ["-c", str(src), "-o", str(obj)]
    out = subprocess.run(args, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(args)}\n{out.stdout}\n{out.stderr}")
    return out.stderr


def build(verbose=False):
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    jobs = []
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as pool:
        for name in SOURCES:
            src = CSRC / name
            obj = objdir / (src.stem + ".o")
            jobs.append(pool.submit(compile_one, src, obj))
        logs = [j.result() for j in jobs]
    if verbose:
        for log in logs:
            if log.strip():
                print(log, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    args = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}",
            *[str(objdir / (Path(s).stem + ".o")) for s in SOURCES], "-o", str(tmp)]
    out = subprocess.run(args, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(args)}\n{out.stdout}\n{out.stderr}")
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
