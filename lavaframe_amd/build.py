"""Build the native library lavaframe_amd/_lib/liblfh.so with hipcc (gfx950).

Every .cpp and .hip file in csrc/ is compiled into one shared library whose
only external interface is the C ABI of include/lfh.h.  Device code is built
with -ffp-contract=off (the traversal's bit-exactness depends on one rounding
per operation) and without packed-FP32 instruction selection.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
OBJ_DIR = PKG / "_lib" / "obj"
LIB = OUT_DIR / "liblfh.so"
ARCH = os.environ.get("LFH_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOSTCXX = os.environ.get("LFH_HOSTCXX", "/opt/rocm/bin/amdclang++")
HOST = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]

COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
          f"-I{PKG.parent / 'include'}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
# Per-source device flags.  The guide kernels keep MFMA accumulators in
# architectural VGPRs: the packing of one layer's results into the next
# layer's operand then reads them directly, without a v_accvgpr_read per
# register (58 per 16-query tile of the inference kernel).
DEVICE_EXTRA = {"kernels_guide.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def _compile(src: Path) -> Path:
    obj = OBJ_DIR / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, *DEVICE_EXTRA.get(src.name, []), "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HOSTCXX, *COMMON, *HOST, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.cpp")) + list(CSRC.glob("*.hip")))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-Wl,--no-undefined", "-L/opt/rocm/lib", "-lamdhip64", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # headless CLI over the C ABI (links the library just built)
    cli = PKG / "cli" / "lfh_render.cpp"
    exe = OUT_DIR / "lfh_render"
    cmd = [HOSTCXX, *COMMON, *HOST, str(cli), "-o", str(exe), f"-L{OUT_DIR}", "-llfh",
           "-Wl,-rpath,$ORIGIN", "-x", "none", str(OBJ_DIR / "host_scene.cpp.o")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cli build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size // 1024} KiB) from {len(srcs)} sources for {ARCH}, and {exe.name}")
    return LIB


def build_oracle(verbose: bool = True) -> None:
    """oracle/_build/liboracle.so (+ oracle/_ref/libref.so where the reference
    sources are present).  Test infrastructure only."""
    odir = PKG.parent / "oracle"
    r = subprocess.run(["make", "-f", str(odir / "Makefile"), "-j8", "all"], capture_output=True, text=True,
                       cwd=str(odir))
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(r.stdout.strip().splitlines()[-1] if r.stdout.strip() else "oracle built")


if __name__ == "__main__":
    build()
    if "--no-oracle" not in sys.argv:
        build_oracle()
