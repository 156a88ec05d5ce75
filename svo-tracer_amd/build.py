"""Build recipe for libsvo.so (HIP, gfx950) -- run `python svo-tracer_amd/build.py`.

Every translation unit is compiled by hipcc for gfx950 with the flags the
path's defined floating-point semantics depend on (-ffp-contract=off; HIP's
default correctly rounded f32 division and square root) and linked in-tree to
svo-tracer_amd/libsvo.so, which is what tests, smoke() and bench.py load.

`python svo-tracer_amd/build.py --sanitize` builds svo-tracer_amd/_asan/libsvo.so instead:
the same sources with AddressSanitizer and UndefinedBehaviorSanitizer in the host code only
(device code is compiled as usual).  The CPU tests run against it through SVO_LIB with the
compiler's ASan runtime preloaded (see the `asan-check` target of oracle/Makefile).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shlex
import shutil
import sys
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libsvo.so"

ARCH = os.environ.get("SVO_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    f"--offload-arch={ARCH}",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
    "-Wall",
    "-Wno-unused-function",
]
DEVICE_ONLY = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# host-code sanitizers; every -fsanitize= sits right after -Xarch_host, so device code is untouched
HOST_SANITIZE = shlex.split("-g -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined "
                            "-Xarch_host -fno-sanitize-recover=undefined -Xarch_host -fno-omit-frame-pointer")


def _compile(src: Path, out: Path = OUT, extra=()) -> Path:
    obj = out / (src.name + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".hip":
        cmd[1:1] = ["-x", "hip"]
        cmd += DEVICE_ONLY
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False, sanitize: bool = False) -> Path:
    out = PKG / "_asan_build" if sanitize else OUT
    lib = PKG / "_asan" / "libsvo.so" if sanitize else LIB
    out.mkdir(exist_ok=True)
    lib.parent.mkdir(exist_ok=True)
    extra = HOST_SANITIZE if sanitize else ()
    sources = sorted(list(CSRC.glob("*.cpp")) + list(CSRC.glob("*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda src: _compile(src, out, extra), sources))
    tmp = lib.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs), "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build(verbose=True, sanitize="--sanitize" in sys.argv[1:])
