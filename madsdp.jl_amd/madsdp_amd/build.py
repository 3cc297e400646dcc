"""Builds libmadsdp_hip.so (the gfx950 C-ABI library) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG.parent / "csrc"
INCLUDE = PKG.parent.parent / "include"
LIB = PKG / "libmadsdp_hip.so"
SOURCES = ["dense.hip", "gemm32.hip", "bf16x3.hip", "eigen.hip", "ops.hip", "qr.hip", "ldlt.hip", "solver.hip",
           "server.hip", "profile.hip"]
ARCH = os.environ.get("MSDP_OFFLOAD_ARCH", "gfx950")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         f"-I{INCLUDE}", f"-I{CSRC}"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and Path(c).exists():
            return c
    return "hipcc"


# host-only AddressSanitizer build of the library's non-kernel code (the
# solver driver, the serving loop's threads, the recorder's pinned ring):
# every -fsanitize= applies to the host compilation only
ASAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
ASAN_LIB = PKG / "libmadsdp_hip_asan.so"


def asan_runtime() -> str:
    """Path of the compiler's shared ASan runtime.  The ASan build loads it as
    a dependency; an uninstrumented Python then needs
    ASAN_OPTIONS=verify_asan_link_order=0 (the runtime is not first in the
    library list)."""
    r = subprocess.run([f"{ROCM}/llvm/bin/clang", "-print-file-name=libclang_rt.asan-x86_64.so"],
                       capture_output=True, text=True)
    return r.stdout.strip()


def build(verbose: bool = False, asan: bool = False) -> Path:
    objdir = PKG.parent.parent / "_build" / ("asan" if asan else "")
    objdir.mkdir(parents=True, exist_ok=True)
    cc = hipcc()
    extra = ASAN_FLAGS if asan else []
    lib = ASAN_LIB if asan else LIB

    def one(src: str) -> Path:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [cc, *FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-4000:]}")
        if verbose:
            print(f"compiled {src}", flush=True)
        return obj

    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(one, SOURCES))
    tmp = lib.with_suffix(".so.tmp")
    link_extra = []
    if asan:  # the shared ASan runtime as a dependency of the library, found by rpath
        link_extra = ["-Xarch_host", "-fsanitize=address", "-shared-libasan",
                      f"-Wl,-rpath,{os.path.dirname(asan_runtime())}"]
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), *link_extra, "-o",
           str(tmp), f"-L{ROCM}/lib", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, lib)
    return lib


ASAN_DRIVER_SRC = PKG.parent.parent / "tests" / "asan" / "driver.cpp"
ASAN_DRIVER = ASAN_DRIVER_SRC.with_suffix("")


def build_asan_driver() -> Path:
    """The host-ASan test driver (tests/asan/driver.cpp) against the ASan
    build of the library: instrumented executable, the ASan runtime loaded
    first as its dependency."""
    rt_dir = os.path.dirname(asan_runtime())
    cmd = [f"{ROCM}/llvm/bin/clang++", "-O1", "-g", "-std=c++17", "-fsanitize=address",
           "-shared-libasan", "-fno-omit-frame-pointer", f"-I{INCLUDE}", str(ASAN_DRIVER_SRC),
           str(ASAN_LIB), f"-Wl,-rpath,{PKG}", f"-Wl,-rpath,{rt_dir}", "-o", str(ASAN_DRIVER)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ASan driver build failed:\n{r.stderr[-4000:]}")
    return ASAN_DRIVER


if __name__ == "__main__":
    if "--asan" in sys.argv:
        print(build(verbose=True, asan=True))
        print(build_asan_driver())
    else:
        print(build(verbose=True))
    sys.exit(0)
