"""Build the gfx950 HIP library (``_lib/libwust_hip.so``) in-tree with hipcc.

Every ``csrc/*.hip`` file is compiled for gfx950 only and linked with
rocSOLVER/rocBLAS (used by the plain comparator path).  Object files go to
``_build/`` next to the package; the shared library goes to ``_lib/``.

``build(debug=True)`` makes ``_lib/libwust_hip_debug.so`` instead: the same
sources with ``WUST_DEBUG_SYNC`` (a device synchronization and error check
after every launch) and the host code instrumented by AddressSanitizer and
UndefinedBehaviorSanitizer (device code is not instrumented).  A program that
loads it must carry the sanitizer runtime itself (tests/sanitize builds such
a driver of the C ABI).
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
LIBDIR = PKG / "_lib"
BUILD = PKG.parent / "_build"
LIBNAME = "libwust_hip.so"
DEBUG_LIBNAME = "libwust_hip_debug.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions stay off (see the engine's numerics notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function", "-Wno-unused-variable",
]
# debug build: synchronizing launch checks; sanitizers on the host pass only.  UBSan's function-type
# check is left out: with it (and identical device code) the kernels' results came out wrong on the
# GPU -- the Gauss-Jordan inverses all zero -- with no sanitizer report.
DEBUG_FLAGS = [
    "-DWUST_DEBUG_SYNC", "-g", "-fno-omit-frame-pointer",
    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize=function",
]


def _hipcc():
    exe = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return exe


def _clangxx():
    exe = os.path.join(ROCM, "llvm", "bin", "clang++")
    if not os.path.exists(exe):
        raise RuntimeError("clang++ not found under ROCM_PATH; the debug library cannot be linked")
    return exe


def sanitizer_runtime_dir() -> str:
    """Directory of clang's shared sanitizer runtimes (rpath of the debug library and its drivers)."""
    res = subprocess.run([_clangxx(), "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = res.stdout.strip()
    if res.returncode != 0 or not os.path.isabs(path):
        raise RuntimeError("clang's AddressSanitizer runtime not found")
    return os.path.dirname(path)


def _compile(src: Path, obj: Path, verbose: bool, extra=()):
    cmd = [_hipcc(), *HIPCC_FLAGS, *extra, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
    return src.name, res.stderr


def build(verbose: bool = False, jobs: int | None = None, debug: bool = False) -> Path:
    """Compile every csrc/*.hip for gfx950 and link the shared library (the debug one when debug)."""
    bdir = BUILD / "debug" if debug else BUILD
    bdir.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objs = [bdir / (s.stem + ".o") for s in sources]
    jobs = jobs or min(len(sources), 8)
    extra = DEBUG_FLAGS if debug else ()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for name, err in ex.map(lambda so: _compile(so[0], so[1], verbose, extra), zip(sources, objs)):
            if verbose and err.strip():
                print(f"[{name}] {err}", file=sys.stderr)
    libname = DEBUG_LIBNAME if debug else LIBNAME
    out = LIBDIR / libname
    tmp = LIBDIR / (libname + ".tmp")
    libs = [f"-L{ROCM}/lib", "-lrocsolver", "-lrocblas", "-lhipfft", f"-Wl,-rpath,{ROCM}/lib"]
    if debug:
        # host link through clang with the shared sanitizer runtime (the objects carry no device link step)
        cmd = [_clangxx(), "-shared", "-fPIC", "-fsanitize=address,undefined", "-shared-libsan", *map(str, objs),
               "-o", str(tmp), *libs, "-lamdhip64", f"-Wl,-rpath,{sanitizer_runtime_dir()}"]
    else:
        cmd = [_hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", *map(str, objs), "-o", str(tmp), *libs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, debug="--debug" in sys.argv))
