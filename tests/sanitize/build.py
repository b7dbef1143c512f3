"""Build the C-ABI driver against the sanitized debug library.

Builds ``waveforminversionust_amd/_lib/libwust_hip_debug.so`` (host code under
AddressSanitizer and UndefinedBehaviorSanitizer, a synchronizing error check after
every kernel launch) and ``tests/sanitize/_bin/abi_driver``, a host program linked
with the same sanitizer runtime.  Run from the repository root:
``python tests/sanitize/build.py``.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
BIN = HERE / "_bin"


def build(verbose: bool = False) -> Path:
    sys.path.insert(0, str(ROOT))
    from waveforminversionust_amd import build as libbuild

    lib = libbuild.build(debug=True)
    BIN.mkdir(exist_ok=True)
    exe = BIN / "abi_driver"
    rocm = libbuild.ROCM
    cmd = [libbuild._clangxx(), "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize=function", "-shared-libsan", "-D__HIP_PLATFORM_AMD__",
           f"-I{ROOT / 'include'}", f"-I{rocm}/include", str(HERE / "abi_driver.cpp"), "-o", str(exe),
           f"-L{lib.parent}", "-lwust_hip_debug", f"-L{rocm}/lib", "-lamdhip64",
           f"-Wl,-rpath,{lib.parent}", f"-Wl,-rpath,{rocm}/lib", f"-Wl,-rpath,{libbuild.sanitizer_runtime_dir()}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"abi_driver build failed:\n{res.stdout}\n{res.stderr}")
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
