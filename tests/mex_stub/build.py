"""Compile the MEX gateway (waveforminversionust_amd/matlab/wust_hip_mex.cpp) against the stub mex.h /
matrix.h of this folder into libwust_mex_stub.so, linked to the engine library, so that the tests can
drive mexFunction without MATLAB.  The engine library must be built first."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
LIB = HERE / "libwust_mex_stub.so"


def build(verbose: bool = False) -> Path:
    engine = ROOT / "waveforminversionust_amd" / "_lib"
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-D__HIP_PLATFORM_AMD__",
           f"-I{HERE}", f"-I{ROOT / 'include'}", f"-I{ROCM}/include",
           str(ROOT / "waveforminversionust_amd" / "matlab" / "wust_hip_mex.cpp"), str(HERE / "mex_stub.cpp"),
           f"-L{engine}", "-lwust_hip", f"-L{ROCM}/lib", "-lamdhip64",
           f"-Wl,-rpath,{engine}", f"-Wl,-rpath,{ROCM}/lib", "-Wl,-rpath,$ORIGIN/../../waveforminversionust_amd/_lib",
           "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"building the MEX stub failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
