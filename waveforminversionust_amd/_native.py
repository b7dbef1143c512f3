"""ctypes binding of the C ABI in ``include/wust_hip.h`` (``_lib/libwust_hip.so``).

Loading fails loudly when the library has not been built: there is no CPU
fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libwust_hip.so"

STATUS_IDS = {
    0: "WUST:OK", 1: "WUST:InvalidInput", 2: "WUST:InvalidOutput", 3: "WUST:HIPError",
    4: "WUST:LibraryError", 5: "WUST:SingularSchurBlock", 6: "WUST:NumericalFailure",
    7: "WUST:ResourceCleanup", 8: "WUST:RuntimeError", 9: "WUST:NotConverged",
}

MODE_REFINED, MODE_C64, MODE_PLAIN = 0, 1, 2
MODES = {"refined": MODE_REFINED, "c64": MODE_C64, "plain": MODE_PLAIN}
FACTOR_GAUSS_JORDAN = 0x100
FACTOR_FIXED_BUDGET = 0x200
FACTORS = {"auto": 0, "fixed": FACTOR_FIXED_BUDGET, "gauss-jordan": FACTOR_GAUSS_JORDAN}


class WustError(RuntimeError):
    """Error raised with one of the reference's ``WUST:*`` identifiers."""

    def __init__(self, identifier: str, message: str):
        super().__init__(f"{identifier}: {message}")
        self.identifier = identifier


class GridDesc(C.Structure):
    _fields_ = [
        ("ny", C.c_int32), ("nx", C.c_int32),
        ("x0", C.c_double), ("dx", C.c_double), ("y0", C.c_double), ("dy", C.c_double),
        ("frequency_hz", C.c_double), ("sign_convention", C.c_double),
        ("pml_strength", C.c_double), ("pml_m", C.c_double),
        ("stencil_vmin", C.c_double), ("stencil_vmax", C.c_double),
        ("dispersion", C.c_int32),
        ("ldr9_time_step_s", C.c_double), ("ldr9_reference_speed_mps", C.c_double),
    ]


class Info(C.Structure):
    _fields_ = [
        ("ny", C.c_int32), ("nx", C.c_int32), ("n", C.c_int32), ("m", C.c_int32), ("npad", C.c_int32),
        ("mode", C.c_int32),
        ("b", C.c_double), ("d", C.c_double), ("e", C.c_double),
        ("assembly_ms", C.c_double), ("factor_ms", C.c_double),
        ("min_pivot_ratio", C.c_double), ("factor_bytes", C.c_int64),
        ("inverse_residual", C.c_double), ("iterative_blocks", C.c_int32), ("factor_fallback", C.c_int32),
        ("gj_residual", C.c_double),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("initial_residual", C.c_double),
        ("final_residual", C.c_double), ("solve_ms", C.c_double),
    ]


EXPORTS = {
    "wust_abi_version": (C.c_int, []),
    "wust_status_id": (C.c_char_p, [C.c_int]),
    "wust_last_error": (C.c_char_p, []),
    "wust_stencil_opt_params": (C.c_int, [C.c_double] * 5 + [C.POINTER(C.c_double)] * 3),
    "wust_helmholtz_create": (C.c_int, [C.POINTER(GridDesc), C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                        C.POINTER(C.c_void_p)]),
    "wust_helmholtz_destroy": (C.c_int, [C.c_void_p]),
    "wust_pool_trim": (C.c_int, []),
    "wust_helmholtz_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "wust_helmholtz_coefficients": (C.c_int, [C.c_void_p] * 11),
    "wust_helmholtz_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_double,
                                       C.c_int, C.c_void_p, C.POINTER(SolveStats)]),
    "wust_helmholtz_virtual_sources": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "wust_helmholtz_mass_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int,
                                            C.c_void_p]),
    "wust_oracle_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                   C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_int,
                                   C.POINTER(C.c_double), C.c_void_p, C.c_void_p]),
    "wust_oracle_scales": (C.c_int, [C.c_void_p, C.c_void_p]),
    "wust_directional_curvature": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                             C.POINTER(C.c_double), C.c_void_p]),
    "wust_oracle_release": (C.c_int, [C.c_void_p]),
    "wust_profile_enable": (C.c_int, [C.c_uint32]),
    "wust_profile_read": (C.c_int, [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "wust_profile_reset": (C.c_int, []),
    "wust_profile_units": (C.c_int, [C.c_int, C.POINTER(C.c_int64)]),
    "wust_profile_stride": (C.c_int, [C.c_uint32]),
    "wust_roi_gradient": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "wust_ncg_direction": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_int,
                                     C.c_int64, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.c_void_p]),
    "wust_linearized_step_length": (C.c_int, [C.c_double, C.c_double, C.POINTER(C.c_double)]),
    "wust_rf_dtft": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_int32, C.c_int32, C.c_void_p, C.c_double,
                               C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.c_void_p]),
    "wust_rf_dtft_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_int32, C.c_int32, C.c_void_p,
                                       C.c_double, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "wust_project_slowness_update": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                               C.c_double, C.c_double, C.c_void_p, C.c_int64, C.c_void_p,
                                               C.POINTER(C.c_double), C.c_void_p]),
}

PROFILE_CLASSES = ["assembly", "schur", "gj_panel", "sweep_operand", "sweep_gemm", "residual", "source_fit",
                   "gradient", "plain_factor_step", "plain_solve_step", "ns_gemm"]

_lib = None


def library_path() -> Path:
    return _LIB_PATH


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise RuntimeError(
            f"HIP library {_LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(there is no CPU fallback)")
    # PyTorch-ROCm ships its own HIP runtime, rocBLAS and rocSOLVER under the
    # same sonames as /opt/rocm.  Importing torch first makes this library bind
    # to those already-loaded copies, so one process never holds two HIP
    # runtimes (which crashes at exit).
    import torch  # noqa: F401
    lib = C.CDLL(str(_LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int):
    if status != 0:
        lib = load()
        msg = lib.wust_last_error().decode(errors="replace")
        raise WustError(STATUS_IDS.get(status, "WUST:Unknown"), msg)
