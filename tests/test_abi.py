"""C-ABI library checks that need no GPU: it loads, exports every declared
entry point, and its host-only pieces and argument validation behave."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import helmholtz_fp64 as H
from waveforminversionust_amd import _native as N

HEADER = Path(__file__).resolve().parents[1] / "include" / "wust_hip.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(wust_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = N.load()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), f"{name} missing from {N.library_path()}"
        assert name in N.EXPORTS, f"{name} has no ctypes prototype"
    assert lib.wust_abi_version() == 3


def test_status_ids_match_reference_identifiers():
    lib = N.load()
    assert lib.wust_status_id(1) == b"WUST:InvalidInput"
    assert lib.wust_status_id(5) == b"WUST:SingularSchurBlock"
    assert lib.wust_status_id(6) == b"WUST:NumericalFailure"
    for k, v in N.STATUS_IDS.items():
        if k in (3, 4):
            continue
        assert lib.wust_status_id(k).decode() == v


@pytest.mark.parametrize("args", [(1400, 1700, 1e5, 1e-3, 1.0), (1400, 1700, 0.8e6, 0.25e-3, 1.0),
                                  (1300, 1800, 5e5, 5e-4, 1.3), (1400, 1600, 0.5e6, 0.5e-3, 1.0)])
def test_stencil_params_host_path_matches_oracle(args):
    lib = N.load()
    b, d, e = C.c_double(), C.c_double(), C.c_double()
    assert lib.wust_stencil_opt_params(*args, C.byref(b), C.byref(d), C.byref(e)) == 0
    rb, rd, re_ = H.stencil_opt_params(*args)
    assert b.value == rb
    np.testing.assert_allclose([d.value, e.value], [rd, re_], rtol=1e-12)


def test_argument_validation_without_device():
    lib = N.load()
    b = C.c_double()
    st = lib.wust_stencil_opt_params(1700, 1400, 1e5, 1e-3, 1.0, C.byref(b), C.byref(b), C.byref(b))
    assert st == 1 and b"stencil_bounds" in lib.wust_last_error()
    h = C.c_void_p()
    assert lib.wust_helmholtz_create(None, None, None, 0, None, C.byref(h)) == 1
    desc = N.GridDesc(ny=2, nx=31, dx=1e-3, dy=1e-3, frequency_hz=1e5, sign_convention=-1, pml_strength=10,
                      pml_m=4e-3, stencil_vmin=1400, stencil_vmax=1700)
    fake = C.c_void_p(1)  # never dereferenced: validation fails first
    assert lib.wust_helmholtz_create(C.byref(desc), fake, None, 0, None, C.byref(h)) == 1
    desc.ny = 31
    desc.stencil_vmin = 1800
    assert lib.wust_helmholtz_create(C.byref(desc), fake, None, 0, None, C.byref(h)) == 1
    assert b"stencil_bounds" in lib.wust_last_error()
    desc.stencil_vmin = 1400
    assert lib.wust_helmholtz_create(C.byref(desc), fake, None, 7, None, C.byref(h)) == 1
    assert b"unknown mode" in lib.wust_last_error()
    assert lib.wust_helmholtz_create(C.byref(desc), fake, None, 0x400, None, C.byref(h)) == 1
    assert b"unknown mode flag" in lib.wust_last_error()
    assert lib.wust_helmholtz_create(C.byref(desc), fake, None, 0x300, None, C.byref(h)) == 1
    assert b"no Newton-Schulz budget" in lib.wust_last_error()
    assert not h.value


def test_python_mirror_rejects_like_the_reference():
    from waveforminversionust_amd.helmholtz import HelmholtzSolver, WustError
    x = np.arange(31) * 1e-3
    c = 1500 * np.ones((31, 31))
    for bad in [dict(stencil_bounds=[1400, 1700]), dict(backend="gpu"), dict(backend="gpu", stencil_bounds=None),
                dict(backend="gpu", stencil_bounds=[1500]), dict(backend="tpu", stencil_bounds=[1400, 1700]),
                dict(backend="gpu", stencil_bounds=[1400, 1700], factor="lu"),
                dict(backend="gpu", stencil_bounds=[1400, 1700], mode="plain", factor="gauss-jordan")]:
        with pytest.raises(WustError) as ei:
            HelmholtzSolver(x, x, c, np.zeros_like(c), 1e5, -1, 10, .004, bad)
        assert ei.value.identifier == "WUST:InvalidInput"


def test_field_shape_guard_on_host():
    """The Python layer hands the engine plain pointers from which it reads ny * nx * R values, so a
    field whose leading dimensions are not the grid's is WUST:InvalidInput before any device work
    (Runtime/solver/applyBlockLU.cu:380-388 rejects such sources); the accepted layout is
    [ny, nx] or [ny, nx, R], stored as (R, nx, ny)."""
    torch = pytest.importorskip("torch")
    from waveforminversionust_amd import helmholtz as G
    ny, nx = 7, 5
    good = np.arange(ny * nx * 3, dtype=np.float64).reshape(ny, nx, 3) * (1 + 1j)
    t = G._field_to_dev(good, "cpu", ny, nx, "bad")
    assert tuple(t.shape) == (3, nx, ny) and t.dtype == torch.complex128
    assert torch.equal(t.permute(2, 1, 0), torch.as_tensor(good))
    assert tuple(G._field_to_dev(good[:, :, 0], "cpu", ny, nx, "bad").shape) == (1, nx, ny)
    for bad in (np.transpose(good, (2, 0, 1)), good[:-1], good[:, :-1], good[:, :, :0], good[..., None],
                np.zeros(ny * nx)):
        with pytest.raises(N.WustError) as ei:
            G._field_to_dev(bad, "cpu", ny, nx, "Inconsistent Block-LU input shapes.")
        assert ei.value.identifier == "WUST:InvalidInput"
    Y = torch.zeros((4, 3), dtype=torch.complex128)    # [n_rx, n_tx] after the transpose
    M = torch.zeros((4, 3), dtype=torch.uint8)
    G._check_obs_slice(Y, M, 3, 4)
    for ntx, nrx in ((4, 3), (3, 5), (2, 4)):
        with pytest.raises(N.WustError):
            G._check_obs_slice(Y, M, ntx, nrx)
