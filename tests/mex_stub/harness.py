"""ctypes driver of the MEX gateway compiled against the stub mex.h (libwust_mex_stub.so): builds
MATLAB-layout arrays (column-major, interleaved complex), calls mexFunction and returns its outputs as
NumPy arrays, raising MexCallError(identifier, message) where MATLAB would raise."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import build as B

_CLS = {0: "double", 1: "logical", 2: "char", 3: "uint64"}


class MexCallError(RuntimeError):
    def __init__(self, identifier, message):
        super().__init__(f"{identifier}: {message}")
        self.identifier = identifier


class Mex:
    def __init__(self):
        from waveforminversionust_amd import _native as N
        N.load()                       # the engine library first (one HIP runtime per process)
        if not B.LIB.exists():
            B.build()
        lib = C.CDLL(str(B.LIB))
        lib.stub_array.restype = C.c_void_p
        lib.stub_array.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_size_t), C.c_void_p]
        lib.stub_string.restype = C.c_void_p
        lib.stub_string.argtypes = [C.c_char_p]
        lib.stub_describe.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_size_t), C.c_int]
        lib.stub_data.restype = C.c_void_p
        lib.stub_data.argtypes = [C.c_void_p]
        lib.stub_bytes.restype = C.c_size_t
        lib.stub_bytes.argtypes = [C.c_void_p]
        lib.stub_destroy.argtypes = [C.c_void_p]
        lib.stub_call.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p), C.c_char_p,
                                  C.c_char_p, C.c_int]
        self.lib = lib

    def array(self, x):
        """A MATLAB array from a Python value: str -> char, bool arrays -> logical, np.uint64 -> uint64,
        anything else -> double (complex when the value is complex)."""
        if isinstance(x, str):
            return self.lib.stub_string(x.encode())
        a = np.asarray(x)
        if a.dtype == np.bool_:
            cls, cx, a = 1, 0, a.astype(np.uint8)
        elif a.dtype == np.uint64:
            cls, cx = 3, 0
        elif np.iscomplexobj(a):
            cls, cx, a = 0, 1, a.astype(np.complex128)
        else:
            cls, cx, a = 0, 0, a.astype(np.float64)
        dims = list(a.shape) if a.ndim >= 2 else ([1, a.size] if a.ndim == 1 else [1, 1])
        buf = np.asfortranarray(a.reshape(dims, order="F") if a.ndim < 2 else a).tobytes(order="F")
        d = (C.c_size_t * len(dims))(*dims)
        return self.lib.stub_array(cls, cx, len(dims), d, C.c_char_p(buf) if buf else None)

    def value(self, p):
        cls, cx, nd = C.c_int(), C.c_int(), C.c_int()
        dims = (C.c_size_t * 8)()
        self.lib.stub_describe(p, C.byref(cls), C.byref(cx), C.byref(nd), dims, 8)
        shape = tuple(dims[i] for i in range(nd.value))
        nbytes = self.lib.stub_bytes(p)
        raw = C.string_at(self.lib.stub_data(p), nbytes) if nbytes else b""
        kind = _CLS[cls.value]
        dt = {"double": np.complex128 if cx.value else np.float64, "logical": np.uint8, "char": np.uint16,
              "uint64": np.uint64}[kind]
        return np.frombuffer(raw, dtype=dt).reshape(shape, order="F").copy()

    def __call__(self, nlhs, *args):
        ins = [self.array(a) for a in args]
        prhs = (C.c_void_p * max(len(ins), 1))(*ins)
        plhs = (C.c_void_p * max(nlhs, 1))()
        ident, msg = C.create_string_buffer(512), C.create_string_buffer(4096)
        try:
            st = self.lib.stub_call(nlhs, plhs, len(ins), prhs, ident, msg, 4096)
        finally:
            for p in ins:
                self.lib.stub_destroy(p)
        outs = []
        for i in range(nlhs):
            if plhs[i]:
                outs.append(self.value(plhs[i]))
                self.lib.stub_destroy(plhs[i])
            else:
                outs.append(None)
        if st != 0:
            raise MexCallError(ident.value.decode(), msg.value.decode(errors="replace"))
        return outs[0] if nlhs == 1 else tuple(outs)
