"""Python side of the mx shim: build mxArrays from numpy values and call a
MEX entry point, getting back numpy outputs or the raised (id, message)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(os.path.dirname(os.path.dirname(HERE)), "matlab-cuda-efief-vs-cpu_amd", "_build")
_shim = None
_mex = {}


class MexError(Exception):
    def __init__(self, ident, msg):
        super().__init__(f"{ident}: {msg}")
        self.ident = ident


def shim():
    global _shim
    if _shim is None:
        L = C.CDLL(os.path.join(OUT, "libmxshim.so"), mode=C.RTLD_GLOBAL)
        L.shim_double.restype = C.c_void_p
        L.shim_double.argtypes = [C.c_size_t, C.c_size_t, C.c_int, C.c_void_p]
        L.shim_int32.restype = C.c_void_p
        L.shim_int32.argtypes = [C.c_size_t, C.c_size_t, C.c_void_p]
        L.shim_nd.restype = C.c_void_p
        L.shim_nd.argtypes = [C.c_int, C.c_size_t, C.POINTER(C.c_size_t), C.c_void_p]
        L.shim_struct.restype = C.c_void_p
        L.shim_struct.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_void_p)]
        L.shim_data.restype = C.c_void_p
        L.shim_data.argtypes = [C.c_void_p]
        L.mxGetM.restype = C.c_size_t
        L.mxGetM.argtypes = [C.c_void_p]
        L.mxGetN.restype = C.c_size_t
        L.mxGetN.argtypes = [C.c_void_p]
        L.shim_is_complex.argtypes = [C.c_void_p]
        L.shim_free.argtypes = [C.c_void_p]
        L.shim_call.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
        L.shim_err_id.restype = C.c_char_p
        L.shim_err_msg.restype = C.c_char_p
        _shim = L
    return _shim


def entry(name):
    if name not in _mex:
        shim()
        L = C.CDLL(os.path.join(OUT, f"mex_{name}.so"))
        _mex[name] = C.cast(L.mexFunction, C.c_void_p)
        _mex[name + "_lib"] = L
    return _mex[name]


def to_mx(v):
    S = shim()
    if isinstance(v, dict):
        names = (C.c_char_p * len(v))(*[k.encode() for k in v])
        vals = (C.c_void_p * len(v))(*[to_mx(x) for x in v.values()])
        return S.shim_struct(len(v), names, vals)
    a = np.asarray(v)
    if a.ndim == 0:
        a = a.reshape(1, 1)
    elif a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.dtype == np.uint8 or a.ndim > 2:
        # N-d or uint8: class 1 double, 2 int32, 4 uint8
        cls = 4 if a.dtype == np.uint8 else (2 if a.dtype == np.int32 else 1)
        a = np.asfortranarray(a if cls != 1 else a.astype(np.float64))
        dims = (C.c_size_t * a.ndim)(*a.shape)
        return S.shim_nd(cls, a.ndim, dims, a.ctypes.data)
    if a.dtype == np.int32:
        a = np.asfortranarray(a)
        return S.shim_int32(a.shape[0], a.shape[1], a.ctypes.data)
    if np.iscomplexobj(a):
        a = np.asfortranarray(a, dtype=np.complex128)
        return S.shim_double(a.shape[0], a.shape[1], 1, a.ctypes.data)
    a = np.asfortranarray(a, dtype=np.float64)
    return S.shim_double(a.shape[0], a.shape[1], 0, a.ctypes.data)


def from_mx(p):
    S = shim()
    m, n = S.mxGetM(p), S.mxGetN(p)
    cplx = S.shim_is_complex(p)
    buf = (C.c_double * (m * n * (2 if cplx else 1))).from_address(S.shim_data(p))
    a = np.frombuffer(buf, dtype=np.complex128 if cplx else np.float64).copy()
    return a.reshape((m, n), order="F")


def call(name, nlhs, *args):
    S = shim()
    f = entry(name)
    prhs = (C.c_void_p * max(1, len(args)))(*[to_mx(a) for a in args])
    plhs = (C.c_void_p * max(1, nlhs))()
    rc = S.shim_call(f, nlhs, plhs, len(args), prhs)
    try:
        if rc:
            raise MexError(S.shim_err_id().decode(), S.shim_err_msg().decode())
        return [from_mx(plhs[i]) for i in range(nlhs)]
    finally:
        for i in range(len(args)):
            S.shim_free(prhs[i])
        if not rc:
            for i in range(nlhs):
                S.shim_free(plhs[i])
