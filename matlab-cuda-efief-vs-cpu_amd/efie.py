"""Python front end of libefie_hip.so (the MI355X EFIE hot path).

Mirrors the reference's host-side MATLAB functions and its two MEX entry
points with the same names, argument order and meaning:

  make_patch_with_ground_mesh   make_patch_with_ground_mesh.m:1-46
  build_rwg_interior            build_rwg_interior.m:1-113
  pick_port_edge_by_point_and_layer  pick_port_edge_by_point_and_layer.m:1-11
  patch_sweep_cuda              patch_sweep_cuda.cu:4-27      -> (Zin, S11)
  gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST  gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:127-148 -> (Z, Vrhs)

All compute goes through the C ABI declared in include/efie.h; there is no
fallback: if the shared library is missing this module raises.  Index
tables are 1-based int32 and matrices column-major, as in MATLAB.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libefie_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

EFIE_OK = 0
MODEL_QUADRATURE, MODEL_CENTROID = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
LU_AUTO, LU_FP32, LU_BF16X6, LU_SYM = 0, 1, 2, 3
NEAR_DIRECT, NEAR_MOMENTS, NEAR_EXTRACT, NEAR_SURROGATE = 0, 1, 2, 3

# physics constants of the reference drivers (main_patch_efie_rwg.m:218-220)
C0 = 299792458.0
MU0 = 4 * np.pi * 1e-7
EPS0 = 1.0 / (MU0 * C0 ** 2)


class EfieError(RuntimeError):
    """Error raised from the C ABI; `ident` is the reference MEX error id."""

    def __init__(self, status, ident, msg):
        super().__init__(f"{ident}: {msg}")
        self.status = status
        self.ident = ident


class Opts(C.Structure):
    _fields_ = [
        ("quad_order", C.c_int), ("near_sing_eps", C.c_double), ("self_subdiv", C.c_int),
        ("self_alpha", C.c_double), ("near_subdiv", C.c_int), ("near_alpha", C.c_double),
        ("near_order", C.c_int), ("model", C.c_int), ("precision", C.c_int), ("batch", C.c_int),
        ("ir_tol", C.c_double), ("ir_max_iter", C.c_int), ("near_mode", C.c_int), ("profile", C.c_int),
        ("lu_kernel", C.c_int),
    ]


KERNEL_CLASSES = ["regular", "near", "self", "poly", "combine", "tslu", "laswp", "panel", "gemm",
                  "trsv", "other", "surrogate", "series"]


class Stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double), ("ms_fill", C.c_double), ("ms_factor", C.c_double),
        ("ms_solve", C.c_double), ("n_near_pairs", C.c_longlong), ("n_green", C.c_longlong),
        ("ir_iters_max", C.c_int), ("ir_resid_max", C.c_double), ("moment_terms", C.c_int),
        ("ms_moments", C.c_double), ("ms_kernel", C.c_double * len(KERNEL_CLASSES)),
        ("n_kernel", C.c_int * len(KERNEL_CLASSES)),
        ("n_sur_pairs", C.c_longlong), ("n_sur_fallback", C.c_longlong), ("n_lu_retry", C.c_int),
        ("n_fp64_retry", C.c_int),
        ("ir_iters_total", C.c_longlong),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("ms_kernel", "n_kernel")}
        d["kernels"] = {name: {"ms": self.ms_kernel[i], "n": self.n_kernel[i]}
                        for i, name in enumerate(KERNEL_CLASSES) if self.n_kernel[i]}
        return d


class SweepItem(C.Structure):
    """efie_sweep_item of efie.h."""
    _fields_ = [("problem", C.c_void_p), ("freqs", C.POINTER(C.c_double)), ("Nf", C.c_int),
                ("patchEdge", C.c_int), ("groundEdge", C.c_int), ("V0", C.c_double), ("spread", C.c_double),
                ("Z0", C.c_double), ("Zin", C.POINTER(C.c_double)), ("S11", C.POINTER(C.c_double)),
                ("release", C.c_int)]


class SurCfg(C.Structure):
    """efie_sur_cfg (include/efie_surrogate.h)."""
    _fields_ = [
        ("n_pairs", C.c_longlong), ("n_val", C.c_int), ("epochs", C.c_int), ("batch", C.c_int),
        ("lr", C.c_double), ("kl_max", C.c_double), ("seed", C.c_ulonglong),
    ]


class SurReport(C.Structure):
    """efie_sur_report (include/efie_surrogate.h)."""
    _fields_ = [
        ("n_train", C.c_longlong), ("n_val", C.c_longlong), ("steps", C.c_int), ("ms_sample", C.c_double),
        ("ms_label", C.c_double), ("ms_train", C.c_double), ("train_loss", C.c_double), ("val_loss", C.c_double),
        ("val_rel_rms", C.c_double), ("val_rel_max", C.c_double), ("base_rel_rms", C.c_double),
        ("base_rel_max", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTS = [
    "efie_opts_default", "efie_opts_preset", "efie_last_error", "efie_last_error_id",
    "efie_mesh_counts", "efie_make_mesh", "efie_rwg_count", "efie_build_rwg", "efie_pick_port",
    "efie_create", "efie_destroy", "efie_set_mesh", "efie_set_rwg", "efie_assemble_zv",
    "efie_sweep", "efie_sweep_many", "efie_zgesv", "efie_zgesv_noalloc", "efie_diag_copy", "efie_diag_pack_plan", "efie_diag_cache_counts", "efie_diag_near_pairs", "efie_diag_quad_points", "efie_rwg_count_device", "efie_build_rwg_device",
    "efie_pick_port_device", "efie_set_prepack", "efie_set_rwg_tables", "efie_assemble_z",
    "efie_sur_cfg_default", "efie_sur_blob_floats", "efie_sur_train", "efie_sur_get", "efie_sur_set",
    "efie_sur_set_exact", "efie_mlp_params", "efie_mlp_forward", "efie_mlp_loss_grad",
]

_lib = None


def lib():
    """Load libefie_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with python matlab-cuda-efief-vs-cpu_amd/build.py")
        L = C.CDLL(LIB_PATH)
        L.efie_opts_default.argtypes = [C.POINTER(Opts)]
        L.efie_opts_default.restype = None
        L.efie_opts_preset.argtypes = [C.POINTER(Opts), C.c_char_p]
        L.efie_last_error.restype = C.c_char_p
        L.efie_last_error_id.restype = C.c_char_p
        L.efie_mesh_counts.argtypes = [C.c_int] * 4 + [_ip, _ip]
        L.efie_make_mesh.argtypes = [C.c_double] * 5 + [C.c_int] * 4 + [_dp, _ip]
        L.efie_rwg_count.argtypes = [_ip, C.c_int, C.c_int, _ip]
        L.efie_build_rwg.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int] + [_ip] * 5 + [_dp] * 6 + [_ip]
        L.efie_pick_port.argtypes = [C.c_int, _dp, _ip, _dp, C.c_int, _ip]
        L.efie_rwg_count_device.argtypes = [C.c_void_p, _ip]
        L.efie_build_rwg_device.argtypes = [C.c_void_p, C.c_int] + [_ip] * 5 + [_dp] * 6 + [_ip]
        L.efie_pick_port_device.argtypes = [C.c_void_p, _dp, C.c_int, _ip]
        L.efie_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        L.efie_destroy.argtypes = [C.c_void_p]
        L.efie_destroy.restype = None
        L.efie_set_mesh.argtypes = [C.c_void_p, _dp, C.c_int, _ip, C.c_int]
        L.efie_set_rwg.argtypes = [C.c_void_p] + [_ip] * 5 + [_dp] * 5 + [C.c_int]
        L.efie_assemble_zv.argtypes = [C.c_void_p] + [C.c_double] * 6 + [_dp, C.c_int, C.POINTER(Opts), _dp, _dp]
        L.efie_set_prepack.argtypes = [C.c_void_p, C.c_void_p]
        L.efie_set_rwg_tables.argtypes = [C.c_void_p, C.c_int] + [_ip] * 4 + [_dp] * 5 + [C.c_int]
        L.efie_assemble_z.argtypes = [C.c_void_p] + [C.c_double] * 6 + [C.POINTER(Opts), _dp]
        L.efie_sweep.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.c_double, C.c_double, C.c_double, C.POINTER(Opts), _dp, _dp, C.POINTER(Stats)]
        L.efie_sweep_many.argtypes = [C.c_void_p, C.POINTER(SweepItem), C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.POINTER(Opts), C.POINTER(Stats)]
        L.efie_zgesv.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, C.POINTER(Stats)]
        L.efie_zgesv_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, C.POINTER(Stats)]
        L.efie_diag_copy.argtypes = [C.c_void_p, C.c_size_t, C.c_int, _dp]
        L.efie_diag_pack_plan.argtypes = [_ip, _ip, C.c_int, C.c_int, C.c_int, C.c_double, _ip, C.c_int, _ip]
        L.efie_diag_cache_counts.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.efie_diag_near_pairs.argtypes = [C.c_void_p, C.POINTER(Opts), C.POINTER(C.c_longlong), _ip, C.c_longlong]
        L.efie_diag_quad_points.argtypes = [C.c_void_p, C.POINTER(Opts), C.c_int, C.POINTER(C.c_longlong), _dp,
                                            C.c_longlong]
        _fp = C.POINTER(C.c_float)
        L.efie_sur_cfg_default.argtypes = [C.POINTER(SurCfg)]
        L.efie_sur_cfg_default.restype = None
        L.efie_sur_blob_floats.restype = C.c_int
        L.efie_mlp_params.restype = C.c_int
        L.efie_sur_train.argtypes = [C.c_void_p, C.POINTER(Opts), C.POINTER(SurCfg), C.POINTER(SurReport)]
        L.efie_sur_get.argtypes = [C.c_void_p, _fp]
        L.efie_sur_set.argtypes = [C.c_void_p, _fp]
        L.efie_sur_set_exact.argtypes = [C.c_void_p, C.c_int]
        L.efie_mlp_forward.argtypes = [C.c_void_p, _fp, _fp, C.c_int, _fp]
        L.efie_mlp_loss_grad.argtypes = [C.c_void_p, _fp, _fp, _fp, C.c_int, _fp, _fp]
        _lib = L
    return _lib


def _check(st):
    if st != EFIE_OK:
        L = lib()
        raise EfieError(st, L.efie_last_error_id().decode(), L.efie_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def opts(preset: str = "accurate", **kw) -> Opts:
    o = Opts()
    _check(lib().efie_opts_preset(C.byref(o), preset.encode()))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


# ----------------------------------------------------------- host tables --

def make_patch_with_ground_mesh(Lp, Wp, Lg, Wg, h, Np_x, Np_y, Ng_x, Ng_y):
    """[V, F] of make_patch_with_ground_mesh.m (F 1-based int32)."""
    Nv, Nt = C.c_int(), C.c_int()
    _check(lib().efie_mesh_counts(Np_x, Np_y, Ng_x, Ng_y, C.byref(Nv), C.byref(Nt)))
    V = np.zeros((Nv.value, 3), order="F")
    F = np.zeros((Nt.value, 3), dtype=np.int32, order="F")
    _check(lib().efie_make_mesh(Lp, Wp, Lg, Wg, h, Np_x, Np_y, Ng_x, Ng_y, _d(V), _i(F)))
    return V, F


@dataclass
class Rwg:
    """The rwg struct of build_rwg_interior.m."""
    edge: np.ndarray
    plusTri: np.ndarray
    minusTri: np.ndarray
    plusSign: np.ndarray
    minusSign: np.ndarray
    rp: np.ndarray
    rm: np.ndarray
    Ap: np.ndarray
    Am: np.ndarray
    len: np.ndarray
    center: np.ndarray
    zSign: np.ndarray

    @property
    def Ne(self) -> int:
        return int(self.plusTri.shape[0])


def _empty_rwg(Ne) -> Rwg:
    return Rwg(edge=np.zeros((Ne, 2), np.int32, order="F"), plusTri=np.zeros(Ne, np.int32),
               minusTri=np.zeros(Ne, np.int32), plusSign=np.zeros(Ne, np.int32),
               minusSign=np.zeros(Ne, np.int32), rp=np.zeros((Ne, 3), order="F"),
               rm=np.zeros((Ne, 3), order="F"), Ap=np.zeros(Ne), Am=np.zeros(Ne), len=np.zeros(Ne),
               center=np.zeros((Ne, 3), order="F"), zSign=np.zeros(Ne, np.int32))


def build_rwg_interior(V, F) -> Rwg:
    V = np.asfortranarray(V, dtype=np.float64)
    F = np.asfortranarray(F, dtype=np.int32)
    Nv, Nt = V.shape[0], F.shape[0]
    ne = C.c_int()
    _check(lib().efie_rwg_count(_i(F), Nt, Nv, C.byref(ne)))
    Ne = ne.value
    r = Rwg(edge=np.zeros((Ne, 2), np.int32, order="F"), plusTri=np.zeros(Ne, np.int32),
            minusTri=np.zeros(Ne, np.int32), plusSign=np.zeros(Ne, np.int32),
            minusSign=np.zeros(Ne, np.int32), rp=np.zeros((Ne, 3), order="F"),
            rm=np.zeros((Ne, 3), order="F"), Ap=np.zeros(Ne), Am=np.zeros(Ne), len=np.zeros(Ne),
            center=np.zeros((Ne, 3), order="F"), zSign=np.zeros(Ne, np.int32))
    _check(lib().efie_build_rwg(_d(V), Nv, _i(F), Nt, Ne, _i(r.edge), _i(r.plusTri), _i(r.minusTri),
                                _i(r.plusSign), _i(r.minusSign), _d(r.rp), _d(r.rm), _d(r.Ap), _d(r.Am),
                                _d(r.len), _d(r.center), _i(r.zSign)))
    return r


def pick_port_edge_by_point_and_layer(rwg: Rwg, p, layerSign) -> int:
    p = np.ascontiguousarray(p, dtype=np.float64)
    e = C.c_int()
    _check(lib().efie_pick_port(rwg.Ne, _d(np.asfortranarray(rwg.center)), _i(rwg.zSign), _d(p),
                                int(layerSign), C.byref(e)))
    return e.value


# ---------------------------------------------------------------- device --

class Prepack(C.Structure):
    """efie_prepack: the prepack struct of efie_precompute_pack.m as plain
    arrays (MATLAB column-major layouts)."""
    _fields_ = [("Nq", C.c_int), ("Nq2", C.c_int), ("NsS", C.c_int), ("NsN", C.c_int),
                ("r_reg", C.c_void_p), ("w_reg", C.c_void_p), ("nearMat", C.c_void_p),
                ("self_r", C.c_void_p), ("self_w", C.c_void_p), ("self_a", C.c_void_p),
                ("near_r", C.c_void_p), ("near_w", C.c_void_p), ("aeqTri", C.c_void_p),
                ("near_alpha", C.c_double)]


class Context:
    """A device context holding one mesh and its RWG tables in HBM."""

    def __init__(self, device: int = 0):
        self._c = C.c_void_p()
        _check(lib().efie_create(C.byref(self._c), device))
        self._keep = []

    def close(self):
        if self._c:
            lib().efie_destroy(self._c)
            self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_mesh(self, V, F):
        V = np.asfortranarray(V, dtype=np.float64)
        F = np.asfortranarray(F, dtype=np.int32)
        _check(lib().efie_set_mesh(self._c, _d(V), V.shape[0], _i(F), F.shape[0]))
        self.Nt = F.shape[0]
        return self

    def set_rwg(self, edge, plusTri, minusTri, plusSign, minusSign, rp, rm, Ap, Am, len_):
        a = [np.asfortranarray(x, dtype=np.int32) for x in (edge, plusTri, minusTri, plusSign, minusSign)]
        b = [np.asfortranarray(x, dtype=np.float64) for x in (rp, rm, Ap, Am, len_)]
        Ne = a[1].shape[0]
        _check(lib().efie_set_rwg(self._c, *[_i(x) for x in a], *[_d(x) for x in b], Ne))
        self.Ne = Ne
        return self

    def build_rwg(self) -> Rwg:
        """build_rwg_interior.m on the device from the mesh of set_mesh; the
        tables become the context's RWG tables (efie_build_rwg_device)."""
        ne = C.c_int()
        _check(lib().efie_rwg_count_device(self._c, C.byref(ne)))
        r = _empty_rwg(ne.value)
        _check(lib().efie_build_rwg_device(self._c, ne.value, _i(r.edge), _i(r.plusTri), _i(r.minusTri),
                                           _i(r.plusSign), _i(r.minusSign), _d(r.rp), _d(r.rm), _d(r.Ap),
                                           _d(r.Am), _d(r.len), _d(r.center), _i(r.zSign)))
        self.Ne = ne.value
        return r

    def pick_port(self, p, layerSign) -> int:
        """pick_port_edge_by_point_and_layer.m on the device over the tables
        of the last build_rwg (efie_pick_port_device)."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        e = C.c_int()
        _check(lib().efie_pick_port_device(self._c, _d(p), int(layerSign), C.byref(e)))
        return e.value

    def set_problem(self, V, F, rwg: Rwg):
        self.set_mesh(V, F)
        return self.set_rwg(rwg.edge, rwg.plusTri, rwg.minusTri, rwg.plusSign, rwg.minusSign,
                            rwg.rp, rwg.rm, rwg.Ap, rwg.Am, rwg.len)

    def set_prepack(self, pk):
        """Install a prepack (dict with the fields of efie_precompute_pack.m:
        r_reg [Nq,3,Nt], w_reg [Nq,Nt], nearMat [Nt,Nt] uint8, self_r
        [Nq2,3,NsS,Nt], self_w [Nq2,NsS,Nt], self_a [NsS,Nt], near_r, near_w,
        aeqTri, near_alpha); None returns to the rule-derived tables."""
        if pk is None:
            _check(lib().efie_set_prepack(self._c, None))
            return self
        arr = {k: np.asfortranarray(pk[k], dtype=np.float64) for k in
               ("r_reg", "w_reg", "self_r", "self_w", "self_a", "near_r", "near_w", "aeqTri")}
        arr["nearMat"] = np.asfortranarray(pk["nearMat"], dtype=np.uint8)
        s = Prepack(Nq=arr["r_reg"].shape[0], Nq2=arr["self_r"].shape[0], NsS=arr["self_r"].shape[2],
                    NsN=arr["near_r"].shape[2], near_alpha=float(pk["near_alpha"]),
                    **{k: v.ctypes.data for k, v in arr.items()})
        _check(lib().efie_set_prepack(self._c, C.byref(s)))
        return self

    def set_rwg_tables(self, Nt, plusTri, minusTri, plusSign, minusSign, rp, rm, Ap, Am, len_):
        """RWG tables without the mesh arrays (the rwg struct of the 8-input
        assemble_ZV_EFIE_RWG_DIFF_GPU_FAST); a prepack must follow."""
        a = [np.asfortranarray(x, dtype=np.int32) for x in (plusTri, minusTri, plusSign, minusSign)]
        b = [np.asfortranarray(x, dtype=np.float64) for x in (rp, rm, Ap, Am, len_)]
        Ne = a[0].shape[0]
        _check(lib().efie_set_rwg_tables(self._c, int(Nt), *[_i(x) for x in a], *[_d(x) for x in b], Ne))
        self.Ne = Ne
        self.Nt = int(Nt)
        return self

    def assemble_z(self, k, omega, mu, eps, o: Optional[Opts] = None):
        k, eps = complex(k), complex(eps)
        Ne = self.Ne
        Z = np.zeros(2 * Ne * Ne)
        o = o or opts()
        _check(lib().efie_assemble_z(self._c, k.real, k.imag, omega, mu, eps.real, eps.imag, C.byref(o), _d(Z)))
        return Z.view(np.complex128).reshape((Ne, Ne), order="F")

    def assemble_zv(self, k, omega, mu, eps, exc, o: Optional[Opts] = None):
        k, eps = complex(k), complex(eps)
        exc = np.ascontiguousarray(exc, dtype=np.float64)
        Ne = self.Ne
        Z = np.zeros(2 * Ne * Ne)
        Vr = np.zeros(2 * Ne)
        o = o or opts()
        _check(lib().efie_assemble_zv(self._c, k.real, k.imag, omega, mu, eps.real, eps.imag, _d(exc),
                                      exc.shape[0], C.byref(o), _d(Z), _d(Vr)))
        return Z.view(np.complex128).reshape((Ne, Ne), order="F"), Vr.view(np.complex128).copy()

    def sweep(self, freqs, mu, eps, patchEdge, groundEdge, V0, spreadRadius, Z0,
              o: Optional[Opts] = None, stats: bool = False):
        freqs = np.ascontiguousarray(np.ravel(freqs), dtype=np.float64)
        Nf = freqs.shape[0]
        eps = complex(eps)
        zin = np.zeros(2 * Nf)
        s11 = np.zeros(2 * Nf)
        st = Stats()
        o = o or opts()
        _check(lib().efie_sweep(self._c, _d(freqs), Nf, mu, eps.real, eps.imag, int(patchEdge), int(groundEdge),
                                V0, spreadRadius, Z0, C.byref(o), _d(zin), _d(s11), C.byref(st)))
        out = (zin.view(np.complex128).copy(), s11.view(np.complex128).copy())
        return (*out, st.as_dict()) if stats else out


    def zgesv(self, A, b, precision=PREC_FP64, stats=False, noalloc=False):
        """x = A \\ b for a batch: A [batch, n, n] (or [n, n]), b [batch, n].
        noalloc: through efie_zgesv_noalloc (no device allocation; the
        library's static scratch array)."""
        A = np.asarray(A, dtype=np.complex128)
        single = A.ndim == 2
        if single:
            A = A[None]
            b = np.asarray(b)[None]
        nb, n = A.shape[0], A.shape[1]
        Af = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))  # column-major per matrix
        bf = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros((nb, n), dtype=np.complex128)
        st = Stats()
        fn = lib().efie_zgesv_noalloc if noalloc else lib().efie_zgesv
        _check(fn(self._c, n, nb, Af.view(np.float64).ctypes.data_as(_dp),
                                bf.view(np.float64).ctypes.data_as(_dp), precision,
                                x.view(np.float64).ctypes.data_as(_dp), C.byref(st)))
        x = x[0] if single else x
        return (x, st.as_dict()) if stats else x


    # ---- near-pair surrogate (include/efie_surrogate.h) ----

    def sur_train(self, o: Optional[Opts] = None, **cfg):
        """Train the near-pair surrogate for the options o and install it;
        returns the training report as a dict."""
        c = SurCfg()
        lib().efie_sur_cfg_default(C.byref(c))
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise AttributeError(k)
            setattr(c, k, v)
        rep = SurReport()
        _check(lib().efie_sur_train(self._c, C.byref(o or opts()), C.byref(c), C.byref(rep)))
        return rep.as_dict()

    def sur_get(self):
        b = np.zeros(lib().efie_sur_blob_floats(), dtype=np.float32)
        _check(lib().efie_sur_get(self._c, b.ctypes.data_as(C.POINTER(C.c_float))))
        return b

    def sur_set(self, blob):
        b = np.ascontiguousarray(blob, dtype=np.float32)
        if b.shape != (lib().efie_sur_blob_floats(),):
            raise ValueError("surrogate blob has the wrong size")
        _check(lib().efie_sur_set(self._c, b.ctypes.data_as(C.POINTER(C.c_float))))

    def sur_set_exact(self, on=True):
        _check(lib().efie_sur_set_exact(self._c, 1 if on else 0))

    def mlp_forward(self, params, X):
        p = np.ascontiguousarray(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 16), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        _check(lib().efie_mlp_forward(self._c, p.ctypes.data_as(fp), X.ctypes.data_as(fp), X.shape[0],
                                      Y.ctypes.data_as(fp)))
        return Y

    def mlp_loss_grad(self, params, X, T):
        p = np.ascontiguousarray(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(lib().efie_mlp_params(), dtype=np.float32)
        loss = C.c_float()
        fp = C.POINTER(C.c_float)
        _check(lib().efie_mlp_loss_grad(self._c, p.ctypes.data_as(fp), X.ctypes.data_as(fp), T.ctypes.data_as(fp),
                                        X.shape[0], C.byref(loss), g.ctypes.data_as(fp)))
        return loss.value, g

    @staticmethod
    def cache_counts():
        """(geometry builds, moment integrations) so far in this process,
        over all contexts (efie_diag.h)."""
        g, m = C.c_longlong(), C.c_longlong()
        _check(lib().efie_diag_cache_counts(C.byref(g), C.byref(m)))
        return g.value, m.value

    def near_pairs(self, o: Optional[Opts] = None):
        """[n, 2] int32 array of the near triangle pairs (0-based t1 < t2,
        ordered by t1 then t2) the fill of o uses (efie_diag.h)."""
        o = o or opts()
        n = C.c_longlong()
        _check(lib().efie_diag_near_pairs(self._c, C.byref(o), C.byref(n), None, 0))
        out = np.zeros((max(1, n.value), 2), dtype=np.int32)
        _check(lib().efie_diag_near_pairs(self._c, C.byref(o), C.byref(n), out.ctypes.data_as(_ip), n.value))
        return out[: n.value]

    def quad_points(self, which, o: Optional[Opts] = None):
        """The fill's quadrature points on the device (efie_diag.h):
        which = "regular" [Nt, Nq, 4], "near" / "self" [Nt, Ns, Nq2, 4] as
        (x, y, z, w) relative to the centroid (flattened per triangle),
        "self_alpha" [Nt, Ns], "radius" [Nt]."""
        sets = {"regular": 0, "near": 1, "self": 2, "self_alpha": 3, "radius": 4}
        o = o or opts()
        n = C.c_longlong()
        _check(lib().efie_diag_quad_points(self._c, C.byref(o), sets[which], C.byref(n), None, 0))
        out = np.zeros(max(1, n.value))
        _check(lib().efie_diag_quad_points(self._c, C.byref(o), sets[which], C.byref(n), _d(out), n.value))
        out = out[: n.value]
        return out.reshape(self.Nt, -1) if which != "radius" else out

    def diag_copy(self, nbytes, reps=10):
        """Average time (ms) of a plain device copy of nbytes (efie_diag.h)."""
        ms = C.c_double()
        _check(lib().efie_diag_copy(self._c, int(nbytes), int(reps), C.byref(ms)))
        return ms.value


# ------------------------------------------------- reference MEX mirrors --

def sweep_many(engine: "Context", items, mu, eps, o: Optional[Opts] = None, stats: bool = False):
    """Several sweeps in one pipelined call (efie_sweep_many): items are
    dicts with keys ctx (the Context holding the problem), freqs, patchEdge,
    groundEdge, V0, spread, Z0 and optionally release.  Returns a list of
    (Zin, S11) per item, and the call's stats when asked."""
    eps = complex(eps)
    n = len(items)
    arr = (SweepItem * max(n, 1))()
    keep = []
    outs = []
    for i, it in enumerate(items):
        fr = np.ascontiguousarray(np.ravel(it["freqs"]), dtype=np.float64)
        Nf = fr.shape[0]
        zin = np.zeros(2 * Nf)
        s11 = np.zeros(2 * Nf)
        keep += [fr, zin, s11]
        outs.append((zin, s11))
        arr[i] = SweepItem(it["ctx"]._c, _d(fr), Nf, int(it["patchEdge"]), int(it["groundEdge"]), float(it["V0"]),
                           float(it["spread"]), float(it["Z0"]), _d(zin), _d(s11), int(it.get("release", 0)))
    st = Stats()
    o = o or opts()
    _check(lib().efie_sweep_many(engine._c, arr, n, mu, eps.real, eps.imag, C.byref(o), C.byref(st)))
    res = [(z.view(np.complex128).copy(), s.view(np.complex128).copy()) for z, s in outs]
    return (res, st.as_dict()) if stats else res


def pack_plan(n, nf, precision=PREC_MIXED, batch=0, free_bytes=250e9):
    """Batches efie_sweep_many forms for items of orders n with nf
    frequencies (efie_diag_pack_plan; no device work): an int array of rows
    (batch, item, first frequency, frequencies, first matrix, batch order)."""
    n = np.ascontiguousarray(n, dtype=np.int32)
    nf = np.ascontiguousarray(nf, dtype=np.int32)
    k = C.c_int()
    _check(lib().efie_diag_pack_plan(_i(n), _i(nf), n.shape[0], precision, batch, free_bytes, None, 0,
                                     C.byref(k)))
    out = np.zeros(6 * max(k.value, 1), dtype=np.int32)
    _check(lib().efie_diag_pack_plan(_i(n), _i(nf), n.shape[0], precision, batch, free_bytes, _i(out), k.value,
                                     C.byref(k)))
    return out[:6 * k.value].reshape(-1, 6)


def patch_sweep_cuda(V, F, edge, plusTri, minusTri, plusSign, minusSign, rp, rm, Ap, Am, len_,
                     freqs, mu, eps, patchEdge, groundEdge, V0, spreadRadius, Z0,
                     o: Optional[Opts] = None, device: int = 0):
    """[Zin, S11] = patch_sweep_cuda(...) with the reference's 20 arguments
    (patch_sweep_cuda.cu:4-27); `o` is the optional extra options argument."""
    with Context(device) as ctx:
        ctx.set_mesh(V, F)
        ctx.set_rwg(edge, plusTri, minusTri, plusSign, minusSign, rp, rm, Ap, Am, len_)
        return ctx.sweep(freqs, mu, eps, patchEdge, groundEdge, V0, spreadRadius, Z0, o)


def gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST(V, F, E, plusTri, minusTri, plusSign, minusSign, rp, rm,
                                        Ap, Am, len_, k, omega, mu, eps, exc,
                                        o: Optional[Opts] = None, device: int = 0):
    """[Z, Vrhs] with the reference's 17 arguments
    (gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:127-148)."""
    with Context(device) as ctx:
        ctx.set_mesh(V, F)
        ctx.set_rwg(E, plusTri, minusTri, plusSign, minusSign, rp, rm, Ap, Am, len_)
        return ctx.assemble_zv(k, omega, mu, eps, exc, o)


def wavenumber(omega, mu, eps):
    """k = omega*sqrt(mu*eps), principal root (main_patch_efie_rwg.m:319)."""
    return omega * np.sqrt(complex(mu) * complex(eps))
