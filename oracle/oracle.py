"""ctypes front end of the CPU FP64 oracle (oracle/efie_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker that the HIP product is compared
against.  Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` may import this module; the product never does.

Arrays follow the reference's MATLAB conventions: index tables are 1-based
int32, matrices are column-major (numpy ``order="F"``), complex values are
complex128.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libefie_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcOpts(C.Structure):
    _fields_ = [
        ("quad_order", C.c_int),
        ("near_sing_eps", C.c_double),
        ("self_subdiv", C.c_int),
        ("self_alpha", C.c_double),
        ("near_subdiv", C.c_int),
        ("near_alpha", C.c_double),
        ("near_order", C.c_int),
        ("extract", C.c_int),
    ]


class OrcProblem(C.Structure):
    _fields_ = [
        ("V", _dp), ("Nv", C.c_int), ("F", _ip), ("Nt", C.c_int), ("Ne", C.c_int),
        ("plusTri", _ip), ("minusTri", _ip), ("plusSign", _ip), ("minusSign", _ip),
        ("rp", _dp), ("rm", _dp), ("Ap", _dp), ("Am", _dp), ("len", _dp),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_mesh_counts.argtypes = [C.c_int] * 4 + [_ip, _ip]
        L.orc_make_mesh.argtypes = [C.c_double] * 5 + [C.c_int] * 4 + [_dp, _ip]
        L.orc_rwg_count.argtypes = [_ip, C.c_int, C.c_int]
        L.orc_build_rwg.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int] + [_ip] * 5 + [_dp] * 6 + [_ip]
        L.orc_quad_rule.argtypes = [C.c_int, _dp, _dp]
        L.orc_triangles_are_near.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_tri_potentials.argtypes = [_dp] * 5
        L.orc_pair_integral.argtypes = [C.POINTER(OrcProblem), C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_double, C.c_double, C.POINTER(OrcOpts), _dp]
        L.orc_assemble_z_rows.argtypes = [C.POINTER(OrcProblem)] + [C.c_double] * 6 + [
            C.POINTER(OrcOpts), _ip, C.c_int, C.c_int, _dp]
        L.orc_assemble_z_centroid.argtypes = [C.POINTER(OrcProblem)] + [C.c_double] * 7 + [_dp]
        L.orc_rhs.argtypes = [C.c_int, _dp, _ip, C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_pick_port.argtypes = [C.c_int, _dp, _ip, _dp, C.c_int]
        L.orc_port_weights.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp]
        L.orc_zgetrf.argtypes = [C.c_int, _dp, _ip]
        L.orc_zgetrs.argtypes = [C.c_int, _dp, _ip, _dp]
        L.orc_sweep.argtypes = [C.POINTER(OrcProblem), _dp, _ip, _dp, C.c_int, C.c_double, C.c_double,
                                C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.POINTER(OrcOpts), C.c_int, C.c_double, C.c_int, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


# option presets of the reference drivers
ACCURATE = dict(quad_order=4, near_sing_eps=2.5, self_subdiv=2, self_alpha=0.20,
                near_subdiv=2, near_alpha=0.15, near_order=4)   # main_patch_efie_rwg.m:281-289
FAST = dict(quad_order=2, near_sing_eps=1.2, self_subdiv=2, self_alpha=0.20,
            near_subdiv=1, near_alpha=0.15, near_order=2)       # main_patch_gpu_mex_sweep_l.m:50-56


def opts(**kw) -> OrcOpts:
    d = dict(ACCURATE)
    d.update(kw)
    return OrcOpts(**d)


def make_mesh(Lp, Wp, Lg, Wg, h, npx, npy, ngx, ngy):
    Nv, Nt = C.c_int(), C.c_int()
    lib().orc_mesh_counts(npx, npy, ngx, ngy, C.byref(Nv), C.byref(Nt))
    V = np.zeros((Nv.value, 3), order="F")
    F = np.zeros((Nt.value, 3), dtype=np.int32, order="F")
    lib().orc_make_mesh(Lp, Wp, Lg, Wg, h, npx, npy, ngx, ngy, _d(V), _i(F))
    return V, F


@dataclass
class Rwg:
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
    def Ne(self):
        return int(self.plusTri.shape[0])


def build_rwg(V, F) -> Rwg:
    V = np.asfortranarray(V, dtype=np.float64)
    F = np.asfortranarray(F, dtype=np.int32)
    Nv, Nt = V.shape[0], F.shape[0]
    Ne = lib().orc_rwg_count(_i(F), Nt, Nv)
    if Ne < 0:
        raise ValueError("F has invalid indices")
    r = Rwg(edge=np.zeros((Ne, 2), np.int32, order="F"),
            plusTri=np.zeros(Ne, np.int32), minusTri=np.zeros(Ne, np.int32),
            plusSign=np.zeros(Ne, np.int32), minusSign=np.zeros(Ne, np.int32),
            rp=np.zeros((Ne, 3), order="F"), rm=np.zeros((Ne, 3), order="F"),
            Ap=np.zeros(Ne), Am=np.zeros(Ne), len=np.zeros(Ne),
            center=np.zeros((Ne, 3), order="F"), zSign=np.zeros(Ne, np.int32))
    rc = lib().orc_build_rwg(_d(V), Nv, _i(F), Nt, Ne, _i(r.edge), _i(r.plusTri), _i(r.minusTri),
                             _i(r.plusSign), _i(r.minusSign), _d(r.rp), _d(r.rm), _d(r.Ap),
                             _d(r.Am), _d(r.len), _d(r.center), _i(r.zSign))
    if rc:
        raise ValueError("opposite vertex not found")
    return r


def quad_rule(order):
    qp = np.zeros(32)
    qw = np.zeros(16)
    n = lib().orc_quad_rule(order, _d(qp), _d(qw))
    return qp[: 2 * n].reshape(2, n).T.copy(), qw[:n].copy()


class Problem:
    """Keeps the arrays alive behind an OrcProblem struct."""

    def __init__(self, V, F, rwg: Rwg):
        self.V = np.asfortranarray(V, dtype=np.float64)
        self.F = np.asfortranarray(F, dtype=np.int32)
        self.rwg = rwg
        self.s = OrcProblem(_d(self.V), self.V.shape[0], _i(self.F), self.F.shape[0], rwg.Ne,
                            _i(rwg.plusTri), _i(rwg.minusTri), _i(rwg.plusSign), _i(rwg.minusSign),
                            _d(rwg.rp), _d(rwg.rm), _d(rwg.Ap), _d(rwg.Am), _d(rwg.len))

    @property
    def Ne(self):
        return self.rwg.Ne

    def near(self, t1, t2, eps):
        return bool(lib().orc_triangles_are_near(_d(self.V), self.V.shape[0], _i(self.F),
                                                 self.F.shape[0], t1, t2, eps))

    def pair_integral(self, kind, m, n, t1, t2, k, o: OrcOpts):
        out = np.zeros(4)
        rc = lib().orc_pair_integral(C.byref(self.s), kind, m, n, t1, t2, complex(k).real,
                                     complex(k).imag, C.byref(o), _d(out))
        if rc:
            raise ValueError("pair integral called with the wrong triangle relation")
        return complex(out[0], out[1]), complex(out[2], out[3])

    def z_rows(self, rows, k, omega, mu, eps, o: OrcOpts, threads=None):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        nr = rows.shape[0]
        Z = np.zeros(2 * nr * self.Ne)
        th = threads or os.cpu_count() or 1
        eps = complex(eps)
        rc = lib().orc_assemble_z_rows(C.byref(self.s), complex(k).real, complex(k).imag, omega, mu,
                                       eps.real, eps.imag, C.byref(o), _i(rows), nr, th, _d(Z))
        if rc:
            raise ValueError("row index out of range")
        return Z.view(np.complex128).reshape((nr, self.Ne), order="F")

    def z_full(self, k, omega, mu, eps, o: OrcOpts, threads=None):
        return self.z_rows(np.arange(1, self.Ne + 1), k, omega, mu, eps, o, threads)

    def z_centroid(self, k, alpha, beta, self_alpha=0.20):
        Z = np.zeros(2 * self.Ne * self.Ne)
        k, alpha, beta = complex(k), complex(alpha), complex(beta)
        lib().orc_assemble_z_centroid(C.byref(self.s), k.real, k.imag, alpha.real, alpha.imag,
                                      beta.real, beta.imag, self_alpha, _d(Z))
        return Z.view(np.complex128).reshape((self.Ne, self.Ne), order="F")

    def rhs(self, patchEdge, groundEdge, V0, spread):
        out = np.zeros(2 * self.Ne)
        rc = lib().orc_rhs(self.Ne, _d(self.rwg.center), _i(self.rwg.zSign), patchEdge, groundEdge,
                           V0, spread, _d(out))
        if rc:
            raise ValueError("spreadRadius too small" if rc == -1 else "port edge out of range")
        return out.view(np.complex128).copy()

    def port_weights(self, patchEdge, groundEdge, spread):
        wP = np.zeros(self.Ne)
        wG = np.zeros(self.Ne)
        lib().orc_port_weights(_d(self.V), self.V.shape[0], _i(self.rwg.edge), self.Ne, patchEdge,
                               groundEdge, spread, _d(wP), _d(wG))
        return wP, wG

    def sweep(self, freqs, mu, eps, patchEdge, groundEdge, V0, spread, Z0, o: OrcOpts,
              centroid=False, self_alpha=0.20, threads=None):
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        Nf = freqs.shape[0]
        zin = np.zeros(2 * Nf)
        s11 = np.zeros(2 * Nf)
        eps = complex(eps)
        th = threads or os.cpu_count() or 1
        rc = lib().orc_sweep(C.byref(self.s), _d(self.rwg.center), _i(self.rwg.zSign), _d(freqs), Nf,
                             mu, eps.real, eps.imag, patchEdge, groundEdge, V0, spread, Z0,
                             C.byref(o), 1 if centroid else 0, self_alpha, th, _d(zin), _d(s11))
        if rc:
            raise RuntimeError(f"oracle sweep failed ({rc})")
        return zin.view(np.complex128).copy(), s11.view(np.complex128).copy()


def tri_potentials(r, v1, v2, v3):
    """(int_T 1/R dS', int_T (r' - rho)/R dS') in closed form (orc_tri_potentials)."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (r, v1, v2, v3)]
    out = np.zeros(4)
    lib().orc_tri_potentials(*[_d(x) for x in a], _d(out))
    return out[0], out[1:].copy()


def pick_port(rwg: Rwg, p, layer):
    p = np.ascontiguousarray(p, dtype=np.float64)
    return lib().orc_pick_port(rwg.Ne, _d(rwg.center), _i(rwg.zSign), _d(p), int(layer))


def zgetrf(A):
    A = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    n = A.shape[0]
    ipiv = np.zeros(n, np.int32)
    info = lib().orc_zgetrf(n, A.ctypes.data_as(_dp), _i(ipiv))
    return A, ipiv, info


def zgetrs(LU, ipiv, b):
    LU = np.asfortranarray(LU, dtype=np.complex128)
    x = np.ascontiguousarray(b, dtype=np.complex128).copy()
    lib().orc_zgetrs(LU.shape[0], LU.ctypes.data_as(_dp), _i(np.ascontiguousarray(ipiv, np.int32)),
                     x.ctypes.data_as(_dp))
    return x


# physics constants (main_patch_efie_rwg.m:218-220)
C0 = 299792458.0
MU0 = 4 * np.pi * 1e-7
EPS0 = 1.0 / (MU0 * C0 ** 2)


def wavenumber(omega, mu, eps):
    """k = omega*sqrt(mu*eps), principal root (main_patch_efie_rwg.m:319)."""
    return omega * np.sqrt(complex(mu) * complex(eps))
