"""Error bound of the device sine/cosine (efie_math.h, Cody-Waite reduction
plus Taylor polynomials), run from its host build against numpy."""
import ctypes as C
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "matlab-cuda-efief-vs-cpu_amd", "_build", "libefie_math_host.so")


def _sincos(x):
    L = C.CDLL(LIB)
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    dp = C.POINTER(C.c_double)
    L.efie_math_sincos(x.ctypes.data_as(dp), x.size, s.ctypes.data_as(dp), c.ctypes.data_as(dp))
    return s, c


def test_sincos_error_bound():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(-400, 400, 20000),
                        np.arange(-64, 65) * np.pi / 4, [0.0, 1e-300, 5e-8, np.pi / 2]])
    s, c = _sincos(x)
    # absolute error within a few ulp of 1 (|sin|, |cos| <= 1)
    # (numpy's own sin of a large argument carries its reduction error too)
    tol = 4e-16 * np.maximum(1, np.abs(x) / 100)
    assert np.all(np.abs(s - np.sin(x)) <= tol)
    assert np.all(np.abs(c - np.cos(x)) <= tol)
    # relative error where the result is not near zero
    big = np.abs(np.sin(x)) > 1e-3
    assert np.max(np.abs(s[big] / np.sin(x[big]) - 1)) <= 2e-13
    # exact special values and odd/even symmetry
    s0, c0 = _sincos(np.array([0.0, -0.3, 0.3]))
    assert s0[0] == 0.0 and c0[0] == 1.0
    assert s0[1] == -s0[2] and c0[1] == c0[2]


def test_small_argument_series():
    """The Taylor series of the residual phase (efie_math.h sincos_small on
    |t| <= 0.75, sincos_small_short on |t| <= 0.2) against numpy.  A series
    missing its last term would be off by more than 1e-14 at the end of its
    range, far outside these bounds."""
    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    rng = np.random.default_rng(1)
    for which, lim in ((0, 0.75), (1, 0.2)):
        x = np.concatenate([rng.uniform(-lim, lim, 50000), [-lim, lim, 0.0, 1e-12]])
        s = np.zeros_like(x)
        c = np.zeros_like(x)
        L.efie_math_sincos_small(x.ctypes.data_as(dp), x.size, s.ctypes.data_as(dp), c.ctypes.data_as(dp),
                                 which)
        assert np.max(np.abs(c - np.cos(x))) <= 3e-16
        nz = np.abs(x) > 0
        assert np.max(np.abs(s[nz] / np.sin(x[nz]) - 1)) <= 4e-16
