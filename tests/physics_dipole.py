"""A centre-fed strip dipole: the physical check of the RWG sign
convention (shared by the CPU oracle test and the GPU test).

A thin PEC strip (length L = 150 mm, width 2 mm, 50 x 1 cells, split-square
triangles oriented as make_patch_with_ground_mesh.m orients them) is fed by
a delta gap on its middle transverse edge: V_m = s_m V0 l_m on that edge
(s_m = +1 when its plus triangle lies on the x < L/2 side), Z x = V,
I = sum s_m x_m l_m, Zin = V0 / I.  A thin half-wave dipole resonates
(Im Zin = 0) at L = 0.47-0.48 lambda with a radiation resistance of about
73 Ohm (textbook: Balanis, Antenna Theory, ch. 4 and 8; a strip of width w
behaves as a wire of radius w/4).  Divergence-conforming RWG signs
(plusSign = minusSign = +1) must reproduce that; the reference's convention
(each half signed by its own triangle's traversal, build_rwg_interior.m:
50-63, so every basis function carries a net charge) radiates nothing.
"""
import numpy as np

L_STRIP, W_STRIP, NX, NY = 0.15, 0.002, 50, 1
C0 = 299792458.0
FREQS = np.arange(0.90e9, 1.0001e9, 0.01e9)


def strip_mesh():
    """V [Nv x 3], F [Nt x 3] (1-based int32) of the strip in z = 0."""
    xs = np.linspace(0.0, L_STRIP, NX + 1)
    ys = np.linspace(-W_STRIP / 2, W_STRIP / 2, NY + 1)
    V = np.array([[x, y, 0.0] for y in ys for x in xs])
    F = []
    for j in range(NY):
        for i in range(NX):
            v1 = j * (NX + 1) + i + 1
            v2, v3, v4 = v1 + 1, v1 + NX + 1, v1 + NX + 2
            F.append([v1, v2, v4])
            F.append([v1, v4, v3])
    return np.asfortranarray(V), np.asfortranarray(np.array(F, dtype=np.int32))


def standard_signs(rwg):
    """A copy of the RWG tables with divergence-conforming signs."""
    import copy
    r = copy.deepcopy(rwg)
    r.plusSign[:] = 1
    r.minusSign[:] = 1
    return r


def delta_gap(V, F, rwg):
    """Gap edges (0-based) and their orientation signs s_m."""
    gap = np.where(np.abs(rwg.center[:, 0] - L_STRIP / 2) < 1e-9)[0]
    cx = np.array([V[F[rwg.plusTri[m] - 1] - 1, 0].mean() for m in gap])
    return gap, np.where(cx < L_STRIP / 2, 1.0, -1.0)


def zin(Z, rwg, gap, s, V0=1.0):
    Vv = np.zeros(rwg.Ne, complex)
    Vv[gap] = s * V0 * rwg.len[gap]
    x = np.linalg.solve(Z, Vv)
    return V0 / np.sum(s * x[gap] * rwg.len[gap])


def resonance(freqs, z):
    """First frequency where Im Zin goes from negative to positive (linear
    interpolation) and Re Zin there; None if there is none."""
    im = z.imag
    for i in range(len(freqs) - 1):
        if im[i] < 0 <= im[i + 1]:
            t = -im[i] / (im[i + 1] - im[i])
            return freqs[i] + t * (freqs[i + 1] - freqs[i]), z.real[i] + t * (z.real[i + 1] - z.real[i])
    return None


def check_standard(f0, r0):
    """The textbook half-wave strip dipole."""
    assert 0.46 <= L_STRIP * f0 / C0 <= 0.49, L_STRIP * f0 / C0
    assert 60.0 <= r0 <= 85.0, r0
