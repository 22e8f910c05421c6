"""Physical validation of the oracle's integrals and of the RWG sign
option on a centre-fed strip dipole (tests/physics_dipole.py): with
divergence-conforming signs the oracle's Z gives the half-wave resonance at
L = 0.46-0.49 lambda with 60-85 Ohm; the reference's sign convention gives
a strip that does not radiate (Re Zin below 1 Ohm across the band)."""
import numpy as np

import physics_dipole as D  # noqa: E402
import workloads as W


def _sweep(orc, P, rwg, gap, s):
    o = orc.opts(**W.OPTS_C1)
    out = []
    for f in D.FREQS:
        w = 2 * np.pi * f
        Z = P.z_full(orc.wavenumber(w, W.MU0, W.EPS0), w, W.MU0, W.EPS0, o)
        out.append(D.zin(Z, rwg, gap, s))
    return np.array(out)


def test_strip_dipole_oracle(orc):
    V, F = D.strip_mesh()
    ref = orc.build_rwg(V, F)
    assert np.all(ref.plusSign == -ref.minusSign)
    std = D.standard_signs(ref)
    gap, s = D.delta_gap(V, F, std)
    assert len(gap) == D.NY
    z_std = _sweep(orc, orc.Problem(V, F, std), std, gap, s)
    res = D.resonance(D.FREQS, z_std)
    assert res is not None
    D.check_standard(*res)
    z_ref = _sweep(orc, orc.Problem(V, F, ref), ref, gap, s)
    assert np.max(np.abs(z_ref.real)) < 1.0
