"""Physical validation on the GPU: a centre-fed strip dipole
(tests/physics_dipole.py) filled by the library (efie_assemble_z, FP64, both
the per-frequency quadrature and the series fill) with divergence-conforming
RWG signs resonates as a half-wave dipole (L = 0.46-0.49 lambda, 60-85 Ohm)
and agrees with the oracle's Zin to 1e-10; with the reference's sign
convention the same strip does not radiate (Re Zin below 1 Ohm)."""
import numpy as np
import pytest

import physics_dipole as D
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("near_mode", [0, 1])
def test_strip_dipole_on_device(efie_mod, orc, near_mode):
    V, F = D.strip_mesh()
    ref = efie_mod.build_rwg_interior(V, F)
    std = D.standard_signs(ref)
    gap, s = D.delta_gap(V, F, std)
    o = efie_mod.opts(**W.OPTS_C1, near_mode=near_mode)
    with efie_mod.Context(0) as ctx:
        out = {}
        for name, rwg in (("std", std), ("ref", ref)):
            ctx.set_problem(V, F, rwg)
            z = []
            for f in D.FREQS:
                w = 2 * np.pi * f
                z.append(D.zin(ctx.assemble_z(w * np.sqrt(W.MU0 * W.EPS0), w, W.MU0, W.EPS0, o), rwg, gap, s))
            out[name] = np.array(z)
    res = D.resonance(D.FREQS, out["std"])
    print(f"near_mode {near_mode}: standard signs resonance {res[0] / 1e9:.4f} GHz "
          f"(L/lambda {D.L_STRIP * res[0] / D.C0:.4f}), R {res[1]:.2f} Ohm; reference signs max Re Zin "
          f"{np.max(np.abs(out['ref'].real)):.3e} Ohm", flush=True)
    D.check_standard(*res)
    assert np.max(np.abs(out["ref"].real)) < 1.0
    # the oracle on the same standard-sign tables, three frequencies
    ostd = D.standard_signs(orc.build_rwg(V, F))
    P = orc.Problem(V, F, ostd)
    oo = orc.opts(**W.OPTS_C1)
    for i in (0, len(D.FREQS) // 2, len(D.FREQS) - 1):
        w = 2 * np.pi * D.FREQS[i]
        zo = D.zin(P.z_full(orc.wavenumber(w, W.MU0, W.EPS0), w, W.MU0, W.EPS0, oo), ostd, gap, s)
        assert abs(out["std"][i] - zo) / abs(zo) <= 1e-10
