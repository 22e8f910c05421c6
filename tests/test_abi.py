"""The C ABI library: loads, exports every declared symbol, and its host-side
table builders (mesh, RWG, port) agree bit-for-bit with the oracle.
No device compute here."""
import ctypes
import os
import re

import numpy as np
import pytest

import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(efie_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(efie_mod):
    L = ctypes.CDLL(efie_mod.LIB_PATH)
    names = _declared("efie.h") + _declared("efie_diag.h") + _declared("efie_surrogate.h")
    assert len(names) >= 17
    for n in names:
        assert hasattr(L, n), n
    for n in efie_mod.EXPORTS:
        assert n in names, n


@pytest.mark.parametrize("cells", [(1, 1, 1, 1), (3, 2, 2, 4), (8, 8, 12, 12), (26, 26, 37, 37)])
def test_host_tables_bit_exact_vs_oracle(efie_mod, orc, cells):
    V, F = efie_mod.make_patch_with_ground_mesh(*W.GEOM, *cells)
    Vo, Fo = orc.make_mesh(*W.GEOM, *cells)
    assert np.array_equal(V, Vo) and np.array_equal(F, Fo)
    r = efie_mod.build_rwg_interior(V, F)
    ro = orc.build_rwg(Vo, Fo)
    for k in ("edge", "plusTri", "minusTri", "plusSign", "minusSign", "zSign", "rp", "rm", "Ap", "Am",
              "len", "center"):
        assert np.array_equal(getattr(r, k), getattr(ro, k)), k
    for p, layer in zip(W.feed_points("c1"), (1, -1)):
        assert efie_mod.pick_port_edge_by_point_and_layer(r, p, layer) == orc.pick_port(ro, p, layer)


def test_host_errors(efie_mod):
    with pytest.raises(efie_mod.EfieError) as e:
        efie_mod.make_patch_with_ground_mesh(*W.GEOM, 0, 1, 1, 1)
    assert e.value.ident == "efie:mesh"
    V, F = efie_mod.make_patch_with_ground_mesh(*W.GEOM, 2, 2, 2, 2)
    Fbad = F.copy()
    Fbad[0, 0] = V.shape[0] + 1
    with pytest.raises(efie_mod.EfieError) as e:
        efie_mod.build_rwg_interior(V, Fbad)
    assert e.value.ident == "mex:F"
    r = efie_mod.build_rwg_interior(V, F)
    with pytest.raises(efie_mod.EfieError) as e:
        efie_mod.pick_port_edge_by_point_and_layer(r, [0, 0, 0], 7)
    assert e.value.ident == "efie:port"


def test_presets(efie_mod):
    a = efie_mod.opts("accurate")
    assert (a.quad_order, a.near_sing_eps, a.self_subdiv, a.self_alpha, a.near_subdiv, a.near_alpha,
            a.near_order) == (4, 2.5, 2, 0.2, 2, 0.15, 4)
    f = efie_mod.opts("fast")
    assert (f.quad_order, f.near_sing_eps, f.near_subdiv, f.near_order) == (2, 1.2, 1, 2)
    assert efie_mod.opts("centroid").model == efie_mod.MODEL_CENTROID
    with pytest.raises(efie_mod.EfieError):
        efie_mod.opts("nope")
