"""GPU: the frequency-independent geometry built on the device.

The quadrature points of the fill (regular rule per triangle, the
subdivided rule on each midpoint child for near and self pairs, relative
to the triangle's centroid, weights times twice the area) are generated by
a kernel from the vertices.  They must equal, bit for bit, the numpy
restatement of map_ref_to_vertices.m and subdivide_triangles.m in
oracle/mirror_np.py (same operations in the same order, no fused
multiply-add on either side).
"""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

RULESETS = [
    dict(),  # the accurate preset: 7-point regular, 6-point on 16 children
    dict(quad_order=2, near_order=2, near_subdiv=1, self_subdiv=1),
    dict(quad_order=1, near_order=1, near_subdiv=3, self_subdiv=3, self_alpha=0.3),
    dict(quad_order=6, near_order=7, near_subdiv=0, self_subdiv=2),
]


@pytest.fixture(scope="module")
def ctx(efie_mod):
    c = efie_mod.Context(0)
    yield c
    c.close()


def _expected(mirror, V, F, o):
    Vt = V[F - 1]  # [Nt, 3 vertices, 3]
    A, B, C = Vt[:, 0], Vt[:, 1], Vt[:, 2]
    cent = (A + B + C) / 3.0
    area2 = np.linalg.norm(np.cross(B - A, C - A), axis=1)
    qp, qw = mirror.quad_rule(o["quad_order"])
    reg = A[:, None, :] + qp[None, :, 0:1] * (B - A)[:, None, :] + qp[None, :, 1:2] * (C - A)[:, None, :]
    reg = np.concatenate([reg - cent[:, None, :], (qw[None, :] * area2[:, None])[..., None]], axis=2)
    sp, sw = mirror.quad_rule(o["near_order"])
    out = {"regular": reg}
    rad = np.max(np.linalg.norm(Vt - cent[:, None, :], axis=2), axis=1)
    rad = np.maximum(rad, np.max(np.linalg.norm(reg[..., :3], axis=2), axis=1))
    for kind, depth in (("near", o["near_subdiv"]), ("self", o["self_subdiv"])):
        pts, alpha = [], []
        for t in range(len(F)):
            p, w, a2p = mirror.points_of(mirror.subdivide(A[t], B[t], C[t], depth), sp, sw)
            pts.append(np.concatenate([p - cent[t], w[..., None]], axis=2))
            alpha.append(o["self_alpha"] * np.sqrt((0.5 * a2p) / np.pi))
        pts = np.array(pts)  # [Nt, Ns, Nq2, 4]
        out[kind] = pts
        rad = np.maximum(rad, np.max(np.linalg.norm(pts[..., :3], axis=3), axis=(1, 2)))
        if kind == "self":
            out["self_alpha"] = np.array(alpha)
    out["radius"] = rad
    return out


@pytest.mark.parametrize("rs", range(len(RULESETS)))
def test_device_quad_points_bit_exact(efie_mod, mirror, ctx, rs):
    V, F = efie_mod.make_patch_with_ground_mesh(*W.mesh_args("c1"))
    r = efie_mod.build_rwg_interior(V, F)
    ctx.set_problem(V, F, r)
    o = dict(W.OPTS_C1)
    o.update(RULESETS[rs])
    exp = _expected(mirror, np.asarray(V), np.asarray(F), o)
    eo = efie_mod.opts(**o)
    Nt = len(F)
    for which in ("regular", "near", "self", "self_alpha", "radius"):
        got = ctx.quad_points(which, eo)
        want = exp[which].reshape(Nt, -1) if which != "radius" else exp[which]
        assert got.shape == want.shape, which
        assert np.array_equal(got, want), (which, np.max(np.abs(got - want)))


RWG_FIELDS = ("edge", "plusTri", "minusTri", "plusSign", "minusSign", "rp", "rm", "Ap", "Am", "len", "center",
              "zSign")


def _mesh(efie_mod, name):
    if name.startswith("c4_"):
        g = W.c4_geometries()[int(name[3:])]
        return efie_mod.make_patch_with_ground_mesh(*g["geom"], *g["cells"])
    if name.startswith("cells"):
        return efie_mod.make_patch_with_ground_mesh(*W.GEOM, *[int(x) for x in name[5:].split("_")])
    return efie_mod.make_patch_with_ground_mesh(*W.mesh_args(name))


@pytest.mark.parametrize("name", ["micro", "tiny", "c1", "c2_6k", "c4_19", "c4_11", "cells6_4_5_5", "cells3_5_4_2",
                                  "cells1_1_1_1"])
def test_device_rwg_tables_equal_host(efie_mod, ctx, name):
    """build_rwg_interior.m on the device (efie_build_rwg_device) against
    the host builder, every table bit for bit, on the workload meshes and on
    meshes with unequal cell counts (whose faces overlap); then the port
    pick against efie_pick_port on both layers."""
    V, F = _mesh(efie_mod, name)
    r = efie_mod.build_rwg_interior(V, F)
    ctx.set_mesh(V, F)
    rd = ctx.build_rwg()
    for f in RWG_FIELDS:
        assert np.array_equal(getattr(rd, f), getattr(r, f)), f
    V = np.asarray(V)
    rng = np.random.default_rng(len(F))
    pts = [V.min(axis=0), V.max(axis=0), V.mean(axis=0)] + list(rng.uniform(V.min(axis=0), V.max(axis=0), (8, 3)))
    pts += [r.center[0], r.center[r.Ne // 2]]  # distance zero to an edge centre
    for p in pts:
        for layer in (1, -1):
            assert ctx.pick_port(p, layer) == efie_mod.pick_port_edge_by_point_and_layer(r, p, layer)
    with pytest.raises(efie_mod.EfieError) as e:
        ctx.pick_port(pts[0], 0)
    assert e.value.ident == "efie:port"


def test_device_rwg_installs_tables(efie_mod, ctx):
    """The device-built tables drive the sweep exactly as the host-built ones
    given through set_problem."""
    V, F = _mesh(efie_mod, "c1")
    ctx.set_mesh(V, F)
    rd = ctx.build_rwg()
    pP, pG = W.feed_points("c1")
    pe, ge = ctx.pick_port(pP, 1), ctx.pick_port(pG, -1)
    fr = W.freqs("c1")[::4]
    o = efie_mod.opts(**W.OPTS_C1, precision=0, near_mode=efie_mod.NEAR_MOMENTS)
    sp = 1.5 * rd.len.mean()
    z1 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, sp, 50.0, o)[0]
    r = efie_mod.build_rwg_interior(V, F)
    assert (pe, ge) == (efie_mod.pick_port_edge_by_point_and_layer(r, pP, 1),
                        efie_mod.pick_port_edge_by_point_and_layer(r, pG, -1))
    ctx.set_problem(V, F, r)
    z2 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, sp, 50.0, o)[0]
    assert np.array_equal(z1, z2)


def test_device_rwg_errors_as_host(efie_mod, ctx):
    """A triangle (1, 2, 1) makes edge (1, 2) interior with no opposite
    vertex: both builders refuse the mesh with the same error."""
    V = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]])
    F = np.array([[1, 2, 1], [2, 4, 3]], dtype=np.int32)
    with pytest.raises(efie_mod.EfieError) as eh:
        efie_mod.build_rwg_interior(V, F)
    ctx.set_mesh(V, F)
    with pytest.raises(efie_mod.EfieError) as ed:
        ctx.build_rwg()
    assert ed.value.ident == eh.value.ident and str(ed.value) == str(eh.value)
    with pytest.raises(efie_mod.EfieError) as ep:
        ctx.pick_port([0.0, 0.0, 0.0], 1)
    assert ep.value.ident == "efie:state"
