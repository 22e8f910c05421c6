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
