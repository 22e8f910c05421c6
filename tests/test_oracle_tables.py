"""Pins of the oracle's mesh, RWG, quadrature and port functions.

The reference ships no tests or fixtures, so every check here is a closed
form, a hand-derived table, or agreement with the independent numpy
restatement (oracle/mirror_np.py).
"""
import math

import numpy as np
import pytest

import workloads as W


def test_mesh_counts_closed_form(orc):
    # make_patch_with_ground_mesh.m: Nv = (Npx+1)(Npy+1)+(Ngx+1)(Ngy+1), Nt = 2(NpxNpy+NgxNgy)
    for cells in [(1, 1, 1, 1), (2, 3, 4, 5), (8, 8, 12, 12), (26, 26, 37, 37)]:
        V, F = orc.make_mesh(*W.GEOM, *cells)
        npx, npy, ngx, ngy = cells
        assert V.shape[0] == (npx + 1) * (npy + 1) + (ngx + 1) * (ngy + 1)
        assert F.shape[0] == 2 * (npx * npy + ngx * ngy)


@pytest.mark.parametrize("cells", [(1, 1, 1, 1), (2, 2, 3, 3), (3, 2, 2, 4), (8, 8, 12, 12)])
def test_mesh_and_rwg_match_numpy_mirror(orc, mirror, cells):
    V, F = orc.make_mesh(*W.GEOM, *cells)
    V2, F2 = mirror.make_mesh(*W.GEOM, *cells)
    assert np.array_equal(V, V2) and np.array_equal(F, F2)
    r = orc.build_rwg(V, F)
    r2 = mirror.build_rwg(V2, F2)
    for k in ("edge", "plusTri", "minusTri", "plusSign", "minusSign", "zSign"):
        assert np.array_equal(getattr(r, k), r2[k]), k
    for k in ("rp", "rm", "Ap", "Am", "len", "center"):
        assert np.array_equal(getattr(r, k), r2[k]), k


def test_rwg_tables_1x1_by_hand(orc):
    # Layer vertices come y-fastest: v1=(x0,y0) v2=(x0,y1) v3=(x1,y0) v4=(x1,y1).
    # Faces [1 2 4],[1 4 3] share edge {1,4}: face 1 runs 4->1 (sign -1), face 2
    # runs 1->4 (sign +1); plus is face 1 (first raw incidence).
    Lp, Wp, Lg, Wg, h = W.GEOM
    V, F = orc.make_mesh(Lp, Wp, Lg, Wg, h, 1, 1, 1, 1)
    assert np.array_equal(F, [[1, 2, 4], [1, 4, 3], [5, 6, 8], [5, 8, 7]])
    assert np.allclose(V[:4], [[-Wp / 2, -Lp / 2, h / 2], [-Wp / 2, Lp / 2, h / 2],
                               [Wp / 2, -Lp / 2, h / 2], [Wp / 2, Lp / 2, h / 2]], rtol=0, atol=0)
    r = orc.build_rwg(V, F)
    assert r.Ne == 2
    assert np.array_equal(r.edge, [[1, 4], [5, 8]])
    assert np.array_equal(r.plusTri, [1, 3]) and np.array_equal(r.minusTri, [2, 4])
    assert np.array_equal(r.plusSign, [-1, -1]) and np.array_equal(r.minusSign, [1, 1])
    assert np.array_equal(r.rp[0], V[1]) and np.array_equal(r.rm[0], V[2])
    assert np.array_equal(r.zSign, [1, -1])
    assert r.len[0] == pytest.approx(math.hypot(Wp, Lp), rel=1e-15)
    assert r.Ap[0] == pytest.approx(Wp * Lp / 2, rel=1e-15)


@pytest.mark.parametrize("cells", [(2, 2, 3, 3), (8, 8, 12, 12), (26, 26, 37, 37)])
def test_rwg_invariants(orc, cells):
    V, F = orc.make_mesh(*W.GEOM, *cells)
    r = orc.build_rwg(V, F)
    npx, npy, ngx, ngy = cells
    # interior edges of an nx x ny split-square grid: 3 nx ny - nx - ny
    assert r.Ne == (3 * npx * npy - npx - npy) + (3 * ngx * ngy - ngx - ngy)
    assert (r.plusTri < r.minusTri).all()
    # consistently oriented layers: each edge is traversed once each way
    assert np.array_equal(r.plusSign, -r.minusSign)
    assert (r.edge[:, 0] < r.edge[:, 1]).all()
    keys = r.edge[:, 0].astype(np.int64) * 10 ** 6 + r.edge[:, 1]
    assert (np.diff(keys) > 0).all()  # lexicographic, unique
    # every interior edge's two triangles contain both its vertices
    for e in range(0, r.Ne, max(1, r.Ne // 200)):
        for t in (r.plusTri[e], r.minusTri[e]):
            assert set(r.edge[e]) <= set(F[t - 1])
    # square cells: both halves of an RWG have the same area
    assert np.allclose(r.Ap, r.Am, rtol=1e-14)


def test_mesh_stride_mismatch_for_unequal_counts(orc):
    # With Np_x == Np_y the faces tile the patch; with Np_x != Np_y the
    # reference's face stride does not match the vertex order and the total
    # triangle area exceeds the patch area (overlapping triangles).
    Lp, Wp, Lg, Wg, h = W.GEOM
    def patch_area(npx, npy):
        V, F = orc.make_mesh(Lp, Wp, Lg, Wg, h, npx, npy, 1, 1)
        Fp = F[: 2 * npx * npy] - 1
        a = np.cross(V[Fp[:, 1]] - V[Fp[:, 0]], V[Fp[:, 2]] - V[Fp[:, 0]])
        return 0.5 * np.linalg.norm(a, axis=1).sum()
    assert patch_area(6, 6) == pytest.approx(Lp * Wp, rel=1e-12)
    assert patch_area(6, 4) > 1.2 * Lp * Wp


def _monomial_errors(qp, qw, d):
    return [abs(np.sum(qw * qp[:, 0] ** a * qp[:, 1] ** (d - a)) -
                math.factorial(a) * math.factorial(d - a) / math.factorial(d + 2)) for a in range(d + 1)]


@pytest.mark.parametrize("order,npts,degree", [(1, 1, 1), (2, 3, 2), (4, 6, 1), (7, 7, 5)])
def test_quadrature_exact_monomials(orc, order, npts, degree):
    """Exact integrals of u^a v^b over the reference triangle, a!b!/(a+b+2)!,
    up to the rule's degree, and a failure one degree higher."""
    qp, qw = orc.quad_rule(order)
    assert qp.shape == (npts, 2)
    # the 6-point rule's 15-digit literals (tri_quad_rule.m:13-14) sum to 0.5 - 2e-15
    tol = 1e-13 if order == 4 else 5e-15
    assert qw.sum() == pytest.approx(0.5, abs=tol)
    for d in range(degree + 1):
        assert max(_monomial_errors(qp, qw, d)) < tol * 0.1
    assert max(_monomial_errors(qp, qw, degree + 1)) > 1e-9


def test_reference_six_point_rule_second_orbit(orc):
    """tri_quad_rule.m:13-21: the second orbit parameter b equals 1-2a (to the
    15 printed digits) instead of the degree-4 value 0.0915762135, so the
    'order 4' rule is exact only to degree 1.  The oracle keeps the
    reference's literals; this pins that it does."""
    qp, qw = orc.quad_rule(4)
    a, b = qp[0, 0], qp[3, 0]
    assert abs(b - (1 - 2 * a)) < 1e-15
    # S3-symmetric orbits: both orbits have centroid (1/3, 1/3)
    for orbit in (qp[:3], qp[3:]):
        assert np.allclose(orbit.mean(axis=0), 1 / 3, rtol=0, atol=1e-15)
    # the rule's u^2 moment differs from the exact 1/12 by a fixed amount
    assert np.sum(qw * qp[:, 0] ** 2) == pytest.approx(0.08078757329498805, rel=1e-13)


def test_near_test_matches_definition(orc):
    V, F = orc.make_mesh(*W.mesh_args("c1"))
    r = orc.build_rwg(V, F)
    P = orc.Problem(V, F, r)
    tri = V[F - 1]
    cent = tri.mean(axis=1)
    area2 = np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
    rng = np.random.default_rng(0)
    for _ in range(300):
        t1, t2 = rng.integers(1, F.shape[0] + 1, 2)
        d = np.linalg.norm(cent[t1 - 1] - cent[t2 - 1])
        l = np.sqrt(4 * (0.5 * area2[[t1 - 1, t2 - 1]]) / np.sqrt(3))
        thr = 2.5 * 0.5 * l.sum()
        if abs(d - thr) > 1e-9 * thr:
            assert P.near(t1, t2, 2.5) == (d < thr)


def test_port_pick_brute_force_and_no_tie(orc):
    V, F = orc.make_mesh(*W.mesh_args("c1"))
    r = orc.build_rwg(V, F)
    pP, pG = W.feed_points("c1")
    for p, layer in ((pP, 1), (pG, -1)):
        e = orc.pick_port(r, p, layer)
        idx = np.flatnonzero(r.zSign == layer)
        d = np.linalg.norm(r.center[idx] - p, axis=1)
        assert e == idx[np.argmin(d)] + 1
        ds = np.sort(d)
        assert ds[1] - ds[0] > 1e-6 * ds[1]  # unique nearest edge, no tie


def test_rhs_normalisation(orc):
    V, F = orc.make_mesh(*W.mesh_args("c1"))
    r = orc.build_rwg(V, F)
    P = orc.Problem(V, F, r)
    pP, pG = W.feed_points("c1")
    pe, ge = orc.pick_port(r, pP, 1), orc.pick_port(r, pG, -1)
    spread = 1.5 * r.len.mean()
    v = P.rhs(pe, ge, 2.0, spread)
    assert np.all(v.imag == 0)
    assert v.real[r.zSign == 1].sum() == pytest.approx(1.0, rel=1e-14)
    assert v.real[r.zSign == -1].sum() == pytest.approx(-1.0, rel=1e-14)
    assert np.all(v.real[r.zSign == 1] > 0) and np.all(v.real[r.zSign == -1] < 0)
    wP, wG = P.port_weights(pe, ge, spread)
    assert np.allclose(wP * 1.0, v.real * (r.zSign == 1), rtol=1e-14, atol=0)
    # the port edge always carries weight exp(0) = 1, so a tiny spread puts
    # the whole excitation on the two port edges (the reference's
    # sum(w) < 1e-15 error is unreachable for a valid port edge)
    v = P.rhs(pe, ge, 1.0, 1e-12)
    assert v[pe - 1] == 0.5 and v[ge - 1] == -0.5
    assert np.count_nonzero(v) == 2


def test_rhs_gaussian_exponent(orc):
    """The excitation falls as exp(-(d/s)^2) of the distance d between edge
    centres on the port's layer (assemble_ZV_EFIE_RWG_DIFF_CPU.m:63).  The
    port edge has weight 1 before normalisation, so -log(v_e / v_port) is
    (d_e / s)^2: at spreads s and 2s it falls exactly fourfold, where
    exp(-d/s) would give a factor 2 and exp(-d^2/s) a factor 2 too."""
    V, F = orc.make_mesh(*W.mesh_args("c1"))
    r = orc.build_rwg(V, F)
    P = orc.Problem(V, F, r)
    pP, _ = W.feed_points("c1")
    pe = orc.pick_port(r, pP, 1)
    ge = orc.pick_port(r, W.feed_points("c1")[1], -1)
    s = 1.5 * r.len.mean()
    on = np.flatnonzero(r.zSign == 1)
    d = np.linalg.norm(r.center[on] - r.center[pe - 1], axis=1)
    keep = (d > 0) & (d < 2.0 * s)  # weights well above underflow at both spreads
    l1 = -np.log(P.rhs(pe, ge, 1.0, s).real[on][keep] / P.rhs(pe, ge, 1.0, s).real[pe - 1])
    l2 = -np.log(P.rhs(pe, ge, 1.0, 2 * s).real[on][keep] / P.rhs(pe, ge, 1.0, 2 * s).real[pe - 1])
    assert keep.sum() >= 10
    assert np.allclose(l1, (d[keep] / s) ** 2, rtol=1e-12, atol=0)
    assert np.allclose(l1 / l2, 4.0, rtol=1e-12, atol=0)


def test_near_pairs_numpy_matches_oracle(orc):
    """The row-wise numpy near test (tests/near_np.py, used to check the
    device's near lists at GPU-test sizes) equals the oracle's per-pair
    triangles_are_near on every pair of the c1 mesh, at both presets."""
    import near_np
    V, F = orc.make_mesh(*W.mesh_args("c1"))
    P = orc.Problem(V, F, orc.build_rwg(V, F))
    Nt = F.shape[0]
    for eps in (2.5, 1.2):
        got = near_np.near_pairs(V, F, eps)
        want = [(t1, t2) for t1 in range(Nt) for t2 in range(t1 + 1, Nt) if P.near(t1 + 1, t2 + 1, eps)]
        assert [tuple(p) for p in got] == want
