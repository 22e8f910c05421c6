"""Pins of the oracle's singularity extraction (orc_tri_potentials and pair
kind 3), a mode with no counterpart in the reference.

- the closed-form potential of an equilateral triangle at its centroid,
  int 1/R = sqrt(3) a ln(2 + sqrt(3)), and the vanishing vector potential
  there (symmetry);
- both potentials against a brute-force polar integration around the
  projection of the field point (the triangle split into three signed
  triangles with a vertex there, each integrated in polar coordinates,
  where the 1/R singularity cancels against the Jacobian; Gauss-Legendre
  panels graded towards the edge foot in angle and towards the height
  scale in radius, the radial integral of the scalar potential taken
  exactly), for points in the plane, above it, outside the triangle and
  next to an edge;
- the extracted self and near pair integrals against the same brute force
  used as the inner integral under the same outer rule (the 4-D integral
  with the exact Green's function, independent of the oracle's code), and
  their gain over the reference's softened subdivision at that rule.
"""
import numpy as np
import pytest

import workloads as W


def _gl(n, a, b):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def _brute_potentials(r, tri, kernel=None, n=20):
    """int_T K(R)/R dS' and int_T K(R) (r' - rho)/R dS' by polar quadrature
    about rho; K = 1 unless given (a function of R, vectorised)."""
    v = [np.asarray(x, float) for x in tri]
    nn = np.cross(v[1] - v[0], v[2] - v[0])
    nn /= np.linalg.norm(nn)
    h = np.dot(r - v[0], nn)
    rho = r - h * nn
    I0 = 0.0
    I1 = np.zeros(3, dtype=complex if kernel is not None else float)
    for i in range(3):
        a, b = v[i] - rho, v[(i + 1) % 3] - rho
        # signed sub-triangle (rho, a, b): orientation against nn
        sgn = np.sign(np.dot(np.cross(a, b), nn))
        if sgn == 0:
            continue
        la, lb = np.linalg.norm(a), np.linalg.norm(b)
        if la < 1e-300 or lb < 1e-300:
            continue
        ea = a / la
        e2 = np.cross(nn, ea)
        tb = np.arctan2(np.dot(b, e2), np.dot(b, ea))
        # angular breakpoints graded towards the foot of the perpendicular
        # from rho to the edge line, where the radial limit varies fastest
        e = b - a
        foot = a - np.dot(a, e) / np.dot(e, e) * e
        dist = np.linalg.norm(foot)
        tf = np.arctan2(np.dot(foot, e2), np.dot(foot, ea))
        lo, hi = min(0.0, tb), max(0.0, tb)
        bp = {lo, hi}
        step = max(dist / np.linalg.norm(e), 1e-12)
        while step < hi - lo:
            for x in (tf - step, tf + step):
                if lo < x < hi:
                    bp.add(x)
            step *= 3.0
        if lo < tf < hi:
            bp.add(tf)
        # and towards both ends, where the radial limit 1/cos grows near a
        # pole just outside the range when rho is close to the edge line
        step = max(dist / max(la, lb), 1e-12)
        while step < 0.5 * (hi - lo):
            bp.add(lo + step)
            bp.add(hi - step)
            step *= 3.0
        bp = sorted(bp)
        th = np.concatenate([_gl(n, bp[k], bp[k + 1])[0] for k in range(len(bp) - 1)])
        wth = np.concatenate([_gl(n, bp[k], bp[k + 1])[1] for k in range(len(bp) - 1)])
        if tb < 0:
            wth = -wth
        for t, wt in zip(th, wth):
            d = np.cos(t) * ea + np.sin(t) * e2
            # distance along d from rho to the edge line a-b
            e = b - a
            num = np.cross(a, e) @ nn
            den = np.cross(d, e) @ nn
            s_max = num / den
            # radial breakpoints graded towards 0 on the scale |h|, where
            # s/R turns from 0 to 1
            rb = [0.0]
            x = abs(h)
            while 0 < x < s_max:
                rb.append(x)
                x *= 3.0
            rb.append(s_max)
            s = np.concatenate([_gl(n, rb[q], rb[q + 1])[0] for q in range(len(rb) - 1)])
            ws = np.concatenate([_gl(n, rb[q], rb[q + 1])[1] for q in range(len(rb) - 1)])
            R = np.sqrt(s * s + h * h)
            if kernel is None:
                # int_0^smax s/sqrt(s^2 + h^2) ds, exactly
                I0 += wt * (np.sqrt(s_max * s_max + h * h) - abs(h))
                I1 += wt * np.sum(ws * s * s / R) * d
            else:
                K = kernel(R)
                I0 += wt * np.sum(ws * s / R * K)
                I1 += wt * np.sum(ws * s * s / R * K) * d
    # the polar sweep from a towards b is positive for sgn > 0 with tb > 0
    return I0, I1


def _equilateral(a=1e-3):
    return [np.array([0.0, 0.0, 0.0]), np.array([a, 0.0, 0.0]), np.array([a / 2, a * np.sqrt(3) / 2, 0.0])]


def test_potentials_closed_form(orc):
    a = 2e-3
    tri = _equilateral(a)
    c = sum(tri) / 3
    I0, I1 = orc.tri_potentials(c, *tri)
    assert abs(I0 - np.sqrt(3) * a * np.log(2 + np.sqrt(3))) <= 1e-14 * I0
    assert np.linalg.norm(I1) <= 1e-15 * a * I0


@pytest.mark.parametrize("where", ["inside", "above", "outside", "outside_above", "near_edge"])
def test_potentials_vs_polar_brute_force(orc, where):
    rng = np.random.default_rng(7)
    tri = [rng.uniform(-1e-3, 1e-3, 3) for _ in range(3)]
    v = np.array(tri)
    nn = np.cross(v[1] - v[0], v[2] - v[0])
    nn /= np.linalg.norm(nn)
    c = v.mean(0)
    pts = {"inside": c + 0.3 * (v[0] - c), "above": c + 0.2 * (v[1] - c) + 4e-4 * nn,
           "outside": c + 2.5 * (v[2] - c), "outside_above": c + 1.7 * (v[0] - c) - 3e-4 * nn,
           "near_edge": 0.5 * (v[0] + v[1]) + 1e-6 * (c - 0.5 * (v[0] + v[1])) + 1e-7 * nn}
    r = pts[where]
    I0, I1 = orc.tri_potentials(r, *tri)
    B0, B1 = _brute_potentials(r, tri)
    assert abs(I0 - B0) <= 1e-11 * abs(B0)
    assert np.linalg.norm(I1 - B1) <= 1e-10 * np.linalg.norm(B1)


def _problem(orc, name):
    V, F = orc.make_mesh(*W.mesh_args(name))
    r = orc.build_rwg(V, F)
    return V, F, r, orc.Problem(V, F, r)


def _exact_pair(orc, V, F, r, m, n, t1, t2, k, depth=3):
    """Outer: 7-point rule on 4^depth children of t1; inner: polar brute
    force with the exact exp(-jkR)/(4 pi R) over t2."""
    import mirror_np as M
    qp, qw = M.quad_rule(7)
    T1, T2 = V[F[t1 - 1] - 1], V[F[t2 - 1] - 1]
    P1, W1, _ = M.points_of(M.subdivide(*T1, depth), qp, qw)
    P1, W1 = P1.reshape(-1, 3), W1.ravel()

    def aff(e, t):
        e -= 1
        if t == r.plusTri[e]:
            s, A, v = r.plusSign[e], r.Ap[e], r.rp[e]
            return s * r.len[e] / (2 * A), v, s * r.len[e] / A
        s, A, v = r.minusSign[e], r.Am[e], r.rm[e]
        return -s * r.len[e] / (2 * A), v, -s * r.len[e] / A

    cm, vm, dm = aff(m, t1)
    cn, vn, dn = aff(n, t2)
    nn = np.cross(T2[1] - T2[0], T2[2] - T2[0])
    nn /= np.linalg.norm(nn)
    iv = 0j
    isc = 0j
    g = lambda R: np.exp(-1j * k * R) / (4 * np.pi)
    for p, w in zip(P1, W1):
        I0, I1 = _brute_potentials(p, T2, kernel=g, n=12)
        rho = p - np.dot(p - T2[0], nn) * nn
        inner_vec = cn * (I1 + (rho - vn) * I0)
        iv += w * np.dot(cm * (p - vm), inner_vec)
        isc += w * dm * dn * I0
    return iv, isc


@pytest.mark.parametrize("kind", ["self", "near"])
def test_extracted_pair_against_brute_force(orc, kind):
    V, F, r, P = _problem(orc, "micro")
    k = 2 * np.pi * 3e9 / W.C0
    if kind == "self":
        # a triangle carrying two different RWG functions
        edges_of = {}
        for e in range(1, r.Ne + 1):
            for t in (r.plusTri[e - 1], r.minusTri[e - 1]):
                edges_of.setdefault(int(t), []).append(e)
        t = next(t for t, es in sorted(edges_of.items()) if len(es) >= 2)
        m, n = edges_of[t][:2]
        t1 = t2 = t
    else:
        m, n = 1, next(e for e in range(2, r.Ne + 1)
                       if r.plusTri[e - 1] != r.plusTri[0] and P.near(r.plusTri[0], r.plusTri[e - 1], 1.2))
        t1, t2 = r.plusTri[m - 1], r.plusTri[n - 1]
    # the same outer rule (7 points on 4 children of the lower triangle) on
    # both sides, so the difference is the inner integral: closed form +
    # quadrature of the smooth remainder against the brute-force integral
    # of the exact kernel
    ex_v, ex_s = _exact_pair(orc, V, F, r, m, n, t1, t2, k, depth=1)
    o = orc.opts(**dict(W.OPTS_C1, self_subdiv=1, near_subdiv=1, near_order=7, extract=1))
    iv, isc = P.pair_integral(3, m, n, t1, t2, k, o)
    err = max(abs(iv - ex_v) / abs(ex_v), abs(isc - ex_s) / abs(ex_s))
    assert err <= (3e-3 if kind == "self" else 1e-6)
    # the reference's softened subdivision at the same rule is far less
    # accurate (what the extraction gains)
    o0 = orc.opts(**dict(W.OPTS_C1, self_subdiv=1, near_subdiv=1, near_order=7))
    sv, ss = P.pair_integral(2 if kind == "self" else 1, m, n, t1, t2, k, o0)
    err0 = max(abs(sv - ex_v) / abs(ex_v), abs(ss - ex_s) / abs(ex_s))
    assert err0 > 10 * err
    # the extraction is symmetric under exchange of the two sides only to
    # discretisation accuracy, so the oracle always integrates outside the
    # lower-numbered triangle (and RWG, on a self pair): swapping the
    # arguments changes nothing
    o = orc.opts(**dict(W.OPTS_C1, extract=1))
    a = P.pair_integral(3, m, n, t1, t2, k, o)
    b = P.pair_integral(3, n, m, t2, t1, k, o)
    assert a == b
