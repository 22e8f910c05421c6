"""Pins of the oracle's softened near and self integrals and of its centroid
model, by quantities the mathematics fixes independently of the oracle code.

The oracle runs with the 7-point rule here: the reference's 6-point rule
(tri_quad_rule.m:13-21) integrates only linear functions exactly, which
would hide the quantity under test behind its own error.

- near pairs (rwg_pair_integrals_near_singular.m:16-19,41-42): with a very
  large softening the integrand is smooth, so a brute-force product rule on a
  finer subdivision converges to the same value; the softening length
  a = near_alpha * min(sqrt(A1/pi), sqrt(A2/pi)) is what that check fixes
  (max instead of min, or area2 instead of the area, fail it);
- self pairs (rwg_pair_integrals_singular.m:28-31,46-50): only the blocks of
  a sub-triangle with itself depend on self_alpha, so the difference of two
  self integrals at two softenings is a sum over those blocks alone, with
  a = self_alpha * sqrt(A_sub/pi) and A_sub = A / 4^self_subdiv;
- centroid model (patch_sweep_cuda.cu:94-153): it equals the quadrature
  scheme with the 1-point rule, no near pairs and an undivided, softened
  self term.
"""
import numpy as np

import workloads as W


def _problem(orc, name):
    V, F = orc.make_mesh(*W.mesh_args(name))
    r = orc.build_rwg(V, F)
    return V, F, r, orc.Problem(V, F, r)


def _f(r, e, t, p):
    """RWG function and divergence of edge e (1-based) on triangle t."""
    e -= 1
    if t == r.plusTri[e]:
        s, A, v = r.plusSign[e], r.Ap[e], r.rp[e]
        return s * r.len[e] / (2 * A) * (p - v), s * r.len[e] / A
    s, A, v = r.minusSign[e], r.Am[e], r.rm[e]
    return s * r.len[e] / (2 * A) * (v - p), -s * r.len[e] / A


def _points(T, depth, order=7):
    import mirror_np as M
    qp, qw = M.quad_rule(order)
    P, Wt, _ = M.points_of(M.subdivide(*T, depth), qp, qw)
    return P.reshape(-1, 3), Wt.ravel()


def _soft_pair(r, m, n, t1, t2, T1, T2, k, a, depth):
    P1, W1 = _points(T1, depth)
    P2, W2 = _points(T2, depth)
    fm, dm = _f(r, m, t1, P1)
    fn, dn = _f(r, n, t2, P2)
    R = np.sqrt(np.sum((P1[:, None] - P2[None]) ** 2, axis=-1) + a * a)
    G = np.exp(-1j * k * R) / (4 * np.pi * R)
    Wt = W1[:, None] * W2[None]
    return np.sum(Wt * (fm @ fn.T) * G), np.sum(Wt * dm * dn * G)


def test_near_softening_length(orc):
    V, F, r, P = _problem(orc, "c1")
    k = 2 * np.pi * 2e9 / W.C0
    # a patch triangle and a ground triangle of different areas that are near
    cen = (V[F[:, 0] - 1] + V[F[:, 1] - 1] + V[F[:, 2] - 1]) / 3
    m = next(e for e in range(1, r.Ne + 1) if cen[r.plusTri[e - 1] - 1, 2] > 0)
    t1 = r.plusTri[m - 1]
    n = next(e for e in range(1, r.Ne + 1)
             if cen[r.plusTri[e - 1] - 1, 2] < 0 and P.near(t1, r.plusTri[e - 1], 2.5))
    t2 = r.plusTri[n - 1]
    T1, T2 = V[F[t1 - 1] - 1], V[F[t2 - 1] - 1]
    A1, A2 = r.Ap[m - 1], r.Ap[n - 1]
    assert abs(A1 - A2) > 0.2 * A1
    alpha = 1e3
    o = orc.opts(**dict(W.OPTS_C1, near_alpha=alpha, near_subdiv=1, near_order=7))
    iv, isc = P.pair_integral(1, m, n, t1, t2, k, o)
    a = alpha * min(np.sqrt(A1 / np.pi), np.sqrt(A2 / np.pi))
    ev, es = _soft_pair(r, m, n, t1, t2, T1, T2, k, a, depth=3)
    assert abs(iv - ev) <= 1e-9 * abs(ev) and abs(isc - es) <= 1e-9 * abs(es)
    # the other plausible softening length is far off
    amax = alpha * max(np.sqrt(A1 / np.pi), np.sqrt(A2 / np.pi))
    bv, _ = _soft_pair(r, m, n, t1, t2, T1, T2, k, amax, depth=1)
    assert abs(iv - bv) > 1e-3 * abs(bv)


def test_self_softening_only_on_coincident_sub_triangles(orc):
    import mirror_np as M
    V, F, r, P = _problem(orc, "c1")
    k = 2 * np.pi * 2.5e9 / W.C0
    t = r.plusTri[10]
    m = 11
    n = next(e for e in range(1, r.Ne + 1) if e != m and t in (r.plusTri[e - 1], r.minusTri[e - 1]))
    T = V[F[t - 1] - 1]
    A = r.Ap[m - 1]
    depth = 2
    res = []
    for alpha in (1e3, 3e3):
        o = orc.opts(**dict(W.OPTS_C1, self_alpha=alpha, self_subdiv=depth, near_order=7))
        res.append(np.array(P.pair_integral(2, m, n, t, t, k, o)))
    d = res[0] - res[1]
    # brute force over the blocks of each sub-triangle with itself
    subs = M.subdivide(*T, depth)
    asub = np.sqrt(A / 4 ** depth / np.pi)
    exp = np.zeros(2, complex)
    for S in subs:
        for alpha, sgn in ((1e3, 1), (3e3, -1)):
            exp += sgn * np.array(_soft_pair(r, m, n, t, t, S, S, k, alpha * asub, depth=1))
    assert np.all(np.abs(d - exp) <= 1e-9 * np.abs(exp))
    # softening with the whole triangle's area would be far off
    wrong = np.zeros(2, complex)
    for S in subs:
        for alpha, sgn in ((1e3, 1), (3e3, -1)):
            wrong += sgn * np.array(_soft_pair(r, m, n, t, t, S, S, k, alpha * np.sqrt(A / np.pi), depth=0))
    assert np.all(np.abs(d - wrong) > 1e-2 * np.abs(wrong))


def test_centroid_model_is_the_one_point_scheme(orc):
    V, F, r, P = _problem(orc, "tiny")
    w = 2 * np.pi * 2.2e9
    eps = W.EPS0 * (1 - 1e-3j)
    k = orc.wavenumber(w, W.MU0, eps)
    for sa in (0.2, 0.35):
        Zc = P.z_centroid(k, 1j * w * W.MU0, 1 / (1j * w * eps), sa)
        o = orc.opts(quad_order=1, near_sing_eps=0.0, self_subdiv=0, self_alpha=sa, near_subdiv=1,
                     near_alpha=0.15, near_order=1)
        Zq = P.z_full(k, w, W.MU0, eps, o)
        assert np.max(np.abs(Zc - Zq) / np.abs(Zq)) <= 1e-12
