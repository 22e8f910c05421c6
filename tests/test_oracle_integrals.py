"""Pins of the oracle's pair integrals, assembly, solve and Zin.

- agreement with the independent numpy restatement on whole matrices,
- convergence of the regular rule to a brute-force fine product rule for
  well-separated triangle pairs (catches a dropped term, a wrong sign or a
  transposed operand in the integrand),
- the near-singular scheme tends to the regular integral as the softening
  goes to zero on separated pairs,
- Z = Z^T, the static (k -> 0) scaling, LAPACK agreement of the LU.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import workloads as W

OPTSETS = [
    dict(quad_order=4, near_sing_eps=1.2, near_subdiv=1, near_order=2),
    dict(quad_order=7, near_sing_eps=1.2, near_subdiv=2, near_order=4),
    dict(quad_order=2, near_sing_eps=2.5, near_subdiv=2, near_order=4),
    dict(quad_order=1, near_sing_eps=0.5, near_subdiv=1, near_order=1, self_subdiv=1),
]


def _problem(orc, name):
    V, F = orc.make_mesh(*W.mesh_args(name))
    r = orc.build_rwg(V, F)
    return V, F, r, orc.Problem(V, F, r)


@pytest.mark.parametrize("ok", range(len(OPTSETS)))
def test_full_z_matches_numpy_mirror(orc, mirror, ok):
    V, F, r, P = _problem(orc, "micro")
    o = dict(W.OPTS_C1)
    o.update(OPTSETS[ok])
    f = 2.5e9
    w = 2 * np.pi * f
    eps = W.EPS0 * (2.2 - 0.01j)  # lossy: exercises complex k
    k = orc.wavenumber(w, W.MU0, eps)
    Z = P.z_full(k, w, W.MU0, eps, orc.opts(**o))
    Zm = mirror.assemble_z(V, F, mirror.build_rwg(V, F), k, w, W.MU0, eps, o)
    assert np.abs(Z - Zm).max() <= 2e-15 * np.abs(Z).max()
    assert np.abs(Z - Z.T).max() <= 2e-15 * np.abs(Z).max()


def test_dispatch_covers_all_three_kinds(orc):
    # the tiny mesh at eps 1.2 has self, near and regular pairs
    V, F, r, P = _problem(orc, "tiny")
    kinds = set()
    for t1 in range(1, F.shape[0] + 1):
        for t2 in range(1, F.shape[0] + 1):
            kinds.add(2 if t1 == t2 else (1 if P.near(t1, t2, 1.2) else 0))
    assert kinds == {0, 1, 2}


def _fine_pair_integral(V, F, r, m, n, t1, t2, k, depth=4):
    """Brute force: both triangles split 4^depth times, 7-point rule on
    every child, plain Green's function.  Independent of the oracle's code."""
    import mirror_np as M
    qp, qw = M.quad_rule(7)
    T1 = V[F[t1 - 1] - 1]
    T2 = V[F[t2 - 1] - 1]
    P1, W1, _ = M.points_of(M.subdivide(*T1, depth), qp, qw)
    P2, W2, _ = M.points_of(M.subdivide(*T2, depth), qp, qw)
    P1, P2, W1, W2 = P1.reshape(-1, 3), P2.reshape(-1, 3), W1.ravel(), W2.ravel()

    def f(e, t, p):
        e -= 1
        if t == r.plusTri[e]:
            s, A, v = r.plusSign[e], r.Ap[e], r.rp[e]
            return s * r.len[e] / (2 * A) * (p - v), s * r.len[e] / A
        s, A, v = r.minusSign[e], r.Am[e], r.rm[e]
        return s * r.len[e] / (2 * A) * (v - p), -s * r.len[e] / A

    fm, dm = f(m, t1, P1)
    fn, dn = f(n, t2, P2)
    R = np.linalg.norm(P1[:, None] - P2[None], axis=-1)
    G = np.exp(-1j * k * R) / (4 * np.pi * R)
    Wt = W1[:, None] * W2[None]
    return np.sum(Wt * (fm @ fn.T) * G), np.sum(Wt * dm * dn * G)


def test_regular_rule_converges_to_fine_product_rule(orc):
    V, F, r, P = _problem(orc, "c1")
    k = 2 * np.pi * 3e9 / W.C0
    o = orc.opts(**W.OPTS_C1)
    # separated pairs: patch/ground, far on the ground plane
    cases = [(5, 300), (40, 420), (300, 520)]
    for m, n in cases:
        for t1 in (r.plusTri[m - 1], r.minusTri[m - 1]):
            for t2 in (r.plusTri[n - 1], r.minusTri[n - 1]):
                assert not P.near(t1, t2, 2.5)
                iv, isc = P.pair_integral(0, m, n, t1, t2, k, o)
                fv, fs = _fine_pair_integral(V, F, r, m, n, t1, t2, k)
                assert abs(iv - fv) <= 1e-5 * abs(fv)
                assert abs(isc - fs) <= 1e-5 * abs(fs)


def test_near_rule_tends_to_regular_without_softening(orc):
    V, F, r, P = _problem(orc, "c1")
    k = 2 * np.pi * 2e9 / W.C0
    m, n = 5, 300
    t1, t2 = r.plusTri[m - 1], r.plusTri[n - 1]
    fv, fs = _fine_pair_integral(V, F, r, m, n, t1, t2, k)
    o = orc.opts(**dict(W.OPTS_C1, near_alpha=0.0, near_subdiv=2, near_order=7))
    iv, isc = P.pair_integral(1, m, n, t1, t2, k, o)
    assert abs(iv - fv) <= 1e-7 * abs(fv) and abs(isc - fs) <= 1e-7 * abs(fs)
    # with the default softening the near rule differs by O(a^2/R^2)
    o = orc.opts(**W.OPTS_C1)
    iv2, _ = P.pair_integral(1, m, n, t1, t2, k, o)
    assert abs(iv2 - fv) > abs(iv - fv)


def test_pair_symmetry_and_kind_errors(orc):
    V, F, r, P = _problem(orc, "c1")
    k = 2 * np.pi * 2.5e9 / W.C0
    o = orc.opts(**W.OPTS_C1)
    m, n = 8, 12
    t1, t2 = r.plusTri[m - 1], r.minusTri[n - 1]
    kind = 1 if P.near(t1, t2, 2.5) else 0
    a = P.pair_integral(kind, m, n, t1, t2, k, o)
    b = P.pair_integral(kind, n, m, t2, t1, k, o)
    assert np.allclose(a, b, rtol=1e-14, atol=0)
    with pytest.raises(ValueError):  # rwg_pair_integrals_regular.m:2-4
        P.pair_integral(0, m, m, t1, t1, k, o)
    with pytest.raises(ValueError):  # rwg_pair_integrals_singular.m:3-5
        P.pair_integral(2, m, n, t1, t2, k, o)


def test_static_limit_scaling(orc):
    # k -> 0: Z = j w mu Ivec0 + Isca0 / (j w eps): Im(Z) w is affine in w^2
    V, F, r, P = _problem(orc, "micro")
    o = orc.opts(**W.OPTS_C1)
    zs = []
    for w in (1e3, 2e3):
        zs.append(P.z_rows([1, 2], 0.0, w, W.MU0, W.EPS0, o) * w)
    # w*Z = j w^2 mu Ivec0 - j Isca0/eps  -> difference of the two is j 3e6 mu Ivec0
    d = (zs[1] - zs[0]) / (1j * 3e6 * W.MU0)
    assert np.all(np.abs(d.imag) <= 1e-12 * np.abs(d))


def test_lu_matches_lapack(orc):
    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 40):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        LU, piv, info = orc.zgetrf(A)
        lu2, piv2 = sla.lu_factor(A)
        assert info == 0
        assert np.array_equal(piv - 1, piv2)          # same pivot rows (izamax on |re|+|im|)
        assert np.allclose(LU, lu2, rtol=0, atol=1e-12 * np.abs(lu2).max())
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        x = orc.zgetrs(LU, piv, b)
        assert np.allclose(x, sla.lu_solve((lu2, piv2), b), rtol=1e-11, atol=0)
    # known answers: identity, permutation, singular
    I = np.eye(5, dtype=complex)
    LU, piv, info = orc.zgetrf(I)
    assert info == 0 and np.array_equal(LU, I) and np.array_equal(piv, np.arange(1, 6))
    Pm = np.eye(4)[[2, 0, 3, 1]].astype(complex)
    LU, piv, info = orc.zgetrf(Pm)
    x = orc.zgetrs(LU, piv, np.arange(4) + 0j)
    assert np.array_equal(x, np.linalg.solve(Pm, np.arange(4) + 0j))
    S = np.ones((3, 3), dtype=complex)
    assert orc.zgetrf(S)[2] == 2


def test_sweep_matches_numpy_pipeline(orc):
    V, F, r, P = _problem(orc, "micro")
    pP, pG = W.feed_points("micro")
    pe, ge = orc.pick_port(r, pP, 1), orc.pick_port(r, pG, -1)
    spread = 1.5 * r.len.mean()
    o = orc.opts(**W.OPTS_C1)
    fr = W.freqs("micro")
    eps = W.EPS0 * (1 - 1e-3j)
    zin, s11 = P.sweep(fr, W.MU0, eps, pe, ge, 1.0, spread, 50.0, o)
    wP, wG = P.port_weights(pe, ge, spread)
    for i, f in enumerate(fr):
        w = 2 * np.pi * f
        k = orc.wavenumber(w, W.MU0, eps)
        Z = P.z_full(k, w, W.MU0, eps, o)
        x = np.linalg.solve(Z, P.rhs(pe, ge, 1.0, spread))
        ip = np.sum(wP * x * r.len) - np.sum(wG * x * r.len)
        assert zin[i] == pytest.approx(1.0 / ip, rel=1e-11)
        assert s11[i] == pytest.approx((zin[i] - 50) / (zin[i] + 50), rel=1e-13)
