"""Independent numpy restatement of the reference EFIE path, for small meshes.

TEST INFRASTRUCTURE ONLY.  It exists to pin the C oracle
(oracle/efie_oracle.c) through a second implementation written with
different data structures (vectorised numpy, MATLAB's `unique(...,'rows')`
reproduced with ``np.unique(axis=0)``), so that a transposed index, a dropped
term or a wrong sign in either one shows up as a disagreement.  Pure numpy,
no ctypes; intended for meshes with at most a few hundred unknowns.
"""
from __future__ import annotations

import numpy as np


def linspace(d1, d2, n):
    """MATLAB linspace: d1 + (0:n1).*(d2-d1)./n1, last element = d2."""
    n1 = n - 1
    y = d1 + (np.arange(n1 + 1) * (d2 - d1)) / n1
    y[-1] = d2
    return y


def make_mesh(Lp, Wp, Lg, Wg, h, npx, npy, ngx, ngy):
    """make_patch_with_ground_mesh.m:1-46 (meshgrid column order for the
    vertices, (Np_x+1) stride for the faces)."""
    def layer(W, L, nx, ny, z, off):
        X, Y = np.meshgrid(linspace(-W / 2, W / 2, nx + 1), linspace(-L / 2, L / 2, ny + 1))
        Vl = np.column_stack([X.ravel(order="F"), Y.ravel(order="F"), np.full(X.size, z)])
        j, i = np.meshgrid(np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
        v1 = off + (j.ravel() - 1) * (nx + 1) + i.ravel()
        v2, v3 = v1 + 1, v1 + (nx + 1)
        v4 = v3 + 1
        Fl = np.empty((2 * v1.size, 3), dtype=np.int64)
        Fl[0::2] = np.column_stack([v1, v2, v4])
        Fl[1::2] = np.column_stack([v1, v4, v3])
        return Vl, Fl

    Vp, Fp = layer(Wp, Lp, npx, npy, +h / 2, 0)
    Vg, Fg = layer(Wg, Lg, ngx, ngy, -h / 2, Vp.shape[0])
    return np.vstack([Vp, Vg]), np.vstack([Fp, Fg]).astype(np.int32)


def tri_geom(V, F):
    r1, r2, r3 = V[F[:, 0] - 1], V[F[:, 1] - 1], V[F[:, 2] - 1]
    area2 = np.linalg.norm(np.cross(r2 - r1, r3 - r1), axis=1)
    return r1, r2, r3, area2, (r1 + r2 + r3) / 3


def build_rwg(V, F):
    """build_rwg_interior.m:1-113 via np.unique over sorted vertex pairs."""
    Nt = F.shape[0]
    d0 = np.stack([F[:, 0], F[:, 1], F[:, 2]], axis=1).reshape(-1)
    d1 = np.stack([F[:, 1], F[:, 2], F[:, 0]], axis=1).reshape(-1)
    raw_tri = np.repeat(np.arange(1, Nt + 1), 3)
    key = np.column_stack([np.minimum(d0, d1), np.maximum(d0, d1)])
    ue, ic, counts = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    ic = ic.ravel()
    interior = np.flatnonzero(counts == 2)
    Ne = interior.size
    out = {k: np.zeros(Ne, np.int32) for k in ("plusTri", "minusTri", "plusSign", "minusSign")}
    # the two raw rows of every interior edge, in raw order
    order = np.argsort(ic, kind="stable")
    first = np.searchsorted(ic[order], interior)
    r1, r2 = order[first], order[first + 1]
    a, b = ue[interior, 0], ue[interior, 1]
    out["plusTri"][:] = raw_tri[r1]
    out["minusTri"][:] = raw_tri[r2]
    out["plusSign"][:] = np.where((d0[r1] == a) & (d1[r1] == b), 1, -1)
    out["minusSign"][:] = np.where((d0[r2] == a) & (d1[r2] == b), 1, -1)
    ra, rb = V[a - 1], V[b - 1]
    length = np.linalg.norm(rb - ra, axis=1)
    center = 0.5 * (ra + rb)

    def opp(tris):
        Ft = F[tris - 1]
        mask = (Ft != a[:, None]) & (Ft != b[:, None])
        assert (mask.sum(1) == 1).all(), "opposite vertex not found"
        o = Ft[mask]
        r1_, r2_, r3_ = V[Ft[:, 0] - 1], V[Ft[:, 1] - 1], V[Ft[:, 2] - 1]
        A = 0.5 * np.linalg.norm(np.cross(r2_ - r1_, r3_ - r1_), axis=1)
        return V[o - 1], A

    rp, Ap = opp(out["plusTri"])
    rm, Am = opp(out["minusTri"])
    zmid = 0.5 * (V[:, 2].max() + V[:, 2].min())
    zSign = np.where(center[:, 2] > zmid, 1, -1).astype(np.int32)
    return dict(edge=np.column_stack([a, b]).astype(np.int32), rp=rp, rm=rm, Ap=Ap, Am=Am,
                len=length, center=center, zSign=zSign, Ne=Ne, **out)


def quad_rule(order):
    if order == 1:
        return np.array([[1 / 3, 1 / 3]]), np.array([0.5])
    if order == 2:
        return np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]), np.full(3, 1 / 6)
    if order == 7:
        s = np.sqrt(15.0)
        a1, a2 = (6 - s) / 21, (6 + s) / 21
        qp = np.array([[1 / 3, 1 / 3], [a1, a1], [1 - 2 * a1, a1], [a1, 1 - 2 * a1],
                       [a2, a2], [1 - 2 * a2, a2], [a2, 1 - 2 * a2]])
        qw = np.array([9 / 80] + [(155 - s) / 2400] * 3 + [(155 + s) / 2400] * 3)
        return qp, qw
    a, b = 0.445948490915965, 0.108103018168070
    w1, w2 = 0.111690794839005, 0.054975871827661
    qp = np.array([[a, a], [a, 1 - 2 * a], [1 - 2 * a, a], [b, b], [b, 1 - 2 * b], [1 - 2 * b, b]])
    return qp, np.array([w1] * 3 + [w2] * 3)


def subdivide(A, B, C, depth):
    """4-way midpoint split, child order [A,AB,CA],[AB,B,BC],[CA,BC,C],[AB,BC,CA]."""
    tris = np.array([[A, B, C]])
    for _ in range(depth):
        A0, B0, C0 = tris[:, 0], tris[:, 1], tris[:, 2]
        AB, BC, CA = 0.5 * (A0 + B0), 0.5 * (B0 + C0), 0.5 * (C0 + A0)
        kids = np.stack([np.stack([A0, AB, CA], 1), np.stack([AB, B0, BC], 1),
                         np.stack([CA, BC, C0], 1), np.stack([AB, BC, CA], 1)], axis=1)
        tris = kids.reshape(-1, 3, 3)
    return tris


def points_of(tris, qp, qw):
    """Quadrature points/weights of every sub-triangle: ([Ns,Nq,3], [Ns,Nq])."""
    A, B, C = tris[:, 0], tris[:, 1], tris[:, 2]
    pts = A[:, None, :] + qp[None, :, 0:1] * (B - A)[:, None, :] + qp[None, :, 1:2] * (C - A)[:, None, :]
    a2 = np.linalg.norm(np.cross(B - A, C - A), axis=1)
    return pts, qw[None, :] * a2[:, None], a2


def rwg_eval(rwg, m, t, r):
    e = m - 1
    if t == rwg["plusTri"][e]:
        s, l, A = rwg["plusSign"][e], rwg["len"][e], rwg["Ap"][e]
        return s * (l / (2 * A)) * (r - rwg["rp"][e]), np.full(r.shape[:-1], s * (l / A))
    if t == rwg["minusTri"][e]:
        s, l, A = rwg["minusSign"][e], rwg["len"][e], rwg["Am"][e]
        return s * (l / (2 * A)) * (rwg["rm"][e] - r), np.full(r.shape[:-1], s * (-l / A))
    return np.zeros_like(r), np.zeros(r.shape[:-1])


def green(k, R):
    return np.exp(-1j * k * R) / (4 * np.pi * R)


def pair_integral(V, F, rwg, m, n, t1, t2, k, opts, geom=None):
    """One [Ivec, Isca] with the self / near / regular dispatch of
    assemble_ZV_EFIE_RWG_DIFF_CPU.m:22-28."""
    r1, r2, r3, area2, cent = geom or tri_geom(V, F)
    T1 = np.array([r1[t1 - 1], r2[t1 - 1], r3[t1 - 1]])
    T2 = np.array([r1[t2 - 1], r2[t2 - 1], r3[t2 - 1]])
    if t1 == t2:
        kind = 2
    else:
        d = np.linalg.norm(cent[t1 - 1] - cent[t2 - 1])
        l1 = np.sqrt(4 * (0.5 * area2[t1 - 1]) / np.sqrt(3))
        l2 = np.sqrt(4 * (0.5 * area2[t2 - 1]) / np.sqrt(3))
        kind = 1 if d < opts["near_sing_eps"] * 0.5 * (l1 + l2) else 0
    if kind == 0:
        qp, qw = quad_rule(opts["quad_order"])
        P1, W1, _ = points_of(T1[None], qp, qw)
        P2, W2, _ = points_of(T2[None], qp, qw)
        # regular rule weights are qw*J with J = area2 of the whole triangle
        P1, P2 = P1[0], P2[0]
        W1, W2 = qw * area2[t1 - 1], qw * area2[t2 - 1]
        R = np.maximum(np.linalg.norm(P1[:, None] - P2[None], axis=-1), 1e-15)
        soft = None
    else:
        qp, qw = quad_rule(opts["near_order"])
        depth = opts["self_subdiv"] if kind == 2 else opts["near_subdiv"]
        P1, W1, a2p = points_of(subdivide(*T1, depth), qp, qw)
        P2, W2, _ = points_of(subdivide(*T2, depth), qp, qw)
        Ns, Nq = W1.shape
        P1, P2, W1, W2 = P1.reshape(-1, 3), P2.reshape(-1, 3), W1.ravel(), W2.ravel()
        R = np.linalg.norm(P1[:, None] - P2[None], axis=-1)
        if kind == 1:
            aeq = min(np.sqrt(0.5 * area2[t1 - 1] / np.pi), np.sqrt(0.5 * area2[t2 - 1] / np.pi))
            R = np.sqrt(R ** 2 + (opts["near_alpha"] * aeq) ** 2)
        else:
            sub = np.repeat(np.arange(Ns), Nq)
            same = sub[:, None] == sub[None, :]
            a_self = opts["self_alpha"] * np.sqrt(0.5 * a2p / np.pi)   # per test sub-triangle
            a_row = np.repeat(a_self, Nq)[:, None]
            R = np.where(same, np.sqrt(R ** 2 + a_row ** 2), np.maximum(R, 1e-15))
    fm, dm = rwg_eval(rwg, m, t1, P1)
    fn, dn = rwg_eval(rwg, n, t2, P2)
    G = green(k, R)
    W = W1[:, None] * W2[None, :]
    Ivec = np.sum(W * (fm @ fn.T) * G)
    Isca = np.sum(W * np.outer(dm, dn) * G)
    return Ivec, Isca


def assemble_z(V, F, rwg, k, omega, mu, eps, opts, rows=None):
    geom = tri_geom(V, F)
    Ne = rwg["Ne"]
    rows = np.arange(1, Ne + 1) if rows is None else np.asarray(rows)
    Z = np.zeros((len(rows), Ne), complex)
    for a, m in enumerate(rows):
        for n in range(1, Ne + 1):
            z = 0j
            for t1 in (rwg["plusTri"][m - 1], rwg["minusTri"][m - 1]):
                for t2 in (rwg["plusTri"][n - 1], rwg["minusTri"][n - 1]):
                    iv, isc = pair_integral(V, F, rwg, m, n, t1, t2, k, opts, geom)
                    z += 1j * omega * mu * iv + isc / (1j * omega * eps)
            Z[a, n - 1] = z
    return Z


def rhs(rwg, patchEdge, groundEdge, V0, spread):
    def g(e0, amp):
        d = np.linalg.norm(rwg["center"] - rwg["center"][e0 - 1], axis=1)
        w = np.exp(-(d / spread) ** 2) * (rwg["zSign"] == rwg["zSign"][e0 - 1])
        return amp * w / w.sum()
    return (g(patchEdge, V0 / 2) + g(groundEdge, -V0 / 2)).astype(complex)


def pick_port(rwg, p, layer):
    idx = np.flatnonzero(rwg["zSign"] == layer)
    d = np.linalg.norm(rwg["center"][idx] - np.asarray(p), axis=1)
    return int(idx[np.argmin(d)] + 1)
