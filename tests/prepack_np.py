"""The `prepack` struct of efie_precompute.m:1-97 packed by
efie_precompute_pack.m:1-59, restated in numpy on the oracle's quadrature
rules and subdivision (oracle/mirror_np.py), in MATLAB's column-major
layouts: the input the prepack MEX entry points and efie_set_prepack take
in the tests.

- nearMat = D < near_sing_eps * (htri1 + htri2) / 2, htri = sqrt(2 area2 / sqrt(3))
  (efie_precompute.m:19-29), diagonal included as the reference has it;
- r_reg(i, :, t) = A + u (B - A) + v (C - A), w_reg(i, t) = qw(i) area2(t)
  (:31-38, map_ref_to_tri.m);
- self / near caches on the 4^depth midpoint children (subdivide_triangles.m
  child order), weights qw2 * 2A_sub, self_a = self_alpha sqrt(A_sub / pi)
  (:49-94);
- aeqTri = sqrt(A / pi) (:41-43).
"""
import numpy as np

import mirror_np as M


def make_prepack(V, F, quad_order=4, near_order=4, near_sing_eps=2.5, self_subdiv=2, self_alpha=0.20,
                 near_subdiv=2, near_alpha=0.15):
    F0 = np.asarray(F, dtype=np.int64) - 1
    r1, r2, r3 = V[F0[:, 0]], V[F0[:, 1]], V[F0[:, 2]]
    Nt = F0.shape[0]
    area2 = np.linalg.norm(np.cross(r2 - r1, r3 - r1), axis=1)
    cent = (r1 + r2 + r3) / 3
    qp, qw = M.quad_rule(quad_order)
    qp2, qw2 = M.quad_rule(near_order)
    htri = np.sqrt(2 * area2 / np.sqrt(3))
    d = cent[:, None, :] - cent[None, :, :]
    D = np.sqrt((d ** 2).sum(-1))
    nearMat = (D < near_sing_eps * 0.5 * (htri[:, None] + htri[None, :])).astype(np.uint8)
    u, v = qp[:, 0], qp[:, 1]
    # [Nt, Nq, 3] -> [Nq, 3, Nt]
    reg = r1[:, None, :] + u[None, :, None] * (r2 - r1)[:, None, :] + v[None, :, None] * (r3 - r1)[:, None, :]
    r_reg = np.transpose(reg, (1, 2, 0)).copy()
    w_reg = qw[:, None] * area2[None, :]

    def sub_cache(depth):
        Ns, Nq2 = 4 ** depth, qp2.shape[0]
        r = np.zeros((Nq2, 3, Ns, Nt))
        w = np.zeros((Nq2, Ns, Nt))
        a2s = np.zeros((Ns, Nt))
        for t in range(Nt):
            tris = M.subdivide(r1[t], r2[t], r3[t], depth)
            pts, ws, a2 = M.points_of(tris, qp2, qw2)      # [Ns, Nq2, 3], [Ns, Nq2]
            r[:, :, :, t] = np.transpose(pts, (1, 2, 0))
            w[:, :, t] = ws.T
            a2s[:, t] = a2
        return r, w, a2s
    self_r, self_w, a2S = sub_cache(self_subdiv)
    near_r, near_w, _ = sub_cache(near_subdiv)
    self_a = self_alpha * np.sqrt(0.5 * a2S / np.pi)
    return dict(r_reg=r_reg, w_reg=w_reg, nearMat=nearMat, self_r=self_r, self_w=self_w, self_a=self_a,
                near_r=near_r, near_w=near_w, aeqTri=np.sqrt(0.5 * area2 / np.pi)[:, None],
                near_alpha=float(near_alpha), self_alpha=float(self_alpha))
