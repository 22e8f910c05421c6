"""numpy restatement of triangles_are_near.m (and precompute_tri_geom.m)
over whole rows of triangle pairs, in the oracle's operation order, so a
mesh's near pairs can be listed at GPU-test sizes (the oracle's per-pair
call is pinned against it in test_oracle_tables.py).  Test helper only."""
import numpy as np


def tri_geom(V, F):
    """Centroids [Nt, 3] and twice the areas [Nt] as the oracle forms them."""
    r1, r2, r3 = V[F[:, 0] - 1], V[F[:, 1] - 1], V[F[:, 2] - 1]
    cent = (r1 + r2 + r3) / 3.0
    e1, e2 = r2 - r1, r3 - r1
    c = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                  e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1)
    area2 = np.sqrt(c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1] + c[:, 2] * c[:, 2])
    return cent, area2


def near_pairs(V, F, eps):
    """[n, 2] int32 (0-based t1 < t2, ordered by t1 then t2): d < eps *
    (l1 + l2) / 2 with l = sqrt(4 A / sqrt(3)) (triangles_are_near.m)."""
    cent, area2 = tri_geom(np.asarray(V, dtype=np.float64), np.asarray(F))
    A = 0.5 * area2
    l = np.sqrt(4 * A / np.sqrt(3.0))
    out = []
    Nt = len(A)
    for t1 in range(Nt - 1):
        d3 = cent[t1] - cent[t1 + 1:]
        d = np.sqrt(d3[:, 0] * d3[:, 0] + d3[:, 1] * d3[:, 1] + d3[:, 2] * d3[:, 2])
        h = 0.5 * (l[t1] + l[t1 + 1:])
        t2 = np.nonzero(d < eps * h)[0] + t1 + 1
        out.append(np.stack([np.full(len(t2), t1), t2], axis=1))
    return np.concatenate(out).astype(np.int32) if out else np.zeros((0, 2), np.int32)
