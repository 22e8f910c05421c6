"""Writes the golden fixtures of tests/golden/ from the CPU FP64 oracle.

    python tests/golden/make_golden.py

The reference holds no golden vectors for this path (no tests, no
fixtures), so the fixtures are the oracle's outputs on the named workloads;
the oracle itself is pinned by the closed-form and cross-implementation
tests in tests/test_oracle_*.py.  Files are small .npz archives (data only).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "matlab-cuda-efief-vs-cpu_amd"))

import oracle as O  # noqa: E402
import workloads as W  # noqa: E402

C1_ROWS = [1, 2, 8, 100, 253, 400, 584]


def c1():
    V, F = O.make_mesh(*W.mesh_args("c1"))
    r = O.build_rwg(V, F)
    P = O.Problem(V, F, r)
    pP, pG = W.feed_points("c1")
    pe, ge = O.pick_port(r, pP, 1), O.pick_port(r, pG, -1)
    spread = 1.5 * r.len.mean()
    fr = W.freqs("c1")
    o = O.opts(**W.OPTS_C1)
    t = time.time()
    zin, s11 = P.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o)
    t_q = time.time() - t
    zin_c, s11_c = P.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o, centroid=True)
    w = 2 * np.pi * fr[0]
    k = O.wavenumber(w, W.MU0, W.EPS0)
    zrows = P.z_rows(C1_ROWS, k, w, W.MU0, W.EPS0, o)
    vrhs = P.rhs(pe, ge, 1.0, spread)
    np.savez_compressed(os.path.join(HERE, "c1.npz"), freqs=fr, zin=zin, s11=s11, zin_centroid=zin_c,
                        s11_centroid=s11_c, rows=np.array(C1_ROWS), zrows=zrows, vrhs=vrhs,
                        ports=np.array([pe, ge]), spread=spread, Ne=r.Ne,
                        edge=r.edge, plusTri=r.plusTri, minusTri=r.minusTri,
                        plusSign=r.plusSign, minusSign=r.minusSign)
    print(f"c1: Ne={r.Ne} ports={pe},{ge} oracle sweep {t_q:.1f}s")


if __name__ == "__main__":
    c1()
