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

# c2_6k: full-matrix solves at the band edges and in the resonance band,
# 32 rows of Z at 2.5 GHz (the two port edges among them)
C2_ZIN_FREQS = (1.6e9, 2.42e9, 3.4e9)
C2_ROW_FREQ = 2.5e9
C2_NROWS = 32


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


def c2_6k(threads=None):
    """Zin/S11 and the solution x of Z x = V at C2_ZIN_FREQS from the
    oracle's full FP64 Z (assemble_ZV_EFIE_RWG_DIFF_CPU.m:9-36) and LAPACK
    zgesv (the `Z \\ Vrhs` of main_patch_efie_rwg.m:333), Zin as
    main_patch_efie_rwg.m:336-340; plus 32 rows of Z at C2_ROW_FREQ and the
    1-norm condition estimate of each solved Z (LAPACK zgecon)."""
    import scipy.linalg as sl

    V, F = O.make_mesh(*W.mesh_args("c2_6k"))
    r = O.build_rwg(V, F)
    P = O.Problem(V, F, r)
    pP, pG = W.feed_points("c2_6k")
    pe, ge = O.pick_port(r, pP, 1), O.pick_port(r, pG, -1)
    spread = 1.5 * r.len.mean()
    o = O.opts(**W.OPTS_C1)
    vrhs = P.rhs(pe, ge, 1.0, spread)
    wP, wG = P.port_weights(pe, ge, spread)
    fz = np.array(C2_ZIN_FREQS)
    zin = np.zeros(len(fz), complex)
    xs = np.zeros((len(fz), r.Ne), complex)
    rcond = np.zeros(len(fz))
    for i, f in enumerate(fz):
        t = time.time()
        w = 2 * np.pi * f
        Z = P.z_full(O.wavenumber(w, W.MU0, W.EPS0), w, W.MU0, W.EPS0, o, threads)
        anorm = np.abs(Z).sum(axis=0).max()
        lu, piv = sl.lu_factor(Z, check_finite=False)
        x = sl.lu_solve((lu, piv), vrhs, check_finite=False)
        rc, info = sl.lapack.zgecon(lu, anorm, norm="1")
        xs[i] = x
        rcond[i] = rc
        zin[i] = 1.0 / (np.sum(wP * x * r.len) - np.sum(wG * x * r.len))
        print(f"c2_6k {f / 1e9:.2f} GHz: Zin={zin[i]:.6f} rcond={rc:.2e} ({time.time() - t:.0f}s)", flush=True)
    s11 = (zin - 50.0) / (zin + 50.0)
    rng = np.random.Generator(np.random.PCG64(6009))
    fixed = np.unique([1, pe, ge, r.Ne])
    pool = np.setdiff1d(np.arange(1, r.Ne + 1), fixed)
    rows = np.sort(np.concatenate([fixed, rng.choice(pool, C2_NROWS - len(fixed), replace=False)]))
    w = 2 * np.pi * C2_ROW_FREQ
    zrows = P.z_rows(rows, O.wavenumber(w, W.MU0, W.EPS0), w, W.MU0, W.EPS0, o, threads)
    np.savez_compressed(os.path.join(HERE, "c2_6k.npz"), zin_freqs=fz, zin=zin, s11=s11, x=xs, rcond=rcond,
                        row_freq=C2_ROW_FREQ, rows=rows.astype(np.int32), zrows=zrows, vrhs=vrhs,
                        ports=np.array([pe, ge]), spread=spread, Ne=r.Ne)
    print(f"c2_6k: Ne={r.Ne} ports={pe},{ge} rows={len(rows)}")


if __name__ == "__main__":
    which = sys.argv[1:] or ["c1"]
    if "c1" in which:
        c1()
    if "c2_6k" in which:
        c2_6k()
