"""GPU tests at the workloads' full sizes.

- c2_6k (Ne = 6009): the FP64 and mixed sweeps' Zin/S11 at 1.6, 2.42 and
  3.4 GHz against the oracle's full-matrix solution held in
  tests/golden/c2_6k.npz (oracle Z + LAPACK zgesv, made by
  tests/golden/make_golden.py), 32 rows of Z at 2.5 GHz against the
  fixture's oracle rows, and efie_zgesv (FP64) on the device-filled Z at
  2.42 GHz against the fixture's solution x;
- the batched uneven sweep (32 seeded geometries, N 1000-8000, 64
  frequencies each): every frequency refined or re-solved, Zin of every
  fourth geometry against the FP64 path, the smallest geometry against the
  oracle's own sweep;
- the learned surrogate at its workload size: trained on 2M pairs, then a
  c2_6k sweep (every near pair inside the envelope), its Z rows against the
  direct near-pair quadrature and its Zin against the FP64 path.

Tolerances: Zin 1e-10 relative against the oracle (both paths; the
condition estimate of each Z is printed beside the error), Z entries 1e-10
relative to the entry.  The solution vector of an ill-conditioned system is
held to the normwise forward-error bound 50 * eps * cond_1(Z) and to an
FP64 backward error.
"""
import os

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ZIN_TOL = 1e-10
ZTOL = 1e-10


@pytest.fixture(scope="module")
def ctx(efie_mod):
    c = efie_mod.Context(0)
    yield c
    c.close()


def _problem(efie_mod, geom, cells, feed):
    V, F = efie_mod.make_patch_with_ground_mesh(*geom, *cells)
    r = efie_mod.build_rwg_interior(V, F)
    pe = efie_mod.pick_port_edge_by_point_and_layer(r, feed[0], 1)
    ge = efie_mod.pick_port_edge_by_point_and_layer(r, feed[1], -1)
    return V, F, r, pe, ge, 1.5 * r.len.mean()


@pytest.fixture(scope="module")
def c2(efie_mod):
    g = np.load(os.path.join(GOLDEN, "c2_6k.npz"))
    V, F, r, pe, ge, spread = _problem(efie_mod, W.CONFIGS["c2_6k"]["geom"], W.CONFIGS["c2_6k"]["cells"],
                                       W.feed_points("c2_6k"))
    assert r.Ne == int(g["Ne"]) and [pe, ge] == list(g["ports"]) and spread == float(g["spread"])
    return g, V, F, r, pe, ge, spread


# ------------------------------------------------------------------ c2_6k --

@pytest.mark.parametrize("path", ["fp64", "mixed"])
def test_c2_6k_sweep_vs_oracle_golden(efie_mod, ctx, c2, path):
    g, V, F, r, pe, ge, spread = c2
    ctx.set_problem(V, F, r)
    o = (efie_mod.opts(**W.OPTS_C1) if path == "fp64" else
         efie_mod.opts(**W.OPTS_C1, precision=efie_mod.PREC_MIXED, near_mode=efie_mod.NEAR_MOMENTS, ir_tol=1e-10))
    zin, s11, st = ctx.sweep(g["zin_freqs"], W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o, stats=True)
    err = np.abs(zin - g["zin"]) / np.abs(g["zin"])
    for f, e, rc in zip(g["zin_freqs"], err, g["rcond"]):
        print(f"{path} {f / 1e9:.2f} GHz: |dZin|/|Zin| = {e:.2e} (cond_1 ~ {1 / rc:.1e})", flush=True)
    assert np.max(err) <= ZIN_TOL
    assert np.max(np.abs(s11 - g["s11"])) <= ZIN_TOL
    if path == "mixed":
        assert st["ir_resid_max"] <= 1e-10


def test_c2_6k_pipelined_refinement_forms(efie_mod, ctx, c2):
    """Three pipelined batches of four frequencies: the first two are refined
    beside the next batch's factorisation (32-row residual strips, solves
    held to 128 registers), the last alone (64-row strips, uncapped solves),
    and the next batch's first update overlaps each factorisation's tail.
    A profiled sweep (one stream, every kernel alone) takes the wide forms
    throughout.  Zin at the oracle's three frequencies (one in each batch)
    against the golden solution, and every frequency equal between the two
    sweeps to the refinement accuracy."""
    g, V, F, r, pe, ge, spread = c2
    ctx.set_problem(V, F, r)
    fz = [float(f) for f in g["zin_freqs"]]
    fr = np.array([fz[0], 1.7e9, 1.8e9, 1.9e9, 2.3e9, fz[1], 2.5e9, 2.6e9, 3.2e9, 3.3e9, fz[2], 3.5e9])
    at = [0, 5, 10]
    o = dict(W.OPTS_C1, precision=efie_mod.PREC_MIXED, near_mode=efie_mod.NEAR_MOMENTS, ir_tol=1e-10, batch=4)
    zp, _, sp = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**o), stats=True)
    zs, _, ss = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**o, profile=1), stats=True)
    for st in (sp, ss):
        assert st["ir_resid_max"] <= 1e-10 and st["n_lu_retry"] == 0 and st["n_fp64_retry"] == 0
    err = np.abs(zp[at] - g["zin"]) / np.abs(g["zin"])
    print("pipelined vs oracle", err, "corrections", sp["ir_iters_total"], ss["ir_iters_total"], flush=True)
    assert np.max(err) <= ZIN_TOL
    assert np.max(np.abs(zp - zs) / np.abs(zs)) <= 1e-9


@pytest.mark.parametrize("near_mode", [0, 1])
def test_c2_6k_rows_vs_oracle_golden(efie_mod, ctx, c2, near_mode):
    g, V, F, r, pe, ge, spread = c2
    ctx.set_problem(V, F, r)
    w = 2 * np.pi * float(g["row_freq"])
    k = w * np.sqrt(W.MU0 * W.EPS0)
    Z, Vr = ctx.assemble_zv(k, w, W.MU0, W.EPS0, [pe, ge, 1.0, spread, 0.2],
                            efie_mod.opts(**W.OPTS_C1, near_mode=near_mode))
    rows = g["rows"] - 1
    err = float(np.max(np.abs(Z[rows] - g["zrows"]) / np.abs(g["zrows"])))
    print(f"near_mode {near_mode}: max entry-relative error of 32 rows {err:.2e}", flush=True)
    assert err <= ZTOL
    assert np.abs(Vr - g["vrhs"]).max() <= 1e-15


def test_c2_6k_zgesv_vs_oracle_golden(efie_mod, ctx, c2):
    """FP64 efie_zgesv on the device-filled Z at 2.42 GHz (the resonance
    band): normwise forward error of x within 50 eps cond_1(Z) of the
    oracle's LAPACK solution, backward error at FP64 level, and the Zin
    computed from x within 1e-10 of the oracle's."""
    g, V, F, r, pe, ge, spread = c2
    ctx.set_problem(V, F, r)
    i = int(np.argmin(np.abs(g["zin_freqs"] - 2.42e9)))
    w = 2 * np.pi * float(g["zin_freqs"][i])
    k = w * np.sqrt(W.MU0 * W.EPS0)
    Z, Vr = ctx.assemble_zv(k, w, W.MU0, W.EPS0, [pe, ge, 1.0, spread, 0.2], efie_mod.opts(**W.OPTS_C1))
    x, st = ctx.zgesv(Z, Vr, precision=efie_mod.PREC_FP64, stats=True)
    xo = g["x"][i]
    cond = 1.0 / float(g["rcond"][i])
    fwd = np.linalg.norm(x - xo, 1) / np.linalg.norm(xo, 1)
    berr = np.linalg.norm(Z @ x - Vr, 1) / (np.abs(Z).sum(axis=0).max() * np.linalg.norm(x, 1))
    # Zin from x as main_patch_efie_rwg.m:336-340 (Gaussian weights of the two port layers)
    mid = r.center
    zmid = 0.5 * (mid[:, 2].min() + mid[:, 2].max())
    lay = np.where(mid[:, 2] > zmid, 1, -1)
    wts = []
    for e0 in (pe, ge):
        d = np.linalg.norm(mid - mid[e0 - 1], axis=1) / spread
        wv = np.exp(-(d * d)) * (lay == lay[e0 - 1])
        wts.append(wv / wv.sum())
    zin = 1.0 / np.sum((wts[0] - wts[1]) * x * r.len)
    zerr = abs(zin - g["zin"][i]) / abs(g["zin"][i])
    print(f"zgesv FP64 at {g['zin_freqs'][i] / 1e9:.2f} GHz: forward {fwd:.2e} (bound {50 * 2.2e-16 * cond:.2e}), "
          f"backward {berr:.2e}, Zin {zerr:.2e}", flush=True)
    assert fwd <= 50 * 2.2e-16 * cond
    assert berr <= 1e-14
    assert zerr <= ZIN_TOL


# ------------------------------------------------------------------- c4 --

def _c4_problems(efie_mod):
    out = []
    for g in W.c4_geometries():
        out.append((g, *_problem(efie_mod, g["geom"], g["cells"], W.c4_feed_points(g["geom"]))))
    return out


@pytest.fixture(scope="module")
def c4(efie_mod):
    return _c4_problems(efie_mod)


def test_c4_full_size(efie_mod, ctx, c4):
    """All 32 geometries x 64 frequencies through the mixed path (FP32
    LDL^T + FP64 refinement to 1e-10) in one efie_sweep_many call (the
    geometries packed into shared batches by order, each series freed after
    its last batch): every frequency either refines to max(ir_tol, 1e-11)
    or is solved again (counted); every fourth geometry's Zin against the
    FP64 path -- within 1e-9 of the plain FP64 path (direct quadrature) at
    the workload's refinement tolerance 1e-10 (a residual of 1e-10 relative
    shows as up to a few 1e-10 in Zin near resonance), and within 1e-10 of
    the FP64 path with the same fill (series) at tolerance 1e-11, which
    isolates the FP32 factorisation's share."""
    fr = np.linspace(*W.C4["freqs"])
    o = dict(W.OPTS_C1, precision=efie_mod.PREC_MIXED, near_mode=efie_mod.NEAR_MOMENTS)
    ns = [p[3].Ne for p in c4]
    assert len(c4) == 32 and min(ns) >= 1000 and max(ns) <= 8000
    ctxs = [efie_mod.Context(0) for _ in c4]
    eng = efie_mod.Context(0)
    try:
        items = []
        for c, (g, V, F, r, pe, ge, spread) in zip(ctxs, c4):
            c.set_problem(V, F, r)
            items.append(dict(ctx=c, freqs=fr, patchEdge=pe, groundEdge=ge, V0=1.0, spread=spread, Z0=50.0,
                              release=1))
        res, st = efie_mod.sweep_many(eng, items, W.MU0, W.EPS0, efie_mod.opts(**o, ir_tol=1e-10), stats=True)
        eng.close()  # its batch buffers, before the comparison sweeps take theirs
        assert all(np.all(np.isfinite(z)) for z, _ in res)
        assert st["ir_resid_max"] <= 1e-10
        print(f"c4: N {min(ns)}-{max(ns)}, {len(c4) * len(fr)} points in {st['ms_total']:.0f} ms, "
              f"{st['n_lu_retry']} re-solved with pivoting", flush=True)
        worst, worst11 = 0.0, 0.0
        for i in range(0, len(c4), 4):
            g, V, F, r, pe, ge, spread = c4[i]
            cx = ctxs[i]
            z64 = cx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**W.OPTS_C1))[0]
            z64m = cx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                            efie_mod.opts(**W.OPTS_C1, near_mode=efie_mod.NEAR_MOMENTS))[0]
            z11 = cx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**o, ir_tol=1e-11))[0]
            e10 = float(np.max(np.abs(res[i][0] - z64) / np.abs(z64)))
            e11 = float(np.max(np.abs(z11 - z64m) / np.abs(z64m)))
            print(f"  geometry {i}: N={r.Ne} |dZin|/|Zin| {e10:.2e} (ir_tol 1e-10), {e11:.2e} (1e-11)", flush=True)
            worst, worst11 = max(worst, e10), max(worst11, e11)
            cx.close()
    finally:
        eng.close()
        for c in ctxs:
            c.close()
    assert worst <= 1e-9
    assert worst11 <= ZIN_TOL


def test_c4_smallest_vs_oracle(efie_mod, orc, ctx, c4):
    """The smallest geometry of the batched workload against the oracle's
    own sweep (its fill and LU) at the band ends and mid-band."""
    i = int(np.argmin([p[3].Ne for p in c4]))
    g, V, F, r, pe, ge, spread = c4[i]
    fr = np.array([1.0e9, 2.5e9, 4.0e9])
    P = orc.Problem(V, F, orc.build_rwg(V, F))
    zo = P.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, orc.opts(**W.OPTS_C1),
                 threads=min(os.cpu_count() or 1, 16))[0]
    ctx.set_problem(V, F, r)
    z64 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**W.OPTS_C1))[0]
    zm = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                   efie_mod.opts(**W.OPTS_C1, precision=efie_mod.PREC_MIXED, near_mode=efie_mod.NEAR_MOMENTS,
                                 ir_tol=1e-11))[0]
    e64 = float(np.max(np.abs(z64 - zo) / np.abs(zo)))
    em = float(np.max(np.abs(zm - zo) / np.abs(zo)))
    print(f"c4 geometry {i} (N={r.Ne}): FP64 {e64:.2e}, mixed {em:.2e} against the oracle", flush=True)
    assert e64 <= ZIN_TOL and em <= ZIN_TOL


# ------------------------------------------------------------------- c5 --

def test_c5_surrogate_at_scale(efie_mod, ctx, c2):
    """The learned near-pair surrogate at its workload: trained on 2M pairs
    (the library's defaults), applied in the c2_6k fill.  Every near pair of
    c2_6k is inside the envelope (no fallback); 32 rows of Z at 2.5 GHz are
    within 1e-3 (entry-relative, largest entry of the row) of the direct
    near-pair quadrature, where the plain product rule alone is off by more
    than 1e-2; Zin at five frequencies within 5e-3 of the FP64 path
    (measured 1.1e-3: the network's record error of ~3e-4 is amplified near
    resonance, which is why the headline path keeps the exact moment series)."""
    g, V, F, r, pe, ge, spread = c2
    rep = ctx.sur_train(efie_mod.opts(**W.OPTS_C1))
    print("surrogate training:", {k: rep[k] for k in ("n_train", "steps", "val_rel_rms", "base_rel_rms")},
          flush=True)
    assert rep["n_train"] == 2_000_000
    assert rep["val_rel_rms"] <= 0.05 * rep["base_rel_rms"]
    ctx.set_problem(V, F, r)
    w = 2 * np.pi * float(g["row_freq"])
    k = w * np.sqrt(W.MU0 * W.EPS0)
    exc = [pe, ge, 1.0, spread, 0.2]
    rows = g["rows"] - 1
    Zs = ctx.assemble_zv(k, w, W.MU0, W.EPS0, exc, efie_mod.opts(**W.OPTS_C1, near_mode=efie_mod.NEAR_SURROGATE))[0]
    Zs = Zs[rows].copy()
    blob = ctx.sur_get()
    scale = np.abs(g["zrows"]).max(axis=1, keepdims=True)
    err = float(np.max(np.abs(Zs - g["zrows"]) / scale))
    fr = np.array([1.6e9, 2.0e9, 2.42e9, 2.8e9, 3.4e9])
    o = dict(W.OPTS_C1, precision=efie_mod.PREC_MIXED, ir_tol=1e-10)
    zs, _, st = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                          efie_mod.opts(**o, near_mode=efie_mod.NEAR_SURROGATE), stats=True)
    z64 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**W.OPTS_C1))[0]
    zerr = float(np.max(np.abs(zs - z64) / np.abs(z64)))
    print(f"c5 at c2_6k: {st['n_sur_pairs']} near pairs by the network, {st['n_sur_fallback']} by quadrature; "
          f"Z rows {err:.2e}; Zin {zerr:.2e}", flush=True)
    assert st["n_sur_pairs"] == st["n_near_pairs"] and st["n_sur_fallback"] == 0
    assert err <= 1e-3
    assert zerr <= 5e-3
    # what the network gains: the plain product rule alone on the same rows
    nb = np.zeros_like(blob)
    nb[:] = blob
    import test_gpu_surrogate as TS
    ctx.sur_set(TS._neutral_blob(W.OPTS_C1))
    Zb = ctx.assemble_zv(k, w, W.MU0, W.EPS0, exc, efie_mod.opts(**W.OPTS_C1, near_mode=efie_mod.NEAR_SURROGATE))[0]
    errb = float(np.max(np.abs(Zb[rows] - g["zrows"]) / scale))
    ctx.sur_set(nb)
    print(f"plain product rule alone: Z rows {errb:.2e}", flush=True)
    assert errb >= 10 * err
