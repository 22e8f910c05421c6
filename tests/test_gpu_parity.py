"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Tolerances (FP64 path): Z entries within 1e-10 of the oracle relative to the
entry (observed ~1e-15), Zin within 1e-10 relative.  The mixed-precision
path (complex FP32 LU + FP64 refinement) is held to the same 1e-10 on Zin.
"""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

ZTOL = 1e-10
ZIN_TOL = 1e-10


@pytest.fixture(scope="module")
def ctx(efie_mod):
    c = efie_mod.Context(0)
    yield c
    c.close()


def _setup(efie_mod, orc, name):
    V, F = efie_mod.make_patch_with_ground_mesh(*W.mesh_args(name))
    r = efie_mod.build_rwg_interior(V, F)
    pP, pG = W.feed_points(name)
    pe = efie_mod.pick_port_edge_by_point_and_layer(r, pP, 1)
    ge = efie_mod.pick_port_edge_by_point_and_layer(r, pG, -1)
    P = orc.Problem(V, F, orc.build_rwg(V, F))
    return V, F, r, pe, ge, 1.5 * r.len.mean(), P


def _entry_rel(A, B):
    return float(np.max(np.abs(A - B) / np.abs(B)))


# ------------------------------------------------------------ dense solve --

@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n", [1, 7, 32, 33, 64, 65, 200, 517])
def test_zgesv_random(efie_mod, ctx, prec, n):
    rng = np.random.default_rng(n)
    nb = 3
    A = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n))
    b = rng.standard_normal((nb, n)) + 1j * rng.standard_normal((nb, n))
    x = ctx.zgesv(A, b, prec)
    xr = np.linalg.solve(A, b[..., None])[..., 0]
    cond = max(np.linalg.cond(a) for a in A)
    assert np.abs(x - xr).max() <= 1e-14 * cond * np.abs(xr).max() * max(n, 10)


def test_zgesv_known_answers(efie_mod, ctx):
    n = 70
    I = np.eye(n, dtype=complex)
    b = np.arange(n) + 1j
    assert np.array_equal(ctx.zgesv(I, b), b)
    Pm = I[np.random.default_rng(3).permutation(n)]
    assert np.array_equal(ctx.zgesv(Pm, b), Pm.T @ b)
    D = np.diag(np.arange(1, n + 1) * (1 + 1j))
    assert np.allclose(ctx.zgesv(D, b), b / np.diag(D), rtol=1e-15, atol=0)
    S = np.ones((n, n), dtype=complex)
    with pytest.raises(efie_mod.EfieError) as e:
        ctx.zgesv(S, b)
    assert e.value.ident == "efie:singular"


# ---------------------------------------------------------------- the fill --

OPTSETS = [
    dict(quad_order=4, near_sing_eps=1.2, near_subdiv=1, near_order=2),
    dict(quad_order=7, near_sing_eps=1.2, near_subdiv=2, near_order=4),
    dict(quad_order=2, near_sing_eps=2.5, near_subdiv=2, near_order=4),
    dict(quad_order=1, near_sing_eps=0.5, near_subdiv=1, near_order=1, self_subdiv=1),
    dict(quad_order=7, near_sing_eps=2.5, near_subdiv=3, near_order=7, self_subdiv=3),
]


@pytest.mark.parametrize("near_mode", [0, 1])
@pytest.mark.parametrize("ok", range(len(OPTSETS)))
def test_full_z_tiny(efie_mod, orc, ctx, ok, near_mode):
    V, F, r, pe, ge, spread, P = _setup(efie_mod, orc, "tiny")
    ctx.set_problem(V, F, r)
    o = dict(W.OPTS_C1)
    o.update(OPTSETS[ok])
    w = 2 * np.pi * 2.5e9
    eps = W.EPS0 * (2.2 - 0.01j)
    k = orc.wavenumber(w, W.MU0, eps)
    Zg, Vg = ctx.assemble_zv(k, w, W.MU0, eps, [pe, ge, 1.0, spread, 0.2],
                             efie_mod.opts(**o, near_mode=near_mode))
    Zo = P.z_full(k, w, W.MU0, eps, orc.opts(**o))
    assert _entry_rel(Zg, Zo) <= ZTOL
    assert np.abs(Vg - P.rhs(pe, ge, 1.0, spread)).max() <= 1e-15


def test_centroid_model(efie_mod, orc, ctx):
    V, F, r, pe, ge, spread, P = _setup(efie_mod, orc, "tiny")
    ctx.set_problem(V, F, r)
    w = 2 * np.pi * 2e9
    eps = W.EPS0 * (1 - 1e-3j)
    k = orc.wavenumber(w, W.MU0, eps)
    for alpha in (0.2, 0.35):
        Zg, _ = ctx.assemble_zv(k, w, W.MU0, eps, [pe, ge, 1.0, spread, alpha], efie_mod.opts("centroid"))
        Zo = P.z_centroid(k, 1j * w * W.MU0, 1 / (1j * w * eps), alpha)
        assert _entry_rel(Zg, Zo) <= ZTOL


@pytest.mark.parametrize("near_mode", [0, 1])
def test_c1_rows_vs_golden(efie_mod, orc, ctx, near_mode):
    g = np.load(W_GOLDEN("c1.npz"))
    V, F, r, pe, ge, spread, P = _setup(efie_mod, orc, "c1")
    assert r.Ne == int(g["Ne"]) and (pe, ge) == tuple(g["ports"])
    for k in ("edge", "plusTri", "minusTri", "plusSign", "minusSign"):
        assert np.array_equal(getattr(r, k), g[k])
    ctx.set_problem(V, F, r)
    w = 2 * np.pi * g["freqs"][0]
    k = orc.wavenumber(w, W.MU0, W.EPS0)
    Zg, Vg = ctx.assemble_zv(k, w, W.MU0, W.EPS0, [pe, ge, 1.0, spread, 0.2],
                             efie_mod.opts(**W.OPTS_C1, near_mode=near_mode))
    rows = g["rows"] - 1
    assert _entry_rel(Zg[rows], g["zrows"]) <= ZTOL
    assert np.abs(Vg - g["vrhs"]).max() <= 1e-15
    # complex symmetry of the whole matrix
    assert np.abs(Zg - Zg.T).max() <= 1e-13 * np.abs(Zg).max()


# ---------------------------------------------------------------- the sweep --

def W_GOLDEN(name):
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)


# Whole Z at the c1 mesh (Ne = 584) against the oracle's whole Z
# (assemble_ZV_EFIE_RWG_DIFF_CPU.m:9-36): one lossy and one lossless
# frequency, every near-pair mode (EXTRACT against the oracle's extraction),
# and the Gaussian excitation at two spreads.
C1_FULL_W = 2 * np.pi * 2.5e9
C1_FULL_EPS = {"lossy": W.EPS0 * (2.2 - 0.01j), "lossless": W.EPS0 * 2.2}


@pytest.fixture(scope="module")
def c1_full(efie_mod, orc):
    V, F, r, pe, ge, spread, P = _setup(efie_mod, orc, "c1")
    out = {}
    for name, eps in C1_FULL_EPS.items():
        k = orc.wavenumber(C1_FULL_W, W.MU0, eps)
        for ex in ((0, 1) if name == "lossy" else (0,)):
            out[name, ex] = P.z_full(k, C1_FULL_W, W.MU0, eps, orc.opts(**W.OPTS_C1, extract=ex))
    return (V, F, r, pe, ge, spread, P), out


@pytest.mark.parametrize("case", [("lossy", 0), ("lossy", 1), ("lossy", 2), ("lossless", 0), ("lossless", 1)])
def test_c1_full_z_vs_oracle(efie_mod, orc, ctx, c1_full, case):
    (V, F, r, pe, ge, spread, P), Zo = c1_full
    name, near_mode = case
    eps = C1_FULL_EPS[name]
    k = orc.wavenumber(C1_FULL_W, W.MU0, eps)
    ctx.set_problem(V, F, r)
    for sp in (spread, 2.5 * spread):
        Zg, Vg = ctx.assemble_zv(k, C1_FULL_W, W.MU0, eps, [pe, ge, 1.0, sp, 0.2],
                                 efie_mod.opts(**W.OPTS_C1, near_mode=near_mode))
        assert np.abs(Vg - P.rhs(pe, ge, 1.0, sp)).max() <= 1e-15
    assert _entry_rel(Zg, Zo[name, 1 if near_mode == efie_mod.NEAR_EXTRACT else 0]) <= ZTOL
    assert np.array_equal(Zg, Zg.T)


@pytest.mark.parametrize("prec,near_mode,batch", [(0, 0, 0), (0, 1, 2), (1, 1, 0), (1, 0, 5)])
def test_c1_sweep_vs_golden(efie_mod, orc, ctx, prec, near_mode, batch):
    g = np.load(W_GOLDEN("c1.npz"))
    V, F, r, pe, ge, spread, _ = _setup(efie_mod, orc, "c1")
    ctx.set_problem(V, F, r)
    o = efie_mod.opts(**W.OPTS_C1, precision=prec, near_mode=near_mode, batch=batch)
    zin, s11, st = ctx.sweep(g["freqs"], W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o, stats=True)
    assert np.max(np.abs(zin - g["zin"]) / np.abs(g["zin"])) <= ZIN_TOL
    assert np.max(np.abs(s11 - g["s11"])) <= ZIN_TOL
    if prec == 1:
        # refinement reaches the FP64 floor of ||V - Z x|| / ||V|| (~1e-12 here)
        assert st["ir_resid_max"] <= 1e-11 and st["ir_iters_max"] <= 8


@pytest.mark.parametrize("prec", [0, 1])
def test_c1_centroid_sweep_vs_golden(efie_mod, orc, ctx, prec):
    """Centroid model, FP64 and mixed.  The centroid Z is symmetric only to
    rounding, so the mixed path factorises its lower triangle (LDL^T) and
    refines against the full matrix: Zin still reaches the FP64 result."""
    g = np.load(W_GOLDEN("c1.npz"))
    V, F, r, pe, ge, spread, _ = _setup(efie_mod, orc, "c1")
    ctx.set_problem(V, F, r)
    o = efie_mod.opts("centroid", precision=prec)
    zin, s11 = ctx.sweep(g["freqs"], W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o)
    assert np.max(np.abs(zin - g["zin_centroid"]) / np.abs(g["zin_centroid"])) <= ZIN_TOL
    assert np.max(np.abs(s11 - g["s11_centroid"])) <= ZIN_TOL


@pytest.mark.parametrize("cells", [(1, 1, 1, 1), (1, 2, 1, 1), (2, 1, 2, 2), (2, 2, 2, 2), (2, 2, 3, 2)])
def test_small_meshes(efie_mod, orc, ctx, cells):
    """Ne = 2, 4, 11, 16 and 21: Z of one or two partial 16 x 16 tiles, fewer
    source triangles than one regular-fill block takes, a symmetric
    factorisation narrower than one 32-column panel.  Full Z in both near
    modes against the oracle (and exactly symmetric), the mixed sweep
    against the FP64 sweep."""
    V, F = efie_mod.make_patch_with_ground_mesh(*W.GEOM, *cells)
    r = efie_mod.build_rwg_interior(V, F)
    pP, pG = W.feed_points("micro")
    pe = efie_mod.pick_port_edge_by_point_and_layer(r, pP, 1)
    ge = efie_mod.pick_port_edge_by_point_and_layer(r, pG, -1)
    P = orc.Problem(V, F, orc.build_rwg(V, F))
    spread = 1.5 * r.len.mean()
    ctx.set_problem(V, F, r)
    w = 2 * np.pi * 2.5e9
    k = orc.wavenumber(w, W.MU0, W.EPS0)
    Zo = P.z_full(k, w, W.MU0, W.EPS0, orc.opts(**W.OPTS_C1))
    for nm in (0, 1):
        Zg, _ = ctx.assemble_zv(k, w, W.MU0, W.EPS0, [pe, ge, 1.0, spread, 0.2],
                                efie_mod.opts(**W.OPTS_C1, near_mode=nm))
        assert _entry_rel(Zg, Zo) <= ZTOL
        assert np.array_equal(Zg, Zg.T)
    fr = np.linspace(1.5e9, 3.5e9, 7)
    z64 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**W.OPTS_C1, precision=0))[0]
    zmx = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                    efie_mod.opts(**W.OPTS_C1, precision=1, near_mode=1))[0]
    assert np.max(np.abs(zmx - z64) / np.abs(z64)) <= ZIN_TOL


def test_lossy_sweep_tiny_vs_oracle(efie_mod, orc, ctx):
    V, F, r, pe, ge, spread, P = _setup(efie_mod, orc, "tiny")
    ctx.set_problem(V, F, r)
    fr = W.freqs("tiny")
    eps = W.EPS0 * 2.59 * (1 - 1e-2j)
    for prec, nm in ((0, 0), (1, 1)):
        zin, s11 = ctx.sweep(fr, W.MU0, eps, pe, ge, 1.0, spread, 50.0,
                             efie_mod.opts(**W.OPTS_C1, precision=prec, near_mode=nm))
        zo, so = P.sweep(fr, W.MU0, eps, pe, ge, 1.0, spread, 50.0, orc.opts(**W.OPTS_C1))
        assert np.max(np.abs(zin - zo) / np.abs(zo)) <= ZIN_TOL


def test_sweep_errors(efie_mod, orc, ctx):
    V, F, r, pe, ge, spread, _ = _setup(efie_mod, orc, "micro")
    ctx.set_problem(V, F, r)
    fr = W.freqs("micro")
    with pytest.raises(efie_mod.EfieError) as e:
        ctx.sweep(fr, W.MU0, W.EPS0, r.Ne + 1, ge, 1.0, spread, 50.0)
    assert e.value.ident == "mex:port"
    with pytest.raises(efie_mod.EfieError) as e:
        ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, 0.0, 50.0)
    assert e.value.ident == "mex:rhs"
    zin, s11 = ctx.sweep(np.array([]), W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0)
    assert zin.size == 0
    # refinement options out of range are rejected before any work
    for bad in (dict(ir_max_iter=-1), dict(ir_tol=0.0), dict(ir_tol=float("nan")), dict(ir_tol=-1e-10)):
        with pytest.raises(efie_mod.EfieError) as e:
            ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(precision=1, **bad))
        assert e.value.ident == "efie:opts"


@pytest.mark.parametrize("lu_kernel", [1, 3])
def test_stalled_refinement_falls_back_to_fp64(efie_mod, orc, ctx, lu_kernel):
    """A mixed sweep whose refinement cannot reach max(ir_tol, 1e-11) (here:
    no correction allowed) returns the FP64 LU's results for those
    frequencies, never the FP32 ones: the symmetric form goes through the
    pivoted FP32 LU first (n_lu_retry), the pivoted form straight to FP64
    (n_fp64_retry)."""
    V, F, r, pe, ge, spread, _ = _setup(efie_mod, orc, "c1")
    ctx.set_problem(V, F, r)
    fr = W.freqs("c1")
    z64 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**W.OPTS_C1))[0]
    zm, _, st = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                          efie_mod.opts(**W.OPTS_C1, precision=1, lu_kernel=lu_kernel, ir_max_iter=0,
                                        ir_tol=1e-12), stats=True)
    print("lu_kernel", lu_kernel, "retries", st["n_lu_retry"], st["n_fp64_retry"], flush=True)
    assert st["n_lu_retry"] == (len(fr) if lu_kernel == 3 else 0)
    assert st["n_fp64_retry"] == len(fr)
    assert np.max(np.abs(zm - z64) / np.abs(z64)) <= 1e-12


def test_zgesv_mixed_falls_back_on_ill_conditioned(efie_mod, ctx):
    """efie_zgesv MIXED on a system the FP32 factorisation cannot resolve
    (2-norm condition 1e10): the refinement stalls, the system is solved
    again by the FP64 LU, and x has FP64 backward error."""
    n = 300
    rng = np.random.default_rng(5)
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = (Q1 * np.logspace(0, -10, n)) @ Q2.conj().T
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, st = ctx.zgesv(A, b, precision=1, stats=True)
    assert st["n_fp64_retry"] == 1
    berr = np.linalg.norm(A @ x - b) / (np.linalg.norm(A, 2) * np.linalg.norm(x) + np.linalg.norm(b))
    assert berr <= 1e-13


# ---------------------------------------------------------- the MEX layer --

def test_mex_entry_points_on_device(efie_mod, orc):
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "mxshim"))
    import mx
    g = np.load(W_GOLDEN("c1.npz"))
    V, F, r, pe, ge, spread, P = _setup(efie_mod, orc, "c1")
    tables = [V, F, r.edge, r.plusTri, r.minusTri, r.plusSign, r.minusSign, r.rp, r.rm, r.Ap, r.Am, r.len]
    opts = {k: float(v) for k, v in W.OPTS_C1.items()}
    zin, s11 = mx.call("patch_sweep_cuda", 2, *tables, g["freqs"], W.MU0, complex(W.EPS0, 0), np.int32(pe),
                       np.int32(ge), 1.0, spread, 50.0, opts)
    assert np.max(np.abs(zin[:, 0] - g["zin"]) / np.abs(g["zin"])) <= ZIN_TOL
    # centroid model of the reference MEX when opts.model = 1 (20-argument call uses quadrature)
    zc, _ = mx.call("patch_sweep_cuda", 2, *tables, g["freqs"], W.MU0, W.EPS0, np.int32(pe), np.int32(ge), 1.0,
                    spread, 50.0, {"model": 1.0})
    assert np.max(np.abs(zc[:, 0] - g["zin_centroid"]) / np.abs(g["zin_centroid"])) <= ZIN_TOL
    w = 2 * np.pi * g["freqs"][0]
    k = w * np.sqrt(W.MU0 * W.EPS0)
    Z, Vr = mx.call("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, *tables, complex(k, 0), w, W.MU0, W.EPS0,
                    np.array([pe, ge, 1.0, spread, 0.2]), opts)
    assert _entry_rel(Z[g["rows"] - 1], g["zrows"]) <= ZTOL
    assert np.abs(Vr[:, 0] - g["vrhs"]).max() <= 1e-15


def test_mex_calls_reuse_geometry(efie_mod, orc):
    """MEX callers pass the same tables on every call: the second identical
    call rebuilds neither the geometry nor the near/self moments (the
    process-wide build counters do not move) and gives the same Z bit for
    bit; a changed table rebuilds."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "mxshim"))
    import mx
    V, F, r, pe, ge, spread, P = _setup(efie_mod, orc, "c1")
    tables = [V, F, r.edge, r.plusTri, r.minusTri, r.plusSign, r.minusSign, r.rp, r.rm, r.Ap, r.Am, r.len]
    opts = {k: float(v) for k, v in W.OPTS_C1.items()}
    opts["near_mode"] = 1.0
    w = 2 * np.pi * 2.5e9
    k = w * np.sqrt(W.MU0 * W.EPS0)
    args = (complex(k, 0), w, W.MU0, W.EPS0, np.array([pe, ge, 1.0, spread, 0.2]), opts)
    Z1, _ = mx.call("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, *tables, *args)
    c1 = efie_mod.Context.cache_counts()
    Z2, _ = mx.call("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, *tables, *args)
    assert efie_mod.Context.cache_counts() == c1
    assert np.array_equal(Z1, Z2)
    # prepack entry point, twice
    pk = _c1_prepack(V, F)
    fr = W.freqs("c1")
    zin1 = mx.call("patch_sweep_fastquad_cuda", 2, *tables, pk, fr, W.MU0, complex(W.EPS0, 0), float(pe),
                   float(ge), 1.0, spread, 50.0)[0]
    c2 = efie_mod.Context.cache_counts()
    zin2 = mx.call("patch_sweep_fastquad_cuda", 2, *tables, pk, fr, W.MU0, complex(W.EPS0, 0), float(pe),
                   float(ge), 1.0, spread, 50.0)[0]
    assert efie_mod.Context.cache_counts() == c2
    assert np.array_equal(zin1, zin2)
    # a moved vertex is a new mesh
    V2 = V.copy()
    V2[0, 0] += 1e-6
    mx.call("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, V2, *tables[1:], *args)
    assert efie_mod.Context.cache_counts()[0] > c2[0]


def _c1_prepack(V, F, **kw):
    import prepack_np
    o = dict(W.OPTS_C1)
    keys = ("quad_order", "near_order", "near_sing_eps", "self_subdiv", "self_alpha", "near_subdiv", "near_alpha")
    a = {k: o[k] for k in keys}
    a.update(kw)
    return prepack_np.make_prepack(V, F, **a)


def test_prepack_entry_points_on_device(efie_mod, orc):
    """patch_sweep_fastquad_cuda, ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST and the
    8-input assemble_ZV_EFIE_RWG_DIFF_GPU_FAST, given the prepack that
    efie_precompute_pack.m makes for the C1 options (restated in
    tests/prepack_np.py): Zin/S11, Z rows and Vrhs against the golden
    vectors, FP64 and mixed."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "mxshim"))
    import mx
    g = np.load(W_GOLDEN("c1.npz"))
    V, F, r, pe, ge, spread, _ = _setup(efie_mod, orc, "c1")
    pk = _c1_prepack(V, F)
    tables = [V, F, r.edge, r.plusTri, r.minusTri, r.plusSign, r.minusSign, r.rp, r.rm, r.Ap, r.Am, r.len]
    for extra in ([], [{"precision": 1.0}]):
        zin, s11 = mx.call("patch_sweep_fastquad_cuda", 2, *tables, pk, g["freqs"], W.MU0, complex(W.EPS0, 0),
                           float(pe), float(ge), 1.0, spread, 50.0, *extra)
        assert np.max(np.abs(zin[:, 0] - g["zin"]) / np.abs(g["zin"])) <= ZIN_TOL
        assert np.max(np.abs(s11[:, 0] - g["s11"])) <= ZIN_TOL
    w = 2 * np.pi * g["freqs"][0]
    k = w * np.sqrt(W.MU0 * W.EPS0)
    Z, Vr = mx.call("ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, *tables, pk, complex(k, 0), w, W.MU0, W.EPS0,
                    np.array([pe, ge, 1.0, spread]))
    assert _entry_rel(Z[g["rows"] - 1], g["zrows"]) <= ZTOL
    assert np.abs(Vr[:, 0] - g["vrhs"]).max() <= 1e-15
    rwg = dict(plusTri=r.plusTri, minusTri=r.minusTri, plusSign=r.plusSign, minusSign=r.minusSign, len=r.len,
               Ap=r.Ap, Am=r.Am, rp=r.rp, rm=r.rm, center=r.center, zSign=r.zSign.astype(np.float64))
    exc = dict(patchEdge=float(pe), groundEdge=float(ge), V0=1.0, spreadRadius=spread)
    Z8, V8 = mx.call("assemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, rwg, pk, complex(k, 0), w, W.MU0, W.EPS0, exc,
                     {"near_alpha": W.OPTS_C1["near_alpha"]})
    assert _entry_rel(Z8[g["rows"] - 1], g["zrows"]) <= ZTOL
    assert np.abs(V8[:, 0] - g["vrhs"]).max() <= 1e-15
    # the self softening under the field name the reference source reads
    pk2 = dict(pk)
    pk2["self_aSoft"] = pk2.pop("self_a")
    Z8b, _ = mx.call("assemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, rwg, pk2, complex(k, 0), w, W.MU0, W.EPS0, exc,
                     {"near_alpha": W.OPTS_C1["near_alpha"]})
    assert np.array_equal(Z8b, Z8)


def test_prepack_drives_the_fill(efie_mod, orc, ctx):
    """The pack's data define Z, not the rule options: a pack made with
    near_alpha 0.3 equals the rules at 0.3 (and not the opts' 0.15); a pack
    with an empty near matrix equals the rules with near_sing_eps 0 (every
    pair of distinct triangles regular); MOMENTS with a pack equals DIRECT
    with it; the sweep with the pack equals the sweep with the rules;
    efie_set_prepack(NULL) returns to the rules; extraction is refused."""
    V, F, r, pe, ge, spread, _ = _setup(efie_mod, orc, "c1")
    ctx.set_problem(V, F, r)
    o = dict(W.OPTS_C1)
    w = 2 * np.pi * 2.7e9
    k = w * np.sqrt(W.MU0 * W.EPS0)
    exc = [pe, ge, 1.0, spread, 0.2]

    def Z(**kw):
        return ctx.assemble_zv(k, w, W.MU0, W.EPS0, exc, efie_mod.opts(**kw))[0]
    base = Z(**o)
    ctx.set_prepack(_c1_prepack(V, F, near_alpha=0.3))
    ZA = Z(**o)
    ctx.set_prepack(None)
    assert _entry_rel(ZA, Z(**dict(o, near_alpha=0.3))) <= 1e-12
    assert _entry_rel(ZA, base) > 1e-6
    pk0 = _c1_prepack(V, F)
    pk0["nearMat"] = np.zeros_like(pk0["nearMat"])
    ctx.set_prepack(pk0)
    Z0 = Z(**o)
    ctx.set_prepack(None)
    assert _entry_rel(Z0, Z(**dict(o, near_sing_eps=0.0))) <= 1e-12
    assert _entry_rel(Z0, base) > 1e-6
    pk = _c1_prepack(V, F)
    ctx.set_prepack(pk)
    Zd = Z(**o)
    assert _entry_rel(Zd, base) <= 1e-12
    assert _entry_rel(Z(**dict(o, near_mode=efie_mod.NEAR_MOMENTS)), Zd) <= 1e-12
    with pytest.raises(efie_mod.EfieError) as e:
        Z(**dict(o, near_mode=efie_mod.NEAR_EXTRACT))
    assert e.value.ident == "efie:opts"
    fr = W.freqs("c1")
    zp = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**o, precision=1, near_mode=1))[0]
    ctx.set_prepack(None)
    zr = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**o))[0]
    assert np.max(np.abs(zp - zr) / np.abs(zr)) <= 1e-10
    assert _entry_rel(Z(**o), base) == 0.0


def test_meshless_problem(efie_mod, orc, ctx):
    """RWG tables without V/F/edges plus a prepack (the 8-input entry
    point's inputs): Z equals the mesh problem's with the same pack; the
    excitation and the sweep need the mesh arrays; without a pack the fill
    is refused."""
    V, F, r, pe, ge, spread, _ = _setup(efie_mod, orc, "c1")
    pk = _c1_prepack(V, F)
    o = efie_mod.opts(**W.OPTS_C1)
    w = 2 * np.pi * 2.2e9
    k = w * np.sqrt(W.MU0 * W.EPS0)
    ctx.set_problem(V, F, r)
    ctx.set_prepack(pk)
    Zm = ctx.assemble_z(k, w, W.MU0, W.EPS0, o)
    ctx.set_prepack(None)
    with efie_mod.Context(0) as c:
        c.set_rwg_tables(F.shape[0], r.plusTri, r.minusTri, r.plusSign, r.minusSign, r.rp, r.rm, r.Ap, r.Am, r.len)
        with pytest.raises(efie_mod.EfieError) as e:
            c.assemble_z(k, w, W.MU0, W.EPS0, o)
        assert e.value.ident == "efie:state"
        c.set_prepack(pk)
        Zl = c.assemble_z(k, w, W.MU0, W.EPS0, o)
        assert _entry_rel(Zl, Zm) <= 1e-12
        with pytest.raises(efie_mod.EfieError) as e:
            c.assemble_zv(k, w, W.MU0, W.EPS0, [pe, ge, 1.0, spread, 0.2], o)
        assert e.value.ident == "efie:state"
        with pytest.raises(efie_mod.EfieError) as e:
            c.sweep(W.freqs("c1"), W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o)
        assert e.value.ident == "efie:state"


@pytest.mark.parametrize("lu_kernel", [1, 2, 3])
def test_c1_sweep_lu_kernels(efie_mod, orc, ctx, lu_kernel):
    """The mixed path with each factorisation form (pivoted LU with the FP32
    MFMA update, with the split-bf16 update, or the unpivoted complex-symmetric
    factorisation) refines to the golden Zin in about as many steps."""
    g = np.load(W_GOLDEN("c1.npz"))
    V, F, r, pe, ge, spread, _ = _setup(efie_mod, orc, "c1")
    ctx.set_problem(V, F, r)
    zin, s11, st = ctx.sweep(g["freqs"], W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                             efie_mod.opts(**W.OPTS_C1, precision=1, lu_kernel=lu_kernel), stats=True)
    assert np.max(np.abs(zin - g["zin"]) / np.abs(g["zin"])) <= ZIN_TOL
    assert np.max(np.abs(s11 - g["s11"])) <= ZIN_TOL
    assert st["ir_resid_max"] <= 1e-11 and st["ir_iters_max"] <= 8
    ref = ctx.sweep(g["freqs"], W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                    efie_mod.opts(**W.OPTS_C1, precision=1, lu_kernel=1), stats=True)[2]
    # SYM starts the refinement from x = 0 (no rhs column through the
    # factorisation), one step more than the pivoted forms
    assert st["ir_iters_max"] <= ref["ir_iters_max"] + (2 if lu_kernel == 3 else 1)


def test_sym_factorisation_c2(efie_mod, ctx):
    """lu_kernel SYM at c2_6k on frequencies around the band where the
    unpivoted factorisation meets its smallest pivots (2.3-2.6 GHz, growth
    about 3e3): the LDL^T factor converges like the pivoted LU, and every
    Zin equals the pivoted path's to the refinement accuracy.  With the
    refinement capped at two steps no frequency reaches the tolerance, so
    every one is solved again with pivoting (and, where two pivoted steps do
    not reach it either, with the FP64 LU): the results keep the converged
    pivoted sweep's accuracy."""
    V, F = efie_mod.make_patch_with_ground_mesh(*W.mesh_args("c2_6k"))
    r = efie_mod.build_rwg_interior(V, F)
    pP, pG = W.feed_points("c2_6k")
    pe = efie_mod.pick_port_edge_by_point_and_layer(r, pP, 1)
    ge = efie_mod.pick_port_edge_by_point_and_layer(r, pG, -1)
    spread = 1.5 * r.len.mean()
    ctx.set_problem(V, F, r)
    fr = np.array([1.6e9, 2.3e9, 2.42e9, 2.5e9, 3.1e9])
    o = dict(W.OPTS_C1, precision=1, near_mode=1, ir_tol=1e-10)
    zp, _, sp = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**o, lu_kernel=1), stats=True)
    zs, _, st = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**o, lu_kernel=3), stats=True)
    print("pivoted", sp["ir_iters_max"], "sym", st["ir_iters_max"], "retries", st["n_lu_retry"], flush=True)
    assert st["n_lu_retry"] == 0
    assert st["ir_iters_max"] <= sp["ir_iters_max"] + 2
    assert st["ir_resid_max"] <= 1e-10
    assert np.max(np.abs(zs - zp) / np.abs(zp)) <= 1e-9
    # forced retry of every frequency
    o2 = dict(o, ir_max_iter=2)
    zs2, _, st2 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**o2, lu_kernel=3),
                            stats=True)
    print("forced: pivoted retries", st2["n_lu_retry"], "fp64 retries", st2["n_fp64_retry"], flush=True)
    assert st2["n_lu_retry"] == len(fr)
    assert np.max(np.abs(zs2 - zp) / np.abs(zp)) <= 1e-9


def test_c2_6k_scale_rows_and_zin(efie_mod, orc, ctx):
    """GPU scale (c2_6k, Ne = 6009): 32 sampled rows of Z at 2.5 GHz
    against the oracle's rows (direct and moment near modes, 1e-10
    entry-relative), and the mixed path's Zin against the FP64 path's at 5
    frequencies across the resonance band (1e-9)."""
    import os
    V, F = efie_mod.make_patch_with_ground_mesh(*W.mesh_args("c2_6k"))
    r = efie_mod.build_rwg_interior(V, F)
    pP, pG = W.feed_points("c2_6k")
    pe = efie_mod.pick_port_edge_by_point_and_layer(r, pP, 1)
    ge = efie_mod.pick_port_edge_by_point_and_layer(r, pG, -1)
    spread = 1.5 * r.len.mean()
    ctx.set_problem(V, F, r)
    w = 2 * np.pi * 2.5e9
    k = w * np.sqrt(W.MU0 * W.EPS0)
    rows = np.linspace(1, r.Ne, 32).astype(np.int32)
    P = orc.Problem(V, F, orc.build_rwg(V, F))
    Zo = P.z_rows(rows, orc.wavenumber(w, W.MU0, W.EPS0), w, W.MU0, W.EPS0, orc.opts(**W.OPTS_C1),
                  threads=min(os.cpu_count() or 1, 16))
    for nm in (efie_mod.NEAR_DIRECT, efie_mod.NEAR_MOMENTS):
        Z, _ = ctx.assemble_zv(k, w, W.MU0, W.EPS0, [pe, ge, 1.0, spread, 0.2],
                               efie_mod.opts(**W.OPTS_C1, near_mode=nm))
        assert _entry_rel(Z[rows - 1], Zo) <= ZTOL
        del Z
    fr = np.array([1.6e9, 2.3e9, 2.42e9, 2.5e9, 3.4e9])
    z64 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, efie_mod.opts(**W.OPTS_C1))[0]
    zm = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                   efie_mod.opts(**W.OPTS_C1, precision=1, near_mode=1, ir_tol=1e-10))[0]
    assert np.max(np.abs(zm - z64) / np.abs(z64)) <= 1e-9


def test_c4_family_small(efie_mod, orc, ctx):
    """Geometries of the batched uneven workload's family (rectangular
    cells, random layer gap and ground size), scaled down, against the
    oracle's sweep: FP64 and mixed paths, one context re-used across
    geometries of different N."""
    fr = np.linspace(1.0e9, 4.0e9, 3)
    for g in W.c4_geometries(n_geom=2, n_range=(120, 260), seed=11):
        V, F = efie_mod.make_patch_with_ground_mesh(*g["geom"], *g["cells"])
        r = efie_mod.build_rwg_interior(V, F)
        pP, pG = W.c4_feed_points(g["geom"])
        pe = efie_mod.pick_port_edge_by_point_and_layer(r, pP, 1)
        ge = efie_mod.pick_port_edge_by_point_and_layer(r, pG, -1)
        spread = 1.5 * r.len.mean()
        P = orc.Problem(V, F, orc.build_rwg(V, F))
        zo, so = P.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, orc.opts(**W.OPTS_C1))
        ctx.set_problem(V, F, r)
        for prec in (0, 1):
            zin, s11 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                                 efie_mod.opts(**W.OPTS_C1, precision=prec, near_mode=1))
            assert np.max(np.abs(zin - zo) / np.abs(zo)) <= ZIN_TOL
            assert np.max(np.abs(s11 - so)) <= ZIN_TOL


# ------------------------------------------- singularity extraction --

@pytest.mark.parametrize("ok", [0, 1, 2])
def test_extract_full_z_tiny(efie_mod, orc, ctx, ok):
    """near_mode EXTRACT (closed-form static part + smooth moment series)
    against the oracle's extraction (orc_pair_extract) on the whole Z."""
    V, F, r, pe, ge, spread, P = _setup(efie_mod, orc, "tiny")
    ctx.set_problem(V, F, r)
    o = dict(W.OPTS_C1)
    o.update(OPTSETS[ok])
    w = 2 * np.pi * 2.5e9
    eps = W.EPS0 * (2.2 - 0.01j)
    k = orc.wavenumber(w, W.MU0, eps)
    Zg, _ = ctx.assemble_zv(k, w, W.MU0, eps, [pe, ge, 1.0, spread, 0.2],
                            efie_mod.opts(**o, near_mode=efie_mod.NEAR_EXTRACT))
    Zo = P.z_full(k, w, W.MU0, eps, orc.opts(**o, extract=1))
    assert _entry_rel(Zg, Zo) <= ZTOL
    assert np.array_equal(Zg, Zg.T)


def test_extract_sweep_tiny(efie_mod, orc, ctx):
    V, F, r, pe, ge, spread, P = _setup(efie_mod, orc, "tiny")
    ctx.set_problem(V, F, r)
    fr = W.freqs("tiny")
    zo, so = P.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, orc.opts(**W.OPTS_C1, extract=1))
    for prec in (0, 1):
        zin, s11 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                             efie_mod.opts(**W.OPTS_C1, precision=prec, near_mode=efie_mod.NEAR_EXTRACT))
        assert np.max(np.abs(zin - zo) / np.abs(zo)) <= ZIN_TOL
        assert np.max(np.abs(s11 - so)) <= ZIN_TOL
