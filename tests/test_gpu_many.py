"""GPU tests of efie_sweep_many: several problems' sweeps in one pipelined
call, problems of similar order packed into shared batches (identity
padding to the batch order).

Every result is held to the one-problem sweep (efie_sweep) on the same
problem and options, and the padded and smallest problems also to the
oracle's own sweep (FP64 fill + LU).  The problems are six seeded patch
geometries of the batched-workload family with N from about 150 to 700, so
that sorting by order packs some of them with padding (a batch takes
problems down to 0.75 of its order) and leaves others alone.

Tolerances (relative, Zin and S11): FP64 paths 1e-10 against each other
(the padded factorisation groups its operations differently, so results
agree to rounding, not bit for bit); mixed paths, refined to a residual of
1e-12 (or re-solved when they stop above 1e-11), 1e-9 against each other
and 1e-10 against the oracle at a refinement tolerance of 1e-11.
"""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

FEW = dict(n_geom=6, n_range=(150, 700), seed=20261204)


def _problems(efie_mod):
    out = []
    for i, g in enumerate(W.c4_geometries(**FEW)):
        V, F = efie_mod.make_patch_with_ground_mesh(*g["geom"], *g["cells"])
        r = efie_mod.build_rwg_interior(V, F)
        pP, pG = W.c4_feed_points(g["geom"])
        pe = efie_mod.pick_port_edge_by_point_and_layer(r, pP, 1)
        ge = efie_mod.pick_port_edge_by_point_and_layer(r, pG, -1)
        # each problem its own frequencies (count and band)
        fr = np.linspace(1.0e9 + 0.1e9 * i, 3.5e9 + 0.08e9 * i, 3 + i)
        out.append(dict(V=V, F=F, r=r, pe=pe, ge=ge, spread=1.5 * r.len.mean(), freqs=fr))
    return out


@pytest.fixture(scope="module")
def probs(efie_mod):
    ps = _problems(efie_mod)
    ctxs = []
    for p in ps:
        c = efie_mod.Context(0)
        c.set_problem(p["V"], p["F"], p["r"])
        p["ctx"] = c
        ctxs.append(c)
    ns = sorted(p["r"].Ne for p in ps)
    print("problem orders:", ns, flush=True)
    yield ps
    for c in ctxs:
        c.close()


@pytest.fixture(scope="module")
def engine(efie_mod):
    c = efie_mod.Context(0)
    yield c
    c.close()


def _items(ps, **kw):
    return [dict(ctx=p["ctx"], freqs=p["freqs"], patchEdge=p["pe"], groundEdge=p["ge"], V0=1.0,
                 spread=p["spread"], Z0=50.0, **kw) for p in ps]


def _single(p, o):
    return p["ctx"].sweep(p["freqs"], W.MU0, W.EPS0, p["pe"], p["ge"], 1.0, p["spread"], 50.0, o)


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def test_packing_pads_some_problems(probs):
    """The problem set exercises a shared batch with padding: two problems
    within 0.75 of each other's order, the smaller one padded."""
    ns = sorted((p["r"].Ne for p in probs), reverse=True)
    assert any(0.75 * ns[i] <= ns[i + 1] < ns[i] for i in range(len(ns) - 1)), ns


OPTION_SETS = {
    "mixed_sym": dict(W.OPTS_C1, precision=1, near_mode=1, ir_tol=1e-12),
    "mixed_pivoted": dict(W.OPTS_C1, precision=1, near_mode=1, ir_tol=1e-12, lu_kernel=1),
    "fp64_series": dict(W.OPTS_C1, precision=0, near_mode=1),
    "fp64_direct": dict(W.OPTS_C1, precision=0, near_mode=0),
    "mixed_direct": dict(W.OPTS_C1, precision=1, near_mode=0, ir_tol=1e-12),
    "centroid_fp64": dict(model=1),
    "centroid_mixed": dict(model=1, precision=1, ir_tol=1e-12),
}


@pytest.mark.parametrize("name", list(OPTION_SETS))
def test_many_equals_single_sweeps(efie_mod, engine, probs, name):
    """One efie_sweep_many call over the six problems against efie_sweep
    of each problem alone, for every path: symmetric and pivoted mixed
    factorisation, FP64, per-frequency and series fills, centroid model."""
    kw = OPTION_SETS[name]
    preset = "centroid" if kw.get("model") == 1 else "accurate"
    o = efie_mod.opts(preset, **kw)
    res, st = efie_mod.sweep_many(engine, _items(probs), W.MU0, W.EPS0, o, stats=True)
    worst = 0.0
    for p, (zin, s11) in zip(probs, res):
        z1, s1 = _single(p, o)
        assert np.all(np.isfinite(zin))
        worst = max(worst, _rel(zin, z1), _rel(s11, s1))
    print(f"{name}: max rel diff vs single sweeps {worst:.2e}; ir_iters_max {st['ir_iters_max']}, "
          f"retries {st['n_lu_retry']}/{st['n_fp64_retry']}", flush=True)
    assert worst <= (1e-9 if kw.get("precision") == 1 else 1e-10)
    if kw.get("precision") == 1:
        assert st["ir_resid_max"] <= max(kw["ir_tol"], 1e-11)


def test_many_lossy_medium(efie_mod, engine, probs):
    """A lossy medium (complex eps: every k complex) through the packed call,
    mixed symmetric path with the series fill, against the one-problem
    sweeps."""
    eps = W.EPS0 * (2.2 - 0.02j)
    o = efie_mod.opts(**W.OPTS_C1, precision=efie_mod.PREC_MIXED, near_mode=efie_mod.NEAR_MOMENTS, ir_tol=1e-12)
    res = efie_mod.sweep_many(engine, _items(probs), W.MU0, eps, o)
    worst = 0.0
    for p, (zin, _) in zip(probs, res):
        z1 = p["ctx"].sweep(p["freqs"], W.MU0, eps, p["pe"], p["ge"], 1.0, p["spread"], 50.0, o)[0]
        worst = max(worst, _rel(zin, z1))
    print(f"lossy: max rel diff vs single sweeps {worst:.2e}", flush=True)
    assert worst <= 1e-9


def test_many_vs_oracle(efie_mod, orc, engine, probs):
    """The smallest problem and the one padded in a shared batch against the
    oracle's sweep, through the mixed headline path (series fill, FP32
    LDL^T, FP64 refinement)."""
    ns = [p["r"].Ne for p in probs]
    srt = sorted(range(len(ns)), key=lambda i: -ns[i])
    padded = next(srt[k + 1] for k in range(len(srt) - 1) if 0.75 * ns[srt[k]] <= ns[srt[k + 1]] < ns[srt[k]])
    smallest = srt[-1]
    o = efie_mod.opts(**W.OPTS_C1, precision=efie_mod.PREC_MIXED, near_mode=efie_mod.NEAR_MOMENTS, ir_tol=1e-11)
    res = efie_mod.sweep_many(engine, _items(probs), W.MU0, W.EPS0, o)
    for i in (padded, smallest):
        p = probs[i]
        P = orc.Problem(p["V"], p["F"], orc.build_rwg(p["V"], p["F"]))
        zo = P.sweep(p["freqs"], W.MU0, W.EPS0, p["pe"], p["ge"], 1.0, p["spread"], 50.0, orc.opts(**W.OPTS_C1))[0]
        e = _rel(res[i][0], zo)
        print(f"problem {i} (N={ns[i]}): |dZin|/|Zin| vs oracle {e:.2e}", flush=True)
        assert e <= 1e-10


def test_many_retries_through_fallbacks(efie_mod, engine, probs):
    """Refinement capped at one correction: the symmetric factorisation's
    frequencies all go to the pivoted FP32 LU, and those it leaves above
    1e-11 to the FP64 LU (each again as a packed call); the results agree
    with the FP64 path."""
    o = efie_mod.opts(**W.OPTS_C1, precision=efie_mod.PREC_MIXED, near_mode=efie_mod.NEAR_MOMENTS,
                      ir_tol=1e-12, ir_max_iter=1)
    res, st = efie_mod.sweep_many(engine, _items(probs), W.MU0, W.EPS0, o, stats=True)
    o64 = efie_mod.opts(**W.OPTS_C1, near_mode=efie_mod.NEAR_MOMENTS)
    nf = sum(len(p["freqs"]) for p in probs)
    print(f"retries: {st['n_lu_retry']} pivoted of {nf}, {st['n_fp64_retry']} FP64", flush=True)
    assert st["n_lu_retry"] == nf
    for p, (zin, _) in zip(probs, res):
        assert _rel(zin, _single(p, o64)[0]) <= 1e-9


def test_many_shared_context_and_release(efie_mod, engine, probs):
    """One problem context in two items (other ports and frequencies), next
    to a second problem; with release set the shared context's series is
    freed after its last batch (the next sweep rebuilds it), without it the
    series stays cached."""
    a, b = probs[0], probs[3]
    r = a["r"]
    pe2 = efie_mod.pick_port_edge_by_point_and_layer(r, np.array([0.0, 0.0, 10.0]), 1)
    items = _items([a, b])
    items.append(dict(ctx=a["ctx"], freqs=np.array([1.7e9, 2.9e9]), patchEdge=pe2, groundEdge=a["ge"], V0=2.0,
                      spread=a["spread"], Z0=75.0))
    o = efie_mod.opts(**W.OPTS_C1, precision=efie_mod.PREC_MIXED, near_mode=efie_mod.NEAR_MOMENTS, ir_tol=1e-12)
    for release in (1, 0):
        for it in items:
            it["release"] = release
        res = efie_mod.sweep_many(engine, items, W.MU0, W.EPS0, o)
        for it, (zin, s11) in zip(items, res):
            z1, s1 = it["ctx"].sweep(it["freqs"], W.MU0, W.EPS0, it["patchEdge"], it["groundEdge"], it["V0"],
                                     it["spread"], it["Z0"], o)
            assert _rel(zin, z1) <= 1e-9 and _rel(s11, s1) <= 1e-9
        # the single sweeps above ran after the call: the first rebuilt
        # the released series, so ask again
        efie_mod.sweep_many(engine, items, W.MU0, W.EPS0, o)
        _, _, st = a["ctx"].sweep(a["freqs"], W.MU0, W.EPS0, a["pe"], a["ge"], 1.0, a["spread"], 50.0, o,
                                  stats=True)
        if release:
            assert st["ms_moments"] > 0
        else:
            assert st["ms_moments"] == 0


def test_many_errors(efie_mod, engine, probs):
    """Argument errors name the item and keep the MEX identifiers; an empty
    call and items without frequencies are fine."""
    o = efie_mod.opts(**W.OPTS_C1, precision=efie_mod.PREC_MIXED, near_mode=efie_mod.NEAR_MOMENTS)
    assert efie_mod.sweep_many(engine, [], W.MU0, W.EPS0, o) == []
    items = _items(probs[:2])
    items[1]["patchEdge"] = probs[1]["r"].Ne + 1
    with pytest.raises(efie_mod.EfieError) as e:
        efie_mod.sweep_many(engine, items, W.MU0, W.EPS0, o)
    assert e.value.ident == "mex:port"
    items = _items(probs[:2])
    items[0]["freqs"] = np.zeros(0)
    res = efie_mod.sweep_many(engine, items, W.MU0, W.EPS0, o)
    assert res[0][0].shape == (0,) and np.all(np.isfinite(res[1][0]))
    items = _items(probs[:1])
    items[0]["spread"] = 0.0
    with pytest.raises(efie_mod.EfieError) as e:
        efie_mod.sweep_many(engine, items, W.MU0, W.EPS0, o)
    assert e.value.ident == "mex:rhs"
