"""Throughput benchmark of the EFIE frequency sweep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2_6k]

One step is one full frequency sweep of the configuration (fill Z, LU
solve, Zin/S11 at every frequency) with the mesh and RWG tables already
resident in HBM.  Two paths are timed:
  fp64  the plain method: per-frequency quadrature of every triangle pair
        (as the reference does) and an FP64 LU;
  fast  the headline path: every triangle pair from its frequency-independent
        series (integrated once per mesh, by the first sweep; Z of a whole
        batch of frequencies formed in one pass), the complex-symmetric Z
        factored in FP32 on the matrix cores without pivoting (frequencies
        where that does not refine are redone with the pivoted FP32 LU, then
        FP64) and FP64 iterative refinement;
  fp64 series fill  the fast path's fill with the FP64 LU: the share of the
        speed-up that the precision alone gives;
  surrogate  as fast, with the near pairs from the learned surrogate (trained
        on the GPU first, outside the timed sweeps).
`value` is the fast path's frequency points per second; the JSON line also
carries the fp64 path's rate, the largest |dZin|/|Zin| between the two, a
roofline entry for the dominant kernel (from a profiled extra sweep) and the
CPU baseline (the oracle on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "matlab-cuda-efief-vs-cpu_amd"))

import efie  # noqa: E402
import workloads as W  # noqa: E402

PEAK = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense TFLOP/s (FP64 matrix = vector)
PEAK_HBM_GBS = 8000.0
FLOP_PER_GREEN = 48   # per Green-function evaluation incl. its moment updates
NB = 32               # LU panel width of the library


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(config):
    V, F = efie.make_patch_with_ground_mesh(*W.mesh_args(config))
    r = efie.build_rwg_interior(V, F)
    pP, pG = W.feed_points(config)
    pe = efie.pick_port_edge_by_point_and_layer(r, pP, 1)
    ge = efie.pick_port_edge_by_point_and_layer(r, pG, -1)
    return V, F, r, pe, ge, 1.5 * r.len.mean(), W.freqs(config)


def timed(ctx, o, fr, pe, ge, spread, steps, warmup, name, first=None):
    """Time `steps` sweeps after `warmup` untimed ones; `first` (a dict)
    receives the stats of the first warmup sweep, which carries the one-off
    per-mesh work (the frequency-independent series) that later sweeps of
    the same mesh reuse."""
    for i in range(warmup):
        st0 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o, stats=True)[2]
        if i == 0 and first is not None:
            first.update(st0)
        log(f"[{name}] warmup {i + 1}/{warmup} done")
    t0 = time.perf_counter()
    out = None
    for i in range(steps):
        out = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o, stats=True)
        log(f"[{name}] step {i + 1}/{steps} done")
    dt = time.perf_counter() - t0  # each sweep returns after its stream has drained
    return dt, out


WB = 8 * NB           # outer block of the LU (trailing update with K = WB)


def lu_gemm_flops(n):
    """Flops of every matrix-core update launched by one LU of size n (rhs
    column included), following the library's blocking: NB-wide sub-panels
    inside WB-wide outer blocks."""
    f = 0.0

    def g(r0, r1, c0, c1, K):
        return 8.0 * max(0, r1 - r0) * max(0, c1 - c0) * K if K > 0 else 0.0
    for j0 in range(0, n, WB):
        jend = min(n, j0 + WB)
        for c0 in range(j0, jend, NB):
            w = min(NB, jend - c0)
            f += g(c0 + w, n, c0 + w, jend, w)           # update inside the block
        for c0 in range(j0, jend, NB):
            w = min(NB, jend - c0)
            f += g(c0 + w, jend, jend, n + 1, w)         # U12 of the block rows
        f += g(jend, n, jend, n + 1, jend - j0)          # trailing update
    return f


def lu_gemm_flops_sym(n, tile=64):
    """Flops of the matrix-core updates of the unpivoted complex-symmetric
    factorisation of size n (no rhs column), on the 64 x 64 tiles the kernel
    computes: every update skips the tiles strictly above the diagonal
    (tile rows rb..rb+63 all above tile columns cb.., i.e. rb + 64 <= cb).
    Inside an outer block the columns are taken by halves split at a
    multiple of NB: left half, the right half's update with K = the left
    half's width, right half."""
    f = 0.0

    def g_lower(r0, r1, c0, c1, K):
        t = 0.0
        for rb in range(r0, r1, tile):
            for cb in range(c0, c1, tile):
                if rb + tile > cb:
                    t += 8.0 * (min(r1, rb + tile) - rb) * (min(c1, cb + tile) - cb) * K
        return t

    def halves(c0, c1):
        if c1 - c0 <= NB:
            return 0.0
        mid = c0 + (c1 - c0 + 2 * NB - 1) // (2 * NB) * NB
        return halves(c0, mid) + g_lower(mid, n, mid, c1, mid - c0) + halves(mid, c1)
    for j0 in range(0, n, WB):
        jend = min(n, j0 + WB)
        f += halves(j0, jend)                            # updates inside the block
        if jend >= n:
            break
        jnext = n if n - jend < WB else jend + WB
        f += g_lower(jend, n, jend, jnext, jend - j0)    # the next block's columns
        f += g_lower(jend, n, jnext, n, jend - j0)       # the rest
    return f


def traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE and scaled by the same counters on a
    copy kernel of known size (profiles/traffic.json), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as fh:
            return json.load(fh)[kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def roofline(st, n, nf, fast, sym=False):
    """Dominant kernel class of a profiled sweep and its achieved rate."""
    ks = st["kernels"]
    name = max(ks, key=lambda k: ks[k]["ms"])
    ms, cnt = ks[name]["ms"], ks[name]["n"]
    if name == "gemm":
        # the mixed sweep factors every frequency without pivoting and those
        # it had to redo (n_lu_retry) once more with pivoting
        retry = st.get("n_lu_retry", 0)
        if fast and sym:
            per = lu_gemm_flops_sym(n)
            work = per * nf + lu_gemm_flops(n) * retry
            desc = (f"8*M*N*K flop per computed update tile ({per:.4e} per unpivoted symmetric factorisation, "
                    f"{lu_gemm_flops(n):.4e} per pivoted LU; {nf} + {retry} redone with pivoting)")
        else:
            work = lu_gemm_flops(n) * nf
            desc = f"8*M*N*K flop per update ({lu_gemm_flops(n):.4e} per LU)"
        prec = "fp32" if fast else "fp64"
        ach = work / (ms * 1e-3) / 1e12
        # the kernel forms each complex product from three real products
        # (3M), so the matrix cores execute 6 M N K of the 8 M N K counted
        return {"kernel": f"k_gemm<{'float' if fast else 'double'}>", "bound": "mfma", "achieved": ach,
                "peak": PEAK[prec], "unit": "TFLOP/s", "frac": ach / PEAK[prec],
                "mfma_busy_frac": 0.75 * ach / PEAK[prec],
                "traffic": traffic("k_gemm" if fast else "k_gemm_f64"),
                "work": f"{desc}, {work/cnt:.3e} flop/launch avg over {cnt} launches, "
                        f"{ms/cnt*1e3:.1f} us/launch avg"}
    if name in ("regular", "near", "self"):
        evals = st["n_green_" + name]
        work = evals * FLOP_PER_GREEN * nf
        ach = work / (ms * 1e-3) / 1e12
        return {"kernel": "k_rec_" + name, "bound": "fp64 valu", "achieved": ach, "peak": PEAK["fp64"],
                "unit": "TFLOP/s", "frac": ach / PEAK["fp64"], "traffic": None,
                "work": f"{evals:.3e} Green evaluations x {FLOP_PER_GREEN} flop per launch, {cnt} launches, "
                        f"{ms/cnt*1e3:.1f} us/launch avg (FP64 vector peak = FP64 matrix peak)"}
    if name == "combine":
        byts = 16.0 * n * n  # Z written once per frequency
        ach = byts * nf / (ms * 1e-3) / 1e9
        return {"kernel": "k_combine", "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS, "traffic": None,
                "work": f"16*N^2 B of Z written per launch, {cnt} launches, {ms/cnt*1e3:.1f} us/launch avg"}
    if name == "series":
        return dict(kernel="k_series_fill", **series_rate(st, n, nf, tri=fast and sym))
    return {"kernel": name, "bound": "mfma", "achieved": None, "peak": None, "unit": None, "frac": None,
            "traffic": None, "work": f"{ms:.1f} ms over {cnt} launches"}


def tri_fill_bytes(n):
    """Bytes per matrix the series fill writes for the symmetric mixed solve:
    the FP64 upper triangle with the diagonal and the lower halves of the
    diagonal 64 x 64 blocks (16 B each), the FP32 lower triangle with the
    diagonal and the upper halves of the diagonal 16 x 16 tiles (8 B)."""
    def halves(b):
        full, rem = divmod(n, b)
        return full * b * (b - 1) // 2 + rem * (rem - 1) // 2
    tri = n * (n + 1) // 2
    return 16.0 * (tri + halves(64)) + 8.0 * (tri + halves(16))


def series_rate(st, n, nf, Nt=None, z32=True, tri=False):
    """HBM rate of the series fill (k_series_fill): per launch the batch's Z
    matrices are written once (16 B FP64 + 8 B FP32 per entry; for the
    symmetric mixed solve only the triangles it reads, tri_fill_bytes) and
    the per-entry series coefficients read once (P + 4 doubles per entry of
    the upper-triangular 16 x 16 tiles)."""
    ks = st["kernels"]["series"]
    P = st["moment_terms"] - 1
    per_f = tri_fill_bytes(n) if tri else (24.0 if z32 else 16.0) * n * n
    T = (n + 15) // 16
    q = 8.0 * (P + 4) * 256 * (T * (T + 1) // 2)
    launches = ks["n"]
    alg = (per_f * nf + q * launches) / launches
    t = ks["ms"] * 1e-3 / launches
    meas = traffic("k_series_fill")
    return {"bound": "hbm", "achieved": alg / t / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": alg / t / 1e9 / PEAK_HBM_GBS, "traffic": meas,
            "work": f"{nf / launches:.1f} frequencies per launch x {per_f:.3e} B of Z written + {q:.3e} B of "
                    f"coefficients read ({alg:.3e} B per launch), {t * 1e3:.2f} ms per launch, {t * 1e3 * launches / nf:.3f} ms "
                    f"per frequency"}


def fill_rates(st, n, Nt, nf, tri=False):
    """Achieved rates of the two fill kernels of a profiled sweep against
    MI355X peaks: the regular records (FP64 VALU work: Green evaluations x
    FLOP_PER_GREEN) and the combine (HBM: the records read once plus the
    FP64 and FP32 Z written once, per frequency; and the bytes rocprof's
    calibrated counters measured for it)."""
    ks = st["kernels"]
    out = {}
    if "series" in ks and ks["series"]["ms"] > 0:
        out["k_series_fill"] = series_rate(st, n, nf, Nt, tri=tri)
    if "regular" in ks and ks["regular"]["ms"] > 0:
        work = st["n_green_regular"] * FLOP_PER_GREEN * nf
        ach = work / (ks["regular"]["ms"] * 1e-3) / 1e12
        out["k_rec_regular"] = {"bound": "fp64 valu", "achieved": ach, "peak": PEAK["fp64"], "unit": "TFLOP/s",
                                "frac": ach / PEAK["fp64"],
                                "work": f"{st['n_green_regular']:.3e} Green evaluations x {FLOP_PER_GREEN} flop "
                                        f"per frequency"}
    if "combine" in ks and ks["combine"]["ms"] > 0:
        lda = ((n + 1) + 31) // 32 * 32
        alg = 128.0 * Nt * (Nt + 1) / 2 + 24.0 * n * lda
        t = ks["combine"]["ms"] * 1e-3 / nf
        meas = traffic("k_combine")
        out["k_combine"] = {"bound": "hbm", "achieved": alg / t / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": alg / t / 1e9 / PEAK_HBM_GBS,
                            "measured_bytes_GBps": (meas / t / 1e9) if meas else None,
                            "work": f"records once ({128 * Nt * (Nt + 1) / 2:.3e} B) + FP64 and FP32 Z once "
                                    f"({24 * n * lda:.3e} B) per frequency, {t * 1e6:.0f} us per launch"}
    return out


def baseline_rows(Ne, threads):
    """The rows of Z the CPU baseline fills (1-based), 48 per thread."""
    return np.linspace(1, Ne, 48 * threads).astype(np.int32)


def cpu_baseline(V, F, Ne, fr, threads, gpu_rows=None):
    """The reference's CPU path on the host cores: the oracle's restatement
    of assemble_ZV_EFIE_RWG_DIFF_CPU filling a sample of rows of Z (rows in
    parallel, as assemble_ZV_EFIE_RWG_DIFF_CPUp.m does with parfor),
    extrapolated to all rows, plus the dense solve of one full-size system
    by LAPACK zgesv (numpy, the routine behind MATLAB's Z\\V), timed at the
    real size.  Both at the same thread count.  gpu_rows: the same rows of
    the GPU's Z at that frequency (the fill of the mixed sweep), checked
    against the oracle's rows (the error of Z the mixed configuration asks
    to report)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ro = O.build_rwg(V, F)
    P = O.Problem(V, F, ro)
    o = O.opts(**W.OPTS_C1)
    fmid = fr[len(fr) // 2]
    w = 2 * np.pi * fmid
    k = O.wavenumber(w, W.MU0, W.EPS0)
    rws = baseline_rows(Ne, threads)
    rows = len(rws)
    t = time.perf_counter()
    Zo = P.z_rows(rws, k, w, W.MU0, W.EPS0, o, threads)
    t_rows = time.perf_counter() - t
    rng = np.random.default_rng(0)
    A = rng.standard_normal((Ne, Ne)) + 1j * rng.standard_normal((Ne, Ne))
    b = rng.standard_normal(Ne) + 1j * rng.standard_normal(Ne)
    t = time.perf_counter()
    np.linalg.solve(A, b)
    t_lu = time.perf_counter() - t
    fill_s = t_rows * Ne / rows
    # one core: 8 rows of the fill, and zgesv at n = 2048 scaled by n^3
    from threadpoolctl import threadpool_limits
    r1 = np.linspace(1, Ne, 8).astype(np.int32)
    t = time.perf_counter()
    P.z_rows(r1, k, w, W.MU0, W.EPS0, o, 1)
    t_rows1 = time.perf_counter() - t
    n1 = min(Ne, 2048)
    with threadpool_limits(limits=1):
        t = time.perf_counter()
        np.linalg.solve(A[:n1, :n1], b[:n1])
        t_lu1 = (time.perf_counter() - t) * (Ne / n1) ** 3
    fill1 = t_rows1 * Ne / len(r1)
    single = {"value": 1.0 / (fill1 + t_lu1), "unit": "freq points/s", "cores": 1,
              "sample": f"{len(r1)} rows of the fill ({t_rows1:.1f} s, x{Ne/len(r1):.0f} -> {fill1:.0f} s) + "
                        f"zgesv at n={n1} on one thread x(Ne/n)^3 -> {t_lu1:.0f} s"}
    zerr = {}
    if gpu_rows is not None:
        zerr = {"z_rows_max_rel_err_gpu_vs_oracle": float(np.max(np.abs(gpu_rows - Zo) / np.abs(Zo))),
                "z_rows_checked": f"{rows} rows x {Ne} at {fmid/1e9:.2f} GHz"}
    return {"value": 1.0 / (fill_s + t_lu), "unit": "freq points/s", "cores": threads, "kind": "port",
            "single_core": single, **zerr,
            "sample": f"oracle (C restatement of assemble_ZV_EFIE_RWG_DIFF_CPU) fill of {rows} of {Ne} rows of Z "
                      f"at {fmid/1e9:.2f} GHz on {threads} threads ({t_rows:.1f} s, x{Ne/rows:.1f} -> "
                      f"{fill_s:.0f} s per frequency) + LAPACK zgesv (numpy) at n={Ne} ({t_lu:.1f} s); "
                      f"one frequency point = {fill_s + t_lu:.0f} s"}


def _c4_problems(n_max=0):
    """The c4 geometries (with n_max > 0 only those of fewer than n_max RWG)."""
    out = []
    for g in W.c4_geometries():
        V, F = efie.make_patch_with_ground_mesh(*g["geom"], *g["cells"])
        r = efie.build_rwg_interior(V, F)
        pP, pG = W.c4_feed_points(g["geom"])
        if n_max > 0 and r.Ne >= n_max:
            continue
        out.append((g, V, F, r, efie.pick_port_edge_by_point_and_layer(r, pP, 1),
                    efie.pick_port_edge_by_point_and_layer(r, pG, -1), 1.5 * r.len.mean()))
    return out


def bench_c4(a):
    """The batched uneven sweep: 32 seeded geometries (N 1000-8000), 64
    frequencies each, mixed path.

    --c4-mode many (default): one efie_sweep_many call on an engine context
    over 32 problem contexts: the 2048 frequency points form one pipeline
    of batches (geometries of similar order share a batch, padded to its
    order), each geometry's series is integrated on the fill stream between
    batches and freed after its last batch.  One untimed pass warms the
    engine (its buffers stay allocated); the timed pass uses new problem
    contexts: the table uploads (set_problem), every geometry's one-off work
    (geometry, near lists, excitation, series) and the call.
    --c4-mode per-geometry: one efie_sweep per geometry, largest first, on
    `--c4-workers` host threads with a context each."""
    probs = _c4_problems(a.c4_max_n)
    fr = np.linspace(*W.C4["freqs"])
    o = efie.opts(**W.OPTS_C1, precision=efie.PREC_MIXED, near_mode=efie.NEAR_MOMENTS, ir_tol=a.ir_tol, ir_max_iter=a.ir_max_iter,
                  lu_kernel=a.lu_kernel, batch=a.batch)
    o64 = efie.opts(**W.OPTS_C1, precision=efie.PREC_FP64, near_mode=efie.NEAR_DIRECT)
    errs = []
    extra = {}
    if a.c4_mode == "many":
        eng = efie.Context(0)
        ctxs = []
        try:
            # warm-up: the whole workload once on other problem contexts
            # (kernels loaded, the engine's batch buffers allocated: freshly
            # recycled device memory can take seconds to hand out)
            warm = [efie.Context(0) for _ in probs]
            try:
                witems = []
                for c, (g, V, F, r, pe, ge, sp) in zip(warm, probs):
                    c.set_problem(V, F, r)
                    witems.append(dict(ctx=c, freqs=fr, patchEdge=pe, groundEdge=ge, V0=1.0, spread=sp, Z0=50.0,
                                       release=1))
                efie.sweep_many(eng, witems, W.MU0, W.EPS0, o)
            finally:
                for c in warm:
                    c.close()
            # timed: new problem contexts (geometry, near lists and series
            # built again), the engine's buffers reused
            ctxs = [efie.Context(0) for _ in probs]
            t0 = time.perf_counter()
            items = []
            for c, (g, V, F, r, pe, ge, sp) in zip(ctxs, probs):
                c.set_problem(V, F, r)
                items.append(dict(ctx=c, freqs=fr, patchEdge=pe, groundEdge=ge, V0=1.0, spread=sp, Z0=50.0,
                                  release=0 if a.c4_reuse else 1))
            res, st = efie.sweep_many(eng, items, W.MU0, W.EPS0, o, stats=True)
            t_all = time.perf_counter() - t0
            if a.c4_reuse:
                # the same call again on the same contexts: series kept, so
                # no builds (what the builds cost inside the pipeline)
                t1 = time.perf_counter()
                efie.sweep_many(eng, items, W.MU0, W.EPS0, o)
                t_re = time.perf_counter() - t1
                extra["series_kept"] = {"ms": 1e3 * t_re, "value": len(fr) * len(probs) / t_re}
            for z, _ in res:
                if not np.all(np.isfinite(z)):
                    raise RuntimeError("non-finite Zin in the c4 sweep")
            extra["stats"] = {k: st[k] for k in ("ms_total", "ms_fill", "ms_factor", "ms_solve", "ms_moments",
                                                 "ir_iters_max", "ir_resid_max", "n_lu_retry", "n_fp64_retry")}
            per = [{"N": int(p[3].Ne)} for p in probs]
            eng.close()  # its batch buffers, before the comparison sweeps take theirs
            if not a.no_fp64:
                # accuracy against the FP64 path on every fourth geometry (untimed)
                for i in range(0, len(probs), 4):
                    g, V, F, r, pe, ge, sp = probs[i]
                    z64 = ctxs[i].sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, sp, 50.0, o64)[0]
                    e = float(np.max(np.abs(res[i][0] - z64) / np.abs(z64)))
                    per[i]["max_rel_err_zin_vs_fp64"] = e
                    errs.append(e)
        finally:
            for c in ctxs:
                c.close()
            eng.close()
        nw = 1
    else:
        per, t_all, nw = _bench_c4_per_geometry(a, probs, fr, o, o64, errs)
    npts = len(fr) * len(probs)
    n3 = float(sum(float(p[3].Ne) ** 3 for p in probs)) * len(fr)
    extra["n3_per_s"] = n3 / t_all
    line = {"metric": f"freq points/s (Z fill+LU), {len(probs)} uneven geometries, mixed path",
            "value": npts / t_all, "unit": "freq points/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": 1e3 * t_all, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "c128 fill, c64 LU + c128 refinement", "data": "synthetic",
            "config": {"workload": f"c4: {len(probs)} geometries (seed {W.C4['seed']}), N {min(p['N'] for p in per)}-"
                                   f"{max(p['N'] for p in per)}, {len(fr)} frequencies 1-4 GHz each",
                       "mode": a.c4_mode, "workers": nw},
            "max_rel_err_zin_vs_fp64": max(errs) if errs else None, **extra, "per_geometry": per}
    print(json.dumps(line))


def _bench_c4_per_geometry(a, probs, fr, o, o64, errs):
    """One efie_sweep per geometry, largest first, handed out to
    `--c4-workers` host threads (one context each).  Tables, geometry
    caches and solver buffers are rebuilt or grown per geometry inside the
    timed region."""
    import threading
    order = sorted(range(len(probs)), key=lambda i: -probs[i][3].Ne)
    nw = max(1, a.c4_workers)
    ctxs = [efie.Context(0) for _ in range(nw)]
    per = [None] * len(probs)
    try:
        for ctx in ctxs:   # warm-up: kernels loaded, streams created
            g, V, F, r, pe, ge, sp = probs[order[-1]]
            ctx.set_problem(V, F, r)
            ctx.sweep(fr[:4], W.MU0, W.EPS0, pe, ge, 1.0, sp, 50.0, o)
        lock = threading.Lock()
        queue = list(order)
        fails = []

        def worker(ctx):
            try:
                while True:
                    with lock:
                        if not queue:
                            return
                        i = queue.pop(0)
                    g, V, F, r, pe, ge, sp = probs[i]
                    t0 = time.perf_counter()
                    ctx.set_problem(V, F, r)
                    zin, _, st = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, sp, 50.0, o, stats=True)
                    dt = time.perf_counter() - t0
                    per[i] = {"N": int(r.Ne), "s": dt, "freq_per_s": len(fr) / dt, "ir_iters_max": st["ir_iters_max"],
                              "n_lu_retry": st["n_lu_retry"], "ms_series_build": st["ms_moments"]}
                    log(f"[c4] geometry {i}: N={r.Ne} {dt:.2f} s")
            except Exception as e:   # reported after the join
                fails.append(e)
        t_all = time.perf_counter()
        ths = [threading.Thread(target=worker, args=(ctx,)) for ctx in ctxs]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        t_all = time.perf_counter() - t_all
        if fails:
            raise fails[0]
        if not a.no_fp64:
            ctx = ctxs[0]
            for g, V, F, r, pe, ge, sp in probs[::4]:
                ctx.set_problem(V, F, r)
                z64 = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, sp, 50.0, o64)[0]
                zm = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, sp, 50.0, o)[0]
                errs.append(float(np.max(np.abs(zm - z64) / np.abs(z64))))
    finally:
        for ctx in ctxs:
            ctx.close()
    return per, t_all, nw


TRACE_CLASSES = ["k_gemm", "k_panel", "k_sym_u12", "k_symv", "k_trsv", "k_series_fill", "k_series_near",
                 "k_series_coeffs", "k_series_regular", "k_series_self", "k_tslu", "k_pad_identity"]


def trace_states(path, window_gap_ms=0.0, range_ms=None):
    """What the GPU runs over the time of a rocprofv3 kernel trace (CSV):
    the union of kernel intervals split by the kernel classes active at
    each moment.  Any moment with a trailing-update GEMM running counts as
    "k_gemm (+ others)"; the rest is listed by the set of classes running
    (what is exposed next to no matrix-core work) or "idle".  With
    window_gap_ms > 0 only the kernels after the last gap of at least that
    many ms with no kernel running (the last timed call) are counted; with
    range_ms = (a, b) only the kernels starting between a and b ms after
    the first one.  Also the summed duration and count of each class."""
    import csv
    import collections
    ev = []
    with open(path) as fh:
        for r in csv.DictReader(fh):
            name = r["Kernel_Name"].split("(")[0].replace("void ", "").replace("efie::", "")
            ev.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), name))
    ev.sort()
    if range_ms:
        t0 = ev[0][0]
        ev = [e for e in ev if range_ms[0] * 1e6 <= e[0] - t0 < range_ms[1] * 1e6]
    if window_gap_ms > 0:
        end, start = ev[0][1], 0
        for i, (a0, b0, _) in enumerate(ev[1:], 1):
            if a0 - end >= window_gap_ms * 1e6:
                start = i
            end = max(end, b0)
        ev = ev[start:]

    def cls(n):
        return next((k for k in TRACE_CLASSES if n.startswith(k)), "other")
    chg = sorted([(a0, 1, cls(n)) for a0, _, n in ev] + [(b0, -1, cls(n)) for _, b0, n in ev])
    act = collections.Counter()
    states = collections.Counter()
    last = chg[0][0]
    for t, d, c in chg:
        if t > last:
            if act["k_gemm"] > 0:
                k = "k_gemm (+ others)"
            elif sum(act.values()) == 0:
                k = "idle"
            else:
                k = " + ".join(sorted(x for x, v in act.items() if v > 0))
            states[k] += t - last
        act[c] += d
        last = t
    summed = collections.defaultdict(lambda: [0, 0])
    for a0, b0, n in ev:
        summed[cls(n)][0] += b0 - a0
        summed[cls(n)][1] += 1
    return {"span_ms": (max(e[1] for e in ev) - ev[0][0]) / 1e6, "kernels": len(ev),
            "states_ms": {k: v / 1e6 for k, v in states.most_common()},
            "summed_ms": {k: v[0] / 1e6 for k, v in sorted(summed.items(), key=lambda x: -x[1][0])},
            "launches": {k: v[1] for k, v in summed.items()}}


TRAFFIC_KEYS = {"k_gemm<float": "k_gemm", "k_gemm<double": "k_gemm_f64", "k_series_fill": "k_series_fill",
                "k_symv_finish": "k_symv_finish", "k_symv_strips": "k_symv_strips",
                "k_trsv_fused<float, false": "k_trsv_fused_upper", "k_trsv_fused<float, true": "k_trsv_fused_lower"}


def pmc_summary(root, legs):
    """HBM bytes per launch from rocprofv3 --pmc runs (one counter per
    run): root/pmc_copy_{FETCH_SIZE,WRITE_SIZE} hold the plain copy kernel
    (bench.py --copy-probe-mb) that calibrates the counters, and
    root/pmc_<leg>_{FETCH_SIZE,WRITE_SIZE} the profiled sweeps.  Each
    counter is scaled by the bytes the copy kernel moves over what it
    reads for it; returns the rows of profiles/traffic.json."""
    import csv
    import glob

    def load(d):
        out = {}
        for f in glob.glob(os.path.join(root, d, "*counter_collection.csv")):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    k = r["Kernel_Name"].replace("void ", "").replace("efie::", "")
                    out.setdefault(k, []).append(float(r["Counter_Value"]))
        return out
    cf, cw = load("pmc_copy_FETCH_SIZE"), load("pmc_copy_WRITE_SIZE")
    copy = next(k for k in cf if k.startswith("k_copy16"))
    nbytes = None
    for f in glob.glob(os.path.join(root, "pmc_copy_FETCH_SIZE*.log")):
        for line in open(f):
            if line.startswith("{") and "copy_bytes" in line:
                nbytes = json.loads(line)["copy_bytes"]
    nbytes = nbytes or (1 << 30)
    fs = nbytes / 1024.0 / (sum(cf[copy]) / len(cf[copy]))
    ws = nbytes / 1024.0 / (sum(cw[copy]) / len(cw[copy]))
    rows = {"_calibration": {"copy_kernel": "efie::k_copy16", "copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                             "fetch_scale": fs, "write_scale": ws}}
    for leg in legs:
        lf, lw = load(f"pmc_{leg}_FETCH_SIZE"), load(f"pmc_{leg}_WRITE_SIZE")
        for k in lf:
            key = next((v for p_, v in TRAFFIC_KEYS.items() if k.startswith(p_)), None)
            if key is None or k not in lw:
                continue
            f_avg = sum(lf[k]) / len(lf[k])
            w_avg = sum(lw[k]) / len(lw[k])
            rows[key] = {"kernel": "efie::" + k.split("(")[0], "leg": leg, "launches": len(lf[k]),
                         "fetch_kib_avg": f_avg, "write_kib_avg": w_avg,
                         "bytes_per_launch": 1024.0 * (f_avg * fs + w_avg * ws)}
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2_6k")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--lu-kernel", type=int, default=0, help="0 auto (symmetric for the mixed sweep), 1 FP32 MFMA pivoted, 2 bf16x6 pivoted, 3 symmetric FP32")
    ap.add_argument("--ir-tol", type=float, default=1e-10)
    ap.add_argument("--ir-max-iter", type=int, default=16, help="refinement corrections at most (the library default)")
    ap.add_argument("--batch", type=int, default=0, help="frequencies per batch (0: the library's choice)")
    ap.add_argument("--c4-reuse", action="store_true",
                    help="c4 many mode: also time a second call on the same contexts (series already built)")
    ap.add_argument("--c4-mode", choices=["many", "per-geometry"], default="many",
                    help="c4: one efie_sweep_many call (default) or one efie_sweep per geometry")
    ap.add_argument("--c4-max-n", type=int, default=0,
                    help="c4: only the geometries of fewer than this many RWG (0: all 32)")
    ap.add_argument("--c4-workers", type=int, default=1,
                    help="host threads (one context each) sweeping c4 geometries side by side")
    ap.add_argument("--nf", type=int, default=0, help="use only the first NF frequencies (profiling runs)")
    ap.add_argument("--profile-leg", choices=["fast", "fp64", "fp64-series"], default=None,
                    help="only one warmup and one sweep of this leg, kernels serialised on one stream as in the "
                         "profiled sweep (for kernel traces: rocprofv3)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run the plain copy kernel on this many MiB (counter calibration)")
    ap.add_argument("--trace-states", default=None, metavar="CSV",
                    help="only analyse a rocprofv3 kernel trace (CSV): GPU time by the kernel classes running")
    ap.add_argument("--trace-window-gap-ms", type=float, default=0.0,
                    help="--trace-states: only the kernels after the last idle gap of this many ms")
    ap.add_argument("--trace-range-ms", type=float, nargs=2, default=None, metavar=("A", "B"),
                    help="--trace-states: only the kernels starting A to B ms after the first one")
    ap.add_argument("--pmc-summary", default=None, metavar="DIR",
                    help="only turn the rocprofv3 --pmc runs under DIR (pmc_copy_*, pmc_<leg>_*) into the rows of "
                         "profiles/traffic.json, merged into that file")
    ap.add_argument("--pmc-legs", nargs="+", default=["fast"], help="--pmc-summary: the profiled legs")
    a = ap.parse_args()
    if a.pmc_summary:
        path = os.path.join(ROOT, "profiles", "traffic.json")
        old = json.load(open(path)) if os.path.exists(path) else {}
        new = pmc_summary(a.pmc_summary, a.pmc_legs)
        old.update(new)
        with open(path, "w") as fh:
            json.dump(old, fh, indent=1)
        print(json.dumps(new, indent=1))
        return
    if a.trace_states:
        print(json.dumps(trace_states(a.trace_states, a.trace_window_gap_ms, a.trace_range_ms), indent=1))
        return
    if a.copy_probe_mb:
        with efie.Context(0) as ctx:
            nbytes = a.copy_probe_mb << 20
            ms = ctx.diag_copy(nbytes, 5)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms, "GB/s": 2 * nbytes / (ms * 1e-3) / 1e9}))
        return
    if a.config == "c4":
        if a.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            sys.exit(2)
        bench_c4(a)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU (asked for {a.gpus}): the sweep would shard by frequency "
              f"across GPUs, which this build does not run.", file=sys.stderr)
        sys.exit(2)
    V, F, r, pe, ge, spread, fr = setup(a.config)
    if a.nf:
        fr = fr[: a.nf]
    n = r.Ne
    if a.profile_leg:
        kw = {"fast": dict(precision=efie.PREC_MIXED, near_mode=efie.NEAR_MOMENTS, ir_tol=a.ir_tol, ir_max_iter=a.ir_max_iter,
                           lu_kernel=a.lu_kernel, batch=a.batch),
              "fp64": dict(precision=efie.PREC_FP64, near_mode=efie.NEAR_DIRECT),
              "fp64-series": dict(precision=efie.PREC_FP64, near_mode=efie.NEAR_MOMENTS)}[a.profile_leg]
        with efie.Context(0) as ctx:
            ctx.set_problem(V, F, r)
            # profile=1: every kernel on one stream, as the profiled sweep of
            # the full run whose per-class times the roofline entries use
            dt, (_, _, st) = timed(ctx, efie.opts(**W.OPTS_C1, **kw, profile=1), fr, pe, ge, spread, 1, 1,
                                   a.profile_leg)
        print(json.dumps({"leg": a.profile_leg, "config": a.config, "Ne": n, "Nf": len(fr), "s": dt,
                          "freq_per_s": len(fr) / dt, "stats": st}))
        return
    log(f"config {a.config}: Ne={n} Nt={F.shape[0]} Nf={len(fr)}")
    o64 = efie.opts(**W.OPTS_C1, precision=efie.PREC_FP64, near_mode=efie.NEAR_DIRECT)
    # refinement to ||V - Zx|| <= 1e-10 ||V||
    ofast = efie.opts(**W.OPTS_C1, precision=efie.PREC_MIXED, near_mode=efie.NEAR_MOMENTS, ir_tol=a.ir_tol, ir_max_iter=a.ir_max_iter,
                      lu_kernel=a.lu_kernel, batch=a.batch)
    with efie.Context(0) as ctx:
        ctx.set_problem(V, F, r)
        first = {}
        dt_fast, (zin_f, s11_f, st_f) = timed(ctx, ofast, fr, pe, ge, spread, a.steps, a.warmup, "fast", first)
        res = {}
        if not a.no_fp64:
            dt_64, (zin_64, s11_64, st_64) = timed(ctx, o64, fr, pe, ge, spread, a.steps, a.warmup, "fp64")
            _, _, st_64p = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                                     efie.opts(**W.OPTS_C1, precision=efie.PREC_FP64, near_mode=efie.NEAR_DIRECT,
                                               profile=1), stats=True)
            res["fp64"] = {"value": len(fr) * a.steps / dt_64, "ms_per_step": 1e3 * dt_64 / a.steps,
                           "stats": st_64, "profile": st_64p["kernels"],
                           "roofline": roofline(st_64p, n, len(fr), False)}
            # the same fill algorithm as the fast path (series), FP64 LU: the
            # speed-up the precision alone gives
            o64m = efie.opts(**W.OPTS_C1, precision=efie.PREC_FP64, near_mode=efie.NEAR_MOMENTS)
            dt_64m, (zin_64m, _, _) = timed(ctx, o64m, fr, pe, ge, spread, a.steps, a.warmup, "fp64 series")
            res["fp64_series_fill"] = {
                "value": len(fr) * a.steps / dt_64m, "ms_per_step": 1e3 * dt_64m / a.steps,
                "max_rel_err_zin_vs_fp64": float(np.max(np.abs(zin_64m - zin_64) / np.abs(zin_64)))}
            log("fp64 legs done")
            e = np.abs(zin_f - zin_64) / np.abs(zin_64)
            res["max_rel_err_zin_fast_vs_fp64"] = float(np.max(e))
            # per-frequency errors: their spread and where the largest falls
            res["zin_err_per_freq"] = {"median": float(np.median(e)), "p90": float(np.quantile(e, 0.9)),
                                       "max_at_ghz": float(fr[int(np.argmax(e))] / 1e9)}
        if not a.no_surrogate:
            # learned near pairs: train on the GPU (one-off, outside the timed
            # sweeps), then the mixed + batched + surrogate sweep
            t0 = time.perf_counter()
            rep = ctx.sur_train(efie.opts(**W.OPTS_C1))
            t_train = time.perf_counter() - t0
            log(f"surrogate trained in {t_train:.2f} s")
            osur = efie.opts(**W.OPTS_C1, precision=efie.PREC_MIXED, near_mode=efie.NEAR_SURROGATE, ir_tol=a.ir_tol, ir_max_iter=a.ir_max_iter,
                             lu_kernel=a.lu_kernel, batch=a.batch)
            dt_s, (zin_s, _, st_s) = timed(ctx, osur, fr, pe, ge, spread, a.steps, a.warmup, "surrogate")
            _, _, st_sp = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0,
                                    efie.opts(**W.OPTS_C1, precision=efie.PREC_MIXED,
                                              near_mode=efie.NEAR_SURROGATE, ir_tol=a.ir_tol, ir_max_iter=a.ir_max_iter,
                                              lu_kernel=a.lu_kernel, batch=a.batch, profile=1), stats=True)
            sur = {"value": len(fr) * a.steps / dt_s, "ms_per_step": 1e3 * dt_s / a.steps,
                   "train_s": t_train, "train": rep, "n_sur_pairs": st_s["n_sur_pairs"],
                   "n_sur_fallback": st_s["n_sur_fallback"],
                   "near_stage_ms": {k: st_sp["kernels"][k]["ms"] for k in ("surrogate", "near", "poly")
                                     if k in st_sp["kernels"]}}
            if not a.no_fp64:
                sur["max_rel_err_zin_vs_fp64"] = float(np.max(np.abs(zin_s - zin_64) / np.abs(zin_64)))
                sur["vs_fp64"] = sur["value"] / res["fp64"]["value"]
            res["surrogate"] = sur
        # profiled pass of the headline path for the roofline
        oprof = efie.opts(**W.OPTS_C1, precision=efie.PREC_MIXED, near_mode=efie.NEAR_MOMENTS, ir_tol=a.ir_tol, ir_max_iter=a.ir_max_iter,
                      lu_kernel=a.lu_kernel, batch=a.batch, profile=1)
        _, _, st_p = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, oprof, stats=True)
        log("profiled pass done")
        threads = min(os.cpu_count() or 1, 16)
        gpu_rows = None
        if not a.no_cpu:
            # the fill of the mixed sweep at the baseline's frequency, rows the
            # baseline fills too
            wm = 2 * np.pi * fr[len(fr) // 2]
            Zg, _ = ctx.assemble_zv(wm * np.sqrt(W.MU0 * W.EPS0), wm, W.MU0, W.EPS0, [pe, ge, 1.0, spread, 0.2],
                                    efie.opts(**W.OPTS_C1, near_mode=efie.NEAR_MOMENTS))
            gpu_rows = Zg[baseline_rows(n, threads) - 1].copy()
            del Zg
    Nt = F.shape[0]
    st_p["Nt"] = Nt
    reg_pairs = Nt * (Nt - 1) // 2 - st_p["n_near_pairs"]
    nq = {4: 6, 7: 7, 2: 3, 1: 1}.get(W.OPTS_C1["quad_order"], 6)
    st_p["n_green_regular"] = reg_pairs * nq * nq
    value = len(fr) * a.steps / dt_fast
    line = {
        "metric": "freq points/s (Z fill+LU) at N~6k, mixed vs FP64; max rel err Zin vs FP64",
        "value": value, "unit": "freq points/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * dt_fast / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128 fill, c64 LU + c128 refinement", "data": "synthetic",
        "config": {"workload": f"{a.config}: patch {W.CONFIGS[a.config]['cells'][0]}x"
                               f"{W.CONFIGS[a.config]['cells'][1]} + ground {W.CONFIGS[a.config]['cells'][2]}x"
                               f"{W.CONFIGS[a.config]['cells'][3]} cells, Ne={n}, Nt={Nt}, Nf={len(fr)} "
                               f"({fr[0]/1e9:.1f}-{fr[-1]/1e9:.1f} GHz), 7-pt regular rule, 6-pt x16 near/self"},
        "roofline": roofline(st_p, n, len(fr), True, sym=a.lu_kernel in (0, 3)),
        "fast": {"stats": st_f,
                 "one_off_per_mesh_ms": first.get("ms_moments", 0.0),
                 "one_off_note": "the frequency-independent series of every triangle pair, integrated by the first "
                                 "sweep of a mesh and reused by later sweeps of it (the timed steps); "
                                 "value_incl_one_off charges it to every step"},
        "value_incl_one_off": len(fr) / (dt_fast / a.steps + 1e-3 * first.get("ms_moments", 0.0)),
        "fast_vs_fp64": (value / res["fp64"]["value"]) if "fp64" in res else None,
        "fast_vs_fp64_same_fill": (value / res["fp64_series_fill"]["value"]) if "fp64_series_fill" in res else None,
        **res,
        "profile": st_p["kernels"],
        "fill": fill_rates(st_p, n, Nt, len(fr), tri=a.lu_kernel in (0, 3)),
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(V, F, n, fr, threads, gpu_rows)
        log("cpu baseline done")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
