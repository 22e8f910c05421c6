"""Throughput benchmark of the EFIE frequency sweep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2_6k]

One step is one full frequency sweep of the configuration (fill Z, LU
solve, Zin/S11 for every frequency) with the mesh and RWG tables already
resident in HBM.  The printed JSON line carries:
  value        frequency points per second of the headline path,
  roofline     the dominant kernel's achieved rate against the MI355X peak,
  cpu_baseline the CPU oracle (this project's C restatement of the
               reference assembler) timed on a bounded sample on the host.
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

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (matrix and vector)
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0
FLOP_PER_GREEN = 48       # flop count convention per Green-function evaluation


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(config):
    V, F = efie.make_patch_with_ground_mesh(*W.mesh_args(config))
    r = efie.build_rwg_interior(V, F)
    pP, pG = W.feed_points(config)
    pe = efie.pick_port_edge_by_point_and_layer(r, pP, 1)
    ge = efie.pick_port_edge_by_point_and_layer(r, pG, -1)
    return V, F, r, pe, ge, 1.5 * r.len.mean(), W.freqs(config)


def run_sweeps(ctx, args, o, fr, pe, ge, spread, steps, warmup):
    for i in range(warmup):
        ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o)
        log(f"warmup {i + 1}/{warmup} done")
    t0 = time.perf_counter()
    out = None
    for i in range(steps):
        out = ctx.sweep(fr, W.MU0, W.EPS0, pe, ge, 1.0, spread, 50.0, o, stats=True)
        log(f"step {i + 1}/{steps} done")
    dt = time.perf_counter() - t0  # each sweep returns after its stream has drained
    return dt, out


def cpu_baseline(V, F, r, Ne, fr, budget_s=20.0):
    """Oracle fill of a few rows of Z at one frequency plus an oracle LU at
    n=1024, both on all host threads, extrapolated to one frequency point."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    ro = O.build_rwg(V, F)
    P = O.Problem(V, F, ro)
    o = O.opts(**W.OPTS_C1)
    w = 2 * np.pi * fr[len(fr) // 2]
    k = O.wavenumber(w, W.MU0, W.EPS0)
    rows = max(threads, 8)
    t = time.perf_counter()
    P.z_rows(np.linspace(1, Ne, rows).astype(np.int32), k, w, W.MU0, W.EPS0, o, threads)
    t_rows = time.perf_counter() - t
    rng = np.random.default_rng(0)
    n_lu = 1024
    A = rng.standard_normal((n_lu, n_lu)) + 1j * rng.standard_normal((n_lu, n_lu))
    t = time.perf_counter()
    O.zgetrf(A)
    t_lu = time.perf_counter() - t
    per_freq = t_rows * Ne / rows + t_lu * (Ne / n_lu) ** 3
    return {"value": 1.0 / per_freq, "unit": "freq points/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle fill of {rows} rows of Z at f={fr[len(fr)//2]/1e9:.3f} GHz "
                      f"({t_rows:.1f}s, scaled x{Ne/rows:.0f}) + oracle zgetrf at n={n_lu} "
                      f"({t_lu:.1f}s, scaled x{(Ne/n_lu)**3:.0f}); fill on {threads} threads, LU on 1",
            "fill_threads": threads}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2_6k")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the frequency sweep would shard "
              f"by frequency across GPUs, which this build does not run.", file=sys.stderr)
        sys.exit(2)
    V, F, r, pe, ge, spread, fr = setup(a.config)
    log(f"config {a.config}: Ne={r.Ne} Nt={F.shape[0]} Nf={len(fr)}")
    o64 = efie.opts(**W.OPTS_C1)
    with efie.Context(0) as ctx:
        ctx.set_problem(V, F, r)
        dt, (zin, s11, st) = run_sweeps(ctx, a, o64, fr, pe, ge, spread, a.steps, a.warmup)
    nf = len(fr) * a.steps
    value = nf / dt
    line = {
        "metric": "freq points/s (Z fill+LU) at N~6k, mixed vs FP64; max rel err Zin vs FP64",
        "value": value, "unit": "freq points/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"{a.config}: patch {W.CONFIGS[a.config]['cells'][:2]} + ground "
                               f"{W.CONFIGS[a.config]['cells'][2:]} cells, Ne={r.Ne}, Nf={len(fr)}, "
                               f"FP64 fill + FP64 LU"},
        "stats": st,
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(V, F, r, r.Ne, fr)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
