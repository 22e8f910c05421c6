"""The MEX entry points (matlab-cuda-efief-vs-cpu_amd/mex/) driven through the
project's mx shim: argument counts, types and error identifiers of the
reference MEX files, checked before any device work."""
import os
import sys

import numpy as np
import pytest

import workloads as W

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "mxshim"))
import mx  # noqa: E402


@pytest.fixture(scope="module")
def args(efie_mod):
    V, F = efie_mod.make_patch_with_ground_mesh(*W.mesh_args("micro"))
    r = efie_mod.build_rwg_interior(V, F)
    pP, pG = W.feed_points("micro")
    pe = efie_mod.pick_port_edge_by_point_and_layer(r, pP, 1)
    ge = efie_mod.pick_port_edge_by_point_and_layer(r, pG, -1)
    tables = [V, F, r.edge, r.plusTri, r.minusTri, r.plusSign, r.minusSign, r.rp, r.rm, r.Ap, r.Am, r.len]
    sweep = tables + [W.freqs("micro"), W.MU0, complex(W.EPS0, 0), np.int32(pe), np.int32(ge), 1.0,
                      1.5 * r.len.mean(), 50.0]
    w = 2 * np.pi * 2e9
    k = w * np.sqrt(W.MU0 * W.EPS0)
    asm = tables + [complex(k, 0), w, W.MU0, W.EPS0, np.array([pe, ge, 1.0, 1.5 * r.len.mean(), 0.2])]
    return dict(sweep=sweep, asm=asm, Ne=r.Ne)


def _err(name, nlhs, a):
    with pytest.raises(mx.MexError) as e:
        mx.call(name, nlhs, *a)
    return e.value.ident


def test_argument_counts(args):
    assert _err("patch_sweep_cuda", 2, args["sweep"][:19]) == "mex:args"
    assert _err("patch_sweep_cuda", 1, args["sweep"]) == "mex:args"
    assert _err("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, args["asm"][:16]) == "mex:args"
    assert _err("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 3, args["asm"]) == "mex:args"


@pytest.mark.parametrize("pos,bad,ident", [
    (0, lambda a: a[:, :2], "mex:V"),
    (1, lambda a: a.astype(np.float64), "mex:F"),
    (3, lambda a: a[:-1], "mex:plusTri"),
    (4, lambda a: a.astype(np.float64), "mex:minusTri"),
    (5, lambda a: a[:-1], "mex:plusSign"),
    (6, lambda a: a.astype(np.float64), "mex:minusSign"),
    (7, lambda a: a[:, :2], "mex:rp"),
    (8, lambda a: a[:-1], "mex:rm"),
    (9, lambda a: a[:-1], "mex:Ap"),
    (10, lambda a: a.astype(np.int32), "mex:Am"),
    (11, lambda a: a[:-1], "mex:len"),
])
def test_table_checks(args, pos, bad, ident):
    a = list(args["sweep"])
    a[pos] = bad(a[pos])
    assert _err("patch_sweep_cuda", 2, a) == ident
    b = list(args["asm"])
    b[pos] = bad(b[pos])
    assert _err("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, b) == ident


def test_edge_ids_differ_per_entry_point(args):
    a = list(args["sweep"])
    a[2] = a[2].astype(np.float64)
    assert _err("patch_sweep_cuda", 2, a) == "mex:edge"
    b = list(args["asm"])
    b[2] = b[2].astype(np.float64)
    assert _err("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, b) == "mex:E"


def test_scalar_checks(args):
    a = list(args["sweep"])
    a[12] = a[12].astype(np.int32)
    assert _err("patch_sweep_cuda", 2, a) == "mex:freqs"
    for pos, v in ((15, np.int32(0)), (16, np.int32(args["Ne"] + 1))):
        a = list(args["sweep"])
        a[pos] = v
        assert _err("patch_sweep_cuda", 2, a) == "mex:port"
    b = list(args["asm"])
    b[16] = b[16][:4]
    assert _err("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, b) == "mex:exc"
    b = list(args["asm"])
    e = b[16].copy()
    e[0] = args["Ne"] + 1
    b[16] = e
    assert _err("gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", 2, b) == "mex:port"
    a = list(args["sweep"]) + [np.array([1.0])]
    assert _err("patch_sweep_cuda", 2, a) == "mex:opts"
