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


# ------------------------------------- the prepack entry points (fastquad) --

@pytest.fixture(scope="module")
def pk_args(efie_mod, args):
    import prepack_np
    V, F = args["sweep"][0], args["sweep"][1]
    pk = prepack_np.make_prepack(V, F)
    r = efie_mod.build_rwg_interior(V, F)
    tables = args["sweep"][:12]
    fq = tables + [pk] + args["sweep"][12:]
    e = args["asm"][16]
    gg = tables + [pk] + args["asm"][12:16] + [e[:4]]
    rwg = dict(plusTri=r.plusTri, minusTri=r.minusTri, plusSign=r.plusSign, minusSign=r.minusSign,
               len=r.len, Ap=r.Ap, Am=r.Am, rp=r.rp, rm=r.rm, center=r.center, zSign=r.zSign.astype(np.float64))
    exc = dict(patchEdge=float(e[0]), groundEdge=float(e[1]), V0=1.0, spreadRadius=float(e[3]))
    a8 = [rwg, pk] + args["asm"][12:16] + [exc, {"near_alpha": 0.15}]
    return dict(fq=fq, gg=gg, a8=a8, pk=pk, Ne=args["Ne"])


FQ, GG, A8 = "patch_sweep_fastquad_cuda", "ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", "assemble_ZV_EFIE_RWG_DIFF_GPU_FAST"


def _with(a, pos, v):
    a = list(a)
    a[pos] = v
    return a


def _pk(pk, **kw):
    q = dict(pk)
    for k, v in kw.items():
        if v is None:
            q.pop(k)
        else:
            q[k] = v
    return q


def test_prepack_argument_counts(pk_args):
    """patch_sweep_fastquad_cuda.cu:473-474, ggassemble...cu:544-546,
    assemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:280-281"""
    assert _err(FQ, 2, pk_args["fq"][:20]) == "mex:args"
    assert _err(FQ, 1, pk_args["fq"]) == "mex:args"
    assert _err(GG, 2, pk_args["gg"][:17]) == "mex:args"
    assert _err(GG, 1, pk_args["gg"]) == "mex:args"
    assert _err(A8, 2, pk_args["a8"][:7]) == "mex:args"
    assert _err(A8, 1, pk_args["a8"]) == "mex:args"


@pytest.mark.parametrize("name", [FQ, GG])
def test_prepack_field_checks(pk_args, name):
    a = pk_args["fq" if name == FQ else "gg"]
    pk = pk_args["pk"]
    assert _err(name, 2, _with(a, 12, np.array([1.0]))) == "mex:prepack"
    assert _err(name, 2, _with(a, 12, _pk(pk, near_r=None))) == "mex:prepack"
    assert _err(name, 2, _with(a, 12, _pk(pk, r_reg=pk["r_reg"][:, :, 0]))) == "mex:r_reg"
    assert _err(name, 2, _with(a, 12, _pk(pk, r_reg=pk["r_reg"][:, :2]))) == "mex:r_reg"
    assert _err(name, 2, _with(a, 12, _pk(pk, w_reg=pk["w_reg"][:-1]))) == "mex:w_reg"
    assert _err(name, 2, _with(a, 12, _pk(pk, nearMat=pk["nearMat"].astype(np.float64)))) == "mex:nearMat"
    assert _err(name, 2, _with(a, 12, _pk(pk, self_w=pk["self_w"][:, :2]))) == "mex:self_w"
    assert _err(name, 2, _with(a, 12, _pk(pk, self_a=pk["self_a"][:2]))) == "mex:self_a"
    assert _err(name, 2, _with(a, 12, _pk(pk, near_r=pk["near_r"][:3]))) == "mex:near_r"
    assert _err(name, 2, _with(a, 12, _pk(pk, aeqTri=pk["aeqTri"][:-1]))) == "mex:aeqTri"
    assert _err(name, 2, _with(a, 12, _pk(pk, near_alpha=None))) == "mex:prepack"


def test_prepack_scalar_checks(pk_args):
    Ne = pk_args["Ne"]
    a = pk_args["fq"]
    assert _err(FQ, 2, _with(a, 13, a[13].astype(np.int32))) == "mex:freqs"
    assert _err(FQ, 2, _with(a, 16, np.int32(Ne + 1))) == "mex:port"
    assert _err(FQ, 2, _with(a, 2, a[2].astype(np.float64))) == "mex:E"
    g = pk_args["gg"]
    assert _err(GG, 2, _with(g, 17, g[17][:3])) == "mex:exc"
    e = g[17].copy()
    e[1] = 0
    assert _err(GG, 2, _with(g, 17, e)) == "mex:port"
    assert _err(GG, 2, _with(g, 14, 0.0)) == "mex:eps"


def test_eight_input_checks(pk_args):
    a = pk_args["a8"]
    rwg, pk, exc = a[0], a[1], a[6]
    assert _err(A8, 2, _with(a, 0, _pk(rwg, plusTri=None))) == "mex:rwg"
    assert _err(A8, 2, _with(a, 0, _pk(rwg, plusTri=rwg["plusTri"].astype(np.float64)))) == "mex:rwg"
    assert _err(A8, 2, _with(a, 0, _pk(rwg, minusSign=rwg["minusSign"].astype(np.float64)))) == "mex:rwg"
    assert _err(A8, 2, _with(a, 0, _pk(rwg, rp=rwg["rp"][:, :2]))) == "mex:rwg"
    assert _err(A8, 2, _with(a, 0, _pk(rwg, center=None))) == "mex:rwg"
    assert _err(A8, 2, _with(a, 1, _pk(pk, w_reg=None))) == "mex:pre"
    assert _err(A8, 2, _with(a, 1, _pk(pk, nearMat=pk["nearMat"].astype(np.float64)))) == "mex:pre"
    assert _err(A8, 2, _with(a, 1, _pk(pk, near_w=None))) == "mex:pre"
    assert _err(A8, 2, _with(a, 1, _pk(pk, w_reg=pk["w_reg"][:, :-1]))) == "mex:pre"
    assert _err(A8, 2, _with(a, 6, _pk(exc, V0=None))) == "mex:exc"
    assert _err(A8, 2, _with(a, 6, _pk(exc, patchEdge=0.0))) == "mex:port"
    # the port edge weighs exp(0) = 1, so only spreadRadius = 0 (0/0) leaves
    # no weights; the reference's sumw < 1e-30 test lets that NaN through
    assert _err(A8, 2, _with(a, 6, _pk(exc, spreadRadius=0.0))) == "mex:rhs"
    # every check passed: without a device the context is the first step
    # that fails
    try:
        mx.call(A8, 2, *a)
    except mx.MexError as e:
        assert e.ident == "efie:hip"
