"""The batch plan of efie_sweep_many (efie_diag_pack_plan, host only): how
the frequencies of many problems are cut into batches and which problems
share a padded batch.  No device work."""
import numpy as np
import pytest

import workloads as W

PER = {0: 16, 1: 24}  # bytes per matrix element: FP64, mixed (FP64 + FP32 copy)


def _check_plan(plan, n, nf, precision, batch, free):
    """Every frequency of every item placed once; batches of order np hold
    problems of order in [0.75 np, np], packed in descending order, within
    the matrix-count and memory caps."""
    seen = {i: np.zeros(nf[i], int) for i in range(len(n))}
    nb_of = {}
    for b, item, f0, k, e0, np_ in plan.tolist():
        assert k > 0 and 0 <= f0 and f0 + k <= nf[item]
        seen[item][f0:f0 + k] += 1
        assert n[item] <= np_
        nb_of.setdefault(b, []).append((e0, k, item, np_))
    for i in range(len(n)):
        assert np.all(seen[i] == 1), (i, seen[i])
    prev_np = None
    for b in sorted(nb_of):
        segs = sorted(nb_of[b])
        np_ = segs[0][3]
        assert all(s[3] == np_ for s in segs)
        # matrices contiguous from 0, the first segment's problem of the batch order
        e = 0
        for e0, k, item, _ in segs:
            assert e0 == e
            e += k
        assert n[segs[0][2]] == np_
        for _, _, item, _ in segs[1:]:
            assert n[item] >= 0.75 * np_
        lda = (np_ + 1 + 31) // 32 * 32
        cap = batch if batch > 0 else min(512, max(64, int(2.4e9 / (np_ * lda))))
        cap = min(cap, max(1, int(0.6 * free / (2 * PER[precision] * np_ * lda))))
        assert e <= cap
        if prev_np is not None:
            assert np_ <= prev_np
        prev_np = np_
    return nb_of


def _c4_orders(efie_mod):
    out = []
    for g in W.c4_geometries():
        V, F = efie_mod.make_patch_with_ground_mesh(*g["geom"], *g["cells"])
        out.append(efie_mod.build_rwg_interior(V, F).Ne)
    return np.array(out)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("free", [280e9, 60e9])
def test_c4_plan(efie_mod, precision, free):
    n = _c4_orders(efie_mod)
    nf = np.full(n.shape[0], 64)
    plan = efie_mod.pack_plan(n, nf, precision=precision, free_bytes=free)
    nb_of = _check_plan(plan, n, nf, precision, 0, free)
    shared = [b for b, segs in nb_of.items() if len({s[2] for s in segs}) > 1]
    print(f"c4 plan ({'mixed' if precision else 'fp64'}, {free / 1e9:.0f} GB free): {len(nb_of)} batches, "
          f"{len(shared)} shared by several geometries")
    assert shared  # the smaller geometries share batches


def test_one_long_sweep_splits_evenly(efie_mod):
    """A single problem longer than one batch is split evenly (201
    frequencies at n = 6009: 4 batches of 51/50/50/50, as efie_sweep)."""
    plan = efie_mod.pack_plan([6009], [201], free_bytes=280e9)
    assert [int(r[3]) for r in plan] == [51, 50, 50, 50]


def test_explicit_batch_and_small_problems(efie_mod):
    rng = np.random.default_rng(5)
    n = rng.integers(20, 3000, 40)
    nf = rng.integers(0, 70, 40)
    for batch in (0, 16, 100):
        plan = efie_mod.pack_plan(n, nf, batch=batch, free_bytes=200e9)
        _check_plan(plan, n, nf, 1, batch, 200e9)
    # tight memory: batches shrink to what two buffers of 60 % hold
    plan = efie_mod.pack_plan(n, nf, free_bytes=2e9)
    _check_plan(plan, n, nf, 1, 0, 2e9)


def test_plan_errors(efie_mod):
    with pytest.raises(efie_mod.EfieError):
        efie_mod.pack_plan([100, 200], [3, -1])
    with pytest.raises(efie_mod.EfieError):
        efie_mod.pack_plan([0], [3])
    assert efie_mod.pack_plan([100], [0]).shape == (0, 6)
