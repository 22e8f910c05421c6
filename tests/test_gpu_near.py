"""The device's near-pair search (centroid cell binning, no Nt x Nt map)
against the row-wise restatement of triangles_are_near.m
(tests/near_np.py, itself equal to the oracle's per-pair test on c1): the
same pairs in the same order on c1, c2_6k and c2_13k at both option
presets' near_sing_eps (2.5 accurate, 1.2 fast)."""
import numpy as np
import pytest

import near_np
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["c1", "c2_6k", "c2_13k"])
def test_near_pairs_on_device(efie_mod, name):
    V, F = efie_mod.make_patch_with_ground_mesh(*W.mesh_args(name))
    r = efie_mod.build_rwg_interior(V, F)
    with efie_mod.Context(0) as ctx:
        ctx.set_problem(V, F, r)
        for eps in (2.5, 1.2):
            got = ctx.near_pairs(efie_mod.opts(**dict(W.OPTS_C1, near_sing_eps=eps)))
            want = near_np.near_pairs(V, F, eps)
            print(f"{name} (Nt={F.shape[0]}) eps {eps}: {len(got)} near pairs", flush=True)
            assert got.shape == want.shape and np.array_equal(got, want)
