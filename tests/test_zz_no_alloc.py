"""The allocation-free dense solve (efie_zgesv_noalloc) against efie_zgesv
through the C ABI: the same systems give the same solutions (both run the
same kernels; the scratch array only replaces where the buffers live), on
two calls in a row, the second smaller, so the second runs on what the
first left in the array.  FP64 results are required to be bit-identical,
MIXED ones within 1e-13 of the solution's norm (both refine to 1e-14); a
system the FP32 factorisation cannot solve goes to the FP64 LU in both; a
size beyond the array is an efie:args error before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(efie_mod):
    c = efie_mod.Context(0)
    yield c
    c.close()


def _systems(n, batch, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))
    A += n ** 0.5 * np.eye(n)
    b = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
    return A, b


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_zgesv_two_calls(efie_mod, ctx, prec):
    for n, batch, seed in ((700, 3, 1), (300, 2, 2)):
        A, b = _systems(n, batch, seed)
        x_ref, st_ref = ctx.zgesv(A, b, precision=prec, stats=True)
        x, st = ctx.zgesv(A, b, precision=prec, stats=True, noalloc=True)
        d = np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref))
        print(f"prec {prec} n {n} batch {batch}: max |dx|/|x| {d:.2e}, identical {np.array_equal(x, x_ref)}",
              flush=True)
        if prec == 0:
            assert np.array_equal(x, x_ref)
        else:
            assert d <= 1e-13
            assert st["n_fp64_retry"] == st_ref["n_fp64_retry"] == 0
        berr = max(np.linalg.norm(A[q] @ x[q] - b[q]) / (np.linalg.norm(A[q], 2) * np.linalg.norm(x[q]))
                   for q in range(batch))
        assert berr <= 1e-14


def test_noalloc_fp64_fallback(efie_mod, ctx):
    n = 300
    rng = np.random.default_rng(5)
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = (Q1 * np.logspace(0, -10, n)) @ Q2.conj().T
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x_ref, st_ref = ctx.zgesv(A, b, precision=1, stats=True)
    x, st = ctx.zgesv(A, b, precision=1, stats=True, noalloc=True)
    assert st["n_fp64_retry"] == st_ref["n_fp64_retry"] == 1
    assert np.array_equal(x, x_ref)


def test_noalloc_size_limit(efie_mod, ctx):
    n = 2100   # 16 n (n + 32) bytes > 64 MiB
    A = np.eye(n, dtype=np.complex128)
    b = np.ones(n, dtype=np.complex128)
    with pytest.raises(efie_mod.EfieError) as e:
        ctx.zgesv(A, b, precision=0, noalloc=True)
    assert e.value.ident == "efie:args"
    with pytest.raises(efie_mod.EfieError) as e:
        ctx.zgesv(np.eye(4)[None].repeat(65, 0), np.ones((65, 4)), precision=0, noalloc=True)
    assert e.value.ident == "efie:args"
