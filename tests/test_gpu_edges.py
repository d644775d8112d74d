"""Edge shapes on the GPU path against the oracle: dims beyond the BF16 kernels'
512 (exact fallback), a single row or center, one dimension, k = n (zero cost: the
reference returns (0, 0, 0) without iterating, solve.h:135-138), identical rows,
and CSR rows with no stored entries under the prior measures."""
import numpy as np
import pytest

from minicore_amd.datagen import gaussian_mixture, poisson_csr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_dense_d_beyond_bf16_kernels(mcg, ctx, mco):
    say("d = 600 Lloyd")
    X, lab, _ = gaussian_mixture(3000, 600, 6, 31, lo=-2, hi=2)
    C0 = X[[int(np.argmax(lab == c)) for c in range(6)]].copy()
    ds = mcg.Dataset.dense(X, ctx)
    for prec in (mcg.PREC_BF16, mcg.PREC_FP32):
        got = ds.lloyd(mcg.SQRL2, C0, eps=1e-10, maxiter=10, precision=prec)
        ref = mco.lloyd(mco.SQRL2, X, C0, eps=1e-10, maxiter=10)
        assert np.array_equal(got["asn"], ref["asn"]) and got["numiter"] == ref["numiter"]
        np.testing.assert_allclose(got["centers"], ref["centers"], rtol=1e-6, atol=1e-6)
    ds.close()


@pytest.mark.parametrize("n,d,k", [(1, 5, 1), (7, 1, 2), (64, 1, 3), (40, 3, 40)])
def test_small_shapes(mcg, ctx, mco, n, d, k):
    say(f"small n={n} d={d} k={k}")
    rng = np.random.default_rng(n * 31 + d)
    X = rng.normal(size=(n, d)).astype(np.float32)
    C0 = X[:k].copy()
    ds = mcg.Dataset.dense(X, ctx)
    a, c = ds.assign(mcg.SQRL2, C0, with_costs=True)
    ra, rc, _ = mco.assign_sqrl2_f32(X, C0)
    assert np.array_equal(a, ra) and np.array_equal(c, rc)
    got = ds.lloyd(mcg.SQRL2, C0, eps=1e-10, maxiter=5)
    ref = mco.lloyd(mco.SQRL2, X, C0, eps=1e-10, maxiter=5)
    assert np.array_equal(got["asn"], ref["asn"])
    assert got["numiter"] == ref["numiter"]
    if k == n:   # every row is a center: zero cost, no iteration (solve.h:135-138)
        assert got["initcost"] == 0. and got["finalcost"] == 0. and got["numiter"] == 0
    ds.close()


def test_identical_rows(mcg, ctx, mco):
    X = np.tile(np.arange(8, dtype=np.float32), (500, 1))
    ds = mcg.Dataset.dense(X, ctx)
    r = ds.lloyd(mcg.SQRL2, X[:3].copy(), maxiter=4)
    assert r["initcost"] == 0. and r["numiter"] == 0 and np.all(r["asn"] == 0)
    c, a, dist = ds.kmeanspp(mcg.SQRL2, 1, seed=3)
    assert len(c) == 1 and np.all(dist == 0.)
    ds.close()


@pytest.mark.parametrize("code", [4, 5, 11, 9, 64])
def test_csr_empty_rows(mcg, ctx, mco, code):
    say(f"csr empty rows msr={code}")
    n, d = 400, 50
    data, idx, ptr = poisson_csr(n, d, 0.1, 32, normalize=False)
    # every 7th row without stored entries
    keep = np.ones(len(data), bool)
    lens = np.diff(ptr.astype(np.int64))
    rows = np.repeat(np.arange(n), lens)
    keep[rows % 7 == 0] = False
    data, idx = data[keep], idx[keep]
    lens = np.where(np.arange(n) % 7 == 0, 0, lens)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    X = mco.CSR(data, idx, ptr, (n, d))
    C0 = X.todense()[[1, 2, 3, 4]] + 0.25
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    got = ds.cost_matrix(code, C0, prior=0.01)
    ref = mco.cost_matrix(code, X, C0, 0.01)
    np.testing.assert_allclose(got, ref, rtol=2e-4, atol=1e-5)
    r = ds.lloyd(code, C0, prior=0.01, maxiter=3)
    assert np.isfinite(r["finalcost"]) and r["finalcost"] <= r["initcost"] * (1 + 1e-6)
    ds.close()
