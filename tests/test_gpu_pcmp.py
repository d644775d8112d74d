"""pcmp (python/pycmp.cpp:249-338): condensed pairwise costs msr(first = row i,
second = row j), i < j.  Every value equals the cost matrix's (same kernels, computed
block by block, with block edges crossed at n > 4096) and is within FP32 tolerance of
the oracle's FP64 costs."""
import numpy as np
import pytest

import minicore_amd as mc
from minicore_amd.datagen import gaussian_mixture, poisson_csr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def condensed(M):
    n = M.shape[0]
    iu = np.triu_indices(n, 1)
    return M[iu]


@pytest.mark.parametrize("msr", [2, 0, 1, 4, 9])
def test_pcmp_dense_equals_cost_matrix(mcg, ctx, mco, msr):
    n, d = 4500, 7   # crosses the 4096-row blocks
    say(f"pcmp dense msr={msr} n={n}")
    X, _, _ = gaussian_mixture(n, d, 5, 3, lo=-3, hi=3)
    if msr in (4, 9):
        X = np.exp(X / 2.).astype(np.float32)
    prior = 0.1 if msr in (4, 9) else 0.
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.pairwise(msr, prior)
    full = ds.cost_matrix(msr, X, prior)
    ds.close()
    assert got.shape == (n * (n - 1) // 2,)
    assert np.array_equal(got, condensed(full))
    sub = np.arange(0, n, 37)
    ref = mco.cost_matrix(msr, X[sub], X[sub], prior)
    pos = {int(v): q for q, v in enumerate(sub)}
    ii, jj = np.triu_indices(len(sub), 1)
    I, J = sub[ii], sub[jj]
    idx = n * I - I * (I + 1) // 2 + (J - I - 1)
    np.testing.assert_allclose(got[idx], ref[ii, jj], rtol=2e-4, atol=1e-4)


@pytest.mark.parametrize("msr", [4, 5, 0])
def test_pcmp_csr_and_binding(mcg, ctx, mco, msr):
    say(f"pcmp csr msr={msr}")
    n, d = 300, 120
    data, idx, ptr = poisson_csr(n, d, 0.1, 21, normalize=False)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    got = ds.pairwise(msr, 0.05)
    X = mco.CSR(data, idx, ptr, (n, d))
    full = ds.cost_matrix(msr, X.todense(), 0.05)
    ds.close()
    assert np.array_equal(got, condensed(full))
    ref = mco.cost_matrix(msr, X, X.todense(), 0.05)
    np.testing.assert_allclose(got, condensed(ref), rtol=2e-4, atol=1e-5)
    import scipy.sparse as sp
    S = sp.csr_matrix((data, idx, ptr), shape=(n, d))
    b = mc.pcmp(mc.CSparseMatrix(S), msr=msr, prior=0.05)
    assert b.dtype == np.float32 and np.array_equal(b, got)


def test_pcmp_dense_binding_dtypes():
    X = np.random.default_rng(4).random((50, 6))
    a = mc.pcmp(X.astype(np.float32), msr=2)
    b = mc.pcmp(X, msr=2)                       # float64 input: float64 output
    c = mc.pcmp(X, msr=2, use_float=1)
    assert a.dtype == np.float32 and b.dtype == np.float64 and c.dtype == np.float32
    assert a.shape == (50 * 49 // 2,)
    i, j = 3, 17
    k = 50 * i - i * (i + 1) // 2 + (j - i - 1)
    assert a[k] == pytest.approx(float(((X[i] - X[j]) ** 2).sum()), rel=1e-5)
    assert len(mc.pcmp(X[:1].astype(np.float32))) == 0
