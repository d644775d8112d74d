"""The pyminicore-compatible entry points on the GPU: return types and values
match the reference's bindings and the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

import minicore_amd as mc
from minicore_amd.datagen import gaussian_mixture, poisson_csr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_hcluster_dense_matches_oracle(mco):
    say("api hcluster dense")
    X, _, _ = gaussian_mixture(20000, 50, 25, 0)
    ids, asn, costs = mc.kmeanspp(X, 25, seed=13)
    assert ids.dtype == np.uint32 and asn.dtype == np.uint8 and costs.dtype == np.float32
    c_ref, a_ref, d_ref = mco.kmeanspp(mco.SQRL2, X, 25, seed=13)
    assert np.array_equal(ids, c_ref) and np.array_equal(asn, a_ref) and np.array_equal(costs, d_ref)
    r = mc.hcluster(X, ids, maxiter=100)
    ref = mco.lloyd(mco.SQRL2, X, X[ids], eps=1e-10, maxiter=100)
    assert set(r) == {"initcost", "finalcost", "numiter", "centers", "costs", "asn"}
    assert r["asn"].dtype == np.uint32 and r["costs"].dtype == np.float32
    assert r["centers"].shape == (25, 50) and r["centers"].dtype == np.float32
    assert r["numiter"] == ref["numiter"] and np.array_equal(r["asn"], ref["asn"])
    assert r["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-6)


def test_hcluster_csr_and_centers_tuple(mco):
    say("api hcluster csr")
    data, idx, ptr = poisson_csr(1500, 300, 0.05, 3, normalize=True)
    m = sp.csr_matrix((data, idx.astype(np.int32), ptr.astype(np.int64)), shape=(1500, 300))
    ids, asn, costs = mc.kmeanspp(m, 8, msr=5, prior=1. / 300, seed=4)
    assert costs.dtype == np.float64
    r = mc.hcluster(m, ids, msr=5, prior=1. / 300, maxiter=10)
    (cd, ci, cp), shape = r["centers"]
    assert list(shape) == [8, 300] and ci.dtype == np.uint32 and cp.dtype == np.uint64
    C = sp.csr_matrix((cd, ci, cp), shape=tuple(shape)).toarray()
    got = mco.assign_f64(mco.MKL, mco.CSR(data, idx, ptr, (1500, 300)), C, 1. / 300)[0]
    assert np.mean(got == r["asn"]) > 0.995


def test_scluster_returns(mco):
    say("api scluster")
    rng = np.random.default_rng(1)
    X = rng.poisson(3., (500, 12)).astype(np.float32)
    r = mc.scluster(X, [0, 1, 2], msr=4, prior=0.5, maxiter=5)
    assert r["costs"].shape == (500, 3) and r["asn"].shape == (500, 3)
    np.testing.assert_allclose(r["asn"].sum(1), 1., rtol=1e-5)
    assert len(r["centers"]) == 2


def test_scluster_median_measures(mco):
    """scluster with L1 / L2: centers from the weighted median and the Weiszfeld
    geometric median of all rows, as the oracle; on CSR rows L1 raises as the
    reference's weighted L1 median does."""
    say("api scluster L1 / L2")
    rng = np.random.default_rng(6)
    X = rng.gamma(4., 1., (600, 10)).astype(np.float32)   # continuous: no tied weights
    X[:200] += 6.
    C0 = X[[1, 300, 500]] + np.float32(.5)
    for code in (0, 1):
        r = mc.scluster(X, C0, msr=code, temp=.3, maxiter=4)
        (cd, ci, cp), shape = r["centers"]
        C = sp.csr_matrix((cd, ci, cp), shape=tuple(shape)).toarray()
        ref = mco.soft(code, X, C0, temp=.3, maxiter=4, eps=1e-8)
        assert r["numiter"] == ref["numiter"]
        assert r["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-4)
        np.testing.assert_allclose(C, ref["centers"], rtol=1e-4, atol=1e-4)
    # centers given as row indices: Weiszfeld started on a data row stays on it
    r = mc.scluster(X, [1, 300, 500], msr=1, temp=.3, maxiter=4)
    (cd, ci, cp), shape = r["centers"]
    C = sp.csr_matrix((cd, ci, cp), shape=tuple(shape)).toarray()
    np.testing.assert_allclose(C, X[[1, 300, 500]], rtol=0, atol=1e-20)
    with pytest.raises(RuntimeError):
        mc.scluster(sp.csr_matrix(X), [1, 300, 500], msr=0, maxiter=2)


def test_cmp_orientation(mco):
    say("api cmp")
    rng = np.random.default_rng(2)
    A = rng.poisson(2., (7, 15)).astype(np.float32) + 0.1
    B = rng.poisson(2., (5, 15)).astype(np.float32) + 0.1
    out = mc.cmp(A, B, msr=5, prior=0.1)          # reverse=True: out[i,j] = msr(first=B_j, second=A_i)
    assert out.shape == (7, 5)
    ref = mco.cost_matrix(mco.MKL, B, A, 0.1).T
    np.testing.assert_allclose(out, ref, rtol=2e-4)
    S = sp.csr_matrix(A)
    out2 = mc.cmp(S, B, msr=5, prior=0.1)         # out[x,y] = msr(first=B_y, second=A_x)
    np.testing.assert_allclose(out2, ref, rtol=2e-4)


def test_coreset_sampler_class(mco):
    say("api coreset sampler")
    rng = np.random.default_rng(3)
    costs = rng.exponential(1., 4000).astype(np.float32)
    asn = rng.integers(0, 10, 4000).astype(np.uint32)
    cs = mc.CoresetSampler()
    cs.make_sampler(10, costs, asn, sens=4)
    p = cs.get_probs()
    assert np.array_equal(p, mco.coreset_probs(4, 10, costs.astype(np.float64), asn))
    w, i = cs.sample(300, seed=5)
    wr, ir = mco.coreset_sample(p, 300, 5)
    assert np.array_equal(i, ir) and np.array_equal(w, wr)
    wu, iu = cs.sample(300, seed=5, unique=True)
    assert len(np.unique(iu)) == 300 and np.all(np.diff(iu.astype(np.int64)) > 0)


def test_reference_coreset_example(mco):
    """The reference's python/examples/coreset_example.py, unchanged but for the
    module name: seeding, an LBK sampler from mc.constants.SENSDICT, 100 draws."""
    say("api coreset example")
    dat = np.random.default_rng(0).random(10000).reshape(1000, 10)
    out, asn, costs = mc.kmeanspp(dat, k=25, msr=2)
    cs = mc.CoresetSampler()
    sensid = mc.constants.SENSDICT["LBK"]
    cs.make_sampler(25, costs=costs, assignments=asn, sens=sensid)
    weights, ids = cs.sample(100)
    assert len(ids) == len(weights) == 100 and (np.asarray(weights) > 0).all()
    cs2 = mc.CoresetSampler()
    cs2.make_sampler(25, costs=costs, assignments=asn, sens=mc.constants.SENSDICT["LFKF"])
    p = cs2.get_probs()
    cnt = np.bincount(np.asarray(asn, dtype=np.int64), minlength=25)
    np.testing.assert_allclose(p, 1. / (cnt[np.asarray(asn, dtype=np.int64)] * (cnt > 0).sum()), rtol=1e-5)
