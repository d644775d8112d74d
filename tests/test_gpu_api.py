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


def test_sparse_matrix_wrapper_overloads(mco):
    """The SparseMatrixWrapper overloads (pycluster.cpp:6-26, pycluster_soft.cpp:4-28,
    pycmp.cpp:16-146, pysmw.cpp:368-381) and kmeanspp(matrix, SumOpts) for both sparse
    types (pysmw.cpp:374-381, pycsparse.cpp:91-99): each against the oracle, and the
    wrapper's results equal the CSparseMatrix path on the same rows."""
    say("api SparseMatrixWrapper")
    n, d = 1200, 250
    data, idx, ptr = poisson_csr(n, d, 0.05, 7, normalize=True)
    m = sp.csr_matrix((data, idx.astype(np.int32), ptr.astype(np.int64)), shape=(n, d))
    X = mco.CSR(data, idx, ptr, (n, d))
    w = mc.SparseMatrixWrapper(m)
    prior = 1. / d
    # kmeanspp keyword form: the same seeding as the CSparseMatrix path
    ids, asn, costs = mc.kmeanspp(w, 6, msr=5, prior=prior, seed=9)
    ids2, asn2, costs2 = mc.kmeanspp(m, 6, msr=5, prior=prior, seed=9)
    assert np.array_equal(ids, ids2) and np.array_equal(asn, asn2) and np.array_equal(costs, costs2)
    assert asn.dtype == np.uint8 and costs.dtype == np.float64 and len(set(ids.tolist())) == 6
    K = mco.cost_matrix(mco.MKL, X, X.todense()[ids.astype(np.int64)], prior)
    np.testing.assert_allclose(costs, K[np.arange(n), asn], rtol=2e-4, atol=1e-7)
    # kmeanspp(matrix, SumOpts): k, measure and prior from the options (random seed: 0)
    so = mc.SumOpts("MKL", k=5, prior=prior)
    for mat in (w, mc.CSparseMatrix(m)):
        i3, a3, c3 = mc._ext().kmeanspp(mat, so)
        assert len(i3) == 5 and len(set(i3.tolist())) == 5 and a3.dtype == np.uint8
        K3 = mco.cost_matrix(mco.MKL, X, X.todense()[i3.astype(np.int64)], prior)
        np.testing.assert_allclose(c3, K3[np.arange(n), a3], rtol=2e-4, atol=1e-7)
    # hcluster / scluster: equal to the CSparseMatrix path
    r = mc.hcluster(w, ids, msr=5, prior=prior, maxiter=10)
    r2 = mc.hcluster(m, ids, msr=5, prior=prior, maxiter=10)
    assert r["numiter"] == r2["numiter"] and np.array_equal(r["asn"], r2["asn"]) and r["finalcost"] == r2["finalcost"]
    (cd, ci, cp), shape = r["centers"]
    C = sp.csr_matrix((cd, ci, cp), shape=tuple(shape)).toarray()
    a_ref, _, gap = mco.assign_f64(mco.MKL, X, C, prior)
    assert np.all(gap[a_ref != r["asn"]] < 1e-4)
    s1 = mc.scluster(w, ids, msr=11, prior=prior, maxiter=3)
    s2 = mc.scluster(m, ids, msr=11, prior=prior, maxiter=3)
    assert np.array_equal(s1["asn"], s2["asn"]) and np.array_equal(s1["costs"], s2["costs"])
    ref = mco.soft(11, X, X.todense()[ids.astype(np.int64)], prior=prior, maxiter=3)
    np.testing.assert_allclose(s1["costs"], ref["costs"], rtol=2e-4)
    # cmp(smw, array): out[x, y] = msr(first = array row y, second = matrix row x); with
    # reverse msr(first = matrix row x, second = array row y); 1-d arrays likewise
    B = np.random.default_rng(2).random((4, d)).astype(np.float32)
    out = mc._ext().cmp(w, B, msr=5, prior=prior)
    np.testing.assert_allclose(out, mco.cost_matrix(mco.MKL, B, X.todense(), prior).T, rtol=2e-4)
    outr = mc._ext().cmp(w, B, msr=5, prior=prior, reverse=True)
    np.testing.assert_allclose(outr, mco.cost_matrix(mco.MKL, X, B, prior), rtol=2e-4)
    out1 = mc._ext().cmp(w, B[1], msr=5, prior=prior, reverse=True)
    np.testing.assert_allclose(out1, outr[:, 1], rtol=1e-6)
    # cmp(smw, smw): out[l, r] = msr(first = rhs row r, second = lhs row l); use_float=0 gives float64
    w2 = mc.SparseMatrixWrapper(m[:30])
    o = mc._ext().cmp(w2, w, msr=4, prior=prior)
    assert o.shape == (30, n) and o.dtype == np.float32
    np.testing.assert_allclose(o, mco.cost_matrix(mco.JSD, X, X.todense()[:30], prior).T, rtol=2e-4, atol=1e-7)
    orv = mc._ext().cmp(w2, w, msr=5, prior=prior, reverse=True, use_float=0)
    assert orv.dtype == np.float64
    Xs = mco.CSR(data[:int(ptr[30])], idx[:int(ptr[30])], ptr[:31].copy(), (30, d))
    # self-distances are 0 in FP64 and a few 1e-7 after the FP32 cancellation of the
    # decomposed terms (costs here are ~1.6)
    np.testing.assert_allclose(orv, mco.cost_matrix(mco.MKL, Xs, X.todense(), prior), rtol=2e-4, atol=2e-6)


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
