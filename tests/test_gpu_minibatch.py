"""Minibatch and coreset-minibatch hard clustering (perform_hard_minibatch_clustering,
solve.h:377-578; hmb_coreset_clustering, solve.h:598-789) on the GPU against the
oracle's restatements: same draws, restarts, coresets and best-centers bookkeeping;
dense SQRL2 bit for bit, CSR JSD within FP32 tolerance."""
import numpy as np
import pytest

import minicore_amd as mc
from minicore_amd.datagen import gaussian_mixture, poisson_csr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _start(X, lab, k):
    return X[[int(np.argmax(lab == c)) for c in range(k)]].copy()


@pytest.mark.parametrize("weighted,with_rep,prec", [(False, False, 1), (True, False, 1), (False, True, 1),
                                                    (True, True, 2)])
def test_minibatch_dense_bitexact(mcg, ctx, mco, weighted, with_rep, prec):
    say(f"minibatch dense weighted={weighted} with_rep={with_rep} precision={prec}")
    n, d, k = 20000, 50, 25
    X, lab, _ = gaussian_mixture(n, d, k, 0)
    w = np.random.default_rng(5).uniform(0.5, 2., n) if weighted else None
    C0 = _start(X, lab, k)
    C0[3] = 1e3          # an empty cluster: restarted at the first check-in
    kw = dict(mbsize=1500, maxiter=23, freq=5, reseed_after=5, seed=17)
    ref = mco.minibatch(mco.SQRL2, X, C0, w=w, with_rep=with_rep, **kw)
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.minibatch(mcg.SQRL2, C0, weights=w, with_replacement=with_rep, precision=prec, **kw)
    ds.close()
    assert got["numiter"] == ref["numiter"] == 23
    assert got["initcost"] == ref["initcost"] and got["finalcost"] == ref["finalcost"]
    assert np.array_equal(got["centers"], ref["centers"])
    assert np.array_equal(got["asn"], ref["asn"])
    assert np.array_equal(got["costs"], ref["costs"].astype(np.float32))
    assert got["finalcost"] < got["initcost"]


def test_minibatch_csr_jsd(mcg, ctx, mco):
    say("minibatch csr jsd")
    n, d, k = 3000, 400, 6
    data, idx, ptr = poisson_csr(n, d, 0.04, 12, normalize=True)
    dense = np.zeros((n, d), np.float32)
    for i in range(n):
        dense[i, idx[ptr[i]:ptr[i + 1]]] = data[ptr[i]:ptr[i + 1]]
    C0 = dense[np.random.default_rng(3).choice(n, k, replace=False)].copy()
    kw = dict(mbsize=600, maxiter=9, freq=4, reseed_after=0, seed=5)
    ref = mco.minibatch(4, mco.CSR(data, idx, ptr, (n, d)), C0, prior=1. / d, **kw)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    got = ds.minibatch(4, C0, prior=1. / d, **kw)
    ds.close()
    assert got["numiter"] == ref["numiter"]
    assert got["initcost"] == pytest.approx(ref["initcost"], rel=2e-4)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=2e-4)
    assert np.mean(got["asn"] == ref["asn"]) > 0.995
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=1e-3, atol=1e-6)


def test_hcluster_minibatch_entry(mcg, mco):
    """hcluster(..., mbsize >= 0) takes the minibatch driver with the reference's
    check-in period ceil(maxiter / ncheckins) and reseed_count."""
    say("api hcluster minibatch")
    X, lab, _ = gaussian_mixture(6000, 16, 8, 2)
    C0 = _start(X, lab, 8)
    r = mc._ext().hcluster(X, C0, maxiter=30, seed=3, mbsize=700, ncheckins=4, reseed_count=2, with_rep=True)
    ref = mco.minibatch(mco.SQRL2, X, C0, mbsize=700, maxiter=30, freq=8, reseed_after=2, with_rep=True, seed=3)
    assert set(r) == {"initcost", "finalcost", "numiter", "centers", "costs", "asn"}
    assert r["numiter"] == ref["numiter"] and r["finalcost"] == ref["finalcost"]
    assert np.array_equal(r["centers"], ref["centers"]) and np.array_equal(r["asn"], ref["asn"])
    r = mc._ext().hcluster(X, C0, maxiter=30, seed=3, mbsize=700, ncheckins=4, reseed_count=2, cs=True)
    ref = mco.cs_minibatch(mco.SQRL2, X, C0, mbsize=700, maxiter=30, freq=8, reseed_after=2, seed=3)
    assert r["numiter"] == ref["numiter"] and r["finalcost"] == ref["finalcost"]
    assert np.array_equal(r["centers"], ref["centers"]) and np.array_equal(r["asn"], ref["asn"])


@pytest.mark.parametrize("weighted", [False, True])
def test_cs_minibatch_dense_bitexact(mcg, ctx, mco, weighted):
    say(f"coreset minibatch dense weighted={weighted}")
    n, d, k = 20000, 50, 25
    X, lab, _ = gaussian_mixture(n, d, k, 0)
    w = np.random.default_rng(6).uniform(0.5, 2., n) if weighted else None
    C0 = _start(X, lab, k)
    C0[5] = 1e3
    kw = dict(mbsize=2000, maxiter=20, freq=5, reseed_after=5, seed=11)
    ref = mco.cs_minibatch(mco.SQRL2, X, C0, w=w, **kw)
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.cs_minibatch(mcg.SQRL2, C0, weights=w, **kw)
    ds.close()
    assert got["numiter"] == ref["numiter"] == 5
    assert got["initcost"] == ref["initcost"] and got["finalcost"] == ref["finalcost"]
    assert np.array_equal(got["centers"], ref["centers"])
    assert np.array_equal(got["asn"], ref["asn"])
    assert np.array_equal(got["costs"], ref["costs"].astype(np.float32))
    assert got["finalcost"] < got["initcost"]


def test_cs_minibatch_csr_jsd(mcg, ctx, mco):
    say("coreset minibatch csr jsd")
    n, d, k = 3000, 400, 6
    data, idx, ptr = poisson_csr(n, d, 0.04, 13, normalize=True)
    dense = np.zeros((n, d), np.float32)
    for i in range(n):
        dense[i, idx[ptr[i]:ptr[i + 1]]] = data[ptr[i]:ptr[i + 1]]
    C0 = dense[np.random.default_rng(8).choice(n, k, replace=False)].copy()
    kw = dict(mbsize=800, maxiter=6, freq=3, reseed_after=0, seed=2)
    ref = mco.cs_minibatch(4, mco.CSR(data, idx, ptr, (n, d)), C0, prior=1. / d, **kw)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    got = ds.cs_minibatch(4, C0, prior=1. / d, **kw)
    ds.close()
    assert got["numiter"] == ref["numiter"]
    assert got["initcost"] == pytest.approx(ref["initcost"], rel=2e-4)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-3)
    assert np.mean(got["asn"] == ref["asn"]) > 0.99
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=5e-3, atol=1e-5)


def test_minibatch_invalid_sizes(mcg, ctx):
    X, _, _ = gaussian_mixture(500, 4, 3, 1)
    ds = mcg.Dataset.dense(X, ctx)
    for kw in (dict(mbsize=0), dict(maxiter=0), dict(freq=0), dict(mbsize=501, with_replacement=False)):
        with pytest.raises(mcg.MCGError):
            ds.minibatch(mcg.SQRL2, X[:3], **kw)
    ds.close()
