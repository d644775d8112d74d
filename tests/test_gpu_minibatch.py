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


def _median_compare(got, ref, tol=1e-5):
    """Median-policy runs: FP32 pair costs on the GPU against FP64 in the oracle, so
    the check-in totals agree to rounding; draws, restarts and the best-centers
    choice follow them, and then centers and assignment are identical."""
    assert got["numiter"] == ref["numiter"]
    assert got["initcost"] == pytest.approx(ref["initcost"], rel=tol)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=tol)
    assert np.array_equal(got["asn"], ref["asn"])
    assert np.array_equal(got["centers"], ref["centers"])


@pytest.mark.parametrize("msr,weighted,with_rep", [(0, False, False), (0, True, False), (9, False, True),
                                                   (9, True, False), (1, False, False), (1, True, True)])
def test_minibatch_median_dense(mcg, ctx, mco, msr, weighted, with_rep):
    """The median updates of solve.h:555-572 (weighted L1 / TVD by member rank, the
    geometric median by member weights) inside the minibatch driver."""
    say(f"minibatch median dense msr={msr} weighted={weighted} with_rep={with_rep}")
    n, d, k = 6000, 16, 10
    X, lab, _ = gaussian_mixture(n, d, k, 4, lo=-6, hi=6)
    if msr == 9:
        X = np.exp(X / 4.).astype(np.float32)
    w = np.random.default_rng(6).uniform(0.5, 2., n) if weighted else None
    C0 = _start(X, lab, k)
    # an empty cluster, restarted at the first check-in (TVD compares distributions,
    # so that center is made a near point mass)
    if msr == 9:
        C0[4] = 1e-3
        C0[4, 0] = 1e3
    else:
        C0[4] = 1e3
    kw = dict(mbsize=900, maxiter=13, freq=4, reseed_after=2, seed=21)
    ref = mco.minibatch(msr, X, C0, w=w, with_rep=with_rep, **kw)
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.minibatch(msr, C0, weights=w, with_replacement=with_rep, **kw)
    ds.close()
    _median_compare(got, ref)


@pytest.mark.parametrize("msr", [0, 1])
def test_minibatch_median_csr_l1(mcg, ctx, mco, msr):
    """CSR rows: the L1 median and the geometric median (FP64 Weiszfeld under the CSR
    rule) inside the minibatch driver."""
    say(f"minibatch median csr msr={msr}")
    n, d, k = 3000, 60, 6
    data, idx, ptr = poisson_csr(n, d, 0.4, 14, normalize=False)
    X = mco.CSR(data, idx, ptr, (n, d))
    C0 = X.todense()[np.random.default_rng(3).choice(n, k, replace=False)].copy()
    if msr == 1:
        C0 += np.float32(.25)
    kw = dict(mbsize=700, maxiter=9, freq=3, reseed_after=1, seed=8)
    ref = mco.minibatch(msr, X, C0, **kw)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    got = ds.minibatch(msr, C0, **kw)
    ds.close()
    agree = np.mean(got["asn"] == ref["asn"])
    say("  agree", agree)
    if agree == 1.0:
        _median_compare(got, ref, tol=1e-4)
    else:   # count data: exact L1 ties between centers may resolve the other way in FP32
        assert agree > 0.99
        assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-3)


@pytest.mark.parametrize("msr", [0, 1])
def test_cs_minibatch_median_dense(mcg, ctx, mco, msr):
    """Coreset minibatch with a median policy.  The weighted Lloyd on one coreset (the
    oracle's draws from the oracle's LBK probabilities) is bit-identical to the
    oracle's, and so is the whole driver (its LBK probabilities come from FP32 costs
    here and FP64 ones in the oracle; on this data no draw depends on the difference)."""
    say(f"cs minibatch median msr={msr}")
    n, d, k = 5000, 12, 8
    X, lab, _ = gaussian_mixture(n, d, k, 5, lo=-6, hi=6)
    C0 = _start(X, lab, k)
    a, c, _ = mco.assign_f64(msr, X, C0, 0.)
    probs = mco.coreset_probs(4, k, c, a)
    wt, ind = mco.coreset_sample(probs, 800, mco.wyrand_at(9, 0))
    o = np.argsort(ind, kind="stable")
    ind, csw = ind[o], wt[o].astype(np.float64)
    for mi in (0, 3):
        ref = mco.lloyd_median(msr, X[ind], C0, w=csw, eps=1e-3, maxiter=mi)
        ds = mcg.Dataset.dense(X[ind], ctx)
        got = ds.lloyd(msr, C0, weights=csw, eps=1e-3, maxiter=mi)
        ds.close()
        _median_compare(got, ref)
    Xs, labs, _ = gaussian_mixture(n, d, k, 6, lo=-40, hi=40)
    C1 = _start(Xs, labs, k) + 0.5
    kw = dict(mbsize=800, maxiter=8, freq=4, reseed_after=1, seed=9)
    ref = mco.cs_minibatch(msr, Xs, C1, **kw)
    ds = mcg.Dataset.dense(Xs, ctx)
    got = ds.cs_minibatch(msr, C1, **kw)
    ds.close()
    _median_compare(got, ref)


def test_minibatch_median_unsupported(mcg, ctx):
    n, d = 400, 10
    data, idx, ptr = poisson_csr(n, d, 0.4, 15, normalize=False)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    C0 = np.ones((3, d), np.float32)
    with pytest.raises(mcg.MCGError) as e:
        ds.minibatch(0, C0, weights=np.ones(n), mbsize=100, maxiter=3, freq=1)
    assert e.value.code == -4
    with pytest.raises(mcg.MCGError) as e:
        ds.minibatch(1, C0, weights=np.ones(n), mbsize=100, maxiter=3, freq=1)
    assert e.value.code == -4
    ds.close()
    dd = mcg.Dataset.dense(np.random.default_rng(2).random((n, d)).astype(np.float32), ctx)
    with pytest.raises(mcg.MCGError) as e:   # rank-indexed weights need mbsize <= n
        dd.minibatch(0, C0, weights=np.ones(n), mbsize=n + 10, maxiter=3, freq=1, with_replacement=True)
    assert e.value.code == -1
    dd.close()
