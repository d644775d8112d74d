"""Median centroid policies on the GPU (L1, TVD, geometric median for L2;
centroid.h:434-647) against the oracle's restatement: whole Lloyd runs through
mcg_lloyd, dense and CSR, weighted and unweighted.  Centers are order statistics
(L1, TVD) or FP64 sums in the oracle's order (geometric median), so for the same
assignments they are bit-identical; costs are FP32 pair sums against FP64."""
import numpy as np
import pytest

from minicore_amd.datagen import gaussian_mixture, poisson_csr

pytestmark = pytest.mark.gpu

L1, L2, TVD = 0, 1, 9


def say(*a):
    print(*a, flush=True)


def _data(msr, n=3000, d=24, k=7, seed=3):
    X, lab, _ = gaussian_mixture(n, d, k, seed=seed, lo=-8, hi=8)
    if msr == TVD:
        X = np.exp(X / 4.).astype(np.float32)
    # starting centers: one row of each component (clusters keep distinct sizes, odd and even)
    C0 = np.stack([X[np.nonzero(lab == q)[0][0]] for q in range(k)]).astype(np.float32)
    return X, C0


def _compare(got, ref, rtol=1e-5):
    assert np.array_equal(got["asn"], ref["asn"])
    assert got["numiter"] == ref["numiter"]
    np.testing.assert_allclose(got["costs"], ref["costs"], rtol=rtol, atol=1e-5)
    assert got["initcost"] == pytest.approx(ref["initcost"], rel=rtol)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=rtol)
    assert np.array_equal(got["centers"], ref["centers"])


@pytest.mark.parametrize("msr", [L1, TVD, L2])
@pytest.mark.parametrize("weighted", [False, True])
def test_lloyd_median_dense(mcg, ctx, mco, msr, weighted):
    say(f"median lloyd dense msr={msr} weighted={weighted}")
    X, C0 = _data(msr)
    w = np.random.default_rng(5).random(len(X)) + 0.25 if weighted else None
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.lloyd(msr, C0, weights=w, eps=1e-10, maxiter=30)
    ref = mco.lloyd_median(msr, X, C0, w=w, eps=1e-10, maxiter=30)
    _compare(got, ref)
    sizes = np.bincount(got["asn"], minlength=len(C0))
    say("  numiter", got["numiter"], "sizes", sizes.tolist())
    ds.close()


@pytest.mark.parametrize("msr", [L1, TVD])
def test_lloyd_median_csr(mcg, ctx, mco, msr):
    say(f"median lloyd csr msr={msr}")
    n, d, k = 2500, 100, 6
    data, idx, ptr = poisson_csr(n, d, 0.3, 11, normalize=False)
    X = mco.CSR(data, idx, ptr, (n, d))
    Xd = X.todense()
    rng = np.random.default_rng(7)
    C0 = Xd[rng.choice(n, k, replace=False)].copy()
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    got = ds.lloyd(msr, C0, eps=1e-10, maxiter=20)
    ref = mco.lloyd_median(msr, X, C0, eps=1e-10, maxiter=20)
    agree = np.mean(got["asn"] == ref["asn"])
    say("  numiter", got["numiter"], ref["numiter"], "agree", agree)
    if agree == 1.0:
        _compare(got, ref, rtol=1e-4)
    else:
        # FP32 pair sums against FP64: a row whose two best centers tie within rounding
        # may go either way (count data tie exactly in L1, and TVD is 1 for disjoint
        # supports); elsewhere the assignment is the oracle's argmin for the GPU centers
        D = mco.cost_matrix(msr, X, got["centers"], 0.)
        srt = np.sort(D, axis=1)
        clear = (srt[:, 1] - srt[:, 0]) > 1e-5 * np.maximum(srt[:, 0], 1e-30)
        assert np.array_equal(got["asn"][clear], D.argmin(1)[clear])
        assert clear.mean() > 0.9
        assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-3)
    ds.close()


def test_median_centers_single_step_csr_bitexact(mcg, ctx, mco):
    """One median M-step on CSR rows from the GPU's own first assignment (maxiter 0:
    E-step, M-step, E-step): accepted, the centers are the oracle's for that
    assignment; reverted, they are the starting centers."""
    n, d, k = 1200, 64, 5
    data, idx, ptr = poisson_csr(n, d, 0.3, 12, normalize=False)
    X = mco.CSR(data, idx, ptr, (n, d))
    C0 = X.todense()[:k].copy()
    for msr in (L1, TVD):
        say(f"median single step csr msr={msr}")
        ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
        a0, c0 = ds.assign(msr, C0, with_costs=True)
        Cm, _, _ = mco.median_centers(msr, X, a0, c0.astype(np.float64), C0)
        got = ds.lloyd(msr, C0, eps=-1., maxiter=0)
        assert np.array_equal(got["centers"], Cm if got["numiter"] == 1 else C0), msr
        ds.close()


def test_geometric_median_csr(mcg, ctx, mco):
    """The geometric median of CSR rows (csc.h:1131-1203 with position-indexed member
    costs: FP64 distances, 1e-80 floor, stop when the FP64 cost is not below the last):
    one M-step from the GPU's own assignment equals the oracle's centers; a full run
    matches the oracle's assignment and cost."""
    say("geometric median csr")
    n, d, k = 1500, 48, 5
    data, idx, ptr = poisson_csr(n, d, 0.3, 21, normalize=False)
    X = mco.CSR(data, idx, ptr, (n, d))
    C0 = X.todense()[:k] + np.float32(.25)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    a0, c0 = ds.assign(L2, C0, with_costs=True)
    Cm, _, _ = mco.median_centers(L2, X, a0, c0.astype(np.float64), C0)
    got = ds.lloyd(L2, C0, eps=-1., maxiter=0)
    assert np.array_equal(got["centers"], Cm if got["numiter"] == 1 else C0)
    got = ds.lloyd(L2, C0, eps=1e-10, maxiter=8)
    ref = mco.lloyd_median(L2, X, C0, eps=1e-10, maxiter=8)
    assert got["numiter"] == ref["numiter"]
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=1e-4, atol=1e-5)
    assert np.mean(got["asn"] == ref["asn"]) > 0.999
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-5)
    ds.close()


def test_median_restart_empty_cluster(mcg, ctx, mco):
    say("median restart")
    X, C0 = _data(L1, n=2000, d=8, k=5, seed=9)
    C0 = np.concatenate([C0, np.full((1, 8), 1e4, np.float32)])   # a center no row takes
    for msr in (L1, L2):
        ds = mcg.Dataset.dense(X, ctx)
        got = ds.lloyd(msr, C0, eps=1e-10, maxiter=10)
        ref = mco.lloyd_median(msr, X, C0, eps=1e-10, maxiter=10)
        assert np.bincount(got["asn"], minlength=6).min() > 0
        _compare(got, ref)
        ds.close()


def test_median_errors(mcg, ctx):
    n, d = 300, 20
    data, idx, ptr = poisson_csr(n, d, 0.3, 13, normalize=False)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    C0 = np.ones((3, d), np.float32)
    with pytest.raises(mcg.MCGError) as e:
        ds.lloyd(L1, C0, weights=np.ones(n), maxiter=2)
    assert e.value.code == -4 and "weighted L1 median" in str(e.value)
    with pytest.raises(mcg.MCGError) as e:
        ds.lloyd(L2, C0, weights=np.ones(n), maxiter=2)
    assert e.value.code == -4 and "weighted geometric median" in str(e.value)
    neg = data.copy()
    neg[3] = -1.
    dn = mcg.Dataset.csr(neg, idx, ptr, (n, d), ctx)
    with pytest.raises(mcg.MCGError) as e:
        dn.lloyd(TVD, C0, maxiter=2)
    assert e.value.code == -1
    X = np.random.default_rng(1).random((n, d)).astype(np.float32)
    dd = mcg.Dataset.dense(X, ctx)
    w = np.ones(n)
    w[5] = 0.
    with pytest.raises(mcg.MCGError) as e:
        dd.lloyd(L1, X[:3], weights=w, maxiter=2)
    assert e.value.code == -1
    for x in (ds, dn, dd):
        x.close()


def test_hcluster_l1_binding(mcg, mco):
    import minicore_amd as mc
    X, C0 = _data(L1, n=1500, d=10, k=4, seed=4)
    r = mc.hcluster(X, C0, msr="L1", maxiter=20)
    ref = mco.lloyd_median(L1, X, C0, eps=1e-10, maxiter=20)
    assert np.array_equal(np.asarray(r["asn"]), ref["asn"])
    assert np.array_equal(np.asarray(r["centers"]), ref["centers"])


@pytest.mark.parametrize("shape", [(1000, 37, 70), (64, 32, 64), (130, 300, 5), (77, 1, 1), (700, 45, 1)])
def test_tiled_kernels_equal_pair_kernel(mcg, ctx, shape):
    """Dense L1 / L2 / TVD costs come from the tiled kernels, CSR ones from the
    pair-by-pair kernel; both sum each pair over the dimensions in order with the same
    smoothing, so a dense matrix and its CSR form give identical costs and assignments
    (ragged n, k and d)."""
    import scipy.sparse as sp
    n, d, k = shape
    say(f"tiled vs pair n={n} d={d} k={k}")
    rng = np.random.default_rng(n + d + k)
    X = rng.normal(size=(n, d)).astype(np.float32)
    X[rng.random((n, d)) < 0.3] = 0.
    C = rng.normal(size=(k, d)).astype(np.float32)
    S = sp.csr_matrix(X)
    dd = mcg.Dataset.dense(X, ctx)
    dc = mcg.Dataset.csr(S.data.astype(np.float32), S.indices.astype(np.uint32), S.indptr.astype(np.uint64),
                         (n, d), ctx)
    A = np.abs(X)
    SA = sp.csr_matrix(A)
    dp = mcg.Dataset.dense(A, ctx)
    dq = mcg.Dataset.csr(SA.data.astype(np.float32), SA.indices.astype(np.uint32), SA.indptr.astype(np.uint64),
                         (n, d), ctx)
    Cp = np.abs(C) + 0.1       # TVD with positive centers
    for msr in (L1, L2, TVD):
        Cm = Cp if msr == TVD else C
        a = (dd if msr != TVD else dp).cost_matrix(msr, Cm, prior=0.01 if msr == TVD else 0.)
        b = (dc if msr != TVD else dq).cost_matrix(msr, Cm, prior=0.01 if msr == TVD else 0.)
        assert np.array_equal(a, b), msr
        pr = 0.01 if msr == TVD else 0.
        ad, cd = (dd if msr != TVD else dp).assign(msr, Cm, prior=pr, with_costs=True)
        ac, cc = (dc if msr != TVD else dq).assign(msr, Cm, prior=pr, with_costs=True)
        assert np.array_equal(ad, ac) and np.array_equal(cd, cc)
        assert np.array_equal(ad, a.argmin(1))
    for x in (dd, dc, dp, dq):
        x.close()
