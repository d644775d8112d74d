"""Bregman measures (MKL, REVERSE_MKL, JSD) and SQRL2 over CSR and dense rows,
hard and soft clustering, on the GPU against the FP64 oracle."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from minicore_amd.datagen import poisson_csr

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MEASURES = [("SQRL2", 2), ("JSD", 4), ("MKL", 5), ("REVERSE_MKL", 11)]


def say(*a):
    print(*a, flush=True)


def csr_data(n, d, density, seed, normalize):
    data, idx, ptr = poisson_csr(n, d, density, seed, sigma=1.5, normalize=normalize)
    return data, idx, ptr


def centers_from_rows(mco, data, idx, ptr, n, d, k, seed):
    rng = np.random.default_rng(seed)
    rows = rng.choice(n, k, replace=False)
    csr = mco.CSR(data, idx, ptr, (n, d))
    dense = csr.todense()
    return dense[rows].copy()


def test_pair26183_gpu(mcg, ctx):
    say("26183-dim pair")
    import json
    z = np.load(os.path.join(GOLD, "testmsr_pair.npz"))
    ref = json.load(open(os.path.join(GOLD, "reference_cases.json")))["msr_pair26183"]
    d = int(z["d"])
    c = np.zeros((1, d), np.float32); c[0, z["rhs_idx"]] = z["rhs_val"]
    ds = mcg.Dataset.csr(z["lhs_val"], z["lhs_idx"], np.array([0, len(z["lhs_idx"])], np.uint64), (1, d), ctx)
    for name, code in MEASURES:
        v = ds.cost_matrix(code, c, prior=ref["prior"])[0, 0]
        assert v == pytest.approx(ref[name], rel=2e-4), name


@pytest.mark.parametrize("normalize", [False, True])
@pytest.mark.parametrize("name,code", MEASURES)
def test_cost_matrix_csr(mcg, ctx, mco, name, code, normalize):
    say(f"csr cost matrix {name} normalize={normalize}")
    n, d, k = 600, 500, 9
    data, idx, ptr = csr_data(n, d, 0.03, 1, normalize)
    prior = 1. / d
    C = centers_from_rows(mco, data, idx, ptr, n, d, k, 2)
    C = (C + 0.5 * C.mean(0)).astype(np.float32)   # dense-ish centers, like cluster means
    ref = mco.cost_matrix(code, mco.CSR(data, idx, ptr, (n, d)), C, prior)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    got = ds.cost_matrix(code, C, prior=prior)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=2e-4, atol=2e-5 * scale)


@pytest.mark.parametrize("name,code", MEASURES)
def test_cost_matrix_dense_rows(mcg, ctx, mco, name, code):
    say(f"dense-row cost matrix {name}")
    rng = np.random.default_rng(4)
    X = rng.poisson(1.5, (400, 60)).astype(np.float32)
    C = rng.poisson(1.5, (7, 60)).astype(np.float32) + 0.25
    ref = mco.cost_matrix(code, X, C, 0.1)
    got = mcg.Dataset.dense(X, ctx).cost_matrix(code, C, prior=0.1)
    np.testing.assert_allclose(got, ref, rtol=2e-4, atol=2e-5 * np.abs(ref).max())


@pytest.mark.parametrize("d,k", [(50, 40), (100, 130), (256, 70)])
@pytest.mark.parametrize("name,code", [("MKL", 5), ("REVERSE_MKL", 11)])
def test_dense_bregman_mfma(mcg, ctx, mco, name, code, d, k):
    """Dense MKL / REVERSE_MKL with d <= 256 take the FP32 MFMA cross term
    (k_breg_mfma, 32x32x2 tiles): the cost matrix within 2e-4 of the FP64 oracle; the
    assignment equals the FP64 argmin where the FP64 top-2 gap exceeds 1e-4, and its
    costs equal the cost matrix's row minima."""
    say(f"dense bregman mfma {name} d={d} k={k}")
    rng = np.random.default_rng(d + k)
    X = rng.poisson(2.0, (3000, d)).astype(np.float32)
    C = (rng.poisson(2.0, (k, d)) + rng.uniform(0.1, 1., (k, d))).astype(np.float32)
    ref = mco.cost_matrix(code, X, C, 0.05)
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.cost_matrix(code, C, prior=0.05)
    np.testing.assert_allclose(got, ref, rtol=2e-4, atol=2e-5 * np.abs(ref).max())
    a, c = ds.assign(code, C, prior=0.05)
    a_ref, _, gap = mco.assign_f64(code, X, C, 0.05)
    bad = a != a_ref
    assert np.all(gap[bad] < 1e-4), gap[bad][:10]
    np.testing.assert_array_equal(c, got[np.arange(len(a)), a])
    assert np.array_equal(a, np.argmin(got, axis=1))
    ds.close()


@pytest.mark.parametrize("name,code", MEASURES[1:])
def test_tiny_prior_direct_path(mcg, ctx, mco, name, code):
    """prior 0 takes the reference's per-pair smoothing (7.0069e-42 floor)."""
    say(f"prior 0 {name}")
    rng = np.random.default_rng(6)
    X = (rng.poisson(2., (100, 30)) + 1).astype(np.float32)
    C = (rng.poisson(2., (5, 30)) + 1).astype(np.float32)
    ref = mco.cost_matrix(code, X, C, 0.)
    got = mcg.Dataset.dense(X, ctx).cost_matrix(code, C, prior=0.)
    np.testing.assert_allclose(got, ref, rtol=5e-4, atol=1e-5 * np.abs(ref).max())


@pytest.mark.parametrize("name,code", MEASURES)
def test_assign_csr_non_ties(mcg, ctx, mco, name, code):
    say(f"csr assign {name}")
    n, d, k = 3000, 800, 12
    data, idx, ptr = csr_data(n, d, 0.02, 3, True)
    prior = 1. / d
    C = centers_from_rows(mco, data, idx, ptr, n, d, k, 5) + 1e-3
    a_ref, c_ref, gap = mco.assign_f64(code, mco.CSR(data, idx, ptr, (n, d)), C, prior)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    a, c = ds.assign(code, C, prior=prior)
    bad = a != a_ref
    assert np.all(gap[bad] < 1e-4), gap[bad][:10]
    np.testing.assert_allclose(c, c_ref, rtol=2e-4, atol=1e-6)


@pytest.mark.parametrize("name,code", [("JSD", 4), ("MKL", 5)])
def test_lloyd_csr(mcg, ctx, mco, name, code):
    say(f"csr lloyd {name}")
    n, d, k = 2000, 400, 6
    data, idx, ptr = csr_data(n, d, 0.03, 7, True)
    prior = 1. / d
    C0 = centers_from_rows(mco, data, idx, ptr, n, d, k, 8)
    ref = mco.lloyd(code, mco.CSR(data, idx, ptr, (n, d)), C0, prior=prior, eps=1e-10, maxiter=8)
    got = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).lloyd(code, C0, prior=prior, eps=1e-10, maxiter=8)
    assert got["initcost"] == pytest.approx(ref["initcost"], rel=1e-4)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-4)
    assert np.mean(got["asn"] == ref["asn"]) > 0.995
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=1e-3, atol=1e-6)


def test_csr_centroids(mcg, ctx, mco):
    say("csr centroids")
    n, d, k = 3000, 5000, 11
    data, idx, ptr = csr_data(n, d, 0.01, 9, False)
    asn = np.random.default_rng(1).integers(0, k, n).astype(np.uint32)
    w = np.random.default_rng(2).uniform(0.5, 2, n)
    prev = np.zeros((k, d), np.float32)
    for ww in (None, w):
        ref, _ = mco.centroids(mco.CSR(data, idx, ptr, (n, d)), asn, k, prev, ww)
        got, _ = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).centroids(asn, k, prev, ww)
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-9)
    # reference fixture: test_centroid.cpp stripes
    data2 = 2. * (np.arange(50) + 1.)
    idx2 = (2 * (np.arange(50) % 5)).astype(np.uint32)
    got, _ = mcg.Dataset.csr(data2.astype(np.float32), idx2, (np.arange(11) * 5).astype(np.uint64), (10, 10),
                             ctx).centroids(np.zeros(10, np.uint32), 1, np.zeros((1, 10)))
    assert list(got[0, ::2].astype(int)) == [47, 49, 51, 53, 55]


@pytest.mark.parametrize("weighted", [False, True])
def test_csr_centroids_skewed(mcg, ctx, mco, weighted):
    """One cluster of 6,000 of 8,000 rows: it spans many 256-row groups of its member
    chunks (the chunk loop, the four per-wave accumulators and the rows loaded ahead),
    the others are small; centers within 1e-6 of the oracle's FP64 means."""
    say(f"csr centroids skewed weighted={weighted}")
    n, d, k = 8000, 3000, 7
    data, idx, ptr = csr_data(n, d, 0.02, 19, False)
    rng = np.random.default_rng(3)
    asn = np.where(rng.random(n) < 0.75, 2, rng.integers(0, k, n)).astype(np.uint32)
    w = rng.uniform(0.5, 2, n) if weighted else None
    prev = np.zeros((k, d), np.float32)
    ref, _ = mco.centroids(mco.CSR(data, idx, ptr, (n, d)), asn, k, prev, w)
    got, _ = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).centroids(asn, k, prev, w)
    assert (asn == 2).sum() > 5000
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("chunks", ["1", "3"])
def test_csr_mstep_chunk_counts(mcg, ctx, mco, monkeypatch, chunks):
    """The CSR M-steps with one member / entry chunk (the centers written by the
    chunk kernel itself, no partial sums) and with three: hard centers within 1e-6 of
    the oracle's FP64 means (weighted and not), soft EM centers within the soft tests'
    bound."""
    say(f"csr m-step chunks={chunks}")
    monkeypatch.setenv("MCG_CSR_CHUNKS", chunks)
    n, d, k = 3000, 2000, 9
    data, idx, ptr = csr_data(n, d, 0.02, 31, False)
    X = mco.CSR(data, idx, ptr, (n, d))
    asn = np.random.default_rng(4).integers(0, k, n).astype(np.uint32)
    asn[asn == 3] = 4                                       # cluster 3 empty: left as it was
    w = np.random.default_rng(5).uniform(0.5, 2, n)
    prev = np.full((k, d), 0.5, np.float32)
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    for ww in (None, w):
        ref, _ = mco.centroids(X, asn, k, prev, ww)
        got, _ = ds.centroids(asn, k, prev, ww)
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-9)
        assert np.array_equal(got[3], prev[3])
        # the rows' window entries found by binary search instead of the window
        # starts: the same entries in the same order, so the same centers bit for bit
        monkeypatch.setenv("MCG_CSR_WSTART", "0")
        got_bs, _ = ds.centroids(asn, k, prev, ww)
        monkeypatch.delenv("MCG_CSR_WSTART")
        assert np.array_equal(got_bs, got)
    C0 = X.todense()[np.random.default_rng(6).choice(n, k, replace=False)] + 0.25
    ref = mco.soft(11, X, C0, prior=1e-3, temp=1., eps=-1., maxiter=2)
    got = ds.soft(11, C0, prior=1e-3, temp=1., eps=-1., maxiter=2)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=2e-4)
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=2e-4, atol=1e-5)
    ds.close()


@pytest.mark.parametrize("k", [100, 130, 300])
def test_soft_em_csr_many_centers(mcg, ctx, mco, k):
    """Soft CSR EM with k = 100 (two centers per lane in the tiled E-step), k = 130
    (two-and-a-bit center quads per lane) and k = 300 (two 256-center slices), with a
    few columns stored in every row, so that stages overflow and the tiles are rebuilt
    with fewer operand rows per stage: costs, responsibilities and centers against the
    oracle."""
    say(f"soft csr k={k}")
    n, d = 1200, 400
    data, idx, ptr = csr_data(n, d, 0.03, 23, False)
    A = sp.csr_matrix((data, idx.astype(np.int64), ptr.astype(np.int64)), shape=(n, d)).tolil()
    A[:, [5, 77, 301]] = np.random.default_rng(2).integers(1, 9, (n, 3)).astype(np.float32)
    A = A.tocsr()
    A.sort_indices()
    data, idx, ptr = A.data.astype(np.float32), A.indices.astype(np.uint32), A.indptr.astype(np.uint64)
    X = mco.CSR(data, idx, ptr, (n, d))
    C0 = X.todense()[np.random.default_rng(k).choice(n, k, replace=False)] + 0.25
    ref = mco.soft(11, X, C0, prior=1e-3, temp=1., eps=-1., maxiter=3)
    got = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).soft(11, C0, prior=1e-3, temp=1., eps=-1., maxiter=3)
    assert got["numiter"] == ref["numiter"] == 3
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=2e-4)
    np.testing.assert_allclose(got["asn"], ref["asn"], atol=2e-4)
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=2e-4, atol=1e-5)


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("name,code", [("SQRL2", 2), ("REVERSE_MKL", 11), ("JSD", 4), ("TVD", 9), ("HELLINGER", 6)])
def test_soft_em(mcg, ctx, mco, name, code, sparse):
    say(f"soft {name} sparse={sparse}")
    if sparse:
        n, d, k = 800, 300, 5
        data, idx, ptr = csr_data(n, d, 0.05, 10, False)
        X = mco.CSR(data, idx, ptr, (n, d))
        C0 = X.todense()[:k] + 0.5
        ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
        prior = 1e-3
    else:
        rng = np.random.default_rng(11)
        X = (rng.poisson(3., (700, 20)) + rng.integers(0, 6, (700, 1))).astype(np.float32)
        C0 = X[:4].copy() + 0.5
        ds = mcg.Dataset.dense(X, ctx)
        prior = 0.5
    # a temperature at the cost scale keeps the responsibilities soft (no underflow to argmin-uniform)
    K0 = mco.cost_matrix(code, X, C0, prior)
    temp = float(2. / np.median(K0)) if code in (2, 9, 6) else 1.
    # eps < 0: both run exactly maxiter iterations
    ref = mco.soft(code, X, C0, prior=prior, temp=temp, eps=-1., maxiter=6)
    got = ds.soft(code, C0, prior=prior, temp=temp, eps=-1., maxiter=6)
    assert got["numiter"] == ref["numiter"] == 6
    assert got["initcost"] == pytest.approx(ref["initcost"], rel=2e-4)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=2e-3)
    np.testing.assert_allclose(got["asn"], ref["asn"], atol=2e-3)
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=2e-3, atol=1e-4)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("code", [0, 1])
def test_soft_em_median_measures(mcg, ctx, mco, code, weighted):
    """Soft L1 / L2 (centroid.h:815-838): the weighted median (heaviest row) and the
    Weiszfeld geometric median of all rows under w_i * r_ic, against the oracle."""
    say(f"soft median code={code} weighted={weighted}")
    rng = np.random.default_rng(21)
    X = rng.normal(size=(900, 12)).astype(np.float32)
    X[:300] += 2.5
    X[300:600] -= 2.5
    w = rng.uniform(.5, 2., 900) if weighted else None
    C0 = X[[1, 400, 800]] + np.float32(.2)
    K0 = mco.cost_matrix(code, X, C0, 0.)
    temp = float(3. / np.median(K0))
    ref = mco.soft(code, X, C0, temp=temp, eps=-1., maxiter=4, w=w)
    got = mcg.Dataset.dense(X, ctx).soft(code, C0, temp=temp, eps=-1., maxiter=4, weights=w)
    assert got["numiter"] == ref["numiter"] == 4
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-4)
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(got["asn"], ref["asn"], atol=1e-4)
    if code == 0:   # every center is a data row
        assert all((X == c).all(1).any() for c in got["centers"])


def test_soft_em_l1_ties_and_zero_weights(mcg, ctx, mco):
    """k = 1: every row ties at weight 1, so the L1 center is the columnwise minimum;
    all-zero weights take the lower_bound rule; the L2 form with all-zero weights
    fails as the reference's NaN check does (MCG_ENAN)."""
    say("soft median ties / zero weights")
    rng = np.random.default_rng(4)
    X = rng.normal(size=(500, 7)).astype(np.float32)
    X[9, 2] = 0.
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.soft(0, X[:1] + 1, maxiter=1)
    np.testing.assert_array_equal(got["centers"][0], X.min(0))
    z = np.zeros(500)
    got = ds.soft(0, X[:2], maxiter=3, weights=z)
    ref = mco.soft(0, X, X[:2], maxiter=3, w=z)
    np.testing.assert_array_equal(got["centers"], ref["centers"])
    with pytest.raises(mcg.MCGError) as e:
        ds.soft(1, X[:2] + 1, maxiter=2, weights=z)
    assert e.value.code == -5
    ds.close()


def test_soft_em_median_csr(mcg, ctx, mco):
    """CSR rows: L2 runs the double-precision Weiszfeld that stops on the unweighted
    sum of distances (csc.h:1131-1203); L1 is refused as the reference throws."""
    say("soft median csr")
    n, d, k = 700, 60, 3
    data, idx, ptr = csr_data(n, d, 0.2, 12, False)
    X = mco.CSR(data, idx, ptr, (n, d))
    Xd = X.todense()
    C0 = Xd[[0, 300, 600]] + np.float32(.3)
    w = np.random.default_rng(5).uniform(.5, 2., n)
    K0 = mco.cost_matrix(1, X, C0, 0.)
    temp = float(3. / np.median(K0))
    ds = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx)
    ref = mco.soft(1, X, C0, temp=temp, eps=-1., maxiter=3, w=w)
    got = ds.soft(1, C0, temp=temp, eps=-1., maxiter=3, weights=w)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-4)
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=1e-3, atol=2e-4)
    with pytest.raises(mcg.MCGError) as e:
        ds.soft(0, C0, maxiter=2)
    assert e.value.code == -4
    ds.close()
