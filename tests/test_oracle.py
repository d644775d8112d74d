"""Pin the CPU oracle against the reference's own known answers, library
routines and invariants (no GPU).  The oracle is then the checker for the
HIP library in the gpu tests."""
import json
import os

import numpy as np
import pytest
from scipy.spatial.distance import jensenshannon
from scipy.stats import chi2, entropy

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = json.load(open(os.path.join(GOLD, "reference_cases.json")))


def smoothed(v, prior):
    v = np.asarray(v, dtype=np.float64)
    return (v + prior) / (v.sum() + prior * len(v))


# ---- a1: msr_with_prior ---------------------------------------------------------
def test_msr_known_cv1_cv4(mco):
    r = REF["msr_known"]
    x, c, p = np.array(r["cv1"], float), np.array(r["cv4"], float), r["prior"]
    for name, code in [("MKL", mco.MKL), ("REVERSE_MKL", mco.REVERSE_MKL), ("JSD", mco.JSD), ("SQRL2", mco.SQRL2)]:
        assert mco.msr_dense(code, x, c, p) == pytest.approx(r[name], rel=1e-13), name
        xi, ci = np.nonzero(x)[0], np.nonzero(c)[0]
        v, _ = mco.msr_sparse(code, x[xi], xi, c[ci], ci, len(x), p)
        assert v == pytest.approx(r[name], rel=1e-13), name


def test_msr_orientation_against_scipy(mco):
    """MKL = KL(center~ || row~), REVERSE_MKL = KL(row~ || center~), JSD = JS divergence (nats)."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        d = int(rng.integers(3, 40))
        x = rng.poisson(2., d).astype(float)
        c = rng.poisson(3., d).astype(float)
        p = float(rng.choice([1e-3, 0.1, 1., 5.]))
        xt, ct = smoothed(x, p), smoothed(c, p)
        assert mco.msr_dense(mco.MKL, x, c, p) == pytest.approx(entropy(ct, xt), rel=1e-10)
        assert mco.msr_dense(mco.REVERSE_MKL, x, c, p) == pytest.approx(entropy(xt, ct), rel=1e-10)
        assert mco.msr_dense(mco.JSD, x, c, p) == pytest.approx(jensenshannon(xt, ct) ** 2, rel=1e-9)
        assert mco.msr_dense(mco.SQRL2, x, c, p) == pytest.approx(((x - c) ** 2).sum(), rel=1e-14)


def test_msr_pair26183(mco):
    z = np.load(os.path.join(GOLD, "testmsr_pair.npz"))
    r = REF["msr_pair26183"]
    d = int(z["d"])
    x = np.zeros(d); x[z["lhs_idx"]] = z["lhs_val"]
    c = np.zeros(d); c[z["rhs_idx"]] = z["rhs_val"]
    for name, code in [("MKL", mco.MKL), ("REVERSE_MKL", mco.REVERSE_MKL), ("JSD", mco.JSD), ("SQRL2", mco.SQRL2)]:
        assert mco.msr_dense(code, x, c, r["prior"]) == pytest.approx(r[name], rel=1e-11), name
        v, _ = mco.msr_sparse(code, z["lhs_val"], z["lhs_idx"], z["rhs_val"], z["rhs_idx"], d, r["prior"])
        assert v == pytest.approx(r[name], rel=1e-11), name
    # reference assertion testmsr.cpp:94-97: both directions positive
    assert mco.msr_dense(mco.MKL, x, c, 1e-5) > 0 and mco.msr_dense(mco.REVERSE_MKL, x, c, 1e-5) > 0


@pytest.mark.parametrize("prior", [1., 50., 1e-2])
def test_msr_self_distance_zero(mco, prior):
    # testmsr.cpp:43-56
    x = np.array(REF["msr_known"]["cv1"], float)
    for code in (mco.MKL, mco.REVERSE_MKL, mco.JSD, mco.SQRL2):
        assert abs(mco.msr_dense(code, x, x, prior)) < 1e-12


def test_shared_zero_counts(mco):
    m = REF["merge_shared_zeros"]
    for a, b, want in m["cases"]:
        ai, bi = np.nonzero(a)[0], np.nonzero(b)[0]
        assert mco.shared_zeros(ai, bi, m["d"]) == want
        _, sz = mco.msr_sparse(mco.JSD, np.array(a, float)[ai], ai, np.array(b, float)[bi], bi, m["d"], 1.)
        assert sz == want


def test_sparse_equals_dense_formula(mco):
    rng = np.random.default_rng(7)
    for _ in range(30):
        d = int(rng.integers(5, 300))
        x = rng.poisson(0.3, d).astype(float)
        c = rng.poisson(0.5, d).astype(float)
        x[rng.integers(d)] += 1
        c[rng.integers(d)] += 1
        p = float(rng.choice([1e-5, 1e-2, 1.]))
        xi, ci = np.nonzero(x)[0], np.nonzero(c)[0]
        for code in (mco.MKL, mco.REVERSE_MKL, mco.JSD, mco.SQRL2):
            a = mco.msr_dense(code, x, c, p)
            b, _ = mco.msr_sparse(code, x[xi], xi, c[ci], ci, d, p)
            assert b == pytest.approx(a, rel=1e-9, abs=1e-12)


# ---- further measures (applicator.h:1673-1760, 1778-1891) ---------------------------
def _smoothed_pair(mco, x, c, prior):
    """(smoothed center a, smoothed row b, lhsum, rhsum) as applicator.h:1642-1656."""
    d = len(x)
    lhsum, rhsum = c.sum() + prior * d, x.sum() + prior * d
    pv = max(prior, 7.0069e-42 * (lhsum + rhsum))
    lhsum, rhsum = c.sum() + pv * d, x.sum() + pv * d
    return (c + pv) / lhsum, (x + pv) / rhsum, lhsum, rhsum


def test_more_measures_sparse_equals_dense(mco):
    """The sparse branch (union terms plus the reference's closed-form shared-zero
    terms) equals the dense formula over all coordinates.  Not for REVERSE_SIS, whose
    sparse branch swaps the vectors but not the increments, or the cosine forms at a
    real prior, whose sparse branch ignores the shared zeros."""
    rng = np.random.default_rng(17)
    for _ in range(25):
        d = int(rng.integers(5, 200))
        x = rng.poisson(0.4, d).astype(float)
        c = rng.poisson(0.6, d).astype(float)
        x[rng.integers(d)] += 1
        c[rng.integers(d)] += 1
        xi, ci = np.nonzero(x)[0], np.nonzero(c)[0]
        for code in mco.MORE_MEASURES:
            if code == mco.RSIS:
                continue
            p = 0. if code in (mco.COSINE_DISTANCE, mco.COSINE_SIMILARITY, mco.PROBABILITY_COSINE_DISTANCE,
                               mco.PROBABILITY_COSINE_SIMILARITY) else float(rng.choice([1e-3, 0.1, 1.]))
            a = mco.msr_dense(code, x, c, p)
            b, _ = mco.msr_sparse(code, x[xi], xi, c[ci], ci, d, p)
            assert b == pytest.approx(a, rel=1e-9, abs=1e-10), code


def test_more_measures_self_and_scipy(mco):
    from scipy.spatial.distance import cityblock, cosine, euclidean
    rng = np.random.default_rng(3)
    x = rng.gamma(1., 1., 40)
    c = rng.gamma(1., 1., 40)
    for code in mco.MORE_MEASURES:
        v = mco.msr_dense(code, x, x, 0.1)
        if code in (mco.COSINE_SIMILARITY, mco.PROBABILITY_COSINE_SIMILARITY):
            assert v == pytest.approx(1., abs=1e-12)
        else:
            assert abs(v) < 2e-7, code      # acos near 1 loses half the digits
    p = 0.05
    a, b, lhs, rhs = _smoothed_pair(mco, x, c, p)
    assert mco.msr_dense(mco.L1, x, c, p) == pytest.approx(cityblock(x, c), rel=1e-12)
    assert mco.msr_dense(mco.L2, x, c, p) == pytest.approx(euclidean(x, c), rel=1e-12)
    assert mco.msr_dense(mco.TVD, x, c, p) == pytest.approx(0.5 * cityblock(a, b), rel=1e-12)
    assert mco.msr_dense(mco.JSM, x, c, p) == pytest.approx(jensenshannon(a, b), rel=1e-9)
    assert mco.msr_dense(mco.PROBABILITY_COSINE_SIMILARITY, x, c, p) == pytest.approx(1 - cosine(a, b), rel=1e-12)
    pv = max(p, 7.0069e-42 * (lhs + rhs))
    cs = 1 - cosine(c + pv, x + pv)
    assert mco.msr_dense(mco.COSINE_SIMILARITY, x, c, p) == pytest.approx(cs, rel=1e-12)
    assert mco.msr_dense(mco.COSINE_DISTANCE, x, c, p) == pytest.approx(np.arccos(cs) / np.pi, rel=1e-12)


def test_more_measures_identities(mco):
    """Itakura-Saito is the Bregman divergence of -log t (center first); SIS is the
    mean IS to the midpoint; UWLLR with equal sums is JSD; LLR scales UWLLR by the
    total mass; Hellinger equals the Bhattacharyya metric on distributions."""
    rng = np.random.default_rng(5)
    x = rng.gamma(2., 1., 30)
    c = rng.gamma(2., 1., 30)
    p = 0.2
    a, b, lhs, rhs = _smoothed_pair(mco, x, c, p)
    breg = lambda u, v: float(np.sum(-np.log(u) + np.log(v) + (u - v) / v))   # phi(u) - phi(v) - phi'(v)(u - v)
    assert mco.msr_dense(mco.IS, x, c, p) == pytest.approx(breg(a, b), rel=1e-10)
    assert mco.msr_dense(mco.REVERSE_IS, x, c, p) == pytest.approx(breg(b, a), rel=1e-10)
    m = 0.5 * (a + b)
    assert mco.msr_dense(mco.SIS, x, c, p) == pytest.approx(0.5 * (breg(a, m) + breg(b, m)), rel=1e-10)
    u = mco.msr_dense(mco.UWLLR, x, c, p)
    assert mco.msr_dense(mco.LLR, x, c, p) == pytest.approx(u * (lhs + rhs), rel=1e-12)
    assert mco.msr_dense(mco.SRULRT, x, c, p) == pytest.approx(np.sqrt(u), rel=1e-12)
    xe = x * (c.sum() / x.sum())                        # equal sums: lambda = 1/2
    assert mco.msr_dense(mco.UWLLR, xe, c, p) == pytest.approx(mco.msr_dense(mco.JSD, xe, c, p), rel=1e-10)
    h = mco.msr_dense(mco.HELLINGER, x, c, p)
    assert h == pytest.approx(mco.msr_dense(mco.BHATTACHARYYA_METRIC, x, c, p), rel=1e-9)
    bc = float(np.sum(np.sqrt(a * b)))
    assert mco.msr_dense(mco.BHATTACHARYYA_DISTANCE, x, c, p) == pytest.approx(-np.log(bc), rel=1e-10)


def test_poisson_divergence(mco):
    """The I-divergence (poisson_bregman, distance.h:576-585): the sum of scipy's
    elementwise kl_div on the prior-shifted vectors; sparse == dense; 0 on itself; and,
    a Bregman divergence, the mean is the best center (cost rises along any
    perturbation of it)."""
    from scipy.special import kl_div
    rng = np.random.default_rng(7)
    d = 40
    x = rng.poisson(2., d).astype(float)
    c = rng.poisson(2., d).astype(float)
    for prior in (0.5, 1e-3):
        ref = float(kl_div(x + prior, c + prior).sum())
        assert mco.msr_dense(mco.POISSON, x, c, prior) == pytest.approx(ref, rel=1e-12)
        xi, ci = np.nonzero(x)[0], np.nonzero(c)[0]
        v, _ = mco.msr_sparse(mco.POISSON, x[xi], xi, c[ci], ci, d, prior)
        assert v == pytest.approx(ref, rel=1e-12)
        assert mco.msr_dense(mco.POISSON, x, x, prior) == 0.
    # zero prior: a coordinate with x > 0 = c is infinite (clamped to FLT_MAX)
    assert mco.msr_dense(mco.POISSON, np.array([1., 0.]), np.array([0., 1.]), 0.) == pytest.approx(3.4028234663852886e38)
    X = rng.gamma(2., 1., (200, 6))
    m = X.mean(0)
    f = lambda cc: sum(mco.msr_dense(mco.POISSON, r, cc, 0.1) for r in X)
    base = f(m)
    for j in range(6):
        for eps in (1e-3, -1e-3):
            e = np.zeros(6); e[j] = eps
            assert f(m + e) > base


def test_sqrl2_f32_is_sequential_fma(mco):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(77).astype(np.float32)
    c = rng.standard_normal(77).astype(np.float32)
    s = np.float32(0)
    for a, b in zip(x, c):
        t = np.float32(a - b)
        s = np.float32(np.float64(t) * np.float64(t) + np.float64(s))  # exact product, one rounding
    assert mco.sqrl2_f32(x, c) == s


# ---- a2: assignment ---------------------------------------------------------------
def test_assign_bruteforce_and_ties(mco):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((300, 9)).astype(np.float32)
    C = rng.standard_normal((12, 9)).astype(np.float32)
    C[7] = C[2]  # duplicate center: ties must go to the lower index
    asn, cost, gap = mco.assign_sqrl2_f32(X, C)
    D = ((X[:, None, :].astype(np.float64) - C[None].astype(np.float64)) ** 2).sum(-1)
    assert np.array_equal(asn, D.argmin(1))
    assert not np.any(asn == 7)
    np.testing.assert_allclose(cost, D.min(1), rtol=1e-5)
    a2, c2, _ = mco.assign_f64(mco.SQRL2, X, C, 0.)
    assert np.array_equal(a2, asn)


def test_assign_csr_matches_dense(mco):
    rng = np.random.default_rng(4)
    Xd = (rng.poisson(0.2, (50, 40))).astype(np.float32)
    Xd[:, 0] += 1
    C = rng.poisson(1.0, (6, 40)).astype(np.float32) + 0.1
    from scipy.sparse import csr_matrix
    Xs = mco.CSR.from_scipy(csr_matrix(Xd))
    for code in (mco.MKL, mco.REVERSE_MKL, mco.JSD, mco.SQRL2):
        a1, c1, _ = mco.assign_f64(code, Xd, C, 0.05)
        a2, c2, _ = mco.assign_f64(code, Xs, C, 0.05)
        assert np.array_equal(a1, a2)
        np.testing.assert_allclose(c1, c2, rtol=1e-12)


# ---- a4: centroids ------------------------------------------------------------------
def test_centroid_reference_means(mco):
    r = REF["centroid_means"]
    dm = np.eye(10, dtype=np.float32) * 10
    Cm, e = mco.centroids(dm, np.zeros(10, np.uint32), 1, np.zeros((1, 10)))
    assert e == 0 and np.all(Cm == r["diag10_mean"])
    csr = mco.CSR(np.ones(10), np.arange(10), np.arange(11), (10, 10))
    Cm, _ = mco.centroids(csr, np.zeros(10, np.uint32), 1, np.zeros((1, 10)))
    np.testing.assert_allclose(Cm, r["csr_identity_mean"], rtol=1e-7)
    data = 2. * (np.arange(50) + 1.)
    idx = 2 * (np.arange(50) % 5)
    csr = mco.CSR(data, idx, np.arange(11) * 5, (10, 10))
    Cm, _ = mco.centroids(csr, np.zeros(10, np.uint32), 1, np.zeros((1, 10)))
    assert list(Cm[0, ::2].astype(int)) == r["csr_stripes_mean_at_even_cols"]
    dm2 = np.zeros((500, 5), np.float32); dm2[:, 0] = 1; dm2[:, 3] = -1.4
    Cm, _ = mco.centroids(dm2, np.zeros(500, np.uint32), 1, np.zeros((1, 5)))
    np.testing.assert_allclose(Cm[0], r["dense_500x5_mean"], rtol=1e-7)


def test_centroids_weighted_and_empty(mco):
    rng = np.random.default_rng(9)
    X = rng.standard_normal((1000, 7)).astype(np.float32)
    asn = rng.integers(0, 5, 1000).astype(np.uint32)
    asn[asn == 3] = 4  # cluster 3 empty
    w = rng.uniform(0.5, 2., 1000)
    prev = np.full((5, 7), 42., np.float32)
    Cm, e = mco.centroids(X, asn, 5, prev, w)
    assert e == 1 and np.all(Cm[3] == 42.)
    for c in (0, 1, 2, 4):
        m = asn == c
        np.testing.assert_allclose(Cm[c], (w[m, None] * X[m]).sum(0) / w[m].sum(), rtol=1e-6, atol=1e-7)


# ---- a3: Lloyd control flow ---------------------------------------------------------
def test_lloyd_rules(mco):
    from minicore_amd.datagen import gaussian_mixture
    X, lab, means = gaussian_mixture(3000, 8, 6, seed=11, lo=-3, hi=3)
    C0 = X[[0, 1, 2, 3, 4, 5]].copy()
    r = mco.lloyd(mco.SQRL2, X, C0, eps=0., maxiter=4)
    # eps = 0: stops after maxiter + 1 accepted iterations unless the cost rises (solve.h:160-170)
    assert r["numiter"] in range(1, 6)
    assert r["finalcost"] <= r["initcost"]
    # the returned assignment is the argmin for the returned centers
    a, c, _ = mco.assign_sqrl2_f32(X, r["centers"])
    assert np.array_equal(a, r["asn"])
    # brute-force FP64 Lloyd reaches the same fixed point from the same start
    Cb = C0.astype(np.float64)
    for _ in range(200):
        D = ((X[:, None, :] - Cb[None]) ** 2).sum(-1)
        ab = D.argmin(1)
        Cn = np.stack([X[ab == j].mean(0) if np.any(ab == j) else Cb[j] for j in range(6)])
        if np.allclose(Cn, Cb, rtol=0, atol=1e-12):
            break
        Cb = Cn
    r2 = mco.lloyd(mco.SQRL2, X, C0, eps=1e-10, maxiter=1000)
    assert np.array_equal(r2["asn"], ab)
    np.testing.assert_allclose(r2["centers"], Cb, rtol=1e-5, atol=1e-5)


# ---- minibatch driver (solve.h:377-578) ------------------------------------------------
def _fixed_total(t):
    t = np.asarray(t, dtype=np.float64)
    mx = float(np.abs(t).max())
    m, e = np.frexp(mx)
    s = 62 - int(e) - int(len(t)).bit_length()
    return float(np.ldexp(float(sum(int(v) for v in np.floor(np.ldexp(t, s)))), -s))


@pytest.mark.parametrize("weighted", [False, True])
def test_minibatch_full_batch_is_lloyd(mco, weighted):
    """mbsize = n without replacement draws every row, so each update is a full
    Lloyd M-step; with a check-in every iteration and no restarts the best check-in
    is the last one: the centers and assignment of Lloyd after the same number of
    updates (mco_lloyd with maxiter T-2 runs T-1 updates)."""
    from minicore_amd.datagen import gaussian_mixture
    X, lab, _ = gaussian_mixture(3000, 8, 5, seed=3, lo=-4, hi=4)
    w = np.random.default_rng(1).uniform(0.5, 2., 3000) if weighted else None
    C0 = X[[int(np.argmax(lab == c)) for c in range(5)]].copy()
    T = 6
    r = mco.minibatch(mco.SQRL2, X, C0, w=w, mbsize=3000, maxiter=T, freq=1, reseed_after=0, with_rep=False, seed=4)
    L = mco.lloyd(mco.SQRL2, X, C0, w=w, eps=-1., maxiter=T - 2)
    assert L["numiter"] == T - 1 and r["numiter"] == T
    assert np.array_equal(r["centers"], L["centers"])
    assert np.array_equal(r["asn"], L["asn"])
    assert r["finalcost"] == pytest.approx(L["finalcost"], rel=1e-6)
    a0, c0, _ = mco.assign_sqrl2_f32(X, C0)
    assert r["initcost"] == _fixed_total(c0.astype(np.float64) * (w if weighted else 1.))


@pytest.mark.parametrize("msr,weighted", [(0, False), (1, False), (1, True)])
def test_minibatch_full_batch_is_median_lloyd(mco, msr, weighted):
    """The median updates of the minibatch driver (solve.h:555-572): with every row
    drawn, no restarts and a check-in every iteration they are the median M-steps of
    Lloyd (L1 unweighted never raises the cost; the geometric median keeps the
    members' weights in both drivers)."""
    from minicore_amd.datagen import gaussian_mixture
    X, lab, _ = gaussian_mixture(2400, 6, 5, seed=5, lo=-4, hi=4)
    w = np.random.default_rng(2).uniform(0.5, 2., 2400) if weighted else None
    C0 = (X[[int(np.argmax(lab == c)) for c in range(5)]] + 0.25).astype(np.float32)
    T = 5
    r = mco.minibatch(msr, X, C0, w=w, mbsize=2400, maxiter=T, freq=1, reseed_after=0, with_rep=False, seed=4)
    L = mco.lloyd_median(msr, X, C0, w=w, eps=-1., maxiter=T - 2)
    assert L["numiter"] == T - 1 and r["numiter"] == T
    assert np.array_equal(r["centers"], L["centers"])
    assert np.array_equal(r["asn"], L["asn"])


@pytest.mark.parametrize("weighted", [False, True])
def test_minibatch_restart_draw(mco, weighted):
    """A center with no rows is restarted at the row drawn over w_i * cost_i with
    the stream's first output; rows move to it when strictly cheaper; with one
    iteration the saved centers are still the input ones (solve.h:494-499)."""
    from minicore_amd.datagen import gaussian_mixture
    X, lab, _ = gaussian_mixture(2000, 6, 4, seed=5, lo=-4, hi=4)
    w = np.random.default_rng(2).uniform(0.5, 2., 2000) if weighted else None
    C0 = np.concatenate([X[[int(np.argmax(lab == c)) for c in range(3)]], np.full((1, 6), 1e3, np.float32)])
    r = mco.minibatch(mco.SQRL2, X, C0, w=w, mbsize=10, maxiter=1, freq=1, reseed_after=0, seed=9)
    _, c0, _ = mco.assign_sqrl2_f32(X, C0)
    wv = c0.astype(np.float64) * (w if weighted else 1.)
    row = mco.fixed_draw(mco.fixed_weights(wv), mco.wyrand_at(9, 0))
    C1 = C0.copy()
    C1[3] = X[row]
    a1, c1, _ = mco.assign_sqrl2_f32(X, C1)
    assert np.array_equal(r["asn"], a1) and r["asn"][row] == 3
    assert np.array_equal(r["costs"], c1.astype(np.float64))
    assert np.array_equal(r["centers"], C0)
    assert r["initcost"] == r["finalcost"] == _fixed_total(c1.astype(np.float64) * (w if weighted else 1.))


@pytest.mark.parametrize("with_rep", [True, False])
def test_minibatch_one_update(mco, with_rep):
    """One minibatch update: mbsize draws rng() mod n (without replacement the first
    distinct ones), sorted, assigned, and the clusters with members take their mean
    (ascending row order, duplicates counted); the returned centers are those of the
    lower of the two check-in costs."""
    from minicore_amd.datagen import gaussian_mixture
    n, seed, m = 1500, 21, 300
    X, lab, _ = gaussian_mixture(n, 5, 4, seed=7, lo=-4, hi=4)
    C0 = X[[int(np.argmax(lab == c)) for c in range(4)]].copy()
    r = mco.minibatch(mco.SQRL2, X, C0, mbsize=m, maxiter=2, freq=100, reseed_after=0, with_rep=with_rep, seed=seed)
    draws, j = [], 0
    while len(draws) < m:
        v = mco.wyrand_at(seed, j) % n
        j += 1
        if with_rep or v not in draws:
            draws.append(v)
    sel = np.sort(np.array(draws, dtype=np.int64))
    Xs = X[sel]
    a, _, _ = mco.assign_sqrl2_f32(Xs, C0)
    C1, _ = mco.centroids(Xs, a, 4, C0)
    _, c0, _ = mco.assign_sqrl2_f32(X, C0)
    a1, c1, _ = mco.assign_sqrl2_f32(X, C1)
    t0, t1 = _fixed_total(c0.astype(np.float64)), _fixed_total(c1.astype(np.float64))
    assert r["initcost"] == t0 and r["finalcost"] == min(t0, t1)
    assert np.array_equal(r["centers"], C1 if t1 < t0 else C0)
    assert np.array_equal(r["asn"], a1)


@pytest.mark.parametrize("weighted", [False, True])
def test_cs_minibatch_one_round(mco, weighted):
    """One outer round of the coreset-minibatch driver is: check-in; LBK
    probabilities from its costs and assignment; a coreset of m draws seeded by the
    stream's first output, sorted, weighted w_i * w_i/(m p_i); Lloyd on it from the
    current centers (eps 1e-3, maxiter 3 in the first rounds); check-in; the
    cheaper check-in's centers are returned."""
    from minicore_amd.datagen import gaussian_mixture
    n, m, seed = 4000, 500, 13
    X, lab, _ = gaussian_mixture(n, 6, 5, seed=2, lo=-4, hi=4)
    w = np.random.default_rng(4).uniform(0.5, 2., n) if weighted else None
    C0 = X[[int(np.argmax(lab == c)) for c in range(5)]].copy()
    r = mco.cs_minibatch(mco.SQRL2, X, C0, w=w, mbsize=m, maxiter=1, freq=1, reseed_after=0, seed=seed)
    assert r["numiter"] == 2
    a0, c0, _ = mco.assign_sqrl2_f32(X, C0)
    probs = mco.coreset_probs(4, 5, c0.astype(np.float64), a0, w)
    wt, ind = mco.coreset_sample(probs, m, mco.wyrand_at(seed, 0), w)
    order = np.argsort(ind, kind="stable")
    ind, wt = ind[order], wt[order]
    csw = (w[ind] if weighted else 1.) * wt.astype(np.float64)
    L = mco.lloyd(mco.SQRL2, X[ind.astype(np.int64)], C0, w=csw, eps=1e-3, maxiter=3)
    a1, c1, _ = mco.assign_sqrl2_f32(X, L["centers"])
    mult = w if weighted else 1.
    t0, t1 = _fixed_total(c0.astype(np.float64) * mult), _fixed_total(c1.astype(np.float64) * mult)
    assert r["initcost"] == t0 and r["finalcost"] == min(t0, t1)
    assert np.array_equal(r["centers"], L["centers"] if t1 < t0 else C0)
    assert np.array_equal(r["asn"], a1)


# ---- a5: soft -------------------------------------------------------------------------
def test_softmax_rows_and_repair(mco):
    rng = np.random.default_rng(2)
    K = rng.uniform(0, 5, (40, 7)).astype(np.float32)
    R = mco.softmax_rows(K, 1.)
    np.testing.assert_allclose(R.sum(1), 1., rtol=1e-6)
    e = np.exp(-K.astype(np.float64))
    np.testing.assert_allclose(R, e / e.sum(1, keepdims=True), rtol=1e-5)
    K2 = np.full((2, 4), 1000., np.float32); K2[0, 1] = 999.; K2[1, [0, 3]] = 998.
    R2 = mco.softmax_rows(K2, 1.)   # every exp underflows: argmin-uniform (centroid.h:781-793)
    assert list(R2[0]) == [0, 1, 0, 0] and list(R2[1]) == [.5, 0, 0, .5]


def test_soft_decreases_and_centers(mco):
    from minicore_amd.datagen import gaussian_mixture
    X, _, means = gaussian_mixture(500, 4, 3, seed=12, lo=-4, hi=4)
    r = mco.soft(mco.SQRL2, X, X[:3], temp=1., maxiter=30)
    assert r["finalcost"] <= r["initcost"]
    np.testing.assert_allclose(r["asn"].sum(1), 1., rtol=1e-5)


def _first_resp(mco, msr, X, C0, temp):
    """Responsibilities of the first soft E-step (as the oracle forms them)."""
    return mco.softmax_rows(mco.cost_matrix(msr, X, C0, 0.).astype(np.float32), temp)


@pytest.mark.parametrize("weighted", [False, True])
def test_soft_l1_heaviest_row(mco, weighted):
    """Soft L1 centers (weighted_median over all rows, kmedian.h:266-301): each
    center is the row of largest w_i * r_ic; with k = 1 every row ties and the center
    is the columnwise minimum (of the rows of largest weight, when weighted)."""
    rng = np.random.default_rng(3)
    X = rng.normal(size=(300, 6)).astype(np.float32)
    w = rng.uniform(.5, 2., 300) if weighted else None
    C0 = X[[3, 50, 120]]
    r = mco.soft(mco.L1, X, C0, temp=.7, maxiter=1, w=w)
    R = _first_resp(mco, mco.L1, X, C0, .7)
    Wt = R * (w.astype(np.float32)[:, None] if weighted else np.float32(1.))
    for c in range(3):
        assert (Wt[:, c] == Wt[:, c].max()).sum() == 1
        np.testing.assert_array_equal(r["centers"][c], X[Wt[:, c].argmax()])
    w1 = np.ones(300)
    w1[[7, 9, 200]] = 3.
    r1 = mco.soft(mco.L1, X, X[:1], maxiter=1, w=w1 if weighted else None)
    expect = X[[7, 9, 200]].min(0) if weighted else X.min(0)
    np.testing.assert_array_equal(r1["centers"][0], expect)


def test_soft_l1_zero_weight_rule(mco):
    """All-zero weights take weighted_median's lower_bound branch with mid = 0: the
    first value >= 0 in sorted order, v, and the next, u, give v == 0 ? (v+u)/2 : u."""
    rng = np.random.default_rng(5)
    X = rng.normal(size=(40, 5)).astype(np.float32)
    X[3, 1] = 0.
    X[:, 4] = -np.abs(X[:, 4])
    X[11, 4] = 2.5                       # a single value >= 0 in column 4
    r = mco.soft(mco.L1, X, X[:2], maxiter=3, w=np.zeros(40))
    for j in range(5):
        v = np.sort(X[:, j][X[:, j] >= 0])
        u = v[1] if len(v) > 1 else v[0]
        want = np.float32(.5 * (float(v[0]) + float(u))) if v[0] == 0 else u
        assert r["centers"][0, j] == want and r["centers"][1, j] == want


@pytest.mark.parametrize("weighted", [False, True])
def test_soft_l2_weiszfeld_optimal(mco, weighted):
    """Soft L2 centers are the geometric medians of all rows under w_i * r_ic
    (blaze_adaptor.h:615-656): each center's weighted cost is within 1e-6 of scipy's
    Nelder-Mead optimum of the same objective."""
    from scipy.optimize import minimize
    rng = np.random.default_rng(9)
    X = rng.normal(size=(250, 3)).astype(np.float32)
    X[:120] += 3.
    w = rng.uniform(.5, 2., 250) if weighted else None
    C0 = X[[0, 200]] + np.float32(.1)
    r = mco.soft(mco.L2, X, C0, temp=2., maxiter=1, w=w)
    R = _first_resp(mco, mco.L2, X, C0, 2.)
    Xd = X.astype(np.float64)
    for c in range(2):
        wc = (R[:, c] * (w.astype(np.float32) if weighted else np.float32(1.))).astype(np.float64)
        f = lambda y: float((wc * np.linalg.norm(Xd - y, axis=1)).sum())
        best = minimize(f, Xd.mean(0), method="Nelder-Mead", options=dict(xatol=1e-10, fatol=1e-12, maxiter=20000))
        assert f(r["centers"][c].astype(np.float64)) <= best.fun * (1 + 1e-6)
    # started on a data row, the floored distance (1e-30) gives that row all the
    # weight: the center stays on it
    r0 = mco.soft(mco.L2, X, X[[0, 200]], temp=2., maxiter=1, w=w)
    np.testing.assert_array_equal(r0["centers"], X[[0, 200]])


def test_soft_l2_csr_unweighted_stop(mco):
    """CSR rows (csc.h:1131-1203): the update uses the weights but stops when the
    UNWEIGHTED sum of distances no longer falls.  With k = 1 (all weights 1) the
    center is the plain geometric median; with skewed weights the iterate is the
    plain-loop Weiszfeld in double that stops on the unweighted sum, which here stops
    earlier than the weighted objective would."""
    rng = np.random.default_rng(2)
    n, d = 120, 4
    Xd = np.abs(rng.normal(size=(n, d))).astype(np.float32)
    Xd[Xd < .4] = 0.
    X = mco.CSR.from_dense(Xd) if hasattr(mco.CSR, "from_dense") else None
    if X is None:
        ptr = np.zeros(n + 1, np.uint64)
        idx, val = [], []
        for i in range(n):
            nz = np.nonzero(Xd[i])[0]
            idx += nz.tolist(); val += Xd[i, nz].tolist()
            ptr[i + 1] = len(idx)
        X = mco.CSR(np.array(val, np.float32), np.array(idx, np.uint32), ptr, (n, d))
    c0 = Xd[:1] + np.float32(.37)
    rd = mco.soft(mco.L2, Xd, c0, maxiter=1)
    rc = mco.soft(mco.L2, X, c0, maxiter=1)
    # the float stop of the dense form ends earlier than the double stop of the CSR form
    np.testing.assert_allclose(rc["centers"], rd["centers"], atol=3e-4)
    f = lambda y: float(np.linalg.norm(Xd.astype(np.float64) - y.astype(np.float64), axis=1).sum())
    assert f(rc["centers"][0]) <= f(rd["centers"][0]) * (1 + 1e-9)
    # skewed weights: a plain Weiszfeld loop in double with the unweighted stop
    w = np.where(np.arange(n) < 10, 50., .1)
    r = mco.soft(mco.L2, X, c0, maxiter=1, w=w)
    wc = w.astype(np.float32).astype(np.float64)
    Xx = Xd.astype(np.float64)

    def loop(stop_weighted):
        c = c0[0].copy()
        prev = np.inf
        for _ in range(100000):
            dist = np.maximum(np.sqrt(((Xx - c.astype(np.float64)) ** 2).sum(1)), 1e-80)
            cur = float((wc * dist).sum() if stop_weighted else dist.sum())
            if cur >= prev:
                break
            prev = cur
            q = wc / dist
            c = ((q[:, None] * Xx).sum(0) / q.sum()).astype(np.float32)
        return c
    np.testing.assert_allclose(r["centers"][0], loop(False), rtol=1e-5, atol=1e-6)
    assert np.abs(loop(True) - loop(False)).max() > 1e-3
    with pytest.raises(NotImplementedError):
        mco.soft(mco.L1, X, Xd[:2], maxiter=1)


# ---- a6: D^2 seeding --------------------------------------------------------------------
def test_kmeanspp_first_center_and_assignments(mco):
    rng = np.random.default_rng(8)
    X = rng.standard_normal((400, 5)).astype(np.float32)
    ctr, asn, dist = mco.kmeanspp(mco.SQRL2, X, 10, seed=77)
    assert ctr[0] == mco.wyrand_at(77, 0) % 400
    assert len(set(ctr.tolist())) == 10
    a, c, _ = mco.assign_sqrl2_f32(X, X[ctr])
    np.testing.assert_array_equal(dist, np.where(np.isin(np.arange(400), ctr), 0., c))
    ok = dist > 0
    assert np.array_equal(asn[ok], a[ok])


def test_kmeanspp_second_center_distribution(mco):
    """The second center is drawn with probability proportional to D^2 (chi-square)."""
    X = np.array([[0.], [1.], [3.], [6.]], np.float32)
    counts = np.zeros(4)
    expected = np.zeros(4)
    trials = 4000
    for s in range(1, trials + 1):
        ctr, _, _ = mco.kmeanspp(mco.SQRL2, X, 2, seed=s)
        counts[ctr[1]] += 1
        d2 = ((X[:, 0] - X[ctr[0], 0]) ** 2)
        expected += d2 / d2.sum()
    m = expected > 0
    stat = ((counts[m] - expected[m]) ** 2 / expected[m]).sum()
    assert stat < chi2.ppf(0.999, m.sum() - 1)


def _local_trials_loops(mco, X, k, L, seed, w):
    """kmeans.h:79-177 with n_local_samples = L, as plain loops over a tiny input: the
    cheapest of L uniform first centers, odists / oasn taken once after it, then per
    round L draws (outputs 0..L-1 of the stream seeded by the round's raw draw) of
    which candidate 0 updates the distances in place and the others start from the
    stale copies; the cheapest total wins, the first on ties.  Totals are correctly
    rounded sums (math.fsum)."""
    import math
    n = len(X)
    cost = lambda i, c: float(mco.sqrl2_f32(X[i], X[c]))
    tot = lambda dv: math.fsum(dv[i] * (w[i] if w is not None else 1.) for i in range(n))
    j = 0
    best = None
    for t in range(L):
        fc = mco.wyrand_at(seed, j) % n
        j += 1
        dv = [0. if i == fc else cost(i, fc) for i in range(n)]
        if best is None or tot(dv) < best[0]:
            best = (tot(dv), fc, dv)
    centers, dist, asn = [best[1]], list(best[2]), [0] * n
    od, oa = list(dist), list(asn)
    for ci in range(1, k):
        rngv = mco.wyrand_at(seed, j)
        j += 1
        fw = mco.fixed_weights(np.array([dist[i] * (w[i] if w is not None else 1.) for i in range(n)]))
        smp = [mco.fixed_draw(fw, mco.wyrand_at(rngv, t)) for t in range(L)]
        dsum, newc = None, smp[0]
        for t in range(L):
            dv, av = (dist, asn) if t == 0 else (list(od), list(oa))
            for i in range(n):
                v = cost(i, smp[t])
                if v < dv[i]:
                    dv[i], av[i] = v, ci
            s = tot(dv)
            if dsum is None:
                dsum, newc = s, smp[t]
            elif s < dsum:
                dist, asn, dsum, newc = dv, av, s, smp[t]
        centers.append(newc)
        asn[newc] = ci
        dist[newc] = 0.
    return np.array(centers), np.array(asn), np.array(dist, np.float32)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("L", [2, 3, 5])
def test_kmeanspp_local_trials_loops(mco, weighted, L):
    """n_local_trials > 1 against the plain-loop statement on tiny inputs; several
    seeds so that later candidates win some rounds (the stale-copy rule matters)."""
    rng = np.random.default_rng(L)
    for seed in range(1, 7):
        X = rng.standard_normal((36, 3)).astype(np.float32)
        w = rng.uniform(0.2, 3., 36) if weighted else None
        c, a, d = mco.kmeanspp_local(mco.SQRL2, X, 6, L, seed=seed, w=w)
        c2, a2, d2 = _local_trials_loops(mco, X, 6, L, seed, w)
        assert np.array_equal(c, c2) and np.array_equal(a, a2) and np.array_equal(d, d2)


def test_kmeanspp_local_trials_invariants(mco):
    """The first center is the cheapest of the L uniform rows; every distance is the
    cost to the assigned center; L candidates never lose to one (the cost after the
    first center is at most any candidate's); local search++ then leaves row minima."""
    rng = np.random.default_rng(4)
    X = rng.standard_normal((300, 4)).astype(np.float32)
    n, L = 300, 4
    c, a, d = mco.kmeanspp_local(mco.SQRL2, X, 8, L, seed=21)
    cands = [mco.wyrand_at(21, t) % n for t in range(L)]
    totals = [float(mco.assign_sqrl2_f32(X, X[[r]])[1].astype(np.float64).sum()) for r in cands]
    assert c[0] == cands[int(np.argmin(totals))]
    own = np.array([mco.sqrl2_f32(X[i], X[c[a[i]]]) for i in range(n)], np.float32)
    assert np.array_equal(d, own)
    c1, a1, d1 = mco.kmeanspp_local(mco.SQRL2, X, 1, L, seed=21)
    assert c1[0] == c[0] and not a1.any()
    cl, al, dl = mco.kmeanspp_local(mco.SQRL2, X, 8, L, seed=21, lspp=15)
    c0, a0, d0 = mco.kmeanspp_local(mco.SQRL2, X, 8, L, seed=21, lspp=0)
    assert np.array_equal(c0, c) and np.array_equal(d0, d)
    ab, cb, _ = mco.assign_sqrl2_f32(X, X[cl.astype(np.int64)])
    assert np.array_equal(al, ab) and np.array_equal(dl, cb)


# ---- fixed-point sampler ------------------------------------------------------------------
@pytest.mark.parametrize("weighted", [False, True])
def test_localsearch_rounds_invariants(mco, weighted):
    """After local search++ the distances and assignment are the row minima over the
    returned centers (first index on ties), the centers are distinct rows, and on a
    mixture the rounds lower the seeding cost."""
    from minicore_amd.datagen import gaussian_mixture
    X, _, _ = gaussian_mixture(2500, 8, 10, seed=6)
    w = np.random.default_rng(1).uniform(0.5, 2., 2500) if weighted else None
    c0, _, d0 = mco.kmeanspp(mco.SQRL2, X, 10, seed=3, w=w)
    c, a, d = mco.kmeanspp_ls(mco.SQRL2, X, 10, 25, seed=3, w=w)
    c_same, a_same, d_same = mco.kmeanspp_ls(mco.SQRL2, X, 10, 0, seed=3, w=w)
    assert np.array_equal(c_same, c0) and np.array_equal(d_same, d0)
    assert len(set(c.tolist())) == 10
    ab, cb, _ = mco.assign_sqrl2_f32(X, X[c.astype(np.int64)])
    assert np.array_equal(a, ab) and np.array_equal(d, cb)
    if not weighted:   # the weighted draw reuses the last round's new costs (lsearchpp.h:53-56)
        assert float(d.astype(np.float64).sum()) < float(d0.astype(np.float64).sum())


def test_fixed_draw_semantics(mco):
    w = np.array([0., 3., 0., 1., 0.])
    fw = mco.fixed_weights(w)
    assert fw[0] == fw[2] == fw[4] == 0 and fw[1] == 3 * fw[3]
    assert mco.fixed_draw(fw, 0) == 1                      # smallest u -> first positive weight
    assert mco.fixed_draw(fw, (1 << 64) - 1) == 3          # largest u -> last positive weight
    draws = [mco.fixed_draw(fw, mco.wyrand_at(5, j)) for j in range(4000)]
    frac = np.mean(np.array(draws) == 1)
    assert abs(frac - 0.75) < 0.03


# ---- a7: coresets -----------------------------------------------------------------------
@pytest.mark.parametrize("weighted", [False, True])
def test_coreset_probs_lfkf(mco, weighted):
    """LFKF (make_gmm_sampler, coreset.h:341-381, alpha_est = 0): p_i = 2T/W_a
    normalised, i.e. (1/W_a) / sum_j 1/W_{a_j}; unweighted 1 / (n_a * nonempty)."""
    rng = np.random.default_rng(12)
    n, k = 3000, 9
    asn = rng.integers(0, k, n).astype(np.uint32)
    asn[asn == 4] = 5                       # an empty cluster
    costs = rng.exponential(1., n)
    w = rng.uniform(.5, 2., n) if weighted else None
    p = mco.coreset_probs(2, k, costs, asn, w)
    wf = w.astype(np.float32).astype(np.float64) if weighted else np.ones(n)
    Wa = np.bincount(asn, weights=wf, minlength=k)
    want = (1. / Wa[asn]) / (1. / Wa[asn]).sum()
    np.testing.assert_allclose(p, want, rtol=1e-6)
    if not weighted:
        cnt = np.bincount(asn, minlength=k)
        np.testing.assert_allclose(p, 1. / (cnt[asn] * (cnt > 0).sum()), rtol=1e-6)
    assert abs(float(p.astype(np.float64).sum()) - 1.) < 1e-5


@pytest.mark.parametrize("sens", [0, 3, 4])
def test_coreset_probs_invariants(mco, sens):
    rng = np.random.default_rng(sens)
    n, k = 2000, 9
    asn = rng.integers(0, k, n).astype(np.uint32)
    costs = rng.exponential(1., n)
    w = rng.uniform(0.5, 3, n)
    for ww in (None, w):
        p = mco.coreset_probs(sens, k, costs, asn, ww)
        assert p.dtype == np.float32 and np.all(p > 0)
        assert abs(p.astype(np.float64).sum() - 1.) < 1e-5
        # within one cluster (and equal weights) sensitivity increases with cost
        m = asn == 0
        if ww is None:
            o = np.argsort(costs[m])
            assert np.all(np.diff(p[m][o]) >= 0)
    # equal costs, equal clusters, unit weights: uniform
    asn2 = (np.arange(900) % 9).astype(np.uint32)
    p2 = mco.coreset_probs(sens, 9, np.ones(900), asn2)
    np.testing.assert_allclose(p2, 1. / 900, rtol=1e-6)


def test_coreset_probs_lbk_hand_cases(mco):
    """LBK (coreset.h:487-533) on 3 points in 2 clusters, worked out by hand from the
    reference's steps with a = 16 ln 2 + 32:
    unweighted, costs (0, 2, 4), clusters (0, 0, 1): mean cost 2, cluster terms
    a + 6 and 4a + 12, so sens = (a + 6, 2a + 6, 6a + 12) over 9a + 24;
    weighted w = (1, 3, 2), costs (1, 0, 2): W = 6, mean cost 5/6, cluster terms
    0.6a + 6 and 4.8a + 12, sens = (1.8a + 6, 0.6a + 6, 7.2a + 12) (the point term
    uses the raw cost).  A wrong a, or a dropped 4W/W_c or 2a S_c/(W_c mean) term,
    changes these ratios."""
    a = 16. * np.log(2.) + 32.
    asn = np.array([0, 0, 1], np.uint32)
    p = mco.coreset_probs(4, 2, [0., 2., 4.], asn)
    s = np.array([a + 6, 2 * a + 6, 6 * a + 12])
    np.testing.assert_allclose(p, s / s.sum(), rtol=2e-7)
    p = mco.coreset_probs(4, 2, [1., 0., 2.], asn, w=[1., 3., 2.])
    s = np.array([1.8 * a + 6, .6 * a + 6, 7.2 * a + 12])
    np.testing.assert_allclose(p, s / s.sum(), rtol=2e-7)
    # alpha comes from the number of centers: the same points labelled with k = 3
    p3 = mco.coreset_probs(4, 3, [0., 2., 4.], asn)
    a3 = 16. * np.log(3.) + 32.
    s3 = np.array([a3 + 6, 2 * a3 + 6, 6 * a3 + 12 + 0.])
    # cluster 2 is empty and assigned to nobody, so only a changes
    np.testing.assert_allclose(p3, s3 / s3.sum(), rtol=2e-7)


def test_coreset_probs_vx_bfl_hand_cases(mco):
    """VX (coreset.h:423-455): s_i = w_i c_i / sum(w c) + 1/count_a.  BFL
    (coreset.h:534-578): p_i = (w_i c_i / sum(w c) + w_i / (W_a count_a)) / 2,
    both normalised.  Worked by hand:
    unweighted costs (1, 1, 2), clusters (0, 0, 1): VX s = (.75, .75, 1.5) -> (.25, .25, .5);
    BFL p = (.25, .25, .75) -> (.2, .2, .6);
    weighted w = (2, 1, 1), costs (1, 2, 1), clusters (0, 1, 1): VX s = (1.4, .9, .7) / 3
    (counts, not weights, in the second term); BFL p = (.7, .325, .225) / 1.25."""
    asn = np.array([0, 0, 1], np.uint32)
    np.testing.assert_allclose(mco.coreset_probs(3, 2, [1., 1., 2.], asn), [.25, .25, .5], rtol=2e-7)
    np.testing.assert_allclose(mco.coreset_probs(0, 2, [1., 1., 2.], asn), [.2, .2, .6], rtol=2e-7)
    asn = np.array([0, 1, 1], np.uint32)
    w = [2., 1., 1.]
    np.testing.assert_allclose(mco.coreset_probs(3, 2, [1., 2., 1.], asn, w=w), np.array([1.4, .9, .7]) / 3.,
                               rtol=2e-7)
    np.testing.assert_allclose(mco.coreset_probs(0, 2, [1., 2., 1.], asn, w=w),
                               np.array([.7, .325, .225]) / 1.25, rtol=2e-7)


def test_coreset_probs_fl_hand_cases(mco):
    """FL (coreset.h:456-486), not normalised: unweighted ceil(n c_i / sum c) + 1, so
    costs (1, 1, 1, 5) (n c / sum = .5, .5, .5, 2.5) give (2, 2, 2, 4) (floor or round
    would not); weighted w_i c_i / sum(w c): w = (2, 1, 1), costs (1, 2, 1) give
    (.4, .4, .2)."""
    p = mco.coreset_probs(1, 2, [1., 1., 1., 5.], np.array([0, 0, 1, 1], np.uint32))
    assert p.tolist() == [2., 2., 2., 4.]
    p = mco.coreset_probs(1, 2, [1., 2., 1.], np.array([0, 1, 1], np.uint32), w=[2., 1., 1.])
    np.testing.assert_allclose(p, [.4, .4, .2], rtol=2e-7)


# One round of local search++ (lsearchpp.h:60-110) on 1-d points, worked out by hand:
# rows whose cost to the drawn row exceeds their distance charge (second best - best)
# to their best center (nothing when the two are equal); the others credit
# (distance - new cost); the center with the smallest charge (the first on ties) is
# swapped for the drawn row when its charge is below the credit.
LS_CASES = [
    # X = 0, 1, 10, 11, 20 seeded at rows (0, 1): every possible draw replaces slot 0
    # (charges (1, 1), credit 441 / 460 / 380 for draws 2 / 3 / 4).
    ([0, 1, 10, 11, 20], [0, 1], {2: [2, 1], 3: [3, 1], 4: [4, 1]}),
    # X = 0..5, 50, 51 seeded at rows (4, 0):
    # draw 6 or 7: charges (48, 24), credit 4324 -> slot 1 replaced;
    # draw 1: charges (824, 16), credit 4 -> kept (charging new cost - distance instead
    #   would give slot 1 a charge of 1 and swap);
    # draws 2, 3: charges (816, 16) / (816, 24), credit 4 -> kept;
    # draw 5: charges (24, 24), credit 185 -> the first of the tied centers, slot 0.
    ([0, 1, 2, 3, 4, 5, 50, 51], [4, 0], {6: [4, 6], 7: [4, 7], 1: [4, 0], 2: [4, 0], 3: [4, 0], 5: [5, 0]}),
    # X = 0, 2, 4 seeded at rows (0, 1): the only draw is row 2, charges (4, 4) equal
    # the credit 4, and a zero gain is no swap.
    ([0, 2, 4], [0, 1], {2: [0, 1]}),
    # X = 0, 1, -1, -2 seeded at rows (0, 1): draw 3 leaves row 2's cost unchanged
    # (1 = 1), so it credits 0 and is not charged: charges (1, 1), credit 4 -> slot 0
    # (charging it, 3 to slot 0, would swap slot 1 instead); draw 2: charges (1, 1),
    # credit 4 -> slot 0.
    ([0, 1, -1, -2], [0, 1], {3: [3, 1], 2: [2, 1]}),
]


@pytest.mark.parametrize("case", range(len(LS_CASES)))
def test_localsearch_one_round_hand_cases(mco, case):
    xs, start, table = LS_CASES[case]
    X = np.array(xs, np.float32)[:, None]
    n = len(xs)
    d0 = np.min((X - X[start].T) ** 2, axis=1).astype(np.float64)
    fw = mco.fixed_weights(d0)
    seen = set()
    seed = 0
    while seen != set(table) and seed < 200000:
        seed += 1
        sel = int(mco.fixed_draw(fw, mco.wyrand_at(seed, 0)))
        if sel in seen:
            continue
        assert sel in table, sel
        seen.add(sel)
        c, a, d = mco.localsearch(mco.SQRL2, X, start, 1, seed=seed, j0=0)
        assert c.tolist() == table[sel], (sel, c.tolist())
        M = (X - X[c].T) ** 2
        assert np.array_equal(a, np.argmin(M, axis=1)) and np.array_equal(d, M.min(axis=1))
    assert seen == set(table)


def test_lloyd_step_means_and_reseed(mco):
    """One lloyd_iteration pass (kmeans.h:359-479): without empty clusters the new
    centers are the weighted means of the brute-force assignment and the cost is the
    FP64 sum against the input centers.  With an empty cluster, worked by hand on
    x = 0, 1, 10, 11 and centers .5, 10.5, 100 (cluster 2 empty): every row costs .25
    to its new mean, the row r that the sampler draws (uniform here) moves to
    cluster 2 and becomes its center, and its partner is left alone in its cluster;
    weighted w = 1, 3, 1, 1: the first mean is .75 and the draw weights are
    (.5625, .1875, .25, .25)."""
    rng = np.random.default_rng(8)
    X = rng.standard_normal((900, 7)).astype(np.float32)
    C0 = X[:12].copy()
    for w in (None, rng.uniform(.5, 2., 900)):
        C, a, cost, draws = mco.lloyd_step(X, C0, w)
        a_ref = np.argmin(((X[:, None, :].astype(np.float64) - C0[None].astype(np.float64)) ** 2).sum(-1), axis=1)
        assert draws == 0 and np.array_equal(a, a_ref)
        ww = np.ones(900) if w is None else w
        for c in range(12):
            m = a_ref == c
            np.testing.assert_allclose(C[c], (ww[m, None] * X[m]).sum(0) / ww[m].sum(), rtol=1e-6, atol=1e-7)
        want = (ww * ((X.astype(np.float64) - C0[a_ref].astype(np.float64)) ** 2).sum(1)).sum()
        assert cost == pytest.approx(want, rel=1e-12)
    X = np.array([[0.], [1.], [10.], [11.]], np.float32)
    C0 = np.array([[.5], [10.5], [100.]], np.float32)
    for w, dw, c0, want_cost in ((None, [.25] * 4, .5, 1.), ([1., 3., 1., 1.], [.5625, .1875, .25, .25], .75, 1.5)):
        r = int(mco.fixed_draw(mco.fixed_weights(dw), mco.wyrand_at(0, 0)))
        C, a, cost, draws = mco.lloyd_step(X, C0, w)
        a_want = np.array([0, 0, 1, 1])
        a_want[r] = 2
        C_want = np.array([c0, 10.5, 100.])
        C_want[2] = X[r, 0]
        C_want[a_want[r ^ 1]] = X[r ^ 1, 0]
        assert draws == 1 and a.tolist() == a_want.tolist() and C[:, 0].tolist() == C_want.tolist()
        assert cost == want_cost


def test_coreset_sample_unbiased(mco):
    rng = np.random.default_rng(21)
    n = 500
    asn = rng.integers(0, 5, n).astype(np.uint32)
    costs = rng.exponential(2., n)
    p = mco.coreset_probs(4, 5, costs, asn)
    tot = []
    for s in range(1, 41):
        wt, idx = mco.coreset_sample(p, 200, s)
        assert idx.max() < n
        np.testing.assert_allclose(wt, 1. / (200 * p[idx].astype(np.float64)), rtol=1e-6)
        tot.append(wt.astype(np.float64).sum())
    # E[sum of weights] = n
    assert abs(np.mean(tot) - n) < 0.05 * n


# ---- median centroid policies (centroid.h:434-647) -------------------------------------
def _median_case(seed=21, n=601, d=9, k=5):
    rng = np.random.default_rng(seed)
    X = (rng.normal(size=(n, d)) * 3 + rng.integers(-5, 5, size=(1, d))).astype(np.float32)
    asn = rng.integers(0, k, n).astype(np.uint32)
    return X, asn, np.ones(n), np.zeros((k, d), np.float32)


def test_l1_median_unweighted_is_numpy_median(mco):
    X, asn, cost, C0 = _median_case()
    asn[asn == 4] = 3
    asn[:7] = 4                             # an odd cluster; the random ones give even sizes too
    asn[7:9] = 0
    C, a, _ = mco.median_centers(mco.L1, X, asn, cost, C0)
    for q in range(5):
        ref = np.median(X[a == q], axis=0).astype(np.float32)   # float32: (a+b)/2 in float
        assert np.array_equal(C[q], ref), q
    assert {int((a == q).sum()) % 2 for q in range(5)} == {0, 1}


def test_l1_median_weighted_is_heaviest_member(mco):
    # weighted_median's weight total stays 0 (kmedian.h:275-285): the heaviest member's values
    X, asn, cost, C0 = _median_case(seed=22)
    rng = np.random.default_rng(3)
    w = rng.random(len(X)) + 0.1
    C, a, _ = mco.median_centers(mco.L1, X, asn, cost, C0, w=w)
    for q in range(5):
        mem = np.nonzero(a == q)[0]
        assert np.array_equal(C[q], X[mem[np.argmax(w[mem])]])
    # tied heaviest weights: per column the smallest of the tied rows
    w2 = np.ones(len(X))
    C2, a2, _ = mco.median_centers(mco.L1, X, asn, cost, C0, w=w2)
    for q in range(5):
        assert np.array_equal(C2[q], X[a2 == q].min(0))


def test_tvd_median_unweighted_and_weighted(mco):
    X, asn, cost, C0 = _median_case(seed=23)
    X = np.abs(X) + 0.5
    C, a, _ = mco.median_centers(mco.TVD, X, asn, cost, C0)
    P = X.astype(np.float64) * (1. / X.astype(np.float64).sum(1, keepdims=True))
    for q in range(5):
        assert np.array_equal(C[q], np.median(P[a == q], axis=0).astype(np.float32))
    # weighted: the member of rank j carries w[j] (the full vector indexed by rank, kmedian.h:340)
    w = np.random.default_rng(4).random(len(X)) + 0.1
    Cw, aw, _ = mco.median_centers(mco.TVD, X, asn, cost, C0, w=w)
    for q in range(5):
        mem = np.nonzero(aw == q)[0]
        j = int(np.argmax(w[:len(mem)]))
        assert np.array_equal(Cw[q], P[mem[j]].astype(np.float32))


def test_single_member_center_is_raw_row(mco):
    X, asn, cost, C0 = _median_case(seed=24)
    X = np.abs(X) + 1
    asn[asn == 2] = 1
    asn[10] = 2
    for msr in (mco.L1, mco.TVD, mco.L2):
        C, a, _ = mco.median_centers(msr, X, asn, cost, C0)
        assert np.array_equal(C[2], X[10]), msr


def test_geometric_median_optimal(mco):
    from scipy.optimize import minimize
    X, asn, cost, C0 = _median_case(seed=25, n=300, d=4, k=3)
    C0 = np.stack([X[asn == q].mean(0) + 0.3 for q in range(3)]).astype(np.float32)
    w = np.random.default_rng(5).random(len(X)) + 0.2
    for ww in (None, w):
        C, a, _ = mco.median_centers(mco.L2, X, asn, cost, C0, w=ww)
        for q in range(3):
            P = X[a == q].astype(np.float64)
            wq = np.ones(len(P)) if ww is None else ww[a == q]
            f = lambda c: float((wq * np.sqrt(((P - c) ** 2).sum(1))).sum())
            opt = minimize(f, P.mean(0), method="Nelder-Mead", options=dict(xatol=1e-9, fatol=1e-12, maxiter=20000))
            assert f(C[q].astype(np.float64)) <= opt.fun * (1 + 1e-6), q
            np.testing.assert_allclose(C[q], opt.x, atol=2e-3)
    # symmetric points: the center of symmetry
    S = np.array([[1, 0], [-1, 0], [0, 1], [0, -1], [2, 2], [-2, -2], [2, -2], [-2, 2]], np.float32) + 5
    Cs, _, _ = mco.median_centers(mco.L2, S, np.zeros(8, np.uint32), np.ones(8), np.array([[5.3, 4.8]], np.float32))
    # (the stop rule compares float costs, so the iterate stops within ~sqrt(ulp) of the optimum)
    np.testing.assert_allclose(Cs[0], [5, 5], atol=1e-3)


def test_median_restart_fills_empty_clusters(mco):
    X, asn, cost, C0 = _median_case(seed=26)
    asn[asn == 3] = 0
    cost = np.random.default_rng(6).random(len(X)) + 0.5
    for msr in (mco.L1, mco.TVD, mco.L2):
        Xp = np.abs(X) + 1 if msr == mco.TVD else X
        C, a, c = mco.median_centers(msr, Xp, asn, cost, C0)
        assert set(np.unique(a)) == set(range(5))
        # the first draw (stream of n, draw 0) over the costs becomes center 3
        fw = mco.fixed_weights(cost)
        r = mco.fixed_draw(fw, int(mco.wyrand_at(len(X), 0)))
        assert a[r] == 3


def _csr(mco, X):
    import scipy.sparse as sp
    return mco.CSR.from_scipy(sp.csr_matrix(X))


def test_csr_l1_median_rules(mco):
    rng = np.random.default_rng(27)
    n, d = 40, 6
    X = rng.integers(1, 9, size=(n, d)).astype(np.float32)
    X[rng.random((n, d)) < 0.2] = 0
    for m in (5, 6, 7, 2):
        asn = np.full(n, 1, np.uint32)
        asn[:m] = 0
        C, a, _ = mco.median_centers(mco.L1, _csr(mco, X), asn, np.ones(n), np.zeros((2, d), np.float32))
        F = np.sort(X[:m], axis=0)
        h = m // 2
        for j in range(d):
            nel = int((X[:m, j] != 0).sum())
            if nel > h:
                want = F[h, j] if m & 1 else np.float32(0.5) * (F[h, j] + F[min(h + 1, m - 1), j])
            elif m & 1 and nel == h:
                want = np.float32(0.5) * X[:m, j][X[:m, j] > 0].min()
            else:
                want = 0
            assert C[0, j] == want, (m, j)


def test_csr_tvd_even_rule_and_dense_odd(mco):
    rng = np.random.default_rng(28)
    n, d = 30, 5
    X = rng.integers(1, 9, size=(n, d)).astype(np.float32)
    asn = np.zeros(n, np.uint32)
    asn[15:] = 1
    C, a, _ = mco.median_centers(mco.TVD, _csr(mco, X), asn, np.ones(n), np.zeros((2, d), np.float32))
    inv = (1. / X.astype(np.float64).sum(1)).astype(np.float32)
    P = X * inv[:, None]
    for q, m in ((0, 15), (1, 15)):
        assert np.array_equal(C[q], np.sort(P[a == q], axis=0)[m // 2])
    # no implicit zeros and odd counts: CSR == dense TVD median up to the float vs double value
    Cd, _, _ = mco.median_centers(mco.TVD, X, asn, np.ones(n), np.zeros((2, d), np.float32))
    np.testing.assert_allclose(C, Cd, rtol=1e-6)


def test_csr_geometric_median_optimal(mco):
    """CSR geometric median (csc.h:1131-1203, position-indexed member costs, FP64,
    1e-80 floor): each center's cost within 1e-7 of scipy's Nelder-Mead optimum of the
    same rows, and at least as low as the dense rule's (float costs) on the same rows."""
    from scipy.optimize import minimize
    rng = np.random.default_rng(17)
    n, d = 90, 4
    Xd = np.abs(rng.normal(size=(n, d))).astype(np.float32)
    Xd[Xd < .3] = 0.
    ptr = np.zeros(n + 1, np.uint64)
    idx, val = [], []
    for i in range(n):
        nz = np.nonzero(Xd[i])[0]
        idx += nz.tolist(); val += Xd[i, nz].tolist()
        ptr[i + 1] = len(idx)
    X = mco.CSR(np.array(val, np.float32), np.array(idx, np.uint32), ptr, (n, d))
    asn = (np.arange(n) % 3).astype(np.uint32)
    C0 = Xd[:3] + np.float32(.2)
    C, _, _ = mco.median_centers(mco.L2, X, asn.copy(), np.ones(n), C0.copy())
    Cd, _, _ = mco.median_centers(mco.L2, Xd, asn.copy(), np.ones(n), C0.copy())
    for c in range(3):
        P = Xd[asn == c].astype(np.float64)
        f = lambda y: float(np.linalg.norm(P - y, axis=1).sum())
        best = minimize(f, P.mean(0), method="Nelder-Mead", options=dict(xatol=1e-11, fatol=1e-13, maxiter=40000))
        assert f(C[c].astype(np.float64)) <= best.fun * (1 + 1e-7)
        assert f(C[c].astype(np.float64)) <= f(Cd[c].astype(np.float64)) * (1 + 1e-9)


def test_lloyd_median_rules(mco):
    from minicore_amd.datagen import gaussian_mixture
    X, lab, means = gaussian_mixture(1500, 6, 5, seed=12, lo=-3, hi=3)
    C0 = X[[0, 1, 2, 3, 4]].copy()
    for msr in (mco.L1, mco.L2):
        r = mco.lloyd_median(msr, X, C0, eps=1e-10, maxiter=50)
        assert r["finalcost"] <= r["initcost"]
        D = mco.cost_matrix(msr, X, r["centers"], 0.)
        assert np.array_equal(D.argmin(1), r["asn"])
        if msr == mco.L1:   # converged: the centers are the medians of their assignment
            C2, _, _ = mco.median_centers(msr, X, r["asn"], r["costs"], r["centers"])
            assert np.array_equal(C2, r["centers"])
