"""Pin the oracle's float64 restatement of dense SQRL2 hard clustering (the
reference's FT = double instantiation, python/pycluster_dense.cpp:85-90) against
exact arithmetic, numpy and the rules of perform_hard_clustering (no GPU)."""
import math
from fractions import Fraction

import numpy as np
import pytest


def test_sqrl2_f64_is_the_fma_chain(mco):
    """s = fma(t, t, s) with t = x - c in double, one rounding per step: emulated
    exactly with rationals (float(Fraction) rounds correctly)."""
    rng = np.random.default_rng(1)
    for d in (1, 7, 64, 300):
        x, c = rng.normal(size=d) * 10., rng.normal(size=d) * 10.
        s = 0.
        for j in range(d):
            t = x[j] - c[j]
            s = float(Fraction(t) * Fraction(t) + Fraction(s))
        got = mco.lib().mco_sqrl2_f64(mco._p(np.ascontiguousarray(x)), mco._p(np.ascontiguousarray(c)), d)
        assert got == s
    # the plain sum of squares is only close: the chain is not that association
    assert got == pytest.approx(float(np.sum((x - c) ** 2)), rel=1e-13)


def test_assign_f64_brute_force(mco):
    rng = np.random.default_rng(2)
    X = rng.normal(size=(700, 13)) * 3.
    C = X[rng.choice(700, 9, replace=False)].copy()
    C = np.vstack([C, C[4]])            # a duplicate of center 4: index 4 must win
    asn, cost = mco.assign_sqrl2_f64(X, C)
    D = ((X[:, None, :] - C[None, :, :]) ** 2).sum(-1)
    ref = D.argmin(1)
    srt = np.sort(D, 1)
    clear = (srt[:, 1] - srt[:, 0]) > 1e-9 * np.maximum(srt[:, 0], 1.)
    assert np.array_equal(asn[clear], ref[clear])
    assert not np.any(asn == 9)
    assert np.allclose(cost, D[np.arange(700), asn], rtol=1e-13)


def test_centroids_f64_chunked_order(mco):
    """Unweighted: members in ascending order, 256-member chunks each summed
    sequentially in double, chunk sums added in order; weighted: numpy means."""
    rng = np.random.default_rng(3)
    n, d, k = 1500, 5, 4
    X = rng.normal(size=(n, d))
    asn = rng.integers(0, k, n).astype(np.uint32)
    asn[:700] = 1                      # cluster 1: several chunks
    C0 = np.full((k, d), 7.)
    C, e = mco.centroids_f64(X, asn, k, C0)
    assert e == 0
    for c in range(k):
        mem = np.nonzero(asn == c)[0]
        tot = [0.] * d
        for s in range(0, len(mem), 256):
            part = [0.] * d
            for p in mem[s:s + 256]:
                for j in range(d):
                    part[j] = part[j] + X[p, j]
            for j in range(d):
                tot[j] = tot[j] + part[j]
        assert np.array_equal(C[c], np.array(tot) / float(len(mem)))
    w = rng.uniform(0.5, 2., n)
    asn[asn == 3] = 0                  # cluster 3 empty: keeps its center
    Cw, e = mco.centroids_f64(X, asn, k, C0, w)
    assert e == 1 and np.array_equal(Cw[3], C0[3])
    for c in range(3):
        m = asn == c
        assert np.allclose(Cw[c], (w[m, None] * X[m]).sum(0) / w[m].sum(), rtol=1e-13, atol=1e-15)


def test_fixed_total_exact(mco):
    """sum of floor(w c 2^s) with s = 62 - exponent(max) - bitlen(n), scaled back:
    integer arithmetic, so the order of the terms does not matter."""
    rng = np.random.default_rng(4)
    c = rng.exponential(size=1000) * 1e3
    w = rng.uniform(0.1, 3., 1000)
    t = w * c
    e = math.frexp(float(np.max(np.abs(t))))[1]
    s = 62 - e - (1000).bit_length()
    T = sum(math.floor(math.ldexp(float(v), s)) for v in t if v > 0)
    assert mco.fixed_total(c, w) == math.ldexp(float(T), -s)
    perm = rng.permutation(1000)
    assert mco.fixed_total(c[perm], w[perm]) == mco.fixed_total(c, w)
    assert mco.fixed_total(c, w) == pytest.approx(math.fsum(t), rel=1e-12)


def _numpy_lloyd(X, C, maxiter):
    """Brute-force FP64 Lloyd (no empty clusters expected): assignments and means
    until the assignment stops changing."""
    asn = None
    for it in range(maxiter + 1):
        D = ((X[:, None, :] - C[None, :, :]) ** 2).sum(-1)
        a = D.argmin(1)
        if asn is not None and np.array_equal(a, asn):
            return C, a
        asn = a
        C = np.stack([X[a == c].mean(0) for c in range(C.shape[0])])
    return C, asn


def test_lloyd_f64_reaches_the_fixed_point(mco):
    rng = np.random.default_rng(5)
    means = rng.uniform(-10, 10, (6, 8))
    X = np.vstack([m + rng.normal(size=(150, 8)) for m in means])
    C0 = X[rng.choice(len(X), 6, replace=False)]
    r = mco.lloyd_f64(X, C0, eps=0., maxiter=200)
    Cn, an = _numpy_lloyd(X, C0.copy(), 200)
    assert np.array_equal(r["asn"], an)
    assert np.allclose(r["centers"], Cn, rtol=1e-12)
    # the final state is a fixed point of assignment + means
    a2, c2 = mco.assign_sqrl2_f64(X, r["centers"])
    assert np.array_equal(a2, r["asn"]) and np.array_equal(c2, r["costs"])
    assert r["finalcost"] == mco.fixed_total(r["costs"])
    assert r["finalcost"] <= r["initcost"]


def test_lloyd_f64_maxiter_rule(mco):
    """Up to maxiter + 1 iterations (solve.h:98-175)."""
    rng = np.random.default_rng(6)
    X = rng.normal(size=(800, 4))
    C0 = X[:12].copy()
    for m in (0, 1, 3):
        assert mco.lloyd_f64(X, C0, eps=0., maxiter=m)["numiter"] == m + 1


def test_lloyd_f64_restart(mco):
    """Two equal starting centers: the second has no members (ties go to the lower
    index), so it restarts at the row drawn over the costs with the first output of
    the stream of n, and that row is forced onto it."""
    rng = np.random.default_rng(7)
    X = rng.normal(size=(300, 3))
    C0 = np.vstack([X[0], X[0], X[5]])
    a0, cost0 = mco.assign_sqrl2_f64(X, C0)
    assert not np.any(a0 == 1)
    fw = np.zeros(300, dtype=np.uint64)
    lib = mco.lib()
    lib.mco_fixed_weights(mco._p(np.ascontiguousarray(cost0)), 300, mco._p(fw))
    row = lib.mco_fixed_draw(mco._p(fw), 300, lib.mco_wyrand_at(300, 0))
    r = mco.lloyd_f64(X, C0, eps=0., maxiter=0)
    assert r["numiter"] == 1
    counts = np.bincount(r["asn"], minlength=3)
    assert counts[1] > 0
    C1 = C0.copy()
    C1[1] = X[row]
    a1, _ = mco.assign_sqrl2_f64(X, C1)
    a1[row] = 1
    means, _ = mco.centroids_f64(X, a1, 3, C1)
    assert np.array_equal(r["centers"], means)
