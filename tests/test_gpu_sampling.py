"""D^2 seeding and coreset sampling on the GPU: bit-exact against the oracle
when fed the same inputs."""
import numpy as np
import pytest

from minicore_amd.datagen import gaussian_mixture, poisson_csr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("n,d,k", [(20000, 50, 25), (30000, 64, 200), (200, 8, 5), (1001, 17, 12),
                                   (8192, 64, 40), (70000, 64, 30)])
def test_kmeanspp_dense_bitexact(mcg, ctx, mco, weighted, n, d, k):
    """Rounds after the first skip points whose center is far from the new one; the
    result must still equal the oracle, which evaluates every point.  200 rows is
    less than one block of the update, d = 17 takes scalar loads."""
    say(f"kmeans++ dense n={n} d={d} k={k} weighted={weighted}")
    X, _, _ = gaussian_mixture(n, d, 25, 0)
    w = np.random.default_rng(3).uniform(0.5, 2, n) if weighted else None
    c_ref, a_ref, d_ref = mco.kmeanspp(mco.SQRL2, X, k, seed=13, w=w)
    c, a, dd = mcg.Dataset.dense(X, ctx).kmeanspp(mcg.SQRL2, k, seed=13, weights=w)
    assert np.array_equal(c, c_ref)
    assert np.array_equal(a, a_ref)
    assert np.array_equal(dd, d_ref)


def test_kmeanspp_duplicate_rows_fallback(mcg, ctx, mco):
    """5 distinct rows repeated: once all are chosen every distance is 0 and each
    later round takes the first index of the maximum (kmeans.h:179-186)."""
    say("kmeans++ duplicate rows")
    base = np.random.default_rng(4).normal(size=(5, 16)).astype(np.float32)
    X = base[np.random.default_rng(5).integers(0, 5, 3000)]
    c_ref, a_ref, d_ref = mco.kmeanspp(mco.SQRL2, X, 9, seed=2)
    c, a, dd = mcg.Dataset.dense(X, ctx).kmeanspp(mcg.SQRL2, 9, seed=2)
    assert np.array_equal(c, c_ref)
    assert np.array_equal(a, a_ref)
    assert np.array_equal(dd, d_ref)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("trials,lspp", [(3, 0), (2, 0), (5, 12)])
def test_kmeanspp_local_trials_bitexact(mcg, ctx, mco, weighted, trials, lspp):
    """n_local_trials > 1 (kmeans.h:79-177): candidates, stale copies, exact totals and
    the continued stream of local search++ all equal the oracle."""
    say(f"kmeans++ local trials={trials} lspp={lspp} weighted={weighted}")
    n, d, k = 12000, 40, 30
    X, _, _ = gaussian_mixture(n, d, 30, 4)
    w = np.random.default_rng(8).uniform(0.5, 2, n) if weighted else None
    c_ref, a_ref, d_ref = mco.kmeanspp_local(mco.SQRL2, X, k, trials, seed=17, w=w, lspp=lspp)
    c, a, dd = mcg.Dataset.dense(X, ctx).kmeanspp(mcg.SQRL2, k, seed=17, weights=w, lspp=lspp,
                                                  n_local_trials=trials)
    assert np.array_equal(c, c_ref)
    assert np.array_equal(a, a_ref)
    assert np.array_equal(dd, d_ref)


def test_kmeanspp_local_trials_csr(mcg, ctx, mco):
    """CSR JSD with local trials: the first center is one of the uniform candidates and
    each distance is the cost to the assigned center."""
    say("kmeans++ local trials csr jsd")
    n, d, k, L = 3000, 300, 10, 3
    data, idx, ptr = poisson_csr(n, d, 0.04, 6, normalize=True)
    X = mco.CSR(data, idx, ptr, (n, d))
    c, a, dist = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).kmeanspp(4, k, prior=1. / d, seed=5,
                                                                        n_local_trials=L)
    assert c[0] in [mco.wyrand_at(5, t) % n for t in range(L)]
    ref = mco.cost_matrix(4, X, X.todense()[c], 1. / d)
    np.testing.assert_allclose(dist, ref[np.arange(n), a], rtol=2e-4, atol=2e-6)
    cr, ar, dr = mco.kmeanspp_local(4, X, k, L, prior=1. / d, seed=5)
    assert c[0] == cr[0]


def test_kmeanspp_local_trials_binding(mco):
    """pyminicore's kmeanspp(..., n_local_trials=4) runs the local-trial seeding."""
    import minicore_amd as mc
    say("api kmeanspp n_local_trials=4")
    X, _, _ = gaussian_mixture(8000, 24, 20, 9)
    ids, asn, costs = mc.kmeanspp(X, 20, seed=31, ntimes=0, n_local_trials=4)
    c_ref, a_ref, d_ref = mco.kmeanspp_local(mco.SQRL2, X, 20, 4, seed=31)
    assert np.array_equal(ids, c_ref) and np.array_equal(asn, a_ref) and np.array_equal(costs, d_ref)


def test_kmeanspp_jsd_pruned_rounds(mcg, ctx, mco):
    """JSD seeding skips rows whose center is far from the new one (sqrt(JSD) is a
    metric); every distance must still be the minimum over the chosen centers."""
    say("kmeans++ csr jsd k=60")
    n, d, k = 6000, 500, 60
    data, idx, ptr = poisson_csr(n, d, 0.05, 9, normalize=True)
    X = mco.CSR(data, idx, ptr, (n, d))
    c, a, dist = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).kmeanspp(4, k, prior=1. / d, seed=11)
    assert len(set(c.tolist())) == k
    ref = mco.cost_matrix(4, X, X.todense()[c], 1. / d)
    np.testing.assert_allclose(dist, ref.min(1), rtol=2e-4, atol=2e-6)
    np.testing.assert_allclose(ref[np.arange(n), a], dist, rtol=2e-4, atol=2e-6)


@pytest.mark.parametrize("code", [4, 5])
def test_kmeanspp_csr_consistent(mcg, ctx, mco, code):
    say(f"kmeans++ csr msr={code}")
    n, d = 2000, 300
    data, idx, ptr = poisson_csr(n, d, 0.04, 2, normalize=True)
    X = mco.CSR(data, idx, ptr, (n, d))
    c, a, dist = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).kmeanspp(code, 10, prior=1. / d, seed=7)
    assert c[0] == mco.wyrand_at(7, 0) % n and len(set(c.tolist())) == 10
    Cd = X.todense()[c]
    ref = mco.cost_matrix(code, X, Cd, 1. / d)
    m = dist > 0
    np.testing.assert_allclose(dist[m], ref.min(1)[m], rtol=2e-4, atol=1e-6)


@pytest.mark.parametrize("weighted", [False, True])
def test_kmeanspp_localsearch_bitexact(mcg, ctx, mco, weighted):
    say(f"kmeans++ with local search++ weighted={weighted}")
    X, _, _ = gaussian_mixture(20000, 32, 40, 2)
    w = np.random.default_rng(2).uniform(0.5, 2, 20000) if weighted else None
    c_ref, a_ref, d_ref = mco.kmeanspp_ls(mco.SQRL2, X, 40, 30, seed=7, w=w)
    c, a, dd = mcg.Dataset.dense(X, ctx).kmeanspp(mcg.SQRL2, 40, seed=7, weights=w, lspp=30)
    assert np.array_equal(c, c_ref)
    assert np.array_equal(a, a_ref)
    assert np.array_equal(dd, d_ref)


def test_kmeanspp_localsearch_csr_consistent(mcg, ctx, mco):
    say("kmeans++ with local search++ csr jsd")
    n, d = 3000, 300
    data, idx, ptr = poisson_csr(n, d, 0.04, 5, normalize=True)
    X = mco.CSR(data, idx, ptr, (n, d))
    c, a, dist = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).kmeanspp(4, 12, prior=1. / d, seed=9, lspp=10)
    assert len(set(c.tolist())) == 12
    ref = mco.cost_matrix(4, X, X.todense()[c], 1. / d)
    np.testing.assert_allclose(dist, ref.min(1), rtol=2e-4, atol=2e-6)
    Ks = np.sort(ref, 1)
    clear = Ks[:, 1] - Ks[:, 0] > 1e-4 * Ks[:, 1]
    assert np.array_equal(a[clear], ref.argmin(1)[clear])


@pytest.mark.parametrize("sens", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("weighted", [False, True])
def test_coreset_probs_and_sample_bitexact(mcg, ctx, mco, sens, weighted):
    say(f"coreset sens={sens} weighted={weighted}")
    rng = np.random.default_rng(sens + 10 * weighted)
    n, k = 50000, 37
    asn = rng.integers(0, k, n).astype(np.uint32)
    costs = rng.exponential(3., n)
    w = rng.uniform(0.2, 4., n) if weighted else None
    p_ref = mco.coreset_probs(sens, k, costs, asn, w)
    p = mcg.coreset_probs(sens, k, costs, asn, w, ctx=ctx)
    assert np.array_equal(p, p_ref)
    if sens == 1:
        return
    wt_ref, idx_ref = mco.coreset_sample(p_ref, 5000, 99, w)
    wt, idx = mcg.coreset_sample(p, 5000, 99, w, ctx=ctx)
    assert np.array_equal(idx, idx_ref)
    assert np.array_equal(wt, wt_ref)
