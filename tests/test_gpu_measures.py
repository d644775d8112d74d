"""The further measures (applicator.h:1673-1760, 1778-1891) on the GPU against the
oracle's FP64 restatement: cost matrices over dense and CSR rows, assignment, the
mean-centroid Lloyd for the measures msr2pol maps to a mean, seeding, and the error
for the measures without a centroid policy."""
import numpy as np
import pytest

from minicore_amd.datagen import poisson_csr

pytestmark = pytest.mark.gpu

MORE = [0, 1, 3, 6, 7, 8, 9, 10, 12, 13, 14, 15, 16, 17, 18, 23, 24, 25, 26, 64]


def say(*a):
    print(*a, flush=True)


def _dense(n, d, seed):
    rng = np.random.default_rng(seed)
    return rng.gamma(1.5, 1., (n, d)).astype(np.float32)


def _check(got, ref, code):
    # FP32 pair sums against FP64; acos near 0 and sqrt near 0 lose relative digits
    atol = 2e-3 if code in (15, 16, 7) else 2e-5
    np.testing.assert_allclose(got, ref, rtol=2e-3, atol=atol, err_msg=f"measure {code}")


@pytest.mark.parametrize("code", MORE)
def test_cost_matrix_dense(mcg, ctx, mco, code):
    say(f"cost matrix dense msr={code}")
    X = _dense(300, 40, 1)
    C = _dense(7, 40, 2)
    got = mcg.Dataset.dense(X, ctx).cost_matrix(code, C, prior=0.1)
    ref = mco.cost_matrix(code, X, C, 0.1)
    _check(got, ref, code)


@pytest.mark.parametrize("code", MORE)
def test_cost_matrix_csr(mcg, ctx, mco, code):
    say(f"cost matrix csr msr={code}")
    n, d = 300, 200
    data, idx, ptr = poisson_csr(n, d, 0.05, 4)
    C = _dense(6, d, 5) * (np.random.default_rng(6).random((6, d)) < 0.3)
    prior = 1. / d
    got = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).cost_matrix(code, C, prior=prior)
    ref = mco.cost_matrix(code, mco.CSR(data, idx, ptr, (n, d)), C, prior)
    _check(got, ref, code)


@pytest.mark.parametrize("code", [6, 13, 10, 23, 15, 64])
def test_assign_and_lloyd_mean_measures(mcg, ctx, mco, code):
    say(f"assign + lloyd msr={code}")
    rng = np.random.default_rng(code)
    means = rng.gamma(2., 1., (5, 30))
    lab = rng.integers(0, 5, 2000)
    X = rng.poisson(means[lab] * 3).astype(np.float32) + 0.1
    C0 = X[[int(np.argmax(lab == c)) for c in range(5)]].copy()
    ds = mcg.Dataset.dense(X, ctx)
    a, c = ds.assign(code, C0, prior=0.05)
    K = mco.cost_matrix(code, X, C0, 0.05)
    Ks = np.sort(K, 1)
    clear = Ks[:, 1] - Ks[:, 0] > 1e-3 * np.maximum(Ks[:, 1], 1e-30)
    assert np.array_equal(a[clear], K.argmin(1)[clear])
    got = ds.lloyd(code, C0, prior=0.05, eps=1e-6, maxiter=10)
    ref = mco.lloyd(code, X, C0, prior=0.05, eps=1e-6, maxiter=10)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=2e-3)
    assert np.mean(got["asn"] == ref["asn"]) > 0.99
    ds.close()


@pytest.mark.parametrize("code", [16, 17, 18])
def test_no_policy_measures_are_cost_only(mcg, ctx, code):
    X = _dense(200, 10, 3)
    ds = mcg.Dataset.dense(X, ctx)
    ds.assign(code, X[:4], prior=0.1)          # costs and assignment work
    with pytest.raises(mcg.MCGError):
        ds.lloyd(code, X[:4], prior=0.1, maxiter=3)
    ds.close()


def test_kmeanspp_further_measure(mcg, ctx, mco):
    say("kmeans++ hellinger csr")
    n, d = 1500, 300
    data, idx, ptr = poisson_csr(n, d, 0.05, 8, normalize=True)
    X = mco.CSR(data, idx, ptr, (n, d))
    c, a, dist = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).kmeanspp(6, 12, prior=1. / d, seed=3)
    assert len(set(c.tolist())) == 12
    ref = mco.cost_matrix(6, X, X.todense()[c], 1. / d)
    np.testing.assert_allclose(dist, ref.min(1), rtol=2e-3, atol=2e-5)


def test_cmp_by_name(mcg, mco):
    import minicore_amd as mc
    X = _dense(50, 12, 9)
    C = _dense(4, 12, 10)
    for name, code in (("HELLINGER", 6), ("ITAKURA_SAITO", 13), ("TVD", 9), ("L1", 0)):
        got = np.asarray(mc._ext().cmp(C, X, msr=name, prior=0.1))
        ref = mco.cost_matrix(code, X, C, 0.1)
        assert got.shape in (ref.shape, ref.T.shape)
        g = got if got.shape == ref.shape else got.T
        _check(g, ref, code)
