"""float64 rows on the GPU: dense SQRL2 hard clustering with every value in FP64
(the reference's FT = double instantiation), bit-exact against the oracle."""
import numpy as np
import pytest

from minicore_amd.datagen import gaussian_mixture

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("n,d,k", [(3000, 17, 11), (20000, 50, 25), (700, 300, 40)])
def test_lloyd_f64_bitexact(mcg, ctx, mco, weighted, n, d, k):
    say(f"lloyd_f64 n={n} d={d} k={k} weighted={weighted}")
    X = gaussian_mixture(n, d, k, 1, dtype=np.float64)[0]
    w = np.random.default_rng(2).uniform(0.5, 2, n) if weighted else None
    C0 = X[np.random.default_rng(3).choice(n, k, replace=False)]
    ref = mco.lloyd_f64(X, C0, w=w, eps=1e-10, maxiter=30)
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.lloyd_f64(C0, weights=w, eps=1e-10, maxiter=30)
    assert got["numiter"] == ref["numiter"]
    assert np.array_equal(got["asn"], ref["asn"])
    assert np.array_equal(got["costs"], ref["costs"])
    assert np.array_equal(got["centers"], ref["centers"])
    assert got["initcost"] == ref["initcost"] and got["finalcost"] == ref["finalcost"]
    ds.close()


def test_lloyd_f64_restart_bitexact(mcg, ctx, mco):
    """Equal starting centers leave clusters empty: the restart draws, the forced rows
    and the means match the oracle."""
    rng = np.random.default_rng(7)
    X = rng.normal(size=(3000, 6))
    C0 = np.vstack([X[0], X[0], X[0], X[5], X[9]])
    ref = mco.lloyd_f64(X, C0, eps=1e-10, maxiter=20)
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.lloyd_f64(C0, eps=1e-10, maxiter=20)
    assert got["numiter"] == ref["numiter"]
    assert np.array_equal(got["asn"], ref["asn"])
    assert np.array_equal(got["centers"], ref["centers"])
    assert got["finalcost"] == ref["finalcost"]
    ds.close()


def test_assign_f64_bitexact(mcg, ctx, mco):
    rng = np.random.default_rng(9)
    X = rng.normal(size=(5000, 33)) * 5.
    C = np.vstack([X[:60], X[10:11]])     # center 60 duplicates center 10: index 10 wins
    ra, rc = mco.assign_sqrl2_f64(X, C)
    ds = mcg.Dataset.dense(X, ctx)
    a, c = ds.assign_f64(C)
    assert np.array_equal(a, ra) and np.array_equal(c, rc)
    assert not np.any(a == 60)
    ds.close()


def test_hcluster_float64_binding(mco):
    """pyminicore's hcluster on a float64 array with SQRL2 runs the FP64 path: float64
    centers and costs equal to the oracle's."""
    import minicore_amd as mc
    X = gaussian_mixture(4000, 20, 8, 4, dtype=np.float64)[0]
    ids = [3, 99, 1000, 2500, 3999, 17, 512, 2048]
    r = mc.hcluster(X, ids, msr=2, maxiter=25)
    ref = mco.lloyd_f64(X, X[ids], eps=1e-10, maxiter=25)
    assert r["centers"].dtype == np.float64 and r["costs"].dtype == np.float64
    assert np.array_equal(r["centers"], ref["centers"])
    assert np.array_equal(np.asarray(r["asn"]), ref["asn"])
    assert r["numiter"] == ref["numiter"]
