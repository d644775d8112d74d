"""Serving layer (mcg_server_*): the request pipeline against the oracle's
restatement of the same steps, results identical alone and in a concurrent mix,
cancellation, injected stage failures and HBM admission."""
import math

import numpy as np
import pytest

from minicore_amd.datagen import gaussian_mixture, poisson_csr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def oracle_pipeline(mco, X, k, seed, frac=0.01, maxiter=9, w=None):
    """The request pipeline on the CPU oracle (dense SQRL2): D^2 seeding,
    LBK sensitivities, m = min(n, max(ceil(frac n), k)) draws from counter 2^32,
    weighted Lloyd on the coreset rows, then assignment of all rows."""
    n = X.shape[0]
    k = min(k, n)
    ids, asn, dist = mco.kmeanspp(mco.SQRL2, X, k, seed=seed, w=w)
    probs = mco.coreset_probs(4, k, dist.astype(np.float64), asn, w)
    m = min(n, max(math.ceil(frac * n), k))
    cw, idx = mco.coreset_sample(probs, m, seed, w, first_draw=1 << 32)
    csw = (w[idx] if w is not None else 1.0) * cw.astype(np.float64)
    sub = X[idx.astype(np.int64)]
    res = mco.lloyd(mco.SQRL2, sub, X[ids.astype(np.int64)], w=csw, eps=0., maxiter=maxiter)
    a, c, _ = mco.assign_sqrl2_f32(X, res["centers"])
    return dict(centers=res["centers"], asn=a, costs=c, m=m, numiter=res["numiter"],
                cost=float(np.dot(c.astype(np.float64), w if w is not None else np.ones(n))))


@pytest.mark.parametrize("weighted", [False, True])
def test_request_matches_oracle_pipeline(mcg, mco, weighted):
    say(f"serve dense pipeline vs oracle weighted={weighted}")
    n, d, k = 6000, 24, 16
    X, _, _ = gaussian_mixture(n, d, k, 5)
    w = np.random.default_rng(9).uniform(0.5, 2., n) if weighted else None
    ref = oracle_pipeline(mco, X, k, seed=21, frac=0.05, w=w)
    with mcg.Server(0, nworkers=2) as srv:
        t = srv.submit(X, k, seed=21, coreset_frac=0.05, weights=w)
        r = srv.wait(t, 120.)
        assert r["state"] == mcg.REQ_DONE, r
        out = srv.outputs(t)
        assert r["coreset_m"] == ref["m"]
        assert r["numiter"] == ref["numiter"]
        np.testing.assert_allclose(out["centers"], ref["centers"], rtol=1e-6, atol=1e-6)
        assert np.array_equal(out["asn"], ref["asn"])
        np.testing.assert_allclose(out["costs"], ref["costs"], rtol=1e-6, atol=1e-6)
        assert r["cost"] == pytest.approx(ref["cost"], rel=1e-9)
        srv.release(t)


def _mix():
    """Requests of the serving mix: dense SQRL2 at d = 64 / 256, sparse JSD."""
    rng = np.random.default_rng(77)
    reqs = []
    for i in range(9):
        kind = i % 3
        n = int(np.exp(rng.uniform(np.log(1500), np.log(40000))))
        k = int(np.exp(rng.uniform(np.log(5), np.log(120))))
        if kind == 2:
            d = 2000
            data, idx, ptr = poisson_csr(n, d, 0.02, 100 + i, normalize=True)
            reqs.append(dict(csr=(data, idx, ptr, (n, d)), k=k, msr=4, prior=1. / d, seed=i + 1))
        else:
            d = 64 if kind == 0 else 256
            X, _, _ = gaussian_mixture(n, d, max(k, 2), 100 + i)
            reqs.append(dict(X=X, k=k, msr=2, seed=i + 1))
    return reqs


def test_results_identical_alone_and_in_mix(mcg):
    reqs = _mix()
    alone = []
    with mcg.Server(0, nworkers=1) as srv:
        for i, q in enumerate(reqs):
            say(f"alone {i}: k={q['k']} msr={q['msr']}")
            t = srv.submit(**q)
            r = srv.wait(t, 300.)
            assert r["state"] == mcg.REQ_DONE, r
            o = srv.outputs(t)
            alone.append((r, {k: v.copy() for k, v in o.items() if v is not None}))
            srv.release(t)
    say("mix: all requests at once on 4 workers")
    with mcg.Server(0, nworkers=4) as srv:
        ts = [srv.submit(**q) for q in reqs]
        for i, t in enumerate(ts):
            r = srv.wait(t, 300.)
            assert r["state"] == mcg.REQ_DONE, (i, r)
            o = srv.outputs(t)
            ra, oa = alone[i]
            assert r["cost"] == ra["cost"], i
            assert r["numiter"] == ra["numiter"] and r["coreset_m"] == ra["coreset_m"], i
            for key in ("centers", "asn", "costs"):
                assert np.array_equal(o[key], oa[key]), (i, key)
            srv.release(t)


def test_csr_request_consistent(mcg, ctx):
    """A sparse JSD request: the returned assignment is the argmin of the cost
    matrix against the returned centers."""
    say("serve csr jsd")
    n, d, k = 3000, 2000, 12
    data, idx, ptr = poisson_csr(n, d, 0.02, 5, normalize=True)
    with mcg.Server(0, nworkers=1) as srv:
        t = srv.submit(csr=(data, idx, ptr, (n, d)), k=k, msr=4, prior=1. / d, seed=3, coreset_frac=0.05)
        r = srv.wait(t, 120.)
        assert r["state"] == mcg.REQ_DONE, r
        o = srv.outputs(t)
    K = mcg.Dataset.csr(data, idx, ptr, (n, d), ctx).cost_matrix(4, o["centers"], prior=1. / d)
    Ks = np.sort(K, axis=1)
    np.testing.assert_allclose(o["costs"], Ks[:, 0], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(K[np.arange(n), o["asn"]], o["costs"], rtol=1e-6, atol=1e-7)
    clear = Ks[:, 1] - Ks[:, 0] > 1e-5 * Ks[:, 1]
    assert np.array_equal(o["asn"][clear], K.argmin(1)[clear])
    assert r["cost"] == pytest.approx(float(o["costs"].astype(np.float64).sum()), rel=1e-9)


def test_cancel_and_failure_injection(mcg):
    say("serve cancel / injected failure")
    X, _, _ = gaussian_mixture(20000, 64, 40, 8)
    Y, _, _ = gaussian_mixture(4000, 32, 10, 9)
    with mcg.Server(0, nworkers=1) as srv:
        t1 = srv.submit(X, 200, seed=1)
        t2 = srv.submit(Y, 10, seed=2)
        t3 = srv.submit(Y, 10, seed=2, fail_at_stage=2)
        t4 = srv.submit(Y, 10, seed=2)
        srv.cancel(t2)          # queued behind t1 on the single worker, or just started
        r1, r2, r3, r4 = (srv.wait(t, 300.) for t in (t1, t2, t3, t4))
        assert r1["state"] == mcg.REQ_DONE
        assert r2["state"] in (mcg.REQ_CANCELLED, mcg.REQ_DONE)
        assert r3["state"] == mcg.REQ_FAILED and "injected failure at stage 2" in r3["error"]
        assert r4["state"] == mcg.REQ_DONE
        for t in (t1, t2, t3, t4):
            srv.release(t)
        with pytest.raises(mcg.MCGError):
            srv.poll(t1)


def test_admission_budget(mcg):
    say("serve admission")
    X, _, _ = gaussian_mixture(5000, 64, 10, 1)
    with mcg.Server(0, nworkers=2, hbm_budget=2 << 20) as srv:
        with pytest.raises(mcg.MCGError) as e:
            srv.submit(X, 10)
        assert e.value.code == -2
    # each request needs ~3.8 MB: with 5 MiB the two workers take them one at a time
    with mcg.Server(0, nworkers=2, hbm_budget=5 << 20) as srv:
        ts = [srv.submit(X, 10, seed=s) for s in range(6)]
        for t in ts:
            assert srv.wait(t, 120.)["state"] == mcg.REQ_DONE
