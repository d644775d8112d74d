"""Dense squared-L2 hot path on the GPU against the CPU oracle (through the C ABI)."""
import numpy as np
import pytest

from minicore_amd.datagen import gaussian_mixture

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def mixture(n, d, k, seed, lo=-10., hi=10., scale=(1., 1.)):
    X, lab, means = gaussian_mixture(n, d, k, seed, lo, hi, scale)
    rng = np.random.default_rng(seed + 100)
    C = (means + 0.3 * rng.standard_normal(means.shape)).astype(np.float32)
    return X, C


CASES = [(20000, 50, 25, 0), (3000, 17, 3, 1), (4000, 64, 64, 2), (5000, 100, 65, 3), (2500, 256, 130, 4),
         (1500, 300, 7, 5), (777, 33, 1, 6)]


@pytest.mark.parametrize("n,d,k,seed", CASES)
def test_assign_exact_and_16bit_bitexact(mcg, ctx, mco, n, d, k, seed):
    say(f"assign n={n} d={d} k={k} seed={seed}")
    X, C = mixture(n, d, k, seed)
    a_ref, c_ref, gap = mco.assign_sqrl2_f32(X, C)
    ds = mcg.Dataset.dense(X, ctx)
    for prec in (mcg.PREC_EXACT, mcg.PREC_BF16, mcg.PREC_FP16):
        a, c = ds.assign(mcg.SQRL2, C, precision=prec)
        assert np.array_equal(a, a_ref), (prec, np.nonzero(a != a_ref)[0][:10])
        assert np.array_equal(c, c_ref), prec
    a, _ = ds.assign(mcg.SQRL2, C, precision=mcg.PREC_FP32)
    bad = a != a_ref
    # plain FP32 may differ only where the oracle's top-2 gap is below 1e-6 relative
    assert np.all(gap[bad] < 1e-6), gap[bad][:10]
    ds.close()


@pytest.mark.parametrize("form", ["BF16", "FP16"])
@pytest.mark.parametrize("n,d,k,seed", [(500, 20, 8, 11), (3000, 64, 40, 12), (20000, 128, 96, 13)])
def test_tuned_form_bitexact(mcg, mco, n, d, k, seed, form):
    """With the E-step tuning on, whichever kernel the timing picks (exact for the
    smallest shape, measured for the others) gives the oracle's assignment; the
    second call reuses the cached choice."""
    say(f"tuned assign n={n} d={d} k={k}")
    X, C = mixture(n, d, k, seed)
    a_ref, c_ref, _ = mco.assign_sqrl2_f32(X, C)
    tc = mcg.Context(0)
    try:
        ds = mcg.Dataset.dense(X, tc)
        for _ in range(2):
            a, c = ds.assign(mcg.SQRL2, C, precision=getattr(mcg, "PREC_" + form))
            assert np.array_equal(a, a_ref)
            assert np.array_equal(c, c_ref)
        ds.close()
    finally:
        tc.close()


@pytest.mark.parametrize("form", ["BF16", "FP16"])
def test_16bit_near_ties_are_rescored(mcg, ctx, mco, form):
    """Points equidistant (in exact arithmetic) from two centers, or sitting on three
    duplicated centers: both refinement paths run and the result stays exact."""
    say(f"near ties {form}")
    rng = np.random.default_rng(0)
    d, k = 128, 40
    C = rng.uniform(-5, 5, (k, d)).astype(np.float32)
    C[5] = C[4]                 # exact duplicate: lowest index must win
    C[9] = C[8]; C[10] = C[8]   # triplicate: needs the full re-score
    pts = []
    for t in range(3000):
        i, j = rng.choice([0, 1, 2, 3, 6, 7, 11, 12, 13], 2, replace=False)
        lam = 0.5 + rng.choice([0., 1e-7, -1e-7, 1e-5, 1e-3])
        pts.append(lam * C[i] + (1 - lam) * C[j] + 1e-4 * rng.standard_normal(d))
    for t in range(500):
        pts.append(C[8] + 0.3 * rng.standard_normal(d))
        pts.append(C[4] + 0.3 * rng.standard_normal(d))
    X = np.array(pts, dtype=np.float32)
    a_ref, c_ref, _ = mco.assign_sqrl2_f32(X, C)
    ds = mcg.Dataset.dense(X, ctx)
    a, c = ds.assign(mcg.SQRL2, C, precision=getattr(mcg, "PREC_" + form))
    pairs, full = ctx.refine_counts()
    say(f"refined pairs={pairs} full={full}")
    assert pairs > 0 and full > 0
    assert np.array_equal(a, a_ref)
    assert np.array_equal(c, c_ref)
    assert not np.any(a == 5) and not np.any(a == 9) and not np.any(a == 10)


@pytest.mark.parametrize("form", ["BF16", "FP16"])
def test_16bit_candidate_overflow(mcg, ctx, mco, form):
    """40 identical centers give every point 40 exact-tie candidates, more than the
    candidate buffer holds (8 per point): the overflow path re-scores over all centers."""
    say(f"candidate overflow {form}")
    prec = getattr(mcg, "PREC_" + form)
    rng = np.random.default_rng(1)
    d = 64
    base = rng.uniform(-3, 3, d).astype(np.float32)
    C = np.concatenate([np.tile(base, (40, 1)), rng.uniform(-3, 3, (5, d)).astype(np.float32)])
    X = (base + 0.5 * rng.standard_normal((2000, d))).astype(np.float32)
    a_ref, c_ref, _ = mco.assign_sqrl2_f32(X, C)
    ds = mcg.Dataset.dense(X, ctx)
    a, c = ds.assign(mcg.SQRL2, C, precision=prec)
    assert np.array_equal(a, a_ref) and np.array_equal(c, c_ref)
    a2, _ = ds.assign(mcg.SQRL2, C, precision=prec)   # flags were reset
    assert np.array_equal(a2, a_ref)


@pytest.mark.parametrize("form", ["BF16", "FP16"])
def test_far_from_origin(mcg, ctx, mco, form):
    """Data offset far from the origin: centering keeps the 16-bit cross term usable."""
    say(f"offset data {form}")
    X, C = mixture(4000, 64, 16, 9, lo=1000., hi=1010.)
    a_ref, c_ref, _ = mco.assign_sqrl2_f32(X, C)
    ds = mcg.Dataset.dense(X, ctx)
    a, c = ds.assign(mcg.SQRL2, C, precision=getattr(mcg, "PREC_" + form))
    assert np.array_equal(a, a_ref)
    assert np.array_equal(c, c_ref)


@pytest.mark.parametrize("scale", [1e-12, 1e-3, 1e7])
def test_fp16_scaling(mcg, ctx, mco, scale):
    """FP16 operands are power-of-two scaled into FP16's range: tiny and huge data
    give the oracle's assignment, also with centers far outside the data (whose
    operands saturate) and a center at the data's mean (operand rows of zeros)."""
    say(f"fp16 scaling {scale}")
    X, C = mixture(3000, 48, 12, 21)
    X = (X * scale).astype(np.float32)
    C = (C * scale).astype(np.float32)
    C[0] = X.mean(0)
    C[1] = C[1] * 1e4
    C[2] = -C[2] * 3e5
    a_ref, c_ref, _ = mco.assign_sqrl2_f32(X, C)
    ds = mcg.Dataset.dense(X, ctx)
    a, c = ds.assign(mcg.SQRL2, C, precision=mcg.PREC_FP16)
    assert np.array_equal(a, a_ref), np.nonzero(a != a_ref)[0][:10]
    assert np.array_equal(c, c_ref)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("d,k", [(70, 37), (20, 2500)])
def test_centroids_bitexact(mcg, ctx, mco, weighted, d, k):
    """k = 2500 puts several clusters on each thread of the single-block count scan and
    leaves dozens of clusters empty; 9000 rows end in a partial 1024-row sort block."""
    say(f"centroids weighted={weighted} d={d} k={k}")
    rng = np.random.default_rng(3)
    X = rng.standard_normal((9000, d)).astype(np.float32)
    asn = rng.integers(0, k, 9000).astype(np.uint32)
    asn[asn == 11] = 12  # one empty cluster
    asn[:700] = 3        # one big cluster (several 256-member chunks)
    w = rng.uniform(0.1, 3., 9000) if weighted else None
    prev = rng.standard_normal((k, d)).astype(np.float32)
    ref, e_ref = mco.centroids(X, asn, k, prev, w)
    ds = mcg.Dataset.dense(X, ctx)
    got, e = ds.centroids(asn, k, prev, w)
    assert e == e_ref
    assert e == 1 if k == 37 else e > 10
    assert np.array_equal(got, ref)


def test_centroid_reference_means_gpu(mcg, ctx):
    say("centroid reference means")
    dm = np.eye(10, dtype=np.float32) * 10
    ds = mcg.Dataset.dense(dm, ctx)
    got, _ = ds.centroids(np.zeros(10, np.uint32), 1, np.zeros((1, 10)))
    assert np.all(got == 1.)
    dm2 = np.zeros((500, 5), np.float32); dm2[:, 0] = 1; dm2[:, 3] = -1.4
    ds = mcg.Dataset.dense(dm2, ctx)
    got, _ = ds.centroids(np.zeros(500, np.uint32), 1, np.zeros((1, 5)))
    np.testing.assert_allclose(got[0], [1, 0, 0, -1.4, 0], rtol=1e-7)


def test_lloyd_matches_oracle(mcg, ctx, mco):
    say("lloyd config-1 shape bf16")
    X, _, means = gaussian_mixture(20000, 50, 25, 0)
    C0 = X[np.random.default_rng(0).choice(20000, 25, replace=False)].copy()
    ref = mco.lloyd(mco.SQRL2, X, C0, eps=1e-10, maxiter=100)
    ds = mcg.Dataset.dense(X, ctx)
    got = ds.lloyd(mcg.SQRL2, C0, eps=1e-10, maxiter=100, precision=mcg.PREC_BF16)
    assert got["numiter"] == ref["numiter"]
    assert got["initcost"] == pytest.approx(ref["initcost"], rel=1e-6)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-6)
    assert np.array_equal(got["asn"], ref["asn"])
    np.testing.assert_allclose(got["centers"], ref["centers"], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(got["costs"], ref["costs"], rtol=1e-6)


def test_lloyd_plain_fp32_form(mcg, ctx, mco):
    """The plain FP32 form may flip near-ties, so its trajectory can differ; the final
    cost must agree and its assignment must be the argmin of its own centers up to ties."""
    say("lloyd config-1 shape fp32")
    X, _, means = gaussian_mixture(20000, 50, 25, 0)
    C0 = X[np.random.default_rng(0).choice(20000, 25, replace=False)].copy()
    ref = mco.lloyd(mco.SQRL2, X, C0, eps=1e-10, maxiter=100)
    got = mcg.Dataset.dense(X, ctx).lloyd(mcg.SQRL2, C0, eps=1e-10, maxiter=100, precision=mcg.PREC_FP32)
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-4)
    a, c, gap = mco.assign_sqrl2_f32(X, got["centers"])
    bad = a != got["asn"]
    assert np.all(gap[bad] < 1e-6)


def test_lloyd_weighted_and_fixed_iterations(mcg, ctx, mco):
    say("lloyd weighted eps=0")
    X, _, _ = gaussian_mixture(6000, 20, 8, 4, lo=-2, hi=2)
    w = np.random.default_rng(1).uniform(0.5, 2., 6000)
    C0 = X[:8].copy()
    ref = mco.lloyd(mco.SQRL2, X, C0, w=w, eps=0., maxiter=5)
    got = mcg.Dataset.dense(X, ctx).lloyd(mcg.SQRL2, C0, weights=w, eps=0., maxiter=5)
    assert got["numiter"] == ref["numiter"]
    assert np.array_equal(got["asn"], ref["asn"])
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-6)


def test_lloyd_empty_cluster_restart(mcg, ctx, mco):
    say("lloyd restart")
    X, _, _ = gaussian_mixture(3000, 10, 4, 5, lo=-5, hi=5)
    C0 = np.concatenate([X[:4], np.full((2, 10), 1e3, np.float32)])   # two centers far away: empty
    ref = mco.lloyd(mco.SQRL2, X, C0, eps=1e-10, maxiter=20)
    got = mcg.Dataset.dense(X, ctx).lloyd(mcg.SQRL2, C0, eps=1e-10, maxiter=20)
    assert np.array_equal(got["asn"], ref["asn"])
    assert got["numiter"] == ref["numiter"]
    assert got["finalcost"] == pytest.approx(ref["finalcost"], rel=1e-6)


@pytest.mark.parametrize("weighted", [False, True])
def test_lloyd_step_reseed(mcg, ctx, mco, weighted):
    """mcg_lloyd_step (lloyd_iteration, kmeans.h:359-479) against the oracle's pass:
    two centers far from the data lose all rows in the first step and are reseeded
    at drawn rows (kmeans.h:431-454); assignment and centers bit for bit, cost within
    1e-12, over three steps."""
    import torch
    say(f"lloyd_step reseed weighted={weighted}")
    X, C = mixture(6000, 40, 24, 31)
    C[-2:] = 1e3
    w = np.random.default_rng(4).uniform(0.5, 2., 6000) if weighted else None
    ds = mcg.Dataset.dense(X, ctx)
    Ct = torch.from_numpy(C).cuda()
    at = torch.empty(6000, dtype=torch.int32, device="cuda")
    Cr = C.copy()
    for step in range(3):
        cost = ds.lloyd_step(mcg.SQRL2, Ct, at, weights=w)
        Cr, ar, cr, draws = mco.lloyd_step(X, Cr, w)
        if step == 0:
            assert draws == 2
        assert np.array_equal(at.cpu().numpy().astype(np.uint32), ar), step
        assert np.array_equal(Ct.cpu().numpy(), Cr), step
        assert cost == pytest.approx(cr, rel=1e-12), step
    ds.close()


def test_cost_matrix_sqrl2(mcg, ctx, mco):
    say("cost matrix sqrl2")
    X, C = mixture(1000, 40, 30, 7)
    ref = mco.cost_matrix(mco.SQRL2, X, C, 0.)
    got = mcg.Dataset.dense(X, ctx).cost_matrix(mcg.SQRL2, C)
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-3)


def test_device_resident_input(mcg, ctx, mco):
    """A torch tensor on the device is borrowed, not copied; results are identical."""
    import torch
    say("device-resident dataset")
    X, C = mixture(5000, 64, 20, 8)
    Xt = torch.from_numpy(X).to("cuda")
    torch.cuda.synchronize()
    ds = mcg.Dataset.dense(Xt, ctx)
    a, c = ds.assign(mcg.SQRL2, C)
    a_ref, c_ref, _ = mco.assign_sqrl2_f32(X, C)
    assert np.array_equal(a, a_ref) and np.array_equal(c, c_ref)


def _lloyd_runs(mcg, make_ds, msr, C0, prior, w, eps, maxiters, graphs):
    c = mcg.Context(0)
    try:
        c.graphs(graphs)
        ds = make_ds(c)
        out = []
        for t in maxiters:
            out.append(ds.lloyd(msr, C0, prior=prior, weights=w, eps=eps, maxiter=t))
            # eager work and synchronisations between the graph replays of two calls
            ds.assign(msr, C0, prior=prior)
        ds.close()
        return out
    finally:
        c.close()


@pytest.mark.parametrize("kind", ["dense_l2", "dense_l2_weighted", "csr_jsd", "dense_mkl"])
def test_lloyd_graph_replay_matches_eager(mcg, kind):
    """The Lloyd loop with its iterations replayed from a captured graph (batches of
    2, 4, ... per host synchronisation) against the same loop launched kernel by
    kernel: every output identical for every iteration count, so each replayed
    iteration reproduces the eager one."""
    from minicore_amd.datagen import poisson_csr
    say(f"lloyd graph replay {kind}")
    w, eps = None, 0.
    if kind.startswith("dense_l2"):
        X, C0 = mixture(6000, 40, 12, 21)
        msr, prior = mcg.SQRL2, 0.
        if kind.endswith("weighted"):
            w = np.random.default_rng(3).uniform(0.5, 2., 6000)

        def make_ds(c):
            return mcg.Dataset.dense(X, c)
    elif kind == "csr_jsd":
        n, d = 3000, 500
        data, idx, ptr = poisson_csr(n, d, 0.03, 9, normalize=True)
        msr, prior = mcg.JSD, 1. / d
        dense = np.zeros((n, d), np.float32)
        for i in range(n):
            dense[i, idx[ptr[i]:ptr[i + 1]]] = data[ptr[i]:ptr[i + 1]]
        C0 = dense[np.random.default_rng(4).choice(n, 8, replace=False)].copy()

        def make_ds(c):
            return mcg.Dataset.csr(data, idx, ptr, (n, d), c)
    else:
        rng = np.random.default_rng(5)
        X = rng.gamma(2., 1., (4000, 30)).astype(np.float32)
        X /= X.sum(1, keepdims=True)
        msr, prior = mcg.MKL, 1e-3
        C0 = X[rng.choice(4000, 7, replace=False)].copy()

        def make_ds(c):
            return mcg.Dataset.dense(X, c)
    maxiters = [0, 1, 2, 3, 5, 9, 40]
    eager = _lloyd_runs(mcg, make_ds, msr, C0, prior, w, eps, maxiters, False)
    replay = _lloyd_runs(mcg, make_ds, msr, C0, prior, w, eps, maxiters, True)
    for t, a, b in zip(maxiters, eager, replay):
        assert a["numiter"] == b["numiter"], t
        assert a["initcost"] == b["initcost"] and a["finalcost"] == b["finalcost"], t
        for key in ("centers", "asn", "costs"):
            assert np.array_equal(a[key], b[key]), (t, key)
    assert eager[-1]["numiter"] >= 3


@pytest.mark.parametrize("n", [3000, 70000])
def test_nan_rows_and_centers(mcg, ctx, mco, n):
    """NaN in a row makes all its costs NaN: the oracle's scan keeps center 0 with a NaN
    cost.  A center with a NaN never wins.  A NaN center 0 keeps every row (its cost is
    NaN for all of them).  Exact form (centers split over blocks at the smaller n) and
    the 16-bit forms against the oracle."""
    say(f"nan rows/centers n={n}")
    X, C = mixture(n, 32, 40, 31)
    X[5, 3] = np.nan
    X[n - 1, 0] = np.nan
    C[7, 2] = np.nan
    C0 = C.copy()
    C0[0, 1] = np.nan
    for cc in (C, C0):
        a_ref, c_ref, _ = mco.assign_sqrl2_f32(X, cc)
        ds = mcg.Dataset.dense(X, ctx)
        for prec in (mcg.PREC_EXACT, mcg.PREC_BF16, mcg.PREC_FP16):
            a, c = ds.assign(mcg.SQRL2, cc, precision=prec)
            assert np.array_equal(a, a_ref), (prec, np.nonzero(a != a_ref)[0][:10])
            np.testing.assert_array_equal(c, c_ref)
        ds.close()
