"""mcg_coreset_probs_noalloc (fixed device scratch, no allocation) against
mcg_coreset_probs through the C ABI: bit-identical probabilities on two calls in a
row, the second smaller; sizes beyond the fixed scratch and bad inputs refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, k, seed, weighted):
    rng = np.random.default_rng(seed)
    asn = rng.integers(0, k, n).astype(np.uint32)
    asn[: min(n, k)] = np.arange(min(n, k), dtype=np.uint32)
    costs = rng.exponential(2., n)
    costs[rng.random(n) < 0.05] = 0.
    w = rng.uniform(0.25, 4., n) if weighted else None
    return costs, asn, w


def _noalloc(mcg, ctx, sens, k, costs, asn, w):
    import torch
    dev = torch.device("cuda:0")
    tc = torch.from_numpy(costs).to(dev)
    ta = torch.from_numpy(asn.astype(np.int32)).to(dev)
    tw = None if w is None else torch.from_numpy(w).to(dev)
    tp = torch.full((len(costs),), float("nan"), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    rc = mcg.load().mcg_coreset_probs_noalloc(ctx.h, sens, len(costs), k, mcg.ptr(tc), mcg.ptr(ta), mcg.ptr(tw),
                                              mcg.ptr(tp))
    mcg.check(rc)
    return tp.cpu().numpy()


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("sens", [0, 1, 2, 3, 4])
def test_noalloc_matches_coreset_probs(mcg, ctx, sens, weighted):
    for n, k, seed in ((300_000, 700, 1), (5_000, 17, 2)):   # two calls in a row, the second smaller
        print(f"noalloc sens={sens} weighted={weighted} n={n} k={k}", flush=True)
        costs, asn, w = _case(n, k, seed + 10 * sens, weighted)
        want = mcg.coreset_probs(sens, k, costs, asn, weights=w, ctx=ctx)
        got = _noalloc(mcg, ctx, sens, k, costs, asn, w)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (n, k)


def test_noalloc_refuses(mcg, ctx):
    import torch
    L = mcg.load()
    dev = torch.device("cuda:0")
    costs = torch.ones(16, dtype=torch.float64, device=dev)
    asn = torch.zeros(16, dtype=torch.int32, device=dev)
    probs = torch.zeros(16, dtype=torch.float32, device=dev)
    args = (mcg.ptr(costs), mcg.ptr(asn), None, mcg.ptr(probs))
    # beyond the fixed scratch: refused before any device work
    assert L.mcg_coreset_probs_noalloc(ctx.h, 4, (1 << 22) + 1, 4, *args) == -1
    assert b"limit" in L.mcg_last_error()
    assert L.mcg_coreset_probs_noalloc(ctx.h, 4, 16, (1 << 16) + 1, *args) == -1
    # host buffers refused
    hc = np.ones(16)
    assert L.mcg_coreset_probs_noalloc(ctx.h, 4, 16, 4, mcg.ptr(hc), args[1], None, args[3]) == -1
    # an assignment >= ncenters is refused (checked on the device before the sums)
    asn[5] = 9
    assert L.mcg_coreset_probs_noalloc(ctx.h, 4, 16, 4, *args) == -1
    assert b"out of range" in L.mcg_last_error()
    # and the entry point still works afterwards
    asn[5] = 1
    costs_h = np.ones(16)
    want = mcg.coreset_probs(4, 4, costs_h, asn.cpu().numpy().astype(np.uint32), ctx=ctx)
    assert L.mcg_coreset_probs_noalloc(ctx.h, 4, 16, 4, *args) == 0
    assert np.array_equal(probs.cpu().numpy(), want)
