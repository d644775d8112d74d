"""Recursive clustering over the seeding and Lloyd entry points
(minicore/recursive_cluster.py:7-32), and pyminicore.get_counthist
(python/pycounthist.cpp:72-155), which it uses to split the rows by label."""
import numpy as np


def _index_dtype(n):
    """Smallest unsigned type for n rows (pyhelpers.h:220-225)."""
    if n > 0xFFFFFFFF:
        return np.uint64
    if n > 0xFFFF:
        return np.uint32
    if n > 0xFF:
        return np.uint16
    return np.uint8


def _hist1d(a):
    a = np.asarray(a)
    order = np.argsort(a, kind="stable")
    vals, starts = np.unique(a[order], return_index=True)
    dt = _index_dtype(len(a))
    bounds = list(starts[1:]) + [len(a)]
    return {v.item(): order[s:e].astype(dt) for v, s, e in zip(vals, starts, bounds)}


def get_counthist(items):
    """A dict from each distinct value of a 1-d array to the (ascending) indices where
    it occurs, as u8/u16/u32/u64 by the array's length; for a 2-d array, one such
    dict per row (the reference's 2-d form reads row 1 for every row,
    pycounthist.cpp:142; here each row is its own)."""
    a = np.asarray(items)
    if a.dtype.kind not in "fiu":
        raise ValueError(f"Unexpected dtype: {a.dtype}")
    if a.ndim == 1:
        return _hist1d(a)
    if a.ndim != 2:
        raise RuntimeError("Expected 2-d Numpy Array")
    return [_hist1d(r) for r in a]


def hiercluster_kmeanspp(dataset, radix=5, msr="JSD", n_local_trials=2, lspp=0, maxiter=3, optimize_centers=True,
                         prior=1e-5, hierdepth=0, mbsize=-1):
    """D^2 seeding of `radix` centers (n_local_trials candidates per center), then, if
    optimize_centers, `maxiter` Lloyd iterations; every cluster with more than radix
    rows is clustered again the same way.  As in the reference, the Lloyd step runs
    with hcluster's default measure and prior (not `msr` / `prior`), and the costs sit
    under the key "costs:"."""
    from . import hcluster, kmeanspp
    centers, assignments, costs = kmeanspp(dataset, k=radix, msr=msr, n_local_trials=n_local_trials, lspp=lspp,
                                           prior=prior)
    if optimize_centers:
        out = hcluster(dataset, centers=centers, maxiter=maxiter, mbsize=mbsize)
        centers = out["centers"]
        assignments = out["asn"]
        costs = out["costs"]
    base = {"dataset": dataset, "radix": radix, "msr": msr, "centerids": centers, "asn": assignments,
            "costs:": costs, "hierdepth": hierdepth}
    if isinstance(centers, np.ndarray) and centers.ndim == 1:
        base["centers"] = dataset[centers].copy()
    else:
        base["centers"] = centers
    children = {}
    for k, v in get_counthist(assignments).items():
        if len(v) > radix:
            children[k] = hiercluster_kmeanspp(dataset[v].copy(), radix=radix, msr=msr,
                                               n_local_trials=n_local_trials, lspp=lspp, prior=prior,
                                               optimize_centers=optimize_centers, maxiter=maxiter,
                                               hierdepth=hierdepth + 1, mbsize=mbsize)
    base["children"] = children
    return base
