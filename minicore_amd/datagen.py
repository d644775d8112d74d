"""Synthetic inputs of the benchmark workloads (numpy default_rng, seeded).

  gaussian_mixture   dense rows: component means U[lo, hi]^d, per-component
                     scales, equal weights (the dense k-means workloads).
  poisson_csr        sparse count rows: nnz_i ~ Poisson(density*d) distinct
                     uniform columns, values Poisson(lambda_j) (zeros -> 1),
                     lambda_j ~ LogNormal(0, sigma), optionally row-normalised.
"""
import numpy as np


def gaussian_mixture(n, d, ncomp, seed, lo=-10., hi=10., scale=(1., 1.), dtype=np.float32):
    rng = np.random.default_rng(seed)
    means = rng.uniform(lo, hi, size=(ncomp, d))
    scales = rng.uniform(scale[0], scale[1], size=ncomp) if scale[0] != scale[1] else np.full(ncomp, scale[0])
    labels = rng.integers(0, ncomp, size=n)
    X = means[labels] + rng.standard_normal((n, d)) * scales[labels, None]
    return X.astype(dtype), labels.astype(np.uint32), means.astype(dtype)


def poisson_csr(n, d, density, seed, sigma=1.5, normalize=False):
    """Returns (data f32, indices u32, indptr u64) with sorted column indices."""
    rng = np.random.default_rng(seed)
    lam = rng.lognormal(0., sigma, size=d)
    nnz_row = np.clip(rng.poisson(density * d, size=n), 1, d)
    indptr = np.zeros(n + 1, dtype=np.uint64)
    indptr[1:] = np.cumsum(nnz_row)
    nnz = int(indptr[-1])
    indices = np.empty(nnz, dtype=np.uint32)
    for i in range(n):
        b, e = int(indptr[i]), int(indptr[i + 1])
        cols = rng.choice(d, size=e - b, replace=False)
        cols.sort()
        indices[b:e] = cols
    vals = rng.poisson(lam[indices]).astype(np.float64)
    vals[vals == 0] = 1.
    if normalize:
        rs = np.add.reduceat(vals, indptr[:-1].astype(np.int64))
        vals = vals / np.repeat(rs, nnz_row)
    return vals.astype(np.float32), indices, indptr


def poisson_csr_fast(n, d, density, seed, sigma=1.5, normalize=False):
    """Vectorised variant of poisson_csr for large n: columns are drawn uniformly
    with replacement and duplicates within a row dropped (so a row holds slightly
    fewer than its Poisson draw of entries: about k^2 / 2d fewer)."""
    rng = np.random.default_rng(seed)
    lam = rng.lognormal(0., sigma, size=d)
    draw = np.clip(rng.poisson(density * d, size=n), 1, d)
    rows = np.repeat(np.arange(n, dtype=np.int64), draw)
    cols = rng.integers(0, d, size=len(rows), dtype=np.int64)
    key = np.unique(rows * d + cols)
    rows, cols = key // d, (key % d).astype(np.uint32)
    nnz_row = np.bincount(rows, minlength=n)
    indptr = np.zeros(n + 1, dtype=np.uint64)
    indptr[1:] = np.cumsum(nnz_row)
    vals = rng.poisson(lam[cols]).astype(np.float64)
    vals[vals == 0] = 1.
    if normalize:
        rs = np.add.reduceat(vals, indptr[:-1].astype(np.int64))
        vals = vals / np.repeat(rs, nnz_row)
    return vals.astype(np.float32), cols, indptr


def counts_csr_gaps(n, d, density, seed, mu_sigma=2.0, normalize=False):
    """Sparse counts at the given density with strictly increasing columns drawn as
    geometric gaps (no sort); values Poisson(mu_j), mu_j ~ LogNormal(0, mu_sigma),
    zero draws -> 1.  For large shapes (10^6 x 32768 at 1%)."""
    rng = np.random.default_rng(seed)
    mu = rng.lognormal(0., mu_sigma, size=d)
    per = int(density * d * 1.5) + 16
    gaps = rng.geometric(density, size=(n, per)).astype(np.int64)
    cols = np.cumsum(gaps, axis=1) - 1 + rng.integers(0, 2, size=(n, 1))
    keep = cols < d
    nnz_row = keep.sum(axis=1)
    nnz_row = np.maximum(nnz_row, 1)
    cols[:, 0] = np.minimum(cols[:, 0], d - 1)
    keep[:, 0] = True
    indices = cols[keep].astype(np.uint32)
    indptr = np.zeros(n + 1, dtype=np.uint64)
    indptr[1:] = np.cumsum(keep.sum(axis=1))
    vals = rng.poisson(mu[indices]).astype(np.float32)
    vals[vals == 0] = 1.
    if normalize:
        rs = np.add.reduceat(vals.astype(np.float64), indptr[:-1].astype(np.int64))
        vals = (vals / np.repeat(rs, np.diff(indptr).astype(np.int64))).astype(np.float32)
    return vals, indices, indptr


def gaussian_mixture_device(n, d, ncomp, seed, lo=-10., hi=10., scale=(0.5, 2.), device="cuda"):
    """Same construction on the device with torch's generator (for the 10^6-row workload)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    means = (torch.rand((ncomp, d), generator=g, device=device, dtype=torch.float32) * (hi - lo) + lo)
    scales = torch.rand((ncomp,), generator=g, device=device, dtype=torch.float32) * (scale[1] - scale[0]) + scale[0]
    labels = torch.randint(0, ncomp, (n,), generator=g, device=device)
    X = torch.randn((n, d), generator=g, device=device, dtype=torch.float32)
    X.mul_(scales[labels, None]).add_(means[labels])
    return X, labels, means
