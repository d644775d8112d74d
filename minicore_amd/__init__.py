"""minicore_amd: the minicore clustering hot path on MI355X (gfx950).

Kernels and drivers live in libmcg.so (C ABI: include/mcg.h).  The compiled
extension `_pyminicore` carries the pyminicore entry points (hcluster,
scluster, cmp, kmeanspp, CoresetSampler, CSparseMatrix) with the reference's
signatures; this module adds the convenience wrappers of the reference's
Python package (minicore/__init__.py:23-261).

There is no CPU fallback: importing the wrappers needs the built extension.
"""
import os

import numpy as np

try:
    from ._lib import _share_hip_runtime
    _share_hip_runtime()
    from . import _pyminicore
except ImportError as e:  # pragma: no cover - exercised when the extension is not built
    _pyminicore = None
    _IMPORT_ERROR = e

__all__ = ["hcluster", "scluster", "kmeanspp", "cluster", "cmp", "pcmp", "CSparseMatrix", "CoresetSampler",
           "SparseMatrixWrapper", "SumOpts",
           "cluster_from_centers", "set_sqrl2_precision", "set_num_threads", "get_num_threads"]

_num_threads = [os.cpu_count() or 1]


def set_num_threads(n):
    """pyminicore.set_num_threads (OpenMP threads of the CPU kernels).  The clustering
    here runs on the GPU, so the value is only kept and returned by get_num_threads."""
    n = int(n)
    if n < 1:
        raise ValueError("number of threads must be positive")
    _num_threads[0] = n


def get_num_threads():
    """pyminicore.get_num_threads: the value last given to set_num_threads."""
    return _num_threads[0]


def _ext():
    if _pyminicore is None:
        raise RuntimeError("minicore_amd._pyminicore is not built (python -m minicore_amd.build): "
                           f"{_IMPORT_ERROR}")
    return _pyminicore


def _is_csr(x):
    try:
        import scipy.sparse as sp
        return isinstance(x, sp.csr_matrix) or (hasattr(sp, "csr_array") and isinstance(x, sp.csr_array))
    except ImportError:  # pragma: no cover
        return False


def _as_matrix(matrix):
    """scipy CSR -> CSparseMatrix (minicore/__init__.py:45-46); dense arrays pass through."""
    if _is_csr(matrix):
        m = matrix
        if not m.has_sorted_indices:
            m = m.sorted_indices()
        return _ext().CSparseMatrix(m)
    return matrix


def CSparseMatrix(sparray):
    return _ext().CSparseMatrix(sparray)


def SparseMatrixWrapper(*args, **kwargs):
    """SparseMatrixWrapper(scipy CSR matrix, skip_empty=False, use_float=True): an owned copy of
    the rows (python/smw.h, pysmw.cpp:31-265), accepted by hcluster, scluster, cmp and kmeanspp."""
    return _ext().SparseMatrixWrapper(*args, **kwargs)


def SumOpts(*args, **kwargs):
    """SumOpts(measure, k=10, prior=0., sm='BFL', outlier_fraction=0., max_rounds=100, soft=False):
    the options object of kmeanspp(matrix, opts) (clustering/mtx2cs.h:13-80, pysmw.cpp:293-366)."""
    return _ext().SumOpts(*args, **kwargs)


def CoresetSampler(*args, **kwargs):
    """CoresetSampler() or CoresetSampler(path): the latter reads a stream written by
    CoresetSampler.write (coreset.h:260-264)."""
    return _ext().CoresetSampler(*args, **kwargs)


def set_sqrl2_precision(p):
    """'bf16' (default) or 'fp16' (exact argmin via near-tie re-scoring), 'fp32' or 'exact'."""
    _ext().set_sqrl2_precision(p)


def kmeanspp(matrix, k, *, msr=2, prior=0., seed=0, ntimes=1, lspp=0, expskips=0, n_local_trials=1,
             weights=None):
    """D^2 seeding (minicore/__init__.py:23-55).  Returns (center ids, assignments, costs)."""
    m = _as_matrix(matrix)
    if isinstance(m, np.ndarray):
        return _ext().kmeanspp(m, k=k, msr=msr, prior=prior, seed=seed, ntimes=ntimes, lspp=lspp,
                               weights=weights, expskips=bool(expskips), n_local_trials=n_local_trials)
    return _ext().kmeanspp(m, msr=msr, k=k, prior=prior, seed=seed, ntimes=ntimes, lspp=lspp,
                           expskips=bool(expskips), n_local_trials=n_local_trials, weights=weights)


def cluster_from_centers(*args, **kwargs):
    return _ext().hcluster(*args, **kwargs)


def hcluster(matrix, centers, *, prior=0., msr=2, weights=None, eps=1e-10, maxiter=1000, mbsize=-1,
             ncheckins=-1, with_rep=False, cs=False):
    """Hard clustering from initial centers (minicore/__init__.py:73-163)."""
    return _ext().hcluster(_as_matrix(matrix), centers=centers, prior=prior, msr=msr, weights=weights, eps=eps,
                           maxiter=maxiter, mbsize=mbsize, ncheckins=ncheckins, with_rep=with_rep)


def scluster(matrix, centers, *, prior=0., msr=2, temp=1., maxiter=1000, savepref="", weights=None, mbsize=-1):
    """Soft clustering from initial centers (minicore/__init__.py:167-231)."""
    return _ext().scluster(_as_matrix(matrix), centers, msr=msr, prior=prior, temp=temp, maxiter=maxiter,
                           savepref=savepref, weights=weights)


def cmp(matrix, data, *, msr=2, prior=0., **kw):
    m = _as_matrix(matrix)
    d = _as_matrix(data)
    return _ext().cmp(m, d, msr=msr, prior=prior, **kw)


def pcmp(matrix, *, msr=2, prior=0., use_float=-1):
    """Condensed pairwise costs msr(row i, row j), i < j (python/pycmp.cpp:249-338)."""
    return _ext().pcmp(_as_matrix(matrix), msr=msr, prior=prior, use_float=use_float)


def cluster(data, *, msr, k, prior=0., seed=0, nmkc=0, ntimes=1, lspp=0, use_exponential_skips=False,
            n_local_trials=1, weights=None, mbsize=-1, clustereps=1e-4, temp=-1., cs=False, with_rep=True,
            outpref="mc.cluster.output", maxiter=50):
    """Seeding then hard (temp <= 0) or soft clustering (minicore/__init__.py:235-261)."""
    soft = temp > 0.
    ids, asn, costs = kmeanspp(data, k, msr=msr, prior=prior, seed=seed, ntimes=ntimes, lspp=lspp,
                               expskips=use_exponential_skips, n_local_trials=n_local_trials, weights=weights)
    if soft:
        return scluster(data, centers=ids, msr=msr, prior=prior, weights=weights, temp=temp, maxiter=maxiter,
                        savepref=outpref)
    return hcluster(data, centers=ids, prior=prior, msr=msr, weights=weights, eps=clustereps, maxiter=maxiter,
                    with_rep=with_rep, cs=cs)


from . import constants  # noqa: E402
from .constants import CSR as csr_tuple, KMCRSV  # noqa: E402
