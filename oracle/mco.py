"""ctypes front end of the CPU oracle (oracle/mco.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product package never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libmco.so")
_lib = None

SQRL2, JSD, MKL, REVERSE_MKL = 2, 4, 5, 11
(L1, L2, JSM, HELLINGER, BHATTACHARYYA_METRIC, BHATTACHARYYA_DISTANCE, TVD, LLR, UWLLR, IS, REVERSE_IS,
 COSINE_DISTANCE, PROBABILITY_COSINE_DISTANCE, COSINE_SIMILARITY, PROBABILITY_COSINE_SIMILARITY, SIS, RSIS, SRLRT,
 SRULRT) = 0, 1, 3, 6, 7, 8, 9, 10, 12, 13, 14, 15, 16, 17, 18, 23, 24, 25, 26
POISSON = 64   # I-divergence (poisson_bregman); no code in the reference
MORE_MEASURES = (L1, L2, JSM, HELLINGER, BHATTACHARYYA_METRIC, BHATTACHARYYA_DISTANCE, TVD, LLR, UWLLR, IS,
                 REVERSE_IS, COSINE_DISTANCE, PROBABILITY_COSINE_DISTANCE, COSINE_SIMILARITY,
                 PROBABILITY_COSINE_SIMILARITY, SIS, RSIS, SRLRT, SRULRT)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _setup(_lib)
    return _lib


P = C.c_void_p
U64 = C.c_uint64
D = C.c_double
I = C.c_int


def _setup(L):
    sig = {
        "mco_wyrand_at": (U64, [U64, U64]),
        "mco_mulhi64": (U64, [U64, U64]),
        "mco_fixed_shift": (I, [D, U64]),
        "mco_fixed_weights": (None, [P, U64, P]),
        "mco_fixed_draw": (U64, [P, U64, U64]),
        "mco_msr_dense": (D, [I, P, P, U64, D, D, D]),
        "mco_msr_sparse": (D, [I, P, P, U64, P, P, U64, U64, D, D, D, P]),
        "mco_sqrl2_f32": (C.c_float, [P, P, U64]),
        "mco_shared_zeros": (U64, [P, U64, P, U64, U64]),
        "mco_cost_matrix_dense": (None, [I, P, U64, U64, P, U64, D, P]),
        "mco_cost_matrix_csr": (None, [I, P, P, P, U64, U64, P, U64, D, P]),
        "mco_assign_sqrl2_f32": (None, [P, U64, U64, P, U64, P, P, P, I]),
        "mco_assign_f64": (None, [I, P, P, P, P, U64, U64, P, U64, D, P, P, P]),
        "mco_centroids_dense": (U64, [P, U64, U64, P, P, U64, U64, P]),
        "mco_centroids_csr": (U64, [P, P, P, U64, U64, P, P, U64, P]),
        "mco_lloyd": (I, [I, P, P, P, P, U64, U64, P, U64, D, P, D, U64, P, P, P, P, P, I]),
        "mco_minibatch": (I, [I, P, P, P, P, U64, U64, P, U64, D, P, U64, U64, U64, U64, I, U64, P, P, P, P, P,
                              I]),
        "mco_cs_minibatch": (I, [I, P, P, P, P, U64, U64, P, U64, D, P, U64, U64, U64, U64, U64, P, P, P, P, P,
                                 I]),
        "mco_soft": (I, [I, P, P, P, P, U64, U64, P, U64, D, D, P, D, U64, P, P, P, P, P]),
        "mco_softmax_rows": (None, [P, U64, U64, D, P]),
        "mco_kmeanspp": (U64, [I, P, P, P, P, U64, U64, U64, D, U64, P, P, P, P, I]),
        "mco_kmeanspp_ls": (U64, [I, P, P, P, P, U64, U64, U64, D, U64, P, U64, P, P, P, I]),
        "mco_lloyd_step": (U64, [P, U64, U64, P, U64, P, P, P]),
        "mco_localsearch": (None, [I, P, P, P, P, U64, U64, U64, D, U64, P, U64, U64, P, P, P]),
        "mco_median_centers": (I, [I, P, P, P, P, U64, U64, P, U64, P, P, P, D]),
        "mco_lloyd_median": (I, [I, P, P, P, P, U64, U64, P, U64, D, P, D, U64, P, P, P, P, P]),
        "mco_kmeanspp_local": (U64, [I, P, P, P, P, U64, U64, U64, D, U64, P, U64, U64, P, P, P]),
        "mco_coreset_probs": (I, [I, U64, U64, P, P, P, P]),
        "mco_coreset_sample": (None, [U64, P, P, U64, U64, U64, P, P]),
        "mco_sqrl2_f64": (D, [P, P, U64]),
        "mco_assign_sqrl2_f64": (None, [P, U64, U64, P, U64, P, P]),
        "mco_centroids_dense_f64": (U64, [P, U64, U64, P, P, U64, U64, P]),
        "mco_fixed_total": (D, [P, P, U64]),
        "mco_lloyd_f64": (I, [P, U64, U64, P, U64, P, D, U64, P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class CSR:
    """Canonical CSR triple used by the oracle: float32 data, uint32 indices, uint64 indptr."""

    def __init__(self, data, indices, indptr, shape):
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.indptr = np.ascontiguousarray(indptr, dtype=np.uint64)
        self.shape = (int(shape[0]), int(shape[1]))

    @classmethod
    def from_scipy(cls, m):
        m = m.tocsr()
        m.sort_indices()
        return cls(m.data, m.indices, m.indptr, m.shape)

    def todense(self):
        out = np.zeros(self.shape, dtype=np.float32)
        for i in range(self.shape[0]):
            b, e = int(self.indptr[i]), int(self.indptr[i + 1])
            out[i, self.indices[b:e]] = self.data[b:e]
        return out


def _mat(X):
    """(X_dense_or_None, val, idx, indptr, n, d) for a dense array or CSR."""
    if isinstance(X, CSR):
        return None, X.data, X.indices, X.indptr, X.shape[0], X.shape[1]
    X = _f32(X)
    return X, None, None, None, X.shape[0], X.shape[1]


def wyrand_at(seed, j):
    return lib().mco_wyrand_at(seed, j)


def fixed_weights(w):
    w = _f64(w)
    out = np.empty(len(w), dtype=np.uint64)
    lib().mco_fixed_weights(_p(w), len(w), _p(out))
    return out


def fixed_draw(fw, r):
    fw = np.ascontiguousarray(fw, dtype=np.uint64)
    return lib().mco_fixed_draw(_p(fw), len(fw), r)


def msr_dense(msr, x, c, prior):
    x = _f64(x)
    c = _f64(c)
    return lib().mco_msr_dense(msr, _p(x), _p(c), len(x), prior, float(x.sum()), float(c.sum()))


def msr_sparse(msr, xv, xi, cv, ci, d, prior):
    xv, cv = _f64(xv), _f64(cv)
    xi = np.ascontiguousarray(xi, dtype=np.uint32)
    ci = np.ascontiguousarray(ci, dtype=np.uint32)
    sz = np.zeros(1, dtype=np.uint64)
    v = lib().mco_msr_sparse(msr, _p(xv), _p(xi), len(xv), _p(cv), _p(ci), len(cv), d, prior,
                             float(xv.sum()), float(cv.sum()), _p(sz))
    return v, int(sz[0])


def shared_zeros(a, b, d):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    return lib().mco_shared_zeros(_p(a), len(a), _p(b), len(b), d)


def sqrl2_f32(x, c):
    x, c = _f32(x), _f32(c)
    return lib().mco_sqrl2_f32(_p(x), _p(c), len(x))


def cost_matrix(msr, X, C, prior):
    Xd, val, idx, ptr, n, d = _mat(X)
    C = _f32(C)
    k = C.shape[0]
    out = np.empty((n, k), dtype=np.float64)
    if val is None:
        lib().mco_cost_matrix_dense(msr, _p(Xd), n, d, _p(C), k, prior, _p(out))
    else:
        lib().mco_cost_matrix_csr(msr, _p(val), _p(idx), _p(ptr), n, d, _p(C), k, prior, _p(out))
    return out


def assign_sqrl2_f32(X, C, nthreads=0):
    X, C = _f32(X), _f32(C)
    n, d = X.shape
    k = C.shape[0]
    asn = np.empty(n, dtype=np.uint32)
    cost = np.empty(n, dtype=np.float32)
    gap = np.empty(n, dtype=np.float64)
    lib().mco_assign_sqrl2_f32(_p(X), n, d, _p(C), k, _p(asn), _p(cost), _p(gap), nthreads)
    return asn, cost, gap


def assign_f64(msr, X, C, prior):
    Xd, val, idx, ptr, n, d = _mat(X)
    C = _f32(C)
    k = C.shape[0]
    asn = np.empty(n, dtype=np.uint32)
    cost = np.empty(n, dtype=np.float64)
    gap = np.empty(n, dtype=np.float64)
    lib().mco_assign_f64(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, _p(C), k, prior,
                         _p(asn), _p(cost), _p(gap))
    return asn, cost, gap


def centroids(X, asn, k, C_prev, w=None, chunk=256):
    Xd, val, idx, ptr, n, d = _mat(X)
    asn = np.ascontiguousarray(asn, dtype=np.uint32)
    C = _f32(C_prev).copy()
    w = _f64(w)
    if val is None:
        e = lib().mco_centroids_dense(_p(Xd), n, d, _p(asn), _p(w), k, chunk, _p(C))
    else:
        e = lib().mco_centroids_csr(_p(val), _p(idx), _p(ptr), n, d, _p(asn), _p(w), k, _p(C))
    return C, int(e)


def minibatch(msr, X, C0, prior=0., w=None, mbsize=100, maxiter=20, freq=5, reseed_after=0, with_rep=False,
              seed=0, nthreads=0):
    """perform_hard_minibatch_clustering (solve.h:377-578) as restated in mco.c."""
    Xd, val, idx, ptr, n, d = _mat(X)
    C = _f32(C0).copy()
    k = C.shape[0]
    w = _f64(w)
    asn = np.empty(n, dtype=np.uint32)
    cost = np.empty(n, dtype=np.float64)
    ic, fc = np.zeros(1), np.zeros(1)
    it = np.zeros(1, dtype=np.uint64)
    rc = lib().mco_minibatch(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, _p(C), k, prior, _p(w), mbsize, maxiter,
                             freq, reseed_after, int(bool(with_rep)), seed, _p(asn), _p(cost), _p(ic), _p(fc), _p(it),
                             nthreads)
    if rc != 0:
        raise ValueError("oracle minibatch: invalid sizes")
    return dict(centers=C, asn=asn, costs=cost, initcost=float(ic[0]), finalcost=float(fc[0]), numiter=int(it[0]))


def cs_minibatch(msr, X, C0, prior=0., w=None, mbsize=100, maxiter=20, freq=5, reseed_after=0, seed=0,
                 nthreads=0):
    """hmb_coreset_clustering (solve.h:598-789) as restated in mco.c."""
    Xd, val, idx, ptr, n, d = _mat(X)
    C = _f32(C0).copy()
    k = C.shape[0]
    w = _f64(w)
    asn = np.empty(n, dtype=np.uint32)
    cost = np.empty(n, dtype=np.float64)
    ic, fc = np.zeros(1), np.zeros(1)
    it = np.zeros(1, dtype=np.uint64)
    rc = lib().mco_cs_minibatch(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, _p(C), k, prior, _p(w), mbsize,
                                maxiter, freq, reseed_after, seed, _p(asn), _p(cost), _p(ic), _p(fc), _p(it), nthreads)
    if rc != 0:
        raise ValueError(f"oracle coreset minibatch failed ({rc})")
    return dict(centers=C, asn=asn, costs=cost, initcost=float(ic[0]), finalcost=float(fc[0]), numiter=int(it[0]))


def lloyd_step(X, C, w=None):
    """One lloyd_iteration pass (dense squared L2): returns (new centers, asn, cost of
    the assignment against C, number of reseed draws)."""
    X = _f32(X)
    n, d = X.shape
    C = np.array(C, dtype=np.float32, order="C")
    w = _f64(w)
    asn = np.empty(n, dtype=np.uint32)
    cost = np.zeros(1)
    draws = lib().mco_lloyd_step(_p(X), n, d, _p(C), C.shape[0], _p(w), _p(asn), _p(cost))
    return C, asn, float(cost[0]), int(draws)


def lloyd(msr, X, C0, prior=0., w=None, eps=1e-10, maxiter=100, nthreads=0):
    Xd, val, idx, ptr, n, d = _mat(X)
    C = _f32(C0).copy()
    k = C.shape[0]
    w = _f64(w)
    asn = np.empty(n, dtype=np.uint32)
    cost = np.empty(n, dtype=np.float64)
    ic, fc = np.zeros(1), np.zeros(1)
    it = np.zeros(1, dtype=np.uint64)
    rc = lib().mco_lloyd(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, _p(C), k, prior, _p(w), eps,
                         maxiter, _p(asn), _p(cost), _p(ic), _p(fc), _p(it), nthreads)
    if rc != 0:
        raise RuntimeError("oracle lloyd failed (NaN center)")
    return dict(centers=C, asn=asn, costs=cost, initcost=float(ic[0]), finalcost=float(fc[0]),
                numiter=int(it[0]))


def _f64c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def assign_sqrl2_f64(X, C):
    """float64 rows: FP64 sequential-FMA argmin (lowest index on ties)."""
    X, C = _f64c(X), _f64c(C)
    n, d = X.shape
    asn = np.empty(n, dtype=np.uint32)
    cost = np.empty(n, dtype=np.float64)
    lib().mco_assign_sqrl2_f64(_p(X), n, d, _p(C), C.shape[0], _p(asn), _p(cost))
    return asn, cost


def centroids_f64(X, asn, k, C_prev, w=None, chunk=256):
    X = _f64c(X)
    n, d = X.shape
    C = _f64c(C_prev).copy()
    w = _f64(w)
    e = lib().mco_centroids_dense_f64(_p(X), n, d, _p(np.ascontiguousarray(asn, dtype=np.uint32)), _p(w), k,
                                      chunk, _p(C))
    return C, int(e)


def fixed_total(cost, w=None):
    cost = _f64c(cost)
    return lib().mco_fixed_total(_p(cost), _p(_f64(w)), cost.shape[0])


def lloyd_f64(X, C0, w=None, eps=1e-10, maxiter=100):
    """perform_hard_clustering on float64 rows, dense SQRL2, everything in FP64."""
    X = _f64c(X)
    n, d = X.shape
    C = _f64c(C0).copy()
    k = C.shape[0]
    w = _f64(w)
    asn = np.empty(n, dtype=np.uint32)
    cost = np.empty(n, dtype=np.float64)
    ic, fc = np.zeros(1), np.zeros(1)
    it = np.zeros(1, dtype=np.uint64)
    rc = lib().mco_lloyd_f64(_p(X), n, d, _p(C), k, _p(w), eps, maxiter, _p(asn), _p(cost), _p(ic), _p(fc),
                             _p(it))
    if rc != 0:
        raise RuntimeError("oracle lloyd_f64 failed (NaN center)")
    return dict(centers=C, asn=asn, costs=cost, initcost=float(ic[0]), finalcost=float(fc[0]),
                numiter=int(it[0]))


def median_centers(msr, X, asn, cost, C0, w=None):
    """One median-policy M-step (L1, TVD, L2) from (asn, cost); returns
    (centers, asn, cost) after the empty-cluster restarts."""
    Xd, val, idx, ptr, n, d = _mat(X)
    C = _f32(C0).copy()
    a = np.ascontiguousarray(asn, dtype=np.uint32).copy()
    c = np.ascontiguousarray(cost, dtype=np.float64).copy()
    w = _f64(w)
    rc = lib().mco_median_centers(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, _p(w), C.shape[0], _p(a), _p(c),
                                  _p(C), 0.)
    if rc != 0:
        raise ValueError("unsupported median case")
    return C, a, c


def lloyd_median(msr, X, C0, prior=0., w=None, eps=1e-10, maxiter=100):
    """perform_hard_clustering for the median policies (L1, TVD, L2)."""
    Xd, val, idx, ptr, n, d = _mat(X)
    C = _f32(C0).copy()
    k = C.shape[0]
    w = _f64(w)
    asn = np.empty(n, dtype=np.uint32)
    cost = np.empty(n, dtype=np.float64)
    ic, fc = np.zeros(1), np.zeros(1)
    it = np.zeros(1, dtype=np.uint64)
    rc = lib().mco_lloyd_median(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, _p(C), k, prior, _p(w), eps,
                                maxiter, _p(asn), _p(cost), _p(ic), _p(fc), _p(it))
    if rc != 0:
        raise ValueError("unsupported median case")
    return dict(centers=C, asn=asn, costs=cost, initcost=float(ic[0]), finalcost=float(fc[0]),
                numiter=int(it[0]))


def soft(msr, X, C0, prior=0., temp=1., w=None, eps=1e-8, maxiter=1000):
    Xd, val, idx, ptr, n, d = _mat(X)
    C = _f32(C0).copy()
    k = C.shape[0]
    w = _f64(w)
    costs = np.empty((n, k), dtype=np.float32)
    resp = np.empty((n, k), dtype=np.float32)
    ic, fc = np.zeros(1), np.zeros(1)
    it = np.zeros(1, dtype=np.uint64)
    rc = lib().mco_soft(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, _p(C), k, prior, temp, _p(w), eps,
                        maxiter, _p(costs), _p(resp), _p(ic), _p(fc), _p(it))
    if rc == -1:
        raise NotImplementedError("weighted L1 median for CSR rows (csc.h:834)")
    if rc:
        raise FloatingPointError("geometric median of a center whose weights are all zero")
    return dict(centers=C, costs=costs, asn=resp, initcost=float(ic[0]), finalcost=float(fc[0]),
                numiter=int(it[0]))


def softmax_rows(costs, temp):
    costs = _f32(costs)
    n, k = costs.shape
    out = np.empty_like(costs)
    lib().mco_softmax_rows(_p(costs), n, k, temp, _p(out))
    return out


def kmeanspp(msr, X, k, prior=0., seed=0, w=None, nthreads=0):
    Xd, val, idx, ptr, n, d = _mat(X)
    w = _f64(w)
    centers = np.zeros(k, dtype=np.uint32)
    asn = np.zeros(n, dtype=np.uint32)
    dist = np.zeros(n, dtype=np.float32)
    kk = lib().mco_kmeanspp(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, k, prior, seed, _p(w),
                            _p(centers), _p(asn), _p(dist), nthreads)
    return centers[:kk], asn, dist


def kmeanspp_ls(msr, X, k, rounds, prior=0., seed=0, w=None, nthreads=0):
    Xd, val, idx, ptr, n, d = _mat(X)
    w = _f64(w)
    centers = np.zeros(k, dtype=np.uint32)
    asn = np.zeros(n, dtype=np.uint32)
    dist = np.zeros(n, dtype=np.float32)
    kk = lib().mco_kmeanspp_ls(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, k, prior, seed, _p(w), rounds,
                               _p(centers), _p(asn), _p(dist), nthreads)
    return centers[:kk], asn, dist


def localsearch(msr, X, centers, rounds, prior=0., seed=0, w=None, j0=0):
    """local search++ rounds from the given center rows (draws j0, j0+1, ... of the
    stream of seed); returns (centers, asn, dist)."""
    Xd, val, idx, ptr, n, d = _mat(X)
    w = _f64(w)
    centers = np.array(centers, dtype=np.uint32)
    asn = np.zeros(n, dtype=np.uint32)
    dist = np.zeros(n, dtype=np.float32)
    lib().mco_localsearch(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, len(centers), prior, seed, _p(w), rounds,
                          j0, _p(centers), _p(asn), _p(dist))
    return centers, asn, dist


def kmeanspp_local(msr, X, k, trials, prior=0., seed=0, w=None, lspp=0):
    """k-means++ with n_local_trials = trials > 1 (kmeans.h:79-177), then lspp rounds of
    local search++."""
    Xd, val, idx, ptr, n, d = _mat(X)
    w = _f64(w)
    centers = np.empty(k, dtype=np.uint32)
    asn = np.empty(n, dtype=np.uint32)
    dist = np.empty(n, dtype=np.float32)
    kk = lib().mco_kmeanspp_local(msr, _p(Xd), _p(val), _p(idx), _p(ptr), n, d, k, prior, seed, _p(w), trials,
                                  lspp, _p(centers), _p(asn), _p(dist))
    return centers[:kk], asn, dist


def coreset_probs(sens, ncenters, costs, asn, w=None):
    costs = _f64(costs)
    asn = np.ascontiguousarray(asn, dtype=np.uint32)
    w = _f64(w)
    n = len(costs)
    probs = np.empty(n, dtype=np.float32)
    rc = lib().mco_coreset_probs(sens, n, ncenters, _p(costs), _p(asn), _p(w), _p(probs))
    if rc:
        raise ValueError("invalid sensitivity method")
    return probs


def coreset_sample(probs, m, seed, w=None, first_draw=0):
    probs = _f32(probs)
    w = _f64(w)
    idx = np.empty(m, dtype=np.uint64)
    wt = np.empty(m, dtype=np.float32)
    lib().mco_coreset_sample(len(probs), _p(probs), _p(w), m, seed, first_draw, _p(idx), _p(wt))
    return wt, idx
