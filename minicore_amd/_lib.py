"""ctypes binding of the libmcg C ABI (include/mcg.h).

The product path: every call goes to the HIP library.  There is no CPU
fallback; a missing or unloadable libmcg.so raises immediately.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmcg.so")

MCG_F32, MCG_F64, MCG_U8, MCG_U16, MCG_U32, MCG_U64, MCG_I32, MCG_I64 = range(8)
SQRL2, JSD, MKL, REVERSE_MKL = 2, 4, 5, 11
PREC_FP32, PREC_BF16, PREC_EXACT, PREC_FP16 = 0, 1, 2, 3
SENS_BFL, SENS_FL, SENS_VX, SENS_LBK = 0, 1, 3, 4

EXPORTS = [
    "mcg_last_error", "mcg_device_count", "mcg_ctx_create", "mcg_ctx_destroy", "mcg_ctx_synchronize",
    "mcg_ctx_stream", "mcg_data_dense", "mcg_data_csr", "mcg_data_free", "mcg_data_shape",
    "mcg_cost_matrix", "mcg_assign_hard", "mcg_centroids_hard", "mcg_lloyd", "mcg_soft_em",
    "mcg_lloyd_f64", "mcg_assign_f64", "mcg_kmeanspp", "mcg_coreset_probs", "mcg_coreset_sample", "mcg_lloyd_step", "mcg_timing_enable",
    "mcg_timing_reset", "mcg_timing_read", "mcg_refine_counts", "mcg_data_rows", "mcg_server_create",
    "mcg_server_destroy", "mcg_submit", "mcg_poll", "mcg_wait", "mcg_cancel", "mcg_release", "mcg_tuning_enable",
    "mcg_graphs_enable", "mcg_minibatch", "mcg_cs_minibatch", "mcg_kmeanspp_ls", "mcg_kmeanspp_local", "mcg_sampler_write",
    "mcg_sampler_read", "mcg_index_coreset_write", "mcg_index_coreset_read", "mcg_pairwise",
    "mcg_coreset_probs_noalloc",
]

REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = range(5)


class Request(C.Structure):
    """struct mcg_request (include/mcg.h)."""
    _fields_ = [
        ("data", C.c_void_p), ("data_dtype", C.c_int),
        ("indices", C.c_void_p), ("idx_dtype", C.c_int),
        ("indptr", C.c_void_p), ("ptr_dtype", C.c_int),
        ("n", C.c_uint64), ("d", C.c_uint64), ("nnz", C.c_uint64), ("ld", C.c_uint64),
        ("weights", C.c_void_p),
        ("msr", C.c_int), ("prior", C.c_double),
        ("k", C.c_uint64), ("seed", C.c_uint64),
        ("sens", C.c_int), ("coreset_frac", C.c_double), ("eps", C.c_double), ("maxiter", C.c_uint64),
        ("precision", C.c_int), ("fail_at_stage", C.c_int),
        ("centers", C.c_void_p), ("asn", C.c_void_p), ("costs", C.c_void_p),
    ]


class Result(C.Structure):
    """struct mcg_result (include/mcg.h)."""
    _fields_ = [
        ("state", C.c_int), ("status", C.c_int),
        ("cost", C.c_double), ("coreset_initcost", C.c_double), ("coreset_finalcost", C.c_double),
        ("coreset_m", C.c_uint64), ("numiter", C.c_uint64),
        ("queue_ms", C.c_double), ("run_ms", C.c_double),
        ("stage_ms", C.c_double * 5),
        ("error", C.c_char * 256),
    ]

_lib = None


def _share_hip_runtime():
    """torch ships its own libamdhip64 (same SONAME as /opt/rocm's).  If libmcg
    loaded first, torch would later load a second HIP runtime into the process,
    which cannot see the GPU.  Importing torch first makes libmcg bind to the
    runtime already loaded."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


class MCGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmcg error {code}: {msg}")
        self.code = code


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: build it with `python -m minicore_amd.build`")
    _share_hip_runtime()
    L = C.CDLL(LIBPATH)
    P, U64, D, I = C.c_void_p, C.c_uint64, C.c_double, C.c_int
    sig = {
        "mcg_last_error": (C.c_char_p, []),
        "mcg_device_count": (I, [P]),
        "mcg_ctx_create": (I, [I, P]),
        "mcg_ctx_destroy": (I, [P]),
        "mcg_ctx_synchronize": (I, [P]),
        "mcg_ctx_stream": (I, [P, P]),
        "mcg_data_dense": (I, [P, P, I, U64, U64, U64, P]),
        "mcg_data_csr": (I, [P, P, I, P, I, P, I, U64, U64, U64, P]),
        "mcg_data_free": (I, [P]),
        "mcg_data_shape": (I, [P, P, P, P]),
        "mcg_cost_matrix": (I, [P, I, D, P, U64, P]),
        "mcg_pairwise": (I, [P, I, D, P]),
        "mcg_assign_hard": (I, [P, I, D, P, U64, I, P, P]),
        "mcg_centroids_hard": (I, [P, P, P, U64, P, P]),
        "mcg_lloyd": (I, [P, I, D, P, U64, P, D, U64, I, P, P, P, P, P]),
        "mcg_lloyd_f64": (I, [P, P, U64, P, D, U64, P, P, P, P, P]),
        "mcg_assign_f64": (I, [P, P, U64, P, P]),
        "mcg_soft_em": (I, [P, I, D, D, P, U64, P, D, U64, P, P, P, P, P]),
        "mcg_kmeanspp": (I, [P, I, D, U64, U64, P, P, P, P]),
        "mcg_coreset_probs": (I, [P, I, U64, U64, P, P, P, P]),
        "mcg_coreset_probs_noalloc": (I, [P, I, U64, U64, P, P, P, P]),
        "mcg_coreset_sample": (I, [P, U64, P, P, U64, U64, U64, P, P]),
        "mcg_lloyd_step": (I, [P, I, D, P, U64, P, I, P, P]),
        "mcg_timing_enable": (I, [P, I]),
        "mcg_tuning_enable": (I, [P, I]),
        "mcg_graphs_enable": (I, [P, I]),
        "mcg_minibatch": (I, [P, I, D, P, U64, P, U64, U64, U64, U64, I, U64, I, P, P, P, P, P]),
        "mcg_cs_minibatch": (I, [P, I, D, P, U64, P, U64, U64, U64, U64, U64, I, P, P, P, P, P]),
        "mcg_kmeanspp_ls": (I, [P, I, D, U64, U64, P, U64, P, P, P]),
        "mcg_kmeanspp_local": (I, [P, I, D, U64, U64, P, U64, U64, P, P, P]),
        "mcg_sampler_write": (I, [C.c_char_p, U64, U64, P, P]),
        "mcg_sampler_read": (I, [C.c_char_p, P, P, P, P, P, U64]),
        "mcg_index_coreset_write": (I, [C.c_char_p, U64, P, I, P, I]),
        "mcg_index_coreset_read": (I, [C.c_char_p, P, P, I, P, I, U64]),
        "mcg_timing_reset": (I, [P]),
        "mcg_timing_read": (I, [P, I, P, P]),
        "mcg_refine_counts": (I, [P, P, P]),
        "mcg_data_rows": (I, [P, P, U64, P]),
        "mcg_server_create": (I, [I, I, U64, P]),
        "mcg_server_destroy": (I, [P]),
        "mcg_submit": (I, [P, C.POINTER(Request), P]),
        "mcg_poll": (I, [P, U64, C.POINTER(Result)]),
        "mcg_wait": (I, [P, U64, D, C.POINTER(Result)]),
        "mcg_cancel": (I, [P, U64]),
        "mcg_release": (I, [P, U64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = load().mcg_last_error()
        raise MCGError(rc, msg.decode() if msg else "")


def ptr(a):
    """Address of a numpy array or a torch tensor (device pointers pass straight through)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    raise TypeError(f"unsupported buffer type {type(a)}")


def device_count():
    n = C.c_int(0)
    check(load().mcg_device_count(C.byref(n)))
    return n.value


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(load().mcg_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device
        self._datasets = weakref.WeakSet()

    def close(self):
        """Frees the datasets still open on this context, then the context (the C ABI
        requires datasets to go first)."""
        if self.h:
            for ds in list(self._datasets):
                ds.close()
            load().mcg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(load().mcg_ctx_synchronize(self.h))

    def tuning(self, on=True):
        """Timing-based choice between bit-equivalent E-step kernels (on by default)."""
        check(load().mcg_tuning_enable(self.h, 1 if on else 0))

    def graphs(self, on=True):
        """Graph replay of the Lloyd iterations (off by default; results are the same).
        While a capture is open, synchronous legacy-stream HIP calls from other threads
        of the process fail, so enable it only where no other thread makes them."""
        check(load().mcg_graphs_enable(self.h, 1 if on else 0))

    def timing(self, on=True):
        """Switch kernel timing on or off and clear the accumulated times."""
        check(load().mcg_timing_enable(self.h, 1 if on else 0))
        check(load().mcg_timing_reset(self.h))

    def timing_sample(self, on):
        """Switch kernel timing on or off for the next calls, keeping the totals."""
        check(load().mcg_timing_enable(self.h, 1 if on else 0))

    def refine_counts(self):
        a, b = C.c_uint64(), C.c_uint64()
        check(load().mcg_refine_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def timing_read(self, which):
        ms, cnt = C.c_double(), C.c_uint64()
        check(load().mcg_timing_read(self.h, which, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def _dtype_code(dt):
    if str(dt).startswith("torch."):
        dt = str(dt)[6:]
    dt = np.dtype(dt)
    table = {np.float32: MCG_F32, np.float64: MCG_F64, np.uint8: MCG_U8, np.uint16: MCG_U16,
             np.uint32: MCG_U32, np.uint64: MCG_U64, np.int32: MCG_I32, np.int64: MCG_I64}
    for k, v in table.items():
        if dt == np.dtype(k):
            return v
    raise TypeError(f"unsupported dtype {dt}")


class Dataset:
    """Rows resident on the device (dense float32 or canonical CSR)."""

    def __init__(self, ctx, h, keep=()):
        self.ctx = ctx
        self.h = h
        self._keep = keep
        ctx._datasets.add(self)
        n, d, nnz = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(load().mcg_data_shape(h, C.byref(n), C.byref(d), C.byref(nnz)))
        self.n, self.d, self.nnz = n.value, d.value, nnz.value

    @classmethod
    def dense(cls, X, ctx=None):
        ctx = ctx or default_context()
        h = C.c_void_p()
        if isinstance(X, np.ndarray):
            if X.dtype not in (np.float32, np.float64):
                X = X.astype(np.float32)
            X = np.ascontiguousarray(X)
            n, d = X.shape
            check(load().mcg_data_dense(ctx.h, ptr(X), _dtype_code(X.dtype), n, d, d, C.byref(h)))
            return cls(ctx, h)
        # torch tensor on the device: borrowed
        n, d = X.shape
        code = MCG_F32 if str(X.dtype) == "torch.float32" else MCG_F64
        check(load().mcg_data_dense(ctx.h, ptr(X), code, n, d, X.stride(0), C.byref(h)))
        return cls(ctx, h, keep=(X,))

    @classmethod
    def csr(cls, data, indices, indptr, shape, ctx=None):
        """CSR rows from numpy arrays or contiguous torch tensors (host or device), any
        accepted dtype; converted and checked on the device."""
        ctx = ctx or default_context()
        h = C.c_void_p()

        def arr(a):
            if hasattr(a, "data_ptr"):
                if not a.is_contiguous():
                    raise ValueError("CSR tensors must be contiguous")
                return a
            return np.ascontiguousarray(a)
        data, indices, indptr = arr(data), arr(indices), arr(indptr)
        n, d = int(shape[0]), int(shape[1])
        check(load().mcg_data_csr(ctx.h, ptr(data), _dtype_code(data.dtype), ptr(indices),
                                  _dtype_code(indices.dtype), ptr(indptr), _dtype_code(indptr.dtype),
                                  n, d, len(data), C.byref(h)))
        return cls(ctx, h)

    def close(self):
        if self.h:
            load().mcg_data_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- operations (host numpy in/out) ----
    def cost_matrix(self, msr, centers, prior=0.):
        centers = np.ascontiguousarray(centers, dtype=np.float32)
        out = np.empty((self.n, centers.shape[0]), dtype=np.float32)
        check(load().mcg_cost_matrix(self.h, msr, prior, ptr(centers), centers.shape[0], ptr(out)))
        return out

    def pairwise(self, msr, prior=0.):
        """Condensed upper triangle of msr(row i, row j), i < j (pcmp)."""
        m = self.n * (self.n - 1) // 2 if self.n > 1 else 0
        out = np.empty(max(m, 1), dtype=np.float32)
        if m:
            check(load().mcg_pairwise(self.h, msr, prior, ptr(out)))
        return out[:m]

    def assign(self, msr, centers, prior=0., precision=PREC_BF16, with_costs=True):
        centers = np.ascontiguousarray(centers, dtype=np.float32)
        asn = np.empty(self.n, dtype=np.uint32)
        costs = np.empty(self.n, dtype=np.float32) if with_costs else None
        check(load().mcg_assign_hard(self.h, msr, prior, ptr(centers), centers.shape[0], precision,
                                     ptr(asn), ptr(costs)))
        return asn, costs

    def centroids(self, asn, k, centers_prev, weights=None):
        asn = np.ascontiguousarray(asn, dtype=np.uint32)
        C_ = np.ascontiguousarray(centers_prev, dtype=np.float32).copy()
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        ne = C.c_uint64()
        check(load().mcg_centroids_hard(self.h, ptr(asn), ptr(w), k, ptr(C_), C.byref(ne)))
        return C_, ne.value

    def lloyd(self, msr, centers, prior=0., weights=None, eps=1e-10, maxiter=100, precision=PREC_BF16):
        Cc = np.ascontiguousarray(centers, dtype=np.float32).copy()
        k = Cc.shape[0]
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        asn = np.empty(self.n, dtype=np.uint32)
        costs = np.empty(self.n, dtype=np.float32)
        ic, fc, it = C.c_double(), C.c_double(), C.c_uint64()
        check(load().mcg_lloyd(self.h, msr, prior, ptr(Cc), k, ptr(w), eps, maxiter, precision, ptr(asn),
                               ptr(costs), C.byref(ic), C.byref(fc), C.byref(it)))
        return dict(centers=Cc, asn=asn, costs=costs, initcost=ic.value, finalcost=fc.value, numiter=it.value)

    def lloyd_f64(self, centers, weights=None, eps=1e-10, maxiter=100):
        """Dense SQRL2 hard clustering of rows given as float64, every value in FP64."""
        Cc = np.ascontiguousarray(centers, dtype=np.float64).copy()
        k = Cc.shape[0]
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        asn = np.empty(self.n, dtype=np.uint32)
        costs = np.empty(self.n, dtype=np.float64)
        ic, fc, it = C.c_double(), C.c_double(), C.c_uint64()
        check(load().mcg_lloyd_f64(self.h, ptr(Cc), k, ptr(w), eps, maxiter, ptr(asn), ptr(costs), C.byref(ic),
                                   C.byref(fc), C.byref(it)))
        return dict(centers=Cc, asn=asn, costs=costs, initcost=ic.value, finalcost=fc.value, numiter=it.value)

    def assign_f64(self, centers):
        """SQRL2 assignment of rows given as float64, in FP64."""
        Cc = np.ascontiguousarray(centers, dtype=np.float64)
        asn = np.empty(self.n, dtype=np.uint32)
        costs = np.empty(self.n, dtype=np.float64)
        check(load().mcg_assign_f64(self.h, ptr(Cc), Cc.shape[0], ptr(asn), ptr(costs)))
        return asn, costs

    def minibatch(self, msr, centers, prior=0., weights=None, mbsize=1000, maxiter=100, freq=10, reseed_after=1,
                  with_replacement=False, seed=0, precision=PREC_BF16):
        """perform_hard_minibatch_clustering: best centers, last check-in asn/costs."""
        Cc = np.ascontiguousarray(centers, dtype=np.float32).copy()
        k = Cc.shape[0]
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        asn = np.empty(self.n, dtype=np.uint32)
        costs = np.empty(self.n, dtype=np.float32)
        ic, fc, it = C.c_double(), C.c_double(), C.c_uint64()
        check(load().mcg_minibatch(self.h, msr, prior, ptr(Cc), k, ptr(w), mbsize, maxiter, freq, reseed_after,
                                   1 if with_replacement else 0, seed, precision, ptr(asn), ptr(costs),
                                   C.byref(ic), C.byref(fc), C.byref(it)))
        return dict(centers=Cc, asn=asn, costs=costs, initcost=ic.value, finalcost=fc.value, numiter=it.value)

    def cs_minibatch(self, msr, centers, prior=0., weights=None, mbsize=1000, maxiter=100, freq=10, reseed_after=1,
                     seed=0, precision=PREC_BF16):
        """hmb_coreset_clustering: best centers, last check-in asn/costs."""
        Cc = np.ascontiguousarray(centers, dtype=np.float32).copy()
        k = Cc.shape[0]
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        asn = np.empty(self.n, dtype=np.uint32)
        costs = np.empty(self.n, dtype=np.float32)
        ic, fc, it = C.c_double(), C.c_double(), C.c_uint64()
        check(load().mcg_cs_minibatch(self.h, msr, prior, ptr(Cc), k, ptr(w), mbsize, maxiter, freq, reseed_after,
                                      seed, precision, ptr(asn), ptr(costs), C.byref(ic), C.byref(fc), C.byref(it)))
        return dict(centers=Cc, asn=asn, costs=costs, initcost=ic.value, finalcost=fc.value, numiter=it.value)

    def lloyd_step(self, msr, centers_dev, asn_dev, prior=0., weights=None, precision=PREC_BF16):
        """One E+M iteration on device tensors (torch) in place; returns the FP64 cost."""
        cost = C.c_double()
        check(load().mcg_lloyd_step(self.h, msr, prior, ptr(centers_dev), centers_dev.shape[0], ptr(weights),
                                    precision, ptr(asn_dev), C.byref(cost)))
        return cost.value

    def soft(self, msr, centers, prior=0., temp=1., weights=None, eps=1e-8, maxiter=1000, with_outputs=True):
        """Soft EM; with_outputs=False leaves the n x k costs and responsibilities on
        the device (returned as None)."""
        Cc = np.ascontiguousarray(centers, dtype=np.float32).copy()
        k = Cc.shape[0]
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        costs = np.empty((self.n, k), dtype=np.float32) if with_outputs else None
        resp = np.empty((self.n, k), dtype=np.float32) if with_outputs else None
        ic, fc, it = C.c_double(), C.c_double(), C.c_uint64()
        check(load().mcg_soft_em(self.h, msr, prior, temp, ptr(Cc), k, ptr(w), eps, maxiter, ptr(costs),
                                 ptr(resp), C.byref(ic), C.byref(fc), C.byref(it)))
        return dict(centers=Cc, costs=costs, asn=resp, initcost=ic.value, finalcost=fc.value, numiter=it.value)

    def rows(self, idx):
        """A new dataset of rows idx (in order) of this one."""
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        h = C.c_void_p()
        check(load().mcg_data_rows(self.h, ptr(idx), len(idx), C.byref(h)))
        return Dataset(self.ctx, h)

    def kmeanspp(self, msr, k, prior=0., seed=0, weights=None, lspp=0, n_local_trials=1):
        """D^2 seeding (n_local_trials candidates per center, then lspp rounds of local
        search++): centers, asn, costs."""
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        kk = min(k, self.n)
        centers = np.empty(kk, dtype=np.uint32)
        asn = np.empty(self.n, dtype=np.uint32)
        costs = np.empty(self.n, dtype=np.float32)
        if n_local_trials > 1:
            check(load().mcg_kmeanspp_local(self.h, msr, prior, k, seed, ptr(w), n_local_trials, lspp, ptr(centers),
                                            ptr(asn), ptr(costs)))
        elif lspp:
            check(load().mcg_kmeanspp_ls(self.h, msr, prior, k, seed, ptr(w), lspp, ptr(centers), ptr(asn),
                                         ptr(costs)))
        else:
            check(load().mcg_kmeanspp(self.h, msr, prior, k, seed, ptr(w), ptr(centers), ptr(asn), ptr(costs)))
        return centers, asn, costs


def coreset_probs(sens, ncenters, costs, asn, weights=None, ctx=None):
    ctx = ctx or default_context()
    costs = np.ascontiguousarray(costs, dtype=np.float64)
    asn = np.ascontiguousarray(asn, dtype=np.uint32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    probs = np.empty(len(costs), dtype=np.float32)
    check(load().mcg_coreset_probs(ctx.h, sens, len(costs), ncenters, ptr(costs), ptr(asn), ptr(w), ptr(probs)))
    return probs


def coreset_sample(probs, m, seed, weights=None, first_draw=0, ctx=None):
    ctx = ctx or default_context()
    probs = np.ascontiguousarray(probs, dtype=np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    idx = np.empty(max(m, 1), dtype=np.uint64)
    wt = np.empty(max(m, 1), dtype=np.float32)
    check(load().mcg_coreset_sample(ctx.h, len(probs), ptr(probs), ptr(w), m, seed, first_draw, ptr(idx), ptr(wt)))
    return wt[:m], idx[:m]


def _path(path):
    return os.fsencode(os.fspath(path))


def sampler_write(path, probs, seed, weights=None):
    """CoresetSampler::write (coreset.h:282-293): gzip stream of n, seed, f32 probs,
    the weights flag (137 / 0) and f32 weights."""
    probs = np.ascontiguousarray(probs, dtype=np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    if w is not None and len(w) != len(probs):
        raise ValueError("weights must match probs")
    check(load().mcg_sampler_write(_path(path), len(probs), int(seed), ptr(probs), ptr(w)))


def sampler_read(path):
    """CoresetSampler::read (coreset.h:306-322): returns (probs f32, seed, weights f32 or None)."""
    L = load()
    n, seed, hw = C.c_uint64(0), C.c_uint64(0), C.c_int(0)
    check(L.mcg_sampler_read(_path(path), C.byref(n), C.byref(seed), C.byref(hw), None, None, 0))
    probs = np.empty(max(n.value, 1), dtype=np.float32)
    w = np.empty(max(n.value, 1), dtype=np.float32) if hw.value else None
    check(L.mcg_sampler_read(_path(path), C.byref(n), C.byref(seed), C.byref(hw), ptr(probs), ptr(w), n.value))
    return probs[:n.value], seed.value, None if w is None else w[:n.value]


def index_coreset_write(path, indices, weights, idx_dtype=np.uint32, w_dtype=np.float32):
    """IndexCoreset<IT, FT>::write (coreset.h:113-123): gzip stream of m, IT indices, FT weights."""
    idx = np.ascontiguousarray(indices, dtype=idx_dtype)
    w = np.ascontiguousarray(weights, dtype=w_dtype)
    if len(idx) != len(w):
        raise ValueError("indices and weights must have the same length")
    check(load().mcg_index_coreset_write(_path(path), len(idx), ptr(idx), _dtype_code(idx.dtype), ptr(w),
                                         _dtype_code(w.dtype)))


def index_coreset_read(path, idx_dtype=np.uint32, w_dtype=np.float32):
    """IndexCoreset<IT, FT>::read (coreset.h:94-106): returns (indices IT, weights FT)."""
    L = load()
    m = C.c_uint64(0)
    ic, wc = _dtype_code(idx_dtype), _dtype_code(w_dtype)
    check(L.mcg_index_coreset_read(_path(path), C.byref(m), None, ic, None, wc, 0))
    idx = np.empty(max(m.value, 1), dtype=idx_dtype)
    w = np.empty(max(m.value, 1), dtype=w_dtype)
    check(L.mcg_index_coreset_read(_path(path), C.byref(m), ptr(idx), ic, ptr(w), wc, m.value))
    return idx[:m.value], w[:m.value]


class Server:
    """Concurrent clustering requests on one GPU (mcg_server_*, include/mcg.h).

    Each request: D^2 seeding, coreset sensitivities, a weighted coreset, weighted
    Lloyd on the coreset, then a hard assignment of all rows.  submit() keeps the
    request's input and output arrays alive until the request is released."""

    def __init__(self, device=0, nworkers=4, hbm_budget=0):
        h = C.c_void_p()
        check(load().mcg_server_create(device, nworkers, hbm_budget, C.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            check(load().mcg_server_destroy(self.h))
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, X=None, k=8, *, csr=None, msr=SQRL2, prior=0., weights=None, seed=0, sens=SENS_LBK,
               coreset_frac=0.01, eps=0., maxiter=9, precision=PREC_BF16, fail_at_stage=0, with_costs=True):
        """X: dense float32/float64 (n, d) numpy array or torch tensor (device tensors are
        borrowed); or csr=(data, indices, indptr, (n, d)).
        Returns a ticket.  Outputs land in self.outputs(ticket) when the request is done."""
        q = Request()
        keep = []
        if csr is not None:
            data, indices, indptr, shape = csr
            data = np.ascontiguousarray(data)
            indices = np.ascontiguousarray(indices)
            indptr = np.ascontiguousarray(indptr)
            keep += [data, indices, indptr]
            n, d = int(shape[0]), int(shape[1])
            q.data, q.data_dtype = ptr(data).value, _dtype_code(data.dtype)
            q.indices, q.idx_dtype = ptr(indices).value, _dtype_code(indices.dtype)
            q.indptr, q.ptr_dtype = ptr(indptr).value, _dtype_code(indptr.dtype)
            q.nnz = len(data)
        elif hasattr(X, "data_ptr") and not isinstance(X, np.ndarray):
            # torch tensor: rows stay where they are (device rows are borrowed)
            if not X.is_contiguous():
                raise ValueError("tensor rows must be contiguous")
            keep.append(X)
            n, d = X.shape
            q.data = X.data_ptr()
            q.data_dtype = MCG_F32 if str(X.dtype) == "torch.float32" else MCG_F64
            q.ld = d
        else:
            X = np.ascontiguousarray(X)
            if X.dtype not in (np.float32, np.float64):
                X = X.astype(np.float32)
            keep.append(X)
            n, d = X.shape
            q.data, q.data_dtype = ptr(X).value, _dtype_code(X.dtype)
            q.ld = d
        q.n, q.d = n, d
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float64)
            keep.append(w)
            q.weights = ptr(w).value
        kk = min(k, n)
        centers = np.empty((kk, d), dtype=np.float32)
        asn = np.empty(n, dtype=np.uint32)
        costs = np.empty(n, dtype=np.float32) if with_costs else None
        keep += [centers, asn, costs]
        q.msr, q.prior, q.k, q.seed, q.sens = msr, prior, k, seed, sens
        q.coreset_frac, q.eps, q.maxiter, q.precision, q.fail_at_stage = coreset_frac, eps, maxiter, precision, \
            fail_at_stage
        q.centers, q.asn = ptr(centers).value, ptr(asn).value
        q.costs = ptr(costs).value if costs is not None else None
        t = C.c_uint64()
        check(load().mcg_submit(self.h, C.byref(q), C.byref(t)))
        self._keep[t.value] = dict(keep=keep, centers=centers, asn=asn, costs=costs)
        return t.value

    @staticmethod
    def _res(r):
        return dict(state=r.state, status=r.status, cost=r.cost, coreset_initcost=r.coreset_initcost,
                    coreset_finalcost=r.coreset_finalcost, coreset_m=r.coreset_m, numiter=r.numiter,
                    queue_ms=r.queue_ms, run_ms=r.run_ms, stage_ms=list(r.stage_ms),
                    error=r.error.decode(errors="replace"))

    def poll(self, ticket):
        r = Result()
        check(load().mcg_poll(self.h, ticket, C.byref(r)))
        return self._res(r)

    def wait(self, ticket, timeout=-1.0):
        r = Result()
        check(load().mcg_wait(self.h, ticket, timeout, C.byref(r)))
        return self._res(r)

    def cancel(self, ticket):
        check(load().mcg_cancel(self.h, ticket))

    def outputs(self, ticket):
        o = self._keep[ticket]
        return dict(centers=o["centers"], asn=o["asn"], costs=o["costs"])

    def release(self, ticket):
        check(load().mcg_release(self.h, ticket))
        self._keep.pop(ticket, None)
