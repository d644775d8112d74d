"""Dataset ingest (mcg_data_csr, mcg_data_dense): every accepted dtype gives the same
device rows, and malformed CSR input is refused with the reference's checks
(indptr bounds and order, column range, sorted unique columns per row)."""
import numpy as np
import pytest

from minicore_amd.datagen import poisson_csr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _costs(mcg, ctx, data, idx, ptr, shape, C):
    ds = mcg.Dataset.csr(data, idx, ptr, shape, ctx)
    K = ds.cost_matrix(mcg.SQRL2, C)
    ds.close()
    return K


def test_csr_dtypes_equivalent(mcg, ctx):
    say("csr dtypes")
    n, d = 700, 300
    data, idx, ptr = poisson_csr(n, d, 0.05, 3)
    C = np.random.default_rng(1).uniform(0, 2, (5, d)).astype(np.float32)
    ref = _costs(mcg, ctx, data.astype(np.float32), idx.astype(np.uint32), ptr.astype(np.uint64), (n, d), C)
    variants = [
        (data.astype(np.float64), idx.astype(np.uint16), ptr.astype(np.int64)),
        (data.astype(np.uint16), idx.astype(np.int32), ptr.astype(np.uint32)),
        (data.astype(np.uint8), idx.astype(np.int64), ptr.astype(np.int32)),
        (data.astype(np.uint32), idx.astype(np.uint64), ptr.astype(np.uint64)),
    ]
    for dv, iv, pv in variants:
        got = _costs(mcg, ctx, dv, iv, pv, (n, d), C)
        assert np.array_equal(got, ref), (dv.dtype, iv.dtype, pv.dtype)


def test_csr_device_arrays(mcg, ctx):
    import torch
    say("csr from device tensors")
    n, d = 500, 200
    data, idx, ptr = poisson_csr(n, d, 0.05, 4)
    C = np.random.default_rng(2).uniform(0, 2, (3, d)).astype(np.float32)
    ref = _costs(mcg, ctx, data, idx, ptr, (n, d), C)
    t = [torch.from_numpy(a).cuda() for a in (data, idx.astype(np.int32), ptr.astype(np.int64))]
    torch.cuda.synchronize()
    ds = mcg.Dataset.csr(*t, (n, d), ctx)
    assert np.array_equal(ds.cost_matrix(mcg.SQRL2, C), ref)
    ds.close()


@pytest.mark.parametrize("case,msg", [
    ("ptr_start", "indptr must start at 0 and end at nnz"),
    ("ptr_end", "indptr must start at 0 and end at nnz"),
    ("ptr_decreasing", "indptr must be non-decreasing"),
    ("col_range", "column index out of range"),
    ("col_negative", "negative index"),
    ("col_unsorted", "column indices must be sorted and unique within a row"),
    ("col_duplicate", "column indices must be sorted and unique within a row"),
])
def test_csr_malformed_refused(mcg, ctx, case, msg):
    say(f"csr malformed {case}")
    data = np.array([1., 2., 3., 4., 5., 6.], np.float32)
    idx = np.array([0, 3, 1, 2, 4, 0], np.int32)
    ptr = np.array([0, 2, 5, 6], np.int64)
    d = 5
    if case == "ptr_start":
        ptr[0] = 1
    elif case == "ptr_end":
        ptr[-1] = 5
    elif case == "ptr_decreasing":
        ptr[1], ptr[2] = 4, 3
    elif case == "col_range":
        idx[4] = 5
    elif case == "col_negative":
        idx[1] = -1
    elif case == "col_unsorted":
        idx[2], idx[3] = 2, 1
    elif case == "col_duplicate":
        idx[3] = 1
    with pytest.raises(mcg.MCGError) as e:
        mcg.Dataset.csr(data, idx, ptr, (3, d), ctx)
    assert e.value.code == -1 and msg in str(e.value), str(e.value)
    # the valid input is accepted; a row boundary is not an ordering violation
    ok = mcg.Dataset.csr(data, np.array([0, 3, 1, 2, 4, 0], np.int32), np.array([0, 2, 5, 6], np.int64), (3, d), ctx)
    ok.close()
