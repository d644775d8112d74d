"""pyminicore-compatible surface: names, signatures and defaults (no GPU)."""
import pytest


def test_extension_exports():
    from minicore_amd import _pyminicore as pm
    for name in ("hcluster", "scluster", "cmp", "kmeanspp", "CoresetSampler", "CSparseMatrix", "SparseMatrixWrapper",
                 "SumOpts"):
        assert hasattr(pm, name), name
    import re

    def default(doc, arg):
        return re.findall(arg + r": [^=]*= ([^,)]+)", doc)

    doc = pm.hcluster.__doc__
    # the three overloads (SparseMatrixWrapper, CSR, dense) with the reference's defaults
    # (pycluster.cpp:6-26, pycluster_dense.cpp:95-108)
    for arg, want in [("prior", "0.0"), ("msr", "2"), ("maxiter", "100"), ("eps", "1e-10"), ("reseed_count", "5"),
                      ("mbsize", "-1"), ("cs", "False")]:
        assert default(doc, arg) == [want] * 3, arg
    sdoc = pm.scluster.__doc__
    assert default(sdoc, "temp") == ["1.0"] * 3 and default(sdoc, "maxiter") == ["1000"] * 3
    kdoc = pm.kmeanspp.__doc__
    # keyword forms: SparseMatrixWrapper (pysmw.cpp:368-373), dense (pydense.cpp:34-77), CSR
    # (pycsparse.cpp:100-104); the two SumOpts forms take (matrix, opts, weights)
    assert default(kdoc, "ntimes") == ["1", "0", "1"] and default(kdoc, "n_local_trials") == ["1", "1", "1"]
    assert len(re.findall(r"opts: .*SumOpts", kdoc)) == 2


def test_csparsematrix_wraps_scipy():
    import numpy as np
    import scipy.sparse as sp
    from minicore_amd import _pyminicore as pm
    m = sp.random(30, 20, density=0.2, format="csr", dtype=np.float32, random_state=0)
    c = pm.CSparseMatrix(m)
    assert c.rows() == 30 and c.columns() == 20 and c.nnz() == m.nnz


def test_wrappers_need_extension(monkeypatch):
    import minicore_amd
    monkeypatch.setattr(minicore_amd, "_pyminicore", None)
    monkeypatch.setattr(minicore_amd, "_IMPORT_ERROR", ImportError("x"), raising=False)
    with pytest.raises(RuntimeError):
        minicore_amd.hcluster([[1.0]], [0])


def test_thread_count_functions():
    """set_num_threads / get_num_threads (pyminicore) keep and return the count."""
    import minicore_amd as mc
    old = mc.get_num_threads()
    mc.set_num_threads(5)
    assert mc.get_num_threads() == 5
    with pytest.raises(ValueError):
        mc.set_num_threads(0)
    mc.set_num_threads(old)



def test_sparse_matrix_wrapper_host_side():
    """SparseMatrixWrapper (pysmw.cpp:31-265) without the device: an owned copy with rows
    sorted by column, skip_empty, float/double storage, tocsr, rowsel, submatrix, sums,
    counts, transpose; blaze archives refused."""
    import numpy as np
    import scipy.sparse as sp
    import minicore_amd as mc
    rng = np.random.default_rng(3)
    A = sp.random(40, 25, density=0.2, random_state=4, format="csr", dtype=np.float64)
    A.data = np.round(A.data * 10, 3)
    A = A.tolil()
    A[7, :] = 0
    A = A.tocsr()
    A.eliminate_zeros()
    # unsorted column order within rows is sorted by the wrapper
    order = np.concatenate([np.arange(A.indptr[i + 1] - 1, A.indptr[i] - 1, -1) for i in range(40)]).astype(np.int64)
    B = sp.csr_matrix((A.data[order], A.indices[order], A.indptr.copy()), shape=A.shape)
    assert not B.has_sorted_indices
    w = mc.SparseMatrixWrapper(B)
    assert w.rows() == 40 and w.columns() == 25 and w.nonzeros() == A.nnz and w.is_float() and not w.is_double()
    (data, idx, ptr), shape = w.tocsr()
    got = sp.csr_matrix((data, idx, ptr), shape=tuple(shape))
    assert np.all(np.diff(idx[ptr[0]:ptr[1]]) > 0)
    np.testing.assert_allclose(got.toarray(), A.toarray().astype(np.float32), rtol=0)
    wd = mc.SparseMatrixWrapper(A, use_float=False)
    assert wd.is_double()
    np.testing.assert_array_equal(wd.rowsel(np.array([3, 0, 3], np.uint32)), A.toarray()[[3, 0, 3]])
    assert w.rowsel(np.array([1, 2])).dtype == np.float32
    we = mc.SparseMatrixWrapper(A, skip_empty=True)
    assert we.rows() == 39
    np.testing.assert_allclose(w.sum(1, False), A.toarray().astype(np.float32).sum(1), rtol=1e-6)
    np.testing.assert_allclose(w.sum(0), A.toarray().sum(0), rtol=1e-5)
    assert abs(wd.sum() - A.sum()) < 1e-9
    assert w.count_nnz() == A.nnz and w.count_nnz(1).tolist() == np.diff(A.indptr).tolist()
    sub = wd.submatrix(np.array([5, 1]), np.array([3, 0, 10]))
    (d2, i2, p2), sh2 = sub.tocsr()
    np.testing.assert_array_equal(sp.csr_matrix((d2, i2, p2), shape=tuple(sh2)).toarray(), A.toarray()[[5, 1]][:, [3, 0, 10]])
    wd.transpose_()
    (d3, i3, p3), sh3 = wd.tocsr()
    np.testing.assert_array_equal(sp.csr_matrix((d3, i3, p3), shape=tuple(sh3)).toarray(), A.toarray().T)
    with pytest.raises(RuntimeError):
        mc.SparseMatrixWrapper("some.blaze")
    with pytest.raises(RuntimeError):
        w.tofile("/tmp/x.blaze")
    so = mc.SumOpts("JSD", k=7, prior=0.5, sm="LBK")
    assert so.k == 7 and so.gamma == 0.5 and so.cs == "LBK" and so.prior == "GAMMA"
    so.cs = 3
    so.lspp = 2
    assert so.cs == "VX" and so.lspp == 2 and mc.SumOpts(4, prior=1.).prior == "DIRICHLET"
