"""pyminicore-compatible surface: names, signatures and defaults (no GPU)."""
import pytest


def test_extension_exports():
    from minicore_amd import _pyminicore as pm
    for name in ("hcluster", "scluster", "cmp", "kmeanspp", "CoresetSampler", "CSparseMatrix"):
        assert hasattr(pm, name), name
    import re

    def default(doc, arg):
        return re.findall(arg + r": [^=]*= ([^,)]+)", doc)

    doc = pm.hcluster.__doc__
    # both overloads (CSR and dense) with the reference's defaults (pycluster_dense.cpp:95-108)
    for arg, want in [("prior", "0.0"), ("msr", "2"), ("maxiter", "100"), ("eps", "1e-10"), ("reseed_count", "5"),
                      ("mbsize", "-1"), ("cs", "False")]:
        assert default(doc, arg) == [want, want], arg
    sdoc = pm.scluster.__doc__
    assert default(sdoc, "temp") == ["1.0", "1.0"] and default(sdoc, "maxiter") == ["1000", "1000"]
    kdoc = pm.kmeanspp.__doc__
    assert default(kdoc, "ntimes") == ["0", "1"] and default(kdoc, "n_local_trials") == ["1", "1"]


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


def test_get_counthist():
    """get_counthist (pycounthist.cpp:72-155): value -> ascending indices, index dtype
    by length (pyhelpers.h:220-225)."""
    import numpy as np
    import minicore_amd as mc
    a = np.array([3, 1, 3, 2, 1, 3], dtype=np.uint32)
    h = mc.get_counthist(a)
    assert set(h) == {1, 2, 3}
    assert h[3].tolist() == [0, 2, 5] and h[1].tolist() == [1, 4] and h[2].tolist() == [3]
    assert h[3].dtype == np.uint8
    b = np.random.default_rng(0).integers(0, 7, 70000)
    hb = mc.get_counthist(b)
    assert all(v.dtype == np.uint32 for v in hb.values())
    for key, idx in hb.items():
        assert np.array_equal(idx, np.nonzero(b == key)[0])
    f = mc.get_counthist(np.array([.5, .25, .5], dtype=np.float32))
    assert f[.5].tolist() == [0, 2] and isinstance(next(iter(f)), float)
    rows = mc.get_counthist(np.array([[1, 2, 1], [0, 0, 5]]))
    assert rows[0][1].tolist() == [0, 2] and rows[1][0].tolist() == [0, 1]
