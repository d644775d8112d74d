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

