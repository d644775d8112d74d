// pybind11 module `_pyminicore`: the pyminicore entry points of the clustering
// hot path (hcluster, scluster, cmp, kmeanspp, CoresetSampler, CSparseMatrix),
// with the reference's names, argument order, defaults and return values,
// running on libmcg (include/mcg.h).
//
// Reference bindings mirrored (paths relative to the dnbaker/minicore tree):
//   hcluster   python/pycluster_dense.cpp:69-110, pycluster_csr.cpp:5-30
//   scluster   python/pycluster_softdense.cpp:65-91, pycluster_softcsr.cpp:4-29
//   cmp        python/pycmp.cpp:147-248, pyarrcmp.cpp:126-145
//   kmeanspp   python/pydense.cpp:34-77, pycsparse.cpp:100-104 (smw.h:179-410)
//   CoresetSampler python/pycs.cpp:4-58
//   CSparseMatrix  python/pycsparse.cpp:15-89, pycsparse.h:33-70
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "mcg.h"

namespace py = pybind11;
using namespace pybind11::literals;

namespace {

mcg_ctx *g_ctx = nullptr;
int g_sqrl2_precision = MCG_PREC_BF16;

void check(int rc) {
    if (rc == MCG_OK) return;
    std::string msg = mcg_last_error();
    if (rc == MCG_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error("libmcg error " + std::to_string(rc) + ": " + msg);
}

mcg_ctx *ctx() {
    if (!g_ctx) check(mcg_ctx_create(0, &g_ctx));
    return g_ctx;
}

struct DataHandle {
    mcg_data *h = nullptr;
    ~DataHandle() { if (h) mcg_data_free(h); }
};
using DataPtr = std::shared_ptr<DataHandle>;

// measure codes: distance.h:24-64 (names as in distance.h / msr2str)
int measure_code(py::object msr) {
    int code;
    if (py::isinstance<py::str>(msr)) {
        std::string s = msr.cast<std::string>();
        std::transform(s.begin(), s.end(), s.begin(), ::toupper);
        // prob2str names (distance.h:273-304) and the enum aliases (distance.h:52-62)
        static const std::map<std::string, int> names = {
            {"L1", 0}, {"L2", 1}, {"SQRL2", 2}, {"JSM", 3}, {"PSM", 3}, {"JSD", 4}, {"PSD", 4}, {"MKL", 5},
            {"HELLINGER", 6}, {"BHATTACHARYYA_METRIC", 7}, {"BHATTACHARYYA_DISTANCE", 8},
            {"TOTAL_VARIATION_DISTANCE", 9}, {"TVD", 9}, {"LLR", 10}, {"WLLR", 10}, {"REVERSE_MKL", 11},
            {"RMKL", 11}, {"UWLLR", 12}, {"ITAKURA_SAITO", 13}, {"IS", 13}, {"REVERSE_ITAKURA_SAITO", 14},
            {"COSINE_DISTANCE", 15}, {"PROBABILITY_COSINE_DISTANCE", 16}, {"COSINE_SIMILARITY", 17},
            {"PROBABILITY_COSINE_SIMILARITY", 18}, {"SYMMETRIC_ITAKURA_SAITO", 23}, {"SIS", 23},
            {"RSYMMETRIC_ITAKURA_SAITO", 24}, {"REVERSE_SYMMETRIC_ITAKURA_SAITO", 24}, {"RSIS", 24}, {"SRLRT", 25},
            {"SRULRT", 26}, {"POISSON", 64}, {"POISSON_BREGMAN", 64}, {"IDIV", 64}};
        auto it = names.find(s);
        if (it == names.end()) throw std::invalid_argument("measure " + s + " has no GPU kernel");
        code = it->second;
    } else {
        code = msr.cast<int>();
    }
    static const std::set<int> ok = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 23, 24, 25, 26,
                                     64};
    if (!ok.count(code)) throw std::invalid_argument("measure " + std::to_string(code) + " has no GPU kernel");
    return code;
}

// ---- CSparseMatrix: borrowed scipy CSR arrays (pycsparse.h:33-70) ------------------
int dtype_code(const py::array &a) {
    const char k = a.dtype().kind();
    const auto sz = a.itemsize();
    if (k == 'f') return sz == 4 ? MCG_F32 : MCG_F64;
    if (k == 'u') return sz == 1 ? MCG_U8 : sz == 2 ? MCG_U16 : sz == 4 ? MCG_U32 : MCG_U64;
    if (k == 'i') {
        if (sz == 4) return MCG_I32;
        if (sz == 8) return MCG_I64;
    }
    throw std::invalid_argument(std::string("unsupported dtype ") + py::str(a.dtype()).cast<std::string>());
}

struct PyCSparseMatrix {
    py::array data, indices, indptr;
    uint64_t nr = 0, nc = 0, nnz = 0;
    DataPtr dev;  // device-resident copy, built on first use

    PyCSparseMatrix(py::array d, py::array i, py::array p, uint64_t r, uint64_t c, uint64_t z)
        : data(py::array::ensure(d, py::array::c_style)), indices(py::array::ensure(i, py::array::c_style)),
          indptr(py::array::ensure(p, py::array::c_style)), nr(r), nc(c), nnz(z) {}
    explicit PyCSparseMatrix(py::object obj)
        : PyCSparseMatrix(obj.attr("data"), obj.attr("indices"), obj.attr("indptr"),
                          py::cast<py::sequence>(obj.attr("shape"))[0].cast<uint64_t>(),
                          py::cast<py::sequence>(obj.attr("shape"))[1].cast<uint64_t>(),
                          obj.attr("nnz").cast<uint64_t>()) {}

    mcg_data *device() {
        if (!dev) {
            auto h = std::make_shared<DataHandle>();
            check(mcg_data_csr(ctx(), data.data(), dtype_code(data), indices.data(), dtype_code(indices),
                               indptr.data(), dtype_code(indptr), nr, nc, nnz, &h->h));
            dev = h;
        }
        return dev->h;
    }
    // dense float32 copy of selected rows
    std::vector<float> dense_rows(const std::vector<int64_t> &rows) const {
        std::vector<float> out(rows.size() * nc, 0.f);
        auto P = [&](uint64_t i) -> uint64_t {
            switch (dtype_code(indptr)) {
            case MCG_U32: return ((const uint32_t *)indptr.data())[i];
            case MCG_I32: return (uint64_t)((const int32_t *)indptr.data())[i];
            case MCG_U64: return ((const uint64_t *)indptr.data())[i];
            case MCG_I64: return (uint64_t)((const int64_t *)indptr.data())[i];
            }
            throw std::invalid_argument("indptr dtype");
        };
        auto J = [&](uint64_t p) -> uint64_t {
            switch (dtype_code(indices)) {
            case MCG_U16: return ((const uint16_t *)indices.data())[p];
            case MCG_U32: return ((const uint32_t *)indices.data())[p];
            case MCG_I32: return (uint64_t)((const int32_t *)indices.data())[p];
            case MCG_U64: return ((const uint64_t *)indices.data())[p];
            case MCG_I64: return (uint64_t)((const int64_t *)indices.data())[p];
            }
            throw std::invalid_argument("indices dtype");
        };
        auto V = [&](uint64_t p) -> float {
            switch (dtype_code(data)) {
            case MCG_F32: return ((const float *)data.data())[p];
            case MCG_F64: return (float)((const double *)data.data())[p];
            case MCG_U16: return (float)((const uint16_t *)data.data())[p];
            case MCG_U32: return (float)((const uint32_t *)data.data())[p];
            case MCG_I32: return (float)((const int32_t *)data.data())[p];
            case MCG_U8: return (float)((const uint8_t *)data.data())[p];
            case MCG_U64: return (float)((const uint64_t *)data.data())[p];
            case MCG_I64: return (float)((const int64_t *)data.data())[p];
            }
            throw std::invalid_argument("data dtype");
        };
        for (size_t q = 0; q < rows.size(); ++q) {
            const int64_t r = rows[q];
            if (r < 0 || (uint64_t)r >= nr) throw std::invalid_argument("center index out of range");
            for (uint64_t p = P(r); p < P(r + 1); ++p) out[q * nc + J(p)] = V(p);
        }
        return out;
    }
    std::vector<float> dense_all() const {
        std::vector<int64_t> rows(nr);
        for (uint64_t i = 0; i < nr; ++i) rows[i] = (int64_t)i;
        return dense_rows(rows);
    }
};

// ---- dense inputs ------------------------------------------------------------------
struct Dense {
    py::array arr;        // kept alive (float32 or float64, C order)
    uint64_t n = 0, d = 0;
    bool f64 = false;
    DataPtr dev;
};

Dense dense_of(py::array a) {
    Dense D;
    if (a.ndim() != 2) throw std::invalid_argument("Expected 2 dimensions");
    if (a.dtype().kind() == 'f' && a.itemsize() == 8) {
        D.arr = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(a);
        D.f64 = true;
    } else {
        D.arr = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(a);
    }
    D.n = (uint64_t)D.arr.shape(0);
    D.d = (uint64_t)D.arr.shape(1);
    auto h = std::make_shared<DataHandle>();
    check(mcg_data_dense(ctx(), D.arr.data(), D.f64 ? MCG_F64 : MCG_F32, D.n, D.d, D.d, &h->h));
    D.dev = h;
    return D;
}

std::vector<float> row_f32(const Dense &D, int64_t r) {
    if (r < 0 || (uint64_t)r >= D.n) throw std::invalid_argument("center index out of range");
    std::vector<float> v(D.d);
    if (D.f64) {
        const double *p = (const double *)D.arr.data() + (uint64_t)r * D.d;
        for (uint64_t j = 0; j < D.d; ++j) v[j] = (float)p[j];
    } else {
        std::memcpy(v.data(), (const float *)D.arr.data() + (uint64_t)r * D.d, D.d * 4);
    }
    return v;
}

// weights: any numeric dtype -> float64 (pycluster_dense.cpp:43-61)
py::object weights_f64(py::object w, uint64_t n) {
    if (w.is_none()) return py::none();
    auto a = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(w);
    if (!a || (uint64_t)a.size() != n) throw std::invalid_argument("weights must have one entry per row");
    return a;
}
const double *wptr(const py::object &w) {
    return w.is_none() ? nullptr : (const double *)py::cast<py::array>(w).data();
}

bool is_index_list(py::object centers) {
    if (py::isinstance<py::array>(centers)) {
        auto a = py::cast<py::array>(centers);
        return a.ndim() == 1;
    }
    if (py::isinstance<py::sequence>(centers)) {
        auto s = py::cast<py::sequence>(centers);
        return s.size() > 0 && !py::isinstance<py::array>(s[0]) && !py::isinstance<py::sequence>(s[0]);
    }
    return false;
}

std::vector<int64_t> index_list(py::object centers) {
    auto a = py::array_t<int64_t, py::array::c_style | py::array::forcecast>::ensure(centers);
    if (!a) throw std::invalid_argument("Invalid: expected numpy array or list of numpy arrays or list of center ids");
    return std::vector<int64_t>(a.data(), a.data() + a.size());
}

// k x d center block from an array, a list of arrays or indices (pyhelpers.h:122-218)
template <class T>
std::vector<T> center_block_t(py::object centers, uint64_t d,
                              const std::function<std::vector<T>(const std::vector<int64_t> &)> &by_index,
                              uint64_t *k_out) {
    std::vector<T> C;
    if (is_index_list(centers)) {
        auto ids = index_list(centers);
        C = by_index(ids);
        *k_out = ids.size();
        return C;
    }
    if (py::isinstance<py::array>(centers)) {
        auto a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(centers);
        if (!a || a.ndim() != 2 || (uint64_t)a.shape(1) != d)
            throw std::invalid_argument("centers must be k x d (d = number of columns)");
        *k_out = (uint64_t)a.shape(0);
        C.assign(a.data(), a.data() + a.size());
        return C;
    }
    if (py::isinstance<py::sequence>(centers)) {
        uint64_t k = 0;
        for (auto item : py::cast<py::sequence>(centers)) {
            auto a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(item);
            if (!a || (uint64_t)a.size() != d) throw std::invalid_argument("each center must have d entries");
            C.insert(C.end(), a.data(), a.data() + d);
            ++k;
        }
        *k_out = k;
        return C;
    }
    throw std::invalid_argument("centers must be a 2d numpy array or sequence containing numpy arrays of the full vectors or center ids");
}

std::vector<float> center_block(py::object centers, uint64_t d,
                                const std::function<std::vector<float>(const std::vector<int64_t> &)> &by_index,
                                uint64_t *k_out) {
    return center_block_t<float>(centers, d, by_index, k_out);
}

// centers2pylist (pyhelpers.h:66-112): ((data f32, idx u32, indptr u64), shape u64[2]) of the non-zeros
py::object centers_csr_tuple(const std::vector<float> &C, uint64_t k, uint64_t d) {
    std::vector<uint64_t> ip(k + 1, 0);
    for (uint64_t c = 0; c < k; ++c) {
        uint64_t z = 0;
        for (uint64_t j = 0; j < d; ++j) z += C[c * d + j] != 0.f;
        ip[c + 1] = ip[c] + z;
    }
    py::array_t<float> data(ip[k]);
    py::array_t<uint32_t> idx(ip[k]);
    py::array_t<uint64_t> indptr(k + 1);
    float *dp = data.mutable_data();
    uint32_t *jp = idx.mutable_data();
    std::memcpy(indptr.mutable_data(), ip.data(), (k + 1) * 8);
    uint64_t q = 0;
    for (uint64_t c = 0; c < k; ++c)
        for (uint64_t j = 0; j < d; ++j)
            if (C[c * d + j] != 0.f) { dp[q] = C[c * d + j]; jp[q] = (uint32_t)j; ++q; }
    py::array_t<uint64_t> shape(2);
    shape.mutable_data()[0] = k;
    shape.mutable_data()[1] = d;
    return py::make_tuple(py::make_tuple(data, idx, indptr), shape);
}

// minibatch arguments of hcluster (python/pycluster.h:31-42)
struct MiniArgs {
    py::ssize_t mbsize, ncheckins, reseed_count;
    bool with_rep;
    uint64_t seed;
    bool cs;
};

// ---- hcluster -------------------------------------------------------------------------
// mbsize < 0: perform_hard_clustering; otherwise perform_hard_minibatch_clustering,
// or hmb_coreset_clustering with cs, with the check-in period ceil(maxiter /
// ncheckins) (ncheckins < 0 means 10).
py::dict run_hard(mcg_data *h, uint64_t n, uint64_t d, std::vector<float> &C, uint64_t k, double prior,
                  int msr, py::object w, double eps, uint64_t maxiter, const MiniArgs &mb) {
    py::array_t<uint32_t> asn(n);
    py::array_t<float> costs(n);
    double ic = 0, fc = 0;
    uint64_t it = 0;
    if (mb.mbsize < 0) {
        check(mcg_lloyd(h, msr, prior, C.data(), k, wptr(w), eps, maxiter, g_sqrl2_precision, asn.mutable_data(),
                        costs.mutable_data(), &ic, &fc, &it));
    } else {
        const uint64_t nck = mb.ncheckins < 0 ? 10 : (uint64_t)mb.ncheckins;
        if (nck == 0) throw std::invalid_argument("ncheckins must be positive");
        const uint64_t freq = (maxiter + nck - 1) / nck;
        const uint64_t reseed = mb.reseed_count < 0 ? 0 : (uint64_t)mb.reseed_count;
        if (mb.cs)
            check(mcg_cs_minibatch(h, msr, prior, C.data(), k, wptr(w), (uint64_t)mb.mbsize, maxiter, freq ? freq : 1,
                                   reseed, mb.seed, g_sqrl2_precision, asn.mutable_data(), costs.mutable_data(), &ic,
                                   &fc, &it));
        else
            check(mcg_minibatch(h, msr, prior, C.data(), k, wptr(w), (uint64_t)mb.mbsize, maxiter, freq ? freq : 1,
                                reseed, mb.with_rep ? 1 : 0, mb.seed, g_sqrl2_precision, asn.mutable_data(),
                                costs.mutable_data(), &ic, &fc, &it));
    }
    return py::dict("initcost"_a = ic, "finalcost"_a = fc, "numiter"_a = it, "costs"_a = costs, "asn"_a = asn);
}

py::object hcluster_dense(py::array dataset, py::object centers, double prior, py::object msr, py::object weights,
                          double eps, uint64_t maxiter, uint64_t seed, py::ssize_t mbsize, py::ssize_t ncheckins,
                          py::ssize_t reseed_count, bool with_rep, bool cs) {
    const int m = measure_code(msr);
    Dense D = dense_of(dataset);
    uint64_t k = 0;
    if (D.f64 && m == MCG_SQRL2 && mbsize < 0) {
        // float64 rows, SQRL2: the FT = double instantiation (pycluster_dense.cpp:85-90),
        // every value in FP64 (mcg_lloyd_f64)
        auto C = center_block_t<double>(centers, D.d, [&](const std::vector<int64_t> &ids) {
            std::vector<double> out;
            for (auto r : ids) {
                if (r < 0 || (uint64_t)r >= D.n) throw std::invalid_argument("center index out of range");
                const double *p = (const double *)D.arr.data() + (uint64_t)r * D.d;
                out.insert(out.end(), p, p + D.d);
            }
            return out;
        }, &k);
        if (k == 0) throw std::invalid_argument("no centers");
        auto w = weights_f64(weights, D.n);
        py::array_t<uint32_t> asn(D.n);
        py::array_t<double> costs(D.n);
        double ic = 0, fc = 0;
        uint64_t it = 0;
        check(mcg_lloyd_f64(D.dev->h, C.data(), k, wptr(w), eps, maxiter, asn.mutable_data(), costs.mutable_data(),
                            &ic, &fc, &it));
        py::array_t<double> ctr({(py::ssize_t)k, (py::ssize_t)D.d});
        std::memcpy(ctr.mutable_data(), C.data(), k * D.d * 8);
        return py::dict("initcost"_a = ic, "finalcost"_a = fc, "numiter"_a = it, "costs"_a = costs, "asn"_a = asn,
                        "centers"_a = ctr);
    }
    auto C = center_block(centers, D.d, [&](const std::vector<int64_t> &ids) {
        std::vector<float> out;
        for (auto r : ids) { auto v = row_f32(D, r); out.insert(out.end(), v.begin(), v.end()); }
        return out;
    }, &k);
    if (k == 0) throw std::invalid_argument("no centers");
    auto w = weights_f64(weights, D.n);
    py::dict r = run_hard(D.dev->h, D.n, D.d, C, k, prior, m, w, eps, maxiter,
                          MiniArgs{mbsize, ncheckins, reseed_count, with_rep, seed, cs});
    if (D.f64) {
        py::array_t<double> ctr({(py::ssize_t)k, (py::ssize_t)D.d});
        for (uint64_t q = 0; q < k * D.d; ++q) ctr.mutable_data()[q] = C[q];
        r["centers"] = ctr;
    } else {
        py::array_t<float> ctr({(py::ssize_t)k, (py::ssize_t)D.d});
        std::memcpy(ctr.mutable_data(), C.data(), k * D.d * 4);
        r["centers"] = ctr;
    }
    return r;
}

template <class M>
py::object hcluster_csr(M &smw, py::object centers, double prior, py::object msr, py::object weights,
                        double eps, uint64_t maxiter, uint64_t seed, py::ssize_t mbsize, py::ssize_t ncheckins,
                        py::ssize_t reseed_count, bool with_rep, bool cs) {
    const int m = measure_code(msr);
    uint64_t k = 0;
    auto C = center_block(centers, smw.nc, [&](const std::vector<int64_t> &ids) { return smw.dense_rows(ids); }, &k);
    if (k == 0) throw std::invalid_argument("no centers");
    auto w = weights_f64(weights, smw.nr);
    py::dict r = run_hard(smw.device(), smw.nr, smw.nc, C, k, prior, m, w, eps, maxiter,
                          MiniArgs{mbsize, ncheckins, reseed_count, with_rep, seed, cs});
    r["centers"] = centers_csr_tuple(C, k, smw.nc);
    return r;
}

// ---- scluster -------------------------------------------------------------------------
py::dict run_soft(mcg_data *h, uint64_t n, uint64_t d, std::vector<float> &C, uint64_t k, double prior, int msr,
                  double temp, uint64_t maxiter, const std::string &savepref, py::object w) {
    py::object costs, asns;
    if (!savepref.empty()) {
        auto memmap = py::module_::import("numpy").attr("memmap");
        std::vector<py::ssize_t> shape{(py::ssize_t)n, (py::ssize_t)k};
        costs = memmap(py::str(savepref + ".costs.f32.npy"), "dtype"_a = "f", "mode"_a = "w+", "shape"_a = shape);
        asns = memmap(py::str(savepref + ".asns.f32.npy"), "dtype"_a = "f", "mode"_a = "w+", "shape"_a = shape);
    } else {
        costs = py::array_t<float>({(py::ssize_t)n, (py::ssize_t)k});
        asns = py::array_t<float>({(py::ssize_t)n, (py::ssize_t)k});
    }
    double ic = 0, fc = 0;
    uint64_t it = 0;
    check(mcg_soft_em(h, msr, prior, temp, C.data(), k, wptr(w), 1e-8, maxiter,
                      (float *)py::cast<py::array>(costs).mutable_data(), (float *)py::cast<py::array>(asns).mutable_data(),
                      &ic, &fc, &it));
    return py::dict("initcost"_a = ic, "finalcost"_a = fc, "numiter"_a = it, "centers"_a = centers_csr_tuple(C, k, d),
                    "costs"_a = costs, "asn"_a = asns);
}

py::object scluster_dense(py::array mat, py::object centers, py::object msr, double prior, double temp,
                          uint64_t maxiter, py::ssize_t mbsize, py::ssize_t, std::string savepref, py::object weights) {
    if (prior < 0.) prior = 0.;  // pycluster_softdense.cpp:71
    if (mbsize > 0) throw std::runtime_error("Not yet completed: minibatch soft clustering");   // solve.h:319-321
    const int m = measure_code(msr);
    Dense D = dense_of(mat);
    uint64_t k = 0;
    // centers given by index keep only the positive entries of the rows (pyhelpers.h:150-157)
    auto C = center_block(centers, D.d, [&](const std::vector<int64_t> &ids) {
        std::vector<float> out;
        for (auto r : ids) {
            auto v = row_f32(D, r);
            for (auto &x : v) if (!(x > 0.f)) x = 0.f;
            out.insert(out.end(), v.begin(), v.end());
        }
        return out;
    }, &k);
    if (k == 0) throw std::invalid_argument("no centers");
    return run_soft(D.dev->h, D.n, D.d, C, k, prior, m, temp, maxiter, savepref, weights_f64(weights, D.n));
}

template <class M>
py::object scluster_csr(M &smw, py::object centers, py::object msr, double prior, double temp,
                        uint64_t maxiter, py::ssize_t mbsize, py::ssize_t, std::string savepref, py::object weights) {
    if (mbsize > 0) throw std::runtime_error("Not yet completed: minibatch soft clustering");   // solve.h:319-321
    const int m = measure_code(msr);
    // csc.h:834 throws NotImplementedError, a std::runtime_error
    if (m == MCG_L1) throw std::runtime_error("Not implemented: weighted L1 median for CSparseMatrix");
    uint64_t k = 0;
    auto C = center_block(centers, smw.nc, [&](const std::vector<int64_t> &ids) { return smw.dense_rows(ids); }, &k);
    if (k == 0) throw std::invalid_argument("no centers");
    return run_soft(smw.device(), smw.nr, smw.nc, C, k, prior, m, temp, maxiter, savepref,
                    weights_f64(weights, smw.nr));
}

// ---- cmp ----------------------------------------------------------------------------------
// out (n_data x k) = msr(first = data row, second = center row)
py::array_t<float> cost_block(mcg_data *h, uint64_t n, const std::vector<float> &C, uint64_t k, int msr, double prior) {
    py::array_t<float> out({(py::ssize_t)n, (py::ssize_t)k});
    check(mcg_cost_matrix(h, msr, prior, C.data(), k, out.mutable_data()));
    return out;
}

py::array transpose_copy(py::array_t<float> a) {
    return py::cast<py::array>(py::module_::import("numpy").attr("ascontiguousarray")(a.attr("T")));
}

// rows2d_reverse: the 2-d array form honours `reverse` (the SparseMatrixWrapper binding,
// pycmp.cpp:84-96); the CSparseMatrix binding ignores it there (pycmp.cpp:191-197)
template <class M>
py::object cmp_csr_array(M &lhs, py::array arr, py::object msr, py::object betaprior, py::object reverse,
                         bool rows2d_reverse = false) {
    const int m = measure_code(msr);
    const double prior = betaprior.cast<double>();
    const bool rev = reverse.cast<bool>();
    auto a = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(arr);
    if (a.ndim() == 1) {
        if ((uint64_t)a.size() != lhs.nc) throw std::invalid_argument("Array must be of the same dimensionality as the matrix");
        std::vector<float> sv(a.data(), a.data() + a.size());
        if (rev) {  // msr(first = matrix row, second = array)
            auto out = cost_block(lhs.device(), lhs.nr, sv, 1, m, prior);
            return out.attr("reshape")(lhs.nr);
        }
        // msr(first = array, second = matrix row)
        Dense D = dense_of(a.reshape({(py::ssize_t)1, (py::ssize_t)lhs.nc}));
        auto out = cost_block(D.dev->h, 1, lhs.dense_all(), lhs.nr, m, prior);
        return out.attr("reshape")(lhs.nr);
    }
    if (a.ndim() != 2) throw std::invalid_argument("NumPy array expected to have 1 or two dimensions.");
    if ((uint64_t)a.shape(1) != lhs.nc) throw std::invalid_argument("Array must be of the same dimensionality as the matrix");
    if (rev && rows2d_reverse) {   // out[x][y] = msr(first = matrix row x, second = data row y)
        std::vector<float> C(a.data(), a.data() + a.size());
        return cost_block(lhs.device(), lhs.nr, C, (uint64_t)a.shape(0), m, prior);
    }
    // out[x][y] = msr(first = data row y, second = matrix row x)   (pycmp.cpp:193-197)
    Dense D = dense_of(a);
    return transpose_copy(cost_block(D.dev->h, D.n, lhs.dense_all(), lhs.nr, m, prior));
}

py::object cmp_csr_csr(PyCSparseMatrix &lhs, PyCSparseMatrix &rhs, py::object msr, py::object betaprior) {
    const int m = measure_code(msr);
    if (lhs.nc != rhs.nc) throw std::invalid_argument("mismatched # columns");
    // out[l][r] = msr(first = rhs row r, second = lhs row l)   (pycmp.cpp:244)
    return transpose_copy(cost_block(rhs.device(), rhs.nr, lhs.dense_all(), lhs.nr, m, betaprior.cast<double>()));
}

py::object cmp_arrays(py::array lhs, py::array rhs, py::object msr, double prior, bool reverse, int use_double) {
    const int m = measure_code(msr);
    if (lhs.ndim() != rhs.ndim() || lhs.ndim() < 1 || lhs.ndim() > 2) throw std::invalid_argument("arrcmp requires arrays of 1 or 2d");
    if (use_double < 0) use_double = std::max(lhs.itemsize(), rhs.itemsize()) > 4;
    const bool one = lhs.ndim() == 1;
    py::array L = one ? py::array(lhs.attr("reshape")(1, -1)) : lhs;
    py::array R = one ? py::array(rhs.attr("reshape")(1, -1)) : rhs;
    auto Lf = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(L);
    auto Rf = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(R);
    if (Lf.shape(1) != Rf.shape(1)) throw std::invalid_argument("mismatched # columns");
    py::array out;
    if (reverse) {  // out[i][j] = msr(first = rhs_j, second = lhs_i)   (pyarrcmp.cpp:16-20)
        Dense D = dense_of(Rf);
        std::vector<float> C(Lf.data(), Lf.data() + Lf.size());
        out = transpose_copy(cost_block(D.dev->h, D.n, C, (uint64_t)Lf.shape(0), m, prior));
    } else {
        Dense D = dense_of(Lf);
        std::vector<float> C(Rf.data(), Rf.data() + Rf.size());
        out = cost_block(D.dev->h, D.n, C, (uint64_t)Rf.shape(0), m, prior);
    }
    if (use_double) out = py::cast<py::array>(out.attr("astype")("float64"));
    if (one) return out.attr("item")(0);
    return out;
}

// ---- pcmp (pycmp.cpp:249-338): condensed upper triangle, out[(i, j)] = msr(first = row i,
// second = row j) ----------------------------------------------------------------------------
py::array pairwise(mcg_data *h, uint64_t n, int msr, double prior) {
    const py::ssize_t np = (py::ssize_t)(n > 1 ? n * (n - 1) / 2 : 0);
    py::array_t<float> out(np);
    if (np) check(mcg_pairwise(h, msr, prior, out.mutable_data()));
    return out;
}

py::array pcmp_csr(PyCSparseMatrix &lhs, py::object msr, py::object betaprior, py::ssize_t) {
    return pairwise(lhs.device(), lhs.nr, measure_code(msr), betaprior.cast<double>());
}

// dense: float32 results, returned as float64 when use_float is 0 or the input is
// wider than 4 bytes and use_float < 0 (the reference then computes in double)
py::array pcmp_dense(py::array mat, py::object msr, py::object betaprior, py::ssize_t use_float) {
    if (mat.ndim() != 2) throw std::invalid_argument("pcmp expects a 2-d numpy matrix");
    const bool luf = use_float < 0 ? mat.itemsize() <= 4 : use_float != 0;
    Dense D = dense_of(mat);
    py::array out = pairwise(D.dev->h, D.n, measure_code(msr), betaprior.cast<double>());
    if (!luf) out = py::cast<py::array>(out.attr("astype")("float64"));
    return out;
}

// ---- kmeanspp (smw.h:179-410) --------------------------------------------------------------
py::object asn_array(const std::vector<uint32_t> &a, uint64_t k) {
    const uint64_t n = a.size();
    if (k <= 256) { py::array_t<uint8_t> r(n); for (uint64_t i = 0; i < n; ++i) r.mutable_data()[i] = (uint8_t)a[i]; return r; }
    if (k <= 63356) { py::array_t<uint16_t> r(n); for (uint64_t i = 0; i < n; ++i) r.mutable_data()[i] = (uint16_t)a[i]; return r; }
    py::array_t<uint32_t> r(n);
    std::memcpy(r.mutable_data(), a.data(), n * 4);
    return r;
}

struct PPResult {
    std::vector<uint32_t> centers, asn;
    std::vector<float> costs;
    double total = 0.;
};

// D^2 seeding, then lspp rounds of local search++ (kmeans.h:209-212)
PPResult run_pp(mcg_data *h, uint64_t n, int m, uint64_t k, double prior, uint64_t seed, const double *w,
                uint64_t lspp, uint64_t trials) {
    PPResult r;
    const uint64_t kk = std::min(k, n);
    r.centers.resize(kk);
    r.asn.resize(n);
    r.costs.resize(n);
    if (trials > 1)
        check(mcg_kmeanspp_local(h, m, prior, k, seed, w, trials, lspp, r.centers.data(), r.asn.data(),
                                 r.costs.data()));
    else if (lspp)
        check(mcg_kmeanspp_ls(h, m, prior, k, seed, w, lspp, r.centers.data(), r.asn.data(), r.costs.data()));
    else
        check(mcg_kmeanspp(h, m, prior, k, seed, w, r.centers.data(), r.asn.data(), r.costs.data()));
    for (float c : r.costs) r.total += c;
    return r;
}

// n_local_trials <= 1 is the single-sample path (kmeans.h:114-186 tests n_local_samples <= 1)
uint64_t check_pp_args(py::ssize_t lspp, bool expskips, py::ssize_t n_local_trials) {
    if (lspp < 0) throw py::value_error("lspp must be >= 0");
    (void)expskips;  // sampling method choice; the GPU path has one exact sampler
    return n_local_trials > 1 ? (uint64_t)n_local_trials : 1;
}

uint64_t pp_seed(uint64_t seed) { return seed ? seed : std::mt19937_64(std::rand())(); }

py::object kmeanspp_dense(py::array arr, int k, py::object msr, py::object prior, py::object seed, py::object ntimes,
                          py::ssize_t lspp, py::object weights, bool expskips, py::ssize_t n_local_trials) {
    const uint64_t trials = check_pp_args(lspp, expskips, n_local_trials);
    const int m = measure_code(msr);
    Dense D = dense_of(arr);
    double pr = prior.cast<double>();
    if (pr < 0.) pr = 1. / (double)D.d;
    uint64_t s = pp_seed(seed.cast<uint64_t>());
    auto w = weights_f64(weights, D.n);
    PPResult best = run_pp(D.dev->h, D.n, m, (uint64_t)k, pr, s, wptr(w), (uint64_t)lspp, trials);
    const unsigned nt = ntimes.cast<unsigned>();
    for (unsigned t = 0; t < nt; ++t) {  // extra runs keep the cheapest (smw.h:385-393)
        PPResult r = run_pp(D.dev->h, D.n, m, (uint64_t)k, pr, s + 0x9e3779b97f4a7c15ULL * (t + 1), wptr(w),
                            (uint64_t)lspp, trials);
        if (r.total < best.total) best = std::move(r);
    }
    py::array_t<uint32_t> ctr(best.centers.size());
    std::memcpy(ctr.mutable_data(), best.centers.data(), best.centers.size() * 4);
    py::array_t<float> costs(D.n);
    std::memcpy(costs.mutable_data(), best.costs.data(), D.n * 4);
    return py::make_tuple(ctr, asn_array(best.asn, (uint64_t)k), costs);
}

template <class M>
py::object kmeanspp_csr(M &smw, py::object msr, py::int_ k, double prior, uint64_t seed, unsigned,
                        py::ssize_t lspp, bool expskips, py::ssize_t n_local_trials, py::object weights) {
    // Extra runs (ntimes) never replace the first in the reference (smw.h:248-256 compares the
    // first run's cost with itself), so the first run is the result.
    const uint64_t trials = check_pp_args(lspp, expskips, n_local_trials);
    const int m = measure_code(msr);
    if (prior < 0.) prior = 1. / (double)smw.nc;
    auto w = weights_f64(weights, smw.nr);
    const uint64_t kk = k.cast<uint64_t>();
    PPResult r = run_pp(smw.device(), smw.nr, m, kk, prior, pp_seed(seed), wptr(w), (uint64_t)lspp, trials);
    py::array_t<uint32_t> ctr(r.centers.size());
    std::memcpy(ctr.mutable_data(), r.centers.data(), r.centers.size() * 4);
    py::array_t<double> costs(smw.nr);
    for (uint64_t i = 0; i < smw.nr; ++i) costs.mutable_data()[i] = r.costs[i];
    return py::make_tuple(ctr, asn_array(r.asn, kk), costs);
}

// ---- CoresetSampler (coreset.h:228-701, pycs.cpp) ----------------------------------------
struct CoresetSampler {
    std::vector<float> probs;
    std::vector<double> weights;  // empty if unweighted
    uint64_t np = 0, b = 0, seed = 137;
    int sens = 0;
    bool ready = false;

    CoresetSampler() = default;
    explicit CoresetSampler(const std::string &path) { read(path); }

    void make_sampler(size_t ncenters, py::array costs, py::array assignments, py::object w, uint64_t seed_,
                      int sens_) {
        if (costs.ndim() != 1) throw std::invalid_argument("costs must be 1-dimensional");
        if (costs.dtype().kind() != 'f') throw std::invalid_argument("Costs is not double or float");
        auto c = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(costs);
        auto a = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>::ensure(assignments);
        np = (uint64_t)c.size();
        if ((uint64_t)a.size() != np) throw std::invalid_argument("assignments must match costs");
        weights.clear();
        if (!w.is_none()) {
            auto wa = py::cast<py::array>(w);
            if (wa.ndim() != 1) throw std::invalid_argument("weightss must be 1-dimensional");
            if (wa.dtype().kind() != 'f') throw std::invalid_argument("Weights can only be double or float");
            auto wd = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(wa);
            weights.assign(wd.data(), wd.data() + wd.size());
        }
        sens = sens_;
        seed = seed_;
        b = ncenters;
        probs.resize(np);
        check(mcg_coreset_probs(ctx(), sens, np, ncenters, c.data(), a.data(), weights.empty() ? nullptr : weights.data(),
                                probs.data()));
        ready = true;
    }

    // CoresetSampler::write / read (coreset.h:267-338): the gzip stream of n, seed,
    // probs, the weights flag and the weights, all f32 as in CoresetSampler<float, ...>
    void write(const std::string &path) const {
        if (!ready) throw std::invalid_argument("Can't write without created sampler. Call make_sampler");
        if (sens == MCG_SENS_FL) throw std::runtime_error("Not implemented: serialization for FL coreset samplers");
        std::vector<float> wf(weights.begin(), weights.end());
        check(mcg_sampler_write(path.c_str(), np, seed, probs.data(), weights.empty() ? nullptr : wf.data()));
    }
    void read(const std::string &path) {
        uint64_t n = 0, sd = 0;
        int hw = 0;
        check(mcg_sampler_read(path.c_str(), &n, &sd, &hw, nullptr, nullptr, 0));
        std::vector<float> p(n), wf(hw ? n : 0);
        check(mcg_sampler_read(path.c_str(), &n, &sd, &hw, p.data(), hw ? wf.data() : nullptr, n));
        probs.swap(p);
        weights.assign(wf.begin(), wf.end());
        np = n;
        seed = sd;
        sens = -1;  // not stored in the stream; the sampler draws like the non-FL methods
        b = 0;
        ready = true;
    }

    py::array_t<float> get_probs() const {
        py::array_t<float> r(probs.size());
        std::memcpy(r.mutable_data(), probs.data(), probs.size() * 4);
        return r;
    }

    py::tuple sample(py::ssize_t size, py::ssize_t seed, bool unique, double) {
        if (!ready) throw std::invalid_argument("Can't sample without created sampler. Call make_ampler");
        if (seed == 0) seed = std::rand();
        const uint64_t n = (uint64_t)size;
        // FL draws n - b points directly (coreset.h:637)
        const uint64_t direct = sens == MCG_SENS_FL ? (n > b ? n - b : 0) : n;
        py::array_t<float> rf(size);
        py::array_t<uint64_t> ri(size);
        std::fill(rf.mutable_data(), rf.mutable_data() + size, 0.f);
        std::fill(ri.mutable_data(), ri.mutable_data() + size, 0ull);
        const double *W = weights.empty() ? nullptr : weights.data();
        if (!unique) {
            if (direct)
                check(mcg_coreset_sample(ctx(), np, probs.data(), W, direct, (uint64_t)seed, 0, ri.mutable_data(),
                                         rf.mutable_data()));
            return py::make_tuple(rf, ri);
        }
        if (direct > np) throw std::invalid_argument("cannot draw more unique points than there are points");
        // draw until `direct` distinct indices (coreset.h:640-672); weight = count * w_i / (n p_i)
        std::map<uint64_t, uint32_t> ctr;
        uint64_t drawn = 0;
        std::vector<uint64_t> idx;
        std::vector<float> wt;
        while (ctr.size() < direct) {
            const uint64_t batch = std::max<uint64_t>(1024, 2 * (direct - ctr.size()));
            idx.resize(batch);
            wt.resize(batch);
            check(mcg_coreset_sample(ctx(), np, probs.data(), W, batch, (uint64_t)seed, drawn, idx.data(), wt.data()));
            for (uint64_t t = 0; t < batch && ctr.size() < direct; ++t) ++ctr[idx[t]];
            drawn += batch;
        }
        uint64_t q = 0;
        for (auto &kv : ctr) {
            const double wi = W ? (double)(float)W[kv.first] : 1.;
            ri.mutable_data()[q] = kv.first;
            rf.mutable_data()[q] = (float)((double)kv.second * (wi / ((double)n * (double)probs[kv.first])));
            ++q;
        }
        return py::make_tuple(rf, ri);
    }
};

// ---- SparseMatrixWrapper (python/smw.h:10-120, pysmw.cpp:31-265): an owned CSR copy -------
// Rows keep their entries sorted by column (csc2sparse, csc.h:1221-1261); skip_empty drops
// rows without entries; values are held as float32 (use_float) or float64.  The device copy
// is built on first use (float64 values are converted to float32 on the device: the GPU path
// computes in FP32).  blaze archives (tofile / fromfile / a path argument) need blaze, which
// is not part of this build: they raise.
struct SparseMatrixWrapper {
    bool f32 = true;
    std::vector<float> vf;
    std::vector<double> vd;
    std::vector<uint32_t> idx;
    std::vector<uint64_t> ptr{0};
    uint64_t nr = 0, nc = 0;
    DataPtr dev;

    SparseMatrixWrapper() = default;
    SparseMatrixWrapper(py::object sp, bool skip_empty, bool use_float) : f32(use_float) {
        if (py::isinstance<py::str>(sp) || !py::hasattr(sp, "indices"))
            throw std::runtime_error("SparseMatrixWrapper from a blaze archive needs blaze, which this build does not "
                                     "include; construct it from a scipy.sparse CSR matrix");
        auto ip = py::array_t<int64_t, py::array::c_style | py::array::forcecast>::ensure(sp.attr("indptr"));
        auto ix = py::array_t<int64_t, py::array::c_style | py::array::forcecast>::ensure(sp.attr("indices"));
        auto dv = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(sp.attr("data"));
        auto shape = py::cast<py::sequence>(sp.attr("shape"));
        const uint64_t rows = shape[0].cast<uint64_t>();
        nc = shape[1].cast<uint64_t>();
        if (!ip || !ix || !dv || (uint64_t)ip.size() != rows + 1)
            throw std::invalid_argument("expected a CSR matrix with indptr of rows + 1 entries");
        const int64_t *P = ip.data(), *J = ix.data();
        const double *V = dv.data();
        std::vector<std::pair<uint32_t, double>> row;
        ptr.assign(1, 0);
        for (uint64_t i = 0; i < rows; ++i) {
            if (P[i] < 0 || P[i + 1] < P[i] || P[i + 1] > (int64_t)dv.size())
                throw std::invalid_argument("indptr out of range or decreasing");
            if (skip_empty && P[i + 1] == P[i]) continue;
            row.clear();
            for (int64_t q = P[i]; q < P[i + 1]; ++q) {
                if (J[q] < 0 || (uint64_t)J[q] >= nc) throw std::invalid_argument("column index out of range");
                row.emplace_back((uint32_t)J[q], V[q]);
            }
            std::stable_sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
            for (auto &e : row) {
                idx.push_back(e.first);
                if (f32) vf.push_back((float)e.second); else vd.push_back(e.second);
            }
            ptr.push_back(idx.size());
        }
        nr = ptr.size() - 1;
    }
    uint64_t nnz() const { return idx.size(); }
    double value(uint64_t q) const { return f32 ? (double)vf[q] : vd[q]; }
    mcg_data *device() {
        if (!dev) {
            if (nr == 0) throw std::invalid_argument("empty matrix");
            auto h = std::make_shared<DataHandle>();
            const void *data = f32 ? (const void *)vf.data() : (const void *)vd.data();
            static const float zero = 0.f;
            check(mcg_data_csr(ctx(), nnz() ? data : &zero, f32 ? MCG_F32 : MCG_F64, nnz() ? idx.data() : (const uint32_t *)&zero,
                               MCG_U32, ptr.data(), MCG_U64, nr, nc, nnz(), &h->h));
            dev = h;
        }
        return dev->h;
    }
    std::vector<float> dense_rows(const std::vector<int64_t> &rows) const {
        std::vector<float> out(rows.size() * nc, 0.f);
        for (size_t q = 0; q < rows.size(); ++q) {
            const int64_t r = rows[q];
            if (r < 0 || (uint64_t)r >= nr) throw std::invalid_argument("center index out of range");
            for (uint64_t e = ptr[r]; e < ptr[r + 1]; ++e) out[q * nc + idx[e]] = (float)value(e);
        }
        return out;
    }
    std::vector<float> dense_all() const {
        std::vector<int64_t> rows(nr);
        for (uint64_t i = 0; i < nr; ++i) rows[i] = (int64_t)i;
        return dense_rows(rows);
    }
    SparseMatrixWrapper select(const std::vector<int64_t> *rows, const std::vector<int64_t> *cols) const {
        SparseMatrixWrapper o;
        o.f32 = f32;
        o.nc = cols ? cols->size() : nc;
        std::vector<int64_t> cmap;
        if (cols) {
            cmap.assign(nc, -1);
            for (size_t q = 0; q < cols->size(); ++q) {
                const int64_t c = (*cols)[q];
                if (c < 0 || (uint64_t)c >= nc) throw std::invalid_argument("column index out of range");
                cmap[c] = (int64_t)q;
            }
        }
        const uint64_t m = rows ? rows->size() : nr;
        std::vector<std::pair<uint32_t, double>> row;
        for (uint64_t q = 0; q < m; ++q) {
            const int64_t r = rows ? (*rows)[q] : (int64_t)q;
            if (r < 0 || (uint64_t)r >= nr) throw std::invalid_argument("row index out of range");
            row.clear();
            for (uint64_t e = ptr[r]; e < ptr[r + 1]; ++e) {
                if (!cols) row.emplace_back(idx[e], value(e));
                else if (cmap[idx[e]] >= 0) row.emplace_back((uint32_t)cmap[idx[e]], value(e));
            }
            std::stable_sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
            for (auto &pr : row) {
                o.idx.push_back(pr.first);
                if (f32) o.vf.push_back((float)pr.second); else o.vd.push_back(pr.second);
            }
            o.ptr.push_back(o.idx.size());
        }
        o.nr = m;
        return o;
    }
    void transpose() {
        SparseMatrixWrapper o;
        o.f32 = f32;
        o.nr = nc;
        o.nc = nr;
        std::vector<uint64_t> cnt(nc + 1, 0);
        for (auto j : idx) cnt[j + 1]++;
        for (uint64_t j = 0; j < nc; ++j) cnt[j + 1] += cnt[j];
        o.ptr = cnt;
        o.idx.resize(nnz());
        if (f32) o.vf.resize(nnz()); else o.vd.resize(nnz());
        for (uint64_t i = 0; i < nr; ++i)
            for (uint64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
                const uint64_t q = cnt[idx[e]]++;
                o.idx[q] = (uint32_t)i;
                if (f32) o.vf[q] = vf[e]; else o.vd[q] = vd[e];
            }
        *this = std::move(o);
    }
    std::string describe() const {
        return "Matrix of " + std::to_string(nr) + "/" + std::to_string(nc) + " elements of " +
               (f32 ? "float32" : "double") + ", " + std::to_string(nnz()) + " nonzeros";
    }
};

std::vector<int64_t> int_list(py::object o) {
    auto a = py::array_t<int64_t, py::array::c_style | py::array::forcecast>::ensure(o);
    if (!a) throw std::invalid_argument("Wrong dtype");
    return std::vector<int64_t>(a.data(), a.data() + a.size());
}

// out[l][r] = msr(first = rhs row r, second = lhs row l), or with reverse msr(first = lhs
// row l, second = rhs row r) (pycmp.cpp:100-146); float32 unless use_float is 0 (or -1 with
// both matrices double), then returned as float64 (the GPU computes in FP32)
py::object cmp_smw_smw(SparseMatrixWrapper &lhs, SparseMatrixWrapper &rhs, py::object msr, py::object betaprior,
                       bool reverse, int use_float) {
    const int m = measure_code(msr);
    if (use_float < 0) use_float = lhs.f32 || rhs.f32;
    if (lhs.nc != rhs.nc) throw std::invalid_argument("mismatched # columns");
    const double prior = betaprior.cast<double>();
    py::array out = reverse ? py::array(cost_block(lhs.device(), lhs.nr, rhs.dense_all(), rhs.nr, m, prior))
                            : transpose_copy(cost_block(rhs.device(), rhs.nr, lhs.dense_all(), lhs.nr, m, prior));
    if (!use_float) out = py::cast<py::array>(out.attr("astype")("float64"));
    return out;
}

// ---- SumOpts (clustering/mtx2cs.h:13-80, pysmw.cpp:293-366): the seeding options the kmeanspp
// overloads read (measure, k, prior value, extra sample tries, lspp, local trials, seed) --------
struct SumOpts {
    int dis = MCG_JSD;
    int prior_kind = 0;           // NONE, DIRICHLET, GAMMA_BETA, FEATURE_SPECIFIC_PRIOR
    double gamma = 1.;
    unsigned k = 10;
    uint64_t seed = 0;
    unsigned extra_sample_tries = 1;
    unsigned lloyd_max_rounds = 1000;
    int lspp = 0;
    int sm = MCG_SENS_BFL;
    bool soft = false, discrete_metric_search = false, use_exponential_skips = false;
    double outlier_fraction = 0.;
    size_t n_local_trials = 1;

    static int sm_of(const std::string &s) {
        static const std::map<std::string, int> names = {{"BFL", MCG_SENS_BFL}, {"FL", MCG_SENS_FL},
                                                         {"LFKF", MCG_SENS_LFKF}, {"VX", MCG_SENS_VX},
                                                         {"LBK", MCG_SENS_LBK}};
        auto it = names.find(s);
        return it == names.end() ? MCG_SENS_BFL : it->second;   // str2sm falls back to BFL
    }
    SumOpts(py::object measure, py::ssize_t k_, double prior_value, py::object sm_, double outlier, size_t rounds,
            bool soft_)
        : dis(measure_code(measure)), prior_kind(prior_value == 0. ? 0 : prior_value == 1. ? 1 : 2), gamma(prior_value),
          k((unsigned)k_), lloyd_max_rounds((unsigned)rounds), soft(soft_), outlier_fraction(outlier) {
        sm = py::isinstance<py::str>(sm_) ? sm_of(sm_.cast<std::string>()) : sm_.cast<int>();
    }
    static const char *sm_name(int v) {
        switch (v) {
        case MCG_SENS_BFL: return "BFL";
        case MCG_SENS_VX: return "VX";
        case MCG_SENS_LFKF: return "LFKF";
        case MCG_SENS_LBK: return "LBK";
        case MCG_SENS_FL: return "FL";
        }
        return "UNKNOWN";
    }
    std::string to_string() const {
        char buf[512];
        std::snprintf(buf, sizeof(buf),
                      "K: %u. Construction method: %d. Seed:%zu. gamma: %g. prior: %d. extrasampletries: %u. dis: %d. "
                      "%s clustering.", k, sm, (size_t)seed, gamma, prior_kind, extra_sample_tries, dis,
                      soft ? "soft" : "hard");
        return buf;
    }
};

// kmeanspp(matrix, opts, weights) (pysmw.cpp:374-381, pycsparse.cpp:91-99): the option
// fields passed to the keyword form, with ntimes = max(extra_sample_tries - 1, 0)
template <class M>
py::object kmeanspp_opts(M &smw, const SumOpts &so, py::object weights) {
    return kmeanspp_csr(smw, py::int_(so.dis), py::int_(so.k), so.gamma, so.seed,
                        (unsigned)std::max<int>((int)so.extra_sample_tries - 1, 0), so.lspp, so.use_exponential_skips,
                        (py::ssize_t)so.n_local_trials, weights);
}

}  // namespace

PYBIND11_MODULE(_pyminicore, m) {
    m.doc() = "pyminicore entry points of the clustering hot path on MI355X (libmcg)";

    py::class_<PyCSparseMatrix>(m, "CSparseMatrix")
        .def(py::init<py::object>(), py::arg("sparray"))
        .def("__str__", [](const PyCSparseMatrix &x) {
            return "CSparseMatrix, " + std::to_string(x.nr) + "x" + std::to_string(x.nc) + ", " + std::to_string(x.nnz);
        })
        .def("columns", [](const PyCSparseMatrix &x) { return x.nc; })
        .def("rows", [](const PyCSparseMatrix &x) { return x.nr; })
        .def("nnz", [](const PyCSparseMatrix &x) { return x.nnz; })
        .def_static("from_items", [](py::array data, py::array idx, py::array indptr, py::sequence shape) {
            return PyCSparseMatrix(data, idx, indptr, shape[0].cast<uint64_t>(), shape[1].cast<uint64_t>(),
                                   (uint64_t)data.size());
        }, py::arg("data"), py::arg("indices"), py::arg("indptr"), py::arg("shape"));

    py::class_<SparseMatrixWrapper>(m, "SparseMatrixWrapper")
        .def(py::init<>())
        .def(py::init<py::object, bool, bool>(), py::arg("sparray"), py::arg("skip_empty") = false,
             py::arg("use_float") = true)
        .def("is_float", [](const SparseMatrixWrapper &x) { return x.f32; })
        .def("is_double", [](const SparseMatrixWrapper &x) { return !x.f32; })
        .def("rows", [](const SparseMatrixWrapper &x) { return x.nr; })
        .def("columns", [](const SparseMatrixWrapper &x) { return x.nc; })
        .def("nonzeros", [](const SparseMatrixWrapper &x) { return x.nnz(); })
        .def("__str__", &SparseMatrixWrapper::describe)
        .def("__repr__", &SparseMatrixWrapper::describe)
        .def("tocsr", [](const SparseMatrixWrapper &x) {   // pysmw.cpp:40-62
            py::array_t<py::ssize_t> indptr(x.nr + 1), shape(2);
            py::array_t<int32_t> indices(x.nnz());
            py::array_t<double> data(x.nnz());
            for (uint64_t i = 0; i <= x.nr; ++i) indptr.mutable_data()[i] = (py::ssize_t)x.ptr[i];
            for (uint64_t q = 0; q < x.nnz(); ++q) {
                indices.mutable_data()[q] = (int32_t)x.idx[q];
                data.mutable_data()[q] = x.value(q);
            }
            shape.mutable_data()[0] = (py::ssize_t)x.nr;
            shape.mutable_data()[1] = (py::ssize_t)x.nc;
            return py::make_tuple(py::make_tuple(data, indices, indptr), shape);
        })
        .def("rowsel", [](const SparseMatrixWrapper &x, py::array idx) -> py::object {   // pysmw.cpp:87-123
            if (idx.dtype().kind() == 'f') throw std::invalid_argument("Unexpected type");
            auto rows = int_list(idx);
            auto v = x.dense_rows(rows);
            if (x.f32) {
                py::array_t<float> out({(py::ssize_t)rows.size(), (py::ssize_t)x.nc});
                std::memcpy(out.mutable_data(), v.data(), v.size() * 4);
                return out;
            }
            py::array_t<double> out({(py::ssize_t)rows.size(), (py::ssize_t)x.nc});
            std::fill(out.mutable_data(), out.mutable_data() + out.size(), 0.);
            for (size_t q = 0; q < rows.size(); ++q)
                for (uint64_t e = x.ptr[rows[q]]; e < x.ptr[rows[q] + 1]; ++e)
                    out.mutable_data()[q * x.nc + x.idx[e]] = x.vd[e];
            return out;
        }, py::arg("idx"))
        .def("submatrix", [](const SparseMatrixWrapper &x, py::object rowsel, py::object columnsel) {   // :127-169
            std::vector<int64_t> r, c;
            if (!rowsel.is_none()) r = int_list(rowsel);
            if (!columnsel.is_none()) c = int_list(columnsel);
            return x.select(rowsel.is_none() ? nullptr : &r, columnsel.is_none() ? nullptr : &c);
        }, py::arg("rowsel") = py::none(), py::arg("columnsel") = py::none())
        .def("transpose_", [](SparseMatrixWrapper &x) { x.transpose(); })
        .def("sum", [](const SparseMatrixWrapper &x, int byrow, bool usefloat) -> py::object {   // :204-237
            if (byrow < -1 || byrow > 1) throw std::invalid_argument("byrow must be -1 (total sum), 0 (by column) or by row (1)");
            std::vector<double> acc(byrow == 1 ? x.nr : byrow == 0 ? x.nc : 1, 0.);
            for (uint64_t i = 0; i < x.nr; ++i)
                for (uint64_t e = x.ptr[i]; e < x.ptr[i + 1]; ++e)
                    acc[byrow == 1 ? i : byrow == 0 ? x.idx[e] : 0] += x.value(e);
            if (byrow == -1) return py::float_(acc[0]);
            if (usefloat) {
                py::array_t<float> out(acc.size());
                for (size_t q = 0; q < acc.size(); ++q) out.mutable_data()[q] = (float)acc[q];
                return out;
            }
            py::array_t<double> out(acc.size());
            std::memcpy(out.mutable_data(), acc.data(), acc.size() * 8);
            return out;
        }, py::arg("kind") = -1, py::arg("usefloat") = true)
        .def("count_nnz", [](const SparseMatrixWrapper &x, int byrow) -> py::object {   // :238-265
            if (byrow <= 0) return py::int_(x.nnz());
            py::array_t<uint64_t> out(x.nr);
            for (uint64_t i = 0; i < x.nr; ++i) out.mutable_data()[i] = x.ptr[i + 1] - x.ptr[i];
            const char *dt = x.nc <= 0xFFu ? "uint8" : x.nc <= 0xFFFFu ? "uint16" : x.nc <= 0xFFFFFFFFu ? "uint32" : "uint64";
            return out.attr("astype")(dt);
        }, py::arg("byrow") = 0)
        .def("tofile", [](const SparseMatrixWrapper &, std::string) {
            throw std::runtime_error("SparseMatrixWrapper.tofile writes a blaze archive; blaze is not part of this build");
        }, py::arg("path"))
        .def("fromfile", [](SparseMatrixWrapper &, std::string) {
            throw std::runtime_error("SparseMatrixWrapper.fromfile reads a blaze archive; blaze is not part of this build");
        }, py::arg("path"));

    py::class_<SumOpts>(m, "SumOpts")
        .def(py::init<py::object, py::ssize_t, double, py::object, double, size_t, bool>(), py::arg("measure") = 0,
             py::arg("k") = 10, py::arg("prior") = 0., py::arg("sm") = "BFL", py::arg("outlier_fraction") = 0.,
             py::arg("max_rounds") = 100, py::arg("soft") = false)
        .def("__str__", &SumOpts::to_string)
        .def("__repr__", &SumOpts::to_string)
        .def_readwrite("lspp", &SumOpts::lspp)
        .def_readwrite("gamma", &SumOpts::gamma)
        .def_readwrite("k", &SumOpts::k)
        .def_readwrite("search_max_rounds", &SumOpts::lloyd_max_rounds)
        .def_readwrite("extra_sample_rounds", &SumOpts::extra_sample_tries)
        .def_readwrite("soft", &SumOpts::soft)
        .def_readwrite("outlier_fraction", &SumOpts::outlier_fraction)
        .def_readwrite("discrete_metric_search", &SumOpts::discrete_metric_search)
        .def_readwrite("use_exponential_skips", &SumOpts::use_exponential_skips)
        .def_property("cs", [](const SumOpts &o) { return std::string(SumOpts::sm_name(o.sm)); },
                      [](SumOpts &o, py::object v) {
                          if (py::isinstance<py::str>(v)) o.sm = SumOpts::sm_of(v.cast<std::string>());
                          else if (py::isinstance<py::int_>(v)) o.sm = v.cast<int>();
                          else throw std::invalid_argument("value must be str or int");
                      })
        .def_property("prior", [](const SumOpts &o) -> std::string {
            switch (o.prior_kind) {
            case 0: return "NONE";
            case 1: return "DIRICHLET";
            case 2: return "GAMMA";
            case 3: return "FSP";
            }
            throw std::invalid_argument("Invalid prior");
        }, [](SumOpts &o, py::object v) {
            if (v.is_none()) { o.prior_kind = 0; return; }
            if (py::isinstance<py::str>(v)) {
                static const std::map<std::string, int> names = {{"NONE", 0}, {"DIRICHLET", 1}, {"GAMMA", 2},
                                                                 {"GAMMA_BETA", 2}, {"FSP", 3},
                                                                 {"FEATURE_SPECIFIC_PRIOR", 3}};
                auto it = names.find(v.cast<std::string>());
                if (it == names.end())
                    throw std::out_of_range("Prior must be NONE, FSP, FEATURE_SPECIFIC_PRIOR, GAMMA, or DIRICHLET");
                o.prior_kind = it->second;
            } else {
                const auto x = v.cast<py::ssize_t>();
                if (x > 3) throw std::out_of_range("x must be <= 3 if an integer, to represent various priors");
                o.prior_kind = (int)x;
            }
        });

    // SparseMatrixWrapper overloads first: a py::array parameter would also accept the object
    m.def("hcluster", &hcluster_csr<SparseMatrixWrapper>, py::arg("smw"), py::arg("centers"), py::arg("prior") = 0.,
          py::arg("msr") = 2, py::arg("weights") = py::none(), py::arg("eps") = 1e-10, py::arg("maxiter") = 100,
          py::arg("seed") = 0, py::arg("mbsize") = py::ssize_t(-1), py::arg("ncheckins") = py::ssize_t(-1),
          py::arg("reseed_count") = py::ssize_t(5), py::arg("with_rep") = false, py::arg("cs") = false);
    m.def("scluster", &scluster_csr<SparseMatrixWrapper>, py::arg("smw"), py::arg("centers"), py::arg("msr") = 2,
          py::arg("prior") = 0., py::arg("temp") = 1., py::arg("maxiter") = 1000, py::arg("mbsize") = py::ssize_t(-1),
          py::arg("mbn") = py::ssize_t(-1), py::arg("savepref") = "", py::arg("weights") = py::none());
    m.def("cmp", [](SparseMatrixWrapper &lhs, py::array arr, py::object msr, py::object prior, py::object reverse) {
        return cmp_csr_array(lhs, arr, msr, prior, reverse, true);
    }, py::arg("matrix"), py::arg("data"), py::arg("msr") = 2, py::arg("prior") = 0., py::arg("reverse") = false);
    m.def("cmp", &cmp_smw_smw, py::arg("matrix"), py::arg("data"), py::arg("msr") = 2, py::arg("prior") = 0.,
          py::arg("reverse") = false, py::arg("use_float") = -1);
    m.def("kmeanspp", &kmeanspp_csr<SparseMatrixWrapper>, py::arg("smw"), py::arg("msr"), py::arg("k"),
          py::arg("prior") = 0., py::arg("seed") = 0, py::arg("ntimes") = 1, py::arg("lspp") = 0,
          py::arg("expskips") = false, py::arg("n_local_trials") = 1, py::arg("weights") = py::none());
    m.def("kmeanspp", &kmeanspp_opts<SparseMatrixWrapper>, py::arg("smw"), py::arg("opts"),
          py::arg("weights") = py::none());
    m.def("kmeanspp", &kmeanspp_opts<PyCSparseMatrix>, py::arg("smw"), py::arg("opts"),
          py::arg("weights") = py::none());

    m.def("hcluster", &hcluster_csr<PyCSparseMatrix>, py::arg("smw"), py::arg("centers"), py::arg("prior") = 0., py::arg("msr") = 2,
          py::arg("weights") = py::none(), py::arg("eps") = 1e-10, py::arg("maxiter") = 100, py::arg("seed") = 0,
          py::arg("mbsize") = py::ssize_t(-1), py::arg("ncheckins") = py::ssize_t(-1),
          py::arg("reseed_count") = py::ssize_t(5), py::arg("with_rep") = false, py::arg("cs") = false);
    m.def("hcluster", &hcluster_dense, py::arg("dataset"), py::arg("centers"), py::arg("prior") = 0., py::arg("msr") = 2,
          py::arg("weights") = py::none(), py::arg("eps") = 1e-10, py::arg("maxiter") = 100, py::arg("seed") = 0,
          py::arg("mbsize") = py::ssize_t(-1), py::arg("ncheckins") = py::ssize_t(-1),
          py::arg("reseed_count") = py::ssize_t(5), py::arg("with_rep") = false, py::arg("cs") = false,
          "Clusters the rows using the centers provided; Lloyd / EM until convergence");

    m.def("scluster", &scluster_csr<PyCSparseMatrix>, py::arg("smw"), py::arg("centers"), py::arg("msr") = 2, py::arg("prior") = 0.,
          py::arg("temp") = 1., py::arg("maxiter") = 1000, py::arg("mbsize") = py::ssize_t(-1),
          py::arg("mbn") = py::ssize_t(-1), py::arg("savepref") = "", py::arg("weights") = py::none());
    m.def("scluster", &scluster_dense, py::arg("matrix"), py::arg("centers"), py::arg("msr") = 2, py::arg("prior") = 0.,
          py::arg("temp") = 1., py::arg("maxiter") = 1000, py::arg("mbsize") = py::ssize_t(-1),
          py::arg("mbn") = py::ssize_t(-1), py::arg("savepref") = "", py::arg("weights") = py::none());

    m.def("cmp", [](PyCSparseMatrix &lhs, py::array arr, py::object msr, py::object prior, py::object reverse) {
        return cmp_csr_array(lhs, arr, msr, prior, reverse);
    }, py::arg("matrix"), py::arg("data"), py::arg("msr") = 2, py::arg("prior") = 0.,
          py::arg("reverse") = false);
    m.def("cmp", &cmp_csr_csr, py::arg("matrix"), py::arg("data"), py::arg("msr") = 2, py::arg("prior") = 0.);
    m.def("cmp", &cmp_arrays, py::arg("lhs"), py::arg("rhs"), py::arg("msr") = 2, py::arg("prior") = 0.,
          py::arg("reverse") = true, py::arg("use_double") = -1);

    m.def("pcmp", &pcmp_csr, py::arg("matrix"), py::arg("msr") = 2, py::arg("prior") = 0.,
          py::arg("use_float") = py::ssize_t(-1));
    m.def("pcmp", &pcmp_dense, py::arg("matrix"), py::arg("msr") = 2, py::arg("prior") = 0.,
          py::arg("use_float") = py::ssize_t(-1));

    m.def("kmeanspp", &kmeanspp_dense, py::arg("matrix"), py::arg("k"), py::arg("msr") = py::int_(2),
          py::arg("prior") = 0., py::arg("seed") = 0, py::arg("ntimes") = 0, py::arg("lspp") = 0,
          py::arg("weights") = py::none(), py::arg("expskips") = false, py::arg("n_local_trials") = 1);
    m.def("kmeanspp", &kmeanspp_csr<PyCSparseMatrix>, py::arg("smw"), py::arg("msr"), py::arg("k"), py::arg("prior") = 0.,
          py::arg("seed") = 0, py::arg("ntimes") = 1, py::arg("lspp") = 0, py::arg("expskips") = false,
          py::arg("n_local_trials") = 1, py::arg("weights") = py::none());

    py::class_<CoresetSampler>(m, "CoresetSampler")
        .def(py::init<>())
        .def(py::init<const std::string &>(), py::arg("path"))
        .def("write", &CoresetSampler::write, py::arg("path"))
        .def("read", &CoresetSampler::read, py::arg("path"))
        .def("make_sampler", &CoresetSampler::make_sampler, py::arg("ncenters"), py::arg("costs"),
             py::arg("assignments"), py::arg("weights") = py::none(), py::arg("seed") = 13, py::arg("sens") = 0)
        .def("get_probs", &CoresetSampler::get_probs)
        .def("sample", &CoresetSampler::sample, py::arg("size"), py::arg("seed") = 0, py::arg("unique") = false,
             py::arg("eps") = 0.1);

    m.def("set_sqrl2_precision", [](const std::string &p) {
        if (p == "bf16") g_sqrl2_precision = MCG_PREC_BF16;
        else if (p == "fp16") g_sqrl2_precision = MCG_PREC_FP16;
        else if (p == "fp32") g_sqrl2_precision = MCG_PREC_FP32;
        else if (p == "exact") g_sqrl2_precision = MCG_PREC_EXACT;
        else throw std::invalid_argument("precision must be bf16, fp16, fp32 or exact");
    }, py::arg("precision"),
          "Cross-term form for dense squared-L2 clustering (default bf16; bf16 and fp16 give the exact argmin)");
}
