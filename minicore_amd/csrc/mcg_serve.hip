// Serving layer of libmcg: a pool of worker threads, each with its own context
// (HIP stream), that runs clustering requests of many callers concurrently on
// one GPU, plus the row-selection entry point the request pipeline needs.
//
// Request pipeline = one round of hmb_coreset_clustering
// (include/minicore/clustering/solve.h:716-762) started from D^2 seeding:
//   1. rows to the device, D^2 seeding on all rows (kmeans.h:49-215)
//   2. sensitivities from the seeding costs / assignments (coreset.h:386-578)
//   3. m weighted coreset draws (coreset.h:625-651), coreset rows gathered
//   4. weighted Lloyd on the coreset from the seeded centers (solve.h:98-175)
//   5. hard assignment of all rows to the final centers (solve.h:234-285)
// Every stage runs on the worker's own stream with deterministic kernels, so
// a request's result does not depend on what runs beside it.
#include "mcg_internal.h"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace mcg {

// ---------------------------------------------------------------------------
// row selection
// ---------------------------------------------------------------------------
// one wave per output row
// (vec: d % 4 == 0 and X 16-byte aligned, so every row is)
__global__ void k_rows_dense(const float *__restrict__ X, uint64_t d, const uint64_t *__restrict__ idx,
                             uint64_t m, int vec, float *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= m) return;
    const float *src = X + idx[i] * d;
    float *dst = out + i * d;
    if (vec) {
        const float4 *s4 = (const float4 *)src;
        float4 *d4 = (float4 *)dst;
        for (uint64_t j = lane; j < d / 4; j += 64) d4[j] = s4[j];
    } else {
        for (uint64_t j = lane; j < d; j += 64) dst[j] = src[j];
    }
}

__global__ void k_rows_len(const uint64_t *__restrict__ ptr, const uint64_t *__restrict__ idx, uint64_t m,
                           uint64_t *__restrict__ len) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) len[i] = ptr[idx[i] + 1] - ptr[idx[i]];
}

// one wave per output row
__global__ void k_rows_csr(const float *__restrict__ val, const uint32_t *__restrict__ col,
                           const uint64_t *__restrict__ ptr, const uint64_t *__restrict__ idx, uint64_t m,
                           const uint64_t *__restrict__ optr, float *__restrict__ oval,
                           uint32_t *__restrict__ ocol) {
    const uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= m) return;
    const uint64_t s = ptr[idx[i]], len = ptr[idx[i] + 1] - s, o = optr[i];
    for (uint64_t j = lane; j < len; j += 64) {
        oval[o + j] = val[s + j];
        ocol[o + j] = col[s + j];
    }
}

void rows_into(const Data &S, const uint64_t *idx_h, const uint64_t *idx_d, uint64_t m, Data &D, hipStream_t st) {
    reset_derived(D);
    D.ctx = S.ctx;
    D.csr = S.csr;
    D.n = m;
    D.d = S.d;
    if (!S.csr) {
        D.nnz = m * S.d;
        D.X = (float *)D.Xown.reserve(m * S.d * sizeof(float));
        const int vec = (S.d % 4 == 0) && ((uintptr_t)S.X % 16 == 0);
        hipLaunchKernelGGL(k_rows_dense, dim3((uint32_t)cdiv(m, 4)), dim3(256), 0, st, S.X, S.d, idx_d, m, vec, D.X);
        HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    DBuf &len = S.ctx->s[S_ROWLEN];
    len.reserve(m * 8);
    hipLaunchKernelGGL(k_rows_len, dim3((uint32_t)cdiv(m, 256)), dim3(256), 0, st, S.ptr.as<const uint64_t>(), idx_d,
                       m, len.as<uint64_t>());
    std::vector<uint64_t> P(m + 1);
    HIP_CHECK(hipMemcpyAsync(P.data() + 1, len.p, m * 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    P[0] = 0;
    for (uint64_t i = 0; i < m; ++i) P[i + 1] += P[i];
    (void)idx_h;
    D.nnz = P[m];
    D.ptr.reserve((m + 1) * 8);
    D.val.reserve(std::max<uint64_t>(D.nnz, 1) * 4);
    D.idx.reserve(std::max<uint64_t>(D.nnz, 1) * 4);
    HIP_CHECK(hipMemcpyAsync(D.ptr.p, P.data(), (m + 1) * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_rows_csr, dim3((uint32_t)cdiv(m, 4)), dim3(256), 0, st, S.val.as<const float>(),
                       S.idx.as<const uint32_t>(), S.ptr.as<const uint64_t>(), idx_d, m, D.ptr.as<const uint64_t>(),
                       D.val.as<float>(), D.idx.as<uint32_t>());
    HIP_CHECK(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// small kernels of the request pipeline
// ---------------------------------------------------------------------------
__global__ void k_f32_to_f64(const float *__restrict__ a, uint64_t n, double *__restrict__ b) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

// coreset row weights: w[idx[i]] * cw[i] (solve.h:759), or cw[i] without row weights
__global__ void k_coreset_weights(const double *__restrict__ w, const uint64_t *__restrict__ idx,
                                  const float *__restrict__ cw, uint64_t m, double *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = (w ? w[idx[i]] : 1.0) * (double)cw[i];
}

// FP64 sum of w_i * c_i: fixed 256-row blocks, then the block sums in order.
__global__ void k_wcost_part(const float *__restrict__ c, const double *__restrict__ w, uint64_t n,
                             double *__restrict__ part) {
    __shared__ double sm[256];
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    sm[threadIdx.x] = i < n ? (double)c[i] * (w ? w[i] : 1.0) : 0.0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void k_wcost_final(const double *__restrict__ part, uint64_t nb, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.;
    for (uint64_t b = 0; b < nb; ++b) s += part[b];
    *out = s;
}

// ---------------------------------------------------------------------------
// server
// ---------------------------------------------------------------------------
using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t ticket = 0;
    mcg_request q{};
    uint64_t footprint = 0;
    std::atomic<int> state{MCG_REQ_QUEUED};
    std::atomic<bool> cancel{false};
    mcg_result res{};
    Clock::time_point t_submit, t_start, t_end;
};

}  // namespace mcg

struct mcg_server {
    int device = 0;
    uint64_t budget = 0, in_use = 0;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<mcg::Req>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<mcg::Req>> reqs;
    std::vector<std::thread> workers;
    std::vector<mcg_ctx *> ctxs;
};

namespace mcg {

static uint64_t coreset_size(const mcg_request &q) {
    const uint64_t k = std::min(q.k, q.n);
    const uint64_t m = (uint64_t)std::ceil(q.coreset_frac * (double)q.n);
    return std::min(q.n, std::max(m, k));
}

// Device working set of a request (bytes): rows, their BF16 copy (dense
// SQRL2), per-row buffers of the seeding / sampling / assignment stages and of
// the near-tie re-score, the coreset rows and the centers.
static uint64_t footprint(const mcg_request &q) {
    const uint64_t n = q.n, d = q.d, k = std::min(q.k, q.n), m = coreset_size(q);
    const bool csr = q.indices != nullptr;
    uint64_t b = 1u << 20;
    b += csr ? q.nnz * 8 + (n + 1) * 8 : n * d * 4;
    if (!csr && q.msr == MCG_SQRL2) b += n * ((d + 63) / 64 * 64) * 2;
    b += n * 160;
    b += csr ? (q.nnz * m / std::max<uint64_t>(n, 1)) * 8 + m * 8 : m * d * 4;
    b += m * 64;
    b += k * d * 4 * 8;
    return b;
}

static void check(int rc) {
    if (rc != MCG_OK) MCG_THROW(rc, mcg_last_error());
}

static void maybe_fail(const Req &r, int stage) {
    if (r.q.fail_at_stage == stage)
        MCG_THROW(MCG_EDEVICE, "injected failure at stage " + std::to_string(stage));
}

// Runs the pipeline; returns false if the request was cancelled.
static bool run_request(Req &r, mcg_ctx *ctx) {
    const mcg_request &q = r.q;
    Ctx &c = ctx->c;
    hipStream_t st = c.stream;
    HIP_CHECK(hipSetDevice(c.device));
    auto cancelled = [&] { return r.cancel.load(std::memory_order_relaxed); };
    // stage boundaries: wall time since the previous boundary (the stages end with a
    // stream synchronisation, so this is the stage's device time plus its host work)
    Clock::time_point tb = Clock::now();
    auto stage_done = [&](int i) {
        const Clock::time_point t = Clock::now();
        r.res.stage_ms[i] = std::chrono::duration<double, std::milli>(t - tb).count();
        tb = t;
    };

    // 1. rows + seeding
    mcg_data *ds_raw = nullptr;
    if (q.indices)
        check(mcg_data_csr(ctx, q.data, q.data_dtype, q.indices, q.idx_dtype, q.indptr, q.ptr_dtype, q.n, q.d, q.nnz,
                           &ds_raw));
    else
        check(mcg_data_dense(ctx, q.data, q.data_dtype, q.n, q.d, q.ld, &ds_raw));
    std::unique_ptr<mcg_data, int (*)(mcg_data *)> ds(ds_raw, mcg_data_free);
    const uint64_t n = q.n, d = q.d, k = std::min(q.k, q.n);
    DBuf Wd, A, Dist, Costd, Probs, Idx, Cw, Csw, Cent, Asub, Afin, Cfin, Part, Tot;
    const double *W = nullptr;
    if (q.weights) {
        Wd.reserve(n * 8);
        HIP_CHECK(hipMemcpyAsync(Wd.p, q.weights, n * 8, is_device_ptr(q.weights) ? hipMemcpyDeviceToDevice
                                                                                  : hipMemcpyHostToDevice, st));
        W = Wd.as<const double>();
    }
    A.reserve(n * 4);
    Dist.reserve(n * 4);
    std::vector<uint32_t> ids(k);
    maybe_fail(r, 1);
    check(mcg_kmeanspp(ds.get(), q.msr, q.prior, k, q.seed, W, ids.data(), A.as<uint32_t>(), Dist.as<float>()));
    stage_done(0);
    if (cancelled()) return false;

    // 2. sensitivities
    maybe_fail(r, 2);
    Costd.reserve(n * 8);
    hipLaunchKernelGGL(k_f32_to_f64, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 256), 4096)), dim3(256), 0, st,
                       Dist.as<const float>(), n, Costd.as<double>());
    Probs.reserve(n * 4);
    check(mcg_coreset_probs(ctx, q.sens, n, k, Costd.as<const double>(), A.as<const uint32_t>(), W, Probs.as<float>()));
    stage_done(1);
    if (cancelled()) return false;

    // 3. coreset draws (from the seed's stream after the first 2^32 draws, which
    //    seeding uses) and the coreset rows
    maybe_fail(r, 3);
    const uint64_t m = coreset_size(q);
    Idx.reserve(m * 8);
    Cw.reserve(m * 4);
    check(mcg_coreset_sample(ctx, n, Probs.as<const float>(), W, m, q.seed, 1ull << 32, Idx.as<uint64_t>(),
                             Cw.as<float>()));
    Csw.reserve(m * 8);
    hipLaunchKernelGGL(k_coreset_weights, dim3((uint32_t)cdiv(m, 256)), dim3(256), 0, st, W,
                       Idx.as<const uint64_t>(), Cw.as<const float>(), m, Csw.as<double>());
    mcg_data *sub_raw = nullptr;
    check(mcg_data_rows(ds.get(), Idx.as<const uint64_t>(), m, &sub_raw));
    std::unique_ptr<mcg_data, int (*)(mcg_data *)> sub(sub_raw, mcg_data_free);
    stage_done(2);
    if (cancelled()) return false;

    // 4. weighted Lloyd on the coreset from the seeded rows
    maybe_fail(r, 4);
    Cent.reserve(k * d * 4);
    {
        std::vector<uint64_t> rows(ids.begin(), ids.end());
        std::vector<uint32_t> cid(k);
        for (uint64_t i = 0; i < k; ++i) cid[i] = (uint32_t)i;
        copy_rows_to_centers(ds->d, rows.data(), cid.data(), k, Cent.as<float>(), st);
        HIP_CHECK(hipStreamSynchronize(st));
    }
    Asub.reserve(m * 4);
    double ci = 0., cf = 0.;
    uint64_t iters = 0;
    check(mcg_lloyd(sub.get(), q.msr, q.prior, Cent.as<float>(), k, Csw.as<const double>(), q.eps, q.maxiter,
                    q.precision, Asub.as<uint32_t>(), nullptr, &ci, &cf, &iters));
    stage_done(3);
    if (cancelled()) return false;

    // 5. all rows to the final centers (one E-step on these rows)
    ds.get()->d.oneshot = true;
    Afin.reserve(n * 4);
    Cfin.reserve(n * 4);
    check(mcg_assign_hard(ds.get(), q.msr, q.prior, Cent.as<const float>(), k, q.precision, Afin.as<uint32_t>(),
                          Cfin.as<float>()));
    const uint64_t nb = cdiv(n, 256);
    Part.reserve(nb * 8);
    Tot.reserve(8);
    hipLaunchKernelGGL(k_wcost_part, dim3((uint32_t)nb), dim3(256), 0, st, Cfin.as<const float>(), W, n,
                       Part.as<double>());
    hipLaunchKernelGGL(k_wcost_final, dim3(1), dim3(64), 0, st, Part.as<const double>(), nb, Tot.as<double>());
    double total = 0.;
    HIP_CHECK(hipMemcpyAsync(&total, Tot.p, 8, hipMemcpyDeviceToHost, st));
    auto put = [&](void *dst, const void *src, size_t bytes) {
        if (!dst) return;
        HIP_CHECK(hipMemcpyAsync(dst, src, bytes, is_device_ptr(dst) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                 st));
    };
    put(q.centers, Cent.p, k * d * 4);
    put(q.asn, Afin.p, n * 4);
    put(q.costs, Cfin.p, n * 4);
    HIP_CHECK(hipStreamSynchronize(st));
    stage_done(4);
    r.res.cost = total;
    r.res.coreset_initcost = ci;
    r.res.coreset_finalcost = cf;
    r.res.coreset_m = m;
    r.res.numiter = iters;
    return true;
}

static double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

static void worker_main(mcg_server *s, mcg_ctx *ctx) {
    (void)hipSetDevice(s->device);
    // the device buffers of this worker's requests are allocated and freed in its
    // stream's order: a request's end then waits for its own stream only (hipFree
    // would wait for every worker's work); the context outlives the worker
    g_dbuf_stream = ctx->c.stream;
    for (;;) {
        std::shared_ptr<Req> r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                if (s->stop) return true;
                if (s->queue.empty()) return false;
                const auto &h = s->queue.front();
                return h->cancel.load() || s->in_use + h->footprint <= s->budget;
            });
            if (s->stop) return;
            r = s->queue.front();
            s->queue.pop_front();
            if (r->cancel.load()) {
                r->t_start = r->t_end = Clock::now();
                r->res.state = MCG_REQ_CANCELLED;
                r->state.store(MCG_REQ_CANCELLED);
                s->cv_done.notify_all();
                s->cv_work.notify_all();
                continue;
            }
            s->in_use += r->footprint;
            r->t_start = Clock::now();
            r->state.store(MCG_REQ_RUNNING);
        }
        int status = MCG_OK;
        bool finished = false;
        std::string err;
        try {
            finished = run_request(*r, ctx);
        } catch (const Error &e) {
            status = e.code;
            err = e.msg;
        } catch (const std::bad_alloc &) {
            status = MCG_ENOMEM;
            err = "host allocation failed";
        } catch (const std::exception &e) {
            status = MCG_EDEVICE;
            err = e.what();
        }
        if (status != MCG_OK) (void)hipStreamSynchronize(ctx->c.stream);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->in_use -= r->footprint;
            r->t_end = Clock::now();
            r->res.status = status;
            r->res.queue_ms = ms_between(r->t_submit, r->t_start);
            r->res.run_ms = ms_between(r->t_start, r->t_end);
            std::snprintf(r->res.error, sizeof(r->res.error), "%s", err.c_str());
            const int stt = status != MCG_OK ? MCG_REQ_FAILED : (finished ? MCG_REQ_DONE : MCG_REQ_CANCELLED);
            r->res.state = stt;
            r->state.store(stt);
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

static int guard_s(const std::function<void()> &f) {
    try {
        f();
        return MCG_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return MCG_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return MCG_EDEVICE;
    }
}

static void fill_result(const Req &r, mcg_result *res) {
    if (!res) return;
    const int st = r.state.load();
    if (st >= MCG_REQ_DONE) {
        *res = r.res;
    } else {
        std::memset(res, 0, sizeof(*res));
    }
    res->state = st;
}

}  // namespace mcg

using namespace mcg;

extern "C" {

int mcg_data_rows(const mcg_data *ds, const uint64_t *idx, uint64_t m, mcg_data **out) {
    return guard_s([&] {
        REQUIRE(ds && idx && out, "null argument");
        REQUIRE(m > 0, "empty row selection");
        const Data &S = ds->d;
        Ctx &c = *S.ctx;
        HIP_CHECK(hipSetDevice(c.device));
        hipStream_t st = c.stream;
        std::vector<uint64_t> h(m);
        DBuf dv;
        const uint64_t *idx_d = idx;
        if (is_device_ptr(idx)) {
            HIP_CHECK(hipMemcpyAsync(h.data(), idx, m * 8, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
        } else {
            std::memcpy(h.data(), idx, m * 8);
            dv.reserve(m * 8);
            HIP_CHECK(hipMemcpyAsync(dv.p, idx, m * 8, hipMemcpyHostToDevice, st));
            idx_d = dv.as<const uint64_t>();
        }
        for (uint64_t i = 0; i < m; ++i) REQUIRE(h[i] < S.n, "row index out of range");
        std::unique_ptr<mcg_data> o(new mcg_data);
        rows_into(S, h.data(), idx_d, m, o->d, st);
        *out = o.release();
    });
}

int mcg_server_create(int device, int nworkers, uint64_t hbm_budget, mcg_server **out) {
    return guard_s([&] {
        REQUIRE(out, "null out");
        REQUIRE(nworkers >= 1 && nworkers <= 64, "nworkers must be in [1, 64]");
        int nd = 0;
        HIP_CHECK(hipGetDeviceCount(&nd));
        REQUIRE(device >= 0 && device < nd, "invalid device ordinal");
        HIP_CHECK(hipSetDevice(device));
        std::unique_ptr<mcg_server> s(new mcg_server);
        s->device = device;
        if (hbm_budget == 0) {
            size_t fr = 0, tot = 0;
            HIP_CHECK(hipMemGetInfo(&fr, &tot));
            hbm_budget = (uint64_t)(0.8 * (double)fr);
        }
        s->budget = hbm_budget;
        for (int w = 0; w < nworkers; ++w) {
            mcg_ctx *c = nullptr;
            const int rc = mcg_ctx_create(device, &c);
            if (rc != MCG_OK) {
                for (auto *x : s->ctxs) mcg_ctx_destroy(x);
                MCG_THROW(rc, mcg_last_error());
            }
            c->c.serving = true;
            s->ctxs.push_back(c);
        }
        mcg_server *raw = s.get();
        for (int w = 0; w < nworkers; ++w) s->workers.emplace_back(worker_main, raw, s->ctxs[w]);
        *out = s.release();
    });
}

int mcg_server_destroy(mcg_server *s) {
    return guard_s([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            for (auto &r : s->queue) {
                r->t_start = r->t_end = Clock::now();
                r->res.state = MCG_REQ_CANCELLED;
                r->state.store(MCG_REQ_CANCELLED);
            }
            s->queue.clear();
        }
        s->cv_work.notify_all();
        s->cv_done.notify_all();
        for (auto &t : s->workers) t.join();
        for (auto *c : s->ctxs) mcg_ctx_destroy(c);
        delete s;
    });
}

int mcg_submit(mcg_server *s, const mcg_request *req, uint64_t *ticket) {
    return guard_s([&] {
        REQUIRE(s && req && ticket, "null argument");
        const mcg_request &q = *req;
        REQUIRE(q.data && q.n > 0 && q.d > 0 && q.k > 0, "request needs rows and k > 0");
        REQUIRE(q.centers && q.asn, "request needs centers and asn outputs");
        REQUIRE(msr_supported(q.msr) && msr_mean_centroid(q.msr), "unsupported measure (needs a mean centroid)");
        REQUIRE(q.coreset_frac > 0. && q.coreset_frac <= 1., "coreset_frac must be in (0, 1]");
        REQUIRE(q.sens == MCG_SENS_BFL || q.sens == MCG_SENS_FL || q.sens == MCG_SENS_VX || q.sens == MCG_SENS_LBK ||
                    q.sens == MCG_SENS_LFKF,
                "unsupported sensitivity method");
        REQUIRE(!q.indices || q.indptr, "CSR request needs indptr");
        auto r = std::make_shared<Req>();
        r->q = q;
        r->footprint = footprint(q);
        std::lock_guard<std::mutex> lk(s->mu);
        REQUIRE(!s->stop, "server is shutting down");
        if (r->footprint > s->budget)
            MCG_THROW(MCG_ENOMEM, "request footprint " + std::to_string(r->footprint) + " B exceeds the HBM budget " +
                                      std::to_string(s->budget) + " B");
        r->ticket = s->next_ticket++;
        r->t_submit = Clock::now();
        s->reqs[r->ticket] = r;
        s->queue.push_back(r);
        *ticket = r->ticket;
        s->cv_work.notify_one();
    });
}

int mcg_poll(mcg_server *s, uint64_t ticket, mcg_result *res) {
    return guard_s([&] {
        REQUIRE(s, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        REQUIRE(it != s->reqs.end(), "unknown ticket");
        fill_result(*it->second, res);
    });
}

int mcg_wait(mcg_server *s, uint64_t ticket, double timeout_s, mcg_result *res) {
    return guard_s([&] {
        REQUIRE(s, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        REQUIRE(it != s->reqs.end(), "unknown ticket");
        std::shared_ptr<Req> r = it->second;
        auto ended = [&] { return r->state.load() >= MCG_REQ_DONE; };
        if (timeout_s < 0) s->cv_done.wait(lk, ended);
        else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ended);
        fill_result(*r, res);
    });
}

int mcg_cancel(mcg_server *s, uint64_t ticket) {
    return guard_s([&] {
        REQUIRE(s, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        REQUIRE(it != s->reqs.end(), "unknown ticket");
        it->second->cancel.store(true);
        s->cv_work.notify_all();
    });
}

int mcg_release(mcg_server *s, uint64_t ticket) {
    return guard_s([&] {
        REQUIRE(s, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        REQUIRE(it != s->reqs.end(), "unknown ticket");
        REQUIRE(it->second->state.load() >= MCG_REQ_DONE, "request has not ended");
        s->reqs.erase(it);
    });
}

}  // extern "C"
