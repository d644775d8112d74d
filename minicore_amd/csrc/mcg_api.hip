// C ABI of libmcg: contexts, datasets, and the host-side drivers of the
// clustering hot path (Lloyd, soft EM, seeding, coresets).  Kernels live in
// the other translation units.
#include "mcg_internal.h"
#include <algorithm>
#include <mutex>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>

namespace mcg {

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }

bool is_device_ptr(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

void Ctx::tic(int which, hipEvent_t &a) {
    a = nullptr;
    if (!timing) return;
    (void)which;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventRecord(a, stream));
}

void Ctx::toc(int which, hipEvent_t a) {
    if (!timing || !a) return;
    hipEvent_t b;
    HIP_CHECK(hipEventCreate(&b));
    HIP_CHECK(hipEventRecord(b, stream));
    pending[which].push_back({a, b});
}

void Ctx::collect() {
    for (int w = 0; w < 2; ++w) {
        for (auto &pr : pending[w]) {
            HIP_CHECK(hipEventSynchronize(pr.second));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
            timed_ms[w] += ms;
            timed_n[w] += 1;
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        pending[w].clear();
    }
}

Ctx::~Ctx() {
    for (int w = 0; w < 2; ++w)
        for (auto &pr : pending[w]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    // the scratch buffers are freed in the stream's order: before the stream goes
    for (DBuf &b : s) b.release();
    if (stream) (void)hipStreamDestroy(stream);
    if (host_pinned) (void)hipHostFree(host_pinned);
}

void *Ctx::pinned(size_t nb) {
    if (nb > host_pinned_bytes) {
        if (host_pinned) (void)hipHostFree(host_pinned);
        host_pinned = nullptr;
        HIP_CHECK(hipHostMalloc(&host_pinned, nb));
        host_pinned_bytes = nb;
    }
    return host_pinned;
}

// Device view of a caller buffer (borrow device memory, stage host memory).
const void *in_dev(Ctx &c, const void *p, size_t bytes, DBuf &stage) {
    if (!p) return nullptr;
    if (is_device_ptr(p)) return p;
    stage.reserve(bytes);
    HIP_CHECK(hipMemcpyAsync(stage.p, p, bytes, hipMemcpyHostToDevice, c.stream));
    return stage.p;
}

// Output target: device pointer used directly; host pointer gets a device
// scratch that out_finish() copies back.
struct Out {
    void *user = nullptr;
    void *dev = nullptr;
    size_t bytes = 0;
    bool host = false;
};
static Out out_dev(Ctx &, void *p, size_t bytes, DBuf &stage) {
    Out o;
    o.user = p;
    o.bytes = bytes;
    if (!p) return o;
    if (is_device_ptr(p)) { o.dev = p; return o; }
    o.host = true;
    o.dev = stage.reserve(bytes);
    return o;
}
static void out_finish(Ctx &c, Out &o) {
    if (o.host && o.user) HIP_CHECK(hipMemcpyAsync(o.user, o.dev, o.bytes, hipMemcpyDeviceToHost, c.stream));
}

void sync(Ctx &c) {
    HIP_CHECK(hipStreamSynchronize(c.stream));
    c.collect();
}

// ---------------------------------------------------------------------------
// element types
// ---------------------------------------------------------------------------
static size_t dt_size(int dt) {
    switch (dt) {
    case MCG_F32: case MCG_U32: case MCG_I32: return 4;
    case MCG_F64: case MCG_U64: case MCG_I64: return 8;
    case MCG_U16: return 2;
    case MCG_U8: return 1;
    }
    MCG_THROW(MCG_EINVAL, "unknown dtype");
}

// ---------------------------------------------------------------------------
// CSR ingest on the device: the caller's dtypes converted to f32 values, u32
// column indices and u64 offsets, with the checks of the CSR invariants
// (err[0]: first/last offset, err[1]: decreasing offsets, err[2]: bit 0 column
// out of range, bit 1 negative index, err[3]: columns not increasing in a row)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rd_num(const void *p, int dt, uint64_t i) {
    switch (dt) {
    case MCG_F32: return ((const float *)p)[i];
    case MCG_F64: return ((const double *)p)[i];
    case MCG_U8: return ((const uint8_t *)p)[i];
    case MCG_U16: return ((const uint16_t *)p)[i];
    case MCG_U32: return ((const uint32_t *)p)[i];
    case MCG_U64: return (double)((const uint64_t *)p)[i];
    case MCG_I32: return ((const int32_t *)p)[i];
    default: return (double)((const int64_t *)p)[i];
    }
}

// index as signed 64-bit (u64 values beyond 2^63 read as negative, i.e. invalid)
__device__ __forceinline__ int64_t rd_idx(const void *p, int dt, uint64_t i) {
    switch (dt) {
    case MCG_U16: return ((const uint16_t *)p)[i];
    case MCG_U32: return ((const uint32_t *)p)[i];
    case MCG_U64: return (int64_t)((const uint64_t *)p)[i];
    case MCG_I32: return ((const int32_t *)p)[i];
    default: return ((const int64_t *)p)[i];
    }
}

__global__ void k_csr_ptr(const void *__restrict__ raw, int dt, uint64_t n, uint64_t nnz, uint64_t *__restrict__ P,
                          uint8_t *__restrict__ start, uint32_t *__restrict__ err) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t v = rd_idx(raw, dt, i);
        P[i] = (uint64_t)v;
        if ((i == 0 && v != 0) || (i == n && (uint64_t)v != nnz)) err[0] = 1u;
        if (i < n) {
            const int64_t nx = rd_idx(raw, dt, i + 1);
            if (v < 0 || nx < v) err[1] = 1u;
            else if (nx > v && (uint64_t)nx <= nnz) start[v] = 1;   // first entry of a non-empty row
        }
    }
}

__global__ void k_csr_entries(const void *__restrict__ rv, int vdt, const void *__restrict__ ri, int idt, uint64_t nnz,
                              uint32_t d, const uint8_t *__restrict__ start, float *__restrict__ val,
                              uint32_t *__restrict__ idx, uint32_t *__restrict__ err) {
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t j = rd_idx(ri, idt, p);
        if (j < 0) atomicOr(&err[2], 2u);
        else if ((uint64_t)j >= d) atomicOr(&err[2], 1u);
        idx[p] = (uint32_t)j;
        val[p] = (float)rd_num(rv, vdt, p);
        if (p > 0 && !start[p] && rd_idx(ri, idt, p - 1) >= j) err[3] = 1u;
    }
}

__global__ void k_f64_to_f32(const double *__restrict__ in, uint64_t n, uint64_t d, uint64_t ld,
                             float *__restrict__ out) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * d; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t i = t / d, j = t % d;
        out[t] = (float)in[i * ld + j];
    }
}

// per-row FP64 sums of the stored values
__global__ void k_rowsum_dense(const float *__restrict__ X, uint64_t n, uint32_t d, double *__restrict__ out) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double s = 0.;
    for (uint32_t j = lane; j < d; j += 64) s += (double)X[row * d + j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[row] = s;
}
__global__ void k_rowsum_csr(const float *__restrict__ val, const uint64_t *__restrict__ ptr, uint64_t n,
                             double *__restrict__ out) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double s = 0.;
    for (uint64_t p = ptr[row] + lane; p < ptr[row + 1]; p += 64) s += (double)val[p];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[row] = s;
}

void rowsums(Data &ds, hipStream_t st) {
    if (ds.rowsum.p) return;
    ds.rowsum.reserve(ds.n * sizeof(double));
    if (ds.csr)
        hipLaunchKernelGGL(k_rowsum_csr, dim3((uint32_t)cdiv(ds.n, 4)), dim3(256), 0, st, ds.val.as<const float>(),
                           ds.ptr.as<const uint64_t>(), ds.n, ds.rowsum.as<double>());
    else
        hipLaunchKernelGGL(k_rowsum_dense, dim3((uint32_t)cdiv(ds.n, 4)), dim3(256), 0, st, ds.X, ds.n,
                           (uint32_t)ds.d, ds.rowsum.as<double>());
}

// weighted FP64 total of per-point costs (fixed order)
__global__ void k_wsum(const float *__restrict__ cost, const double *__restrict__ w, uint64_t n,
                       double *__restrict__ out) {
    __shared__ double sm[1024];
    double s = 0.;
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t b = threadIdx.x * per, e = min<uint64_t>(n, b + per);
    for (uint64_t i = b; i < e; ++i) s += w ? w[i] * (double)cost[i] : (double)cost[i];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}

__global__ void k_wsum_d(const double *__restrict__ v, uint64_t n, double *__restrict__ out) {
    __shared__ double sm[1024];
    double s = 0.;
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t b = threadIdx.x * per, e = min<uint64_t>(n, b + per);
    for (uint64_t i = b; i < e; ++i) s += v[i];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}

// copy data rows into center slots: C[ids[q]] = row rows[q]
__global__ void k_rows_to_centers_dense(const float *__restrict__ X, uint32_t d, const uint64_t *__restrict__ rows,
                                        const uint32_t *__restrict__ ids, uint64_t m, float *__restrict__ C) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m * d; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t q = t / d, j = t % d;
        C[(uint64_t)ids[q] * d + j] = X[rows[q] * d + j];
    }
}
__global__ void k_rows_to_centers_csr(const float *__restrict__ val, const uint32_t *__restrict__ idx,
                                      const uint64_t *__restrict__ ptr, uint32_t d,
                                      const uint64_t *__restrict__ rows, const uint32_t *__restrict__ ids, uint64_t m,
                                      float *__restrict__ C) {
    const uint64_t q = blockIdx.x;
    if (q >= m) return;
    float *c = C + (uint64_t)ids[q] * d;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) c[j] = 0.f;
    __syncthreads();
    for (uint64_t p = ptr[rows[q]] + threadIdx.x; p < ptr[rows[q] + 1]; p += blockDim.x) c[idx[p]] = val[p];
}

void copy_rows_to_centers(Data &ds, const uint64_t *rows_h, const uint32_t *ids_h, uint64_t m, float *C,
                          hipStream_t st) {
    if (!m) return;
    Ctx &c = *ds.ctx;
    char *buf = (char *)c.s[S_MISC2].reserve(m * (sizeof(uint64_t) + sizeof(uint32_t)));
    HIP_CHECK(hipMemcpyAsync(buf, rows_h, m * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(buf + m * sizeof(uint64_t), ids_h, m * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    const uint64_t *rows = (const uint64_t *)buf;
    const uint32_t *ids = (const uint32_t *)(buf + m * sizeof(uint64_t));
    if (ds.csr)
        hipLaunchKernelGGL(k_rows_to_centers_csr, dim3((uint32_t)m), dim3(256), 0, st, ds.val.as<const float>(),
                           ds.idx.as<const uint32_t>(), ds.ptr.as<const uint64_t>(), (uint32_t)ds.d, rows, ids, m, C);
    else
        hipLaunchKernelGGL(k_rows_to_centers_dense, dim3((uint32_t)std::min<uint64_t>(cdiv(m * ds.d, 256), 4096)),
                           dim3(256), 0, st, ds.X, (uint32_t)ds.d, rows, ids, m, C);
    HIP_CHECK(hipStreamSynchronize(st));  // the host arrays may go away
}

// minibatch helpers
__global__ void k_count_asn(const uint32_t *__restrict__ asn, uint64_t n, uint32_t k, uint32_t *__restrict__ cnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (asn[i] < k) atomicAdd(&cnt[asn[i]], 1u);
}

// a row moves to its cheapest restarted center fid[a[i]] when that is strictly
// cheaper than its current cost
__global__ void k_take_lower(const float *__restrict__ v, const uint32_t *__restrict__ a,
                             const uint32_t *__restrict__ fid, uint64_t n, float *__restrict__ cost,
                             uint32_t *__restrict__ asn) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (v[i] < cost[i]) { cost[i] = v[i]; asn[i] = fid[a[i]]; }
}

__global__ void k_gather_rows_f32(const float *__restrict__ C, uint64_t d, const uint32_t *__restrict__ ids,
                                  uint64_t m, float *__restrict__ out) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m * d; t += (uint64_t)gridDim.x * blockDim.x)
        out[t] = C[(uint64_t)ids[t / d] * d + t % d];
}

__global__ void k_f32_to_f64_copy(const float *__restrict__ a, uint64_t n, double *__restrict__ b) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

__global__ void k_gather_f64(const double *__restrict__ w, const uint64_t *__restrict__ idx, uint64_t m,
                             double *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) out[t] = w[idx[t]];
}

__global__ void k_force(uint32_t *__restrict__ asn, float *__restrict__ cost, const uint64_t *__restrict__ rows,
                        const uint32_t *__restrict__ ids, uint64_t m) {
    uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const uint64_t r = rows[q];
    if (asn[r] != ids[q]) { asn[r] = ids[q]; if (cost) cost[r] = 0.f; }
}

// out[0] = *cost, out[1] = number of clusters without members
__global__ void k_step_summary(const double *__restrict__ cost, const uint32_t *__restrict__ counts, uint64_t k,
                               double *__restrict__ out) {
    __shared__ uint32_t sm[256];
    uint32_t z = 0;
    for (uint64_t q = threadIdx.x; q < k; q += blockDim.x) z += counts[q] == 0;
    sm[threadIdx.x] = z;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = *cost; out[1] = (double)sm[0]; }
}

__global__ void k_copy_f32(const float *__restrict__ src, float *__restrict__ dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

static void copy_f32(const float *src, float *dst, uint64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_f32, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 256), 2048)), dim3(256), 0, st, src, dst,
                       n);
}

// ---- device-side control of the Lloyd loop (solve.h:143-170) -------------------
struct LloydCtl {
    float cost;            // accepted cost, in float like the reference's FT (solve.h:120-123)
    float init;            // initial cost
    uint32_t active;       // the loop goes on
    uint32_t accept;       // the last candidate was accepted (read by k_lloyd_commit)
    uint32_t need_host;    // an accepted assignment left a cluster empty: the host restarts it
    uint32_t nan;          // a candidate center held a NaN (the host throws)
    uint32_t nanflag;      // set by k_lloyd_nan for the current candidate
    uint32_t pad;
    unsigned long long iter;
};
static_assert(sizeof(LloydCtl) <= 64, "LloydCtl");

__device__ bool block_any(bool v) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (v) any = 1;
    __syncthreads();
    return any != 0;
}

__global__ void k_lloyd_nan(const float *__restrict__ p, uint64_t n, LloydCtl *__restrict__ ctl) {
    bool bad = false;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        bad |= p[i] != p[i];
    if (bad) ctl->nanflag = 1u;
}

// after the first E+M pass: cost of the input centers; stop at cost 0 (solve.h:131-137)
__global__ void k_lloyd_init(LloydCtl *__restrict__ ctl, const double *__restrict__ cost64,
                             const uint32_t *__restrict__ cnt2, uint64_t k) {
    bool z = false;
    for (uint64_t q = threadIdx.x; q < k; q += blockDim.x) z |= cnt2[q] == 0u;
    const bool empty = block_any(z);
    if (threadIdx.x) return;
    const float cost = (float)*cost64;
    ctl->cost = cost;
    ctl->init = cost;
    ctl->iter = 0;
    ctl->accept = 0;
    ctl->nan = 0;
    ctl->nanflag = 0;
    ctl->need_host = cost != 0.f && empty;
    ctl->active = cost != 0.f && !empty;
}

// stop / accept after an iteration's M-step (solve.h:160-170): a cost increase
// without a restart keeps the old state and stops; otherwise accept, count the
// iteration and stop on old - new < eps * max(old, new) or iter > maxiter.
__global__ void k_lloyd_decide(LloydCtl *__restrict__ ctl, const double *__restrict__ cost64,
                               const uint32_t *__restrict__ cnt2, uint64_t k, double eps, uint64_t maxiter,
                               int restarted) {
    bool z = false;
    for (uint64_t q = threadIdx.x; q < k; q += blockDim.x) z |= cnt2[q] == 0u;
    const bool empty = block_any(z);
    if (threadIdx.x) return;
    const uint32_t nanflag = ctl->nanflag;
    ctl->nanflag = 0;
    ctl->accept = 0;
    if (!ctl->active) return;
    if (nanflag) {
        ctl->nan = 1;
        ctl->active = 0;
        return;
    }
    const float newcost = (float)*cost64;
    if (newcost > ctl->cost && !restarted) {
        ctl->active = 0;
        return;
    }
    ctl->accept = 1;
    ctl->iter += 1;
    const float old = ctl->cost;
    ctl->cost = newcost;
    const float diff = old - newcost;
    const double hi = (double)newcost < (double)old ? (double)old : (double)newcost;
    if ((double)diff < eps * hi || ctl->iter > maxiter) ctl->active = 0;
    else if (empty) { ctl->need_host = 1; ctl->active = 0; }
}

__global__ void k_lloyd_resume(LloydCtl *__restrict__ ctl) {
    ctl->need_host = 0;
    ctl->nanflag = 0;
    ctl->active = 1;
}

// mode 0: an accepted candidate becomes the state (C, A, pc, Cm, cnt);
// mode 1: only the means and counts of the first pass (Cm, cnt)
__global__ void k_lloyd_commit(const LloydCtl *__restrict__ ctl, int mode, float *__restrict__ C,
                               const float *__restrict__ Cn, uint64_t kd, uint32_t *__restrict__ A,
                               const uint32_t *__restrict__ An, uint64_t n, float *__restrict__ pc,
                               const float *__restrict__ pcn, float *__restrict__ Cm, const float *__restrict__ Cm2,
                               uint32_t *__restrict__ cnt, const uint32_t *__restrict__ cnt2, uint64_t k) {
    if (mode == 0 && !ctl->accept) return;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x, t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = t0; i < kd; i += stride) Cm[i] = Cm2[i];
    for (uint64_t i = t0; i < k; i += stride) cnt[i] = cnt2[i];
    if (mode) return;
    for (uint64_t i = t0; i < kd; i += stride) C[i] = Cn[i];
    for (uint64_t i = t0; i < n; i += stride) A[i] = An[i];
    if (pc)
        for (uint64_t i = t0; i < n; i += stride) pc[i] = pcn[i];
}

// One Lloyd iteration captured from the stream and replayed.  Capture is thread
// local, so other threads' contexts keep working meanwhile; a failed capture or
// instantiation leaves the caller to launch the iteration itself.
struct IterGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    bool failed = false;
    ~IterGraph() { reset(); }
    bool ready() const { return ge != nullptr; }
    void reset() {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        ge = nullptr;
        g = nullptr;
    }
    template <class F> void capture(hipStream_t st, F &&body) {
        HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(st, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        if (hipStreamEndCapture(st, &g) != hipSuccess || !g) {
            (void)hipGetLastError();
            g = nullptr;
            failed = true;
            return;
        }
        if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            reset();
            failed = true;
        }
    }
    void launch(hipStream_t st) { HIP_CHECK(hipGraphLaunch(ge, st)); }
};

// ---------------------------------------------------------------------------
// E-step dispatch: asn (device) and per-point best costs (device, may be null
// for dense SQRL2 where the M-step measures the cost).
// ---------------------------------------------------------------------------
static bool dense_l2(const Data &ds, int msr) { return !ds.csr && msr == MCG_SQRL2; }

static void estep(Data &ds, int msr, double prior, const float *C, uint64_t k, int precision, uint32_t *asn,
                  float *best, hipStream_t st) {
    if (dense_l2(ds, msr)) {
        if (best) dense_sqrl2_assign_exact(ds, C, k, asn, best, st);
        else dense_sqrl2_assign(ds, C, k, precision, asn, st);
        return;
    }
    generic_costs(ds, msr, prior, C, k, nullptr, asn, best, st);
}

static void check_msr(int msr) {
    REQUIRE(msr_supported(msr), "unsupported measure code (supported: 0-18, 23-26 of distance.h, and 64 for the "
                                "Poisson I-divergence)");
}

// hard clustering needs a mean-centroid measure (centroid.h:182-207)
static void check_mean_msr(int msr) {
    check_msr(msr);
    REQUIRE(msr_mean_centroid(msr), "this measure's centroid policy (L1/TVD/geometric median, or none) is not "
                                    "on the GPU path; its costs are (mcg_cost_matrix, mcg_assign_hard)");
}

// The hard drivers' measures: a mean policy, or a median policy with what the
// reference supports for it (no weighted or geometric median on CSR rows, and
// non-negative CSR values for the CSR median rules).
static void check_hard_msr(Data &D, int msr, const double *weights) {
    check_msr(msr);
    if (!median_policy(msr)) {
        check_mean_msr(msr);
        return;
    }
    if (!D.csr) return;
    if (weights)
        MCG_THROW(MCG_EUNSUPPORTED, msr == MCG_TVD  ? "Not implemented: weighted TVD median for CSparseMatrix"
                                    : msr == MCG_L2 ? "Not implemented: weighted geometric median for CSparseMatrix"
                                                    : "Not implemented: weighted L1 median for CSparseMatrix");
    if (msr == MCG_L2) return;   // the CSR geometric median takes any values
    HIP_CHECK(hipSetDevice(D.ctx->device));
    REQUIRE(!csr_has_negative(D, D.ctx->stream), "CSR L1/TVD median centers need non-negative values");
}

}  // namespace mcg

using namespace mcg;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *mcg_last_error(void) { return g_err.c_str(); }

int mcg_device_count(int *count) {
    return guard([&] {
        REQUIRE(count, "null count");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) { (void)hipGetLastError(); n = 0; }
        *count = n;
    });
}

// The device's default memory pool keeps freed memory for reuse instead of returning
// it to the driver at synchronisations (once per device).
static void mempool_setup(int device) {
    static std::mutex mu;
    static std::vector<int> done;
    std::lock_guard<std::mutex> lk(mu);
    if (std::find(done.begin(), done.end(), device) != done.end()) return;
    hipMemPool_t pool;
    HIP_CHECK(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t thr = ~0ull;
    HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
    done.push_back(device);
}

int mcg_ctx_create(int device, mcg_ctx **out) {
    return guard([&] {
        REQUIRE(out, "null out");
        int n = 0;
        HIP_CHECK(hipGetDeviceCount(&n));
        REQUIRE(device >= 0 && device < n, "invalid device ordinal");
        HIP_CHECK(hipSetDevice(device));
        // the fixed scratch of the allocation-free entry points, once per device
        noalloc_setup(device);
        std::unique_ptr<mcg_ctx> c(new mcg_ctx);
        c->c.device = device;
        HIP_CHECK(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
        // scratch buffers come from the device's memory pool in the stream's order (their
        // frees and regrowth wait for this stream only); the pool keeps what is freed
        mempool_setup(device);
        for (DBuf &b : c->c.s) b.st = c->c.stream;
        *out = c.release();
    });
}

int mcg_ctx_destroy(mcg_ctx *ctx) {
    return guard([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->c.device);
        (void)hipStreamSynchronize(ctx->c.stream);
        delete ctx;
    });
}

int mcg_ctx_synchronize(mcg_ctx *ctx) {
    return guard([&] { REQUIRE(ctx, "null ctx"); HIP_CHECK(hipStreamSynchronize(ctx->c.stream)); });
}

int mcg_ctx_stream(mcg_ctx *ctx, void **stream) {
    return guard([&] { REQUIRE(ctx && stream, "null argument"); *stream = (void *)ctx->c.stream; });
}

int mcg_data_dense(mcg_ctx *ctx, const void *x, int dtype, uint64_t n, uint64_t d, uint64_t ld, mcg_data **out) {
    return guard([&] {
        REQUIRE(ctx && out && x, "null argument");
        REQUIRE(n > 0 && d > 0, "empty matrix");
        REQUIRE(dtype == MCG_F32 || dtype == MCG_F64, "dense data must be float32 or float64");
        if (ld == 0) ld = d;
        REQUIRE(ld >= d, "ld < d");
        Ctx &c = ctx->c;
        HIP_CHECK(hipSetDevice(c.device));
        std::unique_ptr<mcg_data> ds(new mcg_data);
        Data &D = ds->d;
        D.ctx = &c;
        D.n = n;
        D.d = d;
        D.nnz = n * d;
        const bool dev = is_device_ptr(x);
        if (dev && dtype == MCG_F32 && ld == d) {
            D.X = (float *)x;  // borrowed
        } else {
            D.X = (float *)D.Xown.reserve(n * d * sizeof(float));
            if (dtype == MCG_F32) {
                // host rows travel as one contiguous copy (a pitched copy from pageable
                // memory goes row by row); a pitched source is compacted on the device
                DBuf tmp;
                const void *src = x;
                if (!dev && ld != d) {
                    tmp.reserve(n * ld * 4);
                    HIP_CHECK(hipMemcpyAsync(tmp.p, x, n * ld * 4, hipMemcpyHostToDevice, c.stream));
                    src = tmp.p;
                }
                if (!dev && ld == d)
                    HIP_CHECK(hipMemcpyAsync(D.X, x, n * d * 4, hipMemcpyHostToDevice, c.stream));
                else
                    HIP_CHECK(hipMemcpy2DAsync(D.X, d * 4, src, ld * 4, d * 4, n, hipMemcpyDeviceToDevice, c.stream));
                sync(c);
            } else {
                DBuf tmp;
                const void *src = x;
                if (!dev) {
                    tmp.reserve(n * ld * 8);
                    HIP_CHECK(hipMemcpyAsync(tmp.p, x, n * ld * 8, hipMemcpyHostToDevice, c.stream));
                    src = tmp.p;
                }
                hipLaunchKernelGGL(k_f64_to_f32, dim3((uint32_t)std::min<uint64_t>(cdiv(n * d, 256), 65536)), dim3(256),
                                   0, c.stream, (const double *)src, n, d, ld, D.X);
                // the values as given, for the float64 path (mcg_lloyd_f64, mcg_assign_f64)
                D.Xd.reserve(n * d * 8);
                HIP_CHECK(hipMemcpy2DAsync(D.Xd.p, d * 8, src, ld * 8, d * 8, n, hipMemcpyDeviceToDevice, c.stream));
                sync(c);
            }
        }
        sync(c);
        *out = ds.release();
    });
}

int mcg_data_csr(mcg_ctx *ctx, const void *data, int data_dtype, const void *indices, int idx_dtype,
                 const void *indptr, int ptr_dtype, uint64_t n, uint64_t d, uint64_t nnz, mcg_data **out) {
    return guard([&] {
        REQUIRE(ctx && out && indptr, "null argument");
        REQUIRE(n > 0 && d > 0, "empty matrix");
        REQUIRE(d < (1ull << 32), "d must fit in uint32");
        REQUIRE(nnz == 0 || (data && indices), "null argument");
        REQUIRE(nnz < (1ull << 40), "nnz too large");
        (void)dt_size(data_dtype);
        REQUIRE(idx_dtype == MCG_U16 || idx_dtype == MCG_U32 || idx_dtype == MCG_U64 || idx_dtype == MCG_I32 ||
                    idx_dtype == MCG_I64, "unsupported index dtype");
        REQUIRE(ptr_dtype == MCG_U16 || ptr_dtype == MCG_U32 || ptr_dtype == MCG_U64 || ptr_dtype == MCG_I32 ||
                    ptr_dtype == MCG_I64, "unsupported index dtype");
        Ctx &c = ctx->c;
        HIP_CHECK(hipSetDevice(c.device));
        hipStream_t st = c.stream;
        // the caller's arrays as they are (device memory borrowed, host memory staged),
        // converted and checked on the device
        DBuf s0, s1, s2, flags;
        const void *rp = in_dev(c, indptr, (n + 1) * dt_size(ptr_dtype), s2);
        const void *rv = nnz ? in_dev(c, data, nnz * dt_size(data_dtype), s0) : nullptr;
        const void *ri = nnz ? in_dev(c, indices, nnz * dt_size(idx_dtype), s1) : nullptr;
        std::unique_ptr<mcg_data> ds(new mcg_data);
        Data &D = ds->d;
        D.ctx = &c;
        D.csr = true;
        D.n = n;
        D.d = d;
        D.nnz = nnz;
        D.val.reserve(nnz * 4);
        D.idx.reserve(nnz * 4);
        D.ptr.reserve((n + 1) * 8);
        const uint64_t words = 4 + cdiv(nnz, 4);   // 4 error words, then one start byte per entry
        uint32_t *err = (uint32_t *)flags.reserve(words * 4);
        uint8_t *start = (uint8_t *)(err + 4);
        fill_u32(err, 0u, words, st);
        hipLaunchKernelGGL(k_csr_ptr, dim3((uint32_t)std::min<uint64_t>(cdiv(n + 1, 256), 65536)), dim3(256), 0, st, rp,
                           ptr_dtype, n, nnz, D.ptr.as<uint64_t>(), start, err);
        if (nnz)
            hipLaunchKernelGGL(k_csr_entries, dim3((uint32_t)std::min<uint64_t>(cdiv(nnz, 256), 65536)), dim3(256), 0,
                               st, rv, data_dtype, ri, idx_dtype, nnz, (uint32_t)d, start, D.val.as<float>(),
                               D.idx.as<uint32_t>(), err);
        uint32_t h[4];
        copy_sync(h, err, sizeof(h), hipMemcpyDeviceToHost, st);
        REQUIRE(!h[0], "indptr must start at 0 and end at nnz");
        REQUIRE(!h[1], "indptr must be non-decreasing");
        REQUIRE(!(h[2] & 2u), "negative index");
        REQUIRE(!(h[2] & 1u), "column index out of range");
        REQUIRE(!h[3], "column indices must be sorted and unique within a row");
        *out = ds.release();
    });
}

int mcg_data_free(mcg_data *ds) {
    return guard([&] {
        if (!ds) return;
        (void)hipStreamSynchronize(ds->d.ctx->stream);
        delete ds;
    });
}

int mcg_data_shape(const mcg_data *ds, uint64_t *n, uint64_t *d, uint64_t *nnz) {
    return guard([&] {
        REQUIRE(ds, "null dataset");
        if (n) *n = ds->d.n;
        if (d) *d = ds->d.d;
        if (nnz) *nnz = ds->d.nnz;
    });
}

int mcg_cost_matrix(mcg_data *ds, int msr, double prior, const float *centers, uint64_t k, float *out) {
    return guard([&] {
        REQUIRE(ds && centers && out, "null argument");
        REQUIRE(k > 0, "k must be positive");
        check_msr(msr);
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        HIP_CHECK(hipSetDevice(c.device));
        DBuf sc, so;
        const float *C = (const float *)in_dev(c, centers, k * D.d * 4, sc);
        Out o = out_dev(c, out, D.n * k * 4, so);
        generic_costs(D, msr, prior, C, k, (float *)o.dev, nullptr, nullptr, c.stream);
        out_finish(c, o);
        sync(c);
    });
}

// condensed index of (i, j), i < j: n*i - i*(i+1)/2 + (j - i - 1) (pycmp.cpp:261)
__global__ void k_condense(const float *__restrict__ blk, uint64_t R, uint64_t K, uint64_t i0, uint64_t j0, uint64_t n,
                           float *__restrict__ out) {
    const uint64_t total = R * K;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t a = e / K, b = e - a * K;
        const uint64_t i = i0 + a, j = j0 + b;
        if (j > i) out[n * i - i * (i + 1) / 2 + (j - i - 1)] = blk[e];
    }
}

// pcmp (python/pycmp.cpp:249-338): out[(i, j)] = msr(first = row i, second = row j) for
// i < j in condensed order, by blocks of rows against blocks of rows as centers
// (blocks below the diagonal skipped), each block through the cost-matrix kernels, so
// every value equals the cost matrix's.
int mcg_pairwise(mcg_data *ds, int msr, double prior, float *out) {
    return guard([&] {
        REQUIRE(ds && out, "null argument");
        check_msr(msr);
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n, d = D.d, npairs = n * (n - (n ? 1 : 0)) / 2;
        if (npairs == 0) return;
        DBuf so, bblk, bctr, bidx, bidxc;
        Out o = out_dev(c, out, npairs * 4, so);
        const uint64_t B = 4096;
        float *blk = (float *)bblk.reserve(B * B * 4);
        float *ctr = D.csr ? (float *)bctr.reserve(B * d * 4) : nullptr;
        uint64_t *idx_d = (uint64_t *)bidx.reserve(B * 8);
        std::vector<uint64_t> ids(B);
        std::vector<uint32_t> cids(B);
        mcg_data sub;
        for (uint64_t i0 = 0; i0 + 1 < n; i0 += B) {
            const uint64_t R = std::min(B, n - i0);
            for (uint64_t t = 0; t < R; ++t) ids[t] = i0 + t;
            copy_sync(idx_d, ids.data(), R * 8, hipMemcpyHostToDevice, st);
            rows_into(D, ids.data(), idx_d, R, sub.d, st);
            for (uint64_t j0 = i0; j0 < n; j0 += B) {
                const uint64_t K = std::min(B, n - j0);
                const float *C = D.X + j0 * d;
                if (D.csr) {
                    std::vector<uint64_t> rows(K);
                    for (uint64_t t = 0; t < K; ++t) { rows[t] = j0 + t; cids[t] = (uint32_t)t; }
                    copy_rows_to_centers(D, rows.data(), cids.data(), K, ctr, st);
                    C = ctr;
                }
                generic_costs(sub.d, msr, prior, C, K, blk, nullptr, nullptr, st);
                hipLaunchKernelGGL(k_condense, dim3((uint32_t)std::min<uint64_t>(cdiv(R * K, 256), 65536)), dim3(256),
                                   0, st, blk, R, K, i0, j0, n, (float *)o.dev);
            }
        }
        out_finish(c, o);
        sync(c);
    });
}

int mcg_assign_hard(mcg_data *ds, int msr, double prior, const float *centers, uint64_t k, int precision,
                    uint32_t *asn, float *costs) {
    return guard([&] {
        REQUIRE(ds && centers && asn, "null argument");
        REQUIRE(k > 0, "k must be positive");
        check_msr(msr);
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        HIP_CHECK(hipSetDevice(c.device));
        DBuf sc, sa, sk;
        const float *C = (const float *)in_dev(c, centers, k * D.d * 4, sc);
        Out oa = out_dev(c, asn, D.n * 4, sa);
        Out ok = out_dev(c, costs, D.n * 4, sk);
        if (dense_l2(D, msr)) {
            dense_sqrl2_assign(D, C, k, precision, (uint32_t *)oa.dev, c.stream);
            if (ok.dev) dense_sqrl2_costs(D, C, (const uint32_t *)oa.dev, (float *)ok.dev, c.stream);
        } else {
            generic_costs(D, msr, prior, C, k, nullptr, (uint32_t *)oa.dev, (float *)ok.dev, c.stream);
        }
        out_finish(c, oa);
        out_finish(c, ok);
        sync(c);
    });
}

int mcg_centroids_hard(mcg_data *ds, const uint32_t *asn, const double *weights, uint64_t k, float *centers,
                       uint64_t *n_empty) {
    return guard([&] {
        REQUIRE(ds && asn && centers, "null argument");
        REQUIRE(k > 0, "k must be positive");
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        HIP_CHECK(hipSetDevice(c.device));
        DBuf sa, sw, sc;
        {
            std::vector<uint32_t> h(D.n);
            if (is_device_ptr(asn)) {
                copy_sync(h.data(), asn, D.n * 4, hipMemcpyDeviceToHost, c.stream);
            } else {
                std::memcpy(h.data(), asn, D.n * 4);
            }
            for (uint64_t i = 0; i < D.n; ++i) REQUIRE(h[i] < k, "assignment out of range");
        }
        const uint32_t *A = (const uint32_t *)in_dev(c, asn, D.n * 4, sa);
        const double *W = (const double *)in_dev(c, weights, D.n * 8, sw);
        float *C;
        bool host_c = !is_device_ptr(centers);
        if (host_c) {
            C = (float *)sc.reserve(k * D.d * 4);
            HIP_CHECK(hipMemcpyAsync(C, centers, k * D.d * 4, hipMemcpyHostToDevice, c.stream));
        } else {
            C = centers;
        }
        MStepOut mo = D.csr ? csr_mstep(D, A, W, k, C, c.stream) : dense_mstep(D, A, W, k, C, C, false, c.stream);
        std::vector<uint32_t> cnt(k);
        HIP_CHECK(hipMemcpyAsync(cnt.data(), mo.counts, k * 4, hipMemcpyDeviceToHost, c.stream));
        if (host_c) HIP_CHECK(hipMemcpyAsync(centers, C, k * D.d * 4, hipMemcpyDeviceToHost, c.stream));
        sync(c);
        if (n_empty) *n_empty = (uint64_t)std::count(cnt.begin(), cnt.end(), 0u);
    });
}

// perform_hard_clustering (solve.h:98-175).  For dense SQRL2 the M-step pass also
// measures cost(asn, C) in FP64; elsewhere the E-step's per-point minima are summed.
//
// The stop / accept rules run on the device (k_lloyd_decide), so iterations are
// queued in batches (2, 4, ... up to a size set by the work per iteration) with one
// host synchronisation per batch; iterations queued after the loop has stopped
// compute into the candidate buffers and change nothing.  With graphs enabled on
// the context the iterations after the first replay one captured HIP graph.  An
// accepted assignment that leaves a cluster empty hands the next iteration to the
// host, which runs the reservoir restart of centroid.h:687-752 and then goes back to
// device batches.
// perform_hard_clustering (solve.h:98-175) for the median policies of L1, TVD and
// L2 (centroid.h:434-647): host-driven, one synchronisation per iteration.  The
// M-step works on copies of the accepted assignment and costs (the empty-cluster
// restart moves rows in them); set_centroids_hard reports no restart for these
// policies, so a rising cost always reverts (solve.h:160-164).
static void lloyd_median(Data &D, int msr, double prior, float *centers, uint64_t k, const double *weights,
                         double eps, uint64_t maxiter, uint32_t *asn, float *costs, double *initcost,
                         double *finalcost, uint64_t *numiter) {
    Ctx &c = *D.ctx;
    hipStream_t st = c.stream;
    const uint64_t n = D.n, d = D.d, kd = k * d;
    const int pol = median_policy(msr);
    if (D.csr) {
        if (weights)
            MCG_THROW(MCG_EUNSUPPORTED, pol == 2   ? "Not implemented: weighted TVD median for CSparseMatrix"
                                        : pol == 3 ? "Not implemented: weighted geometric median for CSparseMatrix"
                                                   : "Not implemented: weighted L1 median for CSparseMatrix");
        // the geometric median of CSR rows follows csc.h:1131-1203 with its member
        // costs indexed by position (the reference indexes them by row id)
        if (pol != 3) REQUIRE(!csr_has_negative(D, c.stream), "CSR L1/TVD median centers need non-negative values");
    }
    DBuf sw, sa_out, sk_out;
    const double *W = (const double *)in_dev(c, weights, n * 8, sw);
    if (weights && pol != 3) {
        std::vector<double> hw(n);
        copy_sync(hw.data(), W, n * 8, hipMemcpyDeviceToHost, st);
        for (uint64_t i = 0; i < n; ++i) REQUIRE(hw[i] > 0., "weighted median centers need positive weights");
    }
    float *C = (float *)c.s[S_CENT_A].reserve(kd * 4);
    float *Cc = (float *)c.s[S_CENT_B].reserve(kd * 4);
    uint32_t *A = (uint32_t *)c.s[S_ASN_A].reserve(n * 4);
    uint32_t *An = (uint32_t *)c.s[S_ASN_B].reserve(n * 4);
    float *pc = (float *)c.s[S_COSTS].reserve(2 * n * 4);
    float *pcn = pc + n;
    HIP_CHECK(hipMemcpyAsync(C, centers, kd * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                       : hipMemcpyHostToDevice, st));
    generic_costs(D, msr, prior, C, k, nullptr, A, pc, st);
    float cost = (float)fixed_wsum(c, pc, W, n, st);
    uint64_t iter = 0;
    *initcost = cost;
    if (cost == 0.f) {
        *initcost = 0.;
        *finalcost = 0.;
    } else {
        copy_f32(C, Cc, kd, st);
        for (;;) {
            HIP_CHECK(hipMemcpyAsync(An, A, n * 4, hipMemcpyDeviceToDevice, st));
            copy_f32(pc, pcn, n, st);
            median_mstep(D, msr, W, k, An, pcn, Cc, st);
            generic_costs(D, msr, prior, Cc, k, nullptr, An, pcn, st);
            const float newcost = (float)fixed_wsum(c, pcn, W, n, st);
            if (newcost > cost) break;   // the accepted state stays (the reference re-assigns to it)
            copy_f32(Cc, C, kd, st);
            std::swap(A, An);
            std::swap(pc, pcn);
            ++iter;
            const float old = cost;
            cost = newcost;
            const float diff = old - newcost;
            if ((double)diff < eps * std::max((double)newcost, (double)old) || iter > maxiter) break;
        }
        *finalcost = cost;
    }
    *numiter = iter;
    Out oa = out_dev(c, asn, n * 4, sa_out);
    Out ok = out_dev(c, costs, n * 4, sk_out);
    HIP_CHECK(hipMemcpyAsync(oa.dev, A, n * 4, hipMemcpyDeviceToDevice, st));
    if (ok.dev) HIP_CHECK(hipMemcpyAsync(ok.dev, pc, n * 4, hipMemcpyDeviceToDevice, st));
    out_finish(c, oa);
    out_finish(c, ok);
    HIP_CHECK(hipMemcpyAsync(centers, C, kd * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                       : hipMemcpyDeviceToHost, st));
    sync(c);
}

int mcg_lloyd(mcg_data *ds, int msr, double prior, float *centers, uint64_t k, const double *weights, double eps,
              uint64_t maxiter, int precision, uint32_t *asn, float *costs, double *initcost, double *finalcost,
              uint64_t *numiter) {
    return guard([&] {
        REQUIRE(ds && centers && asn, "null argument");
        REQUIRE(k > 0, "k must be positive");
        check_msr(msr);
        if (median_policy(msr)) {
            HIP_CHECK(hipSetDevice(ds->d.ctx->device));
            lloyd_median(ds->d, msr, prior, centers, k, weights, eps, maxiter, asn, costs, initcost, finalcost,
                         numiter);
            return;
        }
        check_mean_msr(msr);
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n, d = D.d, kd = k * d;
        const bool l2 = dense_l2(D, msr);
        DBuf sw, sa_out, sk_out;
        const double *W = (const double *)in_dev(c, weights, n * 8, sw);
        // accepted state: centers C, assignment A, per-point costs pc, means of A Cm and
        // their counts cnt; candidate state: Cn, An, pcn, means Cm2
        float *C = (float *)c.s[S_CENT_A].reserve(kd * 4);
        float *Cn = (float *)c.s[S_CENT_B].reserve(kd * 4);
        float *Cm = (float *)c.s[S_CENT_C].reserve(kd * 4);
        float *Cm2 = (float *)c.s[S_CENT_D].reserve(kd * 4);
        uint32_t *A = (uint32_t *)c.s[S_ASN_A].reserve(n * 4);
        uint32_t *An = (uint32_t *)c.s[S_ASN_B].reserve(n * 4);
        float *pc = (float *)c.s[S_COSTS].reserve(2 * n * 4);
        float *pcn = pc + n;
        double *tot = (double *)c.s[S_TOT].reserve(4 * sizeof(double));
        char *cbuf = (char *)c.s[S_LCTL].reserve(64 + k * 4);
        LloydCtl *ctl = (LloydCtl *)cbuf;
        uint32_t *cnt = (uint32_t *)(cbuf + 64);
        HIP_CHECK(hipMemcpyAsync(C, centers, kd * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                           : hipMemcpyHostToDevice, st));
        // E-step of Cx into (a, cp), then the FP64 cost and the means of a (empty
        // clusters keep Cx) into Cm2.  Returns the cost and counts locations.
        auto measure = [&](const float *Cx, uint32_t *a, float *cp) -> std::pair<const double *, const uint32_t *> {
            estep(D, msr, prior, Cx, k, precision, a, l2 ? nullptr : cp, st);
            copy_f32(Cx, Cm2, kd, st);
            if (l2) {
                const MStepOut mo = dense_mstep(D, a, W, k, Cx, Cm2, true, st);
                return {mo.total, mo.counts};
            }
            hipLaunchKernelGGL(k_wsum, dim3(1), dim3(1024), 0, st, cp, W, n, tot);
            const MStepOut mo = D.csr ? csr_mstep(D, a, W, k, Cm2, st) : dense_mstep(D, a, W, k, Cx, Cm2, false, st);
            return {tot, mo.counts};
        };
        const uint32_t cgrid = (uint32_t)std::min<uint64_t>(cdiv(2 * kd + 2 * n + k, 256), 2048);
        auto commit = [&](int mode, const uint32_t *cnt2) {
            hipLaunchKernelGGL(k_lloyd_commit, dim3(cgrid), dim3(256), 0, st, ctl, mode, C, Cn, kd, A, An, n,
                               l2 ? nullptr : pc, pcn, Cm, Cm2, cnt, cnt2, k);
        };
        // one iteration from the accepted state (no restart); restarted = the host has
        // already put the candidate centers into Cn
        auto iteration = [&](int restarted) {
            if (!restarted) copy_f32(Cm, Cn, kd, st);
            hipLaunchKernelGGL(k_lloyd_nan, dim3((uint32_t)std::min<uint64_t>(cdiv(kd, 256), 1024)), dim3(256), 0, st,
                               Cn, kd, ctl);
            const auto m = measure(Cn, An, pcn);
            hipLaunchKernelGGL(k_lloyd_decide, dim3(1), dim3(256), 0, st, ctl, m.first, m.second, k, eps, maxiter,
                               restarted);
            commit(0, m.second);
        };
        // the host part of an iteration that restarts empty clusters (next_centers of
        // centroid.h:687-752): draws from the accepted costs, the drawn rows as centers,
        // full re-assignment, the drawn rows forced to their centers, new means in Cn
        auto restart = [&] {
            std::vector<uint32_t> hc(k);
            HIP_CHECK(hipMemcpyAsync(hc.data(), cnt, k * 4, hipMemcpyDeviceToHost, st));
            sync(c);
            std::vector<uint32_t> ids;
            for (uint64_t q = 0; q < k; ++q) if (!hc[q]) ids.push_back((uint32_t)q);
            const uint64_t ne = ids.size();
            REQUIRE(ne > 0, "internal: restart without an empty cluster");
            hipLaunchKernelGGL(k_lloyd_resume, dim3(1), dim3(1), 0, st, ctl);
            if (l2) dense_sqrl2_costs(D, C, A, pc, st);
            std::vector<uint64_t> rows(ne);
            restart_draws(c, pc, n, ne, rows.data(), st);
            copy_f32(C, Cn, kd, st);
            copy_rows_to_centers(D, rows.data(), ids.data(), ne, Cn, st);
            if (l2) dense_sqrl2_assign_exact(D, Cn, k, A, pc, st);
            else generic_costs(D, msr, prior, Cn, k, nullptr, A, pc, st);
            char *buf = (char *)c.s[S_MISC2].reserve(ne * 12);
            HIP_CHECK(hipMemcpyAsync(buf, rows.data(), ne * 8, hipMemcpyHostToDevice, st));
            HIP_CHECK(hipMemcpyAsync(buf + ne * 8, ids.data(), ne * 4, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_force, dim3((uint32_t)cdiv(ne, 256)), dim3(256), 0, st, A, pc,
                               (const uint64_t *)buf, (const uint32_t *)(buf + ne * 8), ne);
            if (D.csr) csr_mstep(D, A, W, k, Cn, st);
            else dense_mstep(D, A, W, k, Cn, Cn, false, st);
            sync(c);
        };

        {
            const auto m = measure(C, A, pc);
            hipLaunchKernelGGL(k_lloyd_init, dim3(1), dim3(256), 0, st, ctl, m.first, m.second, k);
            commit(1, m.second);
        }
        // batch size limit from the work of one iteration
        const double work = (double)(D.csr ? D.nnz : n * d) * (double)k;
        const uint64_t bmax = work <= 0x1p30 ? 16 : work <= 0x1p33 ? 8 : work <= 0x1p35 ? 4 : 1;
        static const bool graphs_env = [] { const char *e = getenv("MCG_GRAPHS"); return !(e && e[0] == '0'); }();
        const bool use_graph = graphs_env && c.graphs && !c.timing;
        IterGraph graph;
        bool eager_done = false;
        uint64_t batch = 2;
        LloydCtl h;
        for (;;) {
            HIP_CHECK(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, st));
            sync(c);
            if (h.nan) MCG_THROW(MCG_ENAN, "Unexpected nan in ctr");
            if (h.need_host) {
                graph.reset();      // the restart may grow scratch buffers the graph points into
                eager_done = false;
                restart();
                iteration(1);
                continue;
            }
            if (!h.active) break;
            const uint64_t span = maxiter - h.iter;  // h.iter <= maxiter while active
            const uint64_t left = span == ~0ull ? span : span + 1;
            const uint64_t nb = std::min<uint64_t>(left, std::min<uint64_t>(batch, bmax));
            for (uint64_t t = 0; t < nb; ++t) {
                if (!eager_done || !use_graph) {
                    iteration(0);
                    eager_done = true;
                } else {
                    if (!graph.ready() && !graph.failed) graph.capture(st, [&] { iteration(0); });
                    if (graph.ready()) graph.launch(st);
                    else iteration(0);
                }
            }
            batch *= 2;
        }
        *initcost = h.init;
        *finalcost = h.cost;
        *numiter = h.iter;
        // outputs
        Out oa = out_dev(c, asn, n * 4, sa_out);
        Out ok = out_dev(c, costs, n * 4, sk_out);
        HIP_CHECK(hipMemcpyAsync(oa.dev, A, n * 4, hipMemcpyDeviceToDevice, st));
        if (ok.dev) {
            if (l2) dense_sqrl2_costs(D, C, A, (float *)ok.dev, st);
            else HIP_CHECK(hipMemcpyAsync(ok.dev, pc, n * 4, hipMemcpyDeviceToDevice, st));
        }
        out_finish(c, oa);
        out_finish(c, ok);
        HIP_CHECK(hipMemcpyAsync(centers, C, kd * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                           : hipMemcpyDeviceToHost, st));
        sync(c);
    });
}

// The check-in of the minibatch drivers (solve.h:445-509, 644-710): every row
// assigned (A, P), each center with <= reseed_after rows restarted at a row drawn
// over w_i * cost_i with the next outputs of the driver's stream (one fixed-point
// prefix for all of them), rows moved to the cheapest restarted center when that is
// strictly cheaper (the reference's sequential scan over the restarted centers,
// as one assignment), and the exact fixed-point total of w_i * cost_i returned.
struct CheckIn {
    Data &D;
    int msr;
    double prior;
    uint64_t k;
    int precision;
    const double *W;                 // device weights or null
    const std::vector<double> &wh;   // host weights (empty without)
    uint64_t reseed_after;
    uint32_t *A, *Ta, *cnt, *fid;
    float *P, *Tmp, *Cf;
    std::vector<uint32_t> hc;
    std::vector<float> hp;

    double run(float *C, uint64_t seed, uint64_t &j, hipStream_t st) {
        Ctx &c = *D.ctx;
        const uint64_t n = D.n, d = D.d;
        const bool l2 = dense_l2(D, msr);
        estep(D, msr, prior, C, k, precision, A, l2 ? nullptr : P, st);
        if (l2) dense_sqrl2_costs(D, C, A, P, st);
        fill_u32(cnt, 0u, k, st);
        const uint32_t pgrid = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 4096);
        hipLaunchKernelGGL(k_count_asn, dim3(pgrid), dim3(256), 0, st, A, n, (uint32_t)k, cnt);
        hc.resize(k);
        copy_sync(hc.data(), cnt, k * 4, hipMemcpyDeviceToHost, st);
        std::vector<uint32_t> found;
        for (uint64_t q = 0; q < k; ++q) if (hc[q] <= reseed_after) found.push_back((uint32_t)q);
        if (!found.empty()) {
            const uint64_t nf = found.size();
            hp.resize(n);
            copy_sync(hp.data(), P, n * 4, hipMemcpyDeviceToHost, st);
            std::vector<double> wv(n);
            for (uint64_t i = 0; i < n; ++i) wv[i] = wh.empty() ? (double)hp[i] : wh[i] * (double)hp[i];
            std::vector<uint64_t> rows(nf), rr(nf);
            for (uint64_t f = 0; f < nf; ++f) rr[f] = wyrand_at(seed, j++);
            draw_fixed_many(wv, rr.data(), nf, rows.data());
            copy_rows_to_centers(D, rows.data(), found.data(), nf, C, st);
            copy_sync(fid, found.data(), nf * 4, hipMemcpyHostToDevice, st);
            hipLaunchKernelGGL(k_gather_rows_f32, dim3((uint32_t)std::min<uint64_t>(cdiv(nf * d, 256), 4096)),
                               dim3(256), 0, st, C, d, fid, nf, Cf);
            estep(D, msr, prior, Cf, nf, precision, Ta, l2 ? nullptr : Tmp, st);
            if (l2) dense_sqrl2_costs(D, Cf, Ta, Tmp, st);
            hipLaunchKernelGGL(k_take_lower, dim3(pgrid), dim3(256), 0, st, Tmp, Ta, fid, n, P, A);
        }
        return fixed_wsum(c, P, W, n, st);
    }
};

// perform_hard_minibatch_clustering (solve.h:377-578) for the mean-centroid
// measures.  One RNG stream (the wyrand counter stream of seed) serves the restart
// draws and the minibatch draws in the reference's order.  Check-ins at iteration
// 0, every freq-th and the last keep the centers of the lowest check-in cost;
// between check-ins mbsize rows are drawn (without replacement: the first mbsize
// distinct draws), sorted, assigned, and every center with members becomes their
// weighted mean.  The draws and the check-in bookkeeping run on the host; rows,
// costs and centers stay on the device.
int mcg_minibatch(mcg_data *ds, int msr, double prior, float *centers, uint64_t k, const double *weights,
                  uint64_t mbsize, uint64_t maxiter, uint64_t freq, uint64_t reseed_after, int with_replacement,
                  uint64_t seed, int precision, uint32_t *asn, float *costs, double *initcost, double *finalcost,
                  uint64_t *numiter) {
    return guard([&] {
        REQUIRE(ds && centers && initcost && finalcost && numiter, "null argument");
        REQUIRE(k > 0, "k must be positive");
        REQUIRE(mbsize > 0 && maxiter > 0 && freq > 0, "mbsize, maxiter and the check-in period must be positive");
        Data &D = ds->d;
        check_hard_msr(D, msr, weights);
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n, d = D.d, kd = k * d;
        REQUIRE(with_replacement || mbsize <= n, "mbsize exceeds the number of rows (draws without replacement)");
        REQUIRE(!(weights && (msr == MCG_L1 || msr == MCG_TVD)) || mbsize <= n,
                "weighted L1 / TVD minibatch updates index the weights by member rank (kmedian.h:340): mbsize "
                "must not exceed the number of rows");
        const bool l2 = dense_l2(D, msr);
        DBuf sw, sa_out, sk_out, bC, bSave, bA, bP, bTmp, bTa, bCnt, bIdx, bWs, bAs, bCs, bCf, bFid;
        const double *W = (const double *)in_dev(c, weights, n * 8, sw);
        std::vector<double> wh;
        if (weights) {
            wh.resize(n);
            if (is_device_ptr(weights)) copy_sync(wh.data(), weights, n * 8, hipMemcpyDeviceToHost, st);
            else std::memcpy(wh.data(), weights, n * 8);
        }
        float *C = (float *)bC.reserve(kd * 4), *Save = (float *)bSave.reserve(kd * 4);
        CheckIn ck{D, msr, prior, k, precision, W, wh, reseed_after, (uint32_t *)bA.reserve(n * 4),
                   (uint32_t *)bTa.reserve(n * 4), (uint32_t *)bCnt.reserve(k * 4), (uint32_t *)bFid.reserve(k * 4),
                   (float *)bP.reserve(n * 4), (float *)bTmp.reserve(n * 4), (float *)bCf.reserve(kd * 4), {}, {}};
        uint64_t *idx_d = (uint64_t *)bIdx.reserve(mbsize * 8);
        double *Ws = W ? (double *)bWs.reserve(mbsize * 8) : nullptr;
        uint32_t *As = (uint32_t *)bAs.reserve(mbsize * 4);
        float *Cs = (float *)bCs.reserve(mbsize * 4);
        HIP_CHECK(hipMemcpyAsync(C, centers, kd * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                           : hipMemcpyHostToDevice, st));
        copy_f32(C, Save, kd, st);
        Data sub;
        std::vector<uint64_t> sel(mbsize);
        std::vector<uint8_t> seen(with_replacement ? 0 : n);
        uint64_t j = 0, iter = 0;
        double init = DBL_MAX, best = DBL_MAX;
        for (;;) {
            if (iter % freq == 0 || iter == maxiter - 1) {
                const double cost = ck.run(C, seed, j, st);
                if (iter == 0) init = best = cost;
                if (cost < best) {
                    best = cost;
                    copy_f32(C, Save, kd, st);
                }
            }
            if (++iter == maxiter) break;
            if (with_replacement) {
                for (uint64_t t = 0; t < mbsize; ++t) sel[t] = wyrand_at(seed, j++) % n;
            } else {
                std::fill(seen.begin(), seen.end(), (uint8_t)0);
                for (uint64_t t = 0; t < mbsize;) {
                    const uint64_t r = wyrand_at(seed, j++) % n;
                    if (!seen[r]) { seen[r] = 1; sel[t++] = r; }
                }
            }
            std::sort(sel.begin(), sel.end());
            copy_sync(idx_d, sel.data(), mbsize * 8, hipMemcpyHostToDevice, st);
            rows_into(D, sel.data(), idx_d, mbsize, sub, st);
            if (Ws)
                hipLaunchKernelGGL(k_gather_f64, dim3((uint32_t)cdiv(mbsize, 256)), dim3(256), 0, st, W, idx_d, mbsize,
                                   Ws);
            estep(sub, msr, prior, C, k, precision, As, l2 ? nullptr : Cs, st);
            if (median_policy(msr)) median_update(sub, msr, Ws, W, k, As, C, st);
            else if (sub.csr) csr_mstep(sub, As, Ws, k, C, st);
            else dense_mstep(sub, As, Ws, k, C, C, false, st);
        }
        Out oa = out_dev(c, asn, n * 4, sa_out);
        Out ok = out_dev(c, costs, n * 4, sk_out);
        if (oa.dev) HIP_CHECK(hipMemcpyAsync(oa.dev, ck.A, n * 4, hipMemcpyDeviceToDevice, st));
        if (ok.dev) HIP_CHECK(hipMemcpyAsync(ok.dev, ck.P, n * 4, hipMemcpyDeviceToDevice, st));
        out_finish(c, oa);
        out_finish(c, ok);
        HIP_CHECK(hipMemcpyAsync(centers, Save, kd * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                              : hipMemcpyDeviceToHost, st));
        sync(c);
        *initcost = init;
        *finalcost = best;
        *numiter = iter;
    });
}

// hmb_coreset_clustering (solve.h:598-789).  ceil(maxiter / freq) + 1 check-ins
// (the saved centers change only when a later check-in is strictly cheaper); after
// each but the last, LBK sensitivities from the check-in's costs and assignment
// (make_sampler without an alias sampler), then freq coresets: m = min(mbsize, n)
// i.i.d. draws seeded by the stream's next output (sample_k with replacement,
// sorted), weights w_i * w_i / (m p_i) (solve.h:771-772 multiplies the row weights
// into coreset weights that already hold them), and perform_hard_clustering on the
// coreset rows from the current centers, eps 1e-3, maxiter 1 + 2 (round < 2).
int mcg_cs_minibatch(mcg_data *ds, int msr, double prior, float *centers, uint64_t k, const double *weights,
                     uint64_t mbsize, uint64_t maxiter, uint64_t freq, uint64_t reseed_after, uint64_t seed,
                     int precision, uint32_t *asn, float *costs, double *initcost, double *finalcost,
                     uint64_t *numiter) {
    return guard([&] {
        REQUIRE(ds && centers && initcost && finalcost && numiter, "null argument");
        REQUIRE(k > 0, "k must be positive");
        REQUIRE(mbsize > 0 && freq > 0, "mbsize and the check-in period must be positive");
        check_hard_msr(ds->d, msr, weights);
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n, d = D.d, kd = k * d;
        const uint64_t rounds = (freq - 1 + maxiter) / freq, m = std::min(mbsize, n);
        DBuf sw, sa_out, sk_out, bC, bSave, bA, bP, bTmp, bTa, bCnt, bCf, bFid, bCost, bProbs, bIdx, bCw, bIdxS,
            bCsw, bAs;
        const double *W = (const double *)in_dev(c, weights, n * 8, sw);
        std::vector<double> wh;
        if (weights) {
            wh.resize(n);
            if (is_device_ptr(weights)) copy_sync(wh.data(), weights, n * 8, hipMemcpyDeviceToHost, st);
            else std::memcpy(wh.data(), weights, n * 8);
        }
        float *C = (float *)bC.reserve(kd * 4), *Save = (float *)bSave.reserve(kd * 4);
        CheckIn ck{D, msr, prior, k, precision, W, wh, reseed_after, (uint32_t *)bA.reserve(n * 4),
                   (uint32_t *)bTa.reserve(n * 4), (uint32_t *)bCnt.reserve(k * 4), (uint32_t *)bFid.reserve(k * 4),
                   (float *)bP.reserve(n * 4), (float *)bTmp.reserve(n * 4), (float *)bCf.reserve(kd * 4), {}, {}};
        double *cost64 = (double *)bCost.reserve(n * 8);
        float *probs = (float *)bProbs.reserve(n * 4);
        uint64_t *idx_d = (uint64_t *)bIdx.reserve(m * 8), *idx_s = (uint64_t *)bIdxS.reserve(m * 8);
        float *cw = (float *)bCw.reserve(m * 4);
        double *csw = (double *)bCsw.reserve(m * 8);
        uint32_t *As = (uint32_t *)bAs.reserve(m * 4);
        HIP_CHECK(hipMemcpyAsync(C, centers, kd * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                           : hipMemcpyHostToDevice, st));
        copy_f32(C, Save, kd, st);
        mcg_data sub;
        std::vector<uint64_t> hi(m);
        std::vector<float> hw(m);
        std::vector<std::pair<uint64_t, double>> pr(m);
        std::vector<uint64_t> si(m);
        std::vector<double> sw_h(m);
        uint64_t j = 0, iter = 0;
        double init = DBL_MAX, best = DBL_MAX;
        const uint32_t pgrid = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 4096);
        for (;;) {
            const double cost = ck.run(C, seed, j, st);
            if (iter == 0) init = best = cost;
            else if (cost < best) {
                best = cost;
                copy_f32(C, Save, kd, st);
            }
            if (++iter > rounds) break;
            hipLaunchKernelGGL(k_f32_to_f64_copy, dim3(pgrid), dim3(256), 0, st, ck.P, n, cost64);
            coreset_probs_run(c, MCG_SENS_LBK, n, k, cost64, ck.A, W, probs, st);
            for (uint64_t jj = 0; jj < freq; ++jj) {
                const uint64_t r = wyrand_at(seed, j++);
                coreset_sample_from(c, n, probs, W, m, r, 0, idx_d, cw, st);
                copy_sync(hi.data(), idx_d, m * 8, hipMemcpyDeviceToHost, st);
                copy_sync(hw.data(), cw, m * 4, hipMemcpyDeviceToHost, st);
                for (uint64_t t = 0; t < m; ++t) pr[t] = {hi[t], (wh.empty() ? 1. : wh[hi[t]]) * (double)hw[t]};
                std::stable_sort(pr.begin(), pr.end(),
                                 [](const std::pair<uint64_t, double> &a, const std::pair<uint64_t, double> &b) {
                                     return a.first < b.first;
                                 });
                for (uint64_t t = 0; t < m; ++t) { si[t] = pr[t].first; sw_h[t] = pr[t].second; }
                copy_sync(idx_s, si.data(), m * 8, hipMemcpyHostToDevice, st);
                copy_sync(csw, sw_h.data(), m * 8, hipMemcpyHostToDevice, st);
                rows_into(D, si.data(), idx_s, m, sub.d, st);
                double ic = 0., fc = 0.;
                uint64_t it = 0;
                const int rc = mcg_lloyd(&sub, msr, prior, C, k, csw, 1e-3, 1 + 2 * (iter < 2), precision, As,
                                         nullptr, &ic, &fc, &it);
                if (rc != MCG_OK) MCG_THROW(rc, g_err);
            }
        }
        Out oa = out_dev(c, asn, n * 4, sa_out);
        Out ok = out_dev(c, costs, n * 4, sk_out);
        if (oa.dev) HIP_CHECK(hipMemcpyAsync(oa.dev, ck.A, n * 4, hipMemcpyDeviceToDevice, st));
        if (ok.dev) HIP_CHECK(hipMemcpyAsync(ok.dev, ck.P, n * 4, hipMemcpyDeviceToDevice, st));
        out_finish(c, oa);
        out_finish(c, ok);
        HIP_CHECK(hipMemcpyAsync(centers, Save, kd * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                              : hipMemcpyDeviceToHost, st));
        sync(c);
        *initcost = init;
        *finalcost = best;
        *numiter = iter;
    });
}

// perform_soft_clustering (solve.h:287-336) with set_centroids_soft (:342-374).
int mcg_soft_em(mcg_data *ds, int msr, double prior, double temp, float *centers, uint64_t k, const double *weights,
                double eps, uint64_t maxiter, float *costs, float *resp, double *initcost, double *finalcost,
                uint64_t *numiter) {
    return guard([&] {
        REQUIRE(ds && centers, "null argument");
        REQUIRE(k > 0, "k must be positive");
        check_msr(msr);
        // set_centroids_full_mean (soft, centroid.h:795-963): every measure takes the
        // responsibility-weighted mean except L1 and L2, which take the weighted median
        // and the weighted geometric median over all rows (centroid.h:815-838)
        const bool med = msr == MCG_L1 || msr == MCG_L2;
        Data &D = ds->d;
        if (msr == MCG_L1 && D.csr) MCG_THROW(MCG_EUNSUPPORTED, "Not implemented: weighted L1 median for CSparseMatrix");
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n, d = D.d;
        DBuf sw, so1, so2;
        const double *W = (const double *)in_dev(c, weights, n * 8, sw);
        float *C = (float *)c.s[S_CENT_A].reserve(k * d * 4);
        float *Cc = (float *)c.s[S_CENT_B].reserve(k * d * 4);
        float *K = (float *)c.s[S_STAGE0].reserve(n * k * 4);
        float *R = (float *)c.s[S_STAGE1].reserve(n * k * 4);
        double *rowdot = (double *)c.s[S_W].reserve(n * 8);
        double *tot = (double *)c.s[S_TOT].reserve(4 * sizeof(double));
        HIP_CHECK(hipMemcpyAsync(C, centers, k * d * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                               : hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(Cc, C, k * d * 4, hipMemcpyDeviceToDevice, st));
        generic_costs(D, msr, prior, Cc, k, K, nullptr, nullptr, st);
        double cost = 1.7976931348623157e308, init = -1.;
        uint64_t iter = 0;
        for (;;) {
            const double oldcost = cost;
            softmax_rows(K, n, k, (float)temp, R, W, rowdot, st);
            hipLaunchKernelGGL(k_wsum_d, dim3(1), dim3(1024), 0, st, rowdot, n, tot);
            if (med) soft_median_mstep(D, msr, R, W, k, Cc, K, st);   // K is rewritten below
            else soft_mstep(D, R, W, k, Cc, st);
            generic_costs(D, msr, prior, Cc, k, K, nullptr, nullptr, st);
            HIP_CHECK(hipMemcpyAsync(&cost, tot, sizeof(double), hipMemcpyDeviceToHost, st));
            sync(c);
            if (init < 0.) init = cost;
            if (oldcost >= cost) HIP_CHECK(hipMemcpyAsync(C, Cc, k * d * 4, hipMemcpyDeviceToDevice, st));
            if (oldcost - cost <= eps * std::max(oldcost, cost) || ++iter == maxiter) break;
        }
        Out ok = out_dev(c, costs, n * k * 4, so1);
        Out orr = out_dev(c, resp, n * k * 4, so2);
        if (ok.dev) HIP_CHECK(hipMemcpyAsync(ok.dev, K, n * k * 4, hipMemcpyDeviceToDevice, st));
        if (orr.dev) HIP_CHECK(hipMemcpyAsync(orr.dev, R, n * k * 4, hipMemcpyDeviceToDevice, st));
        out_finish(c, ok);
        out_finish(c, orr);
        HIP_CHECK(hipMemcpyAsync(centers, C, k * d * 4, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                               : hipMemcpyDeviceToHost, st));
        sync(c);
        if (initcost) *initcost = init;
        if (finalcost) *finalcost = cost;
        if (numiter) *numiter = iter;
    });
}

int mcg_kmeanspp(mcg_data *ds, int msr, double prior, uint64_t k, uint64_t seed, const double *weights,
                 uint32_t *centers, uint32_t *asn, float *costs) {
    return guard([&] {
        REQUIRE(ds && centers, "null argument");
        REQUIRE(k > 0, "k must be positive");
        check_msr(msr);
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n;
        if (k > n) k = n;  // kmeans.h:59-62
        DBuf sw, sa, sk;
        const double *W = (const double *)in_dev(c, weights, n * 8, sw);
        Out oa = out_dev(c, asn, n * 4, sa);
        Out ok = out_dev(c, costs, n * 4, sk);
        uint32_t *A = oa.dev ? (uint32_t *)oa.dev : (uint32_t *)c.s[S_ASN_A].reserve(n * 4);
        float *Dd = ok.dev ? (float *)ok.dev : (float *)c.s[S_COSTS].reserve(n * 4);
        std::vector<uint32_t> ch(k);
        kmeanspp_run(D, msr, prior, k, seed, W, ch.data(), A, Dd, st);
        if (is_device_ptr(centers)) HIP_CHECK(hipMemcpyAsync(centers, ch.data(), k * 4, hipMemcpyHostToDevice, st));
        else std::memcpy(centers, ch.data(), k * 4);
        out_finish(c, oa);
        out_finish(c, ok);
        sync(c);
    });
}

int mcg_kmeanspp_ls(mcg_data *ds, int msr, double prior, uint64_t k, uint64_t seed, const double *weights,
                    uint64_t rounds, uint32_t *centers, uint32_t *asn, float *costs) {
    return guard([&] {
        REQUIRE(ds && centers, "null argument");
        REQUIRE(k > 0, "k must be positive");
        check_msr(msr);
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n;
        if (k > n) k = n;
        REQUIRE(rounds == 0 || (double)n * (double)k * 4. < 64e9, "local search++ keeps an n x k cost matrix: too large");
        DBuf sw, sa, sk;
        const double *W = (const double *)in_dev(c, weights, n * 8, sw);
        Out oa = out_dev(c, asn, n * 4, sa);
        Out ok = out_dev(c, costs, n * 4, sk);
        uint32_t *A = oa.dev ? (uint32_t *)oa.dev : (uint32_t *)c.s[S_ASN_A].reserve(n * 4);
        float *Dd = ok.dev ? (float *)ok.dev : (float *)c.s[S_COSTS].reserve(n * 4);
        std::vector<uint32_t> ch(k);
        kmeanspp_run(D, msr, prior, k, seed, W, ch.data(), A, Dd, st);
        localsearch_run(D, msr, prior, k, seed, W, rounds, k, ch.data(), A, Dd, st);
        if (is_device_ptr(centers)) HIP_CHECK(hipMemcpyAsync(centers, ch.data(), k * 4, hipMemcpyHostToDevice, st));
        else std::memcpy(centers, ch.data(), k * 4);
        out_finish(c, oa);
        out_finish(c, ok);
        sync(c);
    });
}

int mcg_kmeanspp_local(mcg_data *ds, int msr, double prior, uint64_t k, uint64_t seed, const double *weights,
                       uint64_t trials, uint64_t rounds, uint32_t *centers, uint32_t *asn, float *costs) {
    return guard([&] {
        REQUIRE(ds && centers, "null argument");
        REQUIRE(k > 0, "k must be positive");
        REQUIRE(trials >= 1, "n_local_trials must be positive");
        check_msr(msr);
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n;
        if (k > n) k = n;
        REQUIRE(trials == 1 || (double)n * (double)(trials + 3) * 8. < 64e9,
                "local trials keep n_local_trials copies of the distances and assignment: too large");
        REQUIRE(rounds == 0 || (double)n * (double)k * 4. < 64e9, "local search++ keeps an n x k cost matrix: too large");
        DBuf sw, sa, sk;
        const double *W = (const double *)in_dev(c, weights, n * 8, sw);
        Out oa = out_dev(c, asn, n * 4, sa);
        Out ok = out_dev(c, costs, n * 4, sk);
        uint32_t *A = oa.dev ? (uint32_t *)oa.dev : (uint32_t *)c.s[S_ASN_A].reserve(n * 4);
        float *Dd = ok.dev ? (float *)ok.dev : (float *)c.s[S_COSTS].reserve(n * 4);
        std::vector<uint32_t> ch(k);
        uint64_t j0 = k;   // stream outputs the seeding used
        if (trials > 1) {
            kmeanspp_local_run(D, msr, prior, k, seed, W, trials, ch.data(), A, Dd, st);
            j0 = trials + k - 1;
        } else {
            kmeanspp_run(D, msr, prior, k, seed, W, ch.data(), A, Dd, st);
        }
        localsearch_run(D, msr, prior, k, seed, W, rounds, j0, ch.data(), A, Dd, st);
        if (is_device_ptr(centers)) HIP_CHECK(hipMemcpyAsync(centers, ch.data(), k * 4, hipMemcpyHostToDevice, st));
        else std::memcpy(centers, ch.data(), k * 4);
        out_finish(c, oa);
        out_finish(c, ok);
        sync(c);
    });
}

int mcg_coreset_probs(mcg_ctx *ctx, int sens, uint64_t n, uint64_t ncenters, const double *costs,
                      const uint32_t *asn, const double *weights, float *probs) {
    return guard([&] {
        REQUIRE(ctx && costs && asn && probs, "null argument");
        REQUIRE(n > 0 && ncenters > 0, "empty input");
        Ctx &c = ctx->c;
        HIP_CHECK(hipSetDevice(c.device));
        DBuf s1, s2, s3, s4;
        const double *K = (const double *)in_dev(c, costs, n * 8, s1);
        const uint32_t *A = (const uint32_t *)in_dev(c, asn, n * 4, s2);
        const double *W = (const double *)in_dev(c, weights, n * 8, s3);
        {
            std::vector<uint32_t> h(n);
            HIP_CHECK(hipMemcpyAsync(h.data(), A, n * 4, hipMemcpyDeviceToHost, c.stream));
            sync(c);
            for (uint64_t i = 0; i < n; ++i) REQUIRE(h[i] < ncenters, "assignment out of range");
        }
        Out o = out_dev(c, probs, n * 4, s4);
        coreset_probs_run(c, sens, n, ncenters, K, A, W, (float *)o.dev, c.stream);
        out_finish(c, o);
        sync(c);
    });
}

int mcg_coreset_sample(mcg_ctx *ctx, uint64_t n, const float *probs, const double *weights, uint64_t m,
                       uint64_t seed, uint64_t first_draw, uint64_t *indices, float *out_weights) {
    return guard([&] {
        REQUIRE(ctx && probs && indices && out_weights, "null argument");
        REQUIRE(n > 0, "empty input");
        Ctx &c = ctx->c;
        HIP_CHECK(hipSetDevice(c.device));
        DBuf s1, s2, s3, s4;
        const float *P = (const float *)in_dev(c, probs, n * 4, s1);
        const double *W = (const double *)in_dev(c, weights, n * 8, s2);
        Out oi = out_dev(c, indices, (m ? m : 1) * 8, s3);
        Out ow = out_dev(c, out_weights, (m ? m : 1) * 4, s4);
        coreset_sample_from(c, n, P, W, m, seed, first_draw, (uint64_t *)oi.dev, (float *)ow.dev, c.stream);
        out_finish(c, oi);
        out_finish(c, ow);
        sync(c);
    });
}

int mcg_graphs_enable(mcg_ctx *ctx, int on) {
    return guard([&] { REQUIRE(ctx, "null ctx"); ctx->c.graphs = on != 0; });
}

int mcg_tuning_enable(mcg_ctx *ctx, int on) {
    return guard([&] { REQUIRE(ctx, "null ctx"); ctx->c.tuning = on != 0; });
}

int mcg_timing_enable(mcg_ctx *ctx, int on) {
    return guard([&] { REQUIRE(ctx, "null ctx"); ctx->c.timing = on != 0; });
}

int mcg_timing_reset(mcg_ctx *ctx) {
    return guard([&] {
        REQUIRE(ctx, "null ctx");
        sync(ctx->c);
        for (int w = 0; w < 2; ++w) { ctx->c.timed_ms[w] = 0.; ctx->c.timed_n[w] = 0; }
    });
}

int mcg_timing_read(mcg_ctx *ctx, int which, double *total_ms, uint64_t *count) {
    return guard([&] {
        REQUIRE(ctx && (which == 0 || which == 1), "invalid argument");
        sync(ctx->c);
        if (total_ms) *total_ms = ctx->c.timed_ms[which];
        if (count) *count = ctx->c.timed_n[which];
    });
}

// lloyd_iteration (kmeans.h:359-479): one E-step and one M-step on device state.
// An empty cluster takes a row drawn with probability ~ w_j * cost_j (the cost of
// row j to its new mean for dense SQRL2, kmeans.h:433-440; the E-step minimum for
// the other measures), that row moves to it, and the means are formed again, until
// no cluster is empty (kmeans.h:431-454).  Draw e of one call is output e of the
// wyrand stream of seed 0 (lloyd_iteration's default seed) through the fixed-point
// sampler, the drawn row's weight set to zero for the later draws.
int mcg_lloyd_step(mcg_data *ds, int msr, double prior, float *centers, uint64_t k, const double *weights,
                   int precision, uint32_t *asn, double *cost) {
    return guard([&] {
        REQUIRE(ds && centers && asn, "null argument");
        REQUIRE(k > 0, "k must be positive");
        check_mean_msr(msr);
        REQUIRE(is_device_ptr(centers) && is_device_ptr(asn), "mcg_lloyd_step needs device centers and asn");
        Data &D = ds->d;
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n;
        DBuf sw;
        const double *W = (const double *)in_dev(c, weights, n * 8, sw);
        double *res = (double *)c.s[S_MISC3].reserve(2 * sizeof(double));
        float *pc = nullptr;
        MStepOut mo;
        const bool l2 = dense_l2(D, msr);
        // {cost, clusters without members}: written by the M-step's last block straight
        // into pinned host memory (dense SQRL2), else summarised and copied
        double *hout = (double *)c.pinned(2 * sizeof(double));
        if (l2) {
            dense_sqrl2_assign(D, centers, k, precision, asn, st);
            mo = dense_mstep(D, asn, W, k, centers, centers, true, st, hout);
        } else {
            pc = (float *)c.s[S_COSTS].reserve(n * 4);
            double *tot = (double *)c.s[S_TOT].reserve(4 * sizeof(double));
            generic_costs(D, msr, prior, centers, k, nullptr, asn, pc, st);
            hipLaunchKernelGGL(k_wsum, dim3(1), dim3(1024), 0, st, pc, W, n, tot);
            mo = D.csr ? csr_mstep(D, asn, W, k, centers, st) : dense_mstep(D, asn, W, k, centers, centers, false, st);
            mo.total = tot;
            hipLaunchKernelGGL(k_step_summary, dim3(1), dim3(256), 0, st, mo.total, mo.counts, k, res);
            HIP_CHECK(hipMemcpyAsync(hout, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        }
        sync(c);
        const double h[2] = {hout[0], hout[1]};
        if (cost) *cost = h[0];
        if (h[1] == 0.) return;
        // reseed empty clusters (rare): host draws, device moves and means
        std::vector<double> wh;
        if (weights) {
            wh.resize(n);
            copy_sync(wh.data(), W, n * 8, hipMemcpyDeviceToHost, st);
        }
        if (l2) pc = (float *)c.s[S_COSTS].reserve(n * 4);
        std::vector<uint32_t> hc(k);
        std::vector<float> hp(n);
        uint64_t j = 0;
        for (uint64_t round = 0; round <= k; ++round) {
            copy_sync(hc.data(), mo.counts, k * 4, hipMemcpyDeviceToHost, st);
            std::vector<uint32_t> ids;
            for (uint64_t q = 0; q < k; ++q) if (!hc[q]) ids.push_back((uint32_t)q);
            if (ids.empty()) return;
            REQUIRE(round < k, "internal: empty clusters after k reseeds");
            if (l2) dense_sqrl2_costs(D, centers, asn, pc, st);
            copy_sync(hp.data(), pc, n * 4, hipMemcpyDeviceToHost, st);
            std::vector<double> wv(n);
            for (uint64_t i = 0; i < n; ++i) wv[i] = wh.empty() ? (double)hp[i] : wh[i] * (double)hp[i];
            FixedSeq seq(std::move(wv));
            std::vector<uint64_t> rows(ids.size());
            for (size_t e = 0; e < ids.size(); ++e) {
                rows[e] = seq.draw(wyrand_at(0, j++));
                seq.zero(rows[e]);
            }
            copy_rows_to_centers(D, rows.data(), ids.data(), ids.size(), centers, st);
            char *buf = (char *)c.s[S_MISC2].reserve(ids.size() * 12);
            HIP_CHECK(hipMemcpyAsync(buf, rows.data(), ids.size() * 8, hipMemcpyHostToDevice, st));
            HIP_CHECK(hipMemcpyAsync(buf + ids.size() * 8, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_force, dim3((uint32_t)cdiv(ids.size(), 256)), dim3(256), 0, st, asn, pc,
                               (const uint64_t *)buf, (const uint32_t *)(buf + ids.size() * 8), (uint64_t)ids.size());
            mo = D.csr ? csr_mstep(D, asn, W, k, centers, st) : dense_mstep(D, asn, W, k, centers, centers, false, st);
            sync(c);
        }
    });
}

int mcg_refine_counts(mcg_ctx *ctx, uint64_t *pairs, uint64_t *full) {
    return guard([&] {
        REQUIRE(ctx, "null ctx");
        Ctx &c = ctx->c;
        uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (c.refine_cnt) copy_sync(h, c.refine_cnt, sizeof(h), hipMemcpyDeviceToHost, c.stream);
        else sync(c);
        if (pairs) *pairs = h[4];
        if (full) *full = h[1];
    });
}

}  // extern "C"
