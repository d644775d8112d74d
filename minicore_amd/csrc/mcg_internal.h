// Internal declarations shared by the libmcg translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <new>
#include <string>
#include <vector>
#include "../../include/mcg.h"

namespace mcg {

// ---- error plumbing -------------------------------------------------------
void set_error(const std::string &msg);

struct Error {
    int code;
    std::string msg;
};

#define MCG_THROW(code, msg) throw ::mcg::Error{(code), (msg)}
#define HIP_CHECK(expr)                                                                      \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            MCG_THROW(_e == hipErrorOutOfMemory ? MCG_ENOMEM : MCG_EDEVICE,                  \
                      std::string(#expr) + ": " + hipGetErrorString(_e));                    \
    } while (0)
#define REQUIRE(cond, msg) do { if (!(cond)) MCG_THROW(MCG_EINVAL, (msg)); } while (0)

// Runs f, turning exceptions into a status code and the thread's error message.
template <class F>
int guard(F &&f) {
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

// ---- device buffers -------------------------------------------------------
// Stream that device buffers created on this thread are bound to (a server worker
// sets its context's stream for its lifetime; elsewhere none).
inline thread_local hipStream_t g_dbuf_stream = nullptr;

// A device buffer.  Bound to a stream (st), it is allocated and freed in that stream's
// order from the device's memory pool (hipMallocAsync / hipFreeAsync), so freeing it
// does not wait for the work of other streams: hipFree waits for the whole device,
// which would make concurrent server workers wait for each other at every request's
// end.  A bound buffer may only be used by work on its stream, and the stream must
// outlive it.  Unbound, it uses hipMalloc / hipFree.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t st = g_dbuf_stream;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    // (while st is being captured into a graph, the plain calls: a stream-ordered
    // allocation there would belong to the graph)
    bool ordered() const {
        if (!st) return false;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        return !(hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone);
    }
    void release() {
        if (p) {
            if (ordered()) (void)hipFreeAsync(p, st);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
    // grow-only; contents are not preserved
    void *reserve(size_t nb) {
        if (nb > bytes) {
            release();
            if (ordered()) HIP_CHECK(hipMallocAsync(&p, nb ? nb : 1, st));
            else HIP_CHECK(hipMalloc(&p, nb ? nb : 1));
            bytes = nb;
        }
        return p;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Host-or-device classification of a caller pointer.
bool is_device_ptr(const void *p);

struct Ctx;

// A caller buffer made available on the device: borrowed when it already is
// device memory, else staged through a context scratch buffer.
struct DevView {
    const void *dptr = nullptr;
    DBuf staged;
};

// ---- tiled sparse products (mcg_tiles.hip) ------------------------------------
// A sparse matrix reorganised into tiles of tg accumulator indices (twr per wave, 16
// waves) x ts operand rows: per tile the per-index end offsets inside the wave's list
// (u8), the 16 waves' base offsets (u16) and the tile's first entry (u32); entries
// (byte offset of the operand row in the stage buffer, value bits).  cpl: centers per
// lane of the kernel the tiles were built for (0: none built).
struct Tiles {
    bool ready = false;
    int cpl = 0;
    uint32_t twr = 0, tg = 0, ts = 0, nw = 0;   // indices a wave, a block; stage rows; waves a block
    uint64_t nacc = 0, nop = 0;
    uint32_t ngroups = 0, nstages = 0;
    uint32_t maxseg = 0, maxwave = 0;   // largest stage segment of a block / of a wave
    DBuf cnt, wb, off, ent;
};

// Chunks per cluster / column of the CSR M-steps: up to 8, while the Z x k x d FP64
// partial sums stay under about 2 GB (MCG_CSR_CHUNKS = 1..8 overrides, read per call).
inline uint32_t csr_chunks(uint64_t kd) {
    const char *e = getenv("MCG_CSR_CHUNKS");
    const int f = e ? atoi(e) : 0;
    if (f >= 1 && f <= 8) return (uint32_t)f;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(8, (uint64_t)(2e9 / 8.) / kd));
}

// ---- dataset ----------------------------------------------------------------
struct Data {
    Ctx *ctx = nullptr;
    bool csr = false;
    uint64_t n = 0, d = 0, nnz = 0;
    // dense rows (float32, row-major, stride d)
    float *X = nullptr;
    DBuf Xown;
    // rows given as float64: the values as given (row-major, stride d), for the
    // float64 path of dense SQRL2 hard clustering (mcg_f64.hip)
    DBuf Xd;
    // CSR rows (canonical: f32 data, u32 indices, u64 indptr)
    DBuf val, idx, ptr;
    // per-row FP64 sums (csc.h:1008-1016 / blaze sum<rowwise>)
    DBuf rowsum;
    // 16-bit copies (BF16, FP16) for the MFMA cross term: rows centered on the column
    // mean mu, padded to dpad columns; ||x - mu|| per row; max over rows.
    bool base_ready = false, bf16_ready = false, f16_ready = false;
    bool oneshot = false;   // one E-step is asked of these rows (a server's final assignment)
    uint32_t dpad = 0;
    DBuf Xb, mu, xnorm;
    DBuf xdev;         // ||bf16(x - mu) - (x - mu)|| per row (rounded up)
    DBuf Xh;           // FP16 copy of xscale * (x - mu), padded to dpad columns
    DBuf xdev_h;       // ||Xh / xscale - (x - mu)|| per row (rounded up)
    DBuf xstat;        // device floats: [0] xmax = max ||x - mu||, [1] xscale (FP16 form)
    // per-prior CSR transforms (cached for the last prior seen)
    double tf_prior = -1.;
    DBuf tf_logx;      // dense rows: log1p(x / pv) per value (MKL on MFMA), for the prior tf_logx_pv
    double tf_logx_pv = -1.;
    DBuf rowinfo;      // per-row float4 {L, log L, Hx, li}
    std::vector<float> li_unique;  // distinct li values (JSD)
    DBuf li_dev;       // li_unique on the device
    DBuf li_slot;      // per-row index into li_unique
    // column-major copy of the CSR rows (soft M-step), built on first use
    DBuf csc_ptr, csc_row, csc_val;
    // tiles of the rows (E-step) and of the columns (soft M-step), built on first use
    Tiles rtiles, ctiles;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // named scratch buffers
    DBuf s[56];
    DBuf pinned_dummy;
    void *host_pinned = nullptr;
    size_t host_pinned_bytes = 0;
    void *pinned(size_t nb);
    // kernel timing: pending (start, stop) event pairs per timed kernel
    bool timing = false;
    // device counters of the last 16-bit E-step: [4] pair points, [1] listed points
    const uint32_t *refine_cnt = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[2];
    double timed_ms[2] = {0., 0.};
    uint64_t timed_n[2] = {0, 0};
    // E-step kernel chosen per problem shape by timing (see dense_sqrl2_assign)
    bool tuning = true;
    // a server worker: shape choices come from the fixed model, never from timing runs
    // inside a request
    bool serving = false;
    // Lloyd iterations after the first replay one captured HIP graph (off by default:
    // while a capture is open, HIP rejects legacy-stream calls from other threads)
    bool graphs = false;
    void tic(int which, hipEvent_t &a);
    void toc(int which, hipEvent_t a);
    void collect();   // after a stream synchronize: fold pending pairs into the totals
    ~Ctx();
};

// scratch slot ids
enum Slot {
    S_CB = 0, S_CADD, S_CNORM, S_PARAMS, S_FLAGS, S_FLAGCNT, S_ASN_A, S_ASN_B,
    S_HIST, S_BASE, S_PERM, S_CNT, S_ITEMOFF, S_PART, S_PARTW, S_PARTQ,
    S_CCOST, S_TOTAL, S_CENT_A, S_CENT_B, S_CENT_C, S_COSTS, S_CT, S_CPARAM,
    S_MISC0, S_MISC1, S_MISC2, S_MISC3, S_W, S_STAGE0, S_STAGE1, S_STAGE2,
    S_LIS, S_TOT, S_JSDG, S_PAIRS, S_THRESH, S_OVER, S_KEYS, S_OFLAG, S_CINFO, S_CENT_D, S_LCTL, S_PREP, S_ROWLEN, S_MED0, S_MED1, S_MED2, S_MED3, S_CTM, S_JSDT, S_WSTART, S_NSLOTS
};
static_assert(S_NSLOTS <= 56, "scratch slots");

// ---- kernels (launchers) ----------------------------------------------------
// E-step parameters for the BF16 near-tie test, written on the device.
struct EParams {
    float off;      // added to every approximate value to keep it positive
    float cmax;     // max_c ||c - mu||
    float csc;      // FP16 form: power-of-two scale of the center operand (1 for BF16)
    float tau0;     // absolute part of the tie threshold
    float taur;     // relative part (of the approximate best distance)
    float dbmax;    // max_c ||c~ - (c - mu)||, c~ the 16-bit center operand (unscaled)
    float s;        // scale of the MFMA values: x scale * center scale (1 for BF16)
    float inv_s;    // 1 / s (exact: powers of two)
    int c0nan;      // center 0 has a NaN: every row's first cost is NaN
    int pad[3];
};

void prepare_base(Data &ds);  // mu, xnorm, xmax, dpad (once per dataset)
void prepare_bf16(Data &ds);  // + Xb, xdev (once per dataset)
void prepare_f16(Data &ds);   // + Xh, xdev_h, xscale (once per dataset)
// Dense SQRL2 hard assignment into asn (device), by precision.
void dense_sqrl2_assign(Data &ds, const float *C, uint64_t k, int precision, uint32_t *asn,
                        hipStream_t st);
// Exact per-point costs ||x_i - c_asn(i)||^2 (FP32 sequential FMA) into costs (device).
void dense_sqrl2_costs(Data &ds, const float *C, const uint32_t *asn, float *costs, hipStream_t st);
// Dense SQRL2 assignment of every point against every center, FP32 sequential FMA.
void dense_sqrl2_assign_exact(Data &ds, const float *C, uint64_t k, uint32_t *asn, float *costs,
                              hipStream_t st);

// M-step: stable counting sort by asn, FP64 segmented sums; new centers and
// per-cluster costs against Cold.  Writes counts (u32[k]) and total cost
// (f64 at S_TOTAL) on the device.
struct MStepOut {
    uint32_t *counts;
    double *total;   // [0] = cost vs Cold (SQRL2 dense only), [1] = total weight; dense: [2] = empty clusters
};
// host_out: pinned host memory that receives {total[0], total[2]} (or null)
MStepOut dense_mstep(Data &ds, const uint32_t *asn, const double *w, uint64_t k, const float *Cold,
                     float *Cnew, bool want_cost, hipStream_t st, double *host_out = nullptr);
// Stable counting sort of point ids by assignment (ids >= k left out): perm holds
// each cluster's members contiguous and ascending, at off[c], counts[c] of them;
// itemoff the prefix of 256-member chunks.  k <= 16384.
struct SortOut {
    uint32_t *perm, *counts, *off, *itemoff, *imap;   // imap: cluster of each chunk item
};
SortOut sort_by_cluster(Ctx &c, const uint32_t *asn, uint64_t n, uint64_t k, hipStream_t st);
// Exclusive prefix sum of L u32 values (tmp: scratch for block totals).
void scan_exclusive(const uint32_t *in, uint32_t *out, uint64_t L, DBuf &tmp, hipStream_t st);
// Median centroid policies (L1, TVD, L2 -> geometric median), centroid.h:434-647:
// the M-step from the accepted assignment and costs (Ar, pr: working copies that
// the empty-cluster restart updates), previous centers in C, new centers into C.
inline int median_policy(int msr) { return msr == MCG_L1 ? 1 : msr == MCG_TVD ? 2 : msr == MCG_L2 ? 3 : 0; }
void median_mstep(Data &ds, int msr, const double *w, uint64_t k, uint32_t *Ar, float *pr, float *C,
                  hipStream_t st);
// The minibatch update's median centers (solve.h:555-572): wm per-row weights of ds
// (the geometric median), wr the weight vector indexed by member rank (L1, TVD).
void median_update(Data &ds, int msr, const double *wm, const double *wr, uint64_t k, const uint32_t *asn,
                   float *C, hipStream_t st);
// true if a stored CSR value is negative (the CSR median rules assume non-negative data)
bool csr_has_negative(Data &ds, hipStream_t st);
MStepOut csr_mstep(Data &ds, const uint32_t *asn, const double *w, uint64_t k, float *Cnew,
                   hipStream_t st);

// Measures with a kernel here: the four with decomposed kernels, and the further ones
// evaluated pair by pair.
inline bool msr_decomposed(int msr) {
    return msr == MCG_SQRL2 || msr == MCG_JSD || msr == MCG_MKL || msr == MCG_REVERSE_MKL;
}
inline bool msr_supported(int msr) {
    switch (msr) {
    case MCG_L1: case MCG_L2: case MCG_SQRL2: case MCG_JSM: case MCG_JSD: case MCG_MKL: case MCG_HELLINGER:
    case MCG_BHATTACHARYYA_METRIC: case MCG_BHATTACHARYYA_DISTANCE: case MCG_TVD: case MCG_LLR:
    case MCG_REVERSE_MKL: case MCG_UWLLR: case MCG_IS: case MCG_REVERSE_IS: case MCG_COSINE_DISTANCE:
    case MCG_PROBABILITY_COSINE_DISTANCE: case MCG_COSINE_SIMILARITY: case MCG_PROBABILITY_COSINE_SIMILARITY:
    case MCG_SIS: case MCG_RSIS: case MCG_SRLRT: case MCG_SRULRT: case MCG_POISSON:
        return true;
    }
    return false;
}
// msr2pol (centroid.h:182-207): the measures whose centers are weighted means
inline bool msr_mean_centroid(int msr) {
    switch (msr) {
    case MCG_UWLLR: case MCG_LLR: case MCG_MKL: case MCG_JSD: case MCG_SQRL2: case MCG_REVERSE_MKL: case MCG_IS:
    case MCG_REVERSE_IS: case MCG_SIS: case MCG_RSIS: case MCG_SRULRT: case MCG_SRLRT: case MCG_JSM:
    case MCG_COSINE_DISTANCE: case MCG_BHATTACHARYYA_METRIC: case MCG_BHATTACHARYYA_DISTANCE: case MCG_HELLINGER:
    case MCG_POISSON:
        return true;
    }
    return false;
}

// Tiled sparse products (CSR rows): the decomposed costs' cross terms against the
// center table CT (d x kpad) and the soft M-step numerator over the column tiles.
bool tiled_costs_ok(const Data &ds, int msr, uint64_t k);
bool tiled_costs_ready(Data &ds, uint64_t k, hipStream_t st);
void tiled_costs(Data &ds, int msr, const float *CT, uint32_t kpad, uint64_t k, float pv, float logpv,
                 const float4 *rt, const double *cs, float *out_nk, uint32_t *asn, float *best, hipStream_t st);
bool tiled_mstep_ok(const Data &ds, uint64_t k);
bool tiled_mstep_ready(Data &ds, uint64_t k, const uint64_t *cptr, const uint32_t *crow, const float *cval,
                       hipStream_t st);
void tiled_soft_numer(Data &ds, const float *R, uint64_t k, const double *den, float *Cnew, hipStream_t st);

// Generic-measure kernels (dense or CSR rows vs dense centers), FP32.
void generic_costs(Data &ds, int msr, double prior, const float *C, uint64_t k, float *out_nk,
                   uint32_t *asn, float *best, hipStream_t st);

// Soft EM pieces.
void softmax_rows(const float *costs, uint64_t n, uint64_t k, float temp, float *resp,
                  const double *w, double *rowdot, hipStream_t st);
void soft_mstep(Data &ds, const float *resp, const double *w, uint64_t k, float *Cnew,
                hipStream_t st);
// Soft M-step for L1 (dense) and L2: weighted median / Weiszfeld geometric median of
// all rows per center, from the centers in C (updated in place); K (n x k) is scratch.
void soft_median_mstep(Data &ds, int msr, const float *R, const double *w, uint64_t k, float *C, float *K,
                       hipStream_t st);

// Sampling.
void kmeanspp_run(Data &ds, int msr, double prior, uint64_t k, uint64_t seed, const double *w,
                  uint32_t *centers_h, uint32_t *asn_d, float *dist_d, hipStream_t st);
// D^2 seeding with L > 1 local trials per center (kmeans.h:79-92, 114-177).
void kmeanspp_local_run(Data &ds, int msr, double prior, uint64_t k, uint64_t seed, const double *w, uint64_t L,
                        uint32_t *centers_h, uint32_t *asn_d, float *dist_d, hipStream_t st);
// Local search++ rounds after the seeding (centers_h, asn_d, dist_d in/out); the
// draws continue the seeding's stream at output j0.
void localsearch_run(Data &ds, int msr, double prior, uint64_t k, uint64_t seed, const double *w, uint64_t rounds,
                     uint64_t j0, uint32_t *centers_h, uint32_t *asn_d, float *dist_d, hipStream_t st);
void coreset_probs_run(Ctx &ctx, int sens, uint64_t n, uint64_t ncenters, const double *costs_d,
                       const uint32_t *asn_d, const double *w_d, float *probs_d, hipStream_t st);
// The same on a caller-provided device scratch of coreset_scratch_bytes(n, ncenters)
// bytes (256-byte aligned); launches only, no allocation.
uint64_t coreset_scratch_bytes(uint64_t n, uint64_t ncenters);
void coreset_probs_core(int sens, uint64_t n, uint64_t ncenters, const double *costs_d, const uint32_t *asn_d,
                        const double *w_d, float *probs_d, void *scratch, hipStream_t st);
// Set-up of the fixed scratch of the allocation-free entry points for the current
// device (called by mcg_ctx_create with the context's device current).
void noalloc_setup(int device);
void coreset_sample_run(Ctx &ctx, uint64_t n, const float *probs_d, const double *w_d, uint64_t m,
                        uint64_t seed, uint64_t *idx_d, float *wout_d, hipStream_t st);
void coreset_sample_from(Ctx &ctx, uint64_t n, const float *probs_d, const double *w_d, uint64_t m,
                         uint64_t seed, uint64_t first, uint64_t *idx_d, float *wout_d, hipStream_t st);
void restart_draws(Ctx &ctx, const float *costs_d, uint64_t n, uint64_t nempty, uint64_t *rows_h,
                   hipStream_t st);
// Draw j of the wyrand stream of n over device costs (the sampler of restart_draws).
uint64_t restart_draw_device(Ctx &ctx, const float *costs_d, uint64_t n, uint64_t j, hipStream_t st);
// Fixed-point inverse-CDF draws over host weights wv for raw draws r[0..m): the
// weights are converted once, each draw is a search of their prefix sums.
void draw_fixed_many(const std::vector<double> &wv, const uint64_t *r, uint64_t m, uint64_t *out);
// exact sum of w_i v_i (v_i >= 0) at the sampler's fixed-point scale (one host read)
double fixed_wsum(Ctx &c, const float *v, const double *w, uint64_t n, hipStream_t st);
double fixed_wsum(Ctx &c, const double *v, const double *w, uint64_t n, hipStream_t st);
// device view of a caller buffer: device memory borrowed, host memory staged in `stage`
const void *in_dev(Ctx &c, const void *p, size_t bytes, DBuf &stage);
// stream synchronize, then fold pending kernel timings into the totals
void sync(Ctx &c);
class FixedSeq {
  public:
    explicit FixedSeq(std::vector<double> w);
    uint64_t draw(uint64_t r) const;
    void zero(uint64_t i);
  private:
    void build();
    std::vector<double> w_;
    std::vector<uint64_t> fx_, tree_;
    double mx_ = 0.;
    int s_ = 0;
    uint64_t T_ = 0;
};
// Exact sum of w_i * v_i (v float on the device, w nullable): each term in fixed
// point with the sampler's shift, integer sums; independent of the order.
double fixed_wsum(Ctx &ctx, const float *v, const double *w, uint64_t n, hipStream_t st);
// Rows idx (device idx_d; idx_h the same on the host, or null) of S into D, reusing
// D's buffers; D's derived data (BF16 copy, row sums, transforms) is dropped.
void rows_into(const Data &S, const uint64_t *idx_h, const uint64_t *idx_d, uint64_t m, Data &D, hipStream_t st);

// Drop the data derived from a dataset's rows (rebuilt on next use).
inline void reset_derived(Data &D) {
    D.base_ready = false;
    D.bf16_ready = false;
    D.f16_ready = false;
    D.rtiles.ready = false;
    D.rtiles.cpl = 0;
    D.ctiles.ready = false;
    D.ctiles.cpl = 0;
    D.rowsum.release();
    D.tf_prior = -1.;
    D.tf_logx_pv = -1.;
    D.li_unique.clear();
    D.csc_ptr.release();
    D.csc_row.release();
    D.csc_val.release();
}

// Small helpers.
void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st);

void copy_rows_to_centers(Data &ds, const uint64_t *rows_h, const uint32_t *ids_h, uint64_t m,
                          float *C, hipStream_t st);
void rowsums(Data &ds, hipStream_t st);
// CSR per-prior transforms
void csr_transforms(Data &ds, double prior, hipStream_t st);

inline uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Output j (0-based) of the published wyrand generator seeded with seed, in counter
// form: state seed + (j+1)*P0, mixed by the 128-bit product with state ^ P1.
__host__ __device__ inline uint64_t wyrand_at(uint64_t seed, uint64_t j) {
    const uint64_t s = seed + (j + 1) * 0xa0761d6478bd642fULL;
    const uint64_t b = s ^ 0xe7037ed1a0b428dbULL;
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(s, b) ^ (s * b);
#else
    const unsigned __int128 p = (unsigned __int128)s * b;
    return (uint64_t)(p >> 64) ^ (uint64_t)p;
#endif
}

// Blocking copy ordered on st (library code never uses the legacy null stream, which
// would serialise it against every blocking stream of the process).
inline void copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t st) {
    if (!bytes) return;
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind, st));
    HIP_CHECK(hipStreamSynchronize(st));
}

// true while work launched on st is being captured into a graph
inline bool stream_capturing(hipStream_t st) {
    hipStreamCaptureStatus s = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &s) == hipSuccess && s != hipStreamCaptureStatusNone;
}

// ---- D^2 seeding helpers shared by the seeding kernels -------------------------
// Block max of |w_p * dist_p| into *mx (bit patterns of non-negative doubles order
// like their values): the next draw's fixed-point scale, computed in the pass that
// writes the distances.  Every thread of the block calls it.
// Lane l's value to the whole wave; l must be the same on every lane (scalar
// register reads, not LDS permutes).
__device__ __forceinline__ uint32_t bcast_u32(uint32_t v, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ float bcast_f32(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint64_t bcast_u64(uint64_t v, int l) {
    const uint32_t lo = bcast_u32((uint32_t)v, l), hi = bcast_u32((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double bcast_f64(double v, int l) {
    return __longlong_as_double((long long)bcast_u64((uint64_t)__double_as_longlong(v), l));
}

__device__ __forceinline__ void block_max_into(double a, unsigned long long *mx) {
    __shared__ double sm[256];
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    // (the maximum only grows within a round: a block whose maximum is not above the
    // current one skips the atomic, so the blocks' same-address atomics do not queue)
    if (threadIdx.x == 0) {
        const unsigned long long v = (unsigned long long)__double_as_longlong(sm[0]);
        if (v > __hip_atomic_load(mx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(mx, v);
    }
}

// new distance of point p given its cost s to the new center (kmeans.h:189-204)
__device__ __forceinline__ float pp_new_dist(uint64_t p, uint64_t newc, float s, uint32_t ci, int first,
                                             float *__restrict__ dist, uint32_t *__restrict__ asn) {
    if (first) {
        const float v = (p == newc) ? 0.f : s;
        dist[p] = v;
        asn[p] = 0;
        return v;
    }
    if (p == newc) { dist[p] = 0.f; asn[p] = ci; return 0.f; }
    const float dv = dist[p];
    if (dv <= 0.f) return dv;
    if (s < dv) { dist[p] = s; asn[p] = ci; return s; }
    return dv;
}

__device__ __forceinline__ double pp_weighted(float v, const double *__restrict__ w, uint64_t p) {
    double a = (double)v;
    if (w) a = w[p] * a;
    return fabs(a);
}

// One round of D^2 seeding for the measures without the dense-SQRL2 kernel: the
// cost of every row to data row *newc (densified into center_row), folded into
// dist/asn with the weighted maximum into mx.  For JSD (a metric after the square
// root) rounds after the first skip rows whose center is far from the new one,
// using the center-to-center costs of the rows cent[0..ci) (scratch cc, k floats).
// Returns false for priors that need the per-pair kernel (the caller then takes
// the general cost path).
bool generic_pp_round(Data &ds, int msr, double prior, const float *center_row, const uint32_t *cent, uint32_t ci,
                      const uint64_t *newc, float *cc, float *dist, uint32_t *asn, int first, const double *w,
                      unsigned long long *mx, hipStream_t st);

// The single-center tables of the seeding rounds after the first, filled on the
// device by the round's choice kernel (the table of the new center row; JSD's G for
// U <= PP_GMAX distinct li values, else by k_jsd_g in generic_pp_next).
constexpr uint32_t PP_GMAX = 8;
struct PPTables {
    float *CT;           // d floats
    double *cs;          // 2 doubles
    double *G;           // U doubles (JSD)
    const float *lis;    // U distinct li values (JSD)
    uint32_t U;
    double pv;
};
// false when the measure and prior take the general cost path (no tables)
bool generic_pp_tables(Data &ds, int msr, double prior, PPTables &t, hipStream_t st);
// a round after the first whose tables the choice kernel filled: JSD's G when U >
// PP_GMAX, the center-to-center costs (JSD), the update of dist / asn / mx
void generic_pp_next(Data &ds, int msr, const PPTables &t, const uint32_t *cent, uint32_t ci, const uint64_t *newc,
                     float *cc, float *dist, uint32_t *asn, const double *w, unsigned long long *mx,
                     hipStream_t st);

// Table of one center row cr for the decomposed measures (applicator.h:1632-1915 through
// the decompositions): CT[j * kpad + c] and cs[2c], cs[2c + 1].  Computed by threads
// 0..NT-1 of the block (larger blocks: the other threads only take part in the
// barriers); red: 2 x NT doubles of LDS.
template <int NT>
__device__ __forceinline__ void center_table_body(int msr, const float *__restrict__ cr, uint32_t c, uint32_t kpad,
                                                  uint32_t d, double pv, float *__restrict__ CT,
                                                  double *__restrict__ cs, double (*red)[NT]) {
    const uint32_t t = threadIdx.x;
    const bool on = t < NT;
    double s = 0.;
    if (on)
        for (uint32_t j = t; j < d; j += NT) s += (double)cr[j];
    if (on) red[0][t] = s;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if ((int)t < o) red[0][t] += red[0][t + o];
        __syncthreads();
    }
    const double csum = red[0][0];
    __syncthreads();
    const double M = csum + (double)d * pv, Mi = 1. / M;
    double a = 0., b = 0.;
    if (on)
        for (uint32_t j = t; j < d; j += NT) {
            const double cj = (double)cr[j];
            float v;
            // explicit fma: the same arithmetic in every translation unit, whatever its
            // contraction setting
            if (msr == MCG_SQRL2) {
                v = (float)cj;
                a = fma(cj, cj, a);
            } else {
                const double ct = (cj + pv) * Mi;
                if (msr == MCG_REVERSE_MKL) {
                    const double l = log(ct);
                    v = (float)l;
                    a += l;
                } else {
                    v = (float)ct;
                    if (msr == MCG_MKL) {   // JSD reads only the table
                        a = fma(ct, log(ct), a);
                        b += ct;
                    }
                }
            }
            CT[(uint64_t)j * kpad + c] = v;
        }
    if (on) {
        red[0][t] = a;
        red[1][t] = b;
    }
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if ((int)t < o) {
            red[0][t] += red[0][t + o];
            red[1][t] += red[1][t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        cs[2 * c] = red[0][0];
        cs[2 * c + 1] = red[1][0];
    }
    __syncthreads();
}

// JSD's G = sum_j h(l, CT[j * kpad + c]) in FP64 by threads 0..NT-1 of the block:
// lanes over the columns, then the NT / 64 wave sums in a fixed order (pairs, then
// pairs of pairs); ws: NT / 64 doubles of LDS.
template <int NT>
__device__ __forceinline__ double jsd_g_body(const float *__restrict__ CT, uint32_t kpad, uint32_t c, uint32_t d,
                                             double l, double *ws) {
    static_assert(NT == 256 || NT == 1024, "256 or 1024 threads");
    const uint32_t t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    double s = 0.;
    if (t < NT)
        for (uint32_t j = t; j < d; j += NT) {
            const double b = CT[(uint64_t)j * kpad + c];
            const double m = l + b;
            s = fma(0.5, fma(l, log(2. * l / m), b * log(2. * b / m)), s);
        }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0 && t < NT) ws[wave] = s;
    __syncthreads();
    double g = ((ws[0] + ws[1]) + ws[2]) + ws[3];
    if (NT == 1024)
        for (int q = 4; q < 16; q += 4) g += ((ws[q] + ws[q + 1]) + ws[q + 2]) + ws[q + 3];
    __syncthreads();
    return g;
}

}  // namespace mcg

struct mcg_ctx { mcg::Ctx c; };
struct mcg_data { mcg::Data d; };
