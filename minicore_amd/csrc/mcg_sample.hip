// D^2 seeding and coreset sampling.
//
// Replaces coresets::kmeanspp (include/minicore/optim/kmeans.h:49-215), the
// empty-cluster restart draw (clustering/centroid.h:687-752) and
// CoresetSampler::make_probs_* / sample (coreset/coreset.h:340-699).
//
// Sampling "i with probability w_i / sum(w)" is an inverse CDF over integers:
// w_i -> floor(w_i * 2^s) (s from max(w) and n so the total fits in 63 bits),
// u = mulhi(r, T) for a raw 64-bit draw r, and the answer is the first i whose
// inclusive prefix sum exceeds u.  Integer sums are exact, so the index is the
// same for any summation order and equals the host restatement.  Draws come
// from the wyrand stream in counter form: draw j = wyrand output j for seed s.
// Sums in the sensitivity formulas are exact fixed-point integer sums for the
// same reason.
#include "mcg_internal.h"
#include <cstring>
#include <algorithm>
#include <cfloat>
#include <cmath>

#pragma clang fp contract(off)

namespace mcg {


__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

__host__ __device__ inline int bitlen64(uint64_t v) {
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

// shift so that n * max * 2^s < 2^62
__host__ __device__ inline int fixed_shift(double maxw, uint64_t n) {
    if (!(maxw > 0.) || !(maxw <= DBL_MAX)) return 0;
    int e;
    frexp(maxw, &e);
    return 62 - e - bitlen64(n);
}

__host__ __device__ inline uint64_t to_fixed(double w, int s) {
    return w > 0. ? (uint64_t)floor(ldexp(w, s)) : 0ull;
}

__host__ __device__ inline int64_t to_fixed_signed(double v, int s) {
    const double t = ldexp(v, s);
    return (int64_t)(t < 0. ? -floor(-t) : floor(t));
}

__host__ __device__ inline double from_fixed(int64_t v, int s) { return ldexp((double)v, -s); }

// ---- reductions --------------------------------------------------------------
constexpr uint32_t FX_BLK = 1024;   // elements per block sum

// max of |a_i| (a_i = w_i * v_i when w != null, v as float or double)
template <class T>
__global__ void k_absmax(const T *__restrict__ v, const double *__restrict__ w, uint64_t n,
                         unsigned long long *__restrict__ out_bits) {
    __shared__ double sm[256];
    double m = 0.;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        double a = (double)v[i];
        if (w) a = w[i] * a;
        a = fabs(a);
        if (a > m) m = a;
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    // non-negative doubles order like their bit patterns
    if (threadIdx.x == 0) atomicMax(out_bits, (unsigned long long)__double_as_longlong(sm[0]));
}

// block sums of floor(ldexp(w_i v_i, s)); s derived from *maxbits
template <class T>
__global__ void k_fixed_blocksum(const T *__restrict__ v, const double *__restrict__ w, uint64_t n,
                                 const unsigned long long *__restrict__ maxbits, uint64_t *__restrict__ bsum) {
    __shared__ unsigned long long sm[256];
    const int s = fixed_shift(__longlong_as_double((long long)*maxbits), n);
    const uint64_t b0 = (uint64_t)blockIdx.x * FX_BLK;
    unsigned long long acc = 0;
    for (uint32_t q = threadIdx.x; q < FX_BLK; q += blockDim.x) {
        const uint64_t i = b0 + q;
        if (i < n) {
            double a = (double)v[i];
            if (w) a = w[i] * a;
            acc += to_fixed(a, s);
        }
    }
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sm[0];
}

// Inclusive scan of one u64 per thread over a 1024-thread block: a shuffle scan in
// each wave, then the 16 wave totals scanned by the first wave; integer sums, so the
// association does not matter.  sm: 1024 u64 of LDS; sm[1023] holds the block total
// on return.
__device__ __forceinline__ unsigned long long block_scan_u64(unsigned long long x, unsigned long long *sm) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sm[wave] = x;
    __syncthreads();
    if (wave == 0) {
        unsigned long long v = lane < 16 ? sm[lane] : 0ull;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const unsigned long long y = __shfl_up(v, o);
            if (lane >= o) v += y;
        }
        if (lane < 16) sm[16 + lane] = v;
    }
    __syncthreads();
    if (wave > 0) x += sm[16 + wave - 1];
    if (t == 1023) sm[1023] = x;
    __syncthreads();
    return x;
}

// Single block of 1024 threads: choose the index for raw draw r; writes *out.  The
// first index whose fixed-point prefix exceeds u = mulhi(r, total): the block is
// found from a scan of the block sums, then the element from a scan inside the
// block (FX_BLK == 1024, one element per thread).  If the total is 0 the result is 0.
template <class T>
__global__ __launch_bounds__(1024) void k_fixed_select(const T *__restrict__ v, const double *__restrict__ w,
                                                       uint64_t n, const unsigned long long *__restrict__ maxbits,
                                                       const uint64_t *__restrict__ bsum, uint32_t nblk, uint64_t r,
                                                       uint64_t *__restrict__ out) {
    static_assert(FX_BLK == 1024, "one element per thread in the block scan");
    __shared__ unsigned long long sm[1024];
    __shared__ unsigned long long sel_u;
    __shared__ uint32_t sel_t, sel_b;
    const int s = fixed_shift(__longlong_as_double((long long)*maxbits), n);
    const uint32_t t = threadIdx.x;
    // thread t owns block sums [t*per, (t+1)*per)
    const uint32_t per = (nblk + 1023) / 1024;
    const uint32_t b0 = t * per, b1 = min(nblk, b0 + per);
    unsigned long long acc = 0;
    for (uint32_t b = b0; b < b1; ++b) acc += bsum[b];
    const unsigned long long incl = block_scan_u64(acc, sm);
    const unsigned long long tot = sm[1023];
    if (tot == 0) {
        if (t == 0) *out = 0;
        return;
    }
    const uint64_t u = mulhi64(r, tot);
    if (t == 0) {
        sel_t = 0xFFFFFFFFu;
        *out = min<uint64_t>(n - 1, (uint64_t)(nblk - 1) * FX_BLK + FX_BLK - 1);   // replaced below
    }
    __syncthreads();
    // the unique thread whose range holds u: incl - acc <= u < incl
    if (incl > u && incl - acc <= u) {
        unsigned long long run = incl - acc;
        uint32_t blk = b1 - 1;
        for (uint32_t b = b0; b < b1; ++b) {
            if (run + bsum[b] > u) { blk = b; break; }
            run += bsum[b];
        }
        sel_b = blk;
        sel_u = u - run;
        sel_t = t;
    }
    __syncthreads();
    // inside the selected block: one element per thread
    const uint64_t i = (uint64_t)sel_b * FX_BLK + t;
    unsigned long long f = 0;
    if (i < n) {
        double a = (double)v[i];
        if (w) a = w[i] * a;
        f = to_fixed(a, s);
    }
    const unsigned long long pre = block_scan_u64(f, sm);
    if (pre > sel_u && pre - f <= sel_u && f > 0) *out = i;
    // u < total, so some element of the selected block holds it
}

// Host-side draws from device values (the empty-cluster and check-in restarts).
void draw_fixed_many(const std::vector<double> &wv, const uint64_t *r, uint64_t m, uint64_t *out) {
    // the fixed-point weights and their prefix once, then one search per draw: the
    // first index whose inclusive prefix exceeds u
    const uint64_t n = wv.size();
    double mx = 0.;
    for (double x : wv) if (x > mx) mx = x;
    const int s = fixed_shift(mx, n);
    std::vector<uint64_t> pre(n);
    uint64_t T = 0;
    for (uint64_t i = 0; i < n; ++i) { T += to_fixed(wv[i], s); pre[i] = T; }
    for (uint64_t q = 0; q < m; ++q) {
        if (T == 0) { out[q] = 0; continue; }
        const uint64_t u = mulhi64(r[q], T);
        out[q] = (uint64_t)(std::upper_bound(pre.begin(), pre.end(), u) - pre.begin());
        if (out[q] >= n) out[q] = n - 1;
    }
}

// Draws without replacement: draw(r) is draw_fixed_many's draw on the current
// weights, and zero(i) removes row i before the next draw.  A Fenwick tree over the
// fixed-point weights keeps both O(log n); when the removed weight was the maximum,
// the shift and the tree are rebuilt (draw_fixed_many recomputes them per call).
FixedSeq::FixedSeq(std::vector<double> w) : w_(std::move(w)) { build(); }

void FixedSeq::build() {
    const uint64_t n = w_.size();
    mx_ = 0.;
    for (double x : w_) if (x > mx_) mx_ = x;
    s_ = fixed_shift(mx_, n);
    fx_.assign(n, 0);
    tree_.assign(n + 1, 0);
    T_ = 0;
    for (uint64_t i = 0; i < n; ++i) {
        fx_[i] = to_fixed(w_[i], s_);
        T_ += fx_[i];
        tree_[i + 1] += fx_[i];
        const uint64_t p = (i + 1) + ((i + 1) & (~(i + 1) + 1));
        if (p <= n) tree_[p] += tree_[i + 1];
    }
}

uint64_t FixedSeq::draw(uint64_t r) const {
    const uint64_t n = w_.size();
    if (T_ == 0) return 0;
    uint64_t rem = mulhi64(r, T_), pos = 0, step = 1;
    while (step * 2 <= n) step *= 2;
    // the largest pos whose prefix is <= u: the first row whose inclusive prefix exceeds u
    for (; step; step >>= 1)
        if (pos + step <= n && tree_[pos + step] <= rem) { pos += step; rem -= tree_[pos]; }
    return pos < n ? pos : n - 1;
}

void FixedSeq::zero(uint64_t i) {
    const double old = w_[i];
    w_[i] = 0.;
    if (old >= mx_) { build(); return; }
    const uint64_t f = fx_[i];
    fx_[i] = 0;
    T_ -= f;
    for (uint64_t p = i + 1; p < tree_.size(); p += p & (~p + 1)) tree_[p] -= f;
}

template <class V>
static double fixed_wsum_t(Ctx &c, const V *v, const double *w, uint64_t n, hipStream_t st) {
    const uint32_t nblk = (uint32_t)cdiv(n, FX_BLK);
    char *buf = (char *)c.s[S_MISC0].reserve(64 + nblk * 8);
    unsigned long long *mx = (unsigned long long *)buf;
    uint64_t *bsum = (uint64_t *)(buf + 64);
    fill_u32((uint32_t *)mx, 0, 2, st);
    hipLaunchKernelGGL(k_absmax<V>, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 256), 1024)), dim3(256), 0, st, v, w,
                       n, mx);
    hipLaunchKernelGGL(k_fixed_blocksum<V>, dim3(nblk), dim3(256), 0, st, v, w, n,
                       (const unsigned long long *)mx, bsum);
    std::vector<uint64_t> h(nblk + 1);
    HIP_CHECK(hipMemcpyAsync(h.data(), mx, 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(h.data() + 1, bsum, nblk * 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    double mxv;
    std::memcpy(&mxv, h.data(), 8);
    uint64_t T = 0;
    for (uint32_t b = 0; b < nblk; ++b) T += h[b + 1];
    return ldexp((double)T, -fixed_shift(mxv, n));
}

double fixed_wsum(Ctx &c, const float *v, const double *w, uint64_t n, hipStream_t st) {
    return fixed_wsum_t(c, v, w, n, st);
}

double fixed_wsum(Ctx &c, const double *v, const double *w, uint64_t n, hipStream_t st) {
    return fixed_wsum_t(c, v, w, n, st);
}

void restart_draws(Ctx &c, const float *costs_d, uint64_t n, uint64_t nempty, uint64_t *rows_h, hipStream_t st) {
    std::vector<float> cf(n);
    HIP_CHECK(hipMemcpyAsync(cf.data(), costs_d, n * 4, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    std::vector<double> wv(cf.begin(), cf.end());
    std::vector<uint64_t> r(nempty);
    for (uint64_t e = 0; e < nempty; ++e) r[e] = wyrand_at(n, e);
    draw_fixed_many(wv, r.data(), nempty, rows_h);
    (void)c;
}

// ---- D^2 seeding ---------------------------------------------------------------
// dist update for dense SQRL2: sequential FMA over dims of (x_i - x_c)^2.
__global__ __launch_bounds__(256) void k_pp_update_l2(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                      const uint64_t *__restrict__ newc_p, uint32_t ci,
                                                      float *__restrict__ dist, uint32_t *__restrict__ asn,
                                                      int first, const double *__restrict__ w,
                                                      unsigned long long *__restrict__ mx) {
    __shared__ float xs[256][33];
    __shared__ float cs[32];
    const int tid = threadIdx.x;
    const uint64_t newc = *newc_p;
    const uint64_t p0 = (uint64_t)blockIdx.x * 256, p = p0 + tid;
    const float *cr = X + newc * d;
    float s = 0.f;
    for (uint32_t j0 = 0; j0 < d; j0 += 32) {
        __syncthreads();
        for (int q = tid; q < 256 * 32; q += 256) {
            const int r = q >> 5, jj = q & 31;
            const uint64_t pr = p0 + r;
            xs[r][jj] = (pr < n && j0 + jj < d) ? X[pr * d + j0 + jj] : 0.f;
        }
        if (tid < 32) cs[tid] = j0 + tid < d ? cr[j0 + tid] : 0.f;
        __syncthreads();
        const int jn = (int)min<uint32_t>(32, d - j0);
        for (int jj = 0; jj < jn; ++jj) { const float t = xs[tid][jj] - cs[jj]; s = fmaf(t, t, s); }
    }
    double a = 0.;
    if (p < n) a = pp_weighted(pp_new_dist(p, newc, s, ci, first, dist, asn), w, p);
    block_max_into(a, mx);
}

// dist update from precomputed costs to the new center (generic measures)
__global__ __launch_bounds__(256) void k_pp_update_from(const float *__restrict__ cost, uint64_t n,
                                                        const uint64_t *__restrict__ newc_p, uint32_t ci,
                                                        float *__restrict__ dist, uint32_t *__restrict__ asn,
                                                        int first, const double *__restrict__ w,
                                                        unsigned long long *__restrict__ mx) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.;
    if (p < n) a = pp_weighted(pp_new_dist(p, *newc_p, cost[p], ci, first, dist, asn), w, p);
    block_max_into(a, mx);
}

// if dist[newc] == 0, take the first index of the maximum distance (kmeans.h:179-186)
__global__ void k_pp_fallback(const float *__restrict__ dist, uint64_t n, uint64_t *__restrict__ newc_p,
                              uint32_t *__restrict__ centers, uint32_t ci) {
    __shared__ float bv[1024];
    __shared__ uint64_t bi[1024];
    const uint64_t nc = *newc_p;
    if (dist[nc] != 0.f) {
        if (threadIdx.x == 0) centers[ci] = (uint32_t)nc;
        return;
    }
    float m = -1.f;
    uint64_t mi = 0;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x)
        if (dist[i] > m) { m = dist[i]; mi = i; }
    bv[threadIdx.x] = m;
    bi[threadIdx.x] = mi;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const float ov = bv[threadIdx.x + o];
            const uint64_t oi = bi[threadIdx.x + o];
            if (ov > bv[threadIdx.x] || (ov == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
                bv[threadIdx.x] = ov;
                bi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { *newc_p = bi[0]; centers[ci] = (uint32_t)bi[0]; }
}

// FP64 squared distance between dense rows ra and rb by a group of 16 lanes (l16 =
// the lane's index in it; on: the group has a pair): lanes over the dims (float4
// when vec), then a shuffle tree inside the group.  Every lane of the group returns
// the value; all lanes of the wave take part in the shuffles.
__device__ __forceinline__ double row_dist_l2_16(const float *__restrict__ X, uint32_t d, int vec, uint64_t ra,
                                                 uint64_t rb, int l16, bool on) {
    double s = 0.;
    if (on) {
        const float *x = X + ra * d, *y = X + rb * d;
        if (vec) {
            const float4 *x4 = (const float4 *)x, *y4 = (const float4 *)y;
            for (uint32_t j = l16; j < (d >> 2); j += 16) {
                const float4 p = x4[j], q = y4[j];
                double t;
                t = (double)p.x - (double)q.x; s += t * t;
                t = (double)p.y - (double)q.y; s += t * t;
                t = (double)p.z - (double)q.z; s += t * t;
                t = (double)p.w - (double)q.w; s += t * t;
            }
        } else {
            for (uint32_t j = l16; j < d; j += 16) {
                const double t = (double)x[j] - (double)y[j];
                s += t * t;
            }
        }
    }
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o);
    return s;
}

// Squared distances (FP64) between the new center's row and each earlier center's
// row, for rounds with many earlier centers: 16 lanes per earlier center q < ci.
__global__ __launch_bounds__(256) void k_pp_cc(const float *__restrict__ X, uint32_t d, int vec,
                                               const uint32_t *__restrict__ centers, uint32_t ci,
                                               const uint64_t *__restrict__ newc_p, double *__restrict__ cc) {
    const uint32_t q = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l16 = threadIdx.x & 15;
    const bool on = q < ci;
    const double v = row_dist_l2_16(X, d, vec, on ? centers[q] : 0u, *newc_p, l16, on);
    if (on && l16 == 0) cc[q] = v;
}

// Arguments of the fused choice kernel of a seeding round.
struct PickArgs {
    const float *dist;
    const double *w;
    uint64_t n;
    unsigned long long *mx;   // weighted maximum of the last update; cleared for the next
    uint64_t *bsum;           // per-block fixed-point sums
    uint32_t nblk;
    uint32_t *ctr;            // blocks done (the last block resets it)
    uint64_t *newc;
    uint32_t *centers;
    const float *X;           // dense rows (MODE 0; MODE 1 with val == nullptr)
    uint32_t d;
    double *cc;               // MODE 0: FP64 center-to-center squared distances (here when ci <= ccmax)
    uint32_t ccmax;
    int vec;                  // rows as float4
    const float *val;         // MODE 1: CSR rows
    const uint32_t *idx;
    const uint64_t *ptr;
    float *row;               // MODE 1: the new center as a dense row
    int msr;                  // MODE 1 with tables: the measure, prior and tables
    int tables;
    double pv;
    float *CT;
    double *cs;
    const float *lis;
    uint32_t gU;              // JSD: G for these many li values here (0: none)
    double *G;
};

// One seeding round's choice (kmeans.h:115-186) fused with the block sums it needs.
// The blocks of 1024 threads add the fixed-point weights w * dist of chunks of FX_BLK
// rows, chunk b into bsum[b] (a block takes chunks blockIdx.x, + gridDim.x, ...: at
// most a few hundred blocks, since every block's device-scope release and counter
// update are serialised across the GPU); the last block to finish (a device-scope
// counter) reads the chunk sums back through the device-coherent path and draws: the fixed-point draw r over w *
// dist, the first index of the maximum distance instead if the drawn point has
// distance 0.  It writes *newc and centers[ci], clears *mx and the counter, and
// prepares the round's update: MODE 0 (dense SQRL2 rows) the FP64 squared distances
// from the new center's row to the earlier centers' rows; MODE 1 (other measures)
// the new center as a dense row and, with tables, its single-center table and JSD's
// G for the first gU distinct li values.  The draw is that of the block sums and
// select of separate kernels (integer sums: the same index).
constexpr uint32_t PICK_GRID = 128;   // blocks of the choice kernel (at most)

template <int MODE>
__global__ __launch_bounds__(1024) void k_pp_pick(PickArgs a, uint64_t r, uint32_t ci) {
    __shared__ unsigned long long sm[1024];
    __shared__ unsigned long long sel_u;
    __shared__ uint32_t sel_b, last;
    __shared__ uint64_t chosen;
    __shared__ float bv[1024];
    __shared__ uint64_t bi[1024];
    __shared__ double red[2][1024];
    __shared__ double ws[16];
    const uint32_t t = threadIdx.x;
    const int s = fixed_shift(__longlong_as_double((long long)*a.mx), a.n);
    {
        for (uint32_t cb = blockIdx.x; cb < a.nblk; cb += gridDim.x) {
            const uint64_t i = (uint64_t)cb * FX_BLK + t;
            unsigned long long f = 0;
            if (i < a.n) {
                double v = (double)a.dist[i];
                if (a.w) v = a.w[i] * v;
                f = to_fixed(v, s);
            }
            for (int o = 32; o > 0; o >>= 1) f += __shfl_xor(f, o);
            if ((t & 63) == 0) sm[t >> 6] = f;
            __syncthreads();
            if (t == 0) {
                unsigned long long tot = 0;
                for (int q = 0; q < 16; ++q) tot += sm[q];
                a.bsum[cb] = tot;
            }
            __syncthreads();
        }
        if (t == 0) {
            __threadfence();
            last = atomicAdd(a.ctr, 1u) == gridDim.x - 1 ? 1u : 0u;
        }
        __syncthreads();
        if (!last) return;
        __threadfence();
    }
    // the last block: every block's sum is written
    const uint32_t per = (a.nblk + 1023) / 1024;
    const uint32_t b0 = t * per, b1 = min(a.nblk, b0 + per);
    unsigned long long acc = 0;
    for (uint32_t b = b0; b < b1; ++b) acc += __hip_atomic_load(&a.bsum[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long incl = block_scan_u64(acc, sm);
    const unsigned long long tot = sm[1023];
    if (t == 0) chosen = 0;
    __syncthreads();
    if (tot != 0) {
        const uint64_t u = mulhi64(r, tot);
        if (incl > u && incl - acc <= u) {
            unsigned long long run = incl - acc;
            uint32_t blk = b1 - 1;
            for (uint32_t b = b0; b < b1; ++b) {
                const unsigned long long bs = __hip_atomic_load(&a.bsum[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (run + bs > u) { blk = b; break; }
                run += bs;
            }
            sel_b = blk;
            sel_u = u - run;
        }
        __syncthreads();
        const uint64_t i = (uint64_t)sel_b * FX_BLK + t;
        unsigned long long f = 0;
        if (i < a.n) {
            double v = (double)a.dist[i];
            if (a.w) v = a.w[i] * v;
            f = to_fixed(v, s);
        }
        const unsigned long long pre = block_scan_u64(f, sm);
        if (pre > sel_u && pre - f <= sel_u) chosen = i;
        __syncthreads();
    }
    // distance 0 (or all distances 0): first index of the maximum distance
    if (a.dist[chosen] == 0.f) {
        float m = -1.f;
        uint64_t mi = 0;
        for (uint64_t i = t; i < a.n; i += 1024)
            if (a.dist[i] > m) { m = a.dist[i]; mi = i; }
        bv[t] = m;
        bi[t] = mi;
        __syncthreads();
        for (int o = 512; o > 0; o >>= 1) {
            if ((int)t < o) {
                const float ov = bv[t + o];
                const uint64_t oi = bi[t + o];
                if (ov > bv[t] || (ov == bv[t] && oi < bi[t])) { bv[t] = ov; bi[t] = oi; }
            }
            __syncthreads();
        }
        if (t == 0) chosen = bi[0];
        __syncthreads();
    }
    const uint64_t c = chosen;
    if (MODE == 0) {
        // 16 lanes per earlier center q < ci (with more, k_pp_cc follows)
        if (ci <= a.ccmax)
            for (uint32_t q0 = 0; q0 < ci; q0 += 64) {
                const uint32_t q = q0 + (t >> 4);
                const bool on = q < ci;
                const double v = row_dist_l2_16(a.X, a.d, a.vec, on ? a.centers[q] : 0u, c, (int)(t & 15), on);
                if (on && (t & 15) == 0) a.cc[q] = v;
            }
    } else {
        if (!a.val) {
            for (uint32_t j = t; j < a.d; j += 1024) a.row[j] = a.X[c * a.d + j];
        } else {
            for (uint32_t j = t; j < a.d; j += 1024) a.row[j] = 0.f;
            __syncthreads();
            for (uint64_t p = a.ptr[c] + t; p < a.ptr[c + 1]; p += 1024) a.row[a.idx[p]] = a.val[p];
        }
        __syncthreads();
        if (a.tables) {
            center_table_body<1024>(a.msr, a.row, 0u, 1u, a.d, a.pv, a.CT, a.cs, red);
            for (uint32_t u = 0; u < a.gU; ++u) {
                const double g = jsd_g_body<1024>(a.CT, 1u, 0u, a.d, (double)a.lis[u], ws);
                if (t == 0) a.G[u] = g;
            }
        }
    }
    if (t == 0) {
        *a.newc = c;
        a.centers[ci] = (uint32_t)c;
        *a.mx = 0ull;
        *a.ctr = 0u;
    }
}

// Dense SQRL2 update after the first round, one thread per point.  A point whose
// current center a is far from the new one, ||c_a - c_new||^2 > 4 dist (1 + 2^-8),
// cannot get closer (triangle inequality; the margin covers the FP32 rounding of
// dist and of the new cost), so its row is not read and nothing changes.  Other
// points get the exact sequential FMA chain (16-byte loads when vec).
__global__ __launch_bounds__(256) void k_pp_update_l2_pruned(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                             int vec, const uint64_t *__restrict__ newc_p,
                                                             uint32_t ci, const double *__restrict__ cc,
                                                             float *__restrict__ dist, uint32_t *__restrict__ asn,
                                                             const double *__restrict__ w,
                                                             unsigned long long *__restrict__ mx) {
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    double a = 0.;
    if (p < n) {
        const uint64_t newc = *newc_p;
        float v = dist[p];
        if (p == newc) {
            v = 0.f;
            dist[p] = 0.f;
            asn[p] = ci;
        } else if (v > 0.f && !(cc[asn[p]] > 4.0 * (double)v * (1.0 + 1.0 / 256.0))) {
            const float *x = X + p * d, *c = X + newc * d;
            float s = 0.f;
            if (vec) {
                const float4 *x4 = (const float4 *)x, *c4 = (const float4 *)c;
                for (uint32_t j = 0; j < (d >> 2); ++j) {
                    const float4 xa = x4[j], cb = c4[j];
                    float t;
                    t = xa.x - cb.x; s = fmaf(t, t, s);
                    t = xa.y - cb.y; s = fmaf(t, t, s);
                    t = xa.z - cb.z; s = fmaf(t, t, s);
                    t = xa.w - cb.w; s = fmaf(t, t, s);
                }
            } else {
                for (uint32_t j = 0; j < d; ++j) { const float t = x[j] - c[j]; s = fmaf(t, t, s); }
            }
            if (s < v) { v = s; dist[p] = s; asn[p] = ci; }
        }
        a = pp_weighted(v, w, p);
    }
    block_max_into(a, mx);
}

__global__ void k_set_u64(uint64_t *p, uint64_t v) { *p = v; }

// one dense center from data row *newc_p (CSR or dense)
__global__ void k_row_to_dense(const float *__restrict__ X, const float *__restrict__ val,
                               const uint32_t *__restrict__ idx, const uint64_t *__restrict__ ptr, uint32_t d,
                               const uint64_t *__restrict__ newc_p, float *__restrict__ out) {
    const uint64_t r = *newc_p;
    if (!val) {
        for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) out[j] = X[r * d + j];
        return;
    }
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) out[j] = 0.f;
    __syncthreads();
    for (uint64_t p = ptr[r] + threadIdx.x; p < ptr[r + 1]; p += blockDim.x) out[idx[p]] = val[p];
}

void kmeanspp_run(Data &ds, int msr, double prior, uint64_t k, uint64_t seed, const double *w, uint32_t *centers_h,
                  uint32_t *asn_d, float *dist_d, hipStream_t st) {
    Ctx &c = *ds.ctx;
    const uint64_t n = ds.n;
    const uint32_t d = (uint32_t)ds.d;
    const bool l2 = !ds.csr && msr == MCG_SQRL2;
    const uint32_t nblk = (uint32_t)cdiv(n, FX_BLK);
    char *misc = (char *)c.s[S_MISC0].reserve(64 + nblk * 8 + k * 4);
    uint64_t *newc = (uint64_t *)misc;
    unsigned long long *mx = (unsigned long long *)(misc + 8);
    uint64_t *bsum = (uint64_t *)(misc + 64);
    uint32_t *cent = (uint32_t *)(misc + 64 + nblk * 8);
    float *row = nullptr, *cost1 = nullptr;
    uint32_t *asn1 = nullptr;
    if (!l2) {
        row = (float *)c.s[S_STAGE0].reserve(d * sizeof(float));
        cost1 = (float *)c.s[S_STAGE1].reserve(n * sizeof(float));
        asn1 = (uint32_t *)c.s[S_FLAGS].reserve(n * sizeof(uint32_t));
    }
    // update() writes the distances to the new center and their weighted maximum
    // (into mx, which must be clear): the fixed-point scale of the next draw
    // center-to-center costs of the pruned rounds (dense SQRL2: FP64; JSD: FP32)
    void *ccbuf = c.s[S_STAGE2].reserve(k * sizeof(double));
    auto update = [&](uint32_t ci, int first) {
        if (l2) {
            hipLaunchKernelGGL(k_pp_update_l2, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, ds.X, n, d, newc, ci,
                               dist_d, asn_d, first, w, mx);
        } else {
            hipLaunchKernelGGL(k_row_to_dense, dim3(1), dim3(256), 0, st, ds.X,
                               ds.csr ? ds.val.as<const float>() : nullptr,
                               ds.csr ? ds.idx.as<const uint32_t>() : nullptr,
                               ds.csr ? ds.ptr.as<const uint64_t>() : nullptr, d, newc, row);
            if (generic_pp_round(ds, msr, prior, row, cent, ci, newc, (float *)ccbuf, dist_d, asn_d, first, w, mx,
                                 st))
                return;
            generic_costs(ds, msr, prior, row, 1, nullptr, asn1, cost1, st);
            hipLaunchKernelGGL(k_pp_update_from, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, cost1, n, newc, ci,
                               dist_d, asn_d, first, w, mx);
        }
    };
    uint32_t *ctr = (uint32_t *)(misc + 16);
    uint64_t j = 0;
    const uint64_t fc = wyrand_at(seed, j++) % n;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, newc, fc);
    fill_u32((uint32_t *)mx, 0, 3, st);   // mx and the choice kernel's counter; later rounds: it clears both
    update(0, 1);
    centers_h[0] = (uint32_t)fc;
    fill_u32(cent, (uint32_t)fc, 1, st);
    // rounds after the first: the choice kernel (block sums, draw, the round's
    // preparation), then the update; dense SQRL2 prunes by the center-center distances
    PickArgs pa{};
    pa.dist = dist_d;
    pa.w = w;
    pa.n = n;
    pa.mx = mx;
    pa.bsum = bsum;
    pa.nblk = nblk;
    pa.ctr = ctr;
    pa.newc = newc;
    pa.centers = cent;
    pa.X = ds.csr ? nullptr : ds.X;
    pa.d = d;
    PPTables tb{};
    bool tables = false;
    const int vec = (d % 4 == 0) && ((uintptr_t)ds.X % 16 == 0);
    // the choice kernel's one block computes the center-center distances of the
    // first rounds; later rounds spread them over a kernel of their own
    constexpr uint32_t CC_IN_PICK = 128;
    if (l2) {
        pa.cc = (double *)ccbuf;
        pa.ccmax = CC_IN_PICK;
        pa.vec = vec;
    } else {
        pa.val = ds.csr ? ds.val.as<const float>() : nullptr;
        pa.idx = ds.csr ? ds.idx.as<const uint32_t>() : nullptr;
        pa.ptr = ds.csr ? ds.ptr.as<const uint64_t>() : nullptr;
        pa.row = row;
        tables = generic_pp_tables(ds, msr, prior, tb, st);
        pa.tables = tables ? 1 : 0;
        pa.msr = msr;
        pa.pv = tb.pv;
        pa.CT = tb.CT;
        pa.cs = tb.cs;
        pa.lis = tb.lis;
        pa.gU = tb.U <= PP_GMAX ? tb.U : 0u;
        pa.G = tb.G;
    }
    const uint32_t pgrid = std::min<uint32_t>(nblk, PICK_GRID);
    for (uint64_t ci = 1; ci < k; ++ci) {
        const uint64_t r = wyrand_at(seed, j++);
        if (l2) {
            hipLaunchKernelGGL(k_pp_pick<0>, dim3(pgrid), dim3(1024), 0, st, pa, r, (uint32_t)ci);
            if (ci > CC_IN_PICK)
                hipLaunchKernelGGL(k_pp_cc, dim3((uint32_t)cdiv(ci, 16)), dim3(256), 0, st, ds.X, d, vec,
                                   (const uint32_t *)cent, (uint32_t)ci, (const uint64_t *)newc, (double *)ccbuf);
            hipLaunchKernelGGL(k_pp_update_l2_pruned, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, ds.X, n, d,
                               vec, newc, (uint32_t)ci, (const double *)ccbuf, dist_d, asn_d, w, mx);
        } else {
            hipLaunchKernelGGL(k_pp_pick<1>, dim3(pgrid), dim3(1024), 0, st, pa, r, (uint32_t)ci);
            if (tables) {
                generic_pp_next(ds, msr, tb, cent, (uint32_t)ci, newc, (float *)ccbuf, dist_d, asn_d, w, mx, st);
            } else {
                generic_costs(ds, msr, prior, row, 1, nullptr, asn1, cost1, st);
                hipLaunchKernelGGL(k_pp_update_from, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, cost1, n, newc,
                                   (uint32_t)ci, dist_d, asn_d, 0, w, mx);
            }
        }
    }
    HIP_CHECK(hipMemcpyAsync(centers_h, cent, k * 4, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
}

// ---- D^2 seeding with local trials (kmeans.h:79-92, 114-177) ----------------------
// Costs of every row to row *row_p for dense SQRL2: the sequential FMA chain of
// k_pp_update_l2 (rows staged through LDS in 32-column slices), written to out.
__global__ __launch_bounds__(256) void k_cost_row_l2(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                     const uint64_t *__restrict__ row_p, float *__restrict__ out) {
    __shared__ float xs[256][33];
    __shared__ float cs[32];
    const int tid = threadIdx.x;
    const uint64_t p0 = (uint64_t)blockIdx.x * 256, p = p0 + tid;
    const float *cr = X + *row_p * d;
    float s = 0.f;
    for (uint32_t j0 = 0; j0 < d; j0 += 32) {
        __syncthreads();
        for (int q = tid; q < 256 * 32; q += 256) {
            const int r = q >> 5, jj = q & 31;
            const uint64_t pr = p0 + r;
            xs[r][jj] = (pr < n && j0 + jj < d) ? X[pr * d + j0 + jj] : 0.f;
        }
        if (tid < 32) cs[tid] = j0 + tid < d ? cr[j0 + tid] : 0.f;
        __syncthreads();
        const int jn = (int)min<uint32_t>(32, d - j0);
        for (int jj = 0; jj < jn; ++jj) { const float t = xs[tid][jj] - cs[jj]; s = fmaf(t, t, s); }
    }
    if (p < n) out[p] = s;
}

// First-center candidate: out_i = cost_i, 0 for the candidate's own row (setdists).
__global__ void k_local_first(const float *__restrict__ cost, uint64_t n, const uint64_t *__restrict__ row_p,
                              float *__restrict__ out) {
    const uint64_t r = *row_p;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = i == r ? 0.f : cost[i];
}

// One candidate of a round: out_i = cost_i, asn_out_i = ci where cost_i < base_i,
// else base (every row, the candidate's own included).  base and out may alias.
__global__ void k_local_merge(const float *cost, const float *base, const uint32_t *base_a, uint64_t n, uint32_t ci,
                              float *out, uint32_t *out_a) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float v = cost[i], b = base[i];
        const uint32_t a = base_a[i];
        if (v < b) { out[i] = v; out_a[i] = ci; }
        else { out[i] = b; out_a[i] = a; }
    }
}

// The new center: its assignment is its round, its distance 0.
__global__ void k_local_settle(uint64_t newc, uint32_t ci, float *__restrict__ dist, uint32_t *__restrict__ asn) {
    dist[newc] = 0.f;
    asn[newc] = ci;
}

void kmeanspp_local_run(Data &ds, int msr, double prior, uint64_t k, uint64_t seed, const double *w, uint64_t L,
                        uint32_t *centers_h, uint32_t *asn_d, float *dist_d, hipStream_t st) {
    const uint64_t n = ds.n;
    const uint32_t d = (uint32_t)ds.d;
    const bool l2 = !ds.csr && msr == MCG_SQRL2;
    const uint32_t nblk = (uint32_t)cdiv(n, FX_BLK);
    // candidate t > 0 of a round keeps its distances and assignment in slot t - 1
    DBuf bCand, bOd, bOa, bTd, bTa, bRow, bScr, bMisc;
    float *cand = (float *)bCand.reserve(n * 4), *od = (float *)bOd.reserve(n * 4);
    uint32_t *oa = (uint32_t *)bOa.reserve(n * 4), *scr = (uint32_t *)bScr.reserve(n * 4);
    float *td = (float *)bTd.reserve((L - 1) * n * 4);
    uint32_t *ta = (uint32_t *)bTa.reserve((L - 1) * n * 4);
    float *row = (float *)bRow.reserve(d * 4);
    // smp[L] | mx of the draw | mx[L] of the candidates | bsum of the draw | bsum[L][nblk]
    uint64_t *smp = (uint64_t *)bMisc.reserve((2 * L + 2 + (L + 1) * (uint64_t)nblk) * 8);
    unsigned long long *mxs = (unsigned long long *)(smp + L), *mxc = mxs + 2;
    uint64_t *bss = (uint64_t *)(mxc + L), *bsc = bss + nblk;
    const uint32_t pgrid = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 2048);
    const uint32_t agrid = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 1024);
    auto costs_to = [&](uint64_t t) {   // cand = cost of every row to row smp[t]
        if (l2) {
            hipLaunchKernelGGL(k_cost_row_l2, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, ds.X, n, d,
                               (const uint64_t *)(smp + t), cand);
        } else {
            hipLaunchKernelGGL(k_row_to_dense, dim3(1), dim3(256), 0, st, ds.X,
                               ds.csr ? ds.val.as<const float>() : nullptr,
                               ds.csr ? ds.idx.as<const uint32_t>() : nullptr,
                               ds.csr ? ds.ptr.as<const uint64_t>() : nullptr, d, (const uint64_t *)(smp + t), row);
            generic_costs(ds, msr, prior, row, 1, nullptr, scr, cand, st);
        }
    };
    auto sum_into = [&](uint64_t t, const float *v) {   // fixed-point total of w * v into slot t
        hipLaunchKernelGGL(k_absmax<float>, dim3(agrid), dim3(256), 0, st, v, w, n, mxc + t);
        hipLaunchKernelGGL(k_fixed_blocksum<float>, dim3(nblk), dim3(256), 0, st, v, w, n,
                           (const unsigned long long *)(mxc + t), bsc + t * nblk);
    };
    std::vector<uint64_t> h(2 * L + 2 + (L + 1) * (uint64_t)nblk);
    // exact totals of the L slots and the drawn rows; returns the first strict minimum
    auto pick = [&]() {
        copy_sync(h.data(), smp, h.size() * 8, hipMemcpyDeviceToHost, st);
        uint64_t best = 0;
        double bv = 0.;
        for (uint64_t t = 0; t < L; ++t) {
            double mv;
            std::memcpy(&mv, &h[L + 2 + t], 8);
            const uint64_t *bs = h.data() + 2 * L + 2 + nblk + t * nblk;
            uint64_t T = 0;
            for (uint32_t b = 0; b < nblk; ++b) T += bs[b];
            const double v = ldexp((double)T, -fixed_shift(mv, n));
            if (t == 0 || v < bv) { bv = v; best = t; }
        }
        return best;
    };
    uint64_t j = 0;
    // first center: the cheapest of L uniform rows
    fill_u32((uint32_t *)mxc, 0u, 2 * L, st);
    for (uint64_t t = 0; t < L; ++t) {
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, smp + t, wyrand_at(seed, j++) % n);
        costs_to(t);
        float *out = t ? td + (t - 1) * n : dist_d;
        hipLaunchKernelGGL(k_local_first, dim3(pgrid), dim3(256), 0, st, (const float *)cand, n,
                           (const uint64_t *)(smp + t), out);
        sum_into(t, out);
    }
    uint64_t b = pick();
    centers_h[0] = (uint32_t)h[b];
    if (b) HIP_CHECK(hipMemcpyAsync(dist_d, td + (b - 1) * n, n * 4, hipMemcpyDeviceToDevice, st));
    fill_u32(asn_d, 0u, n, st);
    fill_u32(oa, 0u, n, st);
    HIP_CHECK(hipMemcpyAsync(od, dist_d, n * 4, hipMemcpyDeviceToDevice, st));   // never refreshed (kmeans.h:104)
    for (uint64_t ci = 1; ci < k; ++ci) {
        const uint64_t rngv = wyrand_at(seed, j++);
        // L draws over w * dist with outputs 0 .. L-1 of the stream seeded by rngv
        fill_u32((uint32_t *)mxs, 0u, 4 + 2 * L, st);   // the draw's mx (and pad), the candidates'
        hipLaunchKernelGGL(k_absmax<float>, dim3(agrid), dim3(256), 0, st, (const float *)dist_d, w, n, mxs);
        hipLaunchKernelGGL(k_fixed_blocksum<float>, dim3(nblk), dim3(256), 0, st, (const float *)dist_d, w, n,
                           (const unsigned long long *)mxs, bss);
        for (uint64_t t = 0; t < L; ++t)
            hipLaunchKernelGGL(k_fixed_select<float>, dim3(1), dim3(1024), 0, st, (const float *)dist_d, w, n,
                               (const unsigned long long *)mxs, (const uint64_t *)bss, nblk, wyrand_at(rngv, t),
                               smp + t);
        for (uint64_t t = 0; t < L; ++t) {
            costs_to(t);
            float *out = t ? td + (t - 1) * n : dist_d;
            uint32_t *out_a = t ? ta + (t - 1) * n : asn_d;
            hipLaunchKernelGGL(k_local_merge, dim3(pgrid), dim3(256), 0, st, (const float *)cand,
                               (const float *)(t ? od : dist_d), (const uint32_t *)(t ? oa : asn_d), n,
                               (uint32_t)ci, out, out_a);
            sum_into(t, out);
        }
        b = pick();
        const uint64_t newc = h[b];
        centers_h[ci] = (uint32_t)newc;
        if (b) {
            HIP_CHECK(hipMemcpyAsync(dist_d, td + (b - 1) * n, n * 4, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(asn_d, ta + (b - 1) * n, n * 4, hipMemcpyDeviceToDevice, st));
        }
        hipLaunchKernelGGL(k_local_settle, dim3(1), dim3(1), 0, st, newc, (uint32_t)ci, dist_d, asn_d);
    }
    HIP_CHECK(hipStreamSynchronize(st));
}

// ---- local search++ (lsearchpp.h:21-121) ---------------------------------------
// M[i][y] = cost of row i to center row ctrs[y] (the seeding's orientation), n x k
// floats; dense SQRL2 by the sequential FMA chain, other cases by the cost kernels.

// ||x_i - x_r||^2 by the sequential FMA chain, r = *row_p (one thread per row)
__global__ void k_cost_to_row_l2(const float *__restrict__ X, uint64_t n, uint32_t d,
                                 const uint64_t *__restrict__ row_p, float *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *x = X + i * d, *c = X + *row_p * d;
    float s = 0.f;
    for (uint32_t j = 0; j < d; ++j) { const float t = x[j] - c[j]; s = fmaf(t, t, s); }
    out[i] = s;
}

// column y of M
__global__ void k_set_column(float *__restrict__ M, uint64_t n, uint32_t k, uint32_t y, const float *__restrict__ v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        M[i * k + y] = v[i];
}

// row minima (first index on ties)
__global__ void k_row_min(const float *__restrict__ M, uint64_t n, uint32_t k, float *__restrict__ dist,
                          uint32_t *__restrict__ asn) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float *r = M + i * k;
        float m = r[0];
        uint32_t b = 0;
        for (uint32_t y = 1; y < k; ++y) if (r[y] < m) { m = r[y]; b = y; }
        dist[i] = m;
        if (asn) asn[i] = b;
    }
}

// per row: the charge (second best minus best, to the best center) of a row the
// drawn row would serve worse, else the credit (old minus new); their maximum
__global__ void k_ls_terms(const float *__restrict__ M, uint64_t n, uint32_t k, const float *__restrict__ nc,
                           const float *__restrict__ dist, float *__restrict__ t, float *__restrict__ g,
                           uint32_t *__restrict__ top, unsigned long long *__restrict__ mx) {
    double a = 0.;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float ti = 0.f, gi = 0.f;
        uint32_t j1 = 0;
        if (nc[i] > dist[i]) {
            const float *r = M + i * k;
            float v1 = r[0], v2 = r[1];
            uint32_t j2 = 1;
            if (v2 < v1) { const float tv = v1; v1 = v2; v2 = tv; j1 = 1; j2 = 0; }
            for (uint32_t y = 2; y < k; ++y) {
                const float v = r[y];
                if (v < v2) {
                    if (v < v1) { v2 = v1; j2 = j1; v1 = v; j1 = y; }
                    else { v2 = v; j2 = y; }
                }
            }
            (void)j2;
            if (v1 != v2) ti = v2 - v1;
        } else {
            gi = dist[i] - nc[i];
        }
        t[i] = ti;
        g[i] = gi;
        top[i] = j1;
        a = fmax(a, fmax((double)ti, (double)gi));
    }
    block_max_into(a, mx);
}

__global__ void k_ls_sums(const float *__restrict__ t, const float *__restrict__ g, const uint32_t *__restrict__ top,
                          uint64_t n, const unsigned long long *__restrict__ mx, unsigned long long *__restrict__ cf,
                          unsigned long long *__restrict__ gain) {
    const int s = fixed_shift(__longlong_as_double((long long)*mx), n);
    unsigned long long ga = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (t[i] > 0.f) atomicAdd(&cf[top[i]], (unsigned long long)to_fixed((double)t[i], s));
        ga += to_fixed((double)g[i], s);
    }
    for (int o = 32; o > 0; o >>= 1) ga += __shfl_xor(ga, o);
    if ((threadIdx.x & 63) == 0 && ga) atomicAdd(gain, ga);
}

void localsearch_run(Data &ds, int msr, double prior, uint64_t k, uint64_t seed, const double *w, uint64_t rounds,
                     uint64_t j0, uint32_t *centers_h, uint32_t *asn_d, float *dist_d, hipStream_t st) {
    if (rounds == 0 || k < 2) return;
    Ctx &c = *ds.ctx;
    const uint64_t n = ds.n;
    const uint32_t d = (uint32_t)ds.d;
    const bool l2 = !ds.csr && msr == MCG_SQRL2;
    DBuf bM, bNc, bT, bG, bTop, bC, bRow, bTa, bMisc, bIds;
    float *M = (float *)bM.reserve(n * k * 4), *nc = (float *)bNc.reserve(n * 4);
    float *t = (float *)bT.reserve(n * 4), *g = (float *)bG.reserve(n * 4);
    uint32_t *top = (uint32_t *)bTop.reserve(n * 4), *ta = (uint32_t *)bTa.reserve(n * 4);
    float *C = (float *)bC.reserve(k * d * 4), *row = (float *)bRow.reserve(d * 4);
    const uint32_t nblk = (uint32_t)cdiv(n, FX_BLK);
    char *misc = (char *)bMisc.reserve(64 + nblk * 8 + k * 8);
    uint64_t *sel_d = (uint64_t *)misc;
    unsigned long long *mx = (unsigned long long *)(misc + 8), *gain = (unsigned long long *)(misc + 16);
    uint64_t *bsum = (uint64_t *)(misc + 64);
    unsigned long long *cf = (unsigned long long *)(misc + 64 + nblk * 8);
    const uint32_t pgrid = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 2048);
    fill_u32((uint32_t *)nc, 0u, n, st);
    // M from the seeded centers
    {
        std::vector<uint64_t> rows(centers_h, centers_h + k);
        std::vector<uint32_t> ids(k);
        for (uint64_t y = 0; y < k; ++y) ids[y] = (uint32_t)y;
        copy_rows_to_centers(ds, rows.data(), ids.data(), k, C, st);
    }
    for (uint64_t y = 0; y < k; ++y) {
        if (l2) {
            hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, sel_d, (uint64_t)centers_h[y]);
            hipLaunchKernelGGL(k_cost_to_row_l2, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, ds.X, n, d, sel_d, t);
        } else {
            generic_costs(ds, msr, prior, C + y * d, 1, nullptr, ta, t, st);
        }
        hipLaunchKernelGGL(k_set_column, dim3(pgrid), dim3(256), 0, st, M, n, (uint32_t)k, (uint32_t)y, t);
    }
    hipLaunchKernelGGL(k_row_min, dim3(pgrid), dim3(256), 0, st, M, n, (uint32_t)k, dist_d, (uint32_t *)nullptr);
    uint64_t j = j0;   // the seeding used the stream's outputs 0 .. j0-1
    std::vector<unsigned long long> hcf(k + 1);
    for (uint64_t rd = 0; rd < rounds; ++rd) {
        // the draw: over the distances, or over w times the last round's new costs
        // (lsearchpp.h:53-56)
        const float *v = w ? nc : dist_d;
        fill_u32((uint32_t *)mx, 0u, 2, st);
        hipLaunchKernelGGL(k_absmax<float>, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 256), 1024)), dim3(256), 0, st,
                           v, w, n, mx);
        hipLaunchKernelGGL(k_fixed_blocksum<float>, dim3(nblk), dim3(256), 0, st, v, w, n,
                           (const unsigned long long *)mx, bsum);
        hipLaunchKernelGGL(k_fixed_select<float>, dim3(1), dim3(1024), 0, st, v, w, n, (const unsigned long long *)mx,
                           (const uint64_t *)bsum, nblk, wyrand_at(seed, j++), sel_d);
        if (l2) {
            hipLaunchKernelGGL(k_cost_to_row_l2, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, ds.X, n, d, sel_d,
                               nc);
        } else {
            hipLaunchKernelGGL(k_row_to_dense, dim3(1), dim3(256), 0, st, ds.X,
                               ds.csr ? ds.val.as<const float>() : nullptr,
                               ds.csr ? ds.idx.as<const uint32_t>() : nullptr,
                               ds.csr ? ds.ptr.as<const uint64_t>() : nullptr, d, (const uint64_t *)sel_d, row);
            generic_costs(ds, msr, prior, row, 1, nullptr, ta, nc, st);
        }
        fill_u32((uint32_t *)mx, 0u, 2 + 2, st);   // mx and gain
        fill_u32((uint32_t *)cf, 0u, 2 * k, st);
        hipLaunchKernelGGL(k_ls_terms, dim3(pgrid), dim3(256), 0, st, M, n, (uint32_t)k, nc, dist_d, t, g, top, mx);
        hipLaunchKernelGGL(k_ls_sums, dim3(pgrid), dim3(256), 0, st, t, g, top, n, (const unsigned long long *)mx, cf,
                           gain);
        uint64_t sel = 0;
        HIP_CHECK(hipMemcpyAsync(hcf.data(), cf, k * 8, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(&hcf[k], gain, 8, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(&sel, sel_d, 8, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        uint64_t am = 0;
        for (uint64_t y = 1; y < k; ++y) if (hcf[y] < hcf[am]) am = y;
        if (hcf[am] < hcf[k]) {   // charge below credit: swap (lsearchpp.h:104-114)
            centers_h[am] = (uint32_t)sel;
            hipLaunchKernelGGL(k_set_column, dim3(pgrid), dim3(256), 0, st, M, n, (uint32_t)k, (uint32_t)am, nc);
            hipLaunchKernelGGL(k_row_min, dim3(pgrid), dim3(256), 0, st, M, n, (uint32_t)k, dist_d,
                               (uint32_t *)nullptr);
        }
    }
    hipLaunchKernelGGL(k_row_min, dim3(pgrid), dim3(256), 0, st, M, n, (uint32_t)k, dist_d, asn_d);
    HIP_CHECK(hipStreamSynchronize(st));
}

// ---- coreset sensitivities ----------------------------------------------------
struct CsSums {
    unsigned long long maxw, maxc;           // |w|, |w*cost| maxima (bit patterns)
    unsigned long long wtot, ctot;           // fixed-point totals (two's complement)
    unsigned long long maxp, ptot;
};

// wf_i = w ? (double)(float)w_i : 1 (weights are stored as float in CoresetSampler<float>)
__device__ __forceinline__ double wf_of(const double *w, uint64_t i) { return w ? (double)(float)w[i] : 1.; }

__global__ void k_cs_max(const double *__restrict__ cost, const double *__restrict__ w, uint64_t n,
                         CsSums *__restrict__ S) {
    __shared__ double a[256], b[256];
    double ma = 0., mb = 0.;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double wf = wf_of(w, i);
        ma = fmax(ma, fabs(wf));
        mb = fmax(mb, fabs(wf * cost[i]));
    }
    a[threadIdx.x] = ma;
    b[threadIdx.x] = mb;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            a[threadIdx.x] = fmax(a[threadIdx.x], a[threadIdx.x + o]);
            b[threadIdx.x] = fmax(b[threadIdx.x], b[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax(&S->maxw, (unsigned long long)__double_as_longlong(a[0]));
        atomicMax(&S->maxc, (unsigned long long)__double_as_longlong(b[0]));
    }
}

__global__ void k_cs_sums(const double *__restrict__ cost, const uint32_t *__restrict__ asn,
                          const double *__restrict__ w, uint64_t n, CsSums *__restrict__ S,
                          unsigned long long *__restrict__ ws, unsigned long long *__restrict__ cs,
                          unsigned long long *__restrict__ cnt) {
    const int sw = fixed_shift(__longlong_as_double((long long)S->maxw), n);
    const int sc = fixed_shift(__longlong_as_double((long long)S->maxc), n);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double wf = wf_of(w, i);
        const int64_t a = to_fixed_signed(wf, sw), b = to_fixed_signed(wf * cost[i], sc);
        const uint32_t c = asn[i];
        atomicAdd(&ws[c], (unsigned long long)a);
        atomicAdd(&cs[c], (unsigned long long)b);
        atomicAdd(&cnt[c], 1ull);
        atomicAdd(&S->wtot, (unsigned long long)a);
        atomicAdd(&S->ctot, (unsigned long long)b);
    }
}

// The same sums for up to CS_LDS clusters: each block adds its rows into LDS copies
// of the per-cluster sums and its totals by a block reduction, then adds them to the
// global sums (one atomic per block and non-empty cluster instead of five per row,
// two of them on a single address).  Integer sums: the same totals in any order.
constexpr uint32_t CS_LDS = 2048;
__global__ __launch_bounds__(256) void k_cs_sums_lds(const double *__restrict__ cost, const uint32_t *__restrict__ asn,
                                                     const double *__restrict__ w, uint64_t n, uint32_t ncenters,
                                                     CsSums *__restrict__ S, unsigned long long *__restrict__ ws,
                                                     unsigned long long *__restrict__ cs,
                                                     unsigned long long *__restrict__ cnt) {
    __shared__ unsigned long long lw[CS_LDS], lc[CS_LDS], ln[CS_LDS];
    __shared__ unsigned long long rw[256], rc[256];
    for (uint32_t c = threadIdx.x; c < ncenters; c += 256) { lw[c] = 0; lc[c] = 0; ln[c] = 0; }
    __syncthreads();
    const int sw = fixed_shift(__longlong_as_double((long long)S->maxw), n);
    const int sc = fixed_shift(__longlong_as_double((long long)S->maxc), n);
    unsigned long long tw = 0, tc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double wf = wf_of(w, i);
        const int64_t a = to_fixed_signed(wf, sw), b = to_fixed_signed(wf * cost[i], sc);
        const uint32_t c = asn[i];
        atomicAdd(&lw[c], (unsigned long long)a);
        atomicAdd(&lc[c], (unsigned long long)b);
        atomicAdd(&ln[c], 1ull);
        tw += (unsigned long long)a;
        tc += (unsigned long long)b;
    }
    rw[threadIdx.x] = tw;
    rc[threadIdx.x] = tc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) { rw[threadIdx.x] += rw[threadIdx.x + o]; rc[threadIdx.x] += rc[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&S->wtot, rw[0]);
        atomicAdd(&S->ctot, rc[0]);
    }
    for (uint32_t c = threadIdx.x; c < ncenters; c += 256)
        if (ln[c]) {
            atomicAdd(&ws[c], lw[c]);
            atomicAdd(&cs[c], lc[c]);
            atomicAdd(&cnt[c], ln[c]);
        }
}

// unnormalised sensitivities p_i (double), plus max |p|
__global__ void k_cs_sens(int sens, const double *__restrict__ cost, const uint32_t *__restrict__ asn,
                          const double *__restrict__ w, uint64_t n, double alpha, CsSums *__restrict__ S,
                          const unsigned long long *__restrict__ ws, const unsigned long long *__restrict__ cs,
                          const unsigned long long *__restrict__ cnt, double *__restrict__ p, uint64_t nfl,
                          float *__restrict__ probs_fl) {
    __shared__ double red[256];
    const int sw = fixed_shift(__longlong_as_double((long long)S->maxw), n);
    const int sc = fixed_shift(__longlong_as_double((long long)S->maxc), n);
    const double W = from_fixed((int64_t)S->wtot, sw), TC = from_fixed((int64_t)S->ctot, sc);
    double m = 0.;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = asn[i];
        const double wf = wf_of(w, i);
        const double wc = wf * cost[i];
        double v;
        if (sens == MCG_SENS_LBK) {
            const double mean = TC / W, tcinv = alpha / mean, alpha2 = 2. * alpha;
            const double Wc = from_fixed((int64_t)ws[a], sw), Sc = from_fixed((int64_t)cs[a], sc);
            const double adj = alpha2 * Sc / (Wc * mean) + 4. * W / Wc;
            v = tcinv * cost[i] + adj;
        } else if (sens == MCG_SENS_LFKF) {
            // 2 T / W_a (coreset.h:374-375 with alpha_est = 0); the common 2 T cancels
            // in the normalisation
            v = 1. / from_fixed((int64_t)ws[a], sw);
        } else if (sens == MCG_SENS_VX) {
            v = wc / TC + 1. / (double)cnt[a];
        } else if (sens == MCG_SENS_BFL) {
            const double Wc = from_fixed((int64_t)ws[a], sw);
            v = .5 * (wc / TC + wf / (Wc * (double)cnt[a]));
        } else {  // FL: not normalised
            probs_fl[i] = w ? (float)(wc / TC) : (float)(ceil((double)nfl * cost[i] / TC) + 1.);
            continue;
        }
        p[i] = v;
        m = fmax(m, fabs(v));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&S->maxp, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_cs_psum(const double *__restrict__ p, uint64_t n, CsSums *__restrict__ S) {
    const int s = fixed_shift(__longlong_as_double((long long)S->maxp), n);
    __shared__ unsigned long long red[256];
    unsigned long long acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        acc += (unsigned long long)to_fixed_signed(p[i], s);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&S->ptot, red[0]);
}

__global__ void k_cs_norm(const double *__restrict__ p, uint64_t n, const CsSums *__restrict__ S,
                          float *__restrict__ probs) {
    const int s = fixed_shift(__longlong_as_double((long long)S->maxp), n);
    const double P = from_fixed((int64_t)S->ptot, s);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        probs[i] = (float)(p[i] / P);
}

uint64_t coreset_scratch_bytes(uint64_t n, uint64_t ncenters) {
    return cdiv(sizeof(CsSums) + 3 * ncenters * 8, 256) * 256 + n * 8;
}

void coreset_probs_run(Ctx &c, int sens, uint64_t n, uint64_t ncenters, const double *cost, const uint32_t *asn,
                       const double *w, float *probs, hipStream_t st) {
    void *scratch = c.s[S_MISC0].reserve(coreset_scratch_bytes(n, ncenters));
    coreset_probs_core(sens, n, ncenters, cost, asn, w, probs, scratch, st);
}

void coreset_probs_core(int sens, uint64_t n, uint64_t ncenters, const double *cost, const uint32_t *asn,
                        const double *w, float *probs, void *scratch, hipStream_t st) {
    REQUIRE(sens == MCG_SENS_LBK || sens == MCG_SENS_VX || sens == MCG_SENS_BFL || sens == MCG_SENS_FL ||
                sens == MCG_SENS_LFKF,
            "sensitivity method must be BFL (0), FL (1), LFKF (2), VX (3) or LBK (4)");
    // scratch: the sums (cleared by a fill kernel, so that the work can be captured into
    // a graph), then the unnormalised sensitivities at the next 256-byte boundary
    const uint64_t head = cdiv(sizeof(CsSums) + 3 * ncenters * 8, 256) * 256;
    char *buf = (char *)scratch;
    fill_u32((uint32_t *)buf, 0u, (sizeof(CsSums) + 3 * ncenters * 8) / 4, st);
    CsSums *S = (CsSums *)buf;
    unsigned long long *ws = (unsigned long long *)(buf + sizeof(CsSums));
    unsigned long long *cs = ws + ncenters, *cnt = cs + ncenters;
    double *p = (double *)(buf + head);
    const uint32_t g = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 2048);
    hipLaunchKernelGGL(k_cs_max, dim3(g), dim3(256), 0, st, cost, w, n, S);
    if (ncenters <= CS_LDS)
        hipLaunchKernelGGL(k_cs_sums_lds, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 2048), 128)), dim3(256), 0, st,
                           cost, asn, w, n, (uint32_t)ncenters, S, ws, cs, cnt);
    else
        hipLaunchKernelGGL(k_cs_sums, dim3(g), dim3(256), 0, st, cost, asn, w, n, S, ws, cs, cnt);
    const double alpha = 16. * std::log((double)ncenters) + 32.;
    hipLaunchKernelGGL(k_cs_sens, dim3(g), dim3(256), 0, st, sens, cost, asn, w, n, alpha, S, ws, cs, cnt, p, n,
                       probs);
    if (sens == MCG_SENS_FL) return;
    hipLaunchKernelGGL(k_cs_psum, dim3(g), dim3(256), 0, st, p, n, S);
    hipLaunchKernelGGL(k_cs_norm, dim3(g), dim3(256), 0, st, p, n, S, probs);
}

// ---- coreset draws ----------------------------------------------------------------
// inclusive prefix of fixed weights: per-block sums -> host scan (nblk small) -> per-block
// prefixes; each draw binary-searches blocks, then one wave finds the first index of
// its block whose inclusive prefix exceeds the draw: every lane adds 16 consecutive
// weights, a wave scan places the lanes' runs, and the first lane whose run passes
// the draw walks its 16.  (Integer sums: the index of a sequential walk.)
__global__ __launch_bounds__(256) void k_draws(const float *__restrict__ probs, const double *__restrict__ w,
                                               uint64_t n, const unsigned long long *__restrict__ maxbits,
                                               const uint64_t *__restrict__ bpre, uint32_t nblk, uint64_t T,
                                               uint64_t m, uint64_t seed, uint64_t first, uint64_t *__restrict__ idx,
                                               float *__restrict__ wout) {
    static_assert(FX_BLK == 64 * 16, "16 weights per lane");
    const uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (t >= m) return;
    const int s = fixed_shift(__longlong_as_double((long long)*maxbits), n);
    const uint64_t u = mulhi64(wyrand_at(seed, first + t), T);
    // first block with inclusive prefix > u
    uint32_t lo = 0, hi = nblk;
    while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (bpre[mid] > u) hi = mid; else lo = mid + 1; }
    const uint32_t b = lo < nblk ? lo : nblk - 1;
    const uint64_t run0 = b ? bpre[b - 1] : 0;
    const uint64_t i0 = (uint64_t)b * FX_BLK + (uint64_t)lane * 16;
    unsigned long long f[16], sum = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        f[j] = i0 + j < n ? to_fixed((double)probs[i0 + j], s) : 0ull;
        sum += f[j];
    }
    unsigned long long incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const uint64_t hit = __builtin_amdgcn_ballot_w64(run0 + incl > u);
    uint64_t ans = min<uint64_t>(n - 1, (uint64_t)(b + 1) * FX_BLK - 1);
    if (hit) {
        const int L = __builtin_ctzll(hit);
        uint64_t mine = 0;
        if (lane == L) {
            uint64_t run = run0 + incl - sum;
            mine = i0 + 15;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                run += f[j];
                if (run > u) { mine = i0 + j; break; }
            }
        }
        ans = __shfl(mine, L);
    }
    if (lane == 0) {
        idx[t] = ans;
        const double wi = w ? (double)(float)w[ans] : 1.;
        wout[t] = (float)(wi / ((double)m * (double)probs[ans]));
    }
}

void coreset_sample_run(Ctx &c, uint64_t n, const float *probs, const double *w, uint64_t m, uint64_t seed,
                        uint64_t *idx, float *wout, hipStream_t st) {
    coreset_sample_from(c, n, probs, w, m, seed, 0, idx, wout, st);
}

void coreset_sample_from(Ctx &c, uint64_t n, const float *probs, const double *w, uint64_t m, uint64_t seed,
                         uint64_t first, uint64_t *idx, float *wout, hipStream_t st) {
    const uint32_t nblk = (uint32_t)cdiv(n, FX_BLK);
    char *buf = (char *)c.s[S_MISC0].reserve(64 + nblk * 8);
    unsigned long long *mx = (unsigned long long *)buf;
    uint64_t *bsum = (uint64_t *)(buf + 64);
    fill_u32((uint32_t *)mx, 0, 2, st);
    hipLaunchKernelGGL(k_absmax<float>, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 256), 1024)), dim3(256), 0, st,
                       probs, (const double *)nullptr, n, mx);
    hipLaunchKernelGGL(k_fixed_blocksum<float>, dim3(nblk), dim3(256), 0, st, probs, (const double *)nullptr, n,
                       (const unsigned long long *)mx, bsum);
    std::vector<uint64_t> h(nblk);
    HIP_CHECK(hipMemcpyAsync(h.data(), bsum, nblk * 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    uint64_t T = 0;
    for (uint32_t b = 0; b < nblk; ++b) { T += h[b]; h[b] = T; }
    REQUIRE(T > 0, "all sampling probabilities are zero");
    HIP_CHECK(hipMemcpyAsync(bsum, h.data(), nblk * 8, hipMemcpyHostToDevice, st));
    if (m) hipLaunchKernelGGL(k_draws, dim3((uint32_t)cdiv(m, 4)), dim3(256), 0, st, probs, w, n,
                              (const unsigned long long *)mx, (const uint64_t *)bsum, nblk, T, m, seed, first, idx, wout);
    HIP_CHECK(hipStreamSynchronize(st));
}

// One restart draw on the device: draw j of the wyrand stream of n over the costs,
// the fixed-point sampler of draw_fixed_many (all-zero costs give row 0, as there).
uint64_t restart_draw_device(Ctx &c, const float *costs_d, uint64_t n, uint64_t j, hipStream_t st) {
    const uint32_t nblk = (uint32_t)cdiv(n, FX_BLK);
    char *buf = (char *)c.s[S_MISC0].reserve(64 + nblk * 8 + 16);
    unsigned long long *mx = (unsigned long long *)buf;
    uint64_t *bsum = (uint64_t *)(buf + 64);
    uint64_t *idx = (uint64_t *)(buf + 64 + nblk * 8);
    float *wout = (float *)(idx + 1);
    fill_u32((uint32_t *)mx, 0, 2, st);
    hipLaunchKernelGGL(k_absmax<float>, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 256), 1024)), dim3(256), 0, st,
                       costs_d, (const double *)nullptr, n, mx);
    hipLaunchKernelGGL(k_fixed_blocksum<float>, dim3(nblk), dim3(256), 0, st, costs_d, (const double *)nullptr, n,
                       (const unsigned long long *)mx, bsum);
    std::vector<uint64_t> h(nblk);
    copy_sync(h.data(), bsum, nblk * 8, hipMemcpyDeviceToHost, st);
    uint64_t T = 0;
    for (uint32_t b = 0; b < nblk; ++b) { T += h[b]; h[b] = T; }
    if (T == 0) return 0;
    HIP_CHECK(hipMemcpyAsync(bsum, h.data(), nblk * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_draws, dim3(1), dim3(64), 0, st, costs_d, (const double *)nullptr, n,
                       (const unsigned long long *)mx, (const uint64_t *)bsum, nblk, T, (uint64_t)1, n, j, idx, wout);
    uint64_t row = 0;
    copy_sync(&row, idx, 8, hipMemcpyDeviceToHost, st);
    return row;
}

}  // namespace mcg
