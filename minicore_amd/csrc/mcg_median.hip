// Median centroid policies: set_centroids_l1 / set_centroids_tvd / set_centroids_l2
// (include/minicore/clustering/centroid.h:434-647) with l1_median / weighted_median
// / tvd_median (optim/kmedian.h:124-350, util/csc.h:829-950) and the dense
// geometric median (util/blaze_adaptor.h:614-690).
//
// After the stable sort of point ids by cluster (members contiguous, ascending):
//  - L1 / TVD, unweighted, dense: one block per (cluster, 64 columns); a radix
//    select over order-preserving keys of the members' values (8-bit digits, LDS
//    histograms laid out digit-major so the 64 lanes of a wave hit 64 banks), one
//    pass per key byte plus one for the upper middle value of an even count.
//    Reads each member value once per pass; HBM-bound (passes x n x d x 4 bytes).
//  - weighted L1 / TVD: the reference's weighted_median never adds to its weight
//    total (kmedian.h:275-285), so each column takes the value of the heaviest
//    member (TVD: the weight of row j for the member of rank j, kmedian.h:340);
//    one pass, smallest value among equal heaviest weights.
//  - CSR L1 / TVD (unweighted): each (cluster, column) segment's stored values are
//    gathered by a counting sort, and one wave per segment applies csc.h's rules
//    (implicit zeros first) with a wave radix select.
//  - geometric median: Weiszfeld rounds for all clusters at once (each stops on its
//    own rule), FP64 with fixed summation orders so the result is reproducible.
#include "mcg_internal.h"
#include <algorithm>
#include <cfloat>

namespace mcg {

constexpr uint32_t MED_COLS = 64;    // columns per block, one per lane
constexpr uint32_t MED_GROUPS = 4;   // member groups per block of 256 threads
constexpr uint64_t GEO_MAX_ROUNDS = 100000;   // the reference's iteration limit (csc.h geomedian)

// Order-preserving unsigned keys of floating values.
__device__ __forceinline__ uint32_t okey(float v) {
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float okey_value(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
__device__ __forceinline__ uint64_t okey(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ __forceinline__ double okey_value(uint64_t k) {
    return __longlong_as_double((long long)((k >> 63) ? (k & ~(1ull << 63)) : ~k));
}

template <bool TVD> struct MedT { using K = uint32_t; using V = float; };
template <> struct MedT<true> { using K = uint64_t; using V = double; };

// member value at sorted position p (rows gathered in cluster order): x (L1) or
// x * (1 / rowsum) in FP64 (TVD, kmedian.h:349)
template <bool TVD>
__device__ __forceinline__ typename MedT<TVD>::V med_value(const float *__restrict__ Xp, uint32_t d, uint32_t p,
                                                           uint32_t j, const double *__restrict__ invp) {
    const float x = Xp[(uint64_t)p * d + j];
    if constexpr (TVD) return __dmul_rn((double)x, invp[p]);
    else return x;
}

// rows in cluster order: Xp[p] = X[perm[p]], one wave per row
__global__ __launch_bounds__(256) void k_gather_perm(const float *__restrict__ X, uint32_t d,
                                                     const uint32_t *__restrict__ perm, uint64_t n,
                                                     float *__restrict__ Xp) {
    const int lane = threadIdx.x & 63;
    for (uint64_t p = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); p < n; p += (uint64_t)gridDim.x * 4) {
        const float *src = X + (uint64_t)perm[p] * d;
        float *dst = Xp + p * d;
        for (uint32_t j = lane; j < d; j += 64) dst[j] = src[j];
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Per-column median of the members of cluster blockIdx.y (m >= 2): rank h for odd
// m; ranks h-1 and h averaged for even m (0.5 * (a + b) in the value type).  COLS
// columns per block (one per thread of a column group), 256 / COLS member groups.
template <bool TVD, uint32_t COLS>
__global__ __launch_bounds__(256) void k_med_select(const float *__restrict__ Xp, uint32_t d,
                                                    const uint32_t *__restrict__ counts,
                                                    const uint32_t *__restrict__ off,
                                                    const double *__restrict__ invp, uint32_t min_m,
                                                    float *__restrict__ C) {
    using K = typename MedT<TVD>::K;
    using V = typename MedT<TVD>::V;
    constexpr int BITS = 8 * (int)sizeof(K);
    __shared__ uint32_t hist[256 * COLS];
    __shared__ K pref[COLS], key2[COLS];
    __shared__ uint32_t rnk[COLS], eqc[COLS], nd2[COLS];
    __shared__ int need2;
    const uint32_t c = blockIdx.y;
    const uint32_t m = counts[c];
    if (m < min_m) return;
    constexpr uint32_t GROUPS = 256 / COLS;
    const int tid = threadIdx.x, lane = tid % (int)COLS, g = tid / (int)COLS;
    const uint32_t j = blockIdx.x * COLS + lane;
    const bool valid = j < d;
    const uint32_t b0 = off[c];
    const uint32_t h = m >> 1;
    if (tid < (int)COLS) {
        pref[tid] = 0;
        rnk[tid] = (m & 1) ? h : h - 1;
        nd2[tid] = 0;
    }
    if (tid == 0) need2 = 0;
    for (int p = 0; p < BITS / 8; ++p) {
        const int shift = BITS - 8 * (p + 1);
        for (uint32_t q = tid; q < 256 * COLS; q += 256) hist[q] = 0;
        __syncthreads();
        if (valid) {
            const K pr = pref[lane];
            const K hi = p == 0 ? (K)0 : (K)(~(K)0 << (shift + 8));
            // 8 members per step: their loads are independent and in flight together
            constexpr uint32_t U = 8;
            for (uint32_t q0 = g; q0 < m; q0 += U * GROUPS) {
                K key[U];
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t q = q0 + u * GROUPS;
                    key[u] = q < m ? okey(med_value<TVD>(Xp, d, b0 + q, j, invp)) : (K)0;
                }
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    if (q0 + u * GROUPS < m && (key[u] & hi) == pr)
                        atomicAdd(&hist[(uint32_t)((key[u] >> shift) & 255u) * COLS + lane], 1u);
            }
        }
        __syncthreads();
        if (tid < (int)COLS && valid) {
            const uint32_t r = rnk[tid];
            uint32_t cum = 0;
            for (uint32_t b = 0; b < 256; ++b) {
                const uint32_t hb = hist[b * COLS + tid];
                if (cum + hb > r) {
                    pref[tid] |= (K)b << shift;
                    rnk[tid] = r - cum;
                    eqc[tid] = hb;
                    break;
                }
                cum += hb;
            }
        }
        __syncthreads();
    }
    if (!(m & 1)) {
        // rank h: the same key while equal keys remain, else the smallest larger key
        if (tid < (int)COLS && valid) {
            if (rnk[tid] + 1 < eqc[tid]) key2[tid] = pref[tid];
            else { key2[tid] = ~(K)0; nd2[tid] = 1; need2 = 1; }
        }
        __syncthreads();
        if (need2) {
            if (valid && nd2[lane]) {
                const K k1 = pref[lane];
                K best = ~(K)0;
                for (uint32_t q = g; q < m; q += GROUPS) {
                    const K key = okey(med_value<TVD>(Xp, d, b0 + q, j, invp));
                    if (key > k1 && key < best) best = key;
                }
                atomicMin(&key2[lane], best);
            }
            __syncthreads();
        }
    }
    if (tid < (int)COLS && valid) {
        const V v1 = okey_value(pref[tid]);
        V out = v1;
        if (!(m & 1)) out = (V)0.5 * (v1 + okey_value(key2[tid]));
        C[(uint64_t)c * d + j] = (float)out;
    }
}

// Weighted: per column the smallest value among the members of largest weight:
// wm[row] (set_centroids_l1), or wr[rank] where the reference indexes the weight
// vector by member rank (TVD; L1 in the minibatch update).
template <bool TVD>
__global__ __launch_bounds__(256) void k_med_heaviest(const float *__restrict__ Xp, uint32_t d,
                                                      const uint32_t *__restrict__ perm,
                                                      const uint32_t *__restrict__ counts,
                                                      const uint32_t *__restrict__ off,
                                                      const double *__restrict__ invp, const double *__restrict__ wm,
                                                      const double *__restrict__ wr, int byrank, uint32_t min_m,
                                                      float *__restrict__ C) {
    using K = typename MedT<TVD>::K;
    __shared__ double wl[256];
    __shared__ K best[MED_COLS];
    const uint32_t c = blockIdx.y;
    const uint32_t m = counts[c];
    if (m < min_m) return;
    const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
    const uint32_t j = blockIdx.x * MED_COLS + lane;
    const uint32_t *mem = perm + off[c];
    double wmx = -DBL_MAX;
    for (uint32_t q = tid; q < m; q += 256) wmx = fmax(wmx, byrank ? wr[q] : wm[mem[q]]);
    wl[tid] = wmx;
    if (tid < (int)MED_COLS) best[tid] = ~(K)0;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) wl[tid] = fmax(wl[tid], wl[tid + o]);
        __syncthreads();
    }
    const double wmax = wl[0];
    if (j < d) {
        K b = ~(K)0;
        for (uint32_t q = g; q < m; q += MED_GROUPS) {
            const double wq = byrank ? wr[q] : wm[mem[q]];
            if (wq == wmax) {
                const K key = okey(med_value<TVD>(Xp, d, off[c] + q, j, invp));
                b = key < b ? key : b;
            }
        }
        atomicMin(&best[lane], b);
    }
    __syncthreads();
    if (tid < (int)MED_COLS && j < d) C[(uint64_t)c * d + j] = (float)okey_value(best[tid]);
}

// single-member clusters: the center is the member row (raw, also under TVD)
__global__ void k_med_single(const float *__restrict__ X, uint32_t d, const uint32_t *__restrict__ perm,
                             const uint32_t *__restrict__ counts, const uint32_t *__restrict__ off,
                             float *__restrict__ C) {
    const uint32_t c = blockIdx.y;
    if (counts[c] != 1) return;
    const uint64_t r = perm[off[c]];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < d; j += gridDim.x * blockDim.x)
        C[(uint64_t)c * d + j] = X[r * d + j];
}

__global__ void k_med_single_csr(const float *__restrict__ val, const uint32_t *__restrict__ idx,
                                 const uint64_t *__restrict__ ptr, uint32_t d, const uint32_t *__restrict__ perm,
                                 const uint32_t *__restrict__ counts, const uint32_t *__restrict__ off,
                                 float *__restrict__ C) {
    const uint32_t c = blockIdx.x;
    if (counts[c] != 1) return;
    const uint64_t r = perm[off[c]];
    float *ctr = C + (uint64_t)c * d;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) ctr[j] = 0.f;
    __syncthreads();
    for (uint64_t p = ptr[r] + threadIdx.x; p < ptr[r + 1]; p += blockDim.x) ctr[idx[p]] = val[p];
}

// 1 / (row sum), the sum sequential in FP64: one thread per row, 32-dim slices of
// the block's 256 rows staged in LDS with coalesced loads
__global__ __launch_bounds__(256) void k_inv_rowsum_seq(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                        double *__restrict__ inv) {
    __shared__ float xs[256][33];
    const int tid = threadIdx.x;
    const uint64_t p0 = (uint64_t)blockIdx.x * 256;
    double s = 0.;
    for (uint32_t j0 = 0; j0 < d; j0 += 32) {
        __syncthreads();
        for (int q = tid; q < 256 * 32; q += 256) {
            const int r = q >> 5, jj = q & 31;
            const uint64_t pr = p0 + r;
            xs[r][jj] = (pr < n && j0 + jj < d) ? X[pr * d + j0 + jj] : 0.f;
        }
        __syncthreads();
        const int jn = (int)min<uint32_t>(32, d - j0);
        for (int jj = 0; jj < jn; ++jj) s = __dadd_rn(s, (double)xs[tid][jj]);
    }
    if (p0 + tid < n) inv[p0 + tid] = __ddiv_rn(1., s);
}

__global__ void k_inv_rowsum_csr_seq(const float *__restrict__ val, const uint64_t *__restrict__ ptr, uint64_t n,
                                     double *__restrict__ inv) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.;
    for (uint64_t p = ptr[i]; p < ptr[i + 1]; ++p) s = __dadd_rn(s, (double)val[p]);
    inv[i] = __ddiv_rn(1., s);
}

// a row moves to restarted center id when that is strictly cheaper
__global__ void k_take_lower1(const float *__restrict__ v, uint32_t id, uint64_t n, float *__restrict__ pr,
                              uint32_t *__restrict__ Ar) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (v[i] < pr[i]) { pr[i] = v[i]; Ar[i] = id; }
}

__global__ void k_any_negative(const float *__restrict__ v, uint64_t n, uint32_t *__restrict__ flag) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (v[i] < 0.f) *flag = 1u;
}

// ---- CSR: (cluster, column) segments ---------------------------------------------
__global__ __launch_bounds__(256) void k_seg_count(const uint32_t *__restrict__ idx, const uint64_t *__restrict__ ptr,
                                                   uint64_t n, uint32_t d, const uint32_t *__restrict__ A, uint32_t k,
                                                   uint32_t *__restrict__ cnt) {
    const uint64_t row = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const uint32_t a = A[row];
    if (a >= k) return;
    for (uint64_t p = ptr[row] + lane; p < ptr[row + 1]; p += 64) atomicAdd(&cnt[(uint64_t)a * d + idx[p]], 1u);
}

// value of each stored entry into its segment (order inside a segment is free: only
// order statistics are taken).  TVD: float(x) * float(1 / rowsum) (csc.h:908-911).
__global__ __launch_bounds__(256) void k_seg_fill(const float *__restrict__ val, const uint32_t *__restrict__ idx,
                                                  const uint64_t *__restrict__ ptr, uint64_t n, uint32_t d,
                                                  const uint32_t *__restrict__ A, uint32_t k,
                                                  const double *__restrict__ inv, uint32_t *__restrict__ cursor,
                                                  float *__restrict__ out) {
    const uint64_t row = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const uint32_t a = A[row];
    if (a >= k) return;
    const float im = inv ? (float)inv[row] : 1.f;
    for (uint64_t p = ptr[row] + lane; p < ptr[row + 1]; p += 64) {
        const uint32_t pos = atomicAdd(&cursor[(uint64_t)a * d + idx[p]], 1u);
        out[pos] = inv ? val[p] * im : val[p];
    }
}

// t-th smallest (0-based) of v[0..nel), all lanes of the calling wave; hist: 256
// LDS words private to the wave.
__device__ float wave_select(const float *__restrict__ v, uint32_t nel, uint32_t t, uint32_t *hist) {
    const int lane = threadIdx.x & 63;
    if (nel <= 64) {
        const bool in = (uint32_t)lane < nel;
        const float x = in ? v[lane] : 0.f;
        const uint32_t key = in ? okey(x) : 0xFFFFFFFFu;
        uint32_t r = 0;
        for (uint32_t l = 0; l < nel; ++l) {
            const uint32_t o = __shfl(key, (int)l);
            r += (o < key || (o == key && (int)l < lane)) ? 1u : 0u;
        }
        const uint64_t b = __ballot(in && r == t);
        return __shfl(x, __ffsll((long long)b) - 1);
    }
    uint32_t pref = 0, r = t;
    for (int p = 0; p < 4; ++p) {
        const int shift = 24 - 8 * p;
        for (int b = lane; b < 256; b += 64) hist[b] = 0;
        wave_sync();
        const uint32_t hi = p == 0 ? 0u : (~0u << (shift + 8));
        for (uint32_t q = lane; q < nel; q += 64) {
            const uint32_t key = okey(v[q]);
            if ((key & hi) == pref) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        wave_sync();
        uint32_t digit = 0, rr = r;
        if (lane == 0) {
            uint32_t cum = 0;
            for (uint32_t b = 0; b < 256; ++b) {
                const uint32_t hb = hist[b];
                if (cum + hb > r) { digit = b; rr = r - cum; break; }
                cum += hb;
            }
        }
        digit = __shfl(digit, 0);
        r = __shfl(rr, 0);
        pref |= digit << shift;
        wave_sync();
    }
    return okey_value(pref);
}

// One wave per (cluster, column) segment of a cluster with >= 2 members: the rules of
// csc.h:852-884 (L1) and 924-950 (TVD) for non-negative values, with (m - nel)
// implicit zeros before the nel stored values; where the reference reads one past
// the values (m == 2 under the even rule h+1) the index is clamped to the last.
__global__ __launch_bounds__(256) void k_seg_median(const float *__restrict__ vals,
                                                    const uint32_t *__restrict__ segoff,
                                                    const uint32_t *__restrict__ counts, uint32_t d, uint64_t nseg,
                                                    int tvd, uint32_t min_m, float *__restrict__ C) {
    __shared__ uint32_t hist[4][256];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint64_t s = (uint64_t)blockIdx.x * 4 + wv; s < nseg; s += (uint64_t)gridDim.x * 4) {
        const uint32_t c = (uint32_t)(s / d), j = (uint32_t)(s % d);
        const uint32_t m = counts[c];
        if (m < min_m) continue;
        const uint32_t b = segoff[s], nel = segoff[s + 1] - b;
        float out = 0.f;
        if (nel) {
            const float *v = vals + b;
            const uint32_t h = m >> 1, z = m - nel;
            const bool odd = m & 1;
            const float s0 = wave_select(v, nel, 0, hist[wv]);
            auto full = [&](uint32_t r) -> float { return r < z ? 0.f : wave_select(v, nel, r - z, hist[wv]); };
            if (nel > h) {
                if (odd) {
                    out = full(h);
                } else {
                    const uint32_t r1 = h + 1 < m ? h + 1 : m - 1;
                    if (s0 > 0.f || !tvd) out = (float)(0.5 * (double)(full(h) + full(r1)));
                    else out = (float)(0.5 * (double)(full(h - 1) + full(h)));
                }
            } else if (odd && nel == h && s0 > 0.f) {
                out = (float)(0.5 * (double)s0);
            }
        }
        if (lane == 0) C[(uint64_t)c * d + j] = out;
    }
}

// ---- geometric median (Weiszfeld) ----------------------------------------------------
__global__ void k_geo_init(const uint32_t *__restrict__ counts, uint64_t k, float *__restrict__ prev,
                           double *__restrict__ prevd, uint32_t *__restrict__ active) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    prev[c] = FLT_MAX;
    prevd[c] = DBL_MAX;
    active[c] = counts[c] >= 2 ? 1u : 0u;
}

// CSR rows in cluster order as dense rows (Xp cleared beforehand): one wave per row
__global__ __launch_bounds__(256) void k_scatter_perm_csr(const float *__restrict__ val,
                                                          const uint32_t *__restrict__ idx,
                                                          const uint64_t *__restrict__ ptr, uint32_t d,
                                                          const uint32_t *__restrict__ perm, uint64_t n,
                                                          float *__restrict__ Xp) {
    const uint64_t p = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= n) return;
    const uint32_t r = perm[p];
    for (uint64_t q = ptr[r] + (threadIdx.x & 63); q < ptr[r + 1]; q += 64) Xp[p * d + idx[q]] = val[q];
}

// per member (sorted position p): distance to its cluster's current center, FP64
// sequential FMA, rounded to float (kept in FP64 when `exact`), floored;
// tq = w * dist (or dist)
__global__ __launch_bounds__(256) void k_geo_dist(const float *__restrict__ Xp, uint32_t d,
                                                  const uint32_t *__restrict__ perm, uint64_t n,
                                                  const uint32_t *__restrict__ A, const float *__restrict__ C,
                                                  const uint32_t *__restrict__ active, const double *__restrict__ w,
                                                  double floor_, int exact, double *__restrict__ cq,
                                                  double *__restrict__ tq) {
    __shared__ float xs[256][33];
    const int tid = threadIdx.x;
    const uint64_t p0 = (uint64_t)blockIdx.x * 256, p = p0 + tid;
    const uint32_t r = p < n ? perm[p] : 0u;
    const uint32_t c = p < n ? A[r] : 0u;
    const bool on = p < n && active[c];
    const float *cc = C + (uint64_t)c * d;
    double s = 0.;
    // 32-dim slices of the block's 256 gathered rows, staged with coalesced loads
    for (uint32_t j0 = 0; j0 < d; j0 += 32) {
        __syncthreads();
        for (int q = tid; q < 256 * 32; q += 256) {
            const int rr = q >> 5, jj = q & 31;
            const uint64_t pr = p0 + rr;
            xs[rr][jj] = (pr < n && j0 + jj < d) ? Xp[pr * d + j0 + jj] : 0.f;
        }
        __syncthreads();
        if (on) {
            const int jn = (int)min<uint32_t>(32, d - j0);
            for (int jj = 0; jj < jn; ++jj) {
                const double t = __dsub_rn((double)xs[tid][jj], (double)cc[j0 + jj]);
                s = __fma_rn(t, t, s);
            }
        }
    }
    if (!on) return;
    double v = exact ? __dsqrt_rn(s) : (double)(float)__dsqrt_rn(s);
    if (v != v) v = floor_;
    v = v > floor_ ? v : floor_;
    cq[p] = v;
    tq[p] = w ? __dmul_rn(w[r], v) : v;
}

// 256 strided sequential partials added as a binary tree (the order the oracle uses)
__device__ double block_tree_sum(double s, double *sm) {
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] = __dadd_rn(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    const double r = sm[0];
    __syncthreads();
    return r;
}

// per cluster: the round's cost, the stop rule (dense: blaze_adaptor.h:638-646, the
// cost in float; CSR rule: csc.h:1150-1158, the cost in FP64, stopping when it is
// within eps of the previous one or above it), Q = sum q_i
__global__ __launch_bounds__(256) void k_geo_sums(const uint32_t *__restrict__ perm,
                                                  const uint32_t *__restrict__ counts,
                                                  const uint32_t *__restrict__ off, const double *__restrict__ cq,
                                                  double *__restrict__ tq, const double *__restrict__ w,
                                                  double eps, int csr_rule, float *__restrict__ prev,
                                                  double *__restrict__ prevd, uint32_t *__restrict__ active,
                                                  double *__restrict__ Q, uint32_t *__restrict__ nanflag) {
    __shared__ double sm[256];
    __shared__ int stop;
    const uint32_t c = blockIdx.x;
    if (!active[c]) return;
    const uint32_t m = counts[c], b = off[c];
    double s = 0.;
    for (uint32_t q = threadIdx.x; q < m; q += 256) s = __dadd_rn(s, tq[b + q]);
    const double cur = block_tree_sum(s, sm);
    const float curf = (float)cur;
    if (threadIdx.x == 0) {
        stop = 0;
        if (csr_rule) {
            const double dist = fabs(prevd[c] - cur);
            if (dist <= eps || cur > prevd[c]) stop = 1;
            else if (dist != dist) { stop = 1; *nanflag = 1u; }
        } else {
            const float dist = prev[c] - curf;
            if ((double)dist <= eps) stop = 1;
            else if (dist != dist) { stop = 1; *nanflag = 1u; }
        }
        if (stop) active[c] = 0u;
        else { prev[c] = curf; prevd[c] = cur; }
    }
    __syncthreads();
    if (stop) return;
    s = 0.;
    for (uint32_t q = threadIdx.x; q < m; q += 256) {
        const double cv = cq[b + q];
        s = __dadd_rn(s, w ? __ddiv_rn(w[perm[b + q]], cv) : __ddiv_rn(1., cv));
    }
    const double tot = block_tree_sum(s, sm);
    if (threadIdx.x == 0) Q[c] = tot;
    // the members' coefficients of the new center, into tq (its costs are summed)
    const double inv = __ddiv_rn(1., tot);
    for (uint32_t q = threadIdx.x; q < m; q += 256) {
        const double cv = cq[b + q];
        tq[b + q] = w ? __dmul_rn(__ddiv_rn(w[perm[b + q]], cv), inv) : __ddiv_rn(__ddiv_rn(1., cv), tot);
    }
}

// new center of each cluster still active: sum_i coef_i x_i, members split over 4
// groups (member q -> group q mod 4), groups added ((g0 + g1) + (g2 + g3))
__global__ __launch_bounds__(256) void k_geo_update(const float *__restrict__ Xp, uint32_t d,
                                                    const uint32_t *__restrict__ perm,
                                                    const uint32_t *__restrict__ counts,
                                                    const uint32_t *__restrict__ off, const double *__restrict__ coef,
                                                    const uint32_t *__restrict__ active, float *__restrict__ C) {
    __shared__ double part[MED_GROUPS][MED_COLS];
    const uint32_t c = blockIdx.y;
    if (!active[c]) return;
    const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
    const uint32_t j = blockIdx.x * MED_COLS + lane;
    const uint32_t m = counts[c], b = off[c];
    double a = 0.;
    if (j < d)
        for (uint32_t q = g; q < m; q += MED_GROUPS)
            a = __fma_rn(coef[b + q], (double)Xp[(uint64_t)(b + q) * d + j], a);
    part[g][lane] = a;
    __syncthreads();
    if (g == 0 && j < d)
        C[(uint64_t)c * d + j] =
            (float)__dadd_rn(__dadd_rn(part[0][lane], part[1][lane]), __dadd_rn(part[2][lane], part[3][lane]));
}

__global__ void k_count_active(const uint32_t *__restrict__ active, uint64_t k, uint32_t *__restrict__ out) {
    __shared__ uint32_t sm[256];
    uint32_t s = 0;
    for (uint64_t c = threadIdx.x; c < k; c += 256) s += active[c];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}

static void geo_median(Data &D, const float *Xp, const double *w, uint64_t k, const SortOut &so, const uint32_t *A,
                       float *C, hipStream_t st) {
    Ctx &c = *D.ctx;
    const uint64_t n = D.n;
    const uint32_t d = (uint32_t)D.d;
    double *cq = (double *)c.s[S_MED0].reserve(2 * n * sizeof(double));
    double *tq = cq + n;
    char *buf = (char *)c.s[S_MED1].reserve(64 + k * 24);
    uint32_t *flags = (uint32_t *)buf;   // [0] nan, [1] active count
    float *prev = (float *)(buf + 64);
    uint32_t *active = (uint32_t *)(buf + 64 + k * 4);
    double *Q = (double *)(buf + 64 + k * 8);
    double *prevd = (double *)(buf + 64 + k * 16);
    fill_u32(flags, 0u, 2, st);
    hipLaunchKernelGGL(k_geo_init, dim3((uint32_t)cdiv(k, 256)), dim3(256), 0, st, so.counts, k, prev, prevd, active);
    // dense rows: max(d, 1e-40) in the weighted overload, 1e-30 in the unweighted one
    // (blaze_adaptor.h:622,664); CSR rows: 1e-80 in FP64 (csc.h:1148)
    const bool csr_rule = D.csr;
    const double floor_ = csr_rule ? 1e-80 : w ? (double)1e-40f : (double)1e-30f;
    const dim3 gu((uint32_t)cdiv(d, MED_COLS), (uint32_t)k);
    uint64_t rounds = 0, batch = 4;
    uint32_t hf[2];
    while (rounds < GEO_MAX_ROUNDS) {
        const uint64_t nb = std::min<uint64_t>(batch, GEO_MAX_ROUNDS - rounds);
        for (uint64_t t = 0; t < nb; ++t) {
            hipLaunchKernelGGL(k_geo_dist, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, Xp, d, so.perm, n, A, C,
                               active, w, floor_, csr_rule ? 1 : 0, cq, tq);
            hipLaunchKernelGGL(k_geo_sums, dim3((uint32_t)k), dim3(256), 0, st, so.perm, so.counts, so.off, cq, tq, w,
                               0., csr_rule ? 1 : 0, prev, prevd, active, Q, flags);
            hipLaunchKernelGGL(k_geo_update, gu, dim3(256), 0, st, Xp, d, so.perm, so.counts, so.off, tq, active, C);
        }
        rounds += nb;
        hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, st, active, k, flags + 1);
        copy_sync(hf, flags, 8, hipMemcpyDeviceToHost, st);
        if (hf[0]) MCG_THROW(MCG_ENAN, "Optimization failed: nan");
        if (!hf[1]) break;
        batch = std::min<uint64_t>(batch * 2, 64);
    }
}

static void csr_median(Data &D, int tvd, uint64_t k, const SortOut &so, const uint32_t *A, bool single_raw,
                       float *C, hipStream_t st) {
    Ctx &c = *D.ctx;
    const uint64_t n = D.n, nnz = D.nnz;
    const uint32_t d = (uint32_t)D.d;
    const uint64_t nseg = k * d;
    REQUIRE(nseg < (1ull << 31), "k * d must be < 2^31 for CSR median centers");
    REQUIRE(nnz < (1ull << 32), "nnz must be < 2^32 for CSR median centers");
    const float *val = D.val.as<const float>();
    const uint32_t *idx = D.idx.as<const uint32_t>();
    const uint64_t *ptr = D.ptr.as<const uint64_t>();
    char *b0 = (char *)c.s[S_MED0].reserve(n * 8 + nnz * 4 + 16);
    double *inv = tvd ? (double *)b0 : nullptr;
    float *vals = (float *)(b0 + n * 8);
    uint32_t *seg = (uint32_t *)c.s[S_MED1].reserve((3 * (nseg + 1)) * sizeof(uint32_t));
    uint32_t *segoff = seg + (nseg + 1), *cursor = seg + 2 * (nseg + 1);
    if (tvd) hipLaunchKernelGGL(k_inv_rowsum_csr_seq, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, val, ptr, n, inv);
    fill_u32(seg, 0u, nseg + 1, st);
    hipLaunchKernelGGL(k_seg_count, dim3((uint32_t)cdiv(n, 4)), dim3(256), 0, st, idx, ptr, n, d, A, (uint32_t)k, seg);
    scan_exclusive(seg, segoff, nseg + 1, c.s[S_MISC1], st);
    HIP_CHECK(hipMemcpyAsync(cursor, segoff, (nseg + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_seg_fill, dim3((uint32_t)cdiv(n, 4)), dim3(256), 0, st, val, idx, ptr, n, d, A, (uint32_t)k,
                       inv, cursor, vals);
    if (single_raw)
        hipLaunchKernelGGL(k_med_single_csr, dim3((uint32_t)k), dim3(256), 0, st, val, idx, ptr, d, so.perm,
                           so.counts, so.off, C);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(cdiv(nseg, 4), 65536);
    hipLaunchKernelGGL(k_seg_median, dim3(grid), dim3(256), 0, st, vals, segoff, so.counts, d, nseg, tvd,
                       single_raw ? 2u : 1u, C);
}

// Centers from a sorted assignment.  wm: per-row weights of D; wr: the vector the
// reference indexes by member rank (TVD always, L1 when rank_l1).  single_raw: a
// one-member cluster takes its raw row (set_centroids_*); else the rule decides.
static void median_centers_core(Data &D, int pol, const double *wm, const double *wr, uint64_t k,
                                const SortOut &so, const uint32_t *A, bool single_raw, bool rank_l1, float *C,
                                hipStream_t st) {
    Ctx &c = *D.ctx;
    const uint64_t n = D.n;
    const uint32_t d = (uint32_t)D.d;
    const uint32_t min_m = single_raw ? 2u : 1u;
    const double *w = wm;
    if (D.csr && pol == 3) {
        // the geometric median of CSR rows (unweighted; csc.h:1131-1203 with the
        // member costs indexed by position): the members as dense rows in cluster
        // order, then the Weiszfeld rounds under the CSR rule
        REQUIRE(!wm, "Not implemented: weighted geometric median for CSparseMatrix");
        REQUIRE((double)n * (double)d * 4. < 64e9, "the CSR geometric median densifies the rows: n x d too large");
        hipLaunchKernelGGL(k_med_single_csr, dim3((uint32_t)k), dim3(256), 0, st, D.val.as<const float>(),
                           D.idx.as<const uint32_t>(), D.ptr.as<const uint64_t>(), d, so.perm, so.counts, so.off, C);
        float *Xp = (float *)c.s[S_MED3].reserve(n * d * sizeof(float));
        fill_u32((uint32_t *)Xp, 0u, n * d, st);
        hipLaunchKernelGGL(k_scatter_perm_csr, dim3((uint32_t)cdiv(n, 4)), dim3(256), 0, st, D.val.as<const float>(),
                           D.idx.as<const uint32_t>(), D.ptr.as<const uint64_t>(), d, so.perm, n, Xp);
        geo_median(D, Xp, nullptr, k, so, A, C, st);
        return;
    }
    if (D.csr) {
        csr_median(D, pol == 2, k, so, A, single_raw, C, st);
        return;
    }
    const dim3 g1((uint32_t)cdiv(d, 256), (uint32_t)k), gm((uint32_t)cdiv(d, MED_COLS), (uint32_t)k);
    // one member: its raw row (set_centroids_*; the geometric median of one row is
    // that row in both drivers)
    if (single_raw || pol == 3)
        hipLaunchKernelGGL(k_med_single, g1, dim3(256), 0, st, D.X, d, so.perm, so.counts, so.off, C);
    // the rows in cluster order, read contiguously by every pass below
    float *Xp = (float *)c.s[S_MED3].reserve(n * d * sizeof(float));
    hipLaunchKernelGGL(k_gather_perm, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 4), 65536)), dim3(256), 0, st, D.X,
                       d, so.perm, n, Xp);
    if (pol == 3) {
        geo_median(D, Xp, wm, k, so, A, C, st);
        return;
    }
    double *invp = nullptr;
    if (pol == 2) {
        invp = (double *)c.s[S_MED0].reserve(n * sizeof(double));
        hipLaunchKernelGGL(k_inv_rowsum_seq, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, Xp, n, d, invp);
    }
    if (w) {
        const int byrank = pol == 2 || rank_l1;
        if (pol == 2)
            hipLaunchKernelGGL(k_med_heaviest<true>, gm, dim3(256), 0, st, Xp, d, so.perm, so.counts, so.off, invp, wm,
                               wr, byrank, min_m, C);
        else
            hipLaunchKernelGGL(k_med_heaviest<false>, gm, dim3(256), 0, st, Xp, d, so.perm, so.counts, so.off, invp,
                               wm, wr, byrank, min_m, C);
    } else {
        // 16 columns and 16 member groups per block: 16 KB of histograms, so many
        // blocks share a CU and hide the member loads' latency
        const dim3 g16((uint32_t)cdiv(d, 16), (uint32_t)k);
        if (pol == 2)
            hipLaunchKernelGGL((k_med_select<true, 16>), g16, dim3(256), 0, st, Xp, d, so.counts, so.off, invp, min_m,
                               C);
        else
            hipLaunchKernelGGL((k_med_select<false, 16>), g16, dim3(256), 0, st, Xp, d, so.counts, so.off, invp, min_m,
                               C);
    }
}


void median_mstep(Data &D, int msr, const double *w, uint64_t k, uint32_t *Ar, float *pr, float *C,
                  hipStream_t st) {
    Ctx &c = *D.ctx;
    const int pol = median_policy(msr);
    REQUIRE(pol, "internal: not a median policy");
    const uint64_t n = D.n;
    const uint32_t d = (uint32_t)D.d;
    REQUIRE(n < (1ull << 32), "n must fit in uint32 for median centers");
    // empty clusters (centroid.h:442-480): each takes the row drawn over the current
    // costs (draws 0, 1, ... of the wyrand stream of n) and every row strictly closer
    // to it by raw L1 (L2 for the geometric median) moves there; until none is empty
    float *tmp = (float *)c.s[S_MED2].reserve(n * sizeof(float));
    std::vector<uint32_t> hc(k);
    uint64_t draw = 0;
    SortOut so;
    for (;;) {
        so = sort_by_cluster(c, Ar, n, k, st);
        copy_sync(hc.data(), so.counts, k * 4, hipMemcpyDeviceToHost, st);
        std::vector<uint32_t> empty;
        for (uint64_t q = 0; q < k; ++q) if (!hc[q]) empty.push_back((uint32_t)q);
        if (empty.empty()) break;
        for (const uint32_t id : empty) {
            const uint64_t row = restart_draw_device(c, pr, n, draw++, st);
            copy_rows_to_centers(D, &row, &id, 1, C, st);
            generic_costs(D, pol == 3 ? MCG_L2 : MCG_L1, 0., C + (uint64_t)id * d, 1, tmp, nullptr, nullptr, st);
            hipLaunchKernelGGL(k_take_lower1, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 256), 4096)), dim3(256), 0,
                               st, tmp, id, n, pr, Ar);
        }
    }
    median_centers_core(D, pol, w, w, k, so, Ar, true, false, C, st);
}

// The minibatch update's centers (solve.h:555-572): no restart (clusters without
// members keep their center), a one-member cluster by the rule (TVD normalises it),
// weighted L1 / TVD by the weight vector indexed by member rank (wr, the full
// vector), the geometric median by the members' weights (wm, per row of D).
void median_update(Data &D, int msr, const double *wm, const double *wr, uint64_t k, const uint32_t *A, float *C,
                   hipStream_t st) {
    const int pol = median_policy(msr);
    REQUIRE(pol, "internal: not a median policy");
    REQUIRE(D.n < (1ull << 32), "n must fit in uint32 for median centers");
    SortOut so = sort_by_cluster(*D.ctx, A, D.n, k, st);
    median_centers_core(D, pol, wm, wr, k, so, A, false, true, C, st);
}

// non-negative CSR values (the median rules above assume them); true if any is negative
bool csr_has_negative(Data &D, hipStream_t st) {
    Ctx &c = *D.ctx;
    uint32_t *flag = (uint32_t *)c.s[S_MED2].reserve(std::max<uint64_t>(D.n, 1) * sizeof(float));
    fill_u32(flag, 0u, 1, st);
    if (D.nnz)
        hipLaunchKernelGGL(k_any_negative, dim3((uint32_t)std::min<uint64_t>(cdiv(D.nnz, 256), 4096)), dim3(256), 0,
                           st, D.val.as<const float>(), D.nnz, flag);
    uint32_t h = 0;
    copy_sync(&h, flag, 4, hipMemcpyDeviceToHost, st);
    return h != 0;
}

}  // namespace mcg
