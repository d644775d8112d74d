// Hard M-step: weighted means of each cluster's rows, FP64, deterministic, no
// floating-point atomics.
//
// Replaces set_centroids_full_mean / set_center (include/minicore/clustering/
// centroid.h:649-779 and 81-178): the reference collects each cluster's members
// under per-center mutexes, radix-sorts them and sums in data precision
// (dense, blaze::mean) or with OMP atomics into a double buffer (CSR).
//
// Here: (1) a stable counting sort of point ids by assignment (per-block
// histograms, an exclusive scan, an in-order scatter), so each cluster's
// members are contiguous and ascending; (2) the members are cut into chunks of
// MCHUNK, each chunk summed sequentially in FP64 by one block (a thread per
// dim), and (3) chunk sums are added in order.  The association is fixed, so
// results are bit-reproducible and equal the oracle's restatement.
#include "mcg_internal.h"
#include <cfloat>

namespace mcg {

constexpr uint32_t SORT_BLK = 1024;  // points per histogram block
constexpr uint32_t MCHUNK = 256;     // members per partial-sum chunk
constexpr uint32_t MAX_SORT_K = 16384;

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st) {
    if (!n) return;
    uint32_t g = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, st, p, v, n);
}

// Points whose assignment is >= k (never produced by the E-step kernels) are
// left out of every cluster.  Block b counts its SORT_BLK points in an LDS
// histogram, then writes column b of hist (cluster-major, hist[c * nblk + b]).
__global__ __launch_bounds__(256) void k_hist(const uint32_t *__restrict__ asn, uint64_t n, uint32_t nblk,
                                              uint32_t k, uint32_t *__restrict__ hist) {
    extern __shared__ uint32_t h[];
    const uint64_t b = blockIdx.x;
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) h[c] = 0;
    __syncthreads();
    const uint64_t i0 = b * SORT_BLK;
    for (uint32_t q = threadIdx.x; q < SORT_BLK; q += blockDim.x) {
        const uint64_t i = i0 + q;
        if (i < n) {
            const uint32_t a = asn[i];
            if (a < k) atomicAdd(&h[a], 1u);
        }
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) hist[(uint64_t)c * nblk + b] = h[c];
}

// --- exclusive scan of u32 (length L), three phases ------------------------
constexpr uint32_t SCAN_PER = 16;  // elements per thread
__global__ void k_scan_reduce(const uint32_t *__restrict__ in, uint64_t L, uint32_t *__restrict__ bsum) {
    __shared__ uint32_t sm[256];
    const uint64_t base = (uint64_t)blockIdx.x * 256 * SCAN_PER;
    uint32_t s = 0;
    for (uint32_t q = 0; q < SCAN_PER; ++q) {
        uint64_t i = base + (uint64_t)q * 256 + threadIdx.x;
        if (i < L) s += in[i];
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sm[0];
}

// single block, exclusive scan in place of m values (m arbitrary); total -> *tot
__global__ void k_scan_single(uint32_t *__restrict__ v, uint64_t m, uint32_t *__restrict__ tot) {
    __shared__ uint32_t sm[1024];
    const uint64_t per = (m + 1023) / 1024;
    const uint64_t b = threadIdx.x * per, e = min<uint64_t>(m, b + per);
    uint32_t s = 0;
    for (uint64_t i = b; i < e; ++i) s += v[i];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        uint32_t t = threadIdx.x >= (unsigned)o ? sm[threadIdx.x - o] : 0u;
        __syncthreads();
        sm[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? sm[threadIdx.x - 1] : 0u;
    for (uint64_t i = b; i < e; ++i) { uint32_t t = v[i]; v[i] = run; run += t; }
    if (threadIdx.x == 1023 && tot) *tot = sm[1023];
}

__global__ void k_scan_apply(const uint32_t *__restrict__ in, uint64_t L, const uint32_t *__restrict__ boff,
                             uint32_t *__restrict__ out) {
    __shared__ uint32_t sm[256];
    const uint64_t base = (uint64_t)blockIdx.x * 256 * SCAN_PER;
    // thread t owns elements [base + t*SCAN_PER, +SCAN_PER)
    const uint64_t b = base + (uint64_t)threadIdx.x * SCAN_PER;
    uint32_t s = 0;
    for (uint32_t q = 0; q < SCAN_PER; ++q) if (b + q < L) s += in[b + q];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        uint32_t t = threadIdx.x >= (unsigned)o ? sm[threadIdx.x - o] : 0u;
        __syncthreads();
        sm[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = boff[blockIdx.x] + (threadIdx.x ? sm[threadIdx.x - 1] : 0u);
    for (uint32_t q = 0; q < SCAN_PER; ++q) {
        if (b + q < L) { uint32_t t = in[b + q]; out[b + q] = run; run += t; }
    }
}

void scan_exclusive(const uint32_t *in, uint32_t *out, uint64_t L, DBuf &tmp, hipStream_t st) {
    const uint64_t per_blk = 256ull * SCAN_PER;
    const uint64_t nb = cdiv(L, per_blk);
    tmp.reserve((nb + 1) * sizeof(uint32_t));
    hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)nb), dim3(256), 0, st, in, L, tmp.as<uint32_t>());
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, st, tmp.as<uint32_t>(), nb, (uint32_t *)nullptr);
    hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)nb), dim3(256), 0, st, in, L, tmp.as<const uint32_t>(), out);
}

// One block per cluster c: base[c][b] = the exclusive prefix over the histogram blocks
// b of hist[c][b] (the position of block b's first member inside the cluster), and
// counts[c] = the cluster's total.
__global__ __launch_bounds__(256) void k_colscan(const uint32_t *__restrict__ hist, uint32_t nblk,
                                                 uint32_t *__restrict__ base, uint32_t *__restrict__ counts) {
    __shared__ uint32_t sm[256];
    const uint32_t c = blockIdx.x;
    const uint64_t r = (uint64_t)c * nblk;
    const uint32_t per = (nblk + 255) / 256;
    const uint32_t b0 = threadIdx.x * per, b1 = min(nblk, b0 + per);
    uint32_t s = 0;
    for (uint32_t b = b0; b < b1; ++b) s += hist[r + b];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t t = threadIdx.x >= (unsigned)o ? sm[threadIdx.x - o] : 0u;
        __syncthreads();
        sm[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? sm[threadIdx.x - 1] : 0u;
    for (uint32_t b = b0; b < b1; ++b) { const uint32_t t = hist[r + b]; base[r + b] = run; run += t; }
    if (threadIdx.x == 255) counts[c] = sm[255];
}

// One block of 1024 threads: off[c] = exclusive prefix of the cluster counts;
// itemoff = exclusive prefix of the partial-sum chunks ceil(counts / MCHUNK) (k + 1
// entries, the last the total); imap[i] = the cluster of chunk item i.
__global__ __launch_bounds__(1024) void k_counts_items(uint32_t k, const uint32_t *__restrict__ counts,
                                                       uint32_t *__restrict__ off, uint32_t *__restrict__ itemoff,
                                                       uint32_t *__restrict__ imap) {
    __shared__ uint32_t sm[1024], sn[1024];
    const uint32_t per = (k + 1023) / 1024;
    const uint32_t c0 = threadIdx.x * per, c1 = min(k, c0 + per);
    uint32_t s = 0, sc = 0;
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t cnt = counts[c];
        sc += cnt;
        s += (cnt + MCHUNK - 1) / MCHUNK;
    }
    sm[threadIdx.x] = s;
    sn[threadIdx.x] = sc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t t = threadIdx.x >= (unsigned)o ? sm[threadIdx.x - o] : 0u;
        const uint32_t tn = threadIdx.x >= (unsigned)o ? sn[threadIdx.x - o] : 0u;
        __syncthreads();
        sm[threadIdx.x] += t;
        sn[threadIdx.x] += tn;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? sm[threadIdx.x - 1] : 0u;
    uint32_t rn = threadIdx.x ? sn[threadIdx.x - 1] : 0u;
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t cnt = counts[c];
        const uint32_t it = (cnt + MCHUNK - 1) / MCHUNK;
        off[c] = rn;
        rn += cnt;
        itemoff[c] = run;
        for (uint32_t i = 0; i < it; ++i) imap[run + i] = c;
        run += it;
    }
    if (threadIdx.x == 1023) itemoff[k] = sm[1023];
}

// Stable scatter: one wave per histogram block, LDS cursors for all clusters (the
// cluster's offset plus the block's position inside the cluster).  The block's
// assignments are loaded up front (16 per lane) and the cursors set from them, so the
// 16 ranking rounds run without memory waits.
__global__ __launch_bounds__(64) void k_scatter(const uint32_t *__restrict__ asn, uint64_t n, uint32_t nblk,
                                                uint32_t k, uint32_t kbits, const uint32_t *__restrict__ off,
                                                const uint32_t *__restrict__ base, uint32_t *__restrict__ perm) {
    extern __shared__ uint32_t cur[];
    constexpr int R = SORT_BLK / 64;
    const uint32_t b = blockIdx.x;
    const int lane = threadIdx.x;
    const uint64_t i0 = (uint64_t)b * SORT_BLK;
    uint32_t av[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint64_t i = i0 + (uint64_t)r * 64 + lane;
        av[r] = i < n ? asn[i] : 0xFFFFFFFFu;
    }
    // lanes with the same cluster write the same cursor
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (av[r] < k) cur[av[r]] = off[av[r]] + base[(uint64_t)av[r] * nblk + b];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint64_t i = i0 + (uint64_t)r * 64 + lane;
        const uint32_t a = av[r];
        const bool valid = a < k;
        const uint32_t key = valid ? a : 0xFFFFFFFFu;
        // lanes holding the same key: intersect the ballots of each key bit (and of
        // validity); rank = such lanes below this one, tot = all of them
        uint64_t same = valid ? __ballot(valid) : ~__ballot(valid);
        for (uint32_t bit = 0; bit < kbits; ++bit) {
            const bool on = (key >> bit) & 1u;
            const uint64_t m = __ballot(on);
            same &= on ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__builtin_popcountll(same & ((1ull << lane) - 1ull));
        const uint32_t tot = (uint32_t)__builtin_popcountll(same);
        uint32_t pos = 0;
        if (valid) { pos = cur[key] + rank; perm[pos] = (uint32_t)i; }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (valid && rank + 1 == tot) cur[key] = pos + 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// Partial sums: one wave per (cluster, chunk) item; imap[item] is its cluster.
// Lane l owns dims [4l, 4l+4) of each 256-dim pass (16-byte loads when d % 4 == 0,
// else one dim per lane).  Members are walked in order, so each dim's FP64 sum is
// sequential: part[item*d + j] = sum_m w*x (fma) or sum_m x;  partw[item] = sum_m w
// (added in the same member order by every lane; the member count when unweighted);
// partq[item] = sum_m w * ||x - cold||^2 (FP64) when cold != nullptr.
template <bool VEC4>
__global__ __launch_bounds__(64) void k_partial(const float *__restrict__ X, uint32_t d,
                                                const uint32_t *__restrict__ perm,
                                                const uint32_t *__restrict__ counts,
                                                const uint32_t *__restrict__ off,
                                                const uint32_t *__restrict__ itemoff,
                                                const uint32_t *__restrict__ imap, uint32_t k,
                                                const double *__restrict__ w, const float *__restrict__ cold,
                                                double *__restrict__ part, double *__restrict__ partw,
                                                double *__restrict__ partq) {
    __shared__ uint32_t mem[MCHUNK];
    __shared__ double wl[MCHUNK];
    const uint32_t item = blockIdx.x;
    const uint32_t total_items = itemoff[k];
    if (item >= total_items) return;
    const uint32_t c = imap[item];
    const uint32_t mb = off[c] + (item - itemoff[c]) * MCHUNK;
    const uint32_t me = min(off[c] + counts[c], mb + MCHUNK);
    const uint32_t nm = me - mb;
    const int lane = threadIdx.x;
    for (uint32_t q = lane; q < nm; q += 64) {
        const uint32_t pnt = perm[mb + q];
        mem[q] = pnt;
        wl[q] = w ? w[pnt] : 1.;
    }
    __syncthreads();
    double qsum = 0., wsum = 0.;
    constexpr uint32_t DPL = VEC4 ? 4 : 1;   // dims per lane
    for (uint32_t j0 = 0; j0 < d; j0 += 64 * DPL) {
        const uint32_t j = j0 + lane * DPL;
        if (j >= d) break;
        double s[DPL], cj[DPL];
#pragma unroll
        for (uint32_t i = 0; i < DPL; ++i) {
            s[i] = 0.;
            cj[i] = cold ? (double)cold[(uint64_t)c * d + j + i] : 0.;
        }
        auto accumulate = [&](const float (&xv)[DPL], double wq) __attribute__((always_inline)) {
            if (j0 == 0 && w) wsum = __dadd_rn(wsum, wq);
#pragma unroll
            for (uint32_t i = 0; i < DPL; ++i) {
                const double x = (double)xv[i];
                s[i] = w ? __fma_rn(wq, x, s[i]) : __dadd_rn(s[i], x);
                if (cold) { const double t = x - cj[i]; qsum += w ? wq * (t * t) : t * t; }
            }
        };
        if constexpr (VEC4) {
            // software pipeline: the next 4 member rows load while the current 4 are
            // added (in member order, so the sums are unchanged)
            constexpr uint32_t G = 4;
            float4 cur[G];
#pragma unroll
            for (uint32_t g = 0; g < G; ++g)
                cur[g] = g < nm ? *(const float4 *)(X + (uint64_t)mem[g] * d + j) : make_float4(0.f, 0.f, 0.f, 0.f);
            for (uint32_t q0 = 0; q0 < nm; q0 += G) {
                float4 nxt[G];
#pragma unroll
                for (uint32_t g = 0; g < G; ++g) {
                    const uint32_t q = q0 + G + g;
                    nxt[g] = q < nm ? *(const float4 *)(X + (uint64_t)mem[q] * d + j) : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (uint32_t g = 0; g < G; ++g) {
                    if (q0 + g < nm) {
                        const float xv[DPL] = {cur[g].x, cur[g].y, cur[g].z, cur[g].w};
                        accumulate(xv, wl[q0 + g]);
                    }
                }
#pragma unroll
                for (uint32_t g = 0; g < G; ++g) cur[g] = nxt[g];
            }
        } else {
#pragma unroll 4
            for (uint32_t q = 0; q < nm; ++q) {
                const float xv[DPL] = {X[(uint64_t)mem[q] * d + j]};
                accumulate(xv, wl[q]);
            }
        }
#pragma unroll
        for (uint32_t i = 0; i < DPL; ++i) part[(uint64_t)item * d + j + i] = s[i];
    }
    if (cold) {
        for (int o = 32; o > 0; o >>= 1) qsum += __shfl_xor(qsum, o);
        if (lane == 0) partq[item] = qsum;
    }
    // every lane added the same weights in member order (lanes past d skipped the
    // member loop, so lane 0 writes)
    if (lane == 0) partw[item] = w ? wsum : (double)nm;
}


// One block per cluster: add chunk partials in order; mean = sum / weight.
__global__ __launch_bounds__(256) void k_finalize(uint32_t d, const uint32_t *__restrict__ counts,
                                                  const uint32_t *__restrict__ itemoff,
                                                  const double *__restrict__ part,
                                                  const double *__restrict__ partw,
                                                  const double *__restrict__ partq, const float *__restrict__ cold,
                                                  float *__restrict__ cnew, double *__restrict__ ccost,
                                                  double *__restrict__ cw) {
    const uint32_t c = blockIdx.x;
    const uint32_t i0 = itemoff[c], i1 = itemoff[c + 1];
    // the chunks' values are loaded eight at a time, then added in chunk order
    constexpr uint32_t B = 8;
    auto ordered_sum = [&](const double *__restrict__ v, uint64_t stride, uint64_t o) {
        double s = 0.;
        for (uint32_t it = i0; it < i1; it += B) {
            double t[B];
#pragma unroll
            for (uint32_t u = 0; u < B; ++u) t[u] = it + u < i1 ? v[(uint64_t)(it + u) * stride + o] : 0.;
#pragma unroll
            for (uint32_t u = 0; u < B; ++u)
                if (it + u < i1) s = __dadd_rn(s, t[u]);
        }
        return s;
    };
    const double wt = ordered_sum(partw, 1, 0);
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
        if (counts[c] == 0) {
            if (cold && cnew != cold) cnew[(uint64_t)c * d + j] = cold[(uint64_t)c * d + j];
            continue;
        }
        const double s = ordered_sum(part, d, j);
        cnew[(uint64_t)c * d + j] = (float)__ddiv_rn(s, wt);
    }
    if (threadIdx.x == 0) {
        ccost[c] = partq ? ordered_sum(partq, 1, 0) : 0.;
        cw[c] = wt;
    }
}

// Single block: total[0] = sum_c ccost[c], total[1] = sum_c cw[c] (fixed order),
// total[2] = clusters without members; host_out (pinned host memory, or null)
// receives {total[0], total[2]}.
__global__ void k_total(const double *__restrict__ ccost, const double *__restrict__ cw,
                        const uint32_t *__restrict__ counts, uint64_t k, double *__restrict__ total,
                        double *__restrict__ host_out) {
    __shared__ double a[256], b[256];
    __shared__ uint32_t z[256];
    double s = 0., t = 0.;
    uint32_t e = 0;
    for (uint64_t i = threadIdx.x; i < k; i += blockDim.x) { s += ccost[i]; t += cw[i]; e += counts[i] == 0; }
    a[threadIdx.x] = s;
    b[threadIdx.x] = t;
    z[threadIdx.x] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            a[threadIdx.x] += a[threadIdx.x + o];
            b[threadIdx.x] += b[threadIdx.x + o];
            z[threadIdx.x] += z[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        total[0] = a[0];
        total[1] = b[0];
        total[2] = (double)z[0];
        if (host_out) { host_out[0] = a[0]; host_out[1] = (double)z[0]; }
    }
}

SortOut sort_by_cluster(Ctx &c, const uint32_t *asn, uint64_t n, uint64_t k, hipStream_t st) {
    REQUIRE(k <= MAX_SORT_K, "k must be <= 16384 for the GPU M-step");
    REQUIRE(n < (1ull << 32) - 1, "n must fit in uint32");
    const uint32_t nblk = (uint32_t)cdiv(n, SORT_BLK);
    const uint64_t L = (uint64_t)k * nblk;
    uint32_t *hist = (uint32_t *)c.s[S_HIST].reserve(L * sizeof(uint32_t));
    uint32_t *base = (uint32_t *)c.s[S_BASE].reserve(L * sizeof(uint32_t));
    uint32_t *perm = (uint32_t *)c.s[S_PERM].reserve(n * sizeof(uint32_t));
    uint32_t *cnt = (uint32_t *)c.s[S_CNT].reserve(3 * (k + 1) * sizeof(uint32_t));
    uint32_t *off = cnt + (k + 1), *items = cnt + 2 * (k + 1);
    hipLaunchKernelGGL(k_hist, dim3(nblk), dim3(256), (size_t)k * sizeof(uint32_t), st, asn, n, nblk, (uint32_t)k,
                       hist);
    hipLaunchKernelGGL(k_colscan, dim3((uint32_t)k), dim3(256), 0, st, (const uint32_t *)hist, nblk, base, cnt);
    uint32_t *imap = (uint32_t *)c.s[S_ITEMOFF].reserve((n / MCHUNK + k + 1) * sizeof(uint32_t));
    hipLaunchKernelGGL(k_counts_items, dim3(1), dim3(1024), 0, st, (uint32_t)k, (const uint32_t *)cnt, off, items,
                       imap);
    uint32_t kbits = 1;   // bits of the largest cluster id
    while ((1ull << kbits) < k) ++kbits;
    hipLaunchKernelGGL(k_scatter, dim3(nblk), dim3(64), (size_t)k * sizeof(uint32_t), st, asn, n, nblk,
                       (uint32_t)k, kbits, (const uint32_t *)off, (const uint32_t *)base, perm);
    return SortOut{perm, cnt, off, items, imap};
}

MStepOut dense_mstep(Data &ds, const uint32_t *asn, const double *w, uint64_t k, const float *Cold,
                     float *Cnew, bool want_cost, hipStream_t st, double *host_out) {
    Ctx &c = *ds.ctx;
    const uint64_t n = ds.n;
    const uint32_t d = (uint32_t)ds.d;
    SortOut so = sort_by_cluster(c, asn, n, k, st);
    const uint64_t max_items = n / MCHUNK + k + 1;
    double *part = (double *)c.s[S_PART].reserve(max_items * d * sizeof(double));
    double *partw = (double *)c.s[S_PARTW].reserve(max_items * sizeof(double));
    double *partq = (double *)c.s[S_PARTQ].reserve(max_items * sizeof(double));
    double *ccost = (double *)c.s[S_CCOST].reserve(2 * k * sizeof(double));
    double *total = (double *)c.s[S_TOTAL].reserve(4 * sizeof(double));
    const uint32_t *imap = so.imap;
    hipEvent_t ev;
    c.tic(1, ev);
    if (d % 4 == 0)
        hipLaunchKernelGGL(k_partial<true>, dim3((uint32_t)max_items), dim3(64), 0, st, ds.X, d, so.perm, so.counts,
                           so.off, so.itemoff, imap, (uint32_t)k, w, want_cost ? Cold : nullptr, part, partw, partq);
    else
        hipLaunchKernelGGL(k_partial<false>, dim3((uint32_t)max_items), dim3(64), 0, st, ds.X, d, so.perm, so.counts,
                           so.off, so.itemoff, imap, (uint32_t)k, w, want_cost ? Cold : nullptr, part, partw, partq);
    c.toc(1, ev);
    hipLaunchKernelGGL(k_finalize, dim3((uint32_t)k), dim3(256), 0, st, d, so.counts, so.itemoff, part, partw,
                       want_cost ? partq : nullptr, Cold, Cnew, ccost, ccost + k);
    hipLaunchKernelGGL(k_total, dim3(1), dim3(256), 0, st, ccost, ccost + k, so.counts, k, total, host_out);
    return MStepOut{so.counts, total};
}

// ---------------------------------------------------------------------------
// CSR M-step.  After the same sort, one block per (cluster, column window of
// CSR_CW columns, chunk of the cluster's members); 4 waves each own an LDS FP64
// accumulator and take members in groups of 64 in turn; a wave walks its rows in
// order (row entries have distinct columns, so lanes never collide), the 4
// accumulators are added in a fixed order, and the chunks in chunk order.  A row's
// entries in the block's window come from the dataset's window starts (ws), else
// from a binary search of the row.
// ---------------------------------------------------------------------------
constexpr uint32_t CSR_CW = 1024;

// Window starts of every CSR row (column indices ascending within a row): ws[r][w] =
// the number of the row's entries with column < w * CSR_CW, for w = 0..nw.  One wave
// per row; the lane holding an entry writes the starts of the windows between the
// previous entry's window and its own, the first lane past the row's end the rest.
__global__ __launch_bounds__(256) void k_csr_wstart(const uint32_t *__restrict__ idx, const uint64_t *__restrict__ ptr,
                                                    uint64_t n, uint32_t nw, uint32_t *__restrict__ ws) {
    const uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= n) return;
    const uint64_t p0 = ptr[r], len = ptr[r + 1] - p0;
    uint32_t *o = ws + r * (nw + 1);
    for (uint64_t q = lane; q <= len; q += 64) {
        const uint32_t wq = q < len ? idx[p0 + q] / CSR_CW : nw;
        const uint32_t wp = q ? idx[p0 + q - 1] / CSR_CW + 1 : 0;
        for (uint32_t w = wp; w <= wq; ++w) o[w] = (uint32_t)q;
    }
}

__global__ __launch_bounds__(256) void k_csr_mean(const float *__restrict__ val, const uint32_t *__restrict__ idx,
                                                  const uint64_t *__restrict__ ptr, uint32_t d,
                                                  const uint32_t *__restrict__ perm,
                                                  const uint32_t *__restrict__ counts,
                                                  const uint32_t *__restrict__ off, const double *__restrict__ w,
                                                  uint32_t k, double *__restrict__ part,
                                                  const double *__restrict__ cw, float *__restrict__ cnew,
                                                  const uint32_t *__restrict__ ws) {
    __shared__ double acc[4][CSR_CW];
    const uint32_t c = blockIdx.y, z = blockIdx.z, Z = gridDim.z;
    const uint32_t col0 = blockIdx.x * CSR_CW;
    const uint32_t col1 = min(d, col0 + CSR_CW);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t j = tid; j < 4 * CSR_CW; j += 256) (&acc[0][0])[j] = 0.;
    __syncthreads();
    // member chunk z of the cluster's members (contiguous, in sorted order)
    const uint32_t cm = counts[c], chunk = (cm + Z - 1) / Z;
    const uint32_t q0 = min(cm, z * chunk), q1 = min(cm, q0 + chunk);
    const uint32_t m0 = off[c] + q0, nm = q1 - q0;
    for (uint32_t g = wave * 64; g < nm; g += 256) {
        // lanes find the first entry >= col0 of 64 member rows at once
        const uint32_t q = g + lane;
        uint64_t rb = 0, re = 0;
        double wq = 1.;
        if (q < nm) {
            const uint32_t r = perm[m0 + q];
            if (ws) {   // the window's entries [rb, re) from the row's window starts
                const uint64_t p0 = ptr[r];
                const uint32_t *o = ws + (uint64_t)r * (gridDim.x + 1) + blockIdx.x;
                rb = p0 + o[0];
                re = p0 + o[1];
            } else {
                uint64_t lo = ptr[r], hi = ptr[r + 1];
                re = hi;
                while (lo < hi) { uint64_t mid = (lo + hi) >> 1; if (idx[mid] < col0) lo = mid + 1; else hi = mid; }
                rb = lo;
            }
            if (w) wq = w[r];
        }
        const uint32_t nrow = min(64u, nm - g);
        // rows in batches of RB: the first 64 window entries of every row of a batch
        // are loaded before any is added; then the rows are added in order, each in
        // full (a row's entries past the first 64 before the next row)
        constexpr uint32_t RB = 4;
        for (uint32_t t0 = 0; t0 < nrow; t0 += RB) {
            uint64_t bb[RB], ee[RB];
            double ww[RB];
            uint32_t jv[RB];
            float vv[RB];
#pragma unroll
            for (uint32_t u = 0; u < RB; ++u) {
                const uint32_t t = t0 + u < nrow ? t0 + u : t0;
                bb[u] = bcast_u64(rb, (int)t);
                ee[u] = t0 + u < nrow ? bcast_u64(re, (int)t) : bb[u];
                ww[u] = bcast_f64(wq, (int)t);
                const uint64_t p = bb[u] + lane;
                jv[u] = p < ee[u] ? idx[p] : 0xFFFFFFFFu;
                vv[u] = p < ee[u] ? val[p] : 0.f;
            }
#pragma unroll
            for (uint32_t u = 0; u < RB; ++u) {
                if (jv[u] < col1) {
                    double *a = &acc[wave][jv[u] - col0];
                    *a = __dadd_rn(*a, __dmul_rn(ww[u], (double)vv[u]));
                    for (uint64_t p = bb[u] + 64 + lane; p < ee[u]; p += 64) {
                        const uint32_t j = idx[p];
                        if (j >= col1) break;
                        double *a2 = &acc[wave][j - col0];
                        *a2 = __dadd_rn(*a2, __dmul_rn(ww[u], (double)val[p]));
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
        }
    }
    __syncthreads();
    // one chunk (cnew given): the sum is the center's numerator, divided here as the
    // final pass would (empty clusters are left as they are)
    if (cnew) {
        if (cm == 0) return;
        const double wc = cw[c];
        for (uint32_t j = col0 + tid; j < col1; j += 256) {
            const uint32_t jj = j - col0;
            const double s = __dadd_rn(__dadd_rn(acc[0][jj], acc[1][jj]), __dadd_rn(acc[2][jj], acc[3][jj]));
            cnew[(uint64_t)c * d + j] = (float)__ddiv_rn(s, wc);
        }
        return;
    }
    double *pz = part + ((uint64_t)z * k + c) * d;
    for (uint32_t j = col0 + tid; j < col1; j += 256) {
        const uint32_t jj = j - col0;
        pz[j] = __dadd_rn(__dadd_rn(acc[0][jj], acc[1][jj]), __dadd_rn(acc[2][jj], acc[3][jj]));
    }
}

// the chunks' partial sums added in chunk order, divided by the cluster's weight
// (empty clusters are left as they are)
__global__ void k_csr_mean_final(const double *__restrict__ part, uint32_t Z, uint32_t k, uint32_t d,
                                 const uint32_t *__restrict__ counts, const double *__restrict__ cw,
                                 float *__restrict__ cnew) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)k * d) return;
    const uint32_t c = (uint32_t)(e / d);
    if (counts[c] == 0) return;
    double s = part[e];
    for (uint32_t z = 1; z < Z; ++z) s = __dadd_rn(s, part[(uint64_t)z * k * d + e]);
    cnew[e] = (float)__ddiv_rn(s, cw[c]);
}

// per-cluster weight -> cw[c]: one block per cluster, 256 strided partial sums
// added as a binary tree (a fixed order)
__global__ __launch_bounds__(256) void k_cluster_weight(const uint32_t *__restrict__ perm,
                                                        const uint32_t *__restrict__ counts,
                                                        const uint32_t *__restrict__ off,
                                                        const double *__restrict__ w, uint64_t k,
                                                        double *__restrict__ cw) {
    __shared__ double sm[256];
    const uint32_t c = blockIdx.x;
    const uint32_t m0 = off[c], nm = counts[c];
    double s = 0.;
    for (uint32_t q = threadIdx.x; q < nm; q += 256) s = __dadd_rn(s, w ? w[perm[m0 + q]] : 1.);
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] = __dadd_rn(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) cw[c] = sm[0];
}

MStepOut csr_mstep(Data &ds, const uint32_t *asn, const double *w, uint64_t k, float *Cnew, hipStream_t st) {
    Ctx &c = *ds.ctx;
    SortOut so = sort_by_cluster(c, asn, ds.n, k, st);
    double *ccost = (double *)c.s[S_CCOST].reserve(2 * k * sizeof(double));
    double *total = (double *)c.s[S_TOTAL].reserve(4 * sizeof(double));
    hipLaunchKernelGGL(k_cluster_weight, dim3((uint32_t)k), dim3(256), 0, st, so.perm, so.counts, so.off, w, k,
                       ccost + k);
    // up to 8 member chunks per cluster, so a large cluster is spread over blocks
    // (partial sums: Z x k x d FP64, at most about 2 GB); one chunk writes the centers
    // directly
    const uint64_t kd = k * ds.d;
    const uint32_t Z = csr_chunks(kd);
    const uint32_t nw = (uint32_t)cdiv(ds.d, CSR_CW);
    // MCG_CSR_WSTART=0 (read per call) keeps the per-block binary search instead
    const char *wse = getenv("MCG_CSR_WSTART");
    const bool use_ws = !(wse && wse[0] == '0');
    const uint64_t wsbytes = ds.n * (nw + 1) * sizeof(uint32_t);
    const uint32_t *ws = nullptr;
    if (use_ws && wsbytes <= (1ull << 30)) {   // window starts of the rows as they are now
        uint32_t *wsb = (uint32_t *)c.s[S_WSTART].reserve(wsbytes);
        hipLaunchKernelGGL(k_csr_wstart, dim3((uint32_t)cdiv(ds.n, 4)), dim3(256), 0, st, ds.idx.as<const uint32_t>(),
                           ds.ptr.as<const uint64_t>(), ds.n, nw, wsb);
        ws = wsb;
    }
    double *part = Z > 1 ? (double *)c.s[S_PART].reserve((uint64_t)Z * kd * sizeof(double)) : nullptr;
    dim3 g((uint32_t)cdiv(ds.d, CSR_CW), (uint32_t)k, Z);
    hipLaunchKernelGGL(k_csr_mean, g, dim3(256), 0, st, ds.val.as<const float>(), ds.idx.as<const uint32_t>(),
                       ds.ptr.as<const uint64_t>(), (uint32_t)ds.d, so.perm, so.counts, so.off, w, (uint32_t)k, part,
                       (const double *)(ccost + k), Z > 1 ? nullptr : Cnew, ws);
    if (Z > 1)
        hipLaunchKernelGGL(k_csr_mean_final, dim3((uint32_t)cdiv(kd, 256)), dim3(256), 0, st, (const double *)part, Z,
                           (uint32_t)k, (uint32_t)ds.d, so.counts, (const double *)(ccost + k), Cnew);
    return MStepOut{so.counts, total};
}

}  // namespace mcg
