// Dense squared-L2 hard clustering of float64 rows with every value in FP64.
//
// The reference instantiates the dense path with FT = double for float64 arrays
// (python/pycluster_dense.cpp:85-90), so distances, means and the costs of the
// Lloyd control flow (clustering/solve.h:98-175) are doubles there.  This module is
// that path for SQRL2:
//  * k_assign_f64: per pair the sequential FMA chain s = fma(x_j - c_j, x_j - c_j, s)
//    over j in FP64, rows and centers staged through LDS in 16-dim slabs, 16 centers
//    per register tile; the lowest cost wins, ties to the lower index, and the first
//    center's cost sticks when it is NaN (the scan of assign_points_hard).
//  * the hard M-step on the double rows: the stable cluster sort of the FP32 M-step,
//    then per (cluster, 256-member chunk) the sequential FP64 sum of w x, chunks
//    added in order (k_partial_f64, k_finalize_f64): the FP32 path's association.
//  * the control flow compares FP64 costs; each is the exact sum of w_i cost_i at the
//    sampler's fixed-point scale, so it does not depend on summation order.  Empty
//    clusters restart as in centroid.h:687-752.
#include "mcg_internal.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace mcg {

constexpr int F64_TC = 16;   // centers per register tile
constexpr int F64_SD = 16;   // dims per LDS slab
constexpr uint32_t F64_CHUNK = 256;   // members per partial-sum chunk (the FP32 M-step's)

__global__ __launch_bounds__(256) void k_assign_f64(const double *__restrict__ X, uint64_t n, uint32_t d,
                                                    const double *__restrict__ C, uint32_t k,
                                                    uint32_t *__restrict__ asn, double *__restrict__ cost) {
    __shared__ double xs[256][F64_SD + 1];
    __shared__ double cs[F64_TC][F64_SD];
    const int tid = threadIdx.x;
    const uint64_t p0 = (uint64_t)blockIdx.x * 256, p = p0 + tid;
    double best = 0.;
    uint32_t bi = 0xFFFFFFFFu;
    for (uint32_t c0 = 0; c0 < k; c0 += F64_TC) {
        double acc[F64_TC];
#pragma unroll
        for (int c = 0; c < F64_TC; ++c) acc[c] = 0.;
        for (uint32_t j0 = 0; j0 < d; j0 += F64_SD) {
            __syncthreads();
            for (int q = tid; q < 256 * F64_SD; q += 256) {
                const int r = q / F64_SD, jj = q % F64_SD;
                const uint64_t pr = p0 + r;
                xs[r][jj] = (pr < n && j0 + jj < d) ? X[pr * d + j0 + jj] : 0.;
            }
            for (int q = tid; q < F64_TC * F64_SD; q += 256) {
                const int c = q / F64_SD, jj = q % F64_SD;
                cs[c][jj] = (c0 + c < k && j0 + jj < d) ? C[(uint64_t)(c0 + c) * d + j0 + jj] : 0.;
            }
            __syncthreads();
            const int jn = (int)min<uint32_t>(F64_SD, d - j0);
            for (int jj = 0; jj < jn; ++jj) {
                const double xv = xs[tid][jj];
#pragma unroll
                for (int c = 0; c < F64_TC; ++c) { const double t = xv - cs[c][jj]; acc[c] = fma(t, t, acc[c]); }
            }
        }
#pragma unroll
        for (int c = 0; c < F64_TC; ++c)
            if (c0 + c < k && (bi == 0xFFFFFFFFu || acc[c] < best)) { best = acc[c]; bi = c0 + c; }
    }
    if (p < n) {
        asn[p] = bi;
        cost[p] = best;
    }
}

// One wave per (cluster, chunk) item (imap[item] its cluster); lanes over dims, the
// members walked in order (four rows loaded ahead): part[item*d + j] = sum_m w x
// (fma) or sum_m x, partw[item] = sum_m w (the member count unweighted).
__global__ __launch_bounds__(64) void k_partial_f64(const double *__restrict__ X, uint32_t d,
                                                    const uint32_t *__restrict__ perm,
                                                    const uint32_t *__restrict__ counts,
                                                    const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ itemoff,
                                                    const uint32_t *__restrict__ imap, uint32_t k,
                                                    const double *__restrict__ w, double *__restrict__ part,
                                                    double *__restrict__ partw) {
    const uint32_t item = blockIdx.x;
    if (item >= itemoff[k]) return;
    const uint32_t c = imap[item];
    const uint32_t mb = off[c] + (item - itemoff[c]) * F64_CHUNK;
    const uint32_t nm = min(off[c] + counts[c], mb + F64_CHUNK) - mb;
    const uint32_t *mem = perm + mb;
    const int lane = threadIdx.x;
    double ws = 0.;
    for (uint32_t j0 = 0; j0 < d; j0 += 64) {
        const uint32_t j = j0 + lane;
        if (j >= d) break;
        double s = 0.;
        constexpr uint32_t G = 4;
        double cur[G];
#pragma unroll
        for (uint32_t g = 0; g < G; ++g) cur[g] = g < nm ? X[(uint64_t)mem[g] * d + j] : 0.;
        for (uint32_t q0 = 0; q0 < nm; q0 += G) {
            double nxt[G];
#pragma unroll
            for (uint32_t g = 0; g < G; ++g) {
                const uint32_t q = q0 + G + g;
                nxt[g] = q < nm ? X[(uint64_t)mem[q] * d + j] : 0.;
            }
#pragma unroll
            for (uint32_t g = 0; g < G; ++g) {
                if (q0 + g < nm) {
                    if (w) {
                        const double wq = w[mem[q0 + g]];
                        s = __fma_rn(wq, cur[g], s);
                        if (j0 == 0) ws = __dadd_rn(ws, wq);
                    } else {
                        s = __dadd_rn(s, cur[g]);
                    }
                }
            }
#pragma unroll
            for (uint32_t g = 0; g < G; ++g) cur[g] = nxt[g];
        }
        part[(uint64_t)item * d + j] = s;
    }
    if (lane == 0) partw[item] = w ? ws : (double)nm;
}

// One block per cluster: chunk sums added in order, mean = sum / weight (FP64);
// clusters without members keep their center.
__global__ __launch_bounds__(256) void k_finalize_f64(uint32_t d, const uint32_t *__restrict__ counts,
                                                      const uint32_t *__restrict__ itemoff,
                                                      const double *__restrict__ part,
                                                      const double *__restrict__ partw, double *__restrict__ C) {
    const uint32_t c = blockIdx.x;
    if (counts[c] == 0) return;
    const uint32_t i0 = itemoff[c], i1 = itemoff[c + 1];
    double wt = 0.;
    for (uint32_t it = i0; it < i1; ++it) wt = __dadd_rn(wt, partw[it]);
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
        double s = 0.;
        for (uint32_t it = i0; it < i1; ++it) s = __dadd_rn(s, part[(uint64_t)it * d + j]);
        C[(uint64_t)c * d + j] = __ddiv_rn(s, wt);
    }
}

__global__ void k_force_f64(uint32_t *__restrict__ asn, double *__restrict__ cost, const uint64_t *__restrict__ rows,
                            const uint32_t *__restrict__ ids, uint64_t m) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const uint64_t r = rows[q];
    if (asn[r] != ids[q]) { asn[r] = ids[q]; cost[r] = 0.; }
}

__global__ void k_nan_any_f64(const double *__restrict__ v, uint64_t m, uint32_t *__restrict__ flag) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        if (v[i] != v[i]) *flag = 1u;
}

void assign_f64(Data &ds, const double *C, uint64_t k, uint32_t *asn, double *cost, hipStream_t st) {
    REQUIRE(ds.Xd.p, "the float64 path needs rows given as float64");
    REQUIRE(k < (1ull << 32), "k must fit in uint32");
    hipLaunchKernelGGL(k_assign_f64, dim3((uint32_t)cdiv(ds.n, 256)), dim3(256), 0, st, ds.Xd.as<const double>(),
                       ds.n, (uint32_t)ds.d, C, (uint32_t)k, asn, cost);
}

// Means of the double rows by assignment into C (clusters without members keep
// theirs); returns the per-cluster counts (device).
static const uint32_t *mstep_f64(Data &ds, const uint32_t *asn, const double *w, uint64_t k, double *C,
                                 hipStream_t st) {
    Ctx &c = *ds.ctx;
    const uint64_t n = ds.n;
    const uint32_t d = (uint32_t)ds.d;
    SortOut so = sort_by_cluster(c, asn, n, k, st);
    const uint64_t max_items = n / F64_CHUNK + k + 1;
    double *part = (double *)c.s[S_PART].reserve(max_items * d * sizeof(double));
    double *partw = (double *)c.s[S_PARTW].reserve(max_items * sizeof(double));
    hipLaunchKernelGGL(k_partial_f64, dim3((uint32_t)max_items), dim3(64), 0, st, ds.Xd.as<const double>(), d,
                       so.perm, so.counts, so.off, so.itemoff, so.imap, (uint32_t)k, w, part, partw);
    hipLaunchKernelGGL(k_finalize_f64, dim3((uint32_t)k), dim3(256), 0, st, d, so.counts, so.itemoff, part, partw, C);
    return so.counts;
}

}  // namespace mcg

using namespace mcg;

extern "C" {

int mcg_lloyd_f64(mcg_data *ds, double *centers, uint64_t k, const double *weights, double eps, uint64_t maxiter,
                  uint32_t *asn, double *costs, double *initcost, double *finalcost, uint64_t *numiter) {
    return guard([&] {
        REQUIRE(ds && centers && asn && initcost && finalcost && numiter, "null argument");
        REQUIRE(k > 0, "k must be positive");
        Data &D = ds->d;
        REQUIRE(!D.csr && D.Xd.p, "mcg_lloyd_f64 needs dense rows given as float64");
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n, d = D.d, kd = k * d;
        DBuf sw, cb, ab, pb, flag;
        const double *W = (const double *)in_dev(c, weights, n * 8, sw);
        double *C = (double *)cb.reserve(2 * kd * 8), *Cc = C + kd;
        uint32_t *A = (uint32_t *)ab.reserve(2 * n * 4), *An = A + n;
        double *pc = (double *)pb.reserve(2 * n * 8), *pcn = pc + n;
        uint32_t *nanf = (uint32_t *)flag.reserve(4);
        HIP_CHECK(hipMemcpyAsync(C, centers, kd * 8, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                           : hipMemcpyHostToDevice, st));
        assign_f64(D, C, k, A, pc, st);
        double cost = fixed_wsum(c, pc, W, n, st);
        *initcost = cost;
        uint64_t iter = 0;
        if (cost == 0.) {
            *finalcost = 0.;
        } else {
            HIP_CHECK(hipMemcpyAsync(Cc, C, kd * 8, hipMemcpyDeviceToDevice, st));
            std::vector<uint32_t> hc(k);
            for (;;) {
                // means of A, with the empty-cluster restart (centroid.h:687-752): each
                // empty cluster takes the row drawn over the per-point costs (draw e of
                // the stream of n), every row is re-assigned, the drawn rows are forced
                // onto their clusters, then the means
                bool restarted = false;
                {
                    SortOut so = sort_by_cluster(c, A, n, k, st);
                    copy_sync(hc.data(), so.counts, k * 4, hipMemcpyDeviceToHost, st);
                    std::vector<uint32_t> ids;
                    for (uint64_t q = 0; q < k; ++q) if (!hc[q]) ids.push_back((uint32_t)q);
                    if (!ids.empty()) {
                        restarted = true;
                        const uint64_t ne = ids.size();
                        std::vector<double> hp(n);
                        copy_sync(hp.data(), pc, n * 8, hipMemcpyDeviceToHost, st);
                        std::vector<uint64_t> r(ne), rows(ne);
                        for (uint64_t e = 0; e < ne; ++e) r[e] = wyrand_at(n, e);
                        draw_fixed_many(hp, r.data(), ne, rows.data());
                        for (uint64_t e = 0; e < ne; ++e)
                            HIP_CHECK(hipMemcpyAsync(Cc + (uint64_t)ids[e] * d, D.Xd.as<const double>() + rows[e] * d,
                                                     d * 8, hipMemcpyDeviceToDevice, st));
                        assign_f64(D, Cc, k, A, pc, st);
                        DBuf rb;
                        char *buf = (char *)rb.reserve(ne * 12);
                        HIP_CHECK(hipMemcpyAsync(buf, rows.data(), ne * 8, hipMemcpyHostToDevice, st));
                        HIP_CHECK(hipMemcpyAsync(buf + ne * 8, ids.data(), ne * 4, hipMemcpyHostToDevice, st));
                        hipLaunchKernelGGL(k_force_f64, dim3((uint32_t)cdiv(ne, 256)), dim3(256), 0, st, A, pc,
                                           (const uint64_t *)buf, (const uint32_t *)(buf + ne * 8), ne);
                        sync(c);
                    }
                }
                mstep_f64(D, A, W, k, Cc, st);
                fill_u32(nanf, 0u, 1, st);
                hipLaunchKernelGGL(k_nan_any_f64, dim3((uint32_t)std::min<uint64_t>(cdiv(kd, 256), 1024)), dim3(256), 0,
                                   st, Cc, kd, nanf);
                uint32_t hnan = 0;
                copy_sync(&hnan, nanf, 4, hipMemcpyDeviceToHost, st);
                if (hnan) MCG_THROW(MCG_ENAN, "Unexpected nan in ctr");
                assign_f64(D, Cc, k, An, pcn, st);
                const double newcost = fixed_wsum(c, pcn, W, n, st);
                if (newcost > cost && !restarted) break;   // the accepted state (A, pc: assign(C)) stays
                HIP_CHECK(hipMemcpyAsync(C, Cc, kd * 8, hipMemcpyDeviceToDevice, st));
                std::swap(A, An);
                std::swap(pc, pcn);
                ++iter;
                const double old = cost;
                cost = newcost;
                if (old - newcost < eps * std::max(newcost, old) || iter > maxiter) break;
            }
            *finalcost = cost;
        }
        *numiter = iter;
        HIP_CHECK(hipMemcpyAsync(asn, A, n * 4, is_device_ptr(asn) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                 st));
        if (costs)
            HIP_CHECK(hipMemcpyAsync(costs, pc, n * 8, is_device_ptr(costs) ? hipMemcpyDeviceToDevice
                                                                            : hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(centers, C, kd * 8, is_device_ptr(centers) ? hipMemcpyDeviceToDevice
                                                                           : hipMemcpyDeviceToHost, st));
        sync(c);
    });
}

int mcg_assign_f64(mcg_data *ds, const double *centers, uint64_t k, uint32_t *asn, double *costs) {
    return guard([&] {
        REQUIRE(ds && centers && asn, "null argument");
        REQUIRE(k > 0, "k must be positive");
        Data &D = ds->d;
        REQUIRE(!D.csr && D.Xd.p, "mcg_assign_f64 needs dense rows given as float64");
        Ctx &c = *D.ctx;
        hipStream_t st = c.stream;
        HIP_CHECK(hipSetDevice(c.device));
        const uint64_t n = D.n;
        DBuf cb, ab, pb;
        const double *C = (const double *)in_dev(c, centers, k * D.d * 8, cb);
        uint32_t *A = (uint32_t *)ab.reserve(n * 4);
        double *pc = (double *)pb.reserve(n * 8);
        assign_f64(D, C, k, A, pc, st);
        HIP_CHECK(hipMemcpyAsync(asn, A, n * 4, is_device_ptr(asn) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                 st));
        if (costs)
            HIP_CHECK(hipMemcpyAsync(costs, pc, n * 8, is_device_ptr(costs) ? hipMemcpyDeviceToDevice
                                                                            : hipMemcpyDeviceToHost, st));
        sync(c);
    });
}

}  // extern "C"
