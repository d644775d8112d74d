// Point-to-center costs for every measure on the hot path, dense or CSR rows
// against dense centers, FP32 arithmetic with FP64 per-center/per-row terms.
//
// Replaces msr_with_prior (include/minicore/dist/applicator.h:1632-1915) in the
// n x k loops of assign_points_hard (solve.h:234-285), the soft cost matrix
// (solve.h:368-371, python/pycluster_soft.h:38-40) and cmp (python/pycmp.cpp).
//
// With x~ = (x+pv)/L, L = sum(x)+d*pv, li = pv/L (the value of x~ off the row's
// support) and c~ = (c+pv)/M, M = sum(c)+d*pv, each cost splits into per-row
// terms, per-center terms and a sum over the row's stored entries only:
//   SQRL2       = ||c||^2 + sum_supp x (x - 2c)
//   MKL         = H_c - S_c (log pv - log L) - sum_supp c~ log1p(x/pv)
//                 (H_c = sum c~ log c~, S_c = sum c~)
//   REVERSE_MKL = H_x - li * sum_j log c~_j - (1/L) sum_supp x log c~
//   JSD         = G_c(li) + sum_supp [h(x~, c~) - h(li, c~)]
//                 (h = the per-coordinate JSD term, G_c(l) = sum_j h(l, c~_j))
// pv is the prior when prior >= 1e-20 (then applicator.h:1644-1648 always
// returns it); a smaller prior takes the per-pair kernel k_pairs_direct, which
// follows the reference formula coordinate by coordinate.
#include "mcg_internal.h"
#include <algorithm>
#include <cfloat>
#include <cmath>

namespace mcg {

constexpr double DIRECT_PRIOR = 1e-20;

// ---- center tables ----------------------------------------------------------
// One block per center.  CT is d x kpad (transposed so lanes over centers read
// contiguous memory); cs = {a, b} per-center scalars (FP64).
__global__ __launch_bounds__(256) void k_center_tables(int msr, const float *__restrict__ C, uint32_t k,
                                                       uint32_t kpad, uint32_t d, double pv, float *__restrict__ CT,
                                                       double *__restrict__ cs) {
    __shared__ double red[2][256];
    const uint32_t c = blockIdx.x;
    center_table_body<256>(msr, C + (uint64_t)c * d, c, kpad, d, pv, CT, cs, red);
}

// per-row terms: {L, log L, H_x, li}
__global__ void k_row_terms(int msr, const float *__restrict__ X, const float *__restrict__ val,
                            const uint64_t *__restrict__ ptr, uint64_t n, uint32_t d, const double *__restrict__ rsum,
                            double pv, float4 *__restrict__ out) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double L = rsum[row] + (double)d * pv, Li = 1. / L, li = pv * Li;
    double h = 0.;
    uint64_t cnt = 0;
    if (msr == MCG_REVERSE_MKL || msr == MCG_JSD) {
        if (val) {
            for (uint64_t p = ptr[row] + lane; p < ptr[row + 1]; p += 64) {
                const double xt = ((double)val[p] + pv) * Li;
                h += xt * log(xt);
                ++cnt;
            }
        } else {
            for (uint32_t j = lane; j < d; j += 64) {
                const double xt = ((double)X[row * d + j] + pv) * Li;
                h += xt * log(xt);
                ++cnt;
            }
        }
        for (int o = 32; o > 0; o >>= 1) { h += __shfl_xor(h, o); cnt += __shfl_xor(cnt, o); }
        h += (double)(d - cnt) * li * log(li);
    }
    if (lane == 0) out[row] = make_float4((float)L, (float)log(L), (float)h, (float)li);
}

// h(a, b): one coordinate of JSD, 1/2 [a log(2a/(a+b)) + b log(2b/(a+b))]
__device__ __forceinline__ float jsd_h(float a, float b) {
    const float m = a + b;
    return 0.5f * (a * __logf(2.f * a / m) + b * __logf(2.f * b / m));
}

// G[u][c] = sum_j h(li_u, c~_j), FP64 accumulation.  One block per (u, c), blocks
// striding over the pairs.
__global__ __launch_bounds__(256) void k_jsd_g(const float *__restrict__ CT, uint32_t kpad, uint32_t k, uint32_t d,
                                               const float *__restrict__ lis, uint32_t U, double *__restrict__ G) {
    __shared__ double ws[4];
    for (uint64_t pid = blockIdx.x; pid < (uint64_t)U * k; pid += gridDim.x) {
        const uint32_t u = (uint32_t)(pid / k), c = (uint32_t)(pid % k);
        const double g = jsd_g_body<256>(CT, kpad, c, d, (double)lis[u], ws);
        if (threadIdx.x == 0) G[(uint64_t)u * kpad + c] = g;
    }
}

// JSD with few distinct li (rows of nearly equal sums, e.g. normalised rows): the
// zero-baseline term of each (li, column, center), m0 log(m0 / 2) with m0 = li + c~_j,
// is the same for every row of that li, so it is tabulated beside c~_j, one table per
// li (one logarithm per entry and center instead of two).
constexpr uint32_t JSD_TAB_U = 4;
__global__ void k_jsd_pair_table(const float *__restrict__ CT, uint64_t nel, const float *__restrict__ lis,
                                 uint32_t U, float2 *__restrict__ CT2) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nel * U; e += (uint64_t)gridDim.x * blockDim.x) {
        const float li = lis[e / nel];
        const float ct = CT[e % nel];
        const float m0 = li + ct;
        const float l0 = __logf(0.5f * m0);
        CT2[e] = make_float2(ct, m0 * l0);
    }
}

__device__ __forceinline__ float post_kl(float r) {
    r = r < 0.f ? 0.f : r;
    if (r == INFINITY) r = FLT_MAX;
    return r;
}

// One wave per row; lanes over centers in chunks of 64; row entries in batches
// of 64 broadcast from their lanes (scalar reads; LDS permutes for JSD).
template <int MSR>
__global__ __launch_bounds__(256) void k_row_costs(const float *__restrict__ X, const float *__restrict__ val,
                                                   const uint32_t *__restrict__ idx, const uint64_t *__restrict__ ptr,
                                                   uint64_t n, uint32_t d, float pv, float logpv,
                                                   const float4 *__restrict__ rt, const float *__restrict__ CT,
                                                   uint32_t kpad, uint32_t k, const double *__restrict__ cs,
                                                   const double *__restrict__ G, const uint32_t *__restrict__ slot,
                                                   const float2 *__restrict__ CT2, float *__restrict__ out,
                                                   uint32_t *__restrict__ asn, float *__restrict__ best_out) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    uint64_t b, e;
    if (val) { b = ptr[row]; e = ptr[row + 1]; }
    else { b = row * d; e = b + d; }
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
    if (MSR != MCG_SQRL2) r = rt[row];
    const float L = r.x, logL = r.y, Hx = r.z, li = r.w;
    const float Li = 1.f / L;
    const float hli = li * __logf(li);
    float best = 0.f;
    uint32_t bi = 0xFFFFFFFFu;
    for (uint32_t c0 = 0; c0 < k; c0 += 64) {
        const uint32_t c = c0 + lane;
        const uint32_t cc = c < kpad ? c : kpad - 1;
        float acc = 0.f;
        for (uint64_t p0 = b; p0 < e; p0 += 64) {
            const uint64_t p = p0 + lane;
            uint32_t jq = 0;
            float tq = 0.f, tq2 = 0.f;
            if (p < e) {
                const float x = val ? val[p] : X[p];
                jq = val ? idx[p] : (uint32_t)(p - b);
                if (MSR == MCG_SQRL2) { tq = x; }
                else if (MSR == MCG_MKL) { tq = log1pf(x / pv); }
                else if (MSR == MCG_REVERSE_MKL) { tq = x; }
                else { const float xt = (x + pv) * Li; tq = xt; tq2 = xt * __logf(xt); }
            }
            const int cnt = (int)min<uint64_t>(64, e - p0);
            // BT entries at a time: their center-table gathers are issued together,
            // then accumulated in entry order (the same sums as one at a time)
            constexpr int BT = MSR == MCG_JSD ? 8 : 16;
            int t0 = 0;
            if (MSR == MCG_JSD && CT2) {
                // the tabulated zero-baseline term of this row's li: one logarithm per
                // entry and center
                const float2 *CTr = CT2 + (uint64_t)slot[row] * ((uint64_t)d * kpad);
                for (; t0 + BT <= cnt; t0 += BT) {
                    float xv[BT], xl[BT];
                    float2 ct[BT];
#pragma unroll
                    for (int u = 0; u < BT; ++u) {
                        const uint32_t j = (uint32_t)__shfl((int)jq, t0 + u);
                        xv[u] = __shfl(tq, t0 + u);
                        xl[u] = __shfl(tq2, t0 + u);
                        ct[u] = CTr[(uint64_t)j * kpad + cc];
                    }
#pragma unroll
                    for (int u = 0; u < BT; ++u) {
                        const float m1 = xv[u] + ct[u].x;
                        acc += 0.5f * ((xl[u] - hli) - m1 * __logf(0.5f * m1) + ct[u].y);
                    }
                }
                for (int t = t0; t < cnt; ++t) {
                    const uint32_t j = (uint32_t)__shfl((int)jq, t);
                    const float xv = __shfl(tq, t), xl = __shfl(tq2, t);
                    const float2 ct = CTr[(uint64_t)j * kpad + cc];
                    const float m1 = xv + ct.x;
                    acc += 0.5f * ((xl - hli) - m1 * __logf(0.5f * m1) + ct.y);
                }
                continue;
            }
            for (; t0 + BT <= cnt; t0 += BT) {
                float xv[BT], ct[BT], xl[BT];
#pragma unroll
                for (int u = 0; u < BT; ++u) {
                    // JSD is VALU-bound: its broadcasts stay on the LDS pipe
                    const uint32_t j = MSR == MCG_JSD ? (uint32_t)__shfl((int)jq, t0 + u) : bcast_u32(jq, t0 + u);
                    xv[u] = MSR == MCG_JSD ? __shfl(tq, t0 + u) : bcast_f32(tq, t0 + u);
                    if (MSR == MCG_JSD) xl[u] = __shfl(tq2, t0 + u);
                    ct[u] = CT[(uint64_t)j * kpad + cc];
                }
#pragma unroll
                for (int u = 0; u < BT; ++u) {
                    if (MSR == MCG_SQRL2) acc = fmaf(xv[u], xv[u] - 2.f * ct[u], acc);
                    else if (MSR == MCG_MKL || MSR == MCG_REVERSE_MKL) acc = fmaf(xv[u], ct[u], acc);
                    else {
                        const float m1 = xv[u] + ct[u], m0 = li + ct[u];
                        acc += 0.5f * ((xl[u] - hli) - m1 * __logf(0.5f * m1) + m0 * __logf(0.5f * m0));
                    }
                }
            }
            for (int t = t0; t < cnt; ++t) {
                const uint32_t j = MSR == MCG_JSD ? (uint32_t)__shfl((int)jq, t) : bcast_u32(jq, t);
                const float xv = MSR == MCG_JSD ? __shfl(tq, t) : bcast_f32(tq, t);
                const float ct = CT[(uint64_t)j * kpad + cc];
                if (MSR == MCG_SQRL2) acc = fmaf(xv, xv - 2.f * ct, acc);
                else if (MSR == MCG_MKL || MSR == MCG_REVERSE_MKL) acc = fmaf(xv, ct, acc);
                else {
                    const float xl = __shfl(tq2, t);
                    const float m1 = xv + ct, m0 = li + ct;
                    acc += 0.5f * ((xl - hli) - m1 * __logf(0.5f * m1) + m0 * __logf(0.5f * m0));
                }
            }
        }
        float v;
        if (MSR == MCG_SQRL2) {
            v = (float)(cs[2 * cc] + (double)acc);
            v = v < 0.f ? 0.f : v;
        } else if (MSR == MCG_MKL) {
            v = (float)(cs[2 * cc] - cs[2 * cc + 1] * (double)(logpv - logL) - (double)acc);
            v = post_kl(v);
        } else if (MSR == MCG_REVERSE_MKL) {
            v = (float)((double)Hx - (double)li * cs[2 * cc] - (double)acc * (double)Li);
            v = post_kl(v);
        } else {
            v = (float)(G[(uint64_t)slot[row] * kpad + cc] + (double)acc);
            v = post_kl(v);
        }
        if (c < k) {
            if (out) out[row * k + c] = v;
            if (bi == 0xFFFFFFFFu || v < best) { best = v; bi = c; }
        }
    }
    if (asn) {
        for (int o = 32; o > 0; o >>= 1) {
            const float ob = __shfl_xor(best, o);
            const uint32_t oi = __shfl_xor(bi, o);
            const bool take = (oi != 0xFFFFFFFFu) && (bi == 0xFFFFFFFFu || ob < best || (ob == best && oi < bi));
            if (take) { best = ob; bi = oi; }
        }
        if (lane == 0) {
            asn[row] = bi;
            if (best_out) best_out[row] = best;
        }
    }
}

// Cost of one row against one center (column 0 of the center tables), computed by
// a wave: lanes over the row's stored entries, per-lane FP32 partial sums added by
// a fixed shuffle tree, then the per-row / per-center terms of k_row_costs.  Every
// lane returns the value.
struct Cost1 {
    const float *X, *val;
    const uint32_t *idx;
    const uint64_t *ptr;
    uint32_t d;
    float pv, logpv;
    const float4 *rt;
    const float *CT;
    uint32_t kpad;
    const double *cs, *G;
    const uint32_t *slot;
};

// Cost of one row against the single center of P, in two parts: the entry sum over
// [b, e) by a group of GS lanes (gl = the lane's index in its group; groups are
// aligned, so the reduction stays inside; four entries per lane are loaded before
// they are used, and added in entry order), then the per-row and per-center terms
// (r = the row's {L, log L, H_x, li}, g = JSD's G of the row's li).
template <int MSR, int GS>
__device__ __forceinline__ float row_sum1(const Cost1 &P, uint64_t b, uint64_t e, float4 r, int gl) {
    const float Li = 1.f / r.x, li = r.w;
    const float hli = li * __logf(li);
    float acc = 0.f;
    for (uint64_t p0 = b + gl; p0 < e; p0 += 4 * GS) {
        float x[4], ct[4];
        uint32_t j[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t p = p0 + (uint64_t)u * GS;
            const bool ok = p < e;
            x[u] = ok ? (P.val ? P.val[p] : P.X[p]) : 0.f;
            j[u] = ok ? (P.val ? P.idx[p] : (uint32_t)(p - b)) : 0u;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) ct[u] = P.CT[(uint64_t)j[u] * P.kpad];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (p0 + (uint64_t)u * GS >= e) break;
            const float xv = x[u], c = ct[u];
            if (MSR == MCG_SQRL2) acc = fmaf(xv, xv - 2.f * c, acc);
            else if (MSR == MCG_MKL) acc = fmaf(log1pf(xv / P.pv), c, acc);
            else if (MSR == MCG_REVERSE_MKL) acc = fmaf(xv, c, acc);
            else {
                const float xt = (xv + P.pv) * Li;
                const float m1 = xt + c, m0 = li + c;
                acc += 0.5f * ((xt * __logf(xt) - hli) - m1 * __logf(0.5f * m1) + m0 * __logf(0.5f * m0));
            }
        }
    }
    for (int o = GS / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    return acc;
}

template <int MSR>
__device__ __forceinline__ float finish1(const Cost1 &P, float acc, float4 r, double g) {
    const float logL = r.y, Hx = r.z, li = r.w, Li = 1.f / r.x;
    if (MSR == MCG_SQRL2) {
        const float v = (float)(P.cs[0] + (double)acc);
        return v < 0.f ? 0.f : v;
    }
    if (MSR == MCG_MKL) return post_kl((float)(P.cs[0] - P.cs[1] * (double)(P.logpv - logL) - (double)acc));
    if (MSR == MCG_REVERSE_MKL) return post_kl((float)((double)Hx - (double)li * P.cs[0] - (double)acc * (double)Li));
    return post_kl((float)(g + (double)acc));
}

template <int MSR, int GS = 64>
__device__ __forceinline__ float row_cost1(const Cost1 &P, uint64_t row, int lane) {
    uint64_t b, e;
    if (P.val) { b = P.ptr[row]; e = P.ptr[row + 1]; }
    else { b = row * P.d; e = b + P.d; }
    const float4 r = MSR != MCG_SQRL2 ? P.rt[row] : make_float4(1.f, 0.f, 0.f, 0.f);
    const double g = MSR == MCG_JSD ? P.G[(uint64_t)P.slot[row] * P.kpad] : 0.;
    return finish1<MSR>(P, row_sum1<MSR, GS>(P, b, e, r, lane), r, g);
}

// every row against one center (k = 1): asn = 0, best = the cost
template <int MSR>
__global__ __launch_bounds__(256) void k_row_cost1(Cost1 P, uint64_t n, uint32_t *__restrict__ asn,
                                                   float *__restrict__ best_out) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const float v = row_cost1<MSR>(P, row, lane);
    if (lane != 0) return;
    if (asn) asn[row] = 0;
    if (best_out) best_out[row] = v;
}

// D^2 seeding: cost of each earlier center's row cent[a] (a < ci) to the new center
template <int MSR>
__global__ __launch_bounds__(256) void k_pp_cc_gen(Cost1 P, const uint32_t *__restrict__ cent, uint32_t ci,
                                                   float *__restrict__ cc) {
    const uint32_t a = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (a >= ci) return;
    const float v = row_cost1<MSR>(P, cent[a], lane);
    if (lane == 0) cc[a] = v;
}

// D^2 seeding update (kmeans.h:189-204).  A wave takes ROWS (64, or 64 / GS when the
// rows are few, so that more waves share them) rows at a time: each lane tests its
// row for pruning and loads its row's extent and per-row terms, then the rows still
// needed are costed by groups of GS lanes (64 / GS rows at once).  With
// PRUNE (JSD, whose square root is a metric on the smoothed distributions): a row
// whose center is far from the new one, JSD(c_a, c_new) > 4 dist (1 + 2^-6) + 2^-17,
// cannot get closer (triangle inequality for sqrt(JSD); the margins cover the FP32
// evaluation of both costs), so its entries are not read and nothing changes.
template <int MSR, bool PRUNE, int GS, int ROWS>
__global__ __launch_bounds__(256) void k_pp_update_gen(Cost1 P, uint64_t n, const uint64_t *__restrict__ newc_p,
                                                       uint32_t ci, const float *__restrict__ cc,
                                                       float *__restrict__ dist, uint32_t *__restrict__ asn,
                                                       int first, const double *__restrict__ w,
                                                       unsigned long long *__restrict__ mx) {
    // waves stride over tiles of ROWS rows; one block-wide maximum and one atomic per block
    static_assert(ROWS == 64 || ROWS == 64 / GS, "a tile is 64 rows or one row per lane group");
    const uint64_t nwaves = (uint64_t)gridDim.x * 4;
    const int lane = threadIdx.x & 63, g = lane / GS, gl = lane % GS;
    const uint64_t newc = *newc_p;
    double a = 0.;
    for (uint64_t base = ((uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * ROWS; base < n; base += nwaves * ROWS) {
        const uint64_t row = base + lane;
        bool need = lane < ROWS && row < n;
        if (PRUNE && need && !first && row != newc) {
            const float v = dist[row];
            need = v > 0.f && !(cc[asn[row]] > 4.f * v * (1.f + 1.f / 64.f) + 0x1p-17f);
            if (!need) a = fmax(a, pp_weighted(v, w, row));   // unchanged
        }
        // the lane's row: entry range, per-row terms, JSD's G
        uint64_t rb = 0, re = 0;
        float4 rt = make_float4(1.f, 0.f, 0.f, 0.f);
        double rg = 0.;
        if (need) {
            if (P.val) { rb = P.ptr[row]; re = P.ptr[row + 1]; }
            else { rb = row * P.d; re = rb + P.d; }
            if (MSR != MCG_SQRL2) rt = P.rt[row];
            if (MSR == MCG_JSD) rg = P.G[(uint64_t)P.slot[row] * P.kpad];
        }
        uint64_t mask = __ballot(need);
        while (mask) {
            // group g takes the g-th lowest row left in mask
            uint64_t m = mask;
            for (int q = 0; q < g; ++q) m &= m - 1;
            for (int q = 0; q < 64 / GS; ++q) mask &= mask - 1;
            const bool has = m != 0;
            const int src = has ? __ffsll((unsigned long long)m) - 1 : lane;
            // every lane shuffles (uniform), groups without a row get an empty range
            const uint64_t b = __shfl(rb, src), e0 = __shfl(re, src);
            const uint64_t e = has ? e0 : b;
            float4 r;
            r.x = __shfl(rt.x, src);
            r.y = __shfl(rt.y, src);
            r.z = __shfl(rt.z, src);
            r.w = __shfl(rt.w, src);
            const double gg = __shfl(rg, src);
            const float acc = row_sum1<MSR, GS>(P, b, e, r, gl);
            if (has && gl == 0) {
                const uint64_t rr = base + (uint64_t)src;
                a = fmax(a, pp_weighted(pp_new_dist(rr, newc, finish1<MSR>(P, acc, r, gg), ci, first, dist, asn), w,
                                        rr));
            }
        }
    }
    block_max_into(a, mx);
}

// ---- reference-order per-pair kernel (tiny priors) -----------------------------
// FP32 (FT=float) following applicator.h:1642-1760 coordinate by coordinate, with
// pv = max(prior, 7.0069e-42 (lhsum + rhsum)) for each pair.  One wave per row.
__global__ __launch_bounds__(256) void k_pairs_direct(int msr, const float *__restrict__ X,
                                                      const float *__restrict__ val,
                                                      const uint32_t *__restrict__ idx,
                                                      const uint64_t *__restrict__ ptr, uint64_t n, uint32_t d,
                                                      const double *__restrict__ rsum, const float *__restrict__ C,
                                                      uint32_t k, const double *__restrict__ csum, float prior,
                                                      float *__restrict__ out, uint32_t *__restrict__ asn,
                                                      float *__restrict__ best_out) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    float best = 0.f;
    uint32_t bi = 0xFFFFFFFFu;
    const float xs = (float)rsum[row];
    for (uint32_t c0 = 0; c0 < k; c0 += 64) {
        const uint32_t c = c0 + lane;
        const uint32_t cc = c < k ? c : k - 1;
        const float cs = (float)csum[cc];
        const float psum0 = prior * (float)d;
        float lhsum = cs + psum0, rhsum = xs + psum0;
        const float smallest = 7.0069e-42f * (lhsum + rhsum);
        const float pv = fmaxf(prior, smallest);
        const float psum = pv * (float)d;
        lhsum = cs + psum;
        rhsum = xs + psum;
        const float lhrsi = 1.f / lhsum, rhrsi = 1.f / rhsum;
        float lhinc = (pv != 0.f && lhsum != 0.f) ? pv * lhrsi : 0.f;
        float rhinc = (pv != 0.f && rhsum != 0.f) ? pv * rhrsi : 0.f;
        if (lhinc != lhinc) lhinc = 0.f;
        if (rhinc != rhinc) rhinc = 0.f;
        float acc = 0.f;
        uint64_t p = val ? ptr[row] : 0, pe = val ? ptr[row + 1] : 0;
        const float *cr = C + (uint64_t)cc * d;
        for (uint32_t j = 0; j < d; ++j) {
            float x;
            if (val) { x = (p < pe && idx[p] == j) ? val[p++] : 0.f; }
            else x = X[row * d + j];
            const float cj = cr[j];
            if (msr == MCG_SQRL2) { const float t = cj - x; acc = fmaf(t, t, acc); continue; }
            const float a = cj * lhrsi + lhinc, bb = x * rhrsi + rhinc;
            if (msr == MCG_MKL) acc += a * logf(a / bb);
            else if (msr == MCG_REVERSE_MKL) acc += bb * logf(bb / a);
            else { const float m = a + bb; acc += 0.5f * (a * logf(2.f * a / m) + bb * logf(2.f * bb / m)); }
        }
        float v = acc;
        if (msr != MCG_SQRL2) v = post_kl(v);
        if (c < k) {
            if (out) out[row * k + c] = v;
            if (bi == 0xFFFFFFFFu || v < best) { best = v; bi = c; }
        }
    }
    if (asn) {
        for (int o = 32; o > 0; o >>= 1) {
            const float ob = __shfl_xor(best, o);
            const uint32_t oi = __shfl_xor(bi, o);
            const bool take = (oi != 0xFFFFFFFFu) && (bi == 0xFFFFFFFFu || ob < best || (ob == best && oi < bi));
            if (take) { best = ob; bi = oi; }
        }
        if (lane == 0) { asn[row] = bi; if (best_out) best_out[row] = best; }
    }
}

// The further measures, pair by pair in FP32 (FT = float) over every coordinate,
// following applicator.h:1642-1760 with a = smoothed center, b = smoothed row:
// HELLINGER sqrt(sum (sqrt a - sqrt b)^2) / sqrt 2; BHATTACHARYYA from sum sqrt(a b)
// (1 if within 1e-8 of 1; metric sqrt(1 - s), distance -log s); TVD sum |a - b| / 2;
// L1 sum |c - x|; L2 sqrt(sum (c - x)^2); cosine (c + pv, x + pv) or, for the
// probability forms, (a, b), clamped to [0, 1], distances acos / pi; the LLR family
// sum lambda a log(a/m) + (1-lambda) b log(b/m) with lambda = lhsum / (lhsum + rhsum)
// and m = lambda a + (1-lambda) b (LLR and SRLRT times lhsum + rhsum; SRLRT, SRULRT
// square roots); IS sum a/b - log(a/b) - 1 (the -1 per coordinate, not as -d at the
// end, which would cancel in FP32), REVERSE_IS swapped; SIS and REVERSE_SIS
// sum log((a/b + b/a + 2) / 4) / 2; JSM sqrt(JSD).  One wave per row, lanes over
// centers; a CSR row is walked along the dense coordinates.
// One coordinate of the I-divergence: xh log(xh / ch) - xh + ch, written as
// ch (t log1p(u) - u) with t = xh / ch = 1 + u, which keeps the relative error small
// when xh and ch are close; xh = 0 gives ch, ch = 0 < xh gives +inf.
__device__ __forceinline__ float poisson_term(float xh, float ch) {
    if (xh == 0.f) return ch;
    if (ch == 0.f) return INFINITY;
    const float u = (xh - ch) / ch;
    return ch * ((1.f + u) * log1pf(u) - u);
}

__global__ __launch_bounds__(256) void k_pairs_more(int msr, const float *__restrict__ X, const float *__restrict__ val,
                                                    const uint32_t *__restrict__ idx, const uint64_t *__restrict__ ptr,
                                                    uint64_t n, uint32_t d, const double *__restrict__ rsum,
                                                    const float *__restrict__ C, uint32_t k,
                                                    const double *__restrict__ csum, float prior,
                                                    float *__restrict__ out, uint32_t *__restrict__ asn,
                                                    float *__restrict__ best_out) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    float best = 0.f;
    uint32_t bi = 0xFFFFFFFFu;
    const float xs = (float)rsum[row];
    for (uint32_t c0 = 0; c0 < k; c0 += 64) {
        const uint32_t c = c0 + lane;
        const uint32_t cc = c < k ? c : k - 1;
        const float cs = (float)csum[cc];
        const float psum0 = prior * (float)d;
        float lhsum = cs + psum0, rhsum = xs + psum0;
        const float pv = fmaxf(prior, 7.0069e-42f * (lhsum + rhsum));
        const float psum = pv * (float)d;
        lhsum = cs + psum;
        rhsum = xs + psum;
        const float lhrsi = 1.f / lhsum, rhrsi = 1.f / rhsum;
        float lhinc = (pv != 0.f && lhsum != 0.f) ? pv * lhrsi : 0.f;
        float rhinc = (pv != 0.f && rhsum != 0.f) ? pv * rhrsi : 0.f;
        if (lhinc != lhinc) lhinc = 0.f;
        if (rhinc != rhinc) rhinc = 0.f;
        const float lam = lhsum / (lhsum + rhsum), lam1 = 1.f - lam;
        float s = 0.f, s2 = 0.f, s3 = 0.f;
        uint64_t p = val ? ptr[row] : 0, pe = val ? ptr[row + 1] : 0;
        const float *cr = C + (uint64_t)cc * d;
        for (uint32_t j = 0; j < d; ++j) {
            float x;
            if (val) { x = (p < pe && idx[p] == j) ? val[p++] : 0.f; }
            else x = X[row * d + j];
            const float cj = cr[j];
            const float a = fmaf(cj, lhrsi, lhinc), b = fmaf(x, rhrsi, rhinc);
            switch (msr) {
            case MCG_L1: s += fabsf(cj - x); break;
            case MCG_L2: { const float t = cj - x; s = fmaf(t, t, s); } break;
            case MCG_HELLINGER: { const float t = sqrtf(a) - sqrtf(b); s = fmaf(t, t, s); } break;
            case MCG_BHATTACHARYYA_METRIC: case MCG_BHATTACHARYYA_DISTANCE: s += sqrtf(a * b); break;
            case MCG_TVD: s += fabsf(a - b); break;
            case MCG_COSINE_DISTANCE: case MCG_COSINE_SIMILARITY: {
                const float u = cj + pv, v = x + pv;
                s = fmaf(u, v, s); s2 = fmaf(u, u, s2); s3 = fmaf(v, v, s3);
            } break;
            case MCG_PROBABILITY_COSINE_DISTANCE: case MCG_PROBABILITY_COSINE_SIMILARITY:
                s = fmaf(a, b, s); s2 = fmaf(a, a, s2); s3 = fmaf(b, b, s3); break;
            case MCG_JSM: { const float m = a + b; s += 0.5f * (a * logf(2.f * a / m) + b * logf(2.f * b / m)); } break;
            case MCG_LLR: case MCG_UWLLR: case MCG_SRLRT: case MCG_SRULRT: {
                const float m = lam * a + lam1 * b;
                s += (a != 0.f ? lam * a * logf(a / m) : 0.f) + (b != 0.f ? lam1 * b * logf(b / m) : 0.f);
            } break;
            case MCG_IS: { const float r = a / b; s += (r - logf(r)) - 1.f; } break;
            case MCG_REVERSE_IS: { const float r = b / a; s += (r - logf(r)) - 1.f; } break;
            case MCG_POISSON: s += poisson_term(x + prior, cj + prior); break;
            default: s += 0.5f * logf((a / b + b / a + 2.f) * 0.25f); break;   // SIS, RSIS
            }
        }
        float v;
        switch (msr) {
        case MCG_POISSON: v = s < 0.f ? 0.f : (s == INFINITY ? FLT_MAX : s); break;
        case MCG_L1: v = s; break;
        case MCG_L2: v = sqrtf(s); break;
        case MCG_HELLINGER: v = sqrtf(s) * 0.70710678f; break;
        case MCG_BHATTACHARYYA_METRIC: case MCG_BHATTACHARYYA_DISTANCE:
            if (1.f - s < 1e-8f) s = 1.f;
            v = msr == MCG_BHATTACHARYYA_METRIC ? sqrtf(fmaxf(1.f - s, 0.f)) : -logf(s);
            break;
        case MCG_TVD: v = 0.5f * s; break;
        case MCG_COSINE_DISTANCE: case MCG_COSINE_SIMILARITY:
        case MCG_PROBABILITY_COSINE_DISTANCE: case MCG_PROBABILITY_COSINE_SIMILARITY: {
            float r = s / sqrtf(s2 * s3);
            r = r > 1.f ? 1.f : (r < 0.f ? 0.f : r);
            if (msr == MCG_COSINE_DISTANCE || msr == MCG_PROBABILITY_COSINE_DISTANCE) r = acosf(r) * 0.31830988618f;
            v = r;
        } break;
        default: {   // KL family tail (applicator.h:1750-1760)
            float r = s;
            if (msr == MCG_LLR || msr == MCG_SRLRT) r *= (lhsum + rhsum);
            r = r < 0.f ? 0.f : r;
            if (msr == MCG_SRULRT || msr == MCG_SRLRT || msr == MCG_JSM) r = sqrtf(r);
            v = r == INFINITY ? FLT_MAX : r;
        } break;
        }
        if (c < k) {
            if (out) out[row * k + c] = v;
            if (bi == 0xFFFFFFFFu || v < best) { best = v; bi = c; }
        }
    }
    if (asn) {
        for (int o = 32; o > 0; o >>= 1) {
            const float ob = __shfl_xor(best, o);
            const uint32_t oi = __shfl_xor(bi, o);
            const bool take = (oi != 0xFFFFFFFFu) && (bi == 0xFFFFFFFFu || ob < best || (ob == best && oi < bi));
            if (take) { best = ob; bi = oi; }
        }
        if (lane == 0) { asn[row] = bi; if (best_out) best_out[row] = best; }
    }
}

// Dense L1 / L2 costs by 64-row x 64-center tiles (the median policies' E-step):
// 256 threads, thread (tr, tc) owns rows r0 + 4tr + {0..3} and centers c0 + 4tc +
// {0..3}; 32-dim chunks of both staged in LDS dim-major, read as float4.  Each
// pair's sum runs over j in order, as in k_pairs_more, so the values are the same;
// VALU-bound at 2 operations per pair and dimension.  Argmin: a thread scans its
// centers in increasing order (strict <), then the 16 threads of a row combine
// (value, index) with the lower index on ties.
template <bool L2>
__global__ __launch_bounds__(256) void k_tile_lp(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                 const float *__restrict__ C, uint32_t k, float *__restrict__ out,
                                                 uint32_t *__restrict__ asn, float *__restrict__ best_out) {
    constexpr uint32_t TJ = 32, TR = 64, LD = TR + 4;
    __shared__ float xs[TJ * LD], cs[TJ * LD];
    __shared__ float bvs[16 * TR];
    __shared__ uint32_t bis[16 * TR];
    const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
    const uint64_t r0 = (uint64_t)blockIdx.x * TR;
    float best[4];
    uint32_t bidx[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) { best[a] = 0.f; bidx[a] = 0xFFFFFFFFu; }
    const uint32_t sr = tid >> 2, sj = (tid & 3) * 8;   // staging: row / center sr, dims sj..sj+7
    const uint64_t xrow = min<uint64_t>(r0 + sr, n - 1);
    for (uint32_t c0 = 0; c0 < k; c0 += TR) {
        const uint32_t crow = min<uint32_t>(c0 + sr, k - 1);
        float acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
        for (uint32_t j0 = 0; j0 < d; j0 += TJ) {
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                const uint32_t j = j0 + sj + q;
                xs[(sj + q) * LD + sr] = j < d ? X[xrow * d + j] : 0.f;
                cs[(sj + q) * LD + sr] = j < d ? C[(uint64_t)crow * d + j] : 0.f;
            }
            __syncthreads();
            const uint32_t jn = min(TJ, d - j0);
            for (uint32_t j = 0; j < jn; ++j) {
                const float4 xv = *(const float4 *)&xs[j * LD + tr * 4];
                const float4 cv = *(const float4 *)&cs[j * LD + tc * 4];
                const float xa[4] = {xv.x, xv.y, xv.z, xv.w}, cb[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        if constexpr (L2) {
                            const float t = cb[b] - xa[a];
                            acc[a][b] = fmaf(t, t, acc[a][b]);
                        } else {
                            acc[a][b] += fabsf(cb[b] - xa[a]);
                        }
                    }
            }
            __syncthreads();
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const uint64_t row = r0 + tr * 4 + a;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t c = c0 + tc * 4 + b;
                if (c >= k || row >= n) continue;
                const float v = L2 ? sqrtf(acc[a][b]) : acc[a][b];
                if (out) out[row * k + c] = v;
                if (bidx[a] == 0xFFFFFFFFu || v < best[a]) { best[a] = v; bidx[a] = c; }
            }
        }
    }
    if (!asn) return;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        bvs[tc * TR + tr * 4 + a] = best[a];
        bis[tc * TR + tr * 4 + a] = bidx[a];
    }
    __syncthreads();
    if (tid < (int)TR) {
        float b = bvs[tid];
        uint32_t ix = bis[tid];
        for (int t = 1; t < 16; ++t) {
            const float ob = bvs[t * TR + tid];
            const uint32_t oi = bis[t * TR + tid];
            if (oi != 0xFFFFFFFFu && (ix == 0xFFFFFFFFu || ob < b || (ob == b && oi < ix))) { b = ob; ix = oi; }
        }
        const uint64_t row = r0 + tid;
        if (row < n) {
            asn[row] = ix;
            if (best_out) best_out[row] = b;
        }
    }
}

// Dense L1 / L2 costs against one center (the empty-cluster restarts): one thread
// per row, 32-dim slices of the block's 256 rows staged in LDS with coalesced loads;
// the per-pair order of k_tile_lp, so the values are the same.
template <bool L2>
__global__ __launch_bounds__(256) void k_lp_single(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                   const float *__restrict__ C, float *__restrict__ out,
                                                   uint32_t *__restrict__ asn, float *__restrict__ best_out) {
    __shared__ float xs[256][33];
    __shared__ float cs[32];
    const int tid = threadIdx.x;
    const uint64_t p0 = (uint64_t)blockIdx.x * 256, p = p0 + tid;
    float acc = 0.f;
    for (uint32_t j0 = 0; j0 < d; j0 += 32) {
        __syncthreads();
        for (int q = tid; q < 256 * 32; q += 256) {
            const int r = q >> 5, jj = q & 31;
            const uint64_t pr = p0 + r;
            xs[r][jj] = (pr < n && j0 + jj < d) ? X[pr * d + j0 + jj] : 0.f;
        }
        if (tid < 32) cs[tid] = j0 + tid < d ? C[j0 + tid] : 0.f;
        __syncthreads();
        const int jn = (int)min<uint32_t>(32, d - j0);
        for (int jj = 0; jj < jn; ++jj) {
            if constexpr (L2) {
                const float t = cs[jj] - xs[tid][jj];
                acc = fmaf(t, t, acc);
            } else {
                acc += fabsf(cs[jj] - xs[tid][jj]);
            }
        }
    }
    if (p >= n) return;
    const float v = L2 ? sqrtf(acc) : acc;
    if (out) out[p] = v;
    if (asn) asn[p] = 0u;
    if (best_out) best_out[p] = v;
}

// The smoothing of one (row, center) pair as k_pairs_more computes it
// (applicator.h:1642-1656 in FP32): a = c * lhrsi + lhinc, b = x * rhrsi + rhinc.
struct PairSmooth {
    float lhrsi, lhinc, rhrsi, rhinc;
};
__device__ __forceinline__ PairSmooth pair_smooth(float cs, float xs, float prior, uint32_t d) {
    const float psum0 = prior * (float)d;
    float lhsum = cs + psum0, rhsum = xs + psum0;
    const float pv = fmaxf(prior, 7.0069e-42f * (lhsum + rhsum));
    const float psum = pv * (float)d;
    lhsum = cs + psum;
    rhsum = xs + psum;
    PairSmooth r;
    r.lhrsi = 1.f / lhsum;
    r.rhrsi = 1.f / rhsum;
    r.lhinc = (pv != 0.f && lhsum != 0.f) ? pv * r.lhrsi : 0.f;
    r.rhinc = (pv != 0.f && rhsum != 0.f) ? pv * r.rhrsi : 0.f;
    if (r.lhinc != r.lhinc) r.lhinc = 0.f;
    if (r.rhinc != r.rhinc) r.rhinc = 0.f;
    return r;
}

// Dense TVD costs by 64 x 64 tiles (the layout of k_tile_lp): the smoothing of each
// of a thread's 16 pairs is set up per center tile; per dimension two FMAs, a
// difference and an absolute-value add, in k_pairs_more's order and form.
__global__ __launch_bounds__(256) void k_tile_tvd(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                  const double *__restrict__ rsum, const float *__restrict__ C,
                                                  uint32_t k, const double *__restrict__ csum, float prior,
                                                  float *__restrict__ out, uint32_t *__restrict__ asn,
                                                  float *__restrict__ best_out) {
    constexpr uint32_t TJ = 32, TR = 64, LD = TR + 4;
    __shared__ float xs[TJ * LD], cs[TJ * LD];
    __shared__ float bvs[16 * TR];
    __shared__ uint32_t bis[16 * TR];
    const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
    const uint64_t r0 = (uint64_t)blockIdx.x * TR;
    float best[4];
    uint32_t bidx[4];
    float xsum[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        best[a] = 0.f;
        bidx[a] = 0xFFFFFFFFu;
        xsum[a] = (float)rsum[min<uint64_t>(r0 + tr * 4 + a, n - 1)];
    }
    const uint32_t sr = tid >> 2, sj = (tid & 3) * 8;
    const uint64_t xrow = min<uint64_t>(r0 + sr, n - 1);
    for (uint32_t c0 = 0; c0 < k; c0 += TR) {
        const uint32_t crow = min<uint32_t>(c0 + sr, k - 1);
        PairSmooth sm[4][4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float csb = (float)csum[min<uint32_t>(c0 + tc * 4 + b, k - 1)];
#pragma unroll
            for (int a = 0; a < 4; ++a) sm[a][b] = pair_smooth(csb, xsum[a], prior, d);
        }
        float acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
        for (uint32_t j0 = 0; j0 < d; j0 += TJ) {
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                const uint32_t j = j0 + sj + q;
                xs[(sj + q) * LD + sr] = j < d ? X[xrow * d + j] : 0.f;
                cs[(sj + q) * LD + sr] = j < d ? C[(uint64_t)crow * d + j] : 0.f;
            }
            __syncthreads();
            const uint32_t jn = min(TJ, d - j0);
            for (uint32_t j = 0; j < jn; ++j) {
                const float4 xv = *(const float4 *)&xs[j * LD + tr * 4];
                const float4 cv = *(const float4 *)&cs[j * LD + tc * 4];
                const float xa[4] = {xv.x, xv.y, xv.z, xv.w}, cb[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const float av = fmaf(cb[b], sm[a][b].lhrsi, sm[a][b].lhinc);
                        const float bv = fmaf(xa[a], sm[a][b].rhrsi, sm[a][b].rhinc);
                        acc[a][b] += fabsf(av - bv);
                    }
            }
            __syncthreads();
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const uint64_t row = r0 + tr * 4 + a;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t c = c0 + tc * 4 + b;
                if (c >= k || row >= n) continue;
                const float v = 0.5f * acc[a][b];
                if (out) out[row * k + c] = v;
                if (bidx[a] == 0xFFFFFFFFu || v < best[a]) { best[a] = v; bidx[a] = c; }
            }
        }
    }
    if (!asn) return;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        bvs[tc * TR + tr * 4 + a] = best[a];
        bis[tc * TR + tr * 4 + a] = bidx[a];
    }
    __syncthreads();
    if (tid < (int)TR) {
        float b = bvs[tid];
        uint32_t ix = bis[tid];
        for (int t = 1; t < 16; ++t) {
            const float ob = bvs[t * TR + tid];
            const uint32_t oi = bis[t * TR + tid];
            if (oi != 0xFFFFFFFFu && (ix == 0xFFFFFFFFu || ob < b || (ob == b && oi < ix))) { b = ob; ix = oi; }
        }
        const uint64_t row = r0 + tid;
        if (row < n) {
            asn[row] = ix;
            if (best_out) best_out[row] = b;
        }
    }
}

__global__ void k_center_sums(const float *__restrict__ C, uint32_t k, uint32_t d, double *__restrict__ out) {
    const uint64_t c = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= k) return;
    double s = 0.;
    for (uint32_t j = lane; j < d; j += 64) s += (double)C[c * d + j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[c] = s;
}

// Per-prior row transforms: row terms and (JSD) the distinct li values.
void csr_transforms(Data &ds, double pv, hipStream_t st) {
    if (ds.tf_prior == pv && ds.rowinfo.p) return;
    rowsums(ds, st);
    ds.rowinfo.reserve(ds.n * sizeof(float4));
    // row terms for REVERSE_MKL/JSD need H_x; computing it for every measure is cheap
    hipLaunchKernelGGL(k_row_terms, dim3((uint32_t)cdiv(ds.n, 4)), dim3(256), 0, st, (int)MCG_JSD, ds.X,
                       ds.csr ? ds.val.as<const float>() : nullptr, ds.csr ? ds.ptr.as<const uint64_t>() : nullptr,
                       ds.n, (uint32_t)ds.d, ds.rowsum.as<const double>(), pv, ds.rowinfo.as<float4>());
    // distinct li (float) values for JSD
    std::vector<float4> ri(ds.n);
    HIP_CHECK(hipMemcpyAsync(ri.data(), ds.rowinfo.p, ds.n * sizeof(float4), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    std::vector<float> lis(ds.n);
    for (uint64_t i = 0; i < ds.n; ++i) lis[i] = ri[i].w;
    std::vector<float> u = lis;
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    std::vector<uint32_t> sl(ds.n);
    for (uint64_t i = 0; i < ds.n; ++i)
        sl[i] = (uint32_t)(std::lower_bound(u.begin(), u.end(), lis[i]) - u.begin());
    ds.li_unique = u;
    ds.li_dev.reserve(u.size() * sizeof(float));
    copy_sync(ds.li_dev.p, u.data(), u.size() * sizeof(float), hipMemcpyHostToDevice, st);
    ds.li_slot.reserve(ds.n * 4);
    copy_sync(ds.li_slot.p, sl.data(), ds.n * 4, hipMemcpyHostToDevice, st);
    ds.tf_prior = pv;
}

// ---- dense MKL / REVERSE_MKL cross term on MFMA ------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));

// The decomposed costs need, per (row, center), the cross term sum_j a_j * t_cj over all
// d coordinates: a = x (REVERSE_MKL) or log1p(x / pv) (MKL), t = log c~ or c~ (the
// center tables).  For dense rows with d <= 256 it is an FP32 GEMM on
// v_mfma_f32_32x32x2f32 (FP32 products and sums, as the row kernel; the sum is
// formed in the matrix core's order): 32 rows per wave, their operand in registers
// (lane half h supplies coordinates h*DP/2 + s), the centers' tables in LDS in tiles
// of 32 (center-major, copied from the d-major table by k_ct_transpose, so the values
// equal the row kernel's).  The epilogue adds the per-row and per-center terms as
// k_row_costs does (FP64), clamps, writes the cost matrix and/or min-reduces 64-bit keys
// (cost bits above, center below: lowest cost, then lowest center; a NaN cost never
// wins except at center 0, which the row kernel's scan keeps).
__global__ void k_ct_transpose(const float *__restrict__ CT, uint32_t kpad, uint32_t k, uint32_t d, uint32_t dpad,
                               float *__restrict__ Ct) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)kpad * dpad) return;
    const uint32_t c = (uint32_t)(e / dpad), j = (uint32_t)(e % dpad);
    Ct[e] = (c < k && j < d) ? CT[(uint64_t)j * kpad + c] : 0.f;
}

// MKL's row operand log1p(x / pv), the row kernel's expression, once per dataset and prior
__global__ void k_log1p_div(const float *__restrict__ X, uint64_t m, float pv, float *__restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = log1pf(X[i] / pv);
}

template <int MSR, int DP>
__global__ __launch_bounds__(256) void k_breg_mfma(const float *__restrict__ X, uint32_t d, uint64_t n, float pv,
                                                   float logpv, const float4 *__restrict__ rt,
                                                   const float *__restrict__ Ct, const double *__restrict__ cs,
                                                   uint32_t k, uint32_t ntiles, float *__restrict__ out,
                                                   uint32_t *__restrict__ asn, float *__restrict__ best_out) {
    constexpr int KS = DP / 2;
    constexpr int ROWF = DP + 4;              // padded LDS row stride (floats)
    constexpr int CPT = (32 * DP / 4) / 256;  // float4 chunks per thread per tile (2, 4, 8)
    static_assert(CPT <= 8, "tile chunks");
    __shared__ __attribute__((aligned(16))) float lds[2][32 * ROWF];
    __shared__ double lcs[2][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, r32 = lane & 31;
    const uint64_t p = (uint64_t)blockIdx.x * 128u + wave * 32u + r32;
    const uint64_t pc = p < n ? p : n - 1;
    float b[KS];
    const float *xr = X + pc * d;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const uint32_t j = h * KS + s;
        b[s] = j < d ? xr[j] : 0.f;   // MKL: X holds log1p(x / pv) already (k_log1p_div)
    }
    const float4 r = rt[pc];
    // the next tile in registers (CPT float4 per thread), then into LDS; written out
    // without an array captured by a lambda, which would live in scratch memory
    float4 pf0, pf1, pf2, pf3, pf4, pf5, pf6, pf7;
    double csn = 0.;
#define BREG_LD(q, v)                                                                                      \
    if (q < CPT) {                                                                                         \
        const int g = tid + q * 256;                                                                       \
        v = ((const float4 *)Ct)[(size_t)(t1 * 32 + g / (DP / 4)) * (DP / 4) + (g % (DP / 4))];          \
    }
#define BREG_ST(q, v)                                                                                      \
    if (q < CPT) {                                                                                         \
        const int g = tid + q * 256;                                                                       \
        *(float4 *)(&lds[buf1][(g / (DP / 4)) * ROWF + (g % (DP / 4)) * 4]) = v;                          \
    }
    auto load_tile = [&](uint32_t t1) __attribute__((always_inline)) {
        BREG_LD(0, pf0) BREG_LD(1, pf1) BREG_LD(2, pf2) BREG_LD(3, pf3)
        BREG_LD(4, pf4) BREG_LD(5, pf5) BREG_LD(6, pf6) BREG_LD(7, pf7)
        if (tid < 64) csn = cs[(uint64_t)t1 * 64 + tid];   // {a, b} of the tile's 32 centers
    };
    auto store_tile = [&](int buf1) __attribute__((always_inline)) {
        BREG_ST(0, pf0) BREG_ST(1, pf1) BREG_ST(2, pf2) BREG_ST(3, pf3)
        BREG_ST(4, pf4) BREG_ST(5, pf5) BREG_ST(6, pf6) BREG_ST(7, pf7)
        if (tid < 64) lcs[buf1][tid] = csn;
    };
#undef BREG_LD
#undef BREG_ST
    unsigned long long key = ~0ull;
    bool nan0 = false;
    load_tile(0);
    store_tile(0);
    __syncthreads();
    for (uint32_t t = 0; t < ntiles; ++t) {
        const int buf = t & 1;
        if (t + 1 < ntiles) load_tile(t + 1);
        floatx16 acc;
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[e] = 0.f;
        const float *arow = &lds[buf][r32 * ROWF + h * KS];
#pragma unroll
        for (int s = 0; s < KS; s += 4) {
            const float4 a = *(const float4 *)(arow + s);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b[s + 0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b[s + 1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b[s + 2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, b[s + 3], acc, 0, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const uint32_t cl = 8u * (e >> 2) + 4u * h + (e & 3), c = t * 32u + cl;
            if (c >= k) continue;
            // an opaque copy keeps each element's FP64 epilogue with the element (hoisted, the
            // widened values of the whole tile would not fit the registers)
            float av = acc[e];
            asm volatile("" : "+v"(av));
            const double c0 = lcs[buf][2 * cl], c1 = lcs[buf][2 * cl + 1];
            float v;
            if (MSR == MCG_MKL) v = (float)(c0 - c1 * (double)(logpv - r.y) - (double)av);
            else v = (float)((double)r.z - (double)r.w * c0 - (double)av * (double)(1.f / r.x));
            v = v < 0.f ? 0.f : v;
            if (v == INFINITY) v = FLT_MAX;
            if (out && p < n) out[p * k + c] = v;
            if (v != v) {
                if (c == 0) nan0 = true;
            } else {
                const unsigned long long kc = ((unsigned long long)__float_as_uint(v) << 32) | c;
                key = kc < key ? kc : key;
            }
        }
        if (t + 1 < ntiles) store_tile(buf ^ 1);
        __syncthreads();
    }
    if (asn || best_out) {
        const unsigned long long ok = __shfl_xor(key, 32);
        key = ok < key ? ok : key;
        const bool on = __shfl_xor((int)nan0, 32) != 0;
        nan0 = nan0 || on;
        if (h == 0 && p < n) {
            if (asn) asn[p] = nan0 ? 0u : (key == ~0ull ? 0u : (uint32_t)key);
            if (best_out) best_out[p] = (nan0 || key == ~0ull) ? NAN : __uint_as_float((uint32_t)(key >> 32));
        }
    }
}

// Per-center tables (CT d x kpad, per-center scalars, JSD's G per distinct li) and
// the per-row terms for k centers C; returns them with the row arrays of ds.
// Center tables with row stride kpad (>= k): a multiple of 64 for the kernels whose
// lanes run over centers, 1 for the single-center kernels (a dense table of d values).
static Cost1 center_tables(Data &ds, int msr, double pv, const float *C, uint64_t k, uint32_t kpad,
                           hipStream_t st) {
    Ctx &c = *ds.ctx;
    const uint32_t d = (uint32_t)ds.d;
    Cost1 P{};
    P.X = ds.csr ? nullptr : ds.X;
    P.val = ds.csr ? ds.val.as<const float>() : nullptr;
    P.idx = ds.csr ? ds.idx.as<const uint32_t>() : nullptr;
    P.ptr = ds.csr ? ds.ptr.as<const uint64_t>() : nullptr;
    P.d = d;
    P.pv = (float)pv;
    P.logpv = (float)log(pv > 0 ? pv : 1.);
    P.kpad = kpad;
    float *CT = (float *)c.s[S_CT].reserve((uint64_t)d * kpad * sizeof(float));
    double *cs = (double *)c.s[S_CPARAM].reserve(2 * (uint64_t)kpad * sizeof(double));
    hipLaunchKernelGGL(k_center_tables, dim3((uint32_t)k), dim3(256), 0, st, msr, C, (uint32_t)k, kpad, d, pv, CT, cs);
    P.CT = CT;
    P.cs = cs;
    if (msr != MCG_SQRL2) {
        csr_transforms(ds, pv, st);
        P.rt = ds.rowinfo.as<const float4>();
    }
    if (msr == MCG_JSD) {
        const uint32_t U = (uint32_t)ds.li_unique.size();
        REQUIRE(U <= (1u << 20), "JSD: too many distinct row sums");
        double *Gb = (double *)c.s[S_JSDG].reserve((uint64_t)U * kpad * sizeof(double));
        hipLaunchKernelGGL(k_jsd_g, dim3((uint32_t)std::min<uint64_t>((uint64_t)U * k, 65536)), dim3(256), 0, st, CT, kpad,
                           (uint32_t)k, d, ds.li_dev.as<const float>(), U, Gb);
        P.G = Gb;
        P.slot = ds.li_slot.as<const uint32_t>();
    }
    return P;
}

template <int GS, int ROWS>
static void pp_update_launch_r(const Cost1 &P, uint64_t n, int msr, const uint64_t *newc, uint32_t ci,
                               const float *cc, float *dist, uint32_t *asn, int first, const double *w,
                               unsigned long long *mx, hipStream_t st) {
    const dim3 grid((uint32_t)std::min<uint64_t>(cdiv(n, 4 * ROWS), ROWS == 64 ? 2048 : 8192)), blk(256);
    if (msr == MCG_JSD && !first && ci > 0) {
        hipLaunchKernelGGL((k_pp_update_gen<MCG_JSD, true, GS, ROWS>), grid, blk, 0, st, P, n, newc, ci, cc, dist,
                           asn, first, w, mx);
        return;
    }
    switch (msr) {
    case MCG_SQRL2:
        hipLaunchKernelGGL((k_pp_update_gen<MCG_SQRL2, false, GS, ROWS>), grid, blk, 0, st, P, n, newc, ci, cc, dist,
                           asn, first, w, mx);
        break;
    case MCG_MKL:
        hipLaunchKernelGGL((k_pp_update_gen<MCG_MKL, false, GS, ROWS>), grid, blk, 0, st, P, n, newc, ci, cc, dist,
                           asn, first, w, mx);
        break;
    case MCG_REVERSE_MKL:
        hipLaunchKernelGGL((k_pp_update_gen<MCG_REVERSE_MKL, false, GS, ROWS>), grid, blk, 0, st, P, n, newc, ci, cc,
                           dist, asn, first, w, mx);
        break;
    default:
        hipLaunchKernelGGL((k_pp_update_gen<MCG_JSD, false, GS, ROWS>), grid, blk, 0, st, P, n, newc, ci, cc, dist,
                           asn, first, w, mx);
        break;
    }
}

// Fewer than 2^18 rows: a wave takes one row per lane group at a time, so that the
// rows' dependent loads are spread over more waves.
template <int GS>
static void pp_update_launch(const Cost1 &P, uint64_t n, int msr, const uint64_t *newc, uint32_t ci, const float *cc,
                             float *dist, uint32_t *asn, int first, const double *w, unsigned long long *mx,
                             hipStream_t st) {
    if (n < (1ull << 18)) pp_update_launch_r<GS, 64 / GS>(P, n, msr, newc, ci, cc, dist, asn, first, w, mx, st);
    else pp_update_launch_r<GS, 64>(P, n, msr, newc, ci, cc, dist, asn, first, w, mx, st);
}

bool generic_pp_round(Data &ds, int msr, double prior, const float *center_row, const uint32_t *cent, uint32_t ci,
                      const uint64_t *newc, float *cc, float *dist, uint32_t *asn, int first, const double *w,
                      unsigned long long *mx, hipStream_t st) {
    if (!msr_decomposed(msr) || (msr != MCG_SQRL2 && prior < DIRECT_PRIOR)) return false;
    const double pv = msr == MCG_SQRL2 ? 0. : prior;
    const Cost1 P = center_tables(ds, msr, pv, center_row, 1, 1, st);
    const uint64_t n = ds.n;
    if (msr == MCG_JSD && !first && ci > 0)
        hipLaunchKernelGGL(k_pp_cc_gen<MCG_JSD>, dim3((uint32_t)cdiv(ci, 4)), dim3(256), 0, st, P, cent, ci, cc);
    // sparse rows: 8 lanes per row, 8 rows per wave at once; dense rows: a wave per row
    if (ds.csr) pp_update_launch<8>(P, n, msr, newc, ci, cc, dist, asn, first, w, mx, st);
    else pp_update_launch<64>(P, n, msr, newc, ci, cc, dist, asn, first, w, mx, st);
    return true;
}

bool generic_pp_tables(Data &ds, int msr, double prior, PPTables &t, hipStream_t st) {
    if (!msr_decomposed(msr) || (msr != MCG_SQRL2 && prior < DIRECT_PRIOR)) return false;
    Ctx &c = *ds.ctx;
    const double pv = msr == MCG_SQRL2 ? 0. : prior;
    t.pv = pv;
    t.CT = (float *)c.s[S_CT].reserve(ds.d * sizeof(float));
    t.cs = (double *)c.s[S_CPARAM].reserve(2 * sizeof(double));
    t.G = nullptr;
    t.lis = nullptr;
    t.U = 0;
    if (msr != MCG_SQRL2) csr_transforms(ds, pv, st);
    if (msr == MCG_JSD) {
        t.U = (uint32_t)ds.li_unique.size();
        REQUIRE(t.U <= (1u << 20), "JSD: too many distinct row sums");
        t.G = (double *)c.s[S_JSDG].reserve((uint64_t)t.U * sizeof(double));
        t.lis = ds.li_dev.as<const float>();
    }
    return true;
}

void generic_pp_next(Data &ds, int msr, const PPTables &t, const uint32_t *cent, uint32_t ci, const uint64_t *newc,
                     float *cc, float *dist, uint32_t *asn, const double *w, unsigned long long *mx,
                     hipStream_t st) {
    Cost1 P{};
    P.X = ds.csr ? nullptr : ds.X;
    P.val = ds.csr ? ds.val.as<const float>() : nullptr;
    P.idx = ds.csr ? ds.idx.as<const uint32_t>() : nullptr;
    P.ptr = ds.csr ? ds.ptr.as<const uint64_t>() : nullptr;
    P.d = (uint32_t)ds.d;
    P.pv = (float)t.pv;
    P.logpv = (float)log(t.pv > 0 ? t.pv : 1.);
    P.kpad = 1;
    P.CT = t.CT;
    P.cs = t.cs;
    if (msr != MCG_SQRL2) P.rt = ds.rowinfo.as<const float4>();
    if (msr == MCG_JSD) {
        if (t.U > PP_GMAX)
            hipLaunchKernelGGL(k_jsd_g, dim3((uint32_t)std::min<uint64_t>(t.U, 65536)), dim3(256), 0, st, t.CT, 1u, 1u,
                               (uint32_t)ds.d, t.lis, t.U, t.G);
        P.G = t.G;
        P.slot = ds.li_slot.as<const uint32_t>();
        hipLaunchKernelGGL(k_pp_cc_gen<MCG_JSD>, dim3((uint32_t)cdiv(ci, 4)), dim3(256), 0, st, P, cent, ci, cc);
    }
    if (ds.csr) pp_update_launch<8>(P, ds.n, msr, newc, ci, cc, dist, asn, 0, w, mx, st);
    else pp_update_launch<64>(P, ds.n, msr, newc, ci, cc, dist, asn, 0, w, mx, st);
}

// MCG_BREG_MFMA=0 keeps dense MKL / REVERSE_MKL on the row kernel (for comparisons)
static bool dense_breg_mfma_on() {
    static const bool on = [] { const char *e = getenv("MCG_BREG_MFMA"); return !(e && e[0] == '0'); }();
    return on;
}

void generic_costs(Data &ds, int msr, double prior, const float *C, uint64_t k, float *out_nk, uint32_t *asn,
                   float *best, hipStream_t st) {
    Ctx &c = *ds.ctx;
    const uint32_t d = (uint32_t)ds.d;
    const uint64_t n = ds.n;
    const float *X = ds.csr ? nullptr : ds.X;
    const float *val = ds.csr ? ds.val.as<const float>() : nullptr;
    const uint32_t *idx = ds.csr ? ds.idx.as<const uint32_t>() : nullptr;
    const uint64_t *ptr = ds.csr ? ds.ptr.as<const uint64_t>() : nullptr;
    REQUIRE(k < (1ull << 31), "k too large");
    if (!ds.csr && (msr == MCG_L1 || msr == MCG_L2) && k == 1) {
        if (msr == MCG_L2)
            hipLaunchKernelGGL(k_lp_single<true>, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, X, n, d, C, out_nk,
                               asn, best);
        else
            hipLaunchKernelGGL(k_lp_single<false>, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, X, n, d, C, out_nk,
                               asn, best);
        return;
    }
    if (!ds.csr && (msr == MCG_L1 || msr == MCG_L2)) {
        if (msr == MCG_L2)
            hipLaunchKernelGGL(k_tile_lp<true>, dim3((uint32_t)cdiv(n, 64)), dim3(256), 0, st, X, n, d, C, (uint32_t)k,
                               out_nk, asn, best);
        else
            hipLaunchKernelGGL(k_tile_lp<false>, dim3((uint32_t)cdiv(n, 64)), dim3(256), 0, st, X, n, d, C,
                               (uint32_t)k, out_nk, asn, best);
        return;
    }
    if (!msr_decomposed(msr)) {
        rowsums(ds, st);
        double *csum = (double *)c.s[S_CPARAM].reserve(k * sizeof(double));
        hipLaunchKernelGGL(k_center_sums, dim3((uint32_t)cdiv(k, 4)), dim3(256), 0, st, C, (uint32_t)k, d, csum);
        if (!ds.csr && msr == MCG_TVD) {
            hipLaunchKernelGGL(k_tile_tvd, dim3((uint32_t)cdiv(n, 64)), dim3(256), 0, st, X, n, d,
                               ds.rowsum.as<const double>(), C, (uint32_t)k, csum, (float)prior, out_nk, asn, best);
            return;
        }
        hipLaunchKernelGGL(k_pairs_more, dim3((uint32_t)cdiv(n, 4)), dim3(256), 0, st, msr, X, val, idx, ptr, n, d,
                           ds.rowsum.as<const double>(), C, (uint32_t)k, csum, (float)prior, out_nk, asn, best);
        return;
    }
    if (msr != MCG_SQRL2 && prior < DIRECT_PRIOR) {
        rowsums(ds, st);
        double *csum = (double *)c.s[S_CPARAM].reserve(k * sizeof(double));
        hipLaunchKernelGGL(k_center_sums, dim3((uint32_t)cdiv(k, 4)), dim3(256), 0, st, C, (uint32_t)k, d, csum);
        hipLaunchKernelGGL(k_pairs_direct, dim3((uint32_t)cdiv(n, 4)), dim3(256), 0, st, msr, X, val, idx, ptr, n, d,
                           ds.rowsum.as<const double>(), C, (uint32_t)k, csum, (float)prior, out_nk, asn, best);
        return;
    }
    const double pv = msr == MCG_SQRL2 ? 0. : prior;
    const bool single = k == 1 && !out_nk;   // k_row_cost1 reads a dense single-center table
    const uint32_t kpad = single ? 1u : (uint32_t)cdiv(k, 64) * 64;
    const Cost1 T = center_tables(ds, msr, pv, C, k, kpad, st);
    const float *CT = T.CT;
    const double *cs = T.cs, *G = T.G;
    const float4 *rt = T.rt;
    const uint32_t *slot = T.slot;
    const dim3 grid((uint32_t)cdiv(n, 4)), blk(256);
    // JSD rows that all share one li: the zero-baseline terms tabulated beside the table
    const float2 *CT2 = nullptr;
    if (msr == MCG_JSD && !single && !ds.li_unique.empty() && ds.li_unique.size() <= JSD_TAB_U &&
        (uint64_t)ds.li_unique.size() * d * kpad * sizeof(float2) <= (1ull << 30)) {   // tables up to 1 GiB
        const uint64_t nel = (uint64_t)d * kpad, U = ds.li_unique.size();
        float2 *t2 = (float2 *)c.s[S_JSDT].reserve(U * nel * sizeof(float2));
        hipLaunchKernelGGL(k_jsd_pair_table, dim3((uint32_t)std::min<uint64_t>(cdiv(U * nel, 256), 4096)), dim3(256), 0,
                           st, CT, nel, ds.li_dev.as<const float>(), (uint32_t)U, t2);
        CT2 = t2;
    }
    const float fpv = (float)pv, lpv = (float)log(pv > 0 ? pv : 1.);
    if (!single && !ds.csr && (msr == MCG_MKL || msr == MCG_REVERSE_MKL) && d <= 256 && dense_breg_mfma_on()) {
        const uint32_t dpad = d <= 64 ? 64u : d <= 128 ? 128u : 256u;
        float *Ct = (float *)c.s[S_CTM].reserve((uint64_t)kpad * dpad * sizeof(float));
        hipLaunchKernelGGL(k_ct_transpose, dim3((uint32_t)cdiv((uint64_t)kpad * dpad, 256)), dim3(256), 0, st, CT,
                           kpad, (uint32_t)k, d, dpad, Ct);
        const float *A = X;
        if (msr == MCG_MKL) {
            if (ds.tf_logx_pv != pv) {
                ds.tf_logx.reserve(n * d * sizeof(float));
                hipLaunchKernelGGL(k_log1p_div, dim3((uint32_t)std::min<uint64_t>(cdiv(n * d, 256), 8192)), dim3(256),
                                   0, st, X, n * d, fpv, ds.tf_logx.as<float>());
                ds.tf_logx_pv = pv;
            }
            A = ds.tf_logx.as<const float>();
        }
        const dim3 g((uint32_t)cdiv(n, 128));
        const uint32_t nt = kpad / 32;
#define BREG(M, D)                                                                                                 \
    hipLaunchKernelGGL((k_breg_mfma<M, D>), g, dim3(256), 0, st, A, d, n, fpv, lpv, rt, (const float *)Ct, cs,     \
                       (uint32_t)k, nt, out_nk, asn, best)
        hipEvent_t ev;
        c.tic(0, ev);
        if (msr == MCG_MKL) {
            if (dpad == 64) BREG(MCG_MKL, 64); else if (dpad == 128) BREG(MCG_MKL, 128); else BREG(MCG_MKL, 256);
        } else {
            if (dpad == 64) BREG(MCG_REVERSE_MKL, 64);
            else if (dpad == 128) BREG(MCG_REVERSE_MKL, 128);
            else BREG(MCG_REVERSE_MKL, 256);
        }
#undef BREG
        c.toc(0, ev);
        return;
    }
    if (!single && tiled_costs_ok(ds, msr, k) && tiled_costs_ready(ds, k, st)) {
        tiled_costs(ds, msr, CT, kpad, k, fpv, lpv, rt, cs, out_nk, asn, best, st);
        return;
    }
    if (single) {
        const Cost1 P{X, val, idx, ptr, d, fpv, lpv, rt, CT, kpad, cs, G, slot};
        switch (msr) {
        case MCG_SQRL2: hipLaunchKernelGGL(k_row_cost1<MCG_SQRL2>, grid, blk, 0, st, P, n, asn, best); break;
        case MCG_MKL: hipLaunchKernelGGL(k_row_cost1<MCG_MKL>, grid, blk, 0, st, P, n, asn, best); break;
        case MCG_REVERSE_MKL:
            hipLaunchKernelGGL(k_row_cost1<MCG_REVERSE_MKL>, grid, blk, 0, st, P, n, asn, best);
            break;
        default: hipLaunchKernelGGL(k_row_cost1<MCG_JSD>, grid, blk, 0, st, P, n, asn, best); break;
        }
        return;
    }
    switch (msr) {
    case MCG_SQRL2:
        hipLaunchKernelGGL(k_row_costs<MCG_SQRL2>, grid, blk, 0, st, X, val, idx, ptr, n, d, fpv, lpv, rt, CT, kpad,
                           (uint32_t)k, cs, G, slot, CT2, out_nk, asn, best);
        break;
    case MCG_MKL:
        hipLaunchKernelGGL(k_row_costs<MCG_MKL>, grid, blk, 0, st, X, val, idx, ptr, n, d, fpv, lpv, rt, CT, kpad,
                           (uint32_t)k, cs, G, slot, CT2, out_nk, asn, best);
        break;
    case MCG_REVERSE_MKL:
        hipLaunchKernelGGL(k_row_costs<MCG_REVERSE_MKL>, grid, blk, 0, st, X, val, idx, ptr, n, d, fpv, lpv, rt, CT,
                           kpad, (uint32_t)k, cs, G, slot, CT2, out_nk, asn, best);
        break;
    default:
        hipLaunchKernelGGL(k_row_costs<MCG_JSD>, grid, blk, 0, st, X, val, idx, ptr, n, d, fpv, lpv, rt, CT, kpad,
                           (uint32_t)k, cs, G, slot, CT2, out_nk, asn, best);
        break;
    }
}

}  // namespace mcg
