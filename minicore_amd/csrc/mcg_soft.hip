// Soft EM: responsibilities and the soft M-step.
//
// Replaces set_centroids_full_mean (soft) and correct_softmax
// (include/minicore/clustering/centroid.h:781-963): R = softmax(-temp * costs)
// per row (no max subtraction, as blaze's softmax; a row with a NaN becomes
// uniform over its argmin ties), cost = sum_i w_i <costs_i, R_i>, and centers
// = sum_i R_ic w_i x_i / sum_i R_ic w_i (dense) or sum_i R_ic x_i /
// sum_i R_ic w_i (CSR, the reference omits w in that numerator).
// All sums are FP64 in a fixed order (row splits added in order).
#include "mcg_internal.h"
#include <algorithm>
#include <cstring>
#include <cfloat>
#include <cmath>

namespace mcg {

// one wave per row
__global__ void k_softmax(const float *__restrict__ costs, uint64_t n, uint32_t k, float nt,
                          float *__restrict__ resp, const double *__restrict__ w, double *__restrict__ rowdot) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const float *cr = costs + row * k;
    float *r = resp + row * k;
    float s = 0.f;
    for (uint32_t c = lane; c < k; c += 64) { const float e = expf(cr[c] * nt); r[c] = e; s += e; }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    int bad = 0;
    for (uint32_t c = lane; c < k; c += 64) { const float v = r[c] / s; r[c] = v; bad |= (v != v); }
    for (int o = 32; o > 0; o >>= 1) bad |= __shfl_xor(bad, o);
    if (bad) {
        float m = FLT_MAX;
        for (uint32_t c = lane; c < k; c += 64) m = fminf(m, cr[c]);
        for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o));
        int ties = 0;
        for (uint32_t c = lane; c < k; c += 64) ties += (cr[c] == m);
        for (int o = 32; o > 0; o >>= 1) ties += __shfl_xor(ties, o);
        const float per = (float)(1. / (double)ties);
        for (uint32_t c = lane; c < k; c += 64) r[c] = (cr[c] == m) ? per : 0.f;
    }
    float dt = 0.f;
    for (uint32_t c = lane; c < k; c += 64) dt += cr[c] * r[c];
    for (int o = 32; o > 0; o >>= 1) dt += __shfl_xor(dt, o);
    if (lane == 0) rowdot[row] = (double)dt * (w ? w[row] : 1.);
}

void softmax_rows(const float *costs, uint64_t n, uint64_t k, float temp, float *resp, const double *w,
                  double *rowdot, hipStream_t st) {
    hipLaunchKernelGGL(k_softmax, dim3((uint32_t)cdiv(n, 4)), dim3(256), 0, st, costs, n, (uint32_t)k, -temp, resp, w,
                       rowdot);
}

// ---- dense soft M-step -----------------------------------------------------------
constexpr int SM_TC = 32, SM_TD = 64, SM_TR = 64;

__global__ __launch_bounds__(256) void k_soft_dense_part(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                         const float *__restrict__ R, uint32_t k,
                                                         const double *__restrict__ w, uint64_t rows_per,
                                                         double *__restrict__ part, double *__restrict__ dpart) {
    __shared__ double rs[SM_TR][SM_TC + 1];
    __shared__ float xs[SM_TR][SM_TD + 1];
    const uint32_t c0 = blockIdx.x * SM_TC, j0 = blockIdx.y * SM_TD;
    const uint64_t r0 = (uint64_t)blockIdx.z * rows_per, r1 = min<uint64_t>(n, r0 + rows_per);
    const int tid = threadIdx.x, cc = tid >> 3, jj = (tid & 7) * 8;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double den = 0.;
    for (uint64_t rb = r0; rb < r1; rb += SM_TR) {
        __syncthreads();
        for (int q = tid; q < SM_TR * SM_TC; q += 256) {
            const int r = q / SM_TC, c = q % SM_TC;
            const uint64_t i = rb + r;
            rs[r][c] = (i < r1 && c0 + c < k) ? (double)R[i * k + c0 + c] * (w ? w[i] : 1.) : 0.;
        }
        for (int q = tid; q < SM_TR * SM_TD; q += 256) {
            const int r = q / SM_TD, j = q % SM_TD;
            const uint64_t i = rb + r;
            xs[r][j] = (i < r1 && j0 + j < d) ? X[i * d + j0 + j] : 0.f;
        }
        __syncthreads();
        const int nr = (int)min<uint64_t>(SM_TR, r1 - rb);
        for (int r = 0; r < nr; ++r) {
            const double rw = rs[r][cc];
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[q] += rw * (double)xs[r][jj + q];
            den += rw;
        }
    }
    const uint32_t c = c0 + cc;
    if (c >= k) return;
    for (int q = 0; q < 8; ++q) {
        const uint32_t j = j0 + jj + q;
        if (j < d) part[((uint64_t)blockIdx.z * k + c) * d + j] = acc[q];
    }
    if (blockIdx.y == 0 && jj == 0) dpart[(uint64_t)blockIdx.z * k + c] = den;
}

__global__ void k_soft_dense_final(const double *__restrict__ part, const double *__restrict__ dpart, uint32_t S,
                                   uint32_t k, uint32_t d, float *__restrict__ C) {
    const uint32_t c = blockIdx.x;
    double den = 0.;
    for (uint32_t s = 0; s < S; ++s) den += dpart[(uint64_t)s * k + c];
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
        double num = 0.;
        for (uint32_t s = 0; s < S; ++s) num += part[((uint64_t)s * k + c) * d + j];
        C[(uint64_t)c * d + j] = (float)(num / den);
    }
}

// ---- CSR soft M-step (column-major walk: one wave per column) ---------------------
__global__ void k_soft_den_part(const float *__restrict__ R, uint64_t n, uint32_t k, const double *__restrict__ w,
                                uint64_t rows_per, double *__restrict__ dpart) {
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per, r1 = min<uint64_t>(n, r0 + rows_per);
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) {
        double s = 0.;
        for (uint64_t i = r0; i < r1; ++i) s += (double)R[i * k + c] * (w ? w[i] : 1.);
        dpart[(uint64_t)blockIdx.x * k + c] = s;
    }
}

__global__ void k_soft_den_final(const double *__restrict__ dpart, uint32_t S, uint32_t k, double *__restrict__ den) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    double s = 0.;
    for (uint32_t q = 0; q < S; ++q) s += dpart[(uint64_t)q * k + c];
    den[c] = s;
}

// One wave per column j; lane l holds the centers l, l + 64, ... of a group of up to
// 256 centers, so each stored entry's row of R is read once per group (coalesced
// 256-byte pieces) and its row index and value once.  Each (center, column) sum
// runs over the column's entries in order (FP64).
template <int Q>
__global__ __launch_bounds__(256) void k_soft_csc(const uint64_t *__restrict__ cptr, const uint32_t *__restrict__ crow,
                                                  const float *__restrict__ cval, uint32_t d,
                                                  const float *__restrict__ R, uint32_t k, uint32_t c0,
                                                  double *__restrict__ part, const double *__restrict__ den,
                                                  float *__restrict__ C) {
    const uint64_t j = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= d) return;
    // entry chunk z of the column (contiguous, in order)
    const uint64_t cb = cptr[j], ce = cptr[j + 1], Z = gridDim.y, chunk = (ce - cb + Z - 1) / Z;
    const uint64_t b = min(ce, cb + blockIdx.y * chunk), e = min(ce, b + chunk);
    double acc[Q];
    bool ok[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) { acc[q] = 0.; ok[q] = c0 + lane + 64u * q < k; }
    uint64_t p = b;
    // two entries per step: both rows' loads are issued before either is added
    for (; p + 1 < e; p += 2) {
        const uint32_t i0 = crow[p], i1 = crow[p + 1];
        const double x0 = (double)cval[p], x1 = (double)cval[p + 1];
        float r0[Q], r1[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint64_t c = c0 + lane + 64u * q;
            r0[q] = ok[q] ? R[(uint64_t)i0 * k + c] : 0.f;
            r1[q] = ok[q] ? R[(uint64_t)i1 * k + c] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            acc[q] += (double)r0[q] * x0;
            acc[q] += (double)r1[q] * x1;
        }
    }
    if (p < e) {
        const uint32_t i0 = crow[p];
        const double x0 = (double)cval[p];
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (ok[q]) acc[q] += (double)R[(uint64_t)i0 * k + c0 + lane + 64u * q] * x0;
    }
    // one chunk (C given): the sum is the numerator, divided here as the final pass would
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint64_t c = c0 + lane + 64u * q;
        if (!ok[q]) continue;
        if (C) C[c * d + j] = (float)(acc[q] / den[c]);
        else part[((uint64_t)blockIdx.y * k + c) * d + j] = acc[q];
    }
}

// the columns' chunk partials added in chunk order, over the center's denominator
__global__ void k_soft_csc_final(const double *__restrict__ part, uint32_t Z, uint32_t k, uint32_t d,
                                 const double *__restrict__ den, float *__restrict__ C) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)k * d) return;
    double s = part[e];
    for (uint32_t z = 1; z < Z; ++z) s += part[(uint64_t)z * k * d + e];
    C[e] = (float)(s / den[e / d]);
}

static void build_csc(Data &ds) {
    if (ds.csc_ptr.p) return;
    hipStream_t st = ds.ctx->stream;
    const uint64_t n = ds.n, d = ds.d, nnz = ds.nnz;
    std::vector<uint64_t> P(n + 1);
    std::vector<uint32_t> J(nnz);
    std::vector<float> V(nnz);
    copy_sync(P.data(), ds.ptr.p, (n + 1) * 8, hipMemcpyDeviceToHost, st);
    if (nnz) {
        copy_sync(J.data(), ds.idx.p, nnz * 4, hipMemcpyDeviceToHost, st);
        copy_sync(V.data(), ds.val.p, nnz * 4, hipMemcpyDeviceToHost, st);
    }
    std::vector<uint64_t> cp(d + 1, 0);
    for (uint64_t p = 0; p < nnz; ++p) cp[J[p] + 1]++;
    for (uint64_t j = 0; j < d; ++j) cp[j + 1] += cp[j];
    std::vector<uint64_t> cur(cp.begin(), cp.end() - 1);
    std::vector<uint32_t> rows(nnz);
    std::vector<float> cv(nnz);
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t p = P[i]; p < P[i + 1]; ++p) {
            const uint64_t q = cur[J[p]]++;
            rows[q] = (uint32_t)i;
            cv[q] = V[p];
        }
    ds.csc_ptr.reserve((d + 1) * 8);
    ds.csc_row.reserve(nnz * 4 + 4);
    ds.csc_val.reserve(nnz * 4 + 4);
    copy_sync(ds.csc_ptr.p, cp.data(), (d + 1) * 8, hipMemcpyHostToDevice, st);
    if (nnz) {
        copy_sync(ds.csc_row.p, rows.data(), nnz * 4, hipMemcpyHostToDevice, st);
        copy_sync(ds.csc_val.p, cv.data(), nnz * 4, hipMemcpyHostToDevice, st);
    }
}

void soft_mstep(Data &ds, const float *R, const double *w, uint64_t k, float *Cnew, hipStream_t st) {
    Ctx &c = *ds.ctx;
    const uint64_t n = ds.n;
    const uint32_t d = (uint32_t)ds.d;
    if (!ds.csr) {
        const uint32_t tc = (uint32_t)cdiv(k, SM_TC), td = (uint32_t)cdiv(d, SM_TD);
        uint32_t S = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cdiv(n, 1024), cdiv(2048, (uint64_t)tc * td)));
        const uint64_t rows_per = cdiv(n, S);
        S = (uint32_t)cdiv(n, rows_per);
        double *part = (double *)c.s[S_PART].reserve((uint64_t)S * k * d * 8);
        double *dpart = (double *)c.s[S_PARTW].reserve((uint64_t)S * k * 8);
        hipLaunchKernelGGL(k_soft_dense_part, dim3(tc, td, S), dim3(256), 0, st, ds.X, n, d, R, (uint32_t)k, w,
                           rows_per, part, dpart);
        hipLaunchKernelGGL(k_soft_dense_final, dim3((uint32_t)k), dim3(256), 0, st, part, dpart, S, (uint32_t)k, d,
                           Cnew);
        return;
    }
    build_csc(ds);
    const uint64_t rows_per = 4096;
    const uint32_t S = (uint32_t)cdiv(n, rows_per);
    double *dpart = (double *)c.s[S_PARTW].reserve((uint64_t)S * k * 8);
    double *den = (double *)c.s[S_PARTQ].reserve(k * 8);
    hipLaunchKernelGGL(k_soft_den_part, dim3(S), dim3(256), 0, st, R, n, (uint32_t)k, w, rows_per, dpart);
    hipLaunchKernelGGL(k_soft_den_final, dim3((uint32_t)cdiv(k, 256)), dim3(256), 0, st, dpart, S, (uint32_t)k, den);
    if (tiled_mstep_ok(ds, k) && tiled_mstep_ready(ds, k, ds.csc_ptr.as<const uint64_t>(),
                                                   ds.csc_row.as<const uint32_t>(), ds.csc_val.as<const float>(), st)) {
        tiled_soft_numer(ds, R, k, den, Cnew, st);
        return;
    }
    // up to 8 entry chunks per column, so the columns with many entries are spread
    // over waves (partial sums: Z x k x d FP64, at most about 2 GB); one chunk writes
    // the centers directly
    const uint64_t kd = k * (uint64_t)d;
    const uint32_t Z = csr_chunks(kd);
    double *part = Z > 1 ? (double *)c.s[S_PART].reserve((uint64_t)Z * kd * sizeof(double)) : nullptr;
    float *Cd = Z > 1 ? nullptr : Cnew;
    for (uint32_t c0 = 0; c0 < k; c0 += 256) {
        const uint32_t q = (uint32_t)std::min<uint64_t>(4, cdiv(k - c0, 64));
        const dim3 g((uint32_t)cdiv(d, 4), Z), blk(256);
        const uint64_t *cp = ds.csc_ptr.as<const uint64_t>();
        const uint32_t *cr = ds.csc_row.as<const uint32_t>();
        const float *cv = ds.csc_val.as<const float>();
        switch (q) {
        case 1: hipLaunchKernelGGL(k_soft_csc<1>, g, blk, 0, st, cp, cr, cv, d, R, (uint32_t)k, c0, part, den, Cd); break;
        case 2: hipLaunchKernelGGL(k_soft_csc<2>, g, blk, 0, st, cp, cr, cv, d, R, (uint32_t)k, c0, part, den, Cd); break;
        case 3: hipLaunchKernelGGL(k_soft_csc<3>, g, blk, 0, st, cp, cr, cv, d, R, (uint32_t)k, c0, part, den, Cd); break;
        default: hipLaunchKernelGGL(k_soft_csc<4>, g, blk, 0, st, cp, cr, cv, d, R, (uint32_t)k, c0, part, den, Cd); break;
        }
    }
    if (Z > 1)
        hipLaunchKernelGGL(k_soft_csc_final, dim3((uint32_t)cdiv(kd, 256)), dim3(256), 0, st, (const double *)part, Z,
                           (uint32_t)k, d, (const double *)den, Cnew);
}

// ---- soft L1 / L2 centers (set_centroids_full_mean, centroid.h:815-838) ----------
// Per center the weights are wc_i = w_i * r_ic (float) over ALL rows.
//  - L1 (dense rows; kmedian.h:266-301): weighted_median never adds up its weights,
//    so each coordinate is the value of the row of largest weight, the smallest
//    value among rows tied at that weight.  All-zero weights take its lower_bound
//    branch with mid = 0: the first value v >= 0 in sorted order and the next, u,
//    give v == 0 ? (v + u) / 2 : u (u = v when v is last; no value >= 0 gives 0,
//    where the reference reads past the column).
//  - L2 (blaze_adaptor.h:615-656 dense, csc.h:1131-1203 CSR): Weiszfeld from the
//    current center for all centers at once; a center stops when its cost is not
//    below the previous one (dense: sum wc_i * dist_i rounded to float, distances
//    floored at 1e-30f; CSR: the unweighted sum of distances, floored at 1e-80, in
//    double), after at most 99999 updates.  The update sum_i q_i x_i / sum_i q_i,
//    q_i = wc_i / dist_i, runs as the soft M-step with q_i / max_c q as the
//    responsibilities (FP64 sums in a fixed order).

__device__ __forceinline__ uint32_t fkey(float v) {   // order-preserving float -> u32
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_value(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
__device__ __forceinline__ float soft_wc(const double *w, const float *R, uint64_t i, uint32_t k, uint32_t c) {
    return (w ? (float)w[i] : 1.f) * R[i * k + c];
}

// largest key of wc_ic per center (block: rows [r0, r1), threads over centers)
__global__ __launch_bounds__(256) void k_soft_wmax(const float *__restrict__ R, const double *__restrict__ w,
                                                   uint64_t n, uint32_t k, uint64_t rows_per,
                                                   uint32_t *__restrict__ mx) {
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per, r1 = min<uint64_t>(n, r0 + rows_per);
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) {
        uint32_t m = 0;
        for (uint64_t i = r0; i < r1; ++i) m = max(m, fkey(soft_wc(w, R, i, k, c)));
        atomicMax(mx + c, m);
    }
}

// rows at the largest weight: their number and the first of them, per center
__global__ __launch_bounds__(256) void k_soft_ties(const float *__restrict__ R, const double *__restrict__ w,
                                                   uint64_t n, uint32_t k, uint64_t rows_per,
                                                   const uint32_t *__restrict__ mx, uint32_t *__restrict__ cnt,
                                                   uint32_t *__restrict__ first) {
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per, r1 = min<uint64_t>(n, r0 + rows_per);
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) {
        const uint32_t m = mx[c];
        uint32_t t = 0, f = 0xFFFFFFFFu;
        for (uint64_t i = r0; i < r1; ++i)
            if (fkey(soft_wc(w, R, i, k, c)) == m) { if (!t) f = (uint32_t)i; ++t; }
        if (t) { atomicAdd(cnt + c, t); atomicMin(first + c, f); }
    }
}

// columnwise minimum (as keys) over the rows tied at the largest weight, for the
// centers in list[] (grid.x over them, grid.y over row chunks, threads over columns)
__global__ __launch_bounds__(256) void k_soft_tie_min(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                      const float *__restrict__ R, const double *__restrict__ w,
                                                      uint32_t k, const uint32_t *__restrict__ mx,
                                                      const uint32_t *__restrict__ list, uint64_t rows_per,
                                                      uint32_t *__restrict__ ck) {
    const uint32_t c = list[blockIdx.x];
    const uint64_t r0 = (uint64_t)blockIdx.y * rows_per, r1 = min<uint64_t>(n, r0 + rows_per);
    const uint32_t m = mx[c];
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
        uint32_t v = 0xFFFFFFFFu;
        for (uint64_t i = r0; i < r1; ++i)
            if (fkey(soft_wc(w, R, i, k, c)) == m) v = min(v, fkey(X[i * d + j]));
        if (v != 0xFFFFFFFFu) atomicMin(ck + (uint64_t)c * d + j, v);
    }
}

__global__ void k_soft_tie_store(const uint32_t *__restrict__ ck, const uint32_t *__restrict__ list, uint32_t nl,
                                 uint32_t d, float *__restrict__ C) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)nl * d) return;
    const uint32_t c = list[e / d], j = (uint32_t)(e % d);
    C[(uint64_t)c * d + j] = fkey_value(ck[(uint64_t)c * d + j]);
}

// the all-zero-weight rule per column (one thread per column), written to the
// centers in list[]
__global__ void k_soft_zero_rule(const float *__restrict__ X, uint64_t n, uint32_t d,
                                 const uint32_t *__restrict__ list, uint32_t nl, float *__restrict__ C) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d) return;
    float v1 = FLT_MAX, v2 = FLT_MAX;
    uint64_t cnt = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const float x = X[i * d + j];
        if (!(x >= 0.f)) continue;
        ++cnt;
        if (x < v1) { v2 = v1; v1 = x; } else if (x < v2) v2 = x;
    }
    float r = 0.f;
    if (cnt) {
        if (cnt == 1) v2 = v1;
        r = v1 == 0.f ? (float)(.5 * ((double)v1 + (double)v2)) : v2;
    }
    for (uint32_t q = 0; q < nl; ++q) C[(uint64_t)list[q] * d + j] = r;
}

// Weiszfeld terms per center over a chunk of rows: the stopping cost (weighted or
// not) into part[chunk][c], the largest q = wc / dist into qmax[c] (bit pattern)
__global__ __launch_bounds__(256) void k_geo_soft_terms(const float *__restrict__ K, const float *__restrict__ R,
                                                        const double *__restrict__ w, uint64_t n, uint32_t k,
                                                        double floor_d, int weighted, uint64_t rows_per,
                                                        double *__restrict__ part,
                                                        unsigned long long *__restrict__ qmax) {
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per, r1 = min<uint64_t>(n, r0 + rows_per);
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) {
        double cs = 0., qm = 0.;
        for (uint64_t i = r0; i < r1; ++i) {
            const float dd = K[i * k + c];
            const double v = (dd != dd || (double)dd < floor_d) ? floor_d : (double)dd;
            const float wc = soft_wc(w, R, i, k, c);
            cs += weighted ? (double)wc * v : v;
            qm = fmax(qm, (double)wc / v);
        }
        part[(uint64_t)blockIdx.x * k + c] = cs;
        atomicMax(qmax + c, (unsigned long long)__double_as_longlong(qm));
    }
}

__global__ void k_sum_parts(const double *__restrict__ part, uint32_t S, uint32_t k, double *__restrict__ out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    double s = 0.;
    for (uint32_t b = 0; b < S; ++b) s += part[(uint64_t)b * k + c];
    out[c] = s;
}

// K (distances) -> Q = (wc / dist) / qmax_c for the centers still moving, else 0
__global__ void k_geo_soft_q(float *__restrict__ K, const float *__restrict__ R, const double *__restrict__ w,
                             uint64_t n, uint32_t k, double floor_d, const unsigned long long *__restrict__ qmax,
                             const uint32_t *__restrict__ alive) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * k;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / k;
        const uint32_t c = (uint32_t)(e % k);
        float q = 0.f;
        if (alive[c]) {
            const float dd = K[e];
            const double v = (dd != dd || (double)dd < floor_d) ? floor_d : (double)dd;
            q = (float)(((double)soft_wc(w, R, i, k, c) / v) / __longlong_as_double((long long)qmax[c]));
        }
        K[e] = q;
    }
}

__global__ void k_copy_alive(const float *__restrict__ Cn, const uint32_t *__restrict__ alive, uint32_t k, uint32_t d,
                             float *__restrict__ C) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)k * d) return;
    if (alive[e / d]) C[e] = Cn[e];
}

void soft_median_mstep(Data &ds, int msr, const float *R, const double *w, uint64_t k, float *Cc, float *K,
                       hipStream_t st) {
    Ctx &c = *ds.ctx;
    const uint64_t n = ds.n;
    const uint32_t d = (uint32_t)ds.d, kk = (uint32_t)k;
    const uint32_t S = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 1024);
    const uint64_t rows_per = cdiv(n, S);
    const uint32_t kg = (uint32_t)cdiv(k, 256);
    if (msr == MCG_L1) {
        if (ds.csr) MCG_THROW(MCG_EUNSUPPORTED, "Not implemented: weighted L1 median for CSparseMatrix");
        uint32_t *mx = (uint32_t *)c.s[S_MED0].reserve(k * 16);
        uint32_t *cnt = mx + k, *first = cnt + k, *list = first + k;
        fill_u32(mx, 0u, k, st);
        fill_u32(cnt, 0u, k, st);
        fill_u32(first, 0xFFFFFFFFu, k, st);
        hipLaunchKernelGGL(k_soft_wmax, dim3(S), dim3(256), 0, st, R, w, n, kk, rows_per, mx);
        hipLaunchKernelGGL(k_soft_ties, dim3(S), dim3(256), 0, st, R, w, n, kk, rows_per, (const uint32_t *)mx, cnt,
                           first);
        std::vector<uint32_t> h(3 * k);
        copy_sync(h.data(), mx, 3 * k * 4, hipMemcpyDeviceToHost, st);
        std::vector<uint64_t> rows;
        std::vector<uint32_t> ids, tied, zero;
        for (uint32_t q = 0; q < kk; ++q) {
            const uint32_t key = h[q];
            const bool pos = (key & 0x80000000u) && (key & 0x7fffffffu);   // largest weight > 0
            if (!pos) zero.push_back(q);
            else if (h[k + q] == 1) { rows.push_back(h[2 * k + q]); ids.push_back(q); }
            else tied.push_back(q);
        }
        if (!rows.empty()) copy_rows_to_centers(ds, rows.data(), ids.data(), rows.size(), Cc, st);
        if (!tied.empty()) {
            uint32_t *ck = (uint32_t *)c.s[S_MED1].reserve(k * d * 4);
            HIP_CHECK(hipMemcpyAsync(list, tied.data(), tied.size() * 4, hipMemcpyHostToDevice, st));
            for (uint32_t q : tied) fill_u32(ck + (uint64_t)q * d, 0xFFFFFFFFu, d, st);
            hipLaunchKernelGGL(k_soft_tie_min, dim3((uint32_t)tied.size(), S), dim3(256), 0, st, ds.X, n, d, R, w, kk,
                               (const uint32_t *)mx, (const uint32_t *)list, rows_per, ck);
            hipLaunchKernelGGL(k_soft_tie_store, dim3((uint32_t)cdiv(tied.size() * d, 256)), dim3(256), 0, st,
                               (const uint32_t *)ck, (const uint32_t *)list, (uint32_t)tied.size(), d, Cc);
        }
        if (!zero.empty()) {
            uint32_t *zl = (uint32_t *)c.s[S_MED2].reserve(k * 4);
            HIP_CHECK(hipMemcpyAsync(zl, zero.data(), zero.size() * 4, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_soft_zero_rule, dim3((uint32_t)cdiv(d, 64)), dim3(64), 0, st, ds.X, n, d,
                               (const uint32_t *)zl, (uint32_t)zero.size(), Cc);
        }
        HIP_CHECK(hipStreamSynchronize(st));   // the host lists above live until here
        return;
    }
    // L2: Weiszfeld for all centers at once
    if (!ds.csr && n == 1) {   // one row: that row (blaze_adaptor.h:618)
        std::vector<uint64_t> rows(k, 0);
        std::vector<uint32_t> ids(k);
        for (uint32_t q = 0; q < kk; ++q) ids[q] = q;
        copy_rows_to_centers(ds, rows.data(), ids.data(), k, Cc, st);
        HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    const bool csr = ds.csr;
    const double floor_d = csr ? 1e-80 : (double)1e-30f;
    char *m0 = (char *)c.s[S_MED0].reserve(k * 8 * 3 + (uint64_t)S * k * 8);
    double *cost = (double *)m0;
    unsigned long long *qmax = (unsigned long long *)(m0 + k * 8);
    uint32_t *alive_d = (uint32_t *)(m0 + k * 16);
    double *part = (double *)(m0 + k * 24);
    float *Cn = (float *)c.s[S_CENT_C].reserve(k * d * 4);
    std::vector<double> prev(k, csr ? DBL_MAX : (double)FLT_MAX);
    std::vector<uint32_t> alive(k, 1u);
    std::vector<uint64_t> hb(2 * k);
    const uint32_t eg = (uint32_t)std::min<uint64_t>(cdiv(n * k, 256), 4096);
    for (uint64_t it = 0;;) {
        generic_costs(ds, MCG_L2, 0., Cc, k, K, nullptr, nullptr, st);
        fill_u32((uint32_t *)qmax, 0u, 2 * k, st);
        hipLaunchKernelGGL(k_geo_soft_terms, dim3(S), dim3(256), 0, st, (const float *)K, R, w, n, kk, floor_d,
                           csr ? 0 : 1, rows_per, part, qmax);
        hipLaunchKernelGGL(k_sum_parts, dim3(kg), dim3(256), 0, st, (const double *)part, S, kk, cost);
        copy_sync(hb.data(), cost, 2 * k * 8, hipMemcpyDeviceToHost, st);
        bool any = false;
        for (uint32_t q = 0; q < kk; ++q) {
            if (!alive[q]) continue;
            double cur;
            std::memcpy(&cur, &hb[q], 8);
            if (!csr) cur = (double)(float)cur;
            if (!(cur < prev[q])) { alive[q] = 0; continue; }
            prev[q] = cur;
            any = true;
            double qm;
            std::memcpy(&qm, &hb[k + q], 8);
            if (!(qm > 0.))
                MCG_THROW(MCG_ENAN, "Optimization failed: nan (geometric median of a center whose weights are all zero)");
        }
        if (!any || ++it == 100000) break;
        HIP_CHECK(hipMemcpyAsync(alive_d, alive.data(), k * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_geo_soft_q, dim3(eg), dim3(256), 0, st, K, R, w, n, kk, floor_d,
                           (const unsigned long long *)qmax, (const uint32_t *)alive_d);
        soft_mstep(ds, K, nullptr, k, Cn, st);
        hipLaunchKernelGGL(k_copy_alive, dim3((uint32_t)cdiv(k * d, 256)), dim3(256), 0, st, (const float *)Cn,
                           (const uint32_t *)alive_d, kk, d, Cc);
        HIP_CHECK(hipStreamSynchronize(st));   // alive (host) is rewritten next round
    }
}

}  // namespace mcg
