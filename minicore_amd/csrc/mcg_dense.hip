// Dense squared-L2 hard assignment for gfx950.
//
// Replaces the n x k loop of msr_with_prior(SQRL2) calls in assign_points_hard
// (include/minicore/clustering/solve.h:234-285; per pair
// include/minicore/dist/applicator.h:1680-1685).
//
// Three forms:
//  * BF16: ||x-c||^2 = ||x-mu||^2 + ||c-mu||^2 - 2 (x-mu).(c-mu) with the cross term on
//    v_mfma_f32_32x32x16_bf16 (FP32 accumulate).  Points are the MFMA B operand and stay
//    in registers for the whole pass over the centers; centers stream through LDS in
//    tiles of 64.  Each lane keeps a running best/second/third of its column with the
//    center's position in its 32-center sub-tile packed into the low 5 mantissa bits
//    (4 VALU ops per value: v_and_or, two v_med3_u32, v_min_u32).  Points whose
//    best/second gap is not provably larger than the BF16 error bound get their
//    candidate centers re-scored exactly (k_collect16, k_exact_cands, k_refine_final).
//  * FP16: the BF16 form with FP16 operands, power-of-two scaled into FP16's range
//    (the rows by xscale, the centers by a per-E-step scale): 3 more mantissa bits,
//    so an 8x smaller rounding error and far fewer near ties to re-score.  The
//    rounding error that goes into the tie threshold is measured per row / center
//    on the operands actually used (below FP16's normal range they are flushed to 0,
//    so the bound does not depend on how the matrix cores treat denormals).
//  * FP32: the same GEMM form on v_mfma_f32_32x32x2f32, plain compare/select.
//  * EXACT: sum of squared differences per pair as a sequential FMA chain, FP32.
#include "mcg_internal.h"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace mcg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// FP16 bits of an already scaled value: RNE, saturated to +-65504, 0 below the normal
// range (2^-14); NaN stays NaN
__device__ __forceinline__ uint16_t f2h_sat(float v) {
    const float a = fabsf(v);
    if (a != a) return 0x7E00u;
    if (a < 0x1p-14f) return 0u;
    const _Float16 h = (_Float16)copysignf(fminf(a, 65504.f), v);
    return __builtin_bit_cast(uint16_t, h);
}

__device__ __forceinline__ float h2f(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// largest power of two s with s * m <= lim (1 for m = 0), within [2^-100, 2^100]
__device__ __forceinline__ double pow2_scale_d(double m, double lim) {
    if (!(m > 0.)) return 1.;
    int e;
    frexp(lim / m, &e);
    return ldexp(1., min(100, max(-100, e - 1)));
}

// the 16-bit operand vector and its 32x32x16 MFMA
template <bool F16> struct Op16 { typedef bf16x8 V; };
template <> struct Op16<true> { typedef half8 V; };
template <bool F16>
__device__ __forceinline__ floatx16 mfma16(const typename Op16<F16>::V &a, const typename Op16<F16>::V &b,
                                           const floatx16 &c) {
    if constexpr (F16) return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Dataset preparation: column mean mu (FP64 sums), BF16 copy of X - mu padded
// to dpad columns, ||x - mu|| per row (rounded up), max over rows.
// ---------------------------------------------------------------------------
__global__ void k_colsum_partial(const float *__restrict__ X, uint64_t n, uint32_t d,
                                 uint32_t rows_per, double *__restrict__ part) {
    const uint64_t r0 = (uint64_t)blockIdx.y * rows_per;
    const uint64_t r1 = min<uint64_t>(n, r0 + rows_per);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < d; j += gridDim.x * blockDim.x) {
        // four rows in flight; the four running sums are added in a fixed order
        double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
        uint64_t r = r0;
        // non-finite values are left out: mu only centres the 16-bit operands (any
        // finite shift is exact for the bounds), and one such value must not make every
        // row's operand non-finite
        auto fin = [](float v) { return fabsf(v) <= FLT_MAX ? (double)v : 0.; };
        for (; r + 4 <= r1; r += 4) {
            s0 += fin(X[r * d + j]);
            s1 += fin(X[(r + 1) * d + j]);
            s2 += fin(X[(r + 2) * d + j]);
            s3 += fin(X[(r + 3) * d + j]);
        }
        for (; r < r1; ++r) s0 += fin(X[r * d + j]);
        part[(uint64_t)blockIdx.y * d + j] = (s0 + s1) + (s2 + s3);
    }
}

// one block per column: the parts summed by 256 threads, then a fixed tree
__global__ void k_colsum_final(const double *__restrict__ part, uint32_t nparts, uint32_t d,
                               uint64_t n, float *__restrict__ mu) {
    __shared__ double sm[256];
    const uint32_t j = blockIdx.x;
    double s = 0.;
    for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) s += part[(uint64_t)p * d + j];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) mu[j] = (float)(sm[0] / (double)n);
}

// one wave per row: xnorm = ||x - mu||, rounded up
__global__ void k_row_norms(const float *__restrict__ X, uint64_t n, uint32_t d, const float *__restrict__ mu,
                            float *__restrict__ xnorm) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double s = 0.;
    for (uint32_t j = lane; j < d; j += 64) {
        const float v = X[row * d + j] - mu[j];
        s += (double)v * (double)v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) xnorm[row] = (float)(sqrt(s) * (1. + 1e-6));
}

// one wave per row; lane l converts columns [4l, 4l + 4) of each 256-column pass.
// F16: the operand is f2h_sat(scale * (x - mu)), else bf16(x - mu) (scale 1);
// xdev = ||operand / scale - (x - mu)|| (each difference is exact in FP32), rounded up.
template <bool F16>
__global__ void k_make_x16(const float *__restrict__ X, uint64_t n, uint32_t d, uint32_t dpad,
                           const float *__restrict__ mu, const float *__restrict__ scale_p, uint16_t *__restrict__ X16,
                           float *__restrict__ xdev) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const float scale = F16 ? *scale_p : 1.f;
    const float inv = 1.f / scale;
    double e = 0.;
    for (uint32_t j0 = lane * 4; j0 < dpad; j0 += 256) {
        uint16_t bv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t j = j0 + i;
            const float v = j < d ? X[row * d + j] - mu[j] : 0.f;
            bv[i] = F16 ? f2h_sat(v * scale) : f2bf_rne(v);
            const double dv = (F16 ? (double)h2f(bv[i]) * (double)inv : (double)bf2f(bv[i])) - (double)v;
            e += dv * dv;
        }
        const uint2 packed = make_uint2((uint32_t)bv[0] | ((uint32_t)bv[1] << 16), (uint32_t)bv[2] | ((uint32_t)bv[3] << 16));
        *(uint2 *)(X16 + row * dpad + j0) = packed;
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
    if (lane == 0) xdev[row] = (float)(sqrt(e) * (1. + 1e-6));
}

// max of the finite non-negative floats (their bit patterns order like their values)
__global__ void k_max_f32(const float *__restrict__ v, uint64_t n, unsigned int *__restrict__ out_bits) {
    __shared__ float sm[256];
    float m = 0.f;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (v[i] <= FLT_MAX) m = fmaxf(m, v[i]);   // rows with non-finite values are left out
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out_bits, __float_as_uint(sm[0]));
}

static uint32_t pad_dims(uint64_t d) {
    if (d <= 64) return 64;
    if (d <= 128) return 128;
    if (d <= 256) return 256;
    return 512;
}

void prepare_base(Data &ds) {
    if (ds.base_ready) return;
    Ctx &c = *ds.ctx;
    hipStream_t st = c.stream;
    const uint64_t n = ds.n;
    const uint32_t d = (uint32_t)ds.d;
    REQUIRE(d <= 512, "16-bit MFMA forms support d <= 512");
    ds.dpad = pad_dims(d);
    const uint32_t rows_per = 1024;
    const uint32_t nparts = (uint32_t)cdiv(n, rows_per);
    // scratch of the context (a freed buffer would synchronise the device)
    DBuf &part = c.s[S_PREP];
    part.reserve((size_t)nparts * d * sizeof(double) + 256);
    ds.mu.reserve(d * sizeof(float));
    dim3 g1((uint32_t)cdiv(d, 256), nparts);
    hipLaunchKernelGGL(k_colsum_partial, g1, dim3(256), 0, st, ds.X, n, d, rows_per, part.as<double>());
    hipLaunchKernelGGL(k_colsum_final, dim3(d), dim3(256), 0, st, part.as<double>(), nparts, d, n,
                       ds.mu.as<float>());
    ds.xnorm.reserve(n * sizeof(float));
    hipLaunchKernelGGL(k_row_norms, dim3((uint32_t)cdiv(n, 4)), dim3(256), 0, st, ds.X, n, d, ds.mu.as<const float>(),
                       ds.xnorm.as<float>());
    // xmax stays on the device (no host round trip): the center preparation reads it
    ds.xstat.reserve(2 * sizeof(float));
    fill_u32(ds.xstat.as<uint32_t>(), 0u, 2, st);
    hipLaunchKernelGGL(k_max_f32, dim3((uint32_t)std::min<uint64_t>(cdiv(n, 256), 2048)), dim3(256), 0, st,
                       ds.xnorm.as<const float>(), n, ds.xstat.as<unsigned int>());
    ds.base_ready = true;
}

void prepare_bf16(Data &ds) {
    if (ds.bf16_ready) return;
    prepare_base(ds);
    Ctx &c = *ds.ctx;
    hipStream_t st = c.stream;
    ds.Xb.reserve((size_t)ds.n * ds.dpad * sizeof(uint16_t));
    ds.xdev.reserve(ds.n * sizeof(float));
    hipLaunchKernelGGL(k_make_x16<false>, dim3((uint32_t)cdiv(ds.n, 4)), dim3(256), 0, st, ds.X, ds.n, (uint32_t)ds.d,
                       ds.dpad, ds.mu.as<const float>(), (const float *)nullptr, ds.Xb.as<uint16_t>(),
                       ds.xdev.as<float>());
    ds.bf16_ready = true;
}

// xstat[1] = the largest power of two s with s * xmax <= 2^14: every |x_j - mu_j| <=
// ||x - mu|| <= xmax, so the scaled rows stay inside FP16's range
__global__ void k_xscale(float *__restrict__ xstat) { xstat[1] = (float)pow2_scale_d(xstat[0], 16384.); }

void prepare_f16(Data &ds) {
    if (ds.f16_ready) return;
    prepare_base(ds);
    Ctx &c = *ds.ctx;
    hipStream_t st = c.stream;
    hipLaunchKernelGGL(k_xscale, dim3(1), dim3(1), 0, st, ds.xstat.as<float>());
    ds.Xh.reserve((size_t)ds.n * ds.dpad * sizeof(uint16_t));
    ds.xdev_h.reserve(ds.n * sizeof(float));
    hipLaunchKernelGGL(k_make_x16<true>, dim3((uint32_t)cdiv(ds.n, 4)), dim3(256), 0, st, ds.X, ds.n, (uint32_t)ds.d,
                       ds.dpad, ds.mu.as<const float>(), ds.xstat.as<const float>() + 1, ds.Xh.as<uint16_t>(),
                       ds.xdev_h.as<float>());
    ds.f16_ready = true;
}

// ---------------------------------------------------------------------------
// Center preparation (per E-step).
// ---------------------------------------------------------------------------
// one wave per (padded) center: Cb = bf16(-2 (c - mu)), cn = ||c - mu||^2 (FP64)
// One wave per (padded) center: Cb = bf16(-2 (c - mu)), cn = ||c - mu||^2 (FP64),
// cinfo = {||c - mu||, ||bf16(c - mu) - (c - mu)||}, both rounded up (bf16(-2v) is
// exactly -2 bf16(v), so the MFMA operand's error is that of bf16(v)).
__global__ void k_prep_centers_bf16(const float *__restrict__ C, uint64_t k, uint64_t kpad, uint32_t d,
                                    uint32_t dpad, const float *__restrict__ mu,
                                    uint16_t *__restrict__ Cb, double *__restrict__ cn, float2 *__restrict__ cinfo) {
    const uint64_t c = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= kpad) return;
    double s = 0., e = 0.;
    for (uint32_t j = lane; j < dpad; j += 64) {
        const float v = (c < k && j < d) ? C[c * d + j] - mu[j] : 0.f;
        const uint16_t bv = f2bf_rne(-2.f * v);
        Cb[c * dpad + j] = bv;
        s += (double)v * (double)v;
        const double dv = -0.5 * (double)bf2f(bv) - (double)v;
        e += dv * dv;
    }
    for (int o = 32; o > 0; o >>= 1) { s += __shfl_xor(s, o); e += __shfl_xor(e, o); }
    if (lane == 0) {
        cn[c] = s;
        cinfo[c] = make_float2((float)(sqrt(s) * (1. + 1e-6)), (float)(sqrt(e) * (1. + 1e-6)));
    }
}

// FP16 form, step 1 (one wave per padded center): cn = ||c - mu||^2 (FP64)
__global__ void k_center_norms(const float *__restrict__ C, uint64_t k, uint64_t kpad, uint32_t d,
                               const float *__restrict__ mu, double *__restrict__ cn) {
    const uint64_t c = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= kpad) return;
    double s = 0.;
    for (uint32_t j = lane; j < d; j += 64) {
        const float v = c < k ? C[c * d + j] - mu[j] : 0.f;
        s += (double)v * (double)v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) cn[c] = s;
}

// FP16 form, step 3 (one wave per padded center, blocks of 4 waves, after
// k_prep_params): the operand Ch = f2h_sat(-2 csc (c - mu)), cinfo = {||c - mu||,
// ||Ch / (-2 csc) - (c - mu)||}, both rounded up, and dbmax = max over the real
// centers of the latter (one atomic per block)
__global__ __launch_bounds__(256) void k_prep_centers_h(const float *__restrict__ C, uint64_t k, uint64_t kpad,
                                                        uint32_t d, uint32_t dpad, const float *__restrict__ mu,
                                                        EParams *__restrict__ prm, uint16_t *__restrict__ Ch,
                                                        float2 *__restrict__ cinfo) {
    const uint64_t c = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const bool live = c < kpad;   // every wave reaches the block barrier below
    const float sc = prm->csc;
    const double inv = -0.5 / (double)sc;
    double s = 0., e = 0.;
    for (uint32_t j = lane; live && j < dpad; j += 64) {
        const float v = (c < k && j < d) ? C[c * d + j] - mu[j] : 0.f;
        const uint16_t hv = f2h_sat(-2.f * sc * v);
        Ch[c * dpad + j] = hv;
        s += (double)v * (double)v;
        const double dv = (double)h2f(hv) * inv - (double)v;
        e += dv * dv;
    }
    for (int o = 32; o > 0; o >>= 1) { s += __shfl_xor(s, o); e += __shfl_xor(e, o); }
    __shared__ float dmax[4];
    const float dev = (float)(sqrt(e) * (1. + 1e-6));
    if (lane == 0) {
        if (live) cinfo[c] = make_float2((float)(sqrt(s) * (1. + 1e-6)), dev);
        dmax[threadIdx.x >> 6] = c < k ? dev : 0.f;
    }
    __syncthreads();
    // one atomic per block; non-negative floats order like their bit patterns
    if (threadIdx.x == 0)
        atomicMax((unsigned int *)&prm->dbmax, __float_as_uint(fmaxf(fmaxf(dmax[0], dmax[1]), fmaxf(dmax[2], dmax[3]))));
}

// single block: cmax, max center rounding error, offset, near-tie coefficients,
// cadd[c] = S (off + cn[c]).  BF16 (f16 = 0): S = 1 and the center errors come from
// cinfo; FP16: the center scale csc (2 csc cmax <= 2^15, and S times the largest
// approximate value stays below 2^100), S = xscale * csc, and dbmax is left at 0 for
// k_prep_centers_h.
__global__ void k_prep_params(const double *__restrict__ cn, const float2 *__restrict__ cinfo, uint64_t k,
                              uint64_t kpad, const float *__restrict__ xstat, uint32_t d, int f16,
                              float *__restrict__ cadd, EParams *__restrict__ prm, uint32_t *__restrict__ rcnt) {
    __shared__ double sm[256], se[256];
    if (threadIdx.x < 8) rcnt[threadIdx.x] = 0u;   // refinement counters of the E-step that follows
    const float xmax = xstat[0], xscale = f16 ? xstat[1] : 1.f;
    double m = 0., me = 0.;
    for (uint64_t i = threadIdx.x; i < k; i += blockDim.x) {
        if (cn[i] <= DBL_MAX) m = fmax(m, cn[i]);   // non-finite centers never win (cadd below)
        if (!f16 && cn[i] <= DBL_MAX) me = fmax(me, (double)cinfo[i].y);
    }
    sm[threadIdx.x] = m;
    se[threadIdx.x] = me;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
            se[threadIdx.x] = fmax(se[threadIdx.x], se[threadIdx.x + s]);
        }
        __syncthreads();
    }
    const double cmax = sqrt(sm[0]) * (1. + 1e-6);
    const double off = 2.25 * (double)xmax * cmax + 1.;
    // Largest approximate value: off + cmax^2 + 2.1 xmax cmax.
    const double vmax = off + cmax * cmax + 2.1 * (double)xmax * cmax;
    const double e_abs = 400. * vmax * 0x1p-24        // accumulation, offset, packing
                         + 4. * 0x1p-24 * ((double)xmax + cmax) * ((double)xmax + cmax);
    double csc = 1., S = 1.;
    if (f16) {
        csc = fmin(pow2_scale_d(cmax, 16384.), pow2_scale_d((double)xscale * vmax, 0x1p100));
        S = (double)xscale * csc;
    }
    if (threadIdx.x == 0) {
        EParams p;
        p.off = (float)off;
        p.cmax = (float)cmax;
        p.csc = (float)csc;
        p.tau0 = (float)(2. * e_abs);
        p.taur = (float)(1e-4 + 4. * (double)d * 0x1p-24);
        p.dbmax = f16 ? 0.f : (float)se[0];
        p.s = (float)S;
        p.inv_s = (float)(1. / S);
        // a NaN in center 0 makes every row's first cost NaN, which the reference's scan keeps
        p.c0nan = k > 0 && cn[0] != cn[0];
        p.pad[0] = p.pad[1] = p.pad[2] = 0;
        *prm = p;
    }
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < kpad; i += blockDim.x)
        cadd[i] = (i < k && cn[i] <= DBL_MAX) ? (float)(S * (off + cn[i])) : 3.0e38f;
}

// ---------------------------------------------------------------------------
// BF16 MFMA E-step.  KS = dpad / 16 k-steps; NB n-blocks of 32 points per wave.
// K-permutation: lane half h supplies dims [h*dpad/2 + 8s, +8) at step s for both
// operands, so each lane reads contiguous 16-byte chunks.
//
// Center tiles (64 rows) go global -> LDS directly (global_load_lds_dwordx4), with
// the 16-byte chunk index XOR-swizzled by the row so that the 16 rows read by one
// ds_read_b128 phase hit distinct banks.
//
// Main pass: each lane keeps the three smallest packed values of
// its point column and the tiles of the best two; the halves are merged at the end
// and every point is classified:
//   certain : second - best > tau                 -> asn = best
//   pair    : third - best > tau >= second - best -> candidates {best, second}
//   full    : otherwise                           -> listed with threshold best + tau
// The collect pass (k_collect16) then gathers, for each listed point, every center
// whose value is within the threshold; all candidates are scored exactly.
// ---------------------------------------------------------------------------
struct Cand {
    uint32_t p, c;
};

struct RefineBufs {
    Cand *cand;          // candidate (point, center) pairs
    uint32_t cap;        // capacity of cand
    uint32_t *cnt;       // [0] candidates, [1] listed (full) points, [2] overflow points, [3] pair points
    uint32_t *full;      // listed points
    float *thresh;       // per listed point: best + tau (packed-value space)
    uint32_t *over;      // points whose candidates did not fit (each listed once)
    uint32_t *oflag;     // per point: 1 while listed in `over` (reset by the full re-score)
    unsigned long long *key;  // per point: packed (exact cost bits << 32 | center), min-reduced
};

template <int CH>
__device__ __forceinline__ int swz(int row) { return row & ((CH < 16 ? CH : 16) - 1); }

// 16 bytes per lane from base + voff + soff into LDS at dst + 16 * lane, through a
// buffer resource over [base, base + nbytes) (loads past the end return 0).
// Word 3 0x00020000: raw 32-bit data format of the gfx9 buffer descriptor.
__device__ __forceinline__ void lds_dma16(const void *base, uint32_t nbytes, char *dst, uint32_t voff,
                                          uint32_t soff) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)nbytes, 0x00020000);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void *)dst, 16, voff, soff, 0, 0);
}

// Block-wide appends.  Every thread of a 256-thread block calls these together
// with its own count; each returns the thread's first slot.  One global atomic per
// list per block (a wave-level scan, then per-wave offsets through LDS), so the
// appends of the whole grid do not serialize on one counter.  `sm`: 8 u32 of LDS.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
}

// two lists at once: cnt2 = count A (low 16 bits) | count B << 16, block totals < 2^16;
// ctr_p counts the pairs of list A (two entries each)
__device__ __forceinline__ uint2 block_append2(uint32_t cnt2, uint32_t *ctr_a, uint32_t *ctr_b, uint32_t *ctr_p,
                                               uint32_t *sm) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t x = wave_incl_scan(cnt2);
    if (lane == 63) sm[wave] = x;
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int w = 0; w < 4; ++w) { const uint32_t t = sm[w]; sm[w] = run; run += t; }
        const uint32_t ta = run & 0xFFFFu, tb = run >> 16;
        sm[4] = ta ? atomicAdd(ctr_a, ta) : 0u;
        sm[5] = tb ? atomicAdd(ctr_b, tb) : 0u;
        if (ta) atomicAdd(ctr_p, ta / 2u);
    }
    __syncthreads();
    const uint32_t excl = sm[wave] + x - cnt2;
    const uint2 r = make_uint2(sm[4] + (excl & 0xFFFFu), sm[5] + (excl >> 16));
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint32_t block_append1(uint32_t cnt, uint32_t *ctr, uint32_t *sm) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t x = wave_incl_scan(cnt);
    if (lane == 63) sm[wave] = x;
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int w = 0; w < 4; ++w) { const uint32_t t = sm[w]; sm[w] = run; run += t; }
        sm[4] = run ? atomicAdd(ctr, run) : 0u;
    }
    __syncthreads();
    const uint32_t r = sm[4] + sm[wave] + x - cnt;
    __syncthreads();
    return r;
}

// a point whose candidates did not fit goes to the overflow list (once)
__device__ __forceinline__ void cand_overflow(const RefineBufs &R, uint32_t p) {
    if (atomicExch(&R.oflag[p], 1u) == 0u) R.over[atomicAdd(&R.cnt[2], 1u)] = p;
}

template <int KS, int NB, bool F16>
__global__ __launch_bounds__(256, (KS <= 16 ? 2 : 1)) void k_estep16(
    const uint4 *__restrict__ Xb, const uint4 *__restrict__ Cb, const float *__restrict__ cadd,
    const float *__restrict__ xnorm, const float *__restrict__ xdev, const float2 *__restrict__ cinfo,
    const EParams *__restrict__ prm, uint32_t n, uint32_t ntiles, uint32_t kreal, uint32_t *__restrict__ asn,
    RefineBufs R) {
    constexpr int CH = KS * 2;                  // 16-byte chunks per row
    constexpr int TILE = 64 * CH * 16;          // bytes per tile (64 rows)
    constexpr int INSTR = CH / 4;               // LDS loads per wave per tile
    // Two distinct LDS objects (not one indexed array): the tile loop alternates
    // between them statically, so the compiler can tell that the copy into one does
    // not alias the reads of the other and need not wait for it.
    __shared__ __attribute__((aligned(16))) char lds0[TILE];
    __shared__ __attribute__((aligned(16))) char lds1[TILE];
    __shared__ float ladd0[64], ladd1[64];
    __shared__ uint32_t sm_app[8];
    // the wave index is uniform: kept scalar, so the per-wave LDS destinations of the
    // tile copy are SGPRs rather than VGPRs read back with readfirstlane
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int h = lane >> 5, r32 = lane & 31;
    const uint32_t pbase = blockIdx.x * (4u * NB * 32u) + wave * (NB * 32u);

    typedef typename Op16<F16>::V V8;
    V8 b[NB][KS];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        uint32_t p = pbase + nb * 32 + r32;
        p = p < n ? p : n - 1;
        const uint4 *row = Xb + (size_t)p * CH + h * KS;
#pragma unroll
        for (int s = 0; s < KS; ++s) b[nb][s] = __builtin_bit_cast(V8, row[s]);
    }

    // Global->LDS copy of tile t into dst through a buffer resource over the center
    // array (loads past its end return 0).  Load j of this wave fills physical chunk
    // slots P = (4j + wave) * 64 + lane: row P / CH, swizzled chunk P % CH, whose
    // source chunk is (P % CH) ^ swz(row).  The swizzle repeats every PER loads, so
    // the lane's offset is voff[j % PER] and the rest (tile, j * 4 KiB) is scalar.
    constexpr int PER = CH == 64 ? 4 : (CH == 32 ? 2 : 1);
    uint32_t voff[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const uint32_t P = (uint32_t)((i * 4 + wave) * 64 + lane);
        const uint32_t r = P / CH, q = P % CH;
        voff[i] = (r * CH + (q ^ swz<CH>((int)r))) * 16u - (uint32_t)i * (4u * 64u * 16u);
    }
    auto issue_tile = [&](uint32_t t, char *dst) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < INSTR; ++j)
            lds_dma16(Cb, ntiles * (uint32_t)TILE, dst + (j * 4 + wave) * 64 * 16, voff[j % PER],
                      t * (uint32_t)TILE + (uint32_t)j * (4u * 64u * 16u));
    };

    // Running best / second / third per lane as packed values: positive floats whose
    // low 5 mantissa bits hold the position within a 32-center sub-tile, compared as
    // unsigned integers (bit patterns of positive floats order like their values).
    // btile / stile: sub-tile (32 centers) of the best and second.
    uint32_t best[NB], sec[NB], thr[NB];
    uint32_t btile[NB], stile[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        best[nb] = 0x7F7FFFFFu; sec[nb] = 0x7F7FFFFFu; thr[nb] = 0x7F7FFFFFu; btile[nb] = 0; stile[nb] = 0;
    }
    // Lane-constant part of the A-fragment address: the chunk of step s is
    // (h KS + s) ^ sw = ((h KS) ^ sw) ^ s (s < KS, KS a power of two), so the byte
    // address is abase ^ (16 s); row r32 + 32 has the same swizzle as r32.
    const uint32_t abase0 = (uint32_t)(r32 * (CH * 16) + ((h * KS) ^ swz<CH>(r32)) * 16);

    // Software pipeline over 32-center sub-tiles with two accumulator sets in fixed
    // roles: sub-tile 2t (rows 0-31 of tile t) goes to accA while the argmin epilogue
    // of sub-tile 2t - 1 (accB) is interleaved with its MFMAs; sub-tile 2t + 1 goes
    // to accB while accA's epilogue runs.  accB starts as +inf (an epilogue that
    // changes nothing).
    constexpr int E = 16 * NB;                  // epilogue elements per sub-tile
    floatx16 accA[NB], accB[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int e = 0; e < 16; ++e) accB[nb][e] = __uint_as_float(0x7F800000u);
    uint32_t b0[NB], s0[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) { b0[nb] = best[nb]; s0[nb] = sec[nb]; }

    auto epi_elem = [&](const floatx16 (&pv)[NB], int i) __attribute__((always_inline)) {
        const int nb = i / 16, e = i % 16;
        const uint32_t v = (__float_as_uint(pv[nb][e]) & 0xFFFFFFE0u) | (uint32_t)(8 * (e >> 2) + (e & 3));
        // third = med3(second, v, third); second = med3(best, v, second); best = min
        thr[nb] = max(min(sec[nb], v), min(max(sec[nb], v), thr[nb]));
        sec[nb] = max(min(best[nb], v), min(max(best[nb], v), sec[nb]));
        best[nb] = min(best[nb], v);
    };
    // sub-tile bookkeeping once all elements of sub-tile u went in
    auto epi_close = [&](uint32_t u) __attribute__((always_inline)) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            // the new best came from sub-tile u; the second is either the old best
            // (its sub-tile moves over) or from sub-tile u
            const bool nbest = best[nb] != b0[nb], nsec = sec[nb] != s0[nb];
            const uint32_t st_new = (nbest && sec[nb] == b0[nb]) ? btile[nb] : u;
            stile[nb] = (nbest || nsec) ? st_new : stile[nb];
            btile[nb] = nbest ? u : btile[nb];
            b0[nb] = best[nb];
            s0[nb] = sec[nb];
            // consume the results here: keeps the epilogue inside this sub-tile's
            // block, where it is interleaved with the MFMAs (instead of being sunk
            // past the end of the tile and run as one serial stretch)
            asm volatile("" ::"v"(thr[nb]), "v"(btile[nb]), "v"(stile[nb]), "v"(b0[nb]), "v"(s0[nb]));
        }
    };

    // MFMA chain of one sub-tile into acc, with the epilogue of pv interleaved.
    auto subtile = [&](const char *cur, const float *cadd_cur, uint32_t mb, floatx16 (&acc)[NB],
                       const floatx16 (&pv)[NB], uint32_t upv) __attribute__((always_inline)) {
        floatx16 init;   // C operand of the first MFMA of each chain (no copies)
#pragma unroll
        for (int e = 0; e < 16; ++e) init[e] = cadd_cur[mb * 32 + 8 * (e >> 2) + 4 * h + (e & 3)];
        // recomputed per sub-tile (the empty asm keeps the compiler from holding all
        // KS addresses in registers across the loop)
        uint32_t ab = abase0;
        asm volatile("" : "+v"(ab));
        const char *tb = cur + mb * (32u * CH * 16u);
        // Each A fragment is read FD steps ahead of its MFMAs, so its LDS latency is
        // covered by the MFMAs and epilogue work of the steps in between.  The
        // scheduling barrier after each step keeps that order: without it the
        // scheduler sinks every read down to its first use and the wave waits for
        // the LDS at every step.
        constexpr int FD = 2;
        V8 fr[FD + 1];
#pragma unroll
        for (int s = 0; s < FD && s < KS; ++s) fr[s] = *(const V8 *)(tb + (ab ^ (uint32_t)(s * 16)));
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            if (s + FD < KS) fr[(s + FD) % (FD + 1)] = *(const V8 *)(tb + (ab ^ (uint32_t)((s + FD) * 16)));
            const V8 a = fr[s % (FD + 1)];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                acc[nb] = mfma16<F16>(a, b[nb][s], s == 0 ? init : acc[nb]);
#pragma unroll
            for (int i = (s * E) / KS; i < ((s + 1) * E) / KS; ++i) epi_elem(pv, i);
            __builtin_amdgcn_sched_barrier(0);
        }
        epi_close(upv);
    };

    auto tile_step = [&](uint32_t t, const char *cur, const float *cadd_cur, char *nxt,
                         float *cadd_nxt) __attribute__((always_inline)) {
        float addn = 0.f;
        const bool more = t + 1 < ntiles;
        if (more) {
            issue_tile(t + 1, nxt);
            if (tid < 64) addn = cadd[(t + 1) * 64 + tid];
        }
        subtile(cur, cadd_cur, 0u, accA, accB, 2u * t - 1u);
        subtile(cur, cadd_cur, 1u, accB, accA, 2u * t);
        if (more && tid < 64) cadd_nxt[tid] = addn;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };

    issue_tile(0, lds0);
    if (tid < 64) ladd0[tid] = cadd[tid];
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (uint32_t t = 0; t < ntiles; t += 2) {
        tile_step(t, lds0, ladd0, lds1, ladd1);
        if (t + 1 >= ntiles) break;
        tile_step(t + 1, lds1, ladd1, lds0, ladd0);
    }
    // epilogue of the last sub-tile (accB)
#pragma unroll
    for (int i = 0; i < E; ++i) epi_elem(accB, i);
    epi_close(2u * ntiles - 1u);

    const EParams P = *prm;
    bool is_pair[NB], is_full[NB];
    uint32_t pc1[NB], pc2[NB];
    float pthr[NB];
    uint32_t cnt2 = 0;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const uint32_t bb = best[nb], ss = sec[nb];
        const uint32_t gb = btile[nb] * 32u + (bb & 31u) + 4u * h;
        const uint32_t gs = stile[nb] * 32u + (ss & 31u) + 4u * h;
        const float b = __uint_as_float(bb & 0xFFFFFFE0u);
        const float s = __uint_as_float(ss & 0xFFFFFFE0u);
        const float r = __uint_as_float(thr[nb] & 0xFFFFFFE0u);
        const float ob = __shfl_xor(b, 32), os = __shfl_xor(s, 32);
        const uint32_t ogb = __shfl_xor(gb, 32), ogs = __shfl_xor(gs, 32);
        const float orr = __shfl_xor(r, 32);
        const bool oth = (ob < b) || (ob == b && ogb < gb);
        const float v1 = oth ? ob : b;
        const uint32_t c1 = oth ? ogb : gb;
        const float sA = oth ? os : s, rA = oth ? orr : r, bB = oth ? b : ob, sB = oth ? s : os;
        const uint32_t gsA = oth ? ogs : gs, gB = oth ? gb : ogb;
        const bool secB = (bB < sA) || (bB == sA && gB < gsA);
        const float v2 = secB ? bB : sA;
        const uint32_t c2 = secB ? gB : gsA;
        const float v3 = secB ? fminf(sA, sB) : fminf(rA, bB);
        const uint32_t p = pbase + nb * 32 + r32;
        const bool mine = h == 0 && p < n;
        is_pair[nb] = false;
        is_full[nb] = false;
        pthr[nb] = FLT_MAX;
        pc1[nb] = c1;
        pc2[nb] = c2;
        if (mine && (P.c0nan || !(xnorm[p] <= FLT_MAX))) {
            // center 0 NaN: every row keeps center 0 (the scan's first cost is NaN); a row
            // with non-finite values: the exact scan over all centers (overflow list)
            asn[p] = 0u;
            R.oflag[p] = 0u;
            if (!P.c0nan) cand_overflow(R, p);
        } else if (mine) {
            const float xn = xnorm[p], dxn = xdev[p];
            // the values carry the scale P.s (a power of two; 1 for BF16)
            const float dbest = xn * xn + (v1 * P.inv_s - P.off);
            // Rounding of the 16-bit operands: with a = x - mu, b = c - mu and their
            // errors da, db, the approximate difference D~_j - D~_1 is off from the
            // exact one by at most
            //   A |b_j - b_1| + Bc,  A = 2 |da|,  Bc = 2 (|a| + |da|) (|db_1| + |db_j|)
            // (Cauchy-Schwarz on -2 a.b), plus the accumulation terms tau0 and taur D~_1.
            // Loose form: |b_j - b_1| <= |b_1| + cmax and |db_j| <= dbmax.
            const float2 ci = cinfo[c1 < kreal ? c1 : 0];
            const float A = 2.f * (1.f + 1e-3f) * dxn;
            const float Bc = 2.f * (1.f + 1e-3f) * (xn + dxn) * (ci.y + P.dbmax) + P.tau0 + P.taur * fmaxf(dbest, 0.f);
            const float tl = A * (ci.x + P.cmax) + Bc;
            // Tight form: |b_j - b_1| = |c_j - c_1| <= sqrt(D_1) + sqrt(D_j), each
            // exact distance at most its approximate value plus tl.  Every center
            // j ranked at or beyond a value D~ with D~_j - D~_1 - A (R_1 + R_j)
            // non-decreasing in D~_j (true when 4 tl >= A^2) is then decided by its
            // bound at D~.
            const float R1 = sqrtf(fmaxf(dbest, 0.f) + tl);
            const bool mono = 4.f * tl >= A * A;
            auto tau_at = [&](float v) __attribute__((always_inline)) {
                const float dj = xn * xn + (v * P.inv_s - P.off);
                const float t = A * (R1 + sqrtf(fmaxf(dj, 0.f) + tl)) + Bc;
                return (mono ? fminf(t, tl) : tl) * P.s;
            };
            const float tau2 = tau_at(v2), tau3 = tau_at(v3);
            // listed points: a center beyond v1 + tau_list cannot win (its bound at
            // D~_1 + tau_list, with tau_list <= tl)
            const float tlist = mono ? fminf(A * (R1 + sqrtf(fmaxf(dbest, 0.f) + 2.f * tl)) + Bc, tl) : tl;
            const bool ok1 = c1 < kreal;
            asn[p] = ok1 ? c1 : 0u;
            if (!(ok1 && v2 - v1 > tau2)) {
                R.key[p] = ~0ull;
                R.oflag[p] = 0u;   // set again below / by the collect pass if its candidates overflow
                is_pair[nb] = ok1 && c2 < kreal && v3 - v1 > tau3;
                is_full[nb] = !is_pair[nb];
                pthr[nb] = ok1 ? v1 + tlist * P.s : FLT_MAX;   // no finite best: every center is a candidate
            }
        }
        cnt2 += (is_pair[nb] ? 2u : 0u) + (is_full[nb] ? (1u << 16) : 0u);
    }
    // pairs -> two candidates each; the rest -> the list of the collect pass
    const uint2 base = block_append2(cnt2, &R.cnt[0], &R.cnt[1], &R.cnt[4], sm_app);
    uint32_t ia = base.x, ib = base.y;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const uint32_t p = pbase + nb * 32 + r32;
        if (is_pair[nb]) {
            if (ia + 2u <= R.cap) {
                R.cand[ia] = Cand{p, pc1[nb]};
                R.cand[ia + 1] = Cand{p, pc2[nb]};
            } else {
                cand_overflow(R, p);
            }
            ia += 2u;
        }
        if (is_full[nb]) {
            R.full[ib] = p;
            R.thresh[ib] = pthr[nb];
            ib += 1u;
        }
    }
}

// Collect pass over the listed points: every center whose approximate value is
// within the point's threshold becomes a candidate.  Blocks walk point blocks of
// 128 listed points (4 waves x 32, the MFMA B operand in registers) over the tiles
// of their share of the centers (blockIdx.y), streaming the tiles through two LDS
// buffers with the main pass's DMA; one block-wide append per tile.
template <int KS, bool F16>
__global__ __launch_bounds__(256) void k_collect16(const uint4 *__restrict__ Xb, const uint4 *__restrict__ Cb,
                                                      const float *__restrict__ cadd, uint32_t ntiles,
                                                      uint32_t kreal, RefineBufs R) {
    constexpr int CH = KS * 2;
    constexpr int TILE = 64 * CH * 16;
    constexpr int INSTR = CH / 4;
    __shared__ __attribute__((aligned(16))) char lds0[TILE];
    __shared__ __attribute__((aligned(16))) char lds1[TILE];
    __shared__ float ladd0[64], ladd1[64];
    __shared__ uint32_t sm_app[8];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int h = lane >> 5, r32 = lane & 31;
    const uint32_t count = R.cnt[1];
    const uint32_t t_begin = blockIdx.y * ntiles / gridDim.y, t_end = (blockIdx.y + 1) * ntiles / gridDim.y;
    if (t_begin >= t_end) return;   // uniform over the block

    constexpr int PER = CH == 64 ? 4 : (CH == 32 ? 2 : 1);
    uint32_t voff[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const uint32_t P = (uint32_t)((i * 4 + wave) * 64 + lane);
        const uint32_t r = P / CH, q = P % CH;
        voff[i] = (r * CH + (q ^ swz<CH>((int)r))) * 16u - (uint32_t)i * (4u * 64u * 16u);
    }
    auto issue_tile = [&](uint32_t t, char *dst) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < INSTR; ++j)
            lds_dma16(Cb, ntiles * (uint32_t)TILE, dst + (j * 4 + wave) * 64 * 16, voff[j % PER],
                      t * (uint32_t)TILE + (uint32_t)j * (4u * 64u * 16u));
    };
    const uint32_t abase0 = (uint32_t)(r32 * (CH * 16) + ((h * KS) ^ swz<CH>(r32)) * 16);

    typedef typename Op16<F16>::V V8;
    V8 b[KS];
    uint32_t p = 0;
    float thr = 0.f;
    bool valid = false;

    auto tile_step = [&](uint32_t t, const char *cur, const float *cadd_cur, char *nxt,
                         float *cadd_nxt) __attribute__((always_inline)) {
        float addn = 0.f;
        const bool more = t + 1 < t_end;
        if (more) {
            issue_tile(t + 1, nxt);
            if (tid < 64) addn = cadd[(t + 1) * 64 + tid];
        }
        uint32_t want = 0;   // bit mb * 16 + e
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            floatx16 init;
#pragma unroll
            for (int e = 0; e < 16; ++e) init[e] = cadd_cur[mb * 32 + 8 * (e >> 2) + 4 * h + (e & 3)];
            floatx16 acc;
            const char *tb = cur + mb * (32u * CH * 16u);
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const V8 a = *(const V8 *)(tb + (abase0 ^ (uint32_t)(s * 16)));
                acc = mfma16<F16>(a, b[s], s == 0 ? init : acc);
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const float v = __uint_as_float(__float_as_uint(acc[e]) & 0xFFFFFFE0u);
                const uint32_t c = t * 64u + mb * 32u + 8u * (e >> 2) + 4u * h + (e & 3);
                want |= (valid && v <= thr && c < kreal) ? (1u << (mb * 16 + e)) : 0u;
            }
        }
        uint32_t slot = block_append1((uint32_t)__builtin_popcount(want), &R.cnt[0], sm_app);
        bool over = false;
        while (want) {
            const int bit = __builtin_ctz(want);
            want &= want - 1u;
            const int mb = bit >> 4, e = bit & 15;
            const uint32_t c = t * 64u + mb * 32u + 8u * (e >> 2) + 4u * h + (e & 3);
            if (slot < R.cap) R.cand[slot] = Cand{p, c};
            else over = true;
            ++slot;
        }
        if (over) cand_overflow(R, p);
        if (more && tid < 64) cadd_nxt[tid] = addn;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };

    for (uint32_t pb = blockIdx.x; pb * 128u < count; pb += gridDim.x) {   // uniform
        const uint32_t q = pb * 128u + wave * 32u + r32;
        valid = q < count;
        const uint32_t qq = valid ? q : count - 1;
        p = R.full[qq];
        thr = R.thresh[qq];
        const uint4 *row = Xb + (size_t)p * CH + h * KS;
#pragma unroll
        for (int s = 0; s < KS; ++s) b[s] = __builtin_bit_cast(V8, row[s]);
        __syncthreads();
        issue_tile(t_begin, lds0);
        if (tid < 64) ladd0[tid] = cadd[t_begin * 64 + tid];
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (uint32_t t = t_begin; t < t_end; t += 2) {
            tile_step(t, lds0, ladd0, lds1, ladd1);
            if (t + 1 >= t_end) break;
            tile_step(t + 1, lds1, ladd1, lds0, ladd0);
        }
    }
}

// Exact cost of each candidate: s = fma(x_j - c_j, x_j - c_j, s) over j = 0..d-1 in
// FP32 (the oracle's sequential chain), min-reduced per point as (cost bits << 32 |
// center): lowest cost, then lowest index; the result does not depend on the order
// of the atomics.  The first 2 * cnt[4] candidates are the main pass's pairs (a
// point's two candidates side by side), the rest single candidates of the collect
// pass.
__device__ __forceinline__ float exact_sq(const float *__restrict__ x, const float *__restrict__ c, uint32_t d,
                                          int vec) {
    float s = 0.f;
    if (vec) {
        const float4 *x4 = (const float4 *)x, *c4 = (const float4 *)c;
        const uint32_t d4 = d >> 2;
#pragma unroll 8
        for (uint32_t j = 0; j < d4; ++j) {
            const float4 a = x4[j], b = c4[j];
            float t;
            t = a.x - b.x; s = fmaf(t, t, s);
            t = a.y - b.y; s = fmaf(t, t, s);
            t = a.z - b.z; s = fmaf(t, t, s);
            t = a.w - b.w; s = fmaf(t, t, s);
        }
    } else {
        for (uint32_t j = 0; j < d; ++j) { const float t = x[j] - c[j]; s = fmaf(t, t, s); }
    }
    return s;
}


// Pair points, blocks [0, gridDim.x / 2): one wave per 64 pair points (one per lane),
// grid-stride.  The rows and both centers of the wave's points go through LDS in
// slabs of 32 dims, loaded cooperatively (8 lanes per 128-byte row piece when vec,
// else one float per lane) while the previous slab is scored; each lane then runs
// its two FMA chains from LDS (the operations of exact_sq).  Single candidates, the
// other blocks: one thread each, grid-stride (a listed point's candidates share its
// row, which these threads find in cache).
constexpr int XS_SD = 32;   // dims per slab
__global__ __launch_bounds__(64) void k_exact_cands(const float *__restrict__ X, uint32_t d, int vec,
                                                    const float *__restrict__ C, RefineBufs R) {
    const uint32_t cnt = min(R.cnt[0], R.cap);
    const uint32_t np = min(2u * R.cnt[4], cnt) / 2u;   // pair points stored in full
    const uint32_t half = gridDim.x / 2;
    const int lane = threadIdx.x;
    if (blockIdx.x >= half) {
        const uint32_t nsingle = cnt - 2u * np;
        for (uint32_t i = (blockIdx.x - half) * 64u + lane; i < nsingle; i += half * 64u) {
            const Cand e = R.cand[2u * np + i];
            const float s = exact_sq(X + (uint64_t)e.p * d, C + (uint64_t)e.c * d, d, vec);
            atomicMin(&R.key[e.p], ((unsigned long long)__float_as_uint(s) << 32) | e.c);
        }
        return;
    }
    __shared__ float sx[64][XS_SD + 1], sa[64][XS_SD + 1], sb[64][XS_SD + 1];
    const uint32_t nslab = (d + XS_SD - 1) / XS_SD;
    for (uint32_t base = blockIdx.x * 64u; base < np; base += half * 64u) {   // uniform over the wave
        const uint32_t q = min(base + (uint32_t)lane, np - 1u);
        const Cand e0 = R.cand[2u * q], e1 = R.cand[2u * q + 1u];
        // staging: vec -> 8 instructions per operand, lane l takes row 8i + l/8, float4 l%8;
        // else 32 instructions, lane l takes row 2i + l/32, dim l%32
        float4 g[3][8];
        auto load_slab = [&](uint32_t j0) __attribute__((always_inline)) {
            if (vec) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int r = 8 * i + (lane >> 3);
                    const uint32_t j = j0 + 4u * (lane & 7);
                    const uint32_t pr = __shfl(e0.p, r), ca = __shfl(e0.c, r), cb = __shfl(e1.c, r);
                    const bool in = j < d;
                    g[0][i] = in ? *(const float4 *)(X + (uint64_t)pr * d + j) : make_float4(0.f, 0.f, 0.f, 0.f);
                    g[1][i] = in ? *(const float4 *)(C + (uint64_t)ca * d + j) : make_float4(0.f, 0.f, 0.f, 0.f);
                    g[2][i] = in ? *(const float4 *)(C + (uint64_t)cb * d + j) : make_float4(0.f, 0.f, 0.f, 0.f);
                }
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    float v[3][4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int r = 2 * (4 * i + u) + (lane >> 5);
                        const uint32_t j = j0 + (uint32_t)(lane & 31);
                        const uint32_t pr = __shfl(e0.p, r), ca = __shfl(e0.c, r), cb = __shfl(e1.c, r);
                        const bool in = j < d;
                        v[0][u] = in ? X[(uint64_t)pr * d + j] : 0.f;
                        v[1][u] = in ? C[(uint64_t)ca * d + j] : 0.f;
                        v[2][u] = in ? C[(uint64_t)cb * d + j] : 0.f;
                    }
#pragma unroll
                    for (int o = 0; o < 3; ++o) g[o][i] = make_float4(v[o][0], v[o][1], v[o][2], v[o][3]);
                }
            }
        };
        auto store_slab = [&]() __attribute__((always_inline)) {
            float(*dst[3])[XS_SD + 1] = {sx, sa, sb};
#pragma unroll
            for (int o = 0; o < 3; ++o) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if (vec) {
                        float *t = &dst[o][8 * i + (lane >> 3)][4 * (lane & 7)];
                        t[0] = g[o][i].x; t[1] = g[o][i].y; t[2] = g[o][i].z; t[3] = g[o][i].w;
                    } else {
                        const float vv[4] = {g[o][i].x, g[o][i].y, g[o][i].z, g[o][i].w};
#pragma unroll
                        for (int u = 0; u < 4; ++u) dst[o][2 * (4 * i + u) + (lane >> 5)][lane & 31] = vv[u];
                    }
                }
            }
        };
        float s0 = 0.f, s1 = 0.f;
        load_slab(0);
        for (uint32_t sl = 0; sl < nslab; ++sl) {
            __syncthreads();   // the previous slab's reads are done
            store_slab();
            __syncthreads();
            if (sl + 1 < nslab) load_slab((sl + 1) * XS_SD);   // in flight while this slab is scored
            const uint32_t jn = min<uint32_t>(XS_SD, d - sl * XS_SD);
            for (uint32_t j = 0; j < jn; ++j) {
                const float xv = sx[lane][j];
                float t = xv - sa[lane][j];
                s0 = fmaf(t, t, s0);
                t = xv - sb[lane][j];
                s1 = fmaf(t, t, s1);
            }
        }
        if (base + (uint32_t)lane < np) {
            // costs are >= 0 (or NaN), so their bit patterns order like their values
            const unsigned long long k0 = ((unsigned long long)__float_as_uint(s0) << 32) | e0.c;
            const unsigned long long k1 = ((unsigned long long)__float_as_uint(s1) << 32) | e1.c;
            atomicMin(&R.key[e0.p], k0 < k1 ? k0 : k1);
        }
    }
}

// Final assignment of the points the main pass left open, in one launch:
//  * every candidate's point that is not on the overflow list takes the center of its
//    min-reduced key (k_exact_cands has finished: stream order);
//  * the overflow points (candidates did not fit, or non-finite rows) get the exact
//    argmin over all centers: 256 points per block, rows staged through LDS in 32-dim
//    slabs, 16 centers per register tile.
// The two parts write disjoint points (oflag is 1 exactly for the overflow points and
// stays so until the next main pass resets it), so they need no ordering.
__global__ __launch_bounds__(256) void k_refine_final(const float *__restrict__ X, uint32_t d,
                                                      const float *__restrict__ C, uint32_t k, RefineBufs R,
                                                      uint32_t *__restrict__ asn) {
    {
        const uint32_t cnt = min(R.cnt[0], R.cap);
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
            const uint32_t p = R.cand[i].p;
            if (R.oflag[p] == 0u) asn[p] = (uint32_t)(R.key[p] & 0xFFFFFFFFull);
        }
    }
    const uint32_t *list = R.over;
    __shared__ float xs[256][33];
    __shared__ float cs[16][32];
    __shared__ uint32_t pts[256];
    const uint32_t cnt = R.cnt[2];
    const int tid = threadIdx.x;
    for (uint32_t base = blockIdx.x * 256u; base < cnt; base += gridDim.x * 256u) {
        __syncthreads();
        pts[tid] = base + tid < cnt ? list[base + tid] : 0xFFFFFFFFu;
        __syncthreads();
        float best = 0.f;
        uint32_t bi = 0xFFFFFFFFu;
        for (uint32_t c0 = 0; c0 < k; c0 += 16) {
            float acc[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) acc[c] = 0.f;
            for (uint32_t j0 = 0; j0 < d; j0 += 32) {
                __syncthreads();
                for (int q = tid; q < 256 * 32; q += 256) {
                    const int r = q >> 5, jj = q & 31;
                    const uint32_t pr = pts[r];
                    xs[r][jj] = (pr != 0xFFFFFFFFu && j0 + jj < d) ? X[(uint64_t)pr * d + j0 + jj] : 0.f;
                }
                for (int q = tid; q < 16 * 32; q += 256) {
                    const int c = q >> 5, jj = q & 31;
                    cs[c][jj] = (c0 + c < k && j0 + jj < d) ? C[(uint64_t)(c0 + c) * d + j0 + jj] : 0.f;
                }
                __syncthreads();
                const int jn = (int)min<uint32_t>(32, d - j0);
                for (int jj = 0; jj < jn; ++jj) {
                    const float xv = xs[tid][jj];
#pragma unroll
                    for (int c = 0; c < 16; ++c) { const float t = xv - cs[c][jj]; acc[c] = fmaf(t, t, acc[c]); }
                }
            }
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (c0 + c < k && (bi == 0xFFFFFFFFu || acc[c] < best)) { best = acc[c]; bi = c0 + c; }
        }
        if (pts[tid] != 0xFFFFFFFFu) asn[pts[tid]] = bi;
    }
}

// ---------------------------------------------------------------------------
// FP32 MFMA E-step ("plain FP32"): v_mfma_f32_32x32x2f32, 32 points per wave,
// tiles of 32 centers.  KS = dpad/2 steps; lane half h supplies dim h*dpad/2 + s.
// ---------------------------------------------------------------------------
template <int DP>
__global__ __launch_bounds__(256) void k_estep_f32(const float *__restrict__ X, uint32_t d,
                                                   const float *__restrict__ mu,
                                                   const float *__restrict__ Cc,   // kpad x DP, (c - mu) * -2
                                                   const float *__restrict__ cadd, // kpad
                                                   uint32_t n, uint32_t ntiles,
                                                   uint32_t *__restrict__ asn) {
    constexpr int KS = DP / 2;
    constexpr int ROWF = DP + 4;            // padded LDS row stride (floats)
    constexpr int CPT = (32 * DP / 4) / 256; // float4 chunks per thread per tile
    __shared__ __attribute__((aligned(16))) float lds[2][32 * ROWF];
    __shared__ float ladd[2][32];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, r32 = lane & 31;
    const uint32_t p = blockIdx.x * 128u + wave * 32u + r32;
    const uint32_t pc = p < n ? p : n - 1;

    float b[KS];
    const float *xr = X + (uint64_t)pc * d;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const uint32_t j = h * KS + s;
        b[s] = j < d ? xr[j] - mu[j] : 0.f;
    }

    // the next tile in registers (CPT float4 per thread), then into LDS; written out
    // without an array captured by a lambda, which would live in scratch memory
    float4 pf0, pf1, pf2, pf3, pf4, pf5, pf6, pf7;
    float addn = 0.f;
#define F32_LD(q, v)                                                                                       \
    if (q < CPT) {                                                                                         \
        const int g = tid + q * 256;                                                                       \
        v = ((const float4 *)Cc)[(size_t)(t1 * 32 + g / (DP / 4)) * (DP / 4) + (g % (DP / 4))];          \
    }
#define F32_ST(q, v)                                                                                       \
    if (q < CPT) {                                                                                         \
        const int g = tid + q * 256;                                                                       \
        *(float4 *)(&lds[buf1][(g / (DP / 4)) * ROWF + (g % (DP / 4)) * 4]) = v;                          \
    }
    auto load_tile = [&](uint32_t t1) __attribute__((always_inline)) {
        F32_LD(0, pf0) F32_LD(1, pf1) F32_LD(2, pf2) F32_LD(3, pf3)
        F32_LD(4, pf4) F32_LD(5, pf5) F32_LD(6, pf6) F32_LD(7, pf7)
        if (tid < 32) addn = cadd[t1 * 32 + tid];
    };
    auto store_tile = [&](int buf1) __attribute__((always_inline)) {
        F32_ST(0, pf0) F32_ST(1, pf1) F32_ST(2, pf2) F32_ST(3, pf3)
        F32_ST(4, pf4) F32_ST(5, pf5) F32_ST(6, pf6) F32_ST(7, pf7)
        if (tid < 32) ladd[buf1][tid] = addn;
    };
#undef F32_LD
#undef F32_ST

    float best = 0.f;
    uint32_t bi = 0xFFFFFFFFu;
    load_tile(0);
    store_tile(0);
    __syncthreads();
    for (uint32_t t = 0; t < ntiles; ++t) {
        const int buf = t & 1;
        if (t + 1 < ntiles) load_tile(t + 1);
        floatx16 acc;
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[e] = ladd[buf][8 * (e >> 2) + 4 * h + (e & 3)];
        const float *arow = &lds[buf][r32 * ROWF + h * KS];
#pragma unroll
        for (int s = 0; s < KS; s += 4) {
            float4 a = *(const float4 *)(arow + s);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b[s + 0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b[s + 1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b[s + 2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, b[s + 3], acc, 0, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const uint32_t c = t * 32u + 8u * (e >> 2) + 4u * h + (e & 3);
            if (bi == 0xFFFFFFFFu || acc[e] < best) { best = acc[e]; bi = c; }
        }
        if (t + 1 < ntiles) store_tile(buf ^ 1);
        __syncthreads();
    }
    const float ob = __shfl_xor(best, 32);
    const uint32_t oi = __shfl_xor(bi, 32);
    if (ob < best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    if (h == 0 && p < n) asn[p] = bi;
}

__global__ void k_prep_centers_f32(const float *__restrict__ C, uint64_t k, uint64_t kpad, uint32_t d,
                                   uint32_t dpad, const float *__restrict__ mu, float *__restrict__ Cc,
                                   float *__restrict__ cadd) {
    const uint64_t c = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= kpad) return;
    double s = 0.;
    for (uint32_t j = lane; j < dpad; j += 64) {
        float v = (c < k && j < d) ? C[c * d + j] - mu[j] : 0.f;
        Cc[c * dpad + j] = -2.f * v;
        s += (double)v * (double)v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) cadd[c] = c < k ? (float)s : FLT_MAX;
}

// ---------------------------------------------------------------------------
// EXACT form and per-point costs: sequential FMA chain per pair, FP32.
// 256 points per block; X staged through LDS in 32-dim slabs (coalesced reads).
// ---------------------------------------------------------------------------
constexpr int EX_TC = 16;  // centers per register tile

// keys == nullptr: the block's 256 rows against all k centers, written to asn / cost.
// Otherwise blockIdx.y takes the centers [y * per, (y + 1) * per) and min-reduces
// (cost bits << 32 | center) per row into keys (a NaN cost never wins, except at
// center 0, which the scan above keeps: then key 0 and the row's flag), so rows of a
// small n are spread over more blocks; k_exact_keys_out writes the result.
__global__ __launch_bounds__(256) void k_assign_exact(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                      const float *__restrict__ C, uint32_t k, uint32_t per,
                                                      uint32_t *__restrict__ asn, float *__restrict__ cost,
                                                      unsigned long long *__restrict__ keys,
                                                      uint8_t *__restrict__ nanflag) {
    __shared__ float xs[256][33];
    __shared__ float cs[EX_TC][32];
    const int tid = threadIdx.x;
    const uint64_t p0 = (uint64_t)blockIdx.x * 256;
    const uint64_t p = p0 + tid;
    const uint32_t cb = keys ? blockIdx.y * per : 0u;
    const uint32_t ce = keys ? min(k, cb + per) : k;
    float best = 0.f;
    uint32_t bi = 0xFFFFFFFFu;
    bool nan0 = false;
    for (uint32_t c0 = cb; c0 < ce; c0 += EX_TC) {
        float acc[EX_TC];
#pragma unroll
        for (int c = 0; c < EX_TC; ++c) acc[c] = 0.f;
        for (uint32_t j0 = 0; j0 < d; j0 += 32) {
            __syncthreads();
            // stage X[p0..p0+256)[j0..j0+32): 8192 floats, 32 per thread, rows contiguous
            for (int q = tid; q < 256 * 32; q += 256) {
                int r = q >> 5, jj = q & 31;
                uint64_t pr = p0 + r;
                xs[r][jj] = (pr < n && j0 + jj < d) ? X[pr * d + j0 + jj] : 0.f;
            }
            for (int q = tid; q < EX_TC * 32; q += 256) {
                int c = q >> 5, jj = q & 31;
                cs[c][jj] = (c0 + c < ce && j0 + jj < d) ? C[(uint64_t)(c0 + c) * d + j0 + jj] : 0.f;
            }
            __syncthreads();
            const int jn = (int)min<uint32_t>(32, d - j0);
            for (int jj = 0; jj < jn; ++jj) {
                const float xv = xs[tid][jj];
#pragma unroll
                for (int c = 0; c < EX_TC; ++c) { float t = xv - cs[c][jj]; acc[c] = fmaf(t, t, acc[c]); }
            }
        }
#pragma unroll
        for (int c = 0; c < EX_TC; ++c) {
            if (c0 + c >= ce) continue;
            if (keys) {
                if (c0 + c == 0 && acc[c] != acc[c]) nan0 = true;
                if (acc[c] == acc[c] && (bi == 0xFFFFFFFFu || acc[c] < best)) { best = acc[c]; bi = c0 + c; }
            } else if (bi == 0xFFFFFFFFu || acc[c] < best) {
                best = acc[c];
                bi = c0 + c;
            }
        }
    }
    if (p >= n) return;
    if (!keys) {
        asn[p] = bi;
        if (cost) cost[p] = best;
        return;
    }
    if (nan0) {
        nanflag[p] = 1;
        atomicMin(keys + p, 0ull);
    } else if (bi != 0xFFFFFFFFu) {
        // costs are >= 0, so their bit patterns order like their values
        atomicMin(keys + p, ((unsigned long long)__float_as_uint(best) << 32) | bi);
    }
}

__global__ void k_exact_keys_out(const unsigned long long *__restrict__ keys, const uint8_t *__restrict__ nanflag,
                                 uint64_t n, uint32_t *__restrict__ asn, float *__restrict__ cost) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long kk = keys[i];
        // every center NaN (and center 0 not): the scan keeps center 0 with its cost
        const bool none = kk == ~0ull;
        asn[i] = none ? 0u : (uint32_t)kk;
        if (cost) cost[i] = (nanflag[i] || none) ? NAN : __uint_as_float((uint32_t)(kk >> 32));
    }
}

__global__ void k_fill_keys(unsigned long long *__restrict__ keys, uint8_t *__restrict__ nanflag, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        keys[i] = ~0ull;
        nanflag[i] = 0;
    }
}

__global__ __launch_bounds__(256) void k_point_costs(const float *__restrict__ X, uint64_t n, uint32_t d,
                                                     const float *__restrict__ C,
                                                     const uint32_t *__restrict__ asn,
                                                     float *__restrict__ cost) {
    __shared__ float xs[256][33];
    const int tid = threadIdx.x;
    const uint64_t p0 = (uint64_t)blockIdx.x * 256;
    const uint64_t p = p0 + tid;
    const float *cr = C + (uint64_t)(p < n ? asn[p] : 0) * d;
    float s = 0.f;
    for (uint32_t j0 = 0; j0 < d; j0 += 32) {
        __syncthreads();
        for (int q = tid; q < 256 * 32; q += 256) {
            int r = q >> 5, jj = q & 31;
            uint64_t pr = p0 + r;
            xs[r][jj] = (pr < n && j0 + jj < d) ? X[pr * d + j0 + jj] : 0.f;
        }
        __syncthreads();
        const int jn = (int)min<uint32_t>(32, d - j0);
        for (int jj = 0; jj < jn; ++jj) { float t = xs[tid][jj] - cr[j0 + jj]; s = fmaf(t, t, s); }
    }
    if (p < n) cost[p] = s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int KS, int NB, bool F16>
static void launch16(Data &ds, Ctx &c, uint64_t kpad, uint64_t kreal, uint32_t *asn, const RefineBufs &R,
                        hipStream_t st) {
    const uint32_t n = (uint32_t)ds.n;
    const uint32_t per = 4u * NB * 32u;
    const dim3 grid((uint32_t)cdiv(n, per));
    const uint4 *X16 = F16 ? ds.Xh.as<const uint4>() : ds.Xb.as<const uint4>();
    const float *xdev = F16 ? ds.xdev_h.as<const float>() : ds.xdev.as<const float>();
    hipEvent_t ev;
    c.tic(0, ev);
    hipLaunchKernelGGL((k_estep16<KS, NB, F16>), grid, dim3(256), 0, st, X16, c.s[S_CB].as<const uint4>(),
                       c.s[S_CADD].as<const float>(), ds.xnorm.as<const float>(), xdev,
                       c.s[S_CINFO].as<const float2>(), c.s[S_PARAMS].as<const EParams>(), n, (uint32_t)(kpad / 64),
                       (uint32_t)kreal, asn, R);
    c.toc(0, ev);
    // collect pass over the listed points: the center tiles split over up to 4 slices
    // (grid y), so that a few thousand listed points still fill the GPU
    const uint32_t nt = (uint32_t)(kpad / 64), gy = std::min<uint32_t>(nt, 4u);
    hipLaunchKernelGGL((k_collect16<KS, F16>), dim3(2048 / gy, gy), dim3(256), 0, st, X16,
                       c.s[S_CB].as<const uint4>(), c.s[S_CADD].as<const float>(), (uint32_t)(kpad / 64),
                       (uint32_t)kreal, R);
}

template <int DP>
static void launch_f32(Data &ds, Ctx &c, uint64_t kpad, uint32_t *asn, hipStream_t st) {
    const uint32_t n = (uint32_t)ds.n;
    hipEvent_t ev;
    c.tic(0, ev);
    hipLaunchKernelGGL((k_estep_f32<DP>), dim3((uint32_t)cdiv(n, 128)), dim3(256), 0, st, ds.X,
                       (uint32_t)ds.d, ds.mu.as<const float>(), c.s[S_CB].as<const float>(),
                       c.s[S_CADD].as<const float>(), n, (uint32_t)(kpad / 32), asn);
    c.toc(0, ev);
}

void dense_sqrl2_assign_exact(Data &ds, const float *C, uint64_t k, uint32_t *asn, float *costs,
                              hipStream_t st) {
    hipEvent_t ev;
    ds.ctx->tic(0, ev);
    const uint32_t nbx = (uint32_t)cdiv(ds.n, 256);
    // fewer than 256 row blocks: the centers are split over up to 512 / nbx slices
    const uint64_t tiles = cdiv(k, EX_TC);
    const uint64_t S = nbx >= 256 ? 1 : std::min<uint64_t>(cdiv(512, nbx), tiles);
    if (S <= 1) {
        hipLaunchKernelGGL(k_assign_exact, dim3(nbx), dim3(256), 0, st, ds.X, ds.n, (uint32_t)ds.d, C, (uint32_t)k,
                           (uint32_t)k, asn, costs, (unsigned long long *)nullptr, (uint8_t *)nullptr);
    } else {
        const uint32_t per = (uint32_t)(cdiv(tiles, S) * EX_TC);
        const uint32_t ny = (uint32_t)cdiv(k, per);
        char *kb = (char *)ds.ctx->s[S_KEYS].reserve(ds.n * 9 + 8);
        auto *keys = (unsigned long long *)kb;
        auto *nanflag = (uint8_t *)(kb + ds.n * 8);
        const dim3 g1((uint32_t)std::min<uint64_t>(cdiv(ds.n, 256), 4096));
        hipLaunchKernelGGL(k_fill_keys, g1, dim3(256), 0, st, keys, nanflag, ds.n);
        hipLaunchKernelGGL(k_assign_exact, dim3(nbx, ny), dim3(256), 0, st, ds.X, ds.n, (uint32_t)ds.d, C, (uint32_t)k,
                           per, asn, costs, keys, nanflag);
        hipLaunchKernelGGL(k_exact_keys_out, g1, dim3(256), 0, st, (const unsigned long long *)keys,
                           (const uint8_t *)nanflag, ds.n, asn, costs);
    }
    ds.ctx->toc(0, ev);
}

void dense_sqrl2_costs(Data &ds, const float *C, const uint32_t *asn, float *costs, hipStream_t st) {
    hipLaunchKernelGGL(k_point_costs, dim3((uint32_t)cdiv(ds.n, 256)), dim3(256), 0, st, ds.X, ds.n,
                       (uint32_t)ds.d, C, asn, costs);
}

static void dense_sqrl2_assign_mfma(Data &ds, const float *C, uint64_t k, int precision, uint32_t *asn,
                                    hipStream_t st);

// Tuning of the 16-bit refined forms: they and the exact kernel give the same
// assignment bit for bit, so the faster one is used.  Clear cases are decided by the
// work n*k*d (measured: the MFMA form is 1.5x faster at 2^24 and 4x at 2^28.5); in
// between, both are timed once per device and (form, padded d, padded k, log2 n), min
// of two runs each, and the choice is kept for every context of the process; server
// workers never time inside a request and take the exact kernel below 2^30.
// mcg_tuning_enable(ctx, 0) or MCG_TUNE=0 always takes the MFMA form.
static std::mutex g_tune_mu;
static std::vector<std::pair<uint64_t, int>> g_tuned;

static int tuned_form(Data &ds, const float *C, uint64_t k, uint32_t *asn, hipStream_t st, int form16) {
    static const bool tune_on = [] { const char *e = getenv("MCG_TUNE"); return !(e && e[0] == '0'); }();
    if (!tune_on || !ds.ctx->tuning) return form16;
    const double work = (double)ds.n * (double)k * (double)ds.d;
    // one E-step on rows without their 16-bit copy: the copy and its preparation cost
    // more than the exact kernel saves below n*k*d = 2^30 (measured, dataset creation
    // to assignment)
    const bool have16 = form16 == MCG_PREC_FP16 ? ds.f16_ready : ds.bf16_ready;
    if (ds.oneshot && !have16) return work < 0x1p30 ? MCG_PREC_EXACT : form16;
    // a server's datasets are used for a few E-steps (a coreset's Lloyd iterations), so
    // the 16-bit copy and its preparation are paid per request: the exact kernel below
    // n*k*d = 2^30 (measured on the serving mix: the Lloyd stage 2.8 ms per request
    // with the exact kernel against 3.2 ms with the 16-bit form from 2^23)
    if (ds.ctx->serving) return work < 0x1p30 ? MCG_PREC_EXACT : form16;
    if (work >= 0x1p26) return form16;
    if (work <= 0x1p21) return MCG_PREC_EXACT;
    prepare_base(ds);
    uint64_t lg = 0;
    while ((1ull << lg) < ds.n) ++lg;
    const uint64_t key = ((uint64_t)(ds.ctx->device & 0xFF) << 56) | ((uint64_t)form16 << 52) |
                         ((uint64_t)ds.dpad << 40) | (cdiv(k, 64) << 8) | lg;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto &kv : g_tuned)
            if (kv.first == key) return kv.second;
    }
    if (stream_capturing(st)) return MCG_PREC_BF16;  // no timing inside a graph capture
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    auto time_form = [&](int form) {
        float best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            HIP_CHECK(hipEventRecord(e0, st));
            if (form == MCG_PREC_EXACT) dense_sqrl2_assign_exact(ds, C, k, asn, nullptr, st);
            else dense_sqrl2_assign_mfma(ds, C, k, form16, asn, st);
            HIP_CHECK(hipEventRecord(e1, st));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        return best;
    };
    const float t_mfma = time_form(form16), t_exact = time_form(MCG_PREC_EXACT);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int form = t_exact < t_mfma ? MCG_PREC_EXACT : form16;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.push_back({key, form});
    return form;
}

void dense_sqrl2_assign(Data &ds, const float *C, uint64_t k, int precision, uint32_t *asn,
                        hipStream_t st) {
    REQUIRE(ds.n < (1ull << 31), "dense rows must be < 2^31");
    if ((precision == MCG_PREC_BF16 || precision == MCG_PREC_FP16) && ds.d <= 512)
        precision = tuned_form(ds, C, k, asn, st, precision);
    if (precision == MCG_PREC_EXACT || ds.d > 512 || (precision == MCG_PREC_FP32 && ds.d > 256)) {
        dense_sqrl2_assign_exact(ds, C, k, asn, nullptr, st);
        return;
    }
    dense_sqrl2_assign_mfma(ds, C, k, precision, asn, st);
}

static void dense_sqrl2_assign_mfma(Data &ds, const float *C, uint64_t k, int precision, uint32_t *asn,
                                    hipStream_t st) {
    Ctx &c = *ds.ctx;
    const bool f16 = precision == MCG_PREC_FP16;
    prepare_base(ds);  // mu, dpad are shared by the MFMA forms
    if (precision != MCG_PREC_FP32) {
        if (f16) prepare_f16(ds);
        else prepare_bf16(ds);
    }
    const uint32_t d = (uint32_t)ds.d, dpad = ds.dpad;
    if (precision == MCG_PREC_FP32) {
        const uint64_t kpad = cdiv(k, 32) * 32;
        c.s[S_CB].reserve(kpad * dpad * sizeof(float));
        c.s[S_CADD].reserve(kpad * sizeof(float));
        hipLaunchKernelGGL(k_prep_centers_f32, dim3((uint32_t)cdiv(kpad, 4)), dim3(256), 0, st, C, k, kpad, d,
                           dpad, ds.mu.as<const float>(), c.s[S_CB].as<float>(), c.s[S_CADD].as<float>());
        switch (dpad) {
        case 64: launch_f32<64>(ds, c, kpad, asn, st); break;
        case 128: launch_f32<128>(ds, c, kpad, asn, st); break;
        default: launch_f32<256>(ds, c, kpad, asn, st); break;
        }
        return;
    }
    const uint64_t kpad = cdiv(k, 64) * 64;
    const uint64_t n = ds.n;
    // the E-step reads the BF16 centers through a buffer resource (32-bit byte range)
    REQUIRE(kpad * dpad * sizeof(uint16_t) < (1ull << 31), "BF16 form: k * d too large (use precision exact)");
    c.s[S_CB].reserve(kpad * dpad * sizeof(uint16_t));
    c.s[S_CADD].reserve(kpad * sizeof(float));
    c.s[S_CNORM].reserve(kpad * sizeof(double));
    c.s[S_PARAMS].reserve(sizeof(EParams));
    // refinement buffers: candidates (8 per point), listed points, thresholds, keys
    const uint32_t cap = (uint32_t)std::min<uint64_t>(8 * n + 1024, 0xFFFFFFF0ull);
    RefineBufs R;
    R.cand = (Cand *)c.s[S_PAIRS].reserve((uint64_t)cap * sizeof(Cand));
    R.cap = cap;
    R.cnt = (uint32_t *)c.s[S_FLAGCNT].reserve(8 * sizeof(uint32_t));
    R.full = (uint32_t *)c.s[S_FLAGS].reserve(n * sizeof(uint32_t));
    R.thresh = (float *)c.s[S_THRESH].reserve(n * sizeof(float));
    R.over = (uint32_t *)c.s[S_OVER].reserve(n * sizeof(uint32_t));
    R.key = (unsigned long long *)c.s[S_KEYS].reserve(n * sizeof(unsigned long long));
    {
        const size_t before = c.s[S_OFLAG].bytes;
        R.oflag = (uint32_t *)c.s[S_OFLAG].reserve(n * sizeof(uint32_t));
        if (c.s[S_OFLAG].bytes != before) fill_u32(R.oflag, 0u, c.s[S_OFLAG].bytes / 4, st);
    }
    c.s[S_CINFO].reserve(kpad * sizeof(float2));
    if (f16) {
        hipLaunchKernelGGL(k_center_norms, dim3((uint32_t)cdiv(kpad, 4)), dim3(256), 0, st, C, k, kpad, d,
                           ds.mu.as<const float>(), c.s[S_CNORM].as<double>());
        hipLaunchKernelGGL(k_prep_params, dim3(1), dim3(256), 0, st, c.s[S_CNORM].as<const double>(),
                           c.s[S_CINFO].as<const float2>(), k, kpad, ds.xstat.as<const float>(), d, 1,
                           c.s[S_CADD].as<float>(), c.s[S_PARAMS].as<EParams>(), R.cnt);
        hipLaunchKernelGGL(k_prep_centers_h, dim3((uint32_t)cdiv(kpad, 4)), dim3(256), 0, st, C, k, kpad, d, dpad,
                           ds.mu.as<const float>(), c.s[S_PARAMS].as<EParams>(), c.s[S_CB].as<uint16_t>(),
                           c.s[S_CINFO].as<float2>());
    } else {
        hipLaunchKernelGGL(k_prep_centers_bf16, dim3((uint32_t)cdiv(kpad, 4)), dim3(256), 0, st, C, k, kpad, d, dpad,
                           ds.mu.as<const float>(), c.s[S_CB].as<uint16_t>(), c.s[S_CNORM].as<double>(),
                           c.s[S_CINFO].as<float2>());
        hipLaunchKernelGGL(k_prep_params, dim3(1), dim3(256), 0, st, c.s[S_CNORM].as<const double>(),
                           c.s[S_CINFO].as<const float2>(), k, kpad, ds.xstat.as<const float>(), d, 0,
                           c.s[S_CADD].as<float>(), c.s[S_PARAMS].as<EParams>(), R.cnt);
    }
    if (f16) {
        switch (dpad) {
        case 64: launch16<4, 2, true>(ds, c, kpad, k, asn, R, st); break;
        case 128: launch16<8, 2, true>(ds, c, kpad, k, asn, R, st); break;
        case 256: launch16<16, 2, true>(ds, c, kpad, k, asn, R, st); break;
        default: launch16<32, 1, true>(ds, c, kpad, k, asn, R, st); break;
        }
    } else {
        switch (dpad) {
        case 64: launch16<4, 2, false>(ds, c, kpad, k, asn, R, st); break;
        case 128: launch16<8, 2, false>(ds, c, kpad, k, asn, R, st); break;
        case 256: launch16<16, 2, false>(ds, c, kpad, k, asn, R, st); break;
        default: launch16<32, 1, false>(ds, c, kpad, k, asn, R, st); break;
        }
    }
    {
        const int vec = (d % 4 == 0) && ((uintptr_t)ds.X % 16 == 0) && ((uintptr_t)C % 16 == 0);
        hipLaunchKernelGGL(k_exact_cands, dim3(4096), dim3(64), 0, st, ds.X, d, vec, C, R);
    }
    hipLaunchKernelGGL(k_refine_final, dim3(512), dim3(256), 0, st, ds.X, d, C, (uint32_t)k, R, asn);
    // R.cnt[4] (pair points) and R.cnt[1] (listed points) stay on the device for
    // mcg_refine_counts
    c.refine_cnt = R.cnt;
}

}  // namespace mcg
