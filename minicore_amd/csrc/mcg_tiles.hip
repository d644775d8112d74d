// Tiled sparse x dense products for CSR data: the cross term of the decomposed
// Bregman costs (E-step: for row i and center c, the sum over the row's stored entries
// of x * table(column, c)) and the soft M-step numerator (for column j and center c,
// the sum over the column's stored entries of x * R(row, c)).
//
// Both are "accumulator-major" sparse products: accumulator index a (a row for the
// E-step, a column for the M-step) takes sum_p x_p * op[idx_p][c] over its stored
// entries.  The entries are reorganised once per dataset into tiles of TG accumulator
// indices x TS operand rows ("stages").  A block (16 waves) owns one group of TG
// indices (TWR per wave, registers) and one slice of 64 * CPL centers (CPL per lane),
// and walks the stages in order: the operand's TS x (64 * CPL) slice, the stage's
// entries, their per-index end offsets and the waves' base offsets go global -> LDS by
// buffer DMA, double-buffered, so a stored entry costs one LDS read per lane of its
// operand row instead of a gather of the whole table row from L2 / MALL.  A wave loads
// its (at most 64) entries of the stage one per lane and broadcasts each with readlane;
// per index a loop runs over its entries.  Every sum runs over an index's entries in
// stage order, operand index ascending within a stage: the order of the CSR row
// (E-step) or of the column's rows (M-step), independent of the launch shape.
//
// Tiles are built when a product is first asked for.  A stage whose entries exceed
// what a wave or a block stages (64 / 1024) halves TS and the tiles are rebuilt; below
// TS = 8 the product stays with the row / column kernels.
//
// Reference arithmetic: msr_with_prior (include/minicore/dist/applicator.h:1632-1915)
// through the decompositions of mcg_generic.hip; the soft CSR M-step of
// set_centroids_full_mean (include/minicore/clustering/centroid.h:925-939).
#include "mcg_internal.h"
#include <algorithm>
#include <cfloat>
#include <cmath>

namespace mcg {

constexpr int TNW = 16;                        // waves per block (at most; the M-step's count)
constexpr uint32_t WCAP = 64;                  // entries per wave and stage (one per lane)
constexpr uint32_t ECAP = 1024;                // entries per block and stage
constexpr uint32_t STAGE_BYTES = 64 * 1024;    // one operand stage buffer
constexpr uint32_t TS_MAX = 128;               // operand rows per stage (counts are u8)
constexpr uint32_t TS_MIN = 8;
constexpr uint32_t CS_WORDS = 256;             // end-offset words per stage (TG <= 1024)

// lower_bound of v in the sorted a[0..m)
__device__ __forceinline__ uint64_t lbound(const uint32_t *__restrict__ a, uint64_t m, uint32_t v) {
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// entry counts of every (accumulator index, stage), clamped to 255: one wave per index,
// lanes over stages
__global__ __launch_bounds__(256) void k_tiles_count(const uint64_t *__restrict__ ptr, const uint32_t *__restrict__ idx,
                                                      uint64_t nacc, uint32_t nstages, uint32_t ts, uint32_t tg,
                                                      uint8_t *__restrict__ cnt) {
    const uint64_t a = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (a >= nacc) return;
    const uint64_t b = ptr[a], m = ptr[a + 1] - b;
    const uint64_t g = a / tg, al = a % tg;
    for (uint32_t s = lane; s < nstages; s += 64) {
        const uint64_t lo = lbound(idx + b, m, s * ts);
        const uint64_t hi = lbound(idx + b + lo, m - lo, (s + 1) * ts) + lo;
        cnt[((uint64_t)g * nstages + s) * tg + al] = (uint8_t)min<uint64_t>(hi - lo, 255);
    }
}

// per tile (one block of tg threads): the waves' base offsets (exclusive prefix of the
// entry counts of the 16 groups of twr indices), the tile's size and the largest group
__global__ __launch_bounds__(1024) void k_tiles_sizes(const uint8_t *__restrict__ cnt, uint32_t tg, uint32_t twr,
                                                       uint32_t nw, uint16_t *__restrict__ wb, uint32_t *__restrict__ size,
                                                       uint32_t *__restrict__ maxwave) {
    __shared__ uint32_t ws[TNW];
    const uint64_t t = blockIdx.x;
    const uint32_t grp = threadIdx.x / twr;
    uint32_t c = cnt[t * tg + threadIdx.x];
    for (uint32_t o = twr / 2; o > 0; o >>= 1) c += __shfl_xor(c, o);   // stays inside the group
    if (threadIdx.x % twr == 0) ws[grp] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0, mx = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            wb[t * nw + w] = (uint16_t)min<uint32_t>(run, 0xFFFFu);
            run += ws[w];
            mx = max(mx, ws[w]);
        }
        size[t] = run;
        atomicMax(maxwave, mx);
    }
}

// entries of each tile: indices in order, each index's entries of the stage in operand
// order; .x = byte offset of the operand row in the stage buffer (local row * rowbytes).
// The counts are then replaced by the end offsets inside the wave's group.
__global__ __launch_bounds__(1024) void k_tiles_fill(const uint64_t *__restrict__ ptr, const uint32_t *__restrict__ idx,
                                                      const float *__restrict__ val, uint64_t nacc, uint32_t nstages,
                                                      uint32_t ts, uint32_t tg, uint32_t twr, uint32_t nw, uint32_t rowbytes,
                                                      uint8_t *__restrict__ cnt, const uint16_t *__restrict__ wb,
                                                      const uint32_t *__restrict__ off, uint2 *__restrict__ ent) {
    const uint64_t t = blockIdx.x;
    const uint64_t g = t / nstages;
    const uint32_t s = (uint32_t)(t % nstages);
    const uint32_t l = threadIdx.x % twr, grp = threadIdx.x / twr;
    const uint64_t a = g * tg + threadIdx.x;
    const uint32_t c = cnt[t * tg + threadIdx.x];
    // inclusive prefix of the counts inside the group
    uint32_t inc = c;
    for (uint32_t o = 1; o < twr; o <<= 1) {
        const uint32_t v = __shfl_up(inc, o, (int)twr);
        if (l >= o) inc += v;
    }
    cnt[t * tg + threadIdx.x] = (uint8_t)min<uint32_t>(inc, 255);
    if (a >= nacc || c == 0) return;
    const uint64_t pos = (uint64_t)off[t] + wb[t * nw + grp] + (inc - c);
    const uint64_t b = ptr[a], m = ptr[a + 1] - b;
    const uint64_t lo = b + lbound(idx + b, m, s * ts);
    for (uint32_t q = 0; q < c; ++q)
        ent[pos + q] = make_uint2((idx[lo + q] - s * ts) * rowbytes, __float_as_uint(val[lo + q]));
}

__global__ void k_max_u32(const uint32_t *__restrict__ a, uint64_t n, uint32_t *__restrict__ mx) {
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, a[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(mx, m);
}

// Builds T for accumulator lists (ptr, idx, val) of nacc indices over nop operand rows,
// twr indices per wave and rowbytes bytes per staged operand row; TS starts at the
// largest value the stage buffer holds and halves while a stage overflows.
static void build_tiles(Tiles &T, const uint64_t *ptr, const uint32_t *idx, const float *val, uint64_t nacc,
                        uint64_t nop, uint64_t nnz, int cpl, uint32_t twr, uint32_t nw, uint32_t rowbytes,
                        hipStream_t st) {
    REQUIRE(nnz < (1ull << 32), "tiled product: more than 2^32 stored entries");
    T.ready = false;
    T.cpl = cpl;
    T.twr = twr;
    T.nw = nw;
    T.tg = twr * nw;
    T.nacc = nacc;
    T.nop = nop;
    T.ngroups = (uint32_t)cdiv(nacc, T.tg);
    DBuf size, tmp, mx;
    mx.reserve(8);
    for (uint32_t ts = std::min<uint32_t>(TS_MAX, STAGE_BYTES / rowbytes); ts >= TS_MIN; ts /= 2) {
        T.ts = ts;
        T.nstages = (uint32_t)cdiv(nop, ts);
        const uint64_t nt = (uint64_t)T.ngroups * T.nstages;
        T.cnt.reserve(nt * T.tg);
        fill_u32(T.cnt.as<uint32_t>(), 0u, nt * T.tg / 4, st);   // indices beyond nacc count 0
        T.wb.reserve(nt * nw * sizeof(uint16_t));
        T.off.reserve((nt + 1) * sizeof(uint32_t));
        size.reserve((nt + 1) * sizeof(uint32_t));
        fill_u32(mx.as<uint32_t>(), 0u, 2, st);
        hipLaunchKernelGGL(k_tiles_count, dim3((uint32_t)cdiv(nacc, 4)), dim3(256), 0, st, ptr, idx, nacc, T.nstages,
                           ts, T.tg, T.cnt.as<uint8_t>());
        fill_u32(size.as<uint32_t>() + nt, 0u, 1, st);
        hipLaunchKernelGGL(k_tiles_sizes, dim3((uint32_t)nt), dim3(T.tg), 0, st, T.cnt.as<const uint8_t>(), T.tg, twr,
                           nw, T.wb.as<uint16_t>(), size.as<uint32_t>(), mx.as<uint32_t>());
        scan_exclusive(size.as<const uint32_t>(), T.off.as<uint32_t>(), nt + 1, tmp, st);
        hipLaunchKernelGGL(k_max_u32, dim3((uint32_t)std::min<uint64_t>(cdiv(nt, 256), 1024)), dim3(256), 0, st,
                           size.as<const uint32_t>(), nt, mx.as<uint32_t>() + 1);
        uint32_t m2[2];
        copy_sync(m2, mx.p, 8, hipMemcpyDeviceToHost, st);
        T.maxwave = m2[0];
        T.maxseg = m2[1];
        if (T.maxwave > WCAP || T.maxseg > ECAP) continue;
        T.ent.reserve(std::max<uint64_t>(nnz, 1) * sizeof(uint2));
        hipLaunchKernelGGL(k_tiles_fill, dim3((uint32_t)nt), dim3(T.tg), 0, st, ptr, idx, val, nacc, T.nstages, ts, T.tg,
                           twr, nw, rowbytes, T.cnt.as<uint8_t>(), T.wb.as<const uint16_t>(), T.off.as<const uint32_t>(),
                           T.ent.as<uint2>());
        HIP_CHECK(hipStreamSynchronize(st));   // the temporaries are freed on return
        T.ready = true;
        return;
    }
    HIP_CHECK(hipStreamSynchronize(st));
    T.cnt.release();
    T.wb.release();
    T.off.release();
    T.ent.release();
}

// ---- staged stage data ---------------------------------------------------------------------
// Global -> LDS by buffer DMA (no registers), through buffer resources based at the
// stage's first operand row / first entry, so offsets stay 32-bit; reads past a
// resource give 0.
template <int SIZE>
__device__ __forceinline__ void dma(const void *base, uint32_t nbytes, void *dst, uint32_t voff) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)nbytes, 0x00020000);
    auto *l = (__attribute__((address_space(3))) void *)dst;
    if constexpr (SIZE == 16) __builtin_amdgcn_raw_ptr_buffer_load_lds(r, l, 16, voff, 0, 0, 0);
    else __builtin_amdgcn_raw_ptr_buffer_load_lds(r, l, 4, voff, 0, 0, 0);
}

// The block's staging of one stage: the operand rows [s*ts, s*ts + ts) x columns
// [c0, c0 + CW) into Tw (row-major, CW floats a row), the stage's entries into Es, the
// end offsets into Cs and the wave bases into Ws.  MODE 1 and 2: ld and c0 are multiples
// of 4, so 16-byte pieces (1 KB per wave instruction); MODE 2 also: the operand (and the
// last stage's rows past its end) spans under 4 GiB, so one buffer resource covers all of
// it; MODE 0: dwords (256 bytes per wave instruction).
template <int CPL, int MODE, int NW>
struct Stage {
    static constexpr bool ALIGN16 = MODE >= 1;
    static constexpr int CW = 64 * CPL;
    const float *op;
    uint64_t ld, nop;
    uint32_t c0, ts, tg;
    const uint2 *ent;
    const uint8_t *cnt;
    const uint16_t *wb;
    // MODE 2 stages through one buffer resource over the whole operand: the lane's part
    // of the offset is fixed and the stage's and instruction's parts are scalar
    // (soffset), so an instruction costs little more than its LDS destination.  Rows
    // past the operand's end read 0 either way.
    __amdgpu_buffer_rsrc_t wr;
    uint32_t wlane;
    __device__ __forceinline__ Stage(const float *op_, uint64_t ld_, uint64_t nop_, uint32_t c0_, uint32_t ts_,
                                     uint32_t tg_, const uint2 *ent_, const uint8_t *cnt_, const uint16_t *wb_,
                                     int lane)
        : op(op_), ld(ld_), nop(nop_), c0(c0_), ts(ts_), tg(tg_), ent(ent_), cnt(cnt_), wb(wb_) {
        wr = __builtin_amdgcn_make_buffer_rsrc((void *)op, (short)0, (int)(MODE == 2 ? nop * ld * 4 : 0), 0x00020000);
        constexpr int RPI = 4 / CPL, LPR = 64 / RPI;
        wlane = (uint32_t)(((lane / LPR) * ld + c0 + (lane % LPR) * 4) * 4);
    }
    __device__ __forceinline__ void issue(uint64_t t, uint32_t s, uint32_t seg, uint32_t sz, float *Tw, uint2 *Es,
                                          uint32_t *Cs, uint32_t *Ws, int wave, int lane) const {
        const uint64_t r0 = (uint64_t)s * ts;
        const uint32_t rows = (uint32_t)min<uint64_t>(ts, nop - r0);
        const float *base = op + r0 * ld;
        const uint32_t nb = (uint32_t)((uint64_t)rows * ld * 4);
        if constexpr (MODE == 2) {
            constexpr int RPI = 4 / CPL;
            const uint32_t ni = ts / RPI;
            const uint32_t step = (uint32_t)(NW * RPI * ld * 4);
            uint32_t so = (uint32_t)((r0 + (uint64_t)wave * RPI) * ld * 4);
            auto *l = (__attribute__((address_space(3))) char *)(Tw + wave * RPI * CW);
            constexpr uint32_t LSTEP = NW * RPI * CW * 4;
            if (ni == 4 * NW) {   // a full stage of 16 waves: four instructions a wave, unrolled
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(wr, (__attribute__((address_space(3))) void *)(l + j * LSTEP),
                                                             16, wlane, so + (uint32_t)j * step, 0, 0);
            } else {
                for (uint32_t q = wave; q < ni; q += NW) {
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(wr, (__attribute__((address_space(3))) void *)l, 16, wlane, so,
                                                             0, 0);
                    so += step;
                    l += LSTEP;
                }
            }
        } else if constexpr (MODE == 1) {
            constexpr int RPI = 4 / CPL, LPR = 64 / RPI;   // operand rows per instruction, lanes per row
            const uint32_t ni = ts / RPI;
            const int rl = lane / LPR, ch = lane % LPR;
            for (uint32_t q = wave; q < ni; q += NW)
                dma<16>(base, nb, Tw + q * RPI * CW, (uint32_t)(((uint64_t)(q * RPI + rl) * ld + c0 + ch * 4) * 4));
        } else {
            const uint32_t ni = ts * CPL;                  // one instruction: 64 floats of one row
            for (uint32_t q = wave; q < ni; q += NW) {
                const uint32_t r = q / CPL, part = q % CPL;
                dma<4>(base, nb, Tw + q * 64, (uint32_t)(((uint64_t)r * ld + c0 + part * 64 + lane) * 4));
            }
        }
        const uint32_t ninstr = (2 * sz + 63) / 64;
        for (uint32_t i = wave; i < ninstr; i += NW) dma<4>(ent + seg, sz * 8, (uint32_t *)Es + i * 64, (i * 64 + lane) * 4);
        if (wave == NW - 1)
            for (uint32_t i = 0; i < (tg / 4 + 63) / 64; ++i) dma<4>(cnt + t * tg, tg, Cs + i * 64, (i * 64 + lane) * 4);
        if (wave == NW - 2) dma<4>(wb + t * NW, NW * 2, Ws, lane * 4);
    }
};

// One wave's walk of one stage: accumulator a (0..TWR-1) takes the entries
// [end(a-1), end(a)) of the wave's list, in order; fn(a, x, operand row address).
template <int TWR, class Fn>
__device__ __forceinline__ void stage_walk(const float *Tw, const uint2 *Es, const uint32_t *Cs, const uint32_t *Ws,
                                           int wave, int lane, float pv, bool mkl, Fn &&fn) {
    const uint32_t wv = __builtin_amdgcn_readfirstlane(((const uint16_t *)Ws)[wave]);
    uint2 ev = Es[wv + lane];
    // MKL's entry term log1p(x / pv), once per entry (lanes past the list hold no entry)
    if (mkl) ev.y = __float_as_uint(log1pf(__uint_as_float(ev.y) / pv));
    uint32_t cw[TWR / 4];
#pragma unroll
    for (int i = 0; i < TWR / 4; ++i) cw[i] = __builtin_amdgcn_readfirstlane(Cs[wave * (TWR / 4) + i]);
    const char *T = (const char *)Tw;
    // the next entry's offset is read out while this entry's operand row loads (after
    // the last entry, lane e & 63 is read out and goes unused); each accumulator's range
    // comes straight from the end offsets, so nothing but the entry index is carried
    // from entry to entry (the walk is bound by its scalar instructions)
    uint32_t bo = __builtin_amdgcn_readlane(ev.x, 0);
    auto endof = [&](int a) -> uint32_t { return (cw[a >> 2] >> (8 * (a & 3))) & 0xFFu; };
#pragma unroll
    for (int a = 0; a < TWR; ++a) {
        uint32_t e = a ? endof(a - 1) : 0u;
        const uint32_t end = endof(a);
        while (e < end) {
            const float *tp = (const float *)(T + bo);
            const float x = __int_as_float(__builtin_amdgcn_readlane((int)ev.y, e));
            ++e;
            bo = __builtin_amdgcn_readlane(ev.x, e & 63);
            fn(a, x, tp);
        }
    }
}

// The stage loop shared by both products: buffers A / B are separate LDS objects and
// the loop is unrolled by two, so a DMA into one buffer never waits on reads of the other.
// One barrier per stage: the barrier after a stage's DMA has landed is also the point
// every wave passes only after walking the buffer the next DMA refills.
template <int CPL, int TWR, int MODE, int NW, class Walk>
__device__ __forceinline__ void stage_loop(const Stage<CPL, MODE, NW> &S, const uint32_t *__restrict__ off, uint64_t t0,
                                           uint32_t nstages, float *TwA, float *TwB, uint2 *EsA, uint2 *EsB,
                                           uint32_t *CsA, uint32_t *CsB, uint32_t *WsA, uint32_t *WsB, Walk &&walk) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t tend = t0 + nstages;   // off[tend]: the group's end
    // lanes 0 and 1 hold off[t], off[t + 1] of the stage issued next (vector loads, one
    // stage ahead)
    auto load_off = [&](uint32_t s) -> uint32_t { return off[min<uint64_t>(t0 + s + (lane & 1), tend)]; };
    {
        const uint32_t ov = load_off(0);
        __builtin_amdgcn_s_waitcnt(0);
        const uint32_t sg = __builtin_amdgcn_readlane(ov, 0), se = __builtin_amdgcn_readlane(ov, 1);
        S.issue(t0, 0, sg, se - sg, TwA, EsA, CsA, WsA, wave, lane);
    }
    uint32_t offv = load_off(1);
    for (uint32_t s = 0; s < nstages; s += 2) {
        __builtin_amdgcn_s_waitcnt(0);   // stage s has landed
        __syncthreads();
        if (s + 1 < nstages) {
            const uint32_t sg = __builtin_amdgcn_readlane(offv, 0), se = __builtin_amdgcn_readlane(offv, 1);
            S.issue(t0 + s + 1, s + 1, sg, se - sg, TwB, EsB, CsB, WsB, wave, lane);
            offv = load_off(s + 2);
        }
        walk(TwA, EsA, CsA, WsA);
        if (s + 1 >= nstages) break;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (s + 2 < nstages) {
            const uint32_t sg = __builtin_amdgcn_readlane(offv, 0), se = __builtin_amdgcn_readlane(offv, 1);
            S.issue(t0 + s + 2, s + 2, sg, se - sg, TwA, EsA, CsA, WsA, wave, lane);
            offv = load_off(s + 3);
        }
        walk(TwB, EsB, CsB, WsB);
    }
}

#define TILE_LDS                                                                        \
    __shared__ __attribute__((aligned(16))) float TwA[STAGE_BYTES / 4];                \
    __shared__ __attribute__((aligned(16))) float TwB[STAGE_BYTES / 4];                \
    __shared__ __attribute__((aligned(16))) uint2 EsA[ECAP + WCAP];                    \
    __shared__ __attribute__((aligned(16))) uint2 EsB[ECAP + WCAP];                    \
    __shared__ __attribute__((aligned(16))) uint32_t CsA[CS_WORDS];                    \
    __shared__ __attribute__((aligned(16))) uint32_t CsB[CS_WORDS];                    \
    __shared__ __attribute__((aligned(16))) uint32_t WsA[64];                          \
    __shared__ __attribute__((aligned(16))) uint32_t WsB[64];

// ---- E-step ------------------------------------------------------------------------------------
// Cross terms of TG = 16 * TWR rows x 64 * CPL centers per block (TWR = 64 / CPL rows a
// wave, CPL centers a lane), then the per-row / per-center terms of the decomposition
// (as k_row_costs) and the cost matrix and/or the row minimum (64-bit keys: cost bits
// above, center below; lowest cost, then lowest center).  SQRL2, MKL, REVERSE_MKL.
template <int MSR, int CPL, int MODE, int TWR, int NW>
__global__ __launch_bounds__(NW * 64) void k_tile_estep(const uint8_t *__restrict__ cnt, const uint16_t *__restrict__ wb,
                                                     const uint32_t *__restrict__ off, const uint2 *__restrict__ ent,
                                                     uint64_t n, uint32_t nstages, uint32_t ts, uint32_t d,
                                                     const float *__restrict__ CT, uint32_t kpad, uint32_t k,
                                                     uint32_t nslices, float pv, float logpv,
                                                     const float4 *__restrict__ rt, const double *__restrict__ cs,
                                                     float *__restrict__ out, unsigned long long *__restrict__ keys,
                                                     uint8_t *__restrict__ nanflag) {
    constexpr int TG = TWR * NW, CW = 64 * CPL;
    TILE_LDS
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t sl = blockIdx.x % nslices, g = blockIdx.x / nslices;
    const uint32_t c0 = sl * CW;
    const uint64_t row0 = (uint64_t)g * TG + (uint64_t)wave * TWR;
    float acc[TWR][CPL];
#pragma unroll
    for (int a = 0; a < TWR; ++a)
#pragma unroll
        for (int u = 0; u < CPL; ++u) acc[a][u] = 0.f;
    const Stage<CPL, MODE, NW> S(CT, kpad, d, c0, ts, (uint32_t)TG, ent, cnt, wb, lane);
    auto walk = [&](const float *Tw, const uint2 *Es, const uint32_t *Cs, const uint32_t *Ws) {
        stage_walk<TWR>(Tw + lane * CPL, Es, Cs, Ws, wave, lane, pv, MSR == MCG_MKL,
                        [&](int a, float x, const float *tp) {
            float t[CPL];
            if (CPL == 4) { const float4 v = *(const float4 *)tp; t[0] = v.x; t[1] = v.y; t[2] = v.z; t[3] = v.w; }
            else if (CPL == 2) { const float2 v = *(const float2 *)tp; t[0] = v.x; t[1] = v.y; }
            else t[0] = *tp;
            if (MSR == MCG_SQRL2) {
#pragma unroll
                for (int u = 0; u < CPL; ++u) acc[a][u] = fmaf(x, x - 2.f * t[u], acc[a][u]);
            } else {
#pragma unroll
                for (int u = 0; u < CPL; ++u) acc[a][u] = fmaf(x, t[u], acc[a][u]);
            }
        });
    };
    stage_loop<CPL, TWR>(S, off, (uint64_t)g * nstages, nstages, TwA, TwB, EsA, EsB, CsA, CsB, WsA, WsB, walk);
    // epilogue (the per-center terms are re-read per row: cached, and they keep no
    // registers across the row loop)
#pragma unroll
    for (int a = 0; a < TWR; ++a) {
        const uint64_t row = row0 + a;
        if (row >= n) break;
        float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
        if (MSR != MCG_SQRL2) r = rt[row];
        // a NaN cost never wins, except at center 0, which the reference's scan keeps
        // (then the row's flag marks the NaN minimum); the lane's centers ascend, so a
        // strict comparison keeps the lowest center among equal costs
        unsigned long long key = ~0ull;
#pragma unroll
        for (int u = 0; u < CPL; ++u) {
            const uint32_t c = c0 + lane * CPL + u;
            const uint32_t cc = c < kpad ? c : kpad - 1;
            // an opaque copy, so the FP64 widening stays with its row (hoisted, the 64
            // widened accumulators would not fit the registers)
            float av = acc[a][u];
            asm volatile("" : "+v"(av));
            float w;
            if (MSR == MCG_SQRL2) {
                w = (float)(cs[2 * cc] + (double)av);
                w = w <= 0.f ? 0.f : w;
            } else {
                if (MSR == MCG_MKL) w = (float)(cs[2 * cc] - cs[2 * cc + 1] * (double)(logpv - r.y) - (double)av);
                else w = (float)((double)r.z - (double)r.w * cs[2 * cc] - (double)av * (double)(1.f / r.x));
                w = w <= 0.f ? 0.f : w;   // also -0 -> +0, so the key bits order like the values
                if (w == INFINITY) w = FLT_MAX;
            }
            if (c < k) {
                if (out) out[row * k + c] = w;
                unsigned long long ku;
                if (w != w) {
                    ku = c == 0 ? 0ull : ~0ull;
                    if (keys && c == 0) nanflag[row] = 1;
                } else {
                    ku = ((unsigned long long)__float_as_uint(w) << 32) | c;
                }
                key = ku < key ? ku : key;
            }
        }
        if (keys) {
            for (int o = 32; o > 0; o >>= 1) {
                const unsigned long long ok2 = __shfl_xor(key, o);
                key = ok2 < key ? ok2 : key;
            }
            if (lane == 0 && key != ~0ull) atomicMin(keys + row, key);
        }
    }
}

__global__ void k_keys_out(const unsigned long long *__restrict__ keys, const uint8_t *__restrict__ nanflag,
                           uint64_t n, uint32_t *__restrict__ asn, float *__restrict__ best) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long kk = keys[i];
        if (asn) asn[i] = (uint32_t)kk;
        if (best) best[i] = nanflag[i] ? NAN : __uint_as_float((uint32_t)(kk >> 32));
    }
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, unsigned long long v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static bool tiles_enabled() {
    static const bool on = [] { const char *e = getenv("MCG_TILES"); return !(e && e[0] == '0'); }();
    return on;
}

// Shape of the E-step tiles (the row kernel serves k <= 64): centers per lane, rows per
// wave (64 accumulators), waves per block.  (MCG_TILES_ECPL = 2 or 4 overrides the
// centers per lane.)  Twelve waves of 32 rows at four centers a lane (a third less table
// staging, twice the entries a wave and stage) measured the same at 10^6 x 32768 rows,
// k = 256: the walk is bound by its per-entry instruction stream, not by the staging or
// the barrier.
struct EShape { int cpl; uint32_t twr, nw; };
static EShape estep_shape(uint64_t k) {
    static const int fc = [] { const char *e = getenv("MCG_TILES_ECPL"); return e ? atoi(e) : 0; }();
    const int cpl = (fc == 2 || fc == 4) ? fc : (k > 128 ? 4 : 2);
    return {cpl, (uint32_t)(64 / cpl), (uint32_t)TNW};
}

// The tiles pay for their build (a few passes over the entries) and their staging
// when the product is large; JSD's entry terms are bound by their logarithms (two
// per entry and center either way), so it stays with the row kernel.
// (with the context's tuning off they are used whenever they apply)
static bool tiles_worth(const Data &ds, uint64_t k) {
    return !ds.ctx->tuning || (double)ds.nnz * (double)k >= 4e9;
}

bool tiled_costs_ok(const Data &ds, int msr, uint64_t k) {
    return tiles_enabled() && ds.csr && ds.nnz < (1ull << 32) && ds.n < (1ull << 32) && k > 64 &&
           tiles_worth(ds, k) && (msr == MCG_SQRL2 || msr == MCG_MKL || msr == MCG_REVERSE_MKL);
}

// the row tiles for k centers (built on first use); false when no stage size fits
bool tiled_costs_ready(Data &ds, uint64_t k, hipStream_t st) {
    const EShape sh = estep_shape(k);
    Tiles &T = ds.rtiles;
    if (T.cpl != sh.cpl || T.twr != sh.twr || T.nw != sh.nw)
        build_tiles(T, ds.ptr.as<const uint64_t>(), ds.idx.as<const uint32_t>(), ds.val.as<const float>(), ds.n, ds.d,
                    ds.nnz, sh.cpl, sh.twr, sh.nw, 64 * sh.cpl * 4, st);
    return T.ready;
}

void tiled_costs(Data &ds, int msr, const float *CT, uint32_t kpad, uint64_t k, float pv, float logpv,
                 const float4 *rt, const double *cs, float *out_nk, uint32_t *asn, float *best, hipStream_t st) {
    Ctx &c = *ds.ctx;
    const Tiles &T = ds.rtiles;
    const EShape sh = estep_shape(k);
    REQUIRE(T.ready && T.cpl == sh.cpl && T.twr == sh.twr && T.nw == sh.nw, "internal: row tiles not ready");
    const uint32_t nslices = (uint32_t)cdiv(k, 64 * T.cpl);
    unsigned long long *keys = nullptr;
    uint8_t *nanflag = nullptr;
    if (asn || best) {
        char *kb = (char *)c.s[S_KEYS].reserve(ds.n * 9 + 8);
        keys = (unsigned long long *)kb;
        nanflag = (uint8_t *)(kb + ds.n * 8);
        hipLaunchKernelGGL(k_fill_u64, dim3((uint32_t)std::min<uint64_t>(cdiv(ds.n, 256), 4096)), dim3(256), 0, st, keys,
                           ~0ull, ds.n);
        fill_u32((uint32_t *)nanflag, 0u, cdiv(ds.n, 4), st);
    }
    const dim3 grid(T.ngroups * nslices), blk(T.nw * 64);
    const auto *cn = T.cnt.as<const uint8_t>();
    const auto *w = T.wb.as<const uint16_t>();
    const auto *o = T.off.as<const uint32_t>();
    const auto *e = T.ent.as<const uint2>();
    const uint32_t d = (uint32_t)ds.d;
    hipEvent_t ev;
    c.tic(0, ev);
    // one buffer resource over the whole table when it (and the last stage's rows past
    // its end) spans under 4 GiB
    const bool whole = ((uint64_t)d + T.ts) * kpad * 4 < (1ull << 32);
#define TILE_E2(M, P, W, R, N)                                                                                    \
    hipLaunchKernelGGL((k_tile_estep<M, P, W, R, N>), grid, blk, 0, st, cn, w, o, e, ds.n, T.nstages, T.ts, d, CT, kpad,   \
                       (uint32_t)k, nslices, pv, logpv, rt, cs, out_nk, keys, nanflag)
#define TILE_E1(M, P, R, N) do { if (whole) TILE_E2(M, P, 2, R, N); else TILE_E2(M, P, 1, R, N); } while (0)
#define TILE_E4(M) TILE_E1(M, 4, 16, TNW)
    if (T.cpl == 4) {
        switch (msr) {
        case MCG_SQRL2: TILE_E4(MCG_SQRL2); break;
        case MCG_MKL: TILE_E4(MCG_MKL); break;
        default: TILE_E4(MCG_REVERSE_MKL); break;
        }
    } else {
        switch (msr) {
        case MCG_SQRL2: TILE_E1(MCG_SQRL2, 2, 32, TNW); break;
        case MCG_MKL: TILE_E1(MCG_MKL, 2, 32, TNW); break;
        default: TILE_E1(MCG_REVERSE_MKL, 2, 32, TNW); break;
        }
    }
#undef TILE_E4
#undef TILE_E1
#undef TILE_E2
    c.toc(0, ev);
    if (keys)
        hipLaunchKernelGGL(k_keys_out, dim3((uint32_t)std::min<uint64_t>(cdiv(ds.n, 256), 4096)), dim3(256), 0, st,
                           (const unsigned long long *)keys, (const uint8_t *)nanflag, ds.n, asn, best);
}

// ---- soft M-step numerator -----------------------------------------------------------------
// num[j][c] = sum over column j's entries (rows ascending) of R[i][c] * x (FP64): TG =
// 16 * TWR columns (TWR = 32 / CPL a wave) x 64 * CPL centers per block.
template <int CPL, int MODE>
__global__ __launch_bounds__(1024) void k_tile_mstep(const uint8_t *__restrict__ cnt, const uint16_t *__restrict__ wb,
                                                     const uint32_t *__restrict__ off, const uint2 *__restrict__ ent,
                                                     uint32_t d, uint32_t nstages, uint32_t ts, uint64_t n,
                                                     const float *__restrict__ R, uint32_t k, uint32_t nslices,
                                                     uint32_t kpad, double *__restrict__ num) {
    constexpr int TWR = 32 / CPL, TG = TWR * TNW, CW = 64 * CPL;
    constexpr int NW = TNW;
    TILE_LDS
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t sl = blockIdx.x % nslices, g = blockIdx.x / nslices;
    const uint32_t c0 = sl * CW;
    double acc[TWR][CPL];
#pragma unroll
    for (int a = 0; a < TWR; ++a)
#pragma unroll
        for (int u = 0; u < CPL; ++u) acc[a][u] = 0.;
    const Stage<CPL, MODE, NW> S(R, k, n, c0, ts, (uint32_t)TG, ent, cnt, wb, lane);
    auto walk = [&](const float *Tw, const uint2 *Es, const uint32_t *Cs, const uint32_t *Ws) {
        stage_walk<TWR>(Tw + lane * CPL, Es, Cs, Ws, wave, lane, 0.f, false, [&](int a, float x, const float *tp) {
            const double xd = (double)x;
            if (CPL == 4) {
                const float4 v = *(const float4 *)tp;
                acc[a][0] = fma((double)v.x, xd, acc[a][0]);
                acc[a][1] = fma((double)v.y, xd, acc[a][1]);
                acc[a][2] = fma((double)v.z, xd, acc[a][2]);
                acc[a][3] = fma((double)v.w, xd, acc[a][3]);
            } else if (CPL == 2) {
                const float2 v = *(const float2 *)tp;
                acc[a][0] = fma((double)v.x, xd, acc[a][0]);
                acc[a][1] = fma((double)v.y, xd, acc[a][1]);
            } else {
                acc[a][0] = fma((double)*tp, xd, acc[a][0]);
            }
        });
    };
    stage_loop<CPL, TWR>(S, off, (uint64_t)g * nstages, nstages, TwA, TwB, EsA, EsB, CsA, CsB, WsA, WsB, walk);
    const uint64_t j0 = (uint64_t)g * TG + (uint64_t)wave * TWR;
#pragma unroll
    for (int a = 0; a < TWR; ++a) {
        const uint64_t j = j0 + a;
        if (j >= d) break;
#pragma unroll
        for (int u = 0; u < CPL; ++u) num[j * kpad + c0 + lane * CPL + u] = acc[a][u];
    }
}

// C[c][j] = num[j][c] / den[c]
__global__ void k_tile_mstep_final(const double *__restrict__ num, uint32_t kpad, uint32_t k, uint32_t d,
                                   const double *__restrict__ den, float *__restrict__ C) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)k * d) return;
    const uint32_t c = (uint32_t)(e / d), j = (uint32_t)(e % d);
    C[e] = (float)(num[(uint64_t)j * kpad + c] / den[c]);
}

// centers per lane of the M-step tiles (MCG_TILES_MCPL = 1, 2 or 4 overrides)
static int mstep_cpl(uint64_t k) {
    static const int forced = [] { const char *e = getenv("MCG_TILES_MCPL"); return e ? atoi(e) : 0; }();
    if (forced == 1 || forced == 2 || forced == 4) return forced;
    return k > 128 ? 2 : 1;
}

bool tiled_mstep_ok(const Data &ds, uint64_t k) {
    return tiles_enabled() && ds.csr && ds.nnz < (1ull << 32) && ds.n < (1ull << 32) && k > 64 && tiles_worth(ds, k);
}

// the column tiles (built on first use); false when no stage size fits
bool tiled_mstep_ready(Data &ds, uint64_t k, const uint64_t *cptr, const uint32_t *crow, const float *cval,
                       hipStream_t st) {
    const int cpl = mstep_cpl(k);
    Tiles &T = ds.ctiles;
    if (T.cpl != cpl) build_tiles(T, cptr, crow, cval, ds.d, ds.n, ds.nnz, cpl, 32 / cpl, TNW, 64 * cpl * 4, st);
    return T.ready;
}

void tiled_soft_numer(Data &ds, const float *R, uint64_t k, const double *den, float *Cnew, hipStream_t st) {
    Ctx &c = *ds.ctx;
    const Tiles &T = ds.ctiles;
    REQUIRE(T.ready && T.cpl == mstep_cpl(k), "internal: column tiles not ready");
    const uint32_t cw = 64 * T.cpl;
    const uint32_t nslices = (uint32_t)cdiv(k, cw), kpad = nslices * cw;
    double *num = (double *)c.s[S_PART].reserve((uint64_t)ds.d * kpad * sizeof(double));
    const dim3 grid(T.ngroups * nslices), blk(TNW * 64);
    const auto *cn = T.cnt.as<const uint8_t>();
    const auto *w = T.wb.as<const uint16_t>();
    const auto *o = T.off.as<const uint32_t>();
    const auto *e = T.ent.as<const uint2>();
    const uint32_t d = (uint32_t)ds.d;
    const bool a16 = k % 4 == 0;
    hipEvent_t ev;
    c.tic(1, ev);
    const int mode = !a16 ? 0 : ((ds.n + T.ts) * k * 4 < (1ull << 32) ? 2 : 1);
#define TILE_M(P, A)                                                                                             \
    hipLaunchKernelGGL((k_tile_mstep<P, A>), grid, blk, 0, st, cn, w, o, e, d, T.nstages, T.ts, ds.n, R, (uint32_t)k, \
                       nslices, kpad, num)
    switch (T.cpl) {
    case 4: if (mode == 2) TILE_M(4, 2); else if (mode == 1) TILE_M(4, 1); else TILE_M(4, 0); break;
    case 2: if (mode == 2) TILE_M(2, 2); else if (mode == 1) TILE_M(2, 1); else TILE_M(2, 0); break;
    default: if (mode == 2) TILE_M(1, 2); else if (mode == 1) TILE_M(1, 1); else TILE_M(1, 0); break;
    }
#undef TILE_M
    c.toc(1, ev);
    const uint64_t kd = k * ds.d;
    hipLaunchKernelGGL(k_tile_mstep_final, dim3((uint32_t)cdiv(kd, 256)), dim3(256), 0, st, (const double *)num, kpad,
                       (uint32_t)k, (uint32_t)ds.d, den, Cnew);
}

}  // namespace mcg
