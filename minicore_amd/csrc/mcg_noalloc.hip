// Allocation-free entry point: coreset sensitivities on a fixed device scratch.
//
// For callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  The kernels are those
// of mcg_coreset_probs; their scratch (per-cluster fixed-point sums and the
// unnormalised sensitivities) is the file-scope array `scratch` below instead of
// context buffers.  Each GPU holds its own copy of the array, at its own address,
// from the first use of this library on that GPU until the process ends; host
// code reaches it only through the address that hipGetSymbolAddress returns for
// the current device.
#include "mcg_internal.h"
#include <algorithm>
#include <mutex>

namespace mcg {

static constexpr uint64_t NOALLOC_MAX_N = 1ull << 22;        // rows
static constexpr uint64_t NOALLOC_MAX_CENTERS = 1ull << 16;  // clusters
static constexpr uint64_t NOALLOC_BYTES = (48 + 3 * NOALLOC_MAX_CENTERS * 8 + 255) / 256 * 256 + NOALLOC_MAX_N * 8;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static constexpr int MAX_DEVICES = 64;
static std::mutex g_mu;
static void *g_ptr[MAX_DEVICES];   // device address of `scratch` per device (null: not set up)

// check that every assignment is below ncenters; flag[0] becomes nonzero otherwise
__global__ void k_asn_in_range(const uint32_t *__restrict__ asn, uint64_t n, uint64_t ncenters,
                               uint32_t *__restrict__ flag) {
    uint32_t bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        bad |= asn[i] >= ncenters;
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

void noalloc_setup(int device) {
    REQUIRE(device >= 0 && device < MAX_DEVICES, "device id out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return;
    void *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    REQUIRE(p != nullptr, "no device address for the fixed scratch");
    g_ptr[device] = p;
}

static void *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

}  // namespace mcg

using namespace mcg;

extern "C" int mcg_coreset_probs_noalloc(mcg_ctx *ctx, int sens, uint64_t n, uint64_t ncenters,
                                         const double *costs, const uint32_t *asn, const double *weights,
                                         float *probs) {
    return guard([&] {
        REQUIRE(ctx && costs && asn && probs, "null argument");
        REQUIRE(n > 0 && ncenters > 0, "empty input");
        REQUIRE(n <= NOALLOC_MAX_N, "mcg_coreset_probs_noalloc: n above the fixed scratch's limit of 4194304 rows");
        REQUIRE(ncenters <= NOALLOC_MAX_CENTERS,
                "mcg_coreset_probs_noalloc: ncenters above the fixed scratch's limit of 65536");
        REQUIRE(coreset_scratch_bytes(n, ncenters) <= NOALLOC_BYTES, "fixed scratch too small");
        REQUIRE(is_device_ptr(costs) && is_device_ptr(asn) && is_device_ptr(probs) &&
                    (!weights || is_device_ptr(weights)),
                "mcg_coreset_probs_noalloc needs device buffers");
        Ctx &c = ctx->c;
        HIP_CHECK(hipSetDevice(c.device));
        char *base = (char *)scratch_ptr(c.device);
        REQUIRE(base != nullptr, "fixed scratch not set up on this device");
        // the range check uses the scratch's first word; the sums overwrite it after
        uint32_t *flag = (uint32_t *)base;
        fill_u32(flag, 0u, 1, c.stream);
        const uint32_t g = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 2048);
        hipLaunchKernelGGL(k_asn_in_range, dim3(g), dim3(256), 0, c.stream, asn, n, ncenters, flag);
        uint32_t bad = 0;
        copy_sync(&bad, flag, sizeof(bad), hipMemcpyDeviceToHost, c.stream);
        REQUIRE(!bad, "assignment out of range");
        coreset_probs_core(sens, n, ncenters, costs, asn, weights, probs, base, c.stream);
        HIP_CHECK(hipStreamSynchronize(c.stream));
        c.collect();
    });
}
