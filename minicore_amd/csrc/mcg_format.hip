// On-disk formats around the clustering path: the CoresetSampler and IndexCoreset
// streams of include/minicore/coreset/coreset.h.  Host code only (zlib); the
// samplers' probabilities come from mcg_coreset_probs and go back into
// mcg_coreset_sample unchanged.
#include "mcg_internal.h"
#include <zlib.h>
#include <algorithm>
#include <cstring>
#include <memory>

namespace mcg {
namespace {

struct GzFile {
    gzFile f = nullptr;
    GzFile(const char *path, const char *mode) {
        REQUIRE(path && path[0], "null or empty path");
        f = gzopen(path, mode);
        if (!f) MCG_THROW(MCG_EIO, std::string("Failed to open file ") + path);
    }
    ~GzFile() { if (f) gzclose(f); }
    void put(const void *p, uint64_t nb) {
        // gzwrite takes an unsigned count; write in pieces below 1 GiB
        const char *c = (const char *)p;
        while (nb) {
            const unsigned piece = (unsigned)std::min<uint64_t>(nb, 1u << 30);
            if (gzwrite(f, c, piece) != (int)piece) MCG_THROW(MCG_EIO, "Failed to write to file");
            c += piece;
            nb -= piece;
        }
    }
    void get(void *p, uint64_t nb) {
        char *c = (char *)p;
        while (nb) {
            const unsigned piece = (unsigned)std::min<uint64_t>(nb, 1u << 30);
            if (gzread(f, c, piece) != (int)piece) MCG_THROW(MCG_EIO, "Failed to read from file");
            c += piece;
            nb -= piece;
        }
    }
    // reads and drops nb bytes
    void skip(uint64_t nb) {
        std::vector<char> tmp((size_t)std::min<uint64_t>(nb, 1u << 20));
        while (nb) {
            const uint64_t piece = std::min<uint64_t>(nb, tmp.size());
            get(tmp.data(), piece);
            nb -= piece;
        }
    }
    void close() {
        const int rc = gzclose(f);
        f = nullptr;
        if (rc != Z_OK) MCG_THROW(MCG_EIO, "Failed to close file");
    }
};

uint64_t index_bytes(int dt) {
    switch (dt) {
    case MCG_U8: return 1;
    case MCG_U16: return 2;
    case MCG_U32: return 4;
    case MCG_U64: return 8;
    default: MCG_THROW(MCG_EINVAL, "index dtype must be u8, u16, u32 or u64");
    }
}

uint64_t weight_bytes(int dt) {
    switch (dt) {
    case MCG_F32: return 4;
    case MCG_F64: return 8;
    default: MCG_THROW(MCG_EINVAL, "weight dtype must be f32 or f64");
    }
}

// element counts past this are taken as a corrupt header, not an allocation request
constexpr uint64_t MAX_ELEMS = 1ull << 40;

}  // namespace
}  // namespace mcg

using namespace mcg;

extern "C" {

int mcg_sampler_write(const char *path, uint64_t n, uint64_t seed, const float *probs, const float *weights) {
    return guard([&] {
        REQUIRE(probs || n == 0, "null probs");
        GzFile g(path, "wb");
        g.put(&n, 8);
        g.put(&seed, 8);
        g.put(probs, n * 4);
        const uint32_t flag = weights ? 137u : 0u;
        g.put(&flag, 4);
        if (weights) g.put(weights, n * 4);
        g.close();
    });
}

int mcg_sampler_read(const char *path, uint64_t *n, uint64_t *seed, int *has_weights, float *probs,
                     float *weights, uint64_t cap) {
    return guard([&] {
        REQUIRE(n && seed && has_weights, "null output");
        GzFile g(path, "rb");
        uint64_t nn = 0, sd = 0;
        g.get(&nn, 8);
        g.get(&sd, 8);
        REQUIRE(nn <= MAX_ELEMS, "Failed to read from file: implausible sampler size");
        const bool fill = cap >= nn;
        REQUIRE(!fill || probs || nn == 0, "null probs");
        if (fill) g.get(probs, nn * 4);
        else g.skip(nn * 4);
        uint32_t flag = 0;
        g.get(&flag, 4);
        if (flag) {
            if (fill && weights) g.get(weights, nn * 4);
            else g.skip(nn * 4);
        }
        *n = nn;
        *seed = sd;
        *has_weights = flag ? 1 : 0;
    });
}

int mcg_index_coreset_write(const char *path, uint64_t m, const void *indices, int idx_dtype, const void *weights,
                            int w_dtype) {
    return guard([&] {
        const uint64_t ib = index_bytes(idx_dtype), wb = weight_bytes(w_dtype);
        REQUIRE((indices && weights) || m == 0, "null indices or weights");
        GzFile g(path, "wb");
        g.put(&m, 8);
        g.put(indices, m * ib);
        g.put(weights, m * wb);
        g.close();
    });
}

int mcg_index_coreset_read(const char *path, uint64_t *m, void *indices, int idx_dtype, void *weights, int w_dtype,
                           uint64_t cap) {
    return guard([&] {
        REQUIRE(m, "null output");
        const uint64_t ib = index_bytes(idx_dtype), wb = weight_bytes(w_dtype);
        GzFile g(path, "rb");
        uint64_t mm = 0;
        g.get(&mm, 8);
        REQUIRE(mm <= MAX_ELEMS, "Failed to read from file: implausible coreset size");
        if (cap >= mm) {
            REQUIRE((indices && weights) || mm == 0, "null indices or weights");
            g.get(indices, mm * ib);
            g.get(weights, mm * wb);
        } else {
            g.skip(mm * (ib + wb));
        }
        *m = mm;
    });
}

}  // extern "C"
