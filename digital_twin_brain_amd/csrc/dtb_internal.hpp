// Internal helpers of libdtb: status handling, Philox4x32-10, wavefront
// primitives.  gfx950 only (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "dtb.h"

namespace dtb {

void set_error(const char* fmt, ...);

#define DTB_TRY_HIP(expr)                                                                         \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) {                                                                   \
            ::dtb::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return _e == hipErrorOutOfMemory ? DTB_OUT_OF_MEMORY : DTB_RUNTIME_ERROR;             \
        }                                                                                         \
    } while (0)

#define DTB_REQUIRE(cond, ...)                 \
    do {                                       \
        if (!(cond)) {                         \
            ::dtb::set_error(__VA_ARGS__);     \
            return DTB_INVALID_PARAMETER;      \
        }                                      \
    } while (0)

#define DTB_TRY(expr)                  \
    do {                               \
        int _s = (expr);               \
        if (_s != DTB_OK) return _s;   \
    } while (0)

constexpr int kWave = 64;

// ------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11): 10 rounds of two 32x32->64 multiplies with Weyl key bumps.
struct u32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        u32x4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// 24-bit uniform in [0, 1), exact in float32
__host__ __device__ inline float u24(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// Counter word 3 selects the stream.
constexpr uint32_t kStreamNoise = 0x4E4F4953u; // ctr = (i_lo, i_hi, step, NOISE)
constexpr uint32_t kStreamConn = 0x434F4E4Eu;  // ctr = (k, post_lo, post_hi, CONN)
constexpr uint32_t kStreamGamma = 0x47414D4Du; // ctr = (neuron_lo, neuron_hi, draw, GAMMA)
constexpr uint32_t kStreamExt = 0x45585420u;   // ctr = (i_lo, i_hi, step, EXT)

// External Poisson drive: per neuron and step, count = #{j < n : r >= thr[j]}
// with r one 32-bit Philox word and thr[j] = 2^32 P(X <= j) (host-computed,
// so the count is integer-exact everywhere); the count times w arrives on the
// excitatory channels.
constexpr int kMaxExtThr = 16;
struct ExtDrive {
    int32_t n;
    float w;
    uint32_t thr[kMaxExtThr];
};

__host__ __device__ inline uint32_t ext_count(const ExtDrive& e, uint64_t seed, uint64_t step, uint64_t i)
{
    const u32x4 c{(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)step, kStreamExt};
    const uint32_t r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32)).x;
    uint32_t k = 0;
    for (int j = 0; j < e.n; ++j) k += r >= e.thr[j] ? 1u : 0u;
    return k;
}

__host__ __device__ inline float noise_uniform(uint64_t seed, uint64_t step, uint64_t i)
{
    u32x4 c{(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)step, kStreamNoise};
    return u24(philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32)).x);
}

// ------------------------------------------------------------ synapse word
// 32-bit packed synapse target: [post : idx_bits][delay-1 : dbits][kind : 1 @ bit 31]
struct SynPack {
    uint32_t idx_bits;
    uint32_t dbits;
    uint32_t idx_mask;
    uint32_t dmask;
};

inline uint32_t bits_for(uint64_t v)
{
    uint32_t b = 0;
    while ((1ull << b) < v) ++b;
    return b;
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

} // namespace dtb
