// common.hpp -- shared host/device definitions of libtdx (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "../../include/tdx.h"

namespace tdx {

// ---- status plumbing -------------------------------------------------------
int record(int status);
int record_hip(hipError_t e, const char *what, const char *file, int line);
hipStream_t stream();
// compute units of the current device (cached per device)
int device_cus();

#define TDX_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) return tdx::record_hip(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define TDX_CHECK_ARG(cond)                          \
    do {                                             \
        if (!(cond)) return tdx::record(TDX_ERR_ARG); \
    } while (0)

#define TDX_LAUNCH_CHECK() TDX_HIP(hipGetLastError())

// ---- workspace pool: grow-only named slots, reused across calls --------------
enum WsSlot {
    WS_PARTIAL = 0,  // MTTKRP split-K partial sums
    WS_PANEL0,       // factor panels (one slot per mode)
    WS_PANEL1,
    WS_PANEL2,
    WS_PANEL3,
    WS_MISC0,
    WS_MISC1,
    WS_MISC2,
    WS_MISC3,
    WS_STORE,        // CP-ALS: the BF16 / FP8 copy of the tensor for the main phase
    WS_STORE_SCALES, // CP-ALS: its FP8 slice scales
    WS_STORE_BF16,   // CP-ALS, FP8 storage with split-operand opening sweeps: the BF16 copy
    WS_COUNT
};
int workspace(int slot, size_t bytes, void **ptr);
void release_workspace();
// frees the cached CP-ALS plans (als.hip); tdx_release_workspace calls it
void release_cp_plans();
// the fixed device scratch of tdx_mttkrp_ex_noalloc (64 MiB per GPU): its address on the
// current device, taken once per device by hipGetSymbolAddress (mttkrp.hip)
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
int noalloc_scratch(void **ptr);
// FP8 mode-1 slice quantisation in steps (elementwise.hip): slice_amax zeroes and fills amax[I];
// fp8_from_amax writes the scales (max / 448, 1 for an all-zero slice) and the E4M3 codes
int slice_amax(const float *X, uint32_t *amax, long I, long n_rest, hipStream_t s);
int fp8_from_amax(const float *X, const uint32_t *amax, uint8_t *Xq, float *scales, long I, long n_rest,
                  hipStream_t s);

__host__ __device__ inline long ceil_div(long a, long b) { return (a + b - 1) / b; }

// Streaming (nontemporal) loads for data read exactly once per kernel, such as the tensor in
// MTTKRP: on MI355X a read stream of 16-byte nontemporal loads runs at ~7.1 TB/s against
// ~6.3 TB/s with the default cache policy.
__device__ __forceinline__ uint4 ld_stream(const uint4 *p) {
    typedef unsigned v4 __attribute__((ext_vector_type(4)));
    const v4 t = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(p));
    return uint4{t.x, t.y, t.z, t.w};
}
__device__ __forceinline__ float4 ld_stream(const float4 *p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 t = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(p));
    return float4{t.x, t.y, t.z, t.w};
}
__device__ __forceinline__ uint2 ld_stream(const uint2 *p) {
    typedef unsigned v2 __attribute__((ext_vector_type(2)));
    const v2 t = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
    return uint2{t.x, t.y};
}

// ---- device helpers ------------------------------------------------------------
__device__ __forceinline__ uint32_t f2u(float x) { return __float_as_uint(x); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

// IEEE round-to-nearest-even FP32 -> BF16 (NaN kept quiet); identical to the oracle
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float x) {
    uint32_t u = f2u(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// OCP E4M3FN, RNE, saturating at 448, NaN -> 0x7f (software rounding, bit-exact by construction)
__device__ __forceinline__ uint8_t f32_to_e4m3_bits(float x) {
    uint32_t u = f2u(x);
    uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    uint32_t au = u & 0x7fffffffu;
    if (au > 0x7f800000u) return (uint8_t)(sign | 0x7f);
    float ax = u2f(au);
    if (ax >= 464.0f) return (uint8_t)(sign | 0x7e);
    if (ax < 0.015625f) {
        // quantum 2^-9: scale exactly into integer+fraction, round half to even
        float v = ax * 512.0f;  // exact (power of two scaling, no underflow for normals >= 2^-126 * 512)
        float fl = floorf(v);
        float rem = v - fl;
        uint32_t m = (uint32_t)fl;
        if (rem > 0.5f || (rem == 0.5f && (m & 1u))) m++;
        return (uint8_t)(sign | m);
    }
    int e = (int)((au >> 23) & 0xff) - 127;
    uint32_t mant = au & 0x7fffffu;
    uint32_t q = ((uint32_t)(e + 7) << 3) | (mant >> 20);
    uint32_t rem = mant & 0xfffffu;
    if (rem > 0x80000u || (rem == 0x80000u && (q & 1u))) q++;
    if (q > 0x7e) q = 0x7e;
    return (uint8_t)(sign | q);
}

// Philox-4x32-10 (Salmon et al., SC'11) -- same stream definition as the oracle
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c[0];
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ uint32_t philox_word(uint64_t seed, uint32_t strm, uint64_t n) {
    uint64_t blk = n >> 2;
    uint32_t c[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), strm, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return c[n & 3];
}

__device__ __forceinline__ float word_to_uniform(uint32_t w) {
    return ((float)(w >> 9) + 0.5f) * (1.0f / 8388608.0f);
}

}  // namespace tdx
