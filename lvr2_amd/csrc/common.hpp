// Shared device/host helpers for the lvrhip library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace lvrhip {

constexpr int kWave = 64;

// Error carried from the HIP runtime up to the C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define LVRHIP_CHECK(expr)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            throw ::lvrhip::Error(e_ == hipErrorOutOfMemory ? 2 : 3,                              \
                                  std::string(hipGetErrorString(e_)) + " in " + __FILE__ +        \
                                      " at line " + std::to_string(__LINE__) +                   \
                                      (e_ == hipErrorOutOfMemory ? " (out of memory)" : ""));   \
        }                                                                                        \
    } while (0)

#define LVRHIP_LAUNCH_CHECK() LVRHIP_CHECK(hipGetLastError())

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- device helpers ------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits of m below this lane
__device__ __forceinline__ int mbcnt(uint64_t m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t fbits(float x) { return __float_as_uint(x); }
__device__ __forceinline__ float bitsf(uint32_t x) { return __uint_as_float(x); }

// min / max of numbers known not to be NaN (a compare and a select; fminf / fmaxf also
// quiet signalling NaNs, which costs two more instructions each)
__device__ __forceinline__ float min_num(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float max_num(float a, float b) { return b > a ? b : a; }

// wave-uniform value in an SGPR
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float uniformf(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

template <typename T>
__device__ __forceinline__ T shfl(T v, int src)
{
    return __shfl(v, src, kWave);
}
template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int m)
{
    return __shfl_xor(v, m, kWave);
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fminf(v, shfl_xor(v, m));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, shfl_xor(v, m));
    return v;
}

__device__ __forceinline__ int wave_sum_int(int v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor(v, m);
    return v;
}

// Squared distance of the kNN contract: (dx*dx + dy*dy) + dz*dz in fp32 with every
// operation rounded (the library is built with -ffp-contract=off).
__device__ __forceinline__ float d2_contract(float qx, float qy, float qz, float px, float py, float pz)
{
    float dx = qx - px;
    float dy = qy - py;
    float dz = qz - pz;
    return (dx * dx + dy * dy) + dz * dz;
}

__device__ __forceinline__ uint64_t make_key(float d2, uint32_t idx)
{
    return ((uint64_t)fbits(d2) << 32) | (uint64_t)idx;
}

}  // namespace lvrhip
