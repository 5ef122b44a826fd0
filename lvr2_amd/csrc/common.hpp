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

// Cross-lane reads within rows of 16 lanes by DPP (a modifier of the VALU instruction: no LDS
// round trip, unlike a lane shuffle).  Every lane of the wave must be active.
constexpr int DPP_QUAD_XOR1 = 0xB1;     // quad_perm [1, 0, 3, 2]: lane i <- lane i ^ 1
constexpr int DPP_QUAD_XOR2 = 0x4E;     // quad_perm [2, 3, 0, 1]: lane i <- lane i ^ 2
constexpr int DPP_HALF_MIRROR = 0x141;  // lane i <- lane 7 - i within each 8 lanes
constexpr int DPP_ROW_MIRROR = 0x140;   // lane i <- lane 15 - i within each 16 lanes
constexpr int DPP_ROW_SHR = 0x110;      // + n (1..15): lane i <- lane i - n within each 16 lanes
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// op over each aligned group of G lanes, the result on every lane of the group.  Each step
// combines a lane's partial with the partial of the group's other half (xor 1, xor 2, then the
// mirrors, which pair every lane with one of the other half), so a sum adds the same two partials
// in the same order on every lane as the xor butterfly does: the results are identical to it.
template <int G, class Op>
__device__ __forceinline__ float lane_group_reduce(float v, Op op)
{
    static_assert(G >= 1 && G <= kWave && (G & (G - 1)) == 0, "a power-of-two group");
    if constexpr (G >= 2) v = op(v, dpp_f<DPP_QUAD_XOR1>(v));
    if constexpr (G >= 4) v = op(v, dpp_f<DPP_QUAD_XOR2>(v));
    if constexpr (G >= 8) v = op(v, dpp_f<DPP_HALF_MIRROR>(v));
    if constexpr (G >= 16) v = op(v, dpp_f<DPP_ROW_MIRROR>(v));
    if constexpr (G >= 32) v = op(v, shfl_xor(v, 16));
    if constexpr (G >= 64) v = op(v, shfl_xor(v, 32));
    return v;
}

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
// Integer sum over each aligned group of G <= 16 lanes (see lane_group_reduce).
template <int G>
__device__ __forceinline__ int lane_group_sum_int(int v)
{
    static_assert(G >= 1 && G <= 16 && (G & (G - 1)) == 0, "a power-of-two group within a row");
    if constexpr (G >= 2) v += dpp_i<DPP_QUAD_XOR1>(v);
    if constexpr (G >= 4) v += dpp_i<DPP_QUAD_XOR2>(v);
    if constexpr (G >= 8) v += dpp_i<DPP_HALF_MIRROR>(v);
    if constexpr (G >= 16) v += dpp_i<DPP_ROW_MIRROR>(v);
    return v;
}
// Inclusive prefix sum over each aligned group of G <= 16 lanes (lane s of a group adds lanes
// 0..s): shifts within the row (a lane shifted in from outside the row reads 0), a lane taking
// the shifted value only from its own group.
template <int G>
__device__ __forceinline__ int lane_group_scan_int(int v, int s)
{
    static_assert(G >= 1 && G <= 16 && (G & (G - 1)) == 0, "a power-of-two group within a row");
    if constexpr (G >= 2) {
        const int y = __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR + 1, 0xF, 0xF, true);
        if (s >= 1) v += y;
    }
    if constexpr (G >= 4) {
        const int y = __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR + 2, 0xF, 0xF, true);
        if (s >= 2) v += y;
    }
    if constexpr (G >= 8) {
        const int y = __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR + 4, 0xF, 0xF, true);
        if (s >= 4) v += y;
    }
    if constexpr (G >= 16) {
        const int y = __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR + 8, 0xF, 0xF, true);
        if (s >= 8) v += y;
    }
    return v;
}

__device__ __forceinline__ float readlane_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Minimum / maximum over the wave (every lane active), wave-uniform: per row of 16 lanes by DPP,
// then the four rows' results read as scalars.
__device__ __forceinline__ float wave_min(float v)
{
    v = lane_group_reduce<16>(v, [](float a, float b) { return fminf(a, b); });
    return fminf(fminf(readlane_f(v, 0), readlane_f(v, 16)), fminf(readlane_f(v, 32), readlane_f(v, 48)));
}
__device__ __forceinline__ float wave_max(float v)
{
    v = lane_group_reduce<16>(v, [](float a, float b) { return fmaxf(a, b); });
    return fmaxf(fmaxf(readlane_f(v, 0), readlane_f(v, 16)), fmaxf(readlane_f(v, 32), readlane_f(v, 48)));
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
