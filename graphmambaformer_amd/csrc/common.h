// Shared helpers for the libgmfa HIP sources: error plumbing and CDNA
// cross-lane primitives (DPP) used by the SW and chaining kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "gmfa.h"

namespace gmfa {

void set_error(const char *fmt, ...);

#define GMFA_CHECK_HIP(expr)                                                          \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            ::gmfa::set_error("%s failed: %s", #expr, hipGetErrorString(_e));        \
            return GMFA_E_RUNTIME;                                                    \
        }                                                                             \
    } while (0)

#define GMFA_REQUIRE(cond, ...)                                                       \
    do {                                                                              \
        if (!(cond)) {                                                                \
            ::gmfa::set_error(__VA_ARGS__);                                           \
            return GMFA_E_INVALID;                                                    \
        }                                                                             \
    } while (0)

constexpr int kNeg = -(1 << 28);  // minus infinity of the int32 DP

// DPP controls (GFX9 encoding, gfx950 included)
constexpr int kRowShr1 = 0x111;
constexpr int kRowShr2 = 0x112;
constexpr int kRowShr4 = 0x114;
constexpr int kRowShr8 = 0x118;
constexpr int kWaveShl1 = 0x130;
constexpr int kWaveShr1 = 0x138;
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ int dpp(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, CTRL, ROW_MASK, BANK_MASK, false);
}

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Inclusive max-scan over segments of G lanes (G in {16, 32, 64}).  The DPP
// "old" operand is INT_MIN, the identity of signed max, which lets the
// compiler fuse each step into a single v_max_i32_dpp.
template <int G>
__device__ __forceinline__ int seg_scan_max(int x)
{
    constexpr int id = (int)0x80000000;
    x = max(x, dpp<kRowShr1>(id, x));
    x = max(x, dpp<kRowShr2>(id, x));
    x = max(x, dpp<kRowShr4>(id, x));
    x = max(x, dpp<kRowShr8>(id, x));
    if constexpr (G >= 32) x = max(x, dpp<kRowBcast15, 0xa>(id, x));
    if constexpr (G >= 64) x = max(x, dpp<kRowBcast31, 0xc>(id, x));
    return x;
}

// value of lane-1 (kNeg into lane 0 of the wave)
__device__ __forceinline__ int from_prev_lane(int x) { return dpp<kWaveShr1>(kNeg, x); }
// value of lane-1, 0 into lane 0 (DPP bound_ctrl: no "old" operand to set up)
__device__ __forceinline__ int from_prev_lane_or0(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, kWaveShr1, 0xf, 0xf, true);
}
// value of lane+1 (old into lane 63)
__device__ __forceinline__ int from_next_lane(int old, int x) { return dpp<kWaveShl1>(old, x); }

__device__ __forceinline__ uint32_t insert_sign(uint32_t acc, int x)
{
    // (acc << 1) | (x >>> 31): one v_alignbit_b32
    return __builtin_amdgcn_alignbit(acc, (uint32_t)x, 31);
}

}  // namespace gmfa
