/*
 * sad_device.h - gfx950 device helpers shared by the SAD kernels, plus the
 * launch/status plumbing of the C ABI.
 */
#ifndef SAD_DEVICE_H
#define SAD_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "sad_math.h"
#include "../../include/sad_hip.h"

namespace sadh {

/* ------------------------------------------------------------ status ---- */

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
/* Status of the launch just issued (hipGetLastError), recorded for sadh_last_error. */
int launch_status(const char* what);

#define SADH_REQUIRE(cond, msg) \
    do { if (!(cond)) return ::sadh::fail(SADH_ERR_INVALID, msg); } while (0)

inline hipStream_t to_hip(sadh_stream_t s) { return reinterpret_cast<hipStream_t>(s); }
inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

/* ------------------------------------------------------ fast arithmetic --
 * Hardware transcendental approximations (v_exp_f32, v_sqrt_f32, v_rsq_f32,
 * v_rcp_f32; about 1 ulp) for the render and gradient kernels, whose parity bar
 * is a tolerance.  The candidate search never uses them. */
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

/* binary16 -> binary32 with the hardware converter (exact, fp16 denormals on) */
__device__ __forceinline__ float h2f(uint32_t bits16)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)bits16);
}

/* ------------------------------------------------ 8-lane group reductions --
 * Lanes 8g..8g+7 of a wave form one group (one pixel, eight candidates).
 * DPP quad_perm [1,0,3,2], [2,3,0,1] and row_half_mirror combine the group in
 * three steps; every lane of the group ends with the same bits. */
template <int CTRL>
__device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float group8_sum(float v)
{
    v = v + dpp<0xB1>(v);
    v = v + dpp<0x4E>(v);
    v = v + dpp<0x141>(v);
    return v;
}
__device__ __forceinline__ float group8_max(float v)
{
    v = fmaxf(v, dpp<0xB1>(v));
    v = fmaxf(v, dpp<0x4E>(v));
    v = fmaxf(v, dpp<0x141>(v));
    return v;
}

/* Block index -> tile index, grouping consecutive tiles on one XCD.  Workgroups
 * are dealt to the 8 XCDs round-robin, so without the remap neighbouring tiles
 * (which share most candidate sites) land in different L2s. */
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t ntiles)
{
    uint32_t full = ntiles & ~7u;
    if (b >= full) return b;
    return (b & 7u) * (full >> 3) + (b >> 3);
}

}  // namespace sadh

#endif /* SAD_DEVICE_H */
