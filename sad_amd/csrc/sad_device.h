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

/* A device workspace of at least `bytes` for calls on `stream` (scratch.cpp): owned by the library, reused by
 * later calls on the same stream and device. */
int stream_workspace(hipStream_t stream, size_t bytes, void** out);

/* ------------------------------------------------------ fast arithmetic --
 * Hardware transcendental approximations (v_exp_f32, v_sqrt_f32, v_rsq_f32,
 * v_rcp_f32; about 1 ulp) for the render and gradient kernels, whose parity bar
 * is a tolerance.  The candidate search never uses them. */
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

/* --------------------------------------------------- exact cross-tile sums --
 * A float partial v joins a cross-tile sum as two int64 words added with
 * integer atomics: hi = trunc(v), the integer part, and lo = the fraction times
 * 2^40 rounded to the nearest integer (ties to even).  Integer addition is
 * associative, so a sum does not depend on the order in which the tiles add
 * to it.  It is read back as (double)hi + (double)lo * 2^-40, rounded once to
 * float.
 *   range:      partials with |v| < 2^62 are held exactly down to 2^-40;
 *               larger magnitudes saturate at +-2^62 and NaN adds nothing;
 *   headroom:   |lo| < 2^40 per partial, so the lo word takes 2^23 partials
 *               (tiles of one site in one launch) before it could overflow:
 *               images up to 2^31 pixels;
 *   resolution: 2^-40 (9.1e-13) per partial, so a sum of P partials is within
 *               P * 2^-41 of their exact sum; far below what Adam resolves
 *               after the 1/(W*H) gradient normalisation.
 * A site's values are stored as (lo, hi) pairs: gradient lines of
 * FIX_GRAD_STRIDE words (values 0-9 in Site field order, then 4 unused words,
 * 192 B, three 64-byte segments), removal deltas as one pair per site,
 * densify statistics as lines of FIX_STAT_STRIDE words (8 values). */
constexpr uint32_t FIX_GRAD_STRIDE = 24u;
constexpr uint32_t FIX_STAT_STRIDE = 16u;

__device__ __forceinline__ void fix_split(float v, long long* lo, long long* hi)
{
    if (!(v == v)) v = 0.0f;
    v = fminf(fmaxf(v, -0x1p62f), 0x1p62f);
    const float t = truncf(v);
    *hi = (long long)t;
    *lo = (long long)rintf((v - t) * 0x1p40f);
}

/* A pointer loaded from device memory (a job table) is generic to the compiler,
 * so accesses through it compile to generic (flat) memory instructions.
 * Rebuilt as an offset from a kernel argument (which the compiler knows to be
 * global memory), the same address is accessed with global instructions.  A
 * null pointer stays null. */
template <class T, class B>
__device__ __forceinline__ T* rebase(const B* base, T* p)
{
    const char* b = reinterpret_cast<const char*>(base);
    return reinterpret_cast<T*>(const_cast<char*>(b + (reinterpret_cast<uintptr_t>(p) - reinterpret_cast<uintptr_t>(b))));
}

/* one word of fix_split(v): hi (hiw = 1) or lo.  Both are integral floats below
 * 2^62 in magnitude, so one conversion of the selected one gives the same word. */
__device__ __forceinline__ long long fix_word(float v, uint32_t hiw)
{
    if (!(v == v)) v = 0.0f;
    v = fminf(fmaxf(v, -0x1p62f), 0x1p62f);
    const float t = truncf(v);
    return (long long)(hiw ? t : rintf((v - t) * 0x1p40f));
}

__device__ __forceinline__ float fix_value(long long lo, long long hi)
{
    return (float)((double)hi + (double)lo * 0x1p-40);
}

__device__ __forceinline__ void fix_atomic_add(long long* p, long long v)
{
    if (v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

/* One column of (lo, hi) pair lines read as floats: element i of the column is
 * the pair at p + i * stride (the tau passes, score pairs and split read their
 * inputs through this, with the float-array code paths unchanged). */
struct FixCol {
    const long long* p;
    uint32_t stride;
    __device__ __forceinline__ float operator[](uint32_t i) const
    {
        const longlong2 w = *reinterpret_cast<const longlong2*>(p + (size_t)i * stride);
        return fix_value(w.x, w.y);
    }
};

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
    /* every lane of the patterns used here reads an existing lane, so bound_ctrl changes no value; with it set
     * the compiler folds the move into the consuming add or max (v_add_f32_dpp) */
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
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

/* Tau-pass neighbour list of site i (tau_diffuse_kernel's neighbour set,
 * tau_diffusion.cu:27-89): the unique active ids other than i in the candidate
 * list of the site's home cell, as 8 ids in candidate order (SAD_EMPTY for ids
 * that do not count), or NBR_COPY in slot 0 when the pass copies its input
 * through (lambda <= 0 or i inactive). */
#define NBR_COPY 0xFFFFFFFEu

__device__ __forceinline__ void tau_list_only(uint32_t i, const uint4* __restrict__ c0, const uint4* __restrict__ c1,
                                              const SadSite* __restrict__ sites, uint32_t n, float lambda, int width,
                                              int height, float inv_ds, uint4* __restrict__ nbr)
{
    const float2 pos = reinterpret_cast<const float2*>(sites + i)[0];
    if (lambda <= 0.0f || pos.x < 0.0f) {
        nbr[2u * i] = make_uint4(NBR_COPY, SAD_EMPTY, SAD_EMPTY, SAD_EMPTY);
        nbr[2u * i + 1u] = make_uint4(SAD_EMPTY, SAD_EMPTY, SAD_EMPTY, SAD_EMPTY);
        return;
    }
    const int px = sad_clampi((int)(pos.x * inv_ds + 0.5f), 0, width - 1);
    const int py = sad_clampi((int)(pos.y * inv_ds + 0.5f), 0, height - 1);
    const uint4 a = c0[py * width + px], b = c1[py * width + px];
    const uint32_t ids[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t keep[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t nb = ids[k];
        bool ok = nb < n && nb != i;
#pragma unroll
        for (int j = 0; j < k; ++j) ok = ok && ids[j] != nb;
        keep[k] = (ok && sites[nb].pos_x >= 0.0f) ? nb : SAD_EMPTY;
    }
    nbr[2u * i] = make_uint4(keep[0], keep[1], keep[2], keep[3]);
    nbr[2u * i + 1u] = make_uint4(keep[4], keep[5], keep[6], keep[7]);
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
