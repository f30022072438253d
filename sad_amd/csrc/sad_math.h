/*
 * sad_math.h - types and exact arithmetic primitives of the SAD fit/render path.
 *
 * This header is plain C99 and HIP at the same time: every function is marked
 * SAD_HD, which is `__host__ __device__` under hipcc and `static inline` under a
 * C compiler.  The candidate search (JFA flood, VPT) must produce bit-identical
 * candidate lists on the GPU and on a CPU, so every floating-point operation it
 * performs comes from this file, is written with explicit fmaf where a fused
 * multiply-add is wanted, and is compiled with -ffp-contract=off on both sides.
 * IEEE division and sqrtf are correctly rounded on both targets (hipcc lowers
 * fp32 `/` and sqrtf to the div_scale/div_fmas/div_fixup and corrected
 * v_sqrt sequences), so they are used directly.
 *
 * Site model (reference: backends/cuda/src/sad_common.cuh:21-58):
 *   Site       10 fp32 = pos.xy, log_tau, radius, rgb, aniso_dir.xy, log_aniso (40 B)
 *   AdamState  20 fp32 = m/v per parameter group, pos and dir share one v (80 B).
 *              The CUDA struct is 88 B because float2 members are 8-byte aligned;
 *              the torch op contract allocates [N,20] (backends/pytorch/train.py:836,918),
 *              so this build packs the 18 state floats without the alignment hole.
 */
#ifndef SAD_MATH_H
#define SAD_MATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SAD_HD static __host__ __device__ __forceinline__
#else
#define SAD_HD static inline
#endif

#include "../../include/sad_types.h"

/* ---------------------------------------------------------------- bits -- */

SAD_HD uint32_t sad_f2u(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
SAD_HD float sad_u2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }

SAD_HD int sad_clampi(int x, int a, int b) { return x < a ? a : (x > b ? b : x); }
SAD_HD float sad_clampf(float x, float a, float b) { return fmaxf(a, fminf(b, x)); }

/* ----------------------------------------------------------- exp / log --
 * sad_expf: Cody & Waite (1980) argument reduction x = k*ln2 + r, |r| <= ln2/2,
 * with ln2 split into a 16-bit head and a tail, followed by the degree-7 Taylor
 * polynomial of e^r in Horner form (Muller, "Elementary Functions", 3rd ed.,
 * ch. 11).  Truncation error r^8/8! < 8e-9, so the result is within ~1 ulp of
 * e^x.  2^k is applied in two exact power-of-two steps so that gradual
 * underflow rounds once.  Only +, *, fmaf and integer operations are used, so
 * the result is bit-identical on every IEEE binary32 target. */
SAD_HD float sad_pow2i(int k) { return sad_u2f((uint32_t)(k + 127) << 23); }

/* e^xc for xc in [-104, 89] (the caller's clamp): the reduction, the
 * polynomial and 2^k in two exact power-of-two steps.  k is split as
 * floor(k/2) + the rest; any split whose first factor keeps p * 2^k1 normal
 * gives the same, once-rounded product. */
SAD_HD float sad_exp_core(float xc)
{
    const float L2E = 0x1.715476p+0f;   /* log2(e) */
    const float LN2_HI = 0x1.62e400p-1f;   /* ln2 head, 16 significant bits */
    const float LN2_LO = 0x1.7f7d1cp-20f;  /* ln2 - head */
    const float SHIFTER = 12582912.0f;                     /* 1.5 * 2^23 */
    float t = xc * L2E;
#if defined(__HIP_DEVICE_COMPILE__)
    /* v_rndne_f32: the same round-to-nearest-even of t in one instruction */
    float kf = __builtin_rintf(t);
    (void)SHIFTER;
#else
    float kf = (t + SHIFTER) - SHIFTER;          /* round-to-nearest-even of t */
#endif
    float r = fmaf(-kf, LN2_HI, xc);
    r = fmaf(-kf, LN2_LO, r);
    float p = 0x1.a01a02p-13f;   /* 1/5040 */
    p = fmaf(p, r, 0x1.6c16c2p-10f);  /* 1/720 */
    p = fmaf(p, r, 0x1.111112p-7f);   /* 1/120 */
    p = fmaf(p, r, 0x1.555556p-5f);   /* 1/24 */
    p = fmaf(p, r, 0x1.555556p-3f);   /* 1/6 */
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int k = (int)kf;
#if defined(__HIP_DEVICE_COMPILE__)
    /* v_ldexp_f32 scales by 2^k with one IEEE rounding (also into the
     * subnormal range), which is the once-rounded product below; over
     * every float input both forms give the same bits on gfx950 */
    return __builtin_amdgcn_ldexpf(p, k);
#else
    int k1 = k >> 1;                             /* arithmetic shift: floor(k / 2), at least -75 */
    int k2 = k - k1;
    return (p * sad_pow2i(k1)) * sad_pow2i(k2);
#endif
}

SAD_HD float sad_expf(float x)
{
    /* The polynomial path runs on x clamped to [-104, 89] (NaN to -104), so its
     * integer conversions stay in range; the special cases then select their
     * results.  Inside [-103.97.., 88.72..] the clamp changes nothing, and there
     * are no branches for a GPU to diverge on. */
    float e = sad_exp_core(fminf(fmaxf(x, -104.0f), 89.0f));
    e = (x < -103.97208404541015625f) ? 0.0f : e;
    e = (x > 0x1.62e430p+6f) ? INFINITY : e;
    return (x == x) ? e : x + x;               /* NaN */
}

/* sad_expf for x <= 0 or NaN, with the same bits: the softmax's logit - max,
 * which never needs the overflow case. */
SAD_HD float sad_expf_nonpos(float x)
{
    float e = sad_exp_core(fmaxf(x, -104.0f));
    e = (x < -103.97208404541015625f) ? 0.0f : e;
    return (x == x) ? e : x + x;
}

/* sqrtf of x >= 2^-96 (or +inf), correctly rounded.  On the GPU: the hardware
 * square root (within 1 ulp) and one correction step that compares the
 * residuals x - s'*s of the neighbouring floats s' = s -/+ 1 ulp with zero,
 * which is the sequence the compiler emits for sqrtf, less the scaling of tiny
 * inputs and the zero/infinity fix-up that this range does not need.  The
 * soft distance's argument is at least 1e-8.  On a CPU: sqrtf. */
SAD_HD float sad_sqrt_pos(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = sad_u2f(sad_f2u(s) - 1u);
    const float su = sad_u2f(sad_f2u(s) + 1u);
    const float t = (fmaf(-sd, s, x) <= 0.0f) ? sd : s;
    return (fmaf(-su, s, x) > 0.0f) ? su : t;
#else
    return sqrtf(x);
#endif
}

/* sad_logf (x > 0): x = m * 2^e with m in [sqrt(1/2), sqrt(2)), then
 * log(m) = 2 atanh(s), s = (m-1)/(m+1), summed as s + s^3/3 + ... + s^11/11
 * (|s| < 0.1716, truncation < 2e-10 relative).  Used for pow() in the densify
 * score, where CPU and GPU must agree bit for bit. */
SAD_HD float sad_logf(float x)
{
    if (!(x == x)) return x + x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    uint32_t u = sad_f2u(x);
    int e = 0;
    if (u < 0x00800000u) {                       /* subnormal: scale by 2^24 */
        x = x * 16777216.0f;
        u = sad_f2u(x);
        e = -24;
    }
    e += (int)(u >> 23) - 127;
    float m = sad_u2f((u & 0x007FFFFFu) | 0x3F800000u);    /* [1,2) */
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float q = 0x1.745d18p-4f;   /* 1/11 */
    q = fmaf(q, s2, 0x1.c71c72p-4f);  /* 1/9 */
    q = fmaf(q, s2, 0x1.24924ap-3f);  /* 1/7 */
    q = fmaf(q, s2, 0x1.99999ap-3f);  /* 1/5 */
    q = fmaf(q, s2, 0x1.555556p-2f);  /* 1/3 */
    float lm = 2.0f * fmaf(s * s2, q, s);
    const float LN2_HI = 0x1.62e400p-1f;
    const float LN2_LO = 0x1.7f7d1cp-20f;
    float ef = (float)e;
    return fmaf(ef, LN2_HI, fmaf(ef, LN2_LO, lm));
}

/* b^t for an integer t (Adam's bias corrections beta^t): binary powering in
 * double, rounded once to float.  Double products are exact IEEE roundings on
 * every target, and their accumulated error (under 2^-48 relative for t < 2^32)
 * leaves the float result the correctly rounded one except within that
 * distance of a rounding boundary, so hosts of different libm give the same bits. */
SAD_HD float sad_powi_f(float b, uint32_t t)
{
    double r = 1.0, x = (double)b;
    while (t) {
        if (t & 1u) r *= x;
        x *= x;
        t >>= 1;
    }
    return (float)r;
}

/* x^a for x > 0 (pow of a positive base), as exp(a * log x). */
SAD_HD float sad_powf_pos(float x, float a) { return sad_expf(a * sad_logf(x)); }

/* sin/cos for |a| < 2^10: Cody-Waite reduction by pi/2 (head + tail of pi/2),
 * then the Taylor polynomials of sin (degree 9) and cos (degree 8) on
 * |r| <= pi/4 (truncation < 4e-9) and the quadrant identities.  Used by the VPT
 * ring probes (jfa.cu:345-358), whose offsets must match between CPU and GPU. */
SAD_HD void sad_sincosf(float a, float* s_out, float* c_out)
{
    const float TWO_OVER_PI = 0x1.45f306p-1f;
    const float PIO2_HI = 0x1.921fb6p+0f;
    const float PIO2_LO = -0x1.777a5cp-25f;
    const float SHIFTER = 12582912.0f;
    float t = a * TWO_OVER_PI;
    float qf = (t + SHIFTER) - SHIFTER;
    float r = fmaf(-qf, PIO2_HI, a);
    r = fmaf(-qf, PIO2_LO, r);
    float r2 = r * r;
    float sp = -0x1.a01a02p-13f;                  /* -1/5040 */
    sp = fmaf(sp, r2, 0x1.111112p-7f);            /* 1/120 */
    sp = fmaf(sp, r2, -0x1.555556p-3f);           /* -1/6 */
    sp = fmaf(sp * r2, r, r);
    float sp9 = 0x1.71de3ap-19f * (r2 * r2) * (r2 * r2) * r;   /* r^9/9! */
    float sn = sp + sp9;
    float cp = 0x1.a01a02p-16f;                   /* 1/40320 */
    cp = fmaf(cp, r2, -0x1.6c16c2p-10f);          /* -1/720 */
    cp = fmaf(cp, r2, 0x1.555556p-5f);            /* 1/24 */
    cp = fmaf(cp, r2, -0.5f);
    float cs = fmaf(cp, r2, 1.0f);
    int q = ((int)qf) & 3;
    float s, c;
    if (q == 0) { s = sn; c = cs; }
    else if (q == 1) { s = cs; c = -sn; }
    else if (q == 2) { s = -sn; c = -cs; }
    else { s = -cs; c = sn; }
    *s_out = s;
    *c_out = c;
}
SAD_HD float sad_sinf(float a) { float s, c; sad_sincosf(a, &s, &c); return s; }
SAD_HD float sad_cosf(float a) { float s, c; sad_sincosf(a, &s, &c); return c; }

/* ------------------------------------------------------------- fp16 --
 * IEEE binary32 -> binary16, round to nearest even (IEEE 754-2008 5.4.2),
 * including binary16 subnormals; NaN payloads are kept quiet. */
SAD_HD uint16_t sad_f32_to_f16(float f)
{
    uint32_t x = sad_f2u(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u) {
        if (ax > 0x7F800000u) return (uint16_t)(sign | 0x7E00u | ((ax >> 13) & 0x3FFu));
        return (uint16_t)(sign | 0x7C00u);
    }
    if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);      /* >= 65520 -> inf */
    if (ax < 0x38800000u) {                                         /* < 2^-14 */
        if (ax <= 0x33000000u) return (uint16_t)sign;               /* <= 2^-25 -> 0 */
        uint32_t e = ax >> 23;
        uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
        uint32_t shift = 126u - e;
        uint32_t hm = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (hm & 1u))) hm += 1u;
        return (uint16_t)(sign | hm);
    }
    uint32_t h = (ax - 0x38000000u) >> 13;
    uint32_t rem = ax & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
    return (uint16_t)(sign | h);
}

SAD_HD float sad_f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1Fu;
    uint32_t m = (uint32_t)h & 0x3FFu;
    if (e == 0u) {
        float v = (float)m * 5.9604644775390625e-08f;   /* m * 2^-24, exact */
        return sign ? -v : v;
    }
    if (e == 31u) return sad_u2f(sign | 0x7F800000u | (m << 13));
    return sad_u2f(sign | ((e + 112u) << 23) | (m << 13));
}

/* ------------------------------------------------------------ hashes -- */

/* xorshift32 with the (13,17,5) triple, sad_common.cuh:128-133 */
SAD_HD uint32_t sad_xorshift32(uint32_t x)
{
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

/* Knuth multiplicative hash used for the tile hash, sad_common.cuh:168-170 */
SAD_HD uint32_t sad_tile_hash(uint32_t key) { return key * 2654435761u; }

/* Hilbert curve index of (x, y) on a 2^bits grid, sad_cuda.cu:1523-1545 */
SAD_HD uint32_t sad_hilbert_index(uint32_t x, uint32_t y, int bits)
{
    uint32_t xi = x, yi = y, index = 0;
    uint32_t mask = (bits >= 32) ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    for (int i = bits - 1; i >= 0; --i) {
        uint32_t shift = (uint32_t)i;
        uint32_t rx = (xi >> shift) & 1u;
        uint32_t ry = (yi >> shift) & 1u;
        index |= ((3u * rx) ^ ry) << (2u * shift);
        if (ry == 0u) {
            if (rx == 1u) { xi = mask - xi; yi = mask - yi; }
            uint32_t t = xi; xi = yi; yi = t;
        }
    }
    return index;
}

/* candidate cell of pixel (x, y), sad_common.cuh:113-125 */
SAD_HD int sad_cand_index(int x, int y, int outW, int outH, int candW, int candH)
{
    int sw = outW > 0 ? outW : 1;
    int sh = outH > 0 ? outH : 1;
    int cx = sad_clampi((x * candW) / sw, 0, candW - 1);
    int cy = sad_clampi((y * candH) / sh, 0, candH - 1);
    return cy * candW + cx;
}

/* ------------------------------------------------------ soft distance --
 * voronoi_dmix2 (sad_common.cuh:141-165): anisotropic det-1 metric
 * G = e^a u u^T + e^-a (I - u u^T); returns sqrt(max(d^T G d / S^2, 1e-8)) - r/S.
 * The per-site terms l1 = e^a, l2 = 1/l1 and r/S are passed in precomputed so a
 * caller can hoist them out of a per-pixel loop; sad_dmix2 computes them the
 * same way, so both forms give identical bits. */
SAD_HD float sad_dmix2_pre(float px, float py, float dirx, float diry,
                           float l1, float l2, float r_norm,
                           float ux, float uy, float inv_scale_sq)
{
    float dx = ux - px;
    float dy = uy - py;
    float diff2 = fmaf(dx, dx, dy * dy);
    float proj = fmaf(dirx, dx, diry * dy);
    float proj2 = proj * proj;
    float perp2 = fmaxf(diff2 - proj2, 0.0f);
    float d2a = fmaf(l1, proj2, l2 * perp2);
    float d2s = fmaxf(d2a * inv_scale_sq, 1e-8f);
    return sad_sqrt_pos(d2s) - r_norm;
}

SAD_HD float sad_dmix2(float px, float py, float dirx, float diry, float log_aniso,
                       float radius, float ux, float uy, float inv_scale_sq)
{
    float l1 = sad_expf(log_aniso);
    float l2 = 1.0f / l1;
    float r_norm = radius * sqrtf(inv_scale_sq);
    return sad_dmix2_pre(px, py, dirx, diry, l1, l2, r_norm, ux, uy, inv_scale_sq);
}

/* Per-site terms of the candidate-search key, hoisted out of the per-cell loop:
 * the key tau * dmix (jfa.cu:33-35,171-173) is tau * sad_dmix2_pre(..., l1, l2, rn)
 * with tau = max(e^log_tau, 1e-4), l1 = e^log_aniso, l2 = 1/l1,
 * rn = radius * inv_scale and inv_scale = sqrtf(inv_scale_sq).  pos_x < 0 marks
 * an inactive site (key +inf). */
typedef struct SadSearchSite {
    float px, py, dx, dy;
    float l1, l2, rn, tau;
} SadSearchSite;

SAD_HD SadSearchSite sad_search_site(float px, float py, float log_tau, float radius,
                                     float dx, float dy, float log_aniso, float inv_scale)
{
    SadSearchSite s;
    s.px = px; s.py = py; s.dx = dx; s.dy = dy;
    s.l1 = sad_expf(log_aniso);
    s.l2 = 1.0f / s.l1;
    s.rn = radius * inv_scale;
    s.tau = fmaxf(sad_expf(log_tau), 1e-4f);
    return s;
}

SAD_HD float sad_search_key(SadSearchSite s, float ux, float uy, float inv_scale_sq)
{
    if (s.px < 0.0f) return INFINITY;
    return s.tau * sad_dmix2_pre(s.px, s.py, s.dx, s.dy, s.l1, s.l2, s.rn, ux, uy, inv_scale_sq);
}

/* Candidate-search key of a full-precision site (JFA flood, jfa.cu:33-35).
 * Callers have already skipped inactive sites. */
SAD_HD float sad_site_key(const SadSite* s, float ux, float uy, float inv_scale_sq)
{
    float d = sad_dmix2(s->pos_x, s->pos_y, s->dir_x, s->dir_y, s->log_aniso, s->radius,
                        ux, uy, inv_scale_sq);
    float tau = fmaxf(sad_expf(s->log_tau), 1e-4f);
    return tau * d;
}

/* Same key from the fp16-packed site; inactive (pos.x < 0) -> +inf (jfa.cu:150-174). */
SAD_HD float sad_packed_key(SadPackedSite p, float ux, float uy, float inv_scale_sq)
{
    float px = sad_f16_to_f32(p.h[0]);
    if (px < 0.0f) return INFINITY;
    float py = sad_f16_to_f32(p.h[1]);
    float lt = sad_f16_to_f32(p.h[2]);
    float rad = sad_f16_to_f32(p.h[3]);
    float dx = sad_f16_to_f32(p.h[4]);
    float dy = sad_f16_to_f32(p.h[5]);
    float la = sad_f16_to_f32(p.h[6]);
    float d = sad_dmix2(px, py, dx, dy, la, rad, ux, uy, inv_scale_sq);
    float tau = fmaxf(sad_expf(lt), 1e-4f);
    return tau * d;
}

/* ---------------------------------------------------- gradient terms --
 * Derivative of the loss through one (pixel, candidate) pair
 * (backends/cuda/src/gradients.cu:352-420, sad_bindings.cu:145-193), given the
 * site terms tau = e^log_tau, rn = r/S, dir, l1 = e^a, l2 = 1/max(l1, 1e-8),
 * the pixel (ux, uy), the candidate weight w, dL/dpred g and the blended colour
 * pred.  c[] receives the ten contributions in Site field order
 * {pos.x, pos.y, log_tau, radius, r, g, b, dir.x, dir.y, log_aniso}:
 *   dlogit = w g.(col - pred)
 *   dpos   = dlogit tau S^-2 d2^-1/2 (l2 diff + (l1 - l2) dir proj)
 *   dlog_tau = -dlogit dmix tau,  dradius = dlogit tau / S
 *   dlog_aniso = -dlogit tau d2^-1/2 (l1 proj^2 - l2 perp^2) S^-2 / 2
 *   ddir   = -dlogit tau d2^-1/2 (l1 - l2) proj S^-2 diff
 *   dcol   = w g.
 * The CPU oracle and the GPU kernels evaluate this one function, so per-term
 * values agree bit for bit and parity is limited to summation order; the
 * function itself is checked against finite differences of the render. */
/* 1/sqrt(x) for normal x > 0 from +, * and fmaf only, so every IEEE binary32
 * target computes the same bits: the integer seed 0x5F375A86 - (bits(x) >> 1)
 * (Lomont, "Fast inverse square root", 2003; relative error < 3.5e-2) and three
 * Newton-Raphson steps y <- y (3/2 - (x/2) y^2), each of which squares the
 * relative error (3.5e-2 -> 1.8e-3 -> 4.7e-6 -> 3e-11), so the result is within
 * a few ulp of 1/sqrt(x), the rounding of the last step dominating. */
SAD_HD float sad_rsqrt_nr(float x)
{
    float y = sad_u2f(0x5F375A86u - (sad_f2u(x) >> 1));
    const float hx = 0.5f * x;
    y = y * fmaf(-hx, y * y, 1.5f);
    y = y * fmaf(-hx, y * y, 1.5f);
    y = y * fmaf(-hx, y * y, 1.5f);
    return y;
}

/* sqrt(d2) and 1/sqrt(d2) of the derivative terms (d2 >= 1e-8, so normal):
 * the Newton reciprocal square root above, the same bits on the GPU and the
 * CPU, so per-term values are reproducible wherever they are computed.  The
 * forward pass, whose weights must be bit-exact with the correctly rounded
 * sqrtf, does not go through this macro. */
#ifndef SAD_GRAD_SQRT_RSQ
#define SAD_GRAD_SQRT_RSQ(x, sq, isd) do { (isd) = sad_rsqrt_nr(x); (sq) = (x) * (isd); } while (0)
#endif

/* The site's share of sad_grad_terms: its terms, and the products of them that
 * every pair of the site uses (so a caller that runs many pairs of one site
 * forms them once). */
typedef struct SadGradSite {
    float px, py, tau, rn, dirx, diry, l2, cr, cg, cb, inv_scale;
    float a1, a2;     /* (l1 - l2) dir.x, (l1 - l2) dir.y */
    float l1s, l2s;   /* l1 / S^2, l2 / S^2 */
    float ti;         /* tau / S^2 */
    float b;          /* 2 (l1 - l2) / S^2 */
} SadGradSite;

SAD_HD SadGradSite sad_grad_site(float px, float py, float tau, float rn, float dirx, float diry, float l1, float l2,
                                 float cr, float cg, float cb, float iss, float inv_scale)
{
    SadGradSite s;
    s.px = px; s.py = py; s.tau = tau; s.rn = rn; s.dirx = dirx; s.diry = diry; s.l2 = l2;
    s.cr = cr; s.cg = cg; s.cb = cb; s.inv_scale = inv_scale;
    const float dl12 = l1 - l2;
    s.a1 = dl12 * dirx;
    s.a2 = dl12 * diry;
    s.l1s = l1 * iss;
    s.l2s = l2 * iss;
    s.ti = tau * iss;
    s.b = (2.0f * dl12) * iss;
    return s;
}

/* The pair's share: the ten contributions of pixel (ux, uy) with weight w,
 * dL/dpred g and blended colour pred, through site s. */
SAD_HD void sad_grad_pair(const SadGradSite* s, float ux, float uy, float w, float gr, float gg, float gb, float pr,
                          float pg, float pb, float c[10])
{
    c[4] = w * gr;
    c[5] = w * gg;
    c[6] = w * gb;
    const float dlogit = c[4] * (s->cr - pr) + c[5] * (s->cg - pg) + c[6] * (s->cb - pb);
    const float dx = ux - s->px, dy = uy - s->py;
    const float diff2 = fmaf(dx, dx, dy * dy);
    const float proj = fmaf(s->dirx, dx, s->diry * dy);
    const float proj2 = proj * proj;
    const float perp2 = fmaxf(diff2 - proj2, 0.0f);
    const float q = s->l2s * perp2;
    const float d2s = fmaxf(fmaf(s->l1s, proj2, q), 1e-8f);
    float sq, isd;
    SAD_GRAD_SQRT_RSQ(d2s, sq, isd);
    const float dmix = sq - s->rn;
    const float cpos = dlogit * (s->ti * isd);
    c[0] = cpos * fmaf(s->a1, proj, s->l2 * dx);
    c[1] = cpos * fmaf(s->a2, proj, s->l2 * dy);
    const float dt = dlogit * s->tau;
    c[2] = -(dt * dmix);
    c[3] = dt * s->inv_scale;
    const float hq = dt * (-0.5f * isd);
    c[9] = hq * fmaf(s->l1s, proj2, -q);
    const float hc = hq * (s->b * proj);
    c[7] = hc * dx;
    c[8] = hc * dy;
}

SAD_HD void sad_grad_terms(float px, float py, float tau, float rn, float dirx, float diry, float l1, float l2,
                           float cr, float cg, float cb, float ux, float uy, float w, float gr, float gg, float gb,
                           float pr, float pg, float pb, float iss, float inv_scale, float c[10])
{
    const SadGradSite s = sad_grad_site(px, py, tau, rn, dirx, diry, l1, l2, cr, cg, cb, iss, inv_scale);
    sad_grad_pair(&s, ux, uy, w, gr, gg, gb, pr, pg, pb, c);
}

/* removalDeltaForSite (gradients.cu:82-100): change of |pred - tgt|^2 when the
 * site's summed weight wt is removed and the rest renormalised; 1e6 when
 * 1 - wt < 1e-4. */
SAD_HD float sad_removal_delta(float pr, float pg, float pb, float tr, float tg, float tb, float cr, float cg,
                               float cb, float wt)
{
    float keep = 1.0f - wt;
    if (keep < 1e-4f) return 1.0e6f;
    float k0 = (pr - wt * cr) / keep - tr;
    float k1 = (pg - wt * cg) / keep - tg;
    float k2 = (pb - wt * cb) / keep - tb;
    float d0 = pr - tr, d1 = pg - tg, d2 = pb - tb;
    return (k0 * k0 + k1 * k1 + k2 * k2) - (d0 * d0 + d1 * d1 + d2 * d2);
}

SAD_HD SadPackedSite sad_pack_site(const SadSite* s)
{
    SadPackedSite p;
    if (s->pos_x < 0.0f) {
        p.h[0] = sad_f32_to_f16(-1.0f);
        p.h[1] = sad_f32_to_f16(-1.0f);
        for (int i = 2; i < 8; ++i) p.h[i] = 0;
    } else {
        p.h[0] = sad_f32_to_f16(s->pos_x);
        p.h[1] = sad_f32_to_f16(s->pos_y);
        p.h[2] = sad_f32_to_f16(s->log_tau);
        p.h[3] = sad_f32_to_f16(s->radius);
        p.h[4] = sad_f32_to_f16(s->dir_x);
        p.h[5] = sad_f32_to_f16(s->dir_y);
        p.h[6] = sad_f32_to_f16(s->log_aniso);
        p.h[7] = 0;
    }
    return p;
}

/* ------------------------------------------------------ jump schedule --
 * VPT jump step for pass index i: pow2 = next power of two >= max(w,h),
 * stage = min(i, log2(pow2) - 1), step = max(pow2 >> (stage+1), 1)
 * (sad_cuda.cu:1157-1183, candidate_field.swift:10-33). */
SAD_HD uint32_t sad_jump_step(uint32_t step_index, int width, int height)
{
    uint32_t max_dim = (uint32_t)(width > height ? width : height);
    uint32_t pow2 = 1u;
    while (pow2 < max_dim) pow2 <<= 1;
    if (pow2 <= 1u) return 1u;
    uint32_t stages = 0u, t = pow2;
    while (t > 1u) { t >>= 1; stages += 1u; }
    uint32_t stage = step_index >= stages ? stages - 1u : step_index;
    uint32_t step = pow2 >> (stage + 1u);
    return step > 1u ? step : 1u;
}

SAD_HD uint32_t sad_pack_jump_step(uint32_t step_index, int width, int height)
{
    uint32_t j = sad_jump_step(step_index, width, height);
    if (j > 0xFFFFu) j = 0xFFFFu;
    return (j << 16) | (step_index & 0xFFFFu);
}

SAD_HD int sad_hilbert_bits(int width, int height)
{
    int max_dim = width > height ? width : height;
    int n = 1, bits = 0;
    while (n < max_dim) { n <<= 1; bits += 1; }
    return bits > 1 ? bits : 1;
}

/* ------------------------------------------------------------- atan2 --
 * atan(t), |t| <= 1: two half-angle reductions atan(t) = 2 atan(t / (1 +
 * sqrt(1 + t^2))) bring |t| below tan(pi/16) = 0.199, where the Taylor series
 * to t^11 truncates below 7e-11.  atan2 by octant symmetry.  Used to quantise
 * the anisotropy direction of inference sites, where CPU and GPU must agree. */
#define SAD_PI_F 3.14159274f           /* pi rounded to binary32 */

SAD_HD float sad_atan_unit(float t)
{
    float u = t / (1.0f + sqrtf(fmaf(t, t, 1.0f)));
    u = u / (1.0f + sqrtf(fmaf(u, u, 1.0f)));
    const float u2 = u * u;
    float p = -1.0f / 11.0f;
    p = fmaf(p, u2, 1.0f / 9.0f);
    p = fmaf(p, u2, -1.0f / 7.0f);
    p = fmaf(p, u2, 1.0f / 5.0f);
    p = fmaf(p, u2, -1.0f / 3.0f);
    p = fmaf(p, u2, 1.0f);
    return 4.0f * (u * p);
}

SAD_HD float sad_atan2f(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    float a;
    if (ax == 0.0f && ay == 0.0f) a = 0.0f;
    else if (ay <= ax) a = sad_atan_unit(ay / ax);
    else a = 0.5f * SAD_PI_F - sad_atan_unit(ax / ay);
    if (signbit(x)) a = SAD_PI_F - a;
    return signbit(y) ? -a : a;
}

/* -------------------------------------------- quantised inference sites --
 * makePackedInferenceSites (backends/metal/sources/utils/packed_inference.swift:86-199)
 * and decodePackedSite (shaders/sad_common.metal:71-120).  Quantisation:
 * q = uint(clamp((v - min) / scale, 0, 1) * levels + 0.5); NaN inputs quantise
 * to 0.  Decoding: v = min + q * (1 / levels) * max(scale, 0). */
#define SAD_INV_U16 (1.0f / 65535.0f)
#define SAD_INV_U15 (1.0f / 32767.0f)
#define SAD_INV_U11 (1.0f / 2047.0f)
#define SAD_INV_U10 (1.0f / 1023.0f)
#define SAD_TWO_PI_F 6.28318548f       /* 2 pi rounded to binary32 */

SAD_HD uint32_t sad_unorm_q(float v, float minv, float scale, float levels)
{
    if (!(scale > 0.0f)) return 0u;
    const float n = fminf(fmaxf((v - minv) / scale, 0.0f), 1.0f);
    return (uint32_t)(n * levels + 0.5f);
}

SAD_HD uint32_t sad_unorm_q_max(float v, float maxv, float levels)
{
    if (!(maxv > 0.0f)) return 0u;
    const float n = fminf(fmaxf(v / maxv, 0.0f), 1.0f);
    return (uint32_t)(n * levels + 0.5f);
}

/* angle of the normalised direction over [-pi, pi] as unorm16 (packAngle) */
SAD_HD uint32_t sad_pack_angle(float dx, float dy)
{
    const float len2 = dx * dx + dy * dy;
    if (len2 < 1.0e-8f) { dx = 1.0f; dy = 0.0f; }
    else { const float il = 1.0f / sqrtf(len2); dx = dx * il; dy = dy * il; }
    const float a = sad_atan2f(dy, dx);
    const float n = fminf(fmaxf((a + SAD_PI_F) / SAD_TWO_PI_F, 0.0f), 1.0f);
    return (uint32_t)(n * 65535.0f + 0.5f);
}

/* one site; maxx, maxy = W-1, H-1 */
SAD_HD SadInferenceSite sad_pack_inference(const SadSite* s, const SadInferenceQuant* q, float maxx, float maxy)
{
    SadInferenceSite o;
    const int active = s->pos_x >= 0.0f && s->pos_y >= 0.0f;
    const float px = fminf(fmaxf(s->pos_x, 0.0f), maxx), py = fminf(fmaxf(s->pos_y, 0.0f), maxy);
    o.w[0] = ((sad_unorm_q_max(px, maxx, 32767.0f) | (sad_unorm_q_max(py, maxy, 32767.0f) << 15)) & 0x3fffffffu) |
             ((active ? 0u : 1u) << 30);
    o.w[1] = sad_unorm_q(s->col_r, q->r_min, q->r_scale, 2047.0f) |
             (sad_unorm_q(s->col_g, q->g_min, q->g_scale, 2047.0f) << 11) |
             (sad_unorm_q(s->col_b, q->b_min, q->b_scale, 1023.0f) << 22);
    const float lt = isfinite(s->log_tau) ? s->log_tau : q->log_tau_min;
    const float rd = isfinite(s->radius) ? s->radius : q->radius_min;
    o.w[2] = sad_unorm_q(lt, q->log_tau_min, q->log_tau_scale, 65535.0f) |
             (sad_unorm_q(rd, q->radius_min, q->radius_scale, 65535.0f) << 16);
    const float la = isfinite(s->log_aniso) ? s->log_aniso : 0.0f;
    o.w[3] = (sad_pack_angle(s->dir_x, s->dir_y) & 0xffffu) | ((uint32_t)sad_f32_to_f16(la) << 16);
    return o;
}

/* decode; an inactive site comes out with pos = (-1, -1), as every kernel here expects */
SAD_HD SadSite sad_unpack_inference(SadInferenceSite p, const SadInferenceQuant* q, float maxx, float maxy)
{
    SadSite s;
    const uint32_t w0 = p.w[0], w1 = p.w[1], w2 = p.w[2], w3 = p.w[3];
    s.pos_x = ((float)(w0 & 0x7fffu) * SAD_INV_U15) * fmaxf(maxx, 0.0f);
    s.pos_y = ((float)((w0 >> 15) & 0x7fffu) * SAD_INV_U15) * fmaxf(maxy, 0.0f);
    if ((w0 >> 30) & 1u) { s.pos_x = -1.0f; s.pos_y = -1.0f; }
    s.col_r = q->r_min + ((float)(w1 & 0x7ffu) * SAD_INV_U11) * fmaxf(q->r_scale, 0.0f);
    s.col_g = q->g_min + ((float)((w1 >> 11) & 0x7ffu) * SAD_INV_U11) * fmaxf(q->g_scale, 0.0f);
    s.col_b = q->b_min + ((float)((w1 >> 22) & 0x3ffu) * SAD_INV_U10) * fmaxf(q->b_scale, 0.0f);
    s.log_tau = q->log_tau_min + ((float)(w2 & 0xffffu) * SAD_INV_U16) * fmaxf(q->log_tau_scale, 0.0f);
    s.radius = q->radius_min + ((float)((w2 >> 16) & 0xffffu) * SAD_INV_U16) * fmaxf(q->radius_scale, 0.0f);
    const float ang = ((float)(w3 & 0xffffu) * SAD_INV_U16) * SAD_TWO_PI_F - SAD_PI_F;
    float sn, cs;
    sad_sincosf(ang, &sn, &cs);
    s.dir_x = cs;
    s.dir_y = sn;
    s.log_aniso = sad_f16_to_f32((uint16_t)(w3 >> 16));
    return s;
}

/* the ranges: min and max of the finite values over active sites (pos.x and
 * pos.y >= 0); a field with none gets (0, 0); scale = max(max - min, 1e-6) */
SAD_HD void sad_quant_finish(const float mn[5], const float mx[5], SadInferenceQuant* q)
{
    float* o = &q->log_tau_min;
    for (int k = 0; k < 5; ++k) {
        float a = mn[k], b = mx[k];
        if (!isfinite(a) || !isfinite(b)) { a = 0.0f; b = 0.0f; }
        o[2 * k] = a;
        o[2 * k + 1] = fmaxf(b - a, 1.0e-6f);
    }
}

/* ---------------------------------------------------------------- SSIM --
 * 3x3 window statistics of the luma (0.299, 0.587, 0.114) of a rendered and a
 * target image, the window clamped at the borders (computeSSIMKernel,
 * utils.cu:29-87; ssim_window_stats_shared, gradients.cu:44-78), with
 * SSIM_C1 = 1e-4, SSIM_C2 = 9e-4 and 1/9 as 0.111111111f (sad_common.cuh:16-18). */
#define SAD_SSIM_C1 0.0001f
#define SAD_SSIM_C2 0.0009f
#define SAD_SSIM_INV_N 0.111111111f

typedef struct SadSsimStats {
    float mu_x, mu_y, ex2, ey2, exy;
} SadSsimStats;

SAD_HD float sad_luma3(const float* p) { return fmaf(0.114f, p[2], fmaf(0.587f, p[1], 0.299f * p[0])); }

SAD_HD SadSsimStats sad_ssim_window(const float* r, const float* t, int w, int h, int x, int y)
{
    SadSsimStats s = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int dy = -1; dy <= 1; ++dy) {
        const int yy = sad_clampi(y + dy, 0, h - 1);
        for (int dx = -1; dx <= 1; ++dx) {
            const size_t i = 3 * ((size_t)yy * (size_t)w + (size_t)sad_clampi(x + dx, 0, w - 1));
            const float xv = sad_luma3(r + i), yv = sad_luma3(t + i);
            s.mu_x += xv;
            s.mu_y += yv;
            s.ex2 = fmaf(xv, xv, s.ex2);
            s.ey2 = fmaf(yv, yv, s.ey2);
            s.exy = fmaf(xv, yv, s.exy);
        }
    }
    s.mu_x *= SAD_SSIM_INV_N;
    s.mu_y *= SAD_SSIM_INV_N;
    s.ex2 *= SAD_SSIM_INV_N;
    s.ey2 *= SAD_SSIM_INV_N;
    s.exy *= SAD_SSIM_INV_N;
    return s;
}

/* SSIM of a window = (A B) / max(C D, 1e-8) with the four factors out */
SAD_HD float sad_ssim_terms(SadSsimStats s, float* A, float* B, float* C, float* D)
{
    const float sx2 = fmaxf(fmaf(-s.mu_x, s.mu_x, s.ex2), 0.0f);
    const float sy2 = fmaxf(fmaf(-s.mu_y, s.mu_y, s.ey2), 0.0f);
    const float sxy = fmaf(-s.mu_x, s.mu_y, s.exy);
    *A = fmaf(2.0f * s.mu_x, s.mu_y, SAD_SSIM_C1);
    *B = fmaf(2.0f, sxy, SAD_SSIM_C2);
    *C = fmaf(s.mu_y, s.mu_y, s.mu_x * s.mu_x) + SAD_SSIM_C1;
    *D = (sx2 + sy2) + SAD_SSIM_C2;
    return (*A * *B) / fmaxf(*C * *D, 1e-8f);
}

/* d SSIM(window of pixel k) / d x_k, the window centre's own rendered luma,
 * other pixels held fixed (gradients.cu:236-262); yk is the target luma there */
SAD_HD float sad_ssim_dcentre(SadSsimStats s, float xk, float yk)
{
    float A, B, C, D;
    (void)sad_ssim_terms(s, &A, &B, &C, &D);
    const float denom = fmaxf(C * D, 1e-8f), num = A * B;
    const float dmu = SAD_SSIM_INV_N;
    const float dsig = (2.0f * SAD_SSIM_INV_N) * (xk - s.mu_x);
    const float dcov = SAD_SSIM_INV_N * (yk - s.mu_y);
    const float dA = 2.0f * s.mu_y * dmu, dB = 2.0f * dcov, dC = 2.0f * s.mu_x * dmu, dD = dsig;
    return (dA * B + A * dB) / denom - num * (dC * D + C * dD) / (denom * denom);
}

/* ------------------------------------------- gradient-weighted init ----
 * Seed s of the R2 low-discrepancy sequence (g = plastic number) lands at
 * uv = frac(0.5 + s * (1/g, 1/g^2)), pixel position uv * (W-1, H-1).  The
 * seed is accepted when the mask there is positive and the mask-weighted
 * Sobel magnitude of the luma (0.299, 0.587, 0.114) exceeds the threshold
 * (initGradientWeightedKernel, densify_prune.cu:5-83).  The multiply-adds are
 * written as the fused operations a contracting compiler emits for them. */
#define SAD_R2_G 1.32471795724474602596f

SAD_HD float sad_init_luma(const float* target, int w, int h, int x, int y)
{
    const float* p = target + 3 * ((size_t)sad_clampi(y, 0, h - 1) * (size_t)w + (size_t)sad_clampi(x, 0, w - 1));
    return fmaf(0.114f, p[2], fmaf(0.587f, p[1], 0.299f * p[0]));
}

/* -> 1 when seed s is accepted; position and gradient magnitude out */
SAD_HD int sad_init_grad_eval(const float* target, const float* mask, int w, int h, uint32_t seed, float thr,
                              float* px, float* py, int* pix, float* grad)
{
    const float a1 = 1.0f / SAD_R2_G;
    const float a2 = 1.0f / (SAD_R2_G * SAD_R2_G);
    const float fs = (float)seed;
    float ux = fmaf(fs, a1, 0.5f), uy = fmaf(fs, a2, 0.5f);
    ux -= floorf(ux);                 /* fmodf(u, 1) for u >= 0, exactly */
    uy -= floorf(uy);
    const float x = ux * (float)(w - 1), y = uy * (float)(h - 1);
    const int ix = (int)x, iy = (int)y;
    *px = x;
    *py = y;
    *pix = iy * w + ix;
    *grad = 0.0f;
    const float m = mask ? mask[(size_t)iy * (size_t)w + (size_t)ix] : 1.0f;
    if (m <= 0.0f) return 0;
    const float tl = sad_init_luma(target, w, h, ix - 1, iy - 1), t = sad_init_luma(target, w, h, ix, iy - 1);
    const float tr = sad_init_luma(target, w, h, ix + 1, iy - 1), l = sad_init_luma(target, w, h, ix - 1, iy);
    const float r = sad_init_luma(target, w, h, ix + 1, iy), bl = sad_init_luma(target, w, h, ix - 1, iy + 1);
    const float b = sad_init_luma(target, w, h, ix, iy + 1), br = sad_init_luma(target, w, h, ix + 1, iy + 1);
    const float gx = (((((-tl + tr) - 2.0f * l) + 2.0f * r) - bl) + br) / 4.0f;
    const float gy = (((((-tl - 2.0f * t) - tr) + bl) + 2.0f * b) + br) / 4.0f;
    const float g = sqrtf(fmaf(gy, gy, gx * gx)) * m;
    *grad = g;
    return g > thr;
}

/* Site made from an accepted seed (or, pix < 0, from none): log_tau grows with
 * the fourth root of the gradient, taken as two correctly rounded square roots. */
SAD_HD SadSite sad_init_grad_site(const float* target, float px, float py, int pix, float grad, float log_tau0,
                                  float radius0)
{
    SadSite s;
    s.pos_x = pix >= 0 ? px : -1.0f;
    s.pos_y = pix >= 0 ? py : -1.0f;
    s.log_tau = log_tau0 + sqrtf(sqrtf(grad)) * 2.0f;
    s.radius = radius0;
    s.col_r = pix >= 0 ? target[3 * (size_t)pix + 0] : 0.5f;
    s.col_g = pix >= 0 ? target[3 * (size_t)pix + 1] : 0.5f;
    s.col_b = pix >= 0 ? target[3 * (size_t)pix + 2] : 0.5f;
    s.dir_x = 1.0f;
    s.dir_y = 0.0f;
    s.log_aniso = 0.0f;
    return s;
}

/* ------------------------------------------------------ debug images --
 * The Metal CLI's cell and tau pictures (--render-cells, --include-debug-mask).
 *
 * hashColor, shaders/sad_common.metal:542-565: a colour per site id from three
 * integer hashes of it (the id, id * 2654435761, id * 1103515245, each mixed
 * by xor-shifts and a multiply by 0x27d4eb2d), the low 24 bits of each divided
 * by 2^24 - 1.  The cell picture is the render with these colours in place of
 * the sites' own. */
SAD_HD void sad_hash_color(uint32_t id, float rgb[3])
{
    uint32_t h = (id ^ 61u) ^ (id >> 16);
    h += h << 3;
    h ^= h >> 4;
    const uint32_t r = h * 0x27d4eb2du;
    h = id * 2654435761u;
    h = (h ^ 61u) ^ (h >> 16);
    h += h << 3;
    const uint32_t g = h * 0x27d4eb2du;
    h = id * 1103515245u;
    h = (h ^ 61u) ^ (h >> 16);
    h ^= h >> 4;
    const uint32_t b = h * 0x27d4eb2du;
    rgb[0] = (float)(r & 0xFFFFFFu) / 16777215.0f;
    rgb[1] = (float)(g & 0xFFFFFFu) / 16777215.0f;
    rgb[2] = (float)(b & 0xFFFFFFu) / 16777215.0f;
}

/* renderCentroidsTauHeatmap, shaders/encoders/render_encoder.metal:136-190:
 * each candidate site is a dot of radius `dot` (alpha 1 - smoothstep(dot - 1,
 * dot + 1, distance)), coloured blue -> white -> red as tau runs min -> mean ->
 * max with the ramp position raised to 0.2, composited front to back in
 * candidate order over a 0.05 grey. */
SAD_HD float sad_dot_alpha(float dx, float dy, float dot)
{
    const float dist = sqrtf(dx * dx + dy * dy);
    const float t = sad_clampf((dist - (dot - 1.0f)) / ((dot + 1.0f) - (dot - 1.0f)), 0.0f, 1.0f);
    return 1.0f - t * t * (3.0f - 2.0f * t);
}

SAD_HD void sad_heat_color(float tau, float tmin, float tmean, float tmax, float rgb[3])
{
    float t;
    if (tau <= tmean) t = (tmean > tmin) ? (tau - tmin) / (tmean - tmin) : 0.5f;
    else t = (tmax > tmean) ? (tau - tmean) / (tmax - tmean) : 0.5f;
    t = sad_clampf(t, 0.0f, 1.0f);
    const float e = t > 0.0f ? sad_powf_pos(t, 0.2f) : 0.0f;
    if (tau <= tmean) {            /* mix((0, 0, 1), (1, 1, 1), e) */
        rgb[0] = e;
        rgb[1] = e;
        rgb[2] = 1.0f;
    } else {                       /* mix((1, 1, 1), (1, 0, 0), e) */
        rgb[0] = 1.0f;
        rgb[1] = 1.0f - e;
        rgb[2] = 1.0f - e;
    }
}

#endif /* SAD_MATH_H */
