/*
 * sad_types.h - data layouts shared by the SAD HIP library, its C ABI callers
 * and the CPU oracle.  Plain C99.
 *
 * Site       10 fp32 (40 B): backends/cuda/src/sad_common.cuh:21-31, field order kept.
 * AdamState  20 fp32 (80 B): the m/v groups of sad_common.cuh:44-58 packed without
 *            the float2 alignment hole, i.e. the [N,20] float tensor the torch op
 *            contract allocates (backends/pytorch/train.py:836,918).
 * PackedSite 8 x fp16 (16 B): sad_common.cuh:34-41.
 * Pair       (key, value) uint32 pair sorted by key (the reference's uint2).
 * InferenceSite / InferenceQuant: the 16-byte quantised inference format (Metal backend).
 */
#ifndef SAD_TYPES_H
#define SAD_TYPES_H

#include <stdint.h>

#define SAD_K 8                    /* candidates per cell, sad_common.cuh:9 */
#define SAD_EMPTY 0xFFFFFFFFu      /* empty candidate slot / empty hash key */
#define SAD_TILE_HASH 256u         /* sad_common.cuh:10-12 */
#define SAD_TILE_PROBES 8u

typedef struct SadSite {
    float pos_x, pos_y;
    float log_tau;
    float radius;
    float col_r, col_g, col_b;
    float dir_x, dir_y;
    float log_aniso;
} SadSite;

typedef struct SadAdamState {
    float m_pos_x, m_pos_y, v_pos;
    float m_log_tau, v_log_tau;
    float m_radius, v_radius;
    float m_col_r, m_col_g, m_col_b;
    float v_col_r, v_col_g, v_col_b;
    float m_dir_x, m_dir_y, v_dir;
    float m_log_aniso, v_log_aniso;
    float pad0, pad1;
} SadAdamState;

/* 8 x fp16: {pos.x, pos.y, log_tau, radius, dir.x, dir.y, log_aniso, 0}
 * (sad_common.cuh:34-41, candidate_pack.cu:3-27). */
typedef struct SadPackedSite {
    uint16_t h[8];
} SadPackedSite;

typedef struct SadPair {
    uint32_t key;
    uint32_t val;
} SadPair;

/* 16-byte quantised inference site (backends/metal/sources/utils/packed_inference.swift:3-18,
 * shaders/sad_common.metal:18-40):
 *   w[0] = pos.x unorm15 | pos.y unorm15 << 15 | inactive << 30   (over 0..W-1, 0..H-1)
 *   w[1] = r unorm11 | g unorm11 << 11 | b unorm10 << 22          (over the colour ranges)
 *   w[2] = log_tau unorm16 | radius unorm16 << 16                 (over their ranges)
 *   w[3] = angle of aniso_dir unorm16 over [-pi, pi] | fp16 log_aniso << 16 */
typedef struct SadInferenceSite {
    uint32_t w[4];
} SadInferenceSite;

/* ranges of the quantised fields over the active sites: value = min + q * scale */
typedef struct SadInferenceQuant {
    float log_tau_min, log_tau_scale;
    float radius_min, radius_scale;
    float r_min, r_scale;
    float g_min, g_scale;
    float b_min, b_scale;
} SadInferenceQuant;

#endif /* SAD_TYPES_H */
