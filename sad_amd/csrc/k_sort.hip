/*
 * k_sort.hip - stable LSD radix sort of (key, value) uint32 pairs by key,
 * 8-bit digits (launchRadixSortUInt2, radix_sort.cu:279-336).
 *
 * Per digit pass, three launches, each spread over the whole device:
 *   1. per-tile digit histogram (tile = 1024 pairs = 4 chunks of 256),
 *      written digit-major: hist[digit * tiles + tile];
 *   2. one wave per digit turns its row of hist into an inclusive scan over
 *      the tiles; the row's last entry is then the digit's total;
 *   3. scatter: each block forms the 256 digit bases (an exclusive scan of the
 *      row totals) and adds its tile's exclusive row prefix, then ranks each
 *      256-pair chunk stably with wave64 ballots (the 8 digit bits are matched
 *      with 8 ballots, rank = popcount of the matching lower lanes), orders the
 *      waves through LDS counts, and carries a per-digit running offset from
 *      chunk to chunk.
 * Only the caller's hist words are used (256 per tile), so the reference's
 * buffer sizes hold (block_sums is not needed).
 * The reference ranks each element with an O(256) loop over the chunk
 * (radix_sort.cu:244-253); the ballot match is O(8) per element.
 * Output order is exactly that of a stable sort by key, so it is bit-exact
 * against any stable CPU sort.
 */
#include "sad_device.h"

namespace sadh {

constexpr uint32_t SORT_TILE = 1024u;

__global__ __launch_bounds__(256) void k_sort_hist(const uint2* __restrict__ in, uint32_t* __restrict__ hist,
                                                   uint32_t n, uint32_t tiles, uint32_t shift)
{
    __shared__ uint32_t h[256];
    uint32_t t = threadIdx.x;
    h[t] = 0u;
    __syncthreads();
    uint32_t base = blockIdx.x * SORT_TILE;
#pragma unroll
    for (uint32_t c = 0; c < 4u; ++c) {
        uint32_t i = base + c * 256u + t;
        if (i < n) atomicAdd(&h[(in[i].x >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[t * tiles + blockIdx.x] = h[t];
}

/* Inclusive scan of each digit's row hist[d * tiles .. + tiles) in place, one
 * 64-lane wave per digit: chunks of 256 words, 4 consecutive words per lane,
 * a wave scan of the lane totals, and a carry from chunk to chunk. */
__global__ __launch_bounds__(64) void k_sort_rowscan(uint32_t* __restrict__ hist, uint32_t tiles)
{
    uint32_t* row = hist + (size_t)blockIdx.x * tiles;
    const uint32_t lane = threadIdx.x;
    uint32_t carry = 0u;
    for (uint32_t base = 0; base < tiles; base += 256u) {
        const uint32_t i0 = base + 4u * lane;
        uint32_t x[4];
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k) x[k] = (i0 + k < tiles) ? row[i0 + k] : 0u;
        x[1] += x[0];
        x[2] += x[1];
        x[3] += x[2];
        uint32_t v = x[3];
#pragma unroll
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)v, off);
            if (lane >= off) v += o;
        }
        const uint32_t before = carry + v - x[3];
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k)
            if (i0 + k < tiles) row[i0 + k] = before + x[k];
        carry += (uint32_t)__shfl((int)v, 63);
    }
}

__global__ __launch_bounds__(256) void k_sort_scatter(const uint2* __restrict__ in, uint2* __restrict__ out,
                                                      const uint32_t* __restrict__ offs, uint32_t n, uint32_t tiles,
                                                      uint32_t shift)
{
    __shared__ uint32_t running[256];
    __shared__ uint32_t wc[4][256];
    __shared__ uint32_t wtot[4];
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    {
        /* offs holds each digit's inclusive scan over the tiles: its last entry is the digit's total, and the
         * entry before this tile is the pairs of this digit in earlier tiles */
        const uint32_t* row = offs + (size_t)t * tiles;
        const uint32_t tot = row[tiles - 1u];
        const uint32_t prev = blockIdx.x ? row[blockIdx.x - 1u] : 0u;
        uint32_t v = tot;
#pragma unroll
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)v, off);
            if (lane >= off) v += o;
        }
        if (lane == 63u) wtot[w] = v;
        __syncthreads();
        uint32_t base = v - tot;
        for (uint32_t j = 0; j < w; ++j) base += wtot[j];
        running[t] = base + prev;
    }
    const unsigned long long lt = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    const uint32_t base = blockIdx.x * SORT_TILE;
    for (uint32_t c = 0; c < 4u; ++c) {
        wc[0][t] = 0u; wc[1][t] = 0u; wc[2][t] = 0u; wc[3][t] = 0u;
        uint32_t i = base + c * 256u + t;
        bool valid = i < n;
        uint2 v = valid ? in[i] : make_uint2(0u, 0u);
        uint32_t d = (v.x >> shift) & 0xFFu;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (uint32_t bit = 0; bit < 8u; ++bit) {
            bool on = (d >> bit) & 1u;
            unsigned long long bal = __ballot(on);
            peers &= on ? bal : ~bal;
        }
        __syncthreads();
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        bool leader = valid && (peers >> lane) == 1ull;   /* highest lane of its digit group */
        if (leader) wc[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t o = running[d] + rank;
            for (uint32_t j = 0; j < w; ++j) o += wc[j][d];
            out[o] = v;
        }
        __syncthreads();
        running[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
        __syncthreads();
    }
}

int launch_radix_sort(SadPair* data, SadPair* scratch, uint32_t* hist, uint32_t n, uint32_t max_key_exclusive,
                      hipStream_t stream)
{
    if (n == 0) return SADH_OK;
    uint32_t tiles = div_up(n, SORT_TILE);
    int passes = (max_key_exclusive <= (1u << 16)) ? 2 : 4;
    uint2* a = reinterpret_cast<uint2*>(data);
    uint2* b = reinterpret_cast<uint2*>(scratch);
    for (int p = 0; p < passes; ++p) {
        uint32_t shift = 8u * (uint32_t)p;
        hipLaunchKernelGGL(k_sort_hist, dim3(tiles), dim3(256), 0, stream, a, hist, n, tiles, shift);
        hipLaunchKernelGGL(k_sort_rowscan, dim3(256), dim3(64), 0, stream, hist, tiles);
        hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(256), 0, stream, a, b, hist, n, tiles, shift);
        int rc = launch_status("radix_sort_pairs");
        if (rc) return rc;
        uint2* tmp = a; a = b; b = tmp;
    }
    return SADH_OK;
}

}  // namespace sadh

using namespace sadh;

extern "C" int sadh_radix_sort_pairs(SadPair* data, SadPair* scratch, uint32_t* hist, uint32_t* block_sums,
                                     uint32_t padded_count, uint32_t max_key_exclusive, sadh_stream_t stream)
{
    (void)block_sums;
    SADH_REQUIRE(data && scratch && hist, "radix_sort_pairs: null pointer");
    SADH_REQUIRE(data != scratch, "radix_sort_pairs: scratch must not alias data");
    return launch_radix_sort(data, scratch, hist, padded_count, max_key_exclusive, to_hip(stream));
}
