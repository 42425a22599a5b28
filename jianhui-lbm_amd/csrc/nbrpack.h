// nbrpack.h -- the device form of the pull neighbour lists: 16-bit residuals against per-wave
// bases.
//
// The host builds the full table (18 x P uint32 upstream indices, BB_FLAG for bounce-back).
// On the device that table would be 72 of the 224 bytes an FP32 node update moves.  Nodes of
// one wavefront (64 consecutive node ids) have upstream nodes that are close to each other in
// the i-major numbering, so each direction of each wave is stored as
//
//   hdr[w][q-1]   uint32  base = smallest upstream index of the wave's fluid lanes
//   res[q-1][n]   uint16  upstream - base, or RES_BB when x - e_q is solid
//
// and only a wave whose spread reaches 0xFFFF in some direction (one that straddles the
// periodic wrap of a large plane) keeps its 18 x 64 full words in a side table, named by
// hdr[w][18].  A node update then reads 36 bytes of residuals and 1.25 bytes of header
// instead of 72 bytes.  The header is the same for all lanes of a wave, so it is read with
// scalar loads.
#pragma once
#include <cstdint>

#include "handle.h"
#include "lattice.h"

namespace lbm {

// The first node of n's 64-node group: launches over a node range that does not start on a
// group boundary start their threads here (and idle the lanes before the range).
__host__ __device__ __forceinline__ int64_t group_floor(int64_t n) { return n & ~int64_t(63); }

// The 18 upstream words of node n, in the full-table encoding (index, or n | BB_FLAG).  Every
// active lane of the calling wave must lie in the same group of 64 node ids, which holds for
// launches whose workgroups are whole wavefronts over consecutive, 64-aligned node ranges.
__device__ __forceinline__ void load_nbr(const NbrDev &d, int64_t P, int64_t n, uint32_t (&nb)[Q - 1])
{
    const uint32_t w = __builtin_amdgcn_readfirstlane((uint32_t)(n >> 6));
    const uint32_t *__restrict__ h = d.hdr + (int64_t)w * HDR_WORDS;
    const uint32_t slot = h[HDR_WIDE];
    if (slot == HDR_COMPACT) {
        // bases and residuals are loaded unconditionally and combined with a select: a load
        // inside the conditional would be issued (and waited for) one direction at a time
        uint32_t base[Q - 1], r[Q - 1];
#pragma unroll
        for (int q = 1; q < Q; ++q) base[q - 1] = h[q - 1];
#pragma unroll
        for (int q = 1; q < Q; ++q) r[q - 1] = d.res[(int64_t)(q - 1) * P + n];
        const uint32_t self_bb = (uint32_t)n | BB_FLAG;
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            const uint32_t up = base[q - 1] + r[q - 1];
            nb[q - 1] = r[q - 1] == RES_BB ? self_bb : up;
        }
    } else {
        const uint32_t *__restrict__ wp = d.wide + (int64_t)slot * (18 * 64) + (n & 63);
#pragma unroll
        for (int q = 1; q < Q; ++q) nb[q - 1] = wp[(q - 1) * 64];
    }
}

// The first K directions only (K = 6: the axis directions q = 1..6, all a D3Q7 pull needs).
template <int K>
__device__ __forceinline__ void load_nbr_first(const NbrDev &d, int64_t P, int64_t n, uint32_t (&nb)[K])
{
    const uint32_t w = __builtin_amdgcn_readfirstlane((uint32_t)(n >> 6));
    const uint32_t *__restrict__ h = d.hdr + (int64_t)w * HDR_WORDS;
    const uint32_t slot = h[HDR_WIDE];
    if (slot == HDR_COMPACT) {
        uint32_t base[K], r[K];
#pragma unroll
        for (int q = 0; q < K; ++q) base[q] = h[q];
#pragma unroll
        for (int q = 0; q < K; ++q) r[q] = d.res[(int64_t)q * P + n];
        const uint32_t self_bb = (uint32_t)n | BB_FLAG;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const uint32_t up = base[q] + r[q];
            nb[q] = r[q] == RES_BB ? self_bb : up;
        }
    } else {
        const uint32_t *__restrict__ wp = d.wide + (int64_t)slot * (18 * 64) + (n & 63);
#pragma unroll
        for (int q = 0; q < K; ++q) nb[q] = wp[q * 64];
    }
}

}  // namespace lbm
