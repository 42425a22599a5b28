// common.h -- error handling, launch geometry and deterministic FP64 reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "lbm_mi355x.h"

namespace lbm {

constexpr int BLOCK = 256;          // 4 wavefronts of 64
constexpr int WAVE = 64;
constexpr int NUM_XCD = 8;          // MI355X: 8 accelerator complex dies, one L2 each

// Logical tile of a workgroup.  Workgroups are dispatched round-robin over the 8 XCDs
// (id % 8), each XCD with its own L2.
//
// stream_tile: for the population-streaming kernels (single-phase step, colour-gradient
// collide, server step).  The hardware deals workgroups to the XCDs round robin; this order gives
// each XCD runs of 8 consecutive tiles (2048 nodes), so that a tile's y-neighbour rows, a few
// hundred nodes away, are mostly gathered within the same XCD's L2, while all XCDs still stream
// through the same region of each array at once (DRAM page locality).  On the 256^3
// single-phase step: 206.2 us against 207.3 us for plain round robin and 211.1 us for runs of
// 2; a slab-per-XCD order was 5.7 % slower.  On the 192^3 colour-gradient step: 291.3 against
// 294.0 us.  Tiles past the last whole group of 8 x 8 keep the hardware order.
//
// rr_tile: the hardware order (workgroups dealt to the XCDs round robin).  For the D3Q7 solute
// step and the coupled flow + solute step on the 384^3 fibre pack (porosity 0.8) it measured
// faster than runs of 8: 615 against 637 us and 2037 against 2052 us.
__device__ __forceinline__ int64_t rr_tile(int64_t b) { return b; }

// xcd_slab_tile: each XCD sweeps one contiguous slab of the node list, so that the tiles of
// neighbouring x-planes, which gather each other's data, meet in one L2.  For the colour
// streaming kernel (18 gathers of a 32-byte colour state per node) it cut the L2 misses 3x
// (340 to 116 MiB per launch on a 192^3 pack) and the time by 6-13 %.
__device__ __forceinline__ int64_t stream_tile(int64_t b, int64_t grid)
{
    constexpr int64_t RUN = 8;
    const int64_t full = grid / (NUM_XCD * RUN) * (NUM_XCD * RUN);
    if (b >= full) return b;
    const int64_t x = b % NUM_XCD, i = b / NUM_XCD;   // the i-th workgroup dealt to XCD x
    return ((i / RUN) * NUM_XCD + x) * RUN + (i % RUN);
}

__device__ __forceinline__ int64_t xcd_slab_tile(int64_t b, int64_t grid)
{
    const int64_t x = b % NUM_XCD, slot = b / NUM_XCD;
    const int64_t per = grid / NUM_XCD, rem = grid % NUM_XCD;
    return x * per + (x < rem ? x : rem) + slot;
}

// Tile orders a launch can use (chosen per kernel and geometry by the autotuner, or set through
// lbm_set_launch): the hardware order, runs of 8 tiles per XCD, or one slab per XCD.
enum TileOrder : int32_t { ORDER_HW = 0, ORDER_RUNS8 = 1, ORDER_SLAB = 2 };
__device__ __forceinline__ int64_t tile_of(int32_t order, int64_t b, int64_t grid)
{
    return order == ORDER_RUNS8 ? stream_tile(b, grid) : (order == ORDER_SLAB ? xcd_slab_tile(b, grid) : b);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, WAVE);
        v = o > v ? o : v;
    }
    return v;
}

// Block-level reduction of K FP64 values per thread into out[0..K-1] (one writer thread per
// value).  IS_MAX[k] selects max instead of sum.  The tree is fixed, so the result does not
// depend on scheduling.
template <int K>
__device__ __forceinline__ void block_reduce_store(double (&v)[K], const bool (&is_max)[K], double *out)
{
    __shared__ double red[BLOCK / WAVE][K];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double r = is_max[k] ? wave_max(v[k]) : wave_sum(v[k]);
        if (lane == 0) red[wid][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        double r = red[0][k];
        for (int w = 1; w < BLOCK / WAVE; ++w) r = is_max[k] ? (red[w][k] > r ? red[w][k] : r) : r + red[w][k];
        out[k] = r;
    }
}

// Second pass over the per-block partials: fixed assignment and fixed tree, so the sums are
// reproducible run to run.
template <int K>
__global__ void __launch_bounds__(BLOCK) reduce_partials_kernel(const double *__restrict__ partial, int64_t nblk,
                                                                uint32_t max_mask, double *__restrict__ out)
{
    double v[K];
    bool is_max[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        is_max[k] = (max_mask >> k) & 1u;
        v[k] = is_max[k] ? -INFINITY : 0.0;
    }
    for (int64_t b = threadIdx.x; b < nblk; b += BLOCK) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double x = partial[b * K + k];
            v[k] = is_max[k] ? (x > v[k] ? x : v[k]) : v[k] + x;
        }
    }
    block_reduce_store<K>(v, is_max, out);
}

struct Status {
    int code = LBM_OK;
    std::string msg;
};

#define LBM_HIP_TRY(expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            char buf_[512];                                                                     \
            std::snprintf(buf_, sizeof(buf_), "%s failed: %s (%s:%d)", #expr,                  \
                          hipGetErrorString(e_), __FILE__, __LINE__);                           \
            return fail(LBM_ERR_HIP, buf_);                                                     \
        }                                                                                       \
    } while (0)

}  // namespace lbm
