// colour_win.h -- the LDS staging tables of the colour-gradient kernels' staged variant
// (built on the host by build_colour_windows, read by cg_collide_kernel and cg_colour_kernel_win).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "handle.h"
#include "lattice.h"

namespace lbm {

// Per block of the LDS-staged kernels: up to CW_RUNS runs of consecutive node ids
// [lo, lo + len) copied to LDS slots [off, off + len), packed from slot 0 in order; unused runs
// have len 0 and start where the last used one ends.  Most blocks need no search for a read:
// all their nodes lie in one run, at slot n + own, and for each column j of the neighbour table
// (direction q = j + 1) all the upstream nodes that are not solid lie in one run, at slot
// id + d[j].  Such a block has every bit of `fast` set; a solid upstream site reads the node's
// own slot.
constexpr int CW_RUNS = 4;
constexpr int CW_BLOCK = 512;   // nodes per block of the staged kernels
constexpr uint32_t CW_ALL = (1u << (Q - 1)) - 1;
struct ColourWin {
    int32_t lo[CW_RUNS];
    int32_t len[CW_RUNS];
    int32_t off[CW_RUNS];
    uint32_t fast;
    int32_t own;
    int32_t d[Q - 1];
};
static_assert(sizeof(ColourWin) == 128, "one 128-byte record per block");

// node id held by LDS slot s: the run of slot s is the last one that starts at or before it.  A
// slot past the runs, which nothing reads, gets run 0's first node.
__host__ __device__ __forceinline__ int32_t cw_slot_node(const ColourWin &w, int32_t s)
{
    int32_t id = w.lo[0];
#pragma unroll
    for (int k = 0; k < CW_RUNS; ++k)
        if ((uint32_t)(s - w.off[k]) < (uint32_t)w.len[k]) id = w.lo[k] + (s - w.off[k]);
    return id;
}

// The staging tables of the geometry's blocks of CW_BLOCK nodes for `slots` LDS states per block
// (grid = ceil(N / CW_BLOCK) records into out).  Returns the fraction of blocks that read every
// state from their runs.  g.nbr must still hold the unpacked neighbour table.
double build_colour_windows(const SparseGeometry &g, int64_t grid, int slots, std::vector<ColourWin> &out);

}  // namespace lbm
