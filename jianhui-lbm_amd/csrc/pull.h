// pull.h -- the pull-stream gather shared by the single-phase and colour-gradient kernels.
//
// Population q of node n after streaming is the post-collision population q of the upstream
// node n - e_q, or -- when that site is solid -- the post-collision population opp(q) of n
// itself (halfway bounce-back).  The neighbour table gives, per moving direction q = 1..18
// and node, the upstream node index with BB_FLAG set for the bounce-back case (decoded from
// its packed device form by load_nbr), so each load is one select and one gather; the rest
// population needs no index.
#pragma once
#include "collide.h"
#include "lattice.h"
#include "nbrpack.h"
#include "storage.h"

namespace lbm {

enum KMode { K_INIT = 0, K_DIRECT = 1, K_PULL = 2 };

// Gather the 19 pre-collision deviations of node n.  In K_DIRECT mode the buffer holds a
// post-streaming state (a loaded checkpoint) and is read in place.  The 18 neighbour words
// are returned in nb (K_PULL only) for callers that also need the neighbour sites.
template <typename ST, int MODE>
__device__ __forceinline__ void gather_pops(const ST *__restrict__ src, const NbrDev &nbr, int64_t P, int64_t n,
                                            float inv_scale, typename Store<ST>::C (&h)[Q], uint32_t (&nb)[Q - 1])
{
    h[0] = Store<ST>::load(src[n], inv_scale);
    if (MODE == K_DIRECT) {
#pragma unroll
        for (int q = 1; q < Q; ++q) h[q] = Store<ST>::load(src[(int64_t)q * P + n], inv_scale);
        return;
    }
    load_nbr(nbr, P, n, nb);
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        const uint32_t v = nb[q - 1];
        const int64_t row = (v & BB_FLAG) ? (int64_t)opp(q) : (int64_t)q;
        h[q] = Store<ST>::load(src[row * P + (int64_t)(v & ~BB_FLAG)], inv_scale);
    }
}

}  // namespace lbm
