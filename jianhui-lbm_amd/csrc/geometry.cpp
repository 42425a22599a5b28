// geometry.cpp -- fluid-node numbering and pull neighbour lists (host).
//
// Follows the reference's sparse structure: pore voxels are numbered in i-major order, x
// slowest and z fastest (SP:1172-1193), and SupInv holds each node's linear voxel index
// (SP:1261-1277).  Where the reference pushes post-collision populations to the downstream
// node and bounces back in place when that node is solid (SP:1874-1926), the device kernels
// pull: population q of node x comes from x - e_q, or from the opposite population of x
// itself when x - e_q is solid.  Both give the same post-streaming state.
#include <cstring>

#include "handle.h"
#include "lattice.h"

namespace lbm {

int build_sparse_geometry(int nx, int ny, int nz, const int32_t *solid, SparseGeometry &g, std::string &err)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !solid) {
        err = "invalid box size or null geometry";
        return LBM_ERR_ARG;
    }
    const int64_t nvox = (int64_t)nx * ny * nz;
    if (nvox >= ((int64_t)1 << 31)) {
        err = "box has 2^31 or more voxels";
        return LBM_ERR_ARG;
    }
    std::vector<int32_t> idx((size_t)nvox, -1);
    int64_t N = 0;
    for (int64_t v = 0; v < nvox; ++v)
        if (solid[v] == 0) idx[(size_t)v] = (int32_t)N++;
    if (N == 0) {
        err = "geometry has no pore voxels";
        return LBM_ERR_ARG;
    }
    // node indices fit in 31 bits (bit 31 is the bounce-back flag) because nvox < 2^31;
    // element offsets q*P + node are formed in 64-bit arithmetic on the device
    const int64_t P = padded_stride(N);
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.N = N;
    g.P = P;
    g.supinv.assign((size_t)N, 0);
    g.nbr.assign((size_t)(18 * P), 0u);
    for (int64_t v = 0; v < nvox; ++v)
        if (idx[(size_t)v] >= 0) g.supinv[(size_t)idx[(size_t)v]] = (int32_t)v;

    const int64_t nyz = (int64_t)ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
        const int64_t s = g.supinv[(size_t)n];
        const int i = (int)(s / nyz), j = (int)((s % nyz) / nz), k = (int)(s % nz);
        for (int q = 1; q < Q; ++q) {
            int iu = i - EX[q], ju = j - EY[q], ku = k - EZ[q];
            iu = iu < 0 ? iu + nx : (iu >= nx ? iu - nx : iu);
            ju = ju < 0 ? ju + ny : (ju >= ny ? ju - ny : ju);
            ku = ku < 0 ? ku + nz : (ku >= nz ? ku - nz : ku);
            const int32_t up = idx[(size_t)(((int64_t)iu * ny + ju) * nz + ku)];
            g.nbr[(size_t)((q - 1) * P + n)] = up >= 0 ? (uint32_t)up : ((uint32_t)n | BB_FLAG);
        }
    }
    return LBM_OK;
}

}  // namespace lbm
