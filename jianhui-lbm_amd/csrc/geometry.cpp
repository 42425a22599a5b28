// geometry.cpp -- fluid-node numbering and pull neighbour lists (host).
//
// Follows the reference's sparse structure: pore voxels are numbered in i-major order, x
// slowest and z fastest (SP:1172-1193), and SupInv holds each node's linear voxel index
// (SP:1261-1277).  Where the reference pushes post-collision populations to the downstream
// node and bounces back in place when that node is solid (SP:1874-1926), the device kernels
// pull: population q of node x comes from x - e_q, or from the opposite population of x
// itself when x - e_q is solid.  Both give the same post-streaming state.
#include <algorithm>
#include <cstring>

#include "connectivity.h"
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

void reference_rank_ends(const SparseGeometry &g, int nranks, std::vector<int64_t> &ends)
{
    // pore count of each x-plane, and the node id where each plane starts
    const int64_t nyz = (int64_t)g.ny * g.nz;
    std::vector<int64_t> loc_por((size_t)g.nx, 0), first((size_t)g.nx + 1, 0);
    for (int64_t n = 0; n < g.N; ++n) ++loc_por[(size_t)(g.supinv[(size_t)n] / nyz)];
    for (int i = 0; i < g.nx; ++i) first[(size_t)i + 1] = first[(size_t)i] + loc_por[(size_t)i];
    // cut r + 1 where the running plane sum first reaches ave * (r + 1): before or after the plane
    // that crosses it, whichever sum is closer (SP:1089-1102); the last rank keeps the rest
    const double ave = (double)g.N / nranks;
    ends.assign((size_t)nranks, g.N);
    for (int r = 0; r < nranks - 1; ++r) {
        int64_t pre = 0, aft = 0;
        int n_i = 0;
        while (aft < ave * (r + 1) && n_i < g.nx) {
            pre = aft;
            aft += loc_por[(size_t)n_i];
            ++n_i;
        }
        const int disp = ((double)(aft - ave * (r + 1)) > (double)(ave * (r + 1) - pre)) ? n_i - 1 : n_i;
        ends[(size_t)r] = first[(size_t)disp];
    }
}

bool pores_wrap_along(const SparseGeometry &g, int axis)
{
    const int64_t P = g.P;
    return lattice_loop_wraps(g.N, axis, [&g, P](int64_t n, int q) -> int64_t {
        const uint32_t v = g.nbr[(size_t)((q - 1) * P + n)];
        return (v & BB_FLAG) ? -1 : (int64_t)v;
    });
}

int pack_neighbours(const SparseGeometry &g, PackedNbr &out, std::string &err)
{
    const int64_t P = g.P, N = g.N, W = P / 64;
    out.res.assign((size_t)(18 * P), RES_BB);
    out.hdr.assign((size_t)(W * HDR_WORDS), 0u);
    std::vector<uint8_t> is_wide((size_t)W, 0);
    // bases and the compact/wide decision per wave
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < W; ++w) {
        const int64_t n0 = w * 64, n1 = std::min(n0 + 64, N);
        uint32_t *h = &out.hdr[(size_t)(w * HDR_WORDS)];
        bool wide = false;
        for (int q = 0; q < 18; ++q) {
            uint32_t lo = 0xFFFFFFFFu, hi = 0;
            for (int64_t n = n0; n < n1; ++n) {
                const uint32_t v = g.nbr[(size_t)(q * P + n)];
                if (v & BB_FLAG) continue;
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
            if (lo == 0xFFFFFFFFu) lo = hi = 0;   // no fluid upstream in this direction
            h[q] = lo;
            if (hi - lo >= (uint32_t)RES_BB) wide = true;
        }
        is_wide[(size_t)w] = wide;
    }
    int64_t nw = 0;
    for (int64_t w = 0; w < W; ++w) {
        uint32_t *h = &out.hdr[(size_t)(w * HDR_WORDS)];
        h[HDR_WIDE] = is_wide[(size_t)w] ? (uint32_t)nw++ : HDR_COMPACT;
        h[HDR_WIDE + 1] = 0;
    }
    out.n_wide = nw;
    out.wide.assign((size_t)(nw * 18 * 64), 0u);
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < W; ++w) {
        const int64_t n0 = w * 64, n1 = std::min(n0 + 64, N);
        const uint32_t *h = &out.hdr[(size_t)(w * HDR_WORDS)];
        for (int q = 0; q < 18; ++q)
            for (int64_t n = n0; n < n1; ++n) {
                const uint32_t v = g.nbr[(size_t)(q * P + n)];
                if (h[HDR_WIDE] != HDR_COMPACT)
                    out.wide[(size_t)(((int64_t)h[HDR_WIDE] * 18 + q) * 64 + (n - n0))] = v;
                else if (!(v & BB_FLAG))
                    out.res[(size_t)(q * P + n)] = (uint16_t)(v - h[q]);
            }
    }
    // decode everything as the device does and compare
    int64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t n = 0; n < N; ++n) {
        const uint32_t *h = &out.hdr[(size_t)((n >> 6) * HDR_WORDS)];
        for (int q = 0; q < 18; ++q) {
            uint32_t v;
            if (h[HDR_WIDE] == HDR_COMPACT) {
                const uint16_t r = out.res[(size_t)(q * P + n)];
                v = r == RES_BB ? ((uint32_t)n | BB_FLAG) : h[q] + r;
            } else {
                v = out.wide[(size_t)(((int64_t)h[HDR_WIDE] * 18 + q) * 64 + (n & 63))];
            }
            bad += v != g.nbr[(size_t)(q * P + n)];
        }
    }
    if (bad) {
        err = "neighbour table packing failed its decode check (" + std::to_string(bad) + " entries)";
        return LBM_ERR_STATE;
    }
    return LBM_OK;
}

hipError_t DeviceNbr::upload(const PackedNbr &p, int64_t P, hipStream_t s)
{
    release();
    n_wide = p.n_wide;
    const size_t b_res = (size_t)18 * P * sizeof(uint16_t), b_hdr = p.hdr.size() * sizeof(uint32_t);
    const size_t b_wide = std::max<size_t>(p.wide.size(), 1) * sizeof(uint32_t);
    bytes = b_res + b_hdr + b_wide;
    hipError_t e;
    if ((e = hipMalloc(&res, b_res)) != hipSuccess) return e;
    if ((e = hipMalloc(&hdr, b_hdr)) != hipSuccess) return e;
    if ((e = hipMalloc(&wide, b_wide)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(res, p.res.data(), b_res, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(hdr, p.hdr.data(), b_hdr, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if (!p.wide.empty() &&
        (e = hipMemcpyAsync(wide, p.wide.data(), p.wide.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
        return e;
    return hipSuccess;
}

void DeviceNbr::release()
{
    if (res) (void)hipFree(res), res = nullptr;
    if (hdr) (void)hipFree(hdr), hdr = nullptr;
    if (wide) (void)hipFree(wide), wide = nullptr;
    n_wide = 0;
    bytes = 0;
}

}  // namespace lbm
