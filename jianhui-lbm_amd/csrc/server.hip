// server.hip -- continuous-batching permeability server (many independent single-phase solves).
//
// The reference answers "permeability of many sub-samples" by writing one INPUT file per
// sub-sample and running one MPI solver per file (CPP_SCRIPT/HPC_SUBDOMAIN_INPUT*.cpp).  Here
// one process keeps a device arena and a request loop:
//
//   submit()   (caller's thread)  number the sample's pore voxels and build its pull neighbour
//                                 lists with sample-local indices; queue it.
//   loop thread                   admit queued samples whose node count fits a free arena
//                                 segment (first fit, 256-node aligned); run check_interval
//                                 steps, each ONE launch over the tiles of every active sample;
//                                 then per-sample FP64 flux reductions; retire converged
//                                 samples, free their segments, admit more.
//
// Arena layout (row stride CAP nodes, a sample owns rows [base, base + npad)):
//   pops[2]  19 x CAP  populations (deviation form, FP32 or FP64)
//   nbr      18 x CAP  int16 upstream delta: upstream - node in sample-local indices, taken
//                      modulo the sample's node count (so the periodic wraps are short too);
//                      SRV_BB = bounce back.  36 bytes a node instead of 72 for full indices.
//   nbr_hi   18 x CAP  int16, read only for "wide" samples, whose deltas need 32 bits (more
//                      than 32,767 fluid nodes between a node and an upstream node, e.g. x-planes
//                      of more than 32,767 pores): the delta is (nbr_hi << 16) | (uint16) nbr and
//                      bounce back is INT32_MIN.  A tile belongs to one sample, so the choice is
//                      uniform over a workgroup.
// Device tables: slots[] (per-sample parameters), tiles[] (slot, node range of one workgroup).
// Every sample is cut into tiles from its own first node, and its flux is summed over its own
// tiles in tile order, so its result is bitwise the same whatever else shares the launches.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <omp.h>
#include <thread>
#include <vector>

#include "collide.h"
#include "common.h"
#include "connectivity.h"
#include "handle.h"
#include "pull.h"
#include "storage.h"

namespace lbm {

struct SlotDev {
    double gx, gy, gz, sv, sq, rho0;
    int64_t first_step;   // global step of the sample's first (equilibrium) collision
    int64_t count;        // fluid nodes of the sample (the modulus of its neighbour deltas)
    int32_t wide;         // 1: 32-bit deltas (nbr_hi holds their high halves)
    int32_t pad;
};

constexpr int16_t SRV_BB = -32768;
constexpr int32_t SRV_BB32 = INT32_MIN;

// The delta of a link from its low half (and, for a wide sample, its high half); bounce back
// comes out as SRV_BB (narrow) or SRV_BB32 (wide).
__device__ __forceinline__ int32_t srv_delta(int16_t lo, int16_t hi, bool wide)
{
    return wide ? (int32_t)(((uint32_t)(uint16_t)hi << 16) | (uint32_t)(uint16_t)lo) : (int32_t)lo;
}

// Upstream arena node of direction q for node n of a sample whose nodes start at base.
__device__ __forceinline__ int64_t srv_upstream(int32_t d, int64_t n, int64_t base, int64_t count)
{
    int64_t s = (n - base) + (int64_t)d;
    s += s < 0 ? count : 0;
    s -= s >= count ? count : 0;
    return base + s;
}

struct TileDev {
    int32_t slot;
    int32_t pad;
    int64_t begin, end;   // arena node range handled by one workgroup
    int64_t base;         // the sample's first arena node (local index origin)
};

// Addresses: a tile's rows start at the uniform pointer src + base (its sample's first node), and
// each lane adds a 32-bit BYTE offset, (row * CAP + sample-local node) * sizeof(value), which
// fits because the arena holds fewer than 2^32 / (19 * sizeof(value)) nodes.  A uniform base plus
// a 32-bit unsigned offset is the load's own addressing form: one VGPR per address, no 64-bit
// arithmetic.
template <typename T>
__device__ __forceinline__ T *byte_at(T *base, uint32_t off)
{
    return reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off);
}
template <typename T>
__device__ __forceinline__ const T *byte_at(const T *base, uint32_t off)
{
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

// WIDE: the launch covers the tiles of wide samples only (tiles tile0 .. tile0 + gridDim.x - 1;
// the tile table lists narrow samples' tiles first), so a launch over narrow samples carries no
// code or registers for the high halves.
template <typename ST, int BLK, bool WIDE>
__global__ void __launch_bounds__(BLK) srv_step_kernel(const TileDev *__restrict__ tiles, const SlotDev *__restrict__ slots,
                                                       const ST *__restrict__ src, ST *__restrict__ dst,
                                                       const int16_t *__restrict__ nbr,
                                                       const int16_t *__restrict__ nbr_hi, int64_t CAP, int64_t step,
                                                       int64_t tile0)
{
    using C = typename Store<ST>::C;
    constexpr uint32_t SZ = sizeof(ST);
    const TileDev t = tiles[tile0 + stream_tile(blockIdx.x, gridDim.x)];
    const int64_t n = t.begin + threadIdx.x;
    if (n >= t.end) return;
    const SlotDev sl = slots[t.slot];
    const C gx = (C)sl.gx, gy = (C)sl.gy, gz = (C)sl.gz, rho0 = (C)sl.rho0;
    const uint32_t cap = (uint32_t)CAP, ln = (uint32_t)(n - t.base);
    const ST *__restrict__ S = src + t.base;
    C h[Q], m[Q];
    Macro<C> mc;
    if (step == sl.first_step) {   // f = feq(1, 0): the reference's init (SP:1343-1376)
        mc.rho = C(1);
        mc.drho = C(1) - rho0;
        mc.ux = mc.uy = mc.uz = C(0);
        sp_eq_moments(mc, m);
    } else {
        const int16_t *__restrict__ NB = nbr + t.base;
        const int32_t cnt = (int32_t)sl.count;
        constexpr int32_t bbv = WIDE ? SRV_BB32 : (int32_t)SRV_BB;
        h[0] = Store<ST>::load(*byte_at(S, ln * SZ), 1.0f);
        int32_t d[Q - 1];
#pragma unroll
        for (int q = 1; q < Q; ++q) d[q - 1] = *byte_at(NB, ((uint32_t)(q - 1) * cap + ln) * 2u);
        if (WIDE) {
            const int16_t *__restrict__ NH = nbr_hi + t.base;
#pragma unroll
            for (int q = 1; q < Q; ++q)
                d[q - 1] = srv_delta((int16_t)d[q - 1], *byte_at(NH, ((uint32_t)(q - 1) * cap + ln) * 2u), true);
        }
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            // upstream node, sample-local: ln + d reduced modulo the node count (|d| <= count/2)
            int32_t u = (int32_t)ln + d[q - 1];
            u += u < 0 ? cnt : 0;
            u -= u >= cnt ? cnt : 0;
            const uint32_t e = d[q - 1] == bbv ? (uint32_t)opp(q) * cap + ln : (uint32_t)q * cap + (uint32_t)u;
            h[q] = Store<ST>::load(*byte_at(S, e * SZ), 1.0f);
        }
        moments_fwd(h, m);
        mc = macro_from_moments(m, rho0, gx, gy, gz);
    }
    sp_relax(m, mc, gx, gy, gz, (C)sl.sv, (C)sl.sq, h);
    ST *__restrict__ D = dst + t.base;
#pragma unroll
    for (int q = 0; q < Q; ++q) Store<ST>::store_nt(byte_at(D, ((uint32_t)q * cap + ln) * SZ), h[q], 1.0f);
}

// FP64 arena: 64-bit element indices, which keep this kernel at 70 VGPRs (7 waves per SIMD)
// where the byte-offset form below needs 84 (5 waves)
template <typename ST, int BLK, bool WIDE>
__global__ void __launch_bounds__(BLK) srv_step_kernel_wide(const TileDev *__restrict__ tiles, const SlotDev *__restrict__ slots,
                                                       const ST *__restrict__ src, ST *__restrict__ dst,
                                                       const int16_t *__restrict__ nbr,
                                                       const int16_t *__restrict__ nbr_hi, int64_t CAP, int64_t step,
                                                       int64_t tile0)
{
    using C = typename Store<ST>::C;
    const TileDev t = tiles[tile0 + stream_tile(blockIdx.x, gridDim.x)];
    const int64_t n = t.begin + threadIdx.x;
    if (n >= t.end) return;
    const SlotDev sl = slots[t.slot];
    const C gx = (C)sl.gx, gy = (C)sl.gy, gz = (C)sl.gz, rho0 = (C)sl.rho0;
    C h[Q], m[Q];
    Macro<C> mc;
    if (step == sl.first_step) {   // f = feq(1, 0): the reference's init (SP:1343-1376)
        mc.rho = C(1);
        mc.drho = C(1) - rho0;
        mc.ux = mc.uy = mc.uz = C(0);
        sp_eq_moments(mc, m);
    } else {
        h[0] = Store<ST>::load(src[n], 1.0f);
        constexpr int32_t bbv = WIDE ? SRV_BB32 : (int32_t)SRV_BB;
        int32_t d[Q - 1];
#pragma unroll
        for (int q = 1; q < Q; ++q) d[q - 1] = nbr[(int64_t)(q - 1) * CAP + n];
        if (WIDE) {
#pragma unroll
            for (int q = 1; q < Q; ++q) d[q - 1] = srv_delta((int16_t)d[q - 1], nbr_hi[(int64_t)(q - 1) * CAP + n], true);
        }
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            const bool bb = d[q - 1] == bbv;
            const int64_t up = srv_upstream(d[q - 1], n, t.base, sl.count);
            h[q] = Store<ST>::load(src[(bb ? (int64_t)opp(q) : (int64_t)q) * CAP + (bb ? n : up)], 1.0f);
        }
        moments_fwd(h, m);
        mc = macro_from_moments(m, rho0, gx, gy, gz);
    }
    sp_relax(m, mc, gx, gy, gz, (C)sl.sv, (C)sl.sq, h);
#pragma unroll
    for (int q = 0; q < Q; ++q) Store<ST>::store_nt(&dst[(int64_t)q * CAP + n], h[q], 1.0f);
}

// per-tile flux partials Q[3] and a non-finite flag, from the current (pulled) state
template <typename ST, int BLK>
__global__ void __launch_bounds__(BLK) srv_flux_kernel(const TileDev *__restrict__ tiles, const SlotDev *__restrict__ slots,
                                                       const ST *__restrict__ src, const int16_t *__restrict__ nbr,
                                                       const int16_t *__restrict__ nbr_hi, int64_t CAP,
                                                       double *__restrict__ partial)
{
    using C = typename Store<ST>::C;
    const TileDev t = tiles[blockIdx.x];
    const int64_t n = t.begin + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (n < t.end) {
        const SlotDev sl = slots[t.slot];
        C h[Q], m[Q];
        h[0] = Store<ST>::load(src[n], 1.0f);
        const bool wide = sl.wide != 0;
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            const int64_t i = (int64_t)(q - 1) * CAP + n;
            const int32_t d = srv_delta(nbr[i], wide ? nbr_hi[i] : (int16_t)0, wide);
            const bool bb = d == (wide ? SRV_BB32 : (int32_t)SRV_BB);
            const int64_t up = srv_upstream(d, n, t.base, sl.count);
            h[q] = Store<ST>::load(src[(bb ? (int64_t)opp(q) : (int64_t)q) * CAP + (bb ? n : up)], 1.0f);
        }
        moments_fwd(h, m);
        const Macro<C> mc = macro_from_moments(m, (C)sl.rho0, (C)sl.gx, (C)sl.gy, (C)sl.gz);
        v[0] = (double)mc.ux;
        v[1] = (double)mc.uy;
        v[2] = (double)mc.uz;
        v[3] = (isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2])) ? 0.0 : 1.0;
    }
    // block reduction in a fixed tree (BLK/64 waves)
    __shared__ double red[BLK / WAVE][4];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double r = wave_sum(v[k]);
        if (lane == 0) red[wid][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double r = red[0][threadIdx.x];
        for (int w = 1; w < BLK / WAVE; ++w) r += red[w][threadIdx.x];
        partial[(int64_t)blockIdx.x * 4 + threadIdx.x] = r;
    }
}

// one workgroup per active sample: sum its tiles' NV partials in tile order
template <int NV>
__global__ void __launch_bounds__(BLOCK) srv_slot_sum_kernel(const double *__restrict__ partial,
                                                             const int32_t *__restrict__ tile_lo,
                                                             const int32_t *__restrict__ tile_hi, double *__restrict__ out)
{
    const int a = blockIdx.x;
    const int lo = tile_lo[a], hi = tile_hi[a];
    double v[NV];
    bool is_max[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        v[k] = 0.0;
        is_max[k] = false;
    }
    for (int t = lo + (int)threadIdx.x; t < hi; t += BLOCK)
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += partial[(int64_t)t * NV + k];
    block_reduce_store<NV>(v, is_max, out + (int64_t)a * NV);
}

// ---------------------------------------------------------------------------- Anderson acceleration
// Per sample, the steady state is the fixed point of the K-step map x -> g(x) (K = the
// acceleration period; x is the stored post-collision state of the sample's nodes).  Every K
// steps, with r = g - x:
//   SAVE    (end of the warm-up)      X = x
//   FIRST   (one period later)        R_prev = r, G_prev = g, X = g  (plain step)
//   RESID                             dR_h = r - R_prev, dG_h = g - G_prev into ring slot h,
//                                     dot products dR_i.dR_h, dR_i.r, r.r for the host
//   UPDATE  (after the host's solve)  x = g - sum_i gamma_i dG_i, written to the state and to X
// gamma minimises |r - dR gamma| (Walker & Ni's form of Anderson mixing).  The dot products are
// reduced per tile and then per sample in tile order, so they are the same whatever shares
// the launch.
constexpr int AA_MAX = 8;
constexpr int AA_NV = 2 * AA_MAX + 1;   // dR_i.dR_head, dR_i.r, r.r
enum AaMode : int32_t { AA_NONE = 0, AA_SAVE = 1, AA_FIRST = 2, AA_RESID = 3 };

struct AaDev {
    int32_t mode;
    int32_t head;    // ring slot written by this RESID
    int32_t count;   // valid ring slots after this RESID (0..m-1 are candidates)
    int32_t pad;
    double gamma[AA_MAX];
};

template <int BLK, int K>
__device__ __forceinline__ void tile_reduce(double (&v)[K], double *out)
{
    __shared__ double red[BLK / WAVE][K];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double r = wave_sum(v[k]);
        if (lane == 0) red[wid][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double r = red[0][threadIdx.x];
        for (int w = 1; w < BLK / WAVE; ++w) r += red[w][threadIdx.x];
        out[threadIdx.x] = r;
    }
}

template <typename ST, int BLK>
__global__ void __launch_bounds__(BLK) srv_aa_resid_kernel(const TileDev *__restrict__ tiles,
                                                           const AaDev *__restrict__ aa, const ST *__restrict__ pops,
                                                           ST *__restrict__ X, ST *__restrict__ RP, ST *__restrict__ GP,
                                                           ST *__restrict__ DR, ST *__restrict__ DG, int64_t CAP,
                                                           double *__restrict__ partial)
{
    using C = typename Store<ST>::C;
    const TileDev t = tiles[blockIdx.x];
    const AaDev a = aa[t.slot];
    const int64_t n = t.begin + threadIdx.x;
    const int64_t V = (int64_t)Q * CAP;
    double acc[AA_NV];
#pragma unroll
    for (int k = 0; k < AA_NV; ++k) acc[k] = 0.0;
    if (a.mode != AA_NONE && n < t.end) {
        for (int q = 0; q < Q; ++q) {
            const int64_t i = (int64_t)q * CAP + n;
            const ST gs = pops[i];
            if (a.mode == AA_SAVE) {
                X[i] = gs;
                continue;
            }
            const C g = (C)gs, r = g - (C)X[i];
            if (a.mode == AA_RESID) {
                const C dr = r - (C)RP[i], dg = g - (C)GP[i];
                DR[a.head * V + i] = (ST)dr;
                DG[a.head * V + i] = (ST)dg;
                const double drs = (double)(ST)dr;
#pragma unroll
                for (int h = 0; h < AA_MAX; ++h) {
                    if (h < a.count) {
                        const double v = h == a.head ? drs : (double)DR[h * V + i];
                        acc[h] += v * drs;
                        acc[AA_MAX + h] += v * (double)r;
                    }
                }
                acc[2 * AA_MAX] += (double)r * (double)r;
            } else {
                X[i] = gs;   // FIRST: the next iterate is g itself
            }
            RP[i] = (ST)r;
            GP[i] = gs;
        }
    }
    tile_reduce<BLK, AA_NV>(acc, partial + (int64_t)blockIdx.x * AA_NV);
}

template <typename ST, int BLK>
__global__ void __launch_bounds__(BLK) srv_aa_update_kernel(const TileDev *__restrict__ tiles,
                                                            const AaDev *__restrict__ aa, ST *__restrict__ pops,
                                                            ST *__restrict__ X, const ST *__restrict__ DG, int64_t CAP)
{
    const TileDev t = tiles[blockIdx.x];
    const AaDev a = aa[t.slot];
    const int64_t n = t.begin + threadIdx.x;
    if (a.mode != AA_RESID || n >= t.end) return;
    const int64_t V = (int64_t)Q * CAP;
    for (int q = 0; q < Q; ++q) {
        const int64_t i = (int64_t)q * CAP + n;
        double x = (double)pops[i];
#pragma unroll
        for (int h = 0; h < AA_MAX; ++h)
            if (h < a.count) x -= a.gamma[h] * (double)DG[h * V + i];
        pops[i] = (ST)x;
        X[i] = (ST)x;
    }
}

// ---------------------------------------------------------------------------- sample preparation
// A submitted sample's pore numbering (i-major, x slowest: build_sparse_geometry's) and its pull
// deltas, built on the device from its pore flags when it is admitted: a three-pass scan numbers
// the pores (count per tile of PREP_TILE voxels, scan of the tile counts, numbering inside each
// tile), then one thread per voxel writes each pore's 18 deltas into the sample's arena rows --
// the shortest of up - n, up - n - N, up - n + N, as build_deltas does on the host.
constexpr int PREP_TILE = 2048;   // voxels per scan tile: 256 threads x 8
constexpr int PREP_PER = PREP_TILE / 256;
constexpr int32_t PREP_WIDE = 1;  // a delta does not fit 16 bits
constexpr int32_t PREP_COUNT = 2; // the pore count differs from the host's

// inclusive scan of one int per thread over a 256-thread block; returns the block total
__device__ __forceinline__ int32_t block_scan_256(int32_t &v)
{
    __shared__ int32_t part[256 / WAVE];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const int32_t o = __shfl_up(v, off, WAVE);
        if (lane >= off) v += o;
    }
    if (lane == WAVE - 1) part[wid] = v;
    __syncthreads();
    int32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 256 / WAVE; ++w) {
        before += w < wid ? part[w] : 0;
        total += part[w];
    }
    v += before;
    __syncthreads();
    return total;
}

__global__ void __launch_bounds__(256) srv_prep_count_kernel(const uint8_t *__restrict__ pore, int64_t V,
                                                             int32_t *__restrict__ tile_cnt)
{
    const int64_t v0 = (int64_t)blockIdx.x * PREP_TILE + (int64_t)threadIdx.x * PREP_PER;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < PREP_PER; ++k) c += (v0 + k < V) ? (int32_t)pore[v0 + k] : 0;
    const int32_t total = block_scan_256(c);
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = total;
}

// exclusive scan of the tile counts in place (one block, any number of tiles), and the total
__global__ void __launch_bounds__(256) srv_prep_scan_kernel(int32_t *__restrict__ tile, int64_t ntiles,
                                                            int64_t N_host, int32_t *__restrict__ flag)
{
    int32_t carry = 0;
    for (int64_t t0 = 0; t0 < ntiles; t0 += 256) {
        const int64_t t = t0 + threadIdx.x;
        const int32_t c = t < ntiles ? tile[t] : 0;
        int32_t v = c;
        const int32_t total = block_scan_256(v);
        if (t < ntiles) tile[t] = carry + v - c;
        carry += total;
    }
    if (threadIdx.x == 0 && (int64_t)carry != N_host) atomicOr(flag, PREP_COUNT);
}

__global__ void __launch_bounds__(256) srv_prep_number_kernel(const uint8_t *__restrict__ pore, int64_t V,
                                                              const int32_t *__restrict__ tile_off,
                                                              int32_t *__restrict__ idx)
{
    const int64_t v0 = (int64_t)blockIdx.x * PREP_TILE + (int64_t)threadIdx.x * PREP_PER;
    uint8_t p[PREP_PER];
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < PREP_PER; ++k) {
        p[k] = (v0 + k < V) ? pore[v0 + k] : 0;
        c += p[k];
    }
    int32_t incl = c;
    (void)block_scan_256(incl);
    int32_t n = tile_off[blockIdx.x] + incl - c;
#pragma unroll
    for (int k = 0; k < PREP_PER; ++k)
        if (v0 + k < V) idx[v0 + k] = p[k] ? n++ : -1;
}

// WIDE = false: 16-bit deltas into nbr (SRV_BB for bounce back), PREP_WIDE raised if one does
// not fit; WIDE = true: 32-bit deltas, low halves into nbr and high halves into nbr_hi
template <bool WIDE>
__global__ void __launch_bounds__(256) srv_prep_links_kernel(const int32_t *__restrict__ idx, int32_t nx, int32_t ny,
                                                             int32_t nz, int64_t N, int16_t *__restrict__ nbr,
                                                             int16_t *__restrict__ nbr_hi, int64_t CAP, int64_t base,
                                                             int32_t *__restrict__ flag)
{
    const int64_t nyz = (int64_t)ny * nz, V = (int64_t)nx * nyz;
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    const int32_t n = idx[v];
    if (n < 0) return;
    const int32_t i = (int32_t)(v / nyz), j = (int32_t)((v / nz) % ny), k = (int32_t)(v % nz);
    bool bad = false;
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        int iu = i - EX[q], ju = j - EY[q], ku = k - EZ[q];
        iu = iu < 0 ? iu + nx : (iu >= nx ? iu - nx : iu);
        ju = ju < 0 ? ju + ny : (ju >= ny ? ju - ny : ju);
        ku = ku < 0 ? ku + nz : (ku >= nz ? ku - nz : ku);
        const int32_t up = idx[((int64_t)iu * ny + ju) * nz + ku];
        int32_t d = SRV_BB32;
        if (up >= 0) {
            const int64_t dd = (int64_t)up - n;
            int64_t best = dd;
            if (llabs(dd - N) < llabs(best)) best = dd - N;
            if (llabs(dd + N) < llabs(best)) best = dd + N;
            d = (int32_t)best;
        }
        const int64_t e = (int64_t)(q - 1) * CAP + base + n;
        if (WIDE) {
            nbr[e] = (int16_t)(uint16_t)((uint32_t)d & 0xFFFFu);
            nbr_hi[e] = (int16_t)(uint16_t)((uint32_t)d >> 16);
        } else {
            const bool fits = d == SRV_BB32 || (d >= -32767 && d <= 32767);
            bad |= !fits;
            nbr[e] = d == SRV_BB32 ? SRV_BB : (fits ? (int16_t)d : (int16_t)0);
        }
    }
    if (!WIDE && bad) atomicOr(flag, PREP_WIDE);
}

// ---------------------------------------------------------------------------- host
namespace {
using Clock = std::chrono::steady_clock;

struct Prepared {
    int64_t ticket = 0;
    int64_t prior_steps = 0;     // steps already spent in an FP32 arena before escalation
    int prior_acc = 0;
    std::vector<float> carried;  // 19 x npad FP32 state to continue from (escalation), else start at rest
    int64_t N = 0, npad = 0;
    // a submitted sample: its pore flags ([i][j][k], 1 = pore), numbered and linked on the device
    // at admission; shared with the connectivity test, which may still read them then
    std::shared_ptr<const std::vector<uint8_t>> pore;
    int nx = 0, ny = 0, nz = 0;
    bool wrap_pending = false;   // its connectivity test has not reported yet: it does not retire
    // a sample handed over from another arena (escalation): its 18 x npad upstream deltas
    // (SRV_BB = bounce back), read back from that arena
    std::shared_ptr<int16_t[]> nbr;
    // wide sample: the high halves of its 32-bit deltas, 18 x npad (null for a narrow one)
    std::shared_ptr<int16_t[]> nbr_hi;
    bool wide = false;
    double nu = 0, g[3] = {0, 0, 0};
    int dir = 1;
    int64_t volume = 0;
    double t_submit = 0;
};

struct Active {
    Prepared req;
    int64_t base = 0;
    int64_t first_step = 0;
    int64_t steps = 0;
    double k_prev = 0;
    bool has_prev = false;
    double t_admit = 0;
    // Anderson acceleration (host side of the per-sample state)
    int aa_stage = 0;          // 0 before SAVE, 1 before FIRST, 2 running
    int64_t aa_next = 0;       // sample step count of the next acceleration event
    int aa_head = 0, aa_count = 0, aa_updates = 0;
    double gram[AA_MAX][AA_MAX] = {};
    double r2_min = INFINITY;
    double k_ev = 0;           // k at the previous event
    bool has_k_ev = false;
    int small_events = 0;      // consecutive events whose k changed by <= tol
    bool no_wrap = false;      // pores found not to wrap along dir: to FP64 at the next check
};

// gamma = (G + lambda I)^-1 b for the n x n leading block, Gaussian elimination with partial
// pivoting; false if the system is singular or the result not finite
bool solve_small(const double (&G)[AA_MAX][AA_MAX], const double *b, int n, double *x)
{
    double A[AA_MAX][AA_MAX + 1];
    double dmax = 0.0;
    for (int i = 0; i < n; ++i) dmax = std::max(dmax, std::fabs(G[i][i]));
    const double lam = 1e-12 * dmax;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) A[i][j] = G[i][j] + (i == j ? lam : 0.0);
        A[i][n] = b[i];
    }
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
        if (!(std::fabs(A[p][c]) > 0.0)) return false;
        if (p != c)
            for (int j = 0; j <= n; ++j) std::swap(A[p][j], A[c][j]);
        for (int r = c + 1; r < n; ++r) {
            const double f = A[r][c] / A[c][c];
            for (int j = c; j <= n; ++j) A[r][j] -= f * A[c][j];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double v = A[i][n];
        for (int j = i + 1; j < n; ++j) v -= A[i][j] * x[j];
        x[i] = v / A[i][i];
        if (!std::isfinite(x[i])) return false;
    }
    return true;
}

struct Segment {
    int64_t start, len;
};

// a queued sample's connectivity test (run on a host worker, beside the device work)
struct PrepJob {
    int64_t ticket = 0;
    int nx = 0, ny = 0, nz = 0;
    std::shared_ptr<const std::vector<uint8_t>> pore;   // 1 = pore voxel, [i][j][k]
    int dir = 1;
};

// Whether the pores wrap round the periodic box along axis (connectivity.h), from the pore flags:
// the numbering in one pass, the links computed as the union-find asks for them.
bool pore_flags_wrap(int nx, int ny, int nz, const uint8_t *pore, int axis)
{
    const int64_t nyz = (int64_t)ny * nz, V = (int64_t)nx * nyz;
    std::vector<int32_t> idx((size_t)V), sup;
    int32_t N = 0;
    for (int64_t v = 0; v < V; ++v) {
        idx[(size_t)v] = pore[v] ? N++ : -1;
        if (pore[v]) sup.push_back((int32_t)v);
    }
    return lattice_loop_wraps(N, axis, [&](int64_t n, int q) -> int64_t {
        const int64_t s = sup[(size_t)n];
        int i = (int)(s / nyz), j = (int)((s / nz) % ny), k = (int)(s % nz);
        i -= EX[q];
        j -= EY[q];
        k -= EZ[q];
        i = i < 0 ? i + nx : (i >= nx ? i - nx : i);
        j = j < 0 ? j + ny : (j >= ny ? j - ny : j);
        k = k < 0 ? k + nz : (k >= nz ? k - nz : k);
        return idx[(size_t)(((int64_t)i * ny + j) * nz + k)];
    });
}

// The sample's pore numbering (i-major, x slowest: the same as build_sparse_geometry) and its
// pull links as deltas: for each link the shortest of up - n, up - n - N, up - n + N (the device
// adds it to n and reduces modulo N).  16-bit deltas when every one fits (out_hi left null);
// otherwise a second pass writes 32-bit deltas as low halves in out and high halves in out_hi.
// False if the box has no pores.
bool build_deltas(int nx, int ny, int nz, const uint8_t *pore, int64_t &N, int64_t &P, std::shared_ptr<int16_t[]> &out,
                  std::shared_ptr<int16_t[]> &out_hi)
{
    const int64_t nyz = (int64_t)ny * nz;
    // numbering: pores per x-plane, their prefix, then each plane's pores in order
    std::vector<int64_t> first((size_t)nx + 1, 0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i) {
        int64_t c = 0;
        for (int64_t v = i * nyz; v < (i + 1) * nyz; ++v) c += pore[v];
        first[(size_t)i + 1] = c;
    }
    for (int i = 0; i < nx; ++i) first[(size_t)i + 1] += first[(size_t)i];
    N = first[(size_t)nx];
    P = padded_stride(N);
    if (N == 0) return false;
    std::unique_ptr<int32_t[]> idx(new int32_t[(size_t)(nx * nyz)]);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i) {
        int64_t c = first[(size_t)i];
        for (int64_t v = i * nyz; v < (i + 1) * nyz; ++v) idx[(size_t)v] = pore[v] ? (int32_t)c++ : -1;
    }
    // every entry is written below (a delta, or SRV_BB for a solid site and for the padding), so
    // the rows are allocated uninitialised and first touched by the threads that fill them
    out = std::shared_ptr<int16_t[]>(new int16_t[(size_t)(18 * P)]);
    out_hi.reset();
    int16_t *o = out.get();
    const int64_t NN = N;
    // the shortest delta of each link of node n at (i, j, k), or SRV_BB32 for bounce back
    auto link = [&](int i, int j, int k, int32_t n, int q) -> int32_t {
        int iu = i - EX[q], ju = j - EY[q], ku = k - EZ[q];
        iu = iu < 0 ? iu + nx : (iu >= nx ? iu - nx : iu);
        ju = ju < 0 ? ju + ny : (ju >= ny ? ju - ny : ju);
        ku = ku < 0 ? ku + nz : (ku >= nz ? ku - nz : ku);
        const int32_t up = idx[(size_t)(iu * nyz + (int64_t)ju * nz + ku)];
        if (up < 0) return SRV_BB32;
        const int64_t d = (int64_t)up - n;
        int64_t best = d;
        if (std::llabs(d - NN) < std::llabs(best)) best = d - NN;
        if (std::llabs(d + NN) < std::llabs(best)) best = d + NN;
        return (int32_t)best;   // |best| <= N / 2 < 2^30
    };
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int32_t n = idx[(size_t)(i * nyz + (int64_t)j * nz + k)];
                if (n < 0) continue;
                for (int q = 1; q < Q; ++q) {
                    const int32_t d = link(i, j, k, n, q);
                    int16_t w = SRV_BB;
                    if (d != SRV_BB32) {
                        if (d < -32767 || d > 32767)
                            bad = 1;
                        else
                            w = (int16_t)d;
                    }
                    o[(size_t)((q - 1) * P + n)] = w;
                }
            }
    if (bad) {
        // a wide sample: every link as a 32-bit delta split into halves
        out_hi = std::shared_ptr<int16_t[]>(new int16_t[(size_t)(18 * P)]);
        int16_t *oh = out_hi.get();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    const int32_t n = idx[(size_t)(i * nyz + (int64_t)j * nz + k)];
                    if (n < 0) continue;
                    for (int q = 1; q < Q; ++q) {
                        const uint32_t d = (uint32_t)link(i, j, k, n, q);
                        o[(size_t)((q - 1) * P + n)] = (int16_t)(uint16_t)(d & 0xFFFFu);
                        oh[(size_t)((q - 1) * P + n)] = (int16_t)(uint16_t)(d >> 16);
                    }
                }
        for (int q = 0; q < 18; ++q)
            for (int64_t n = N; n < P; ++n) {
                o[(size_t)(q * P + n)] = 0;
                oh[(size_t)(q * P + n)] = (int16_t)(uint16_t)((uint32_t)SRV_BB32 >> 16);
            }
    } else {
        for (int q = 0; q < 18; ++q)
            for (int64_t n = N; n < P; ++n) o[(size_t)(q * P + n)] = SRV_BB;
    }
    return true;
}

}  // namespace

class Server : public lbm_handle {
public:
    Server() : lbm_handle(SERVER) {}
    ~Server() override { shutdown(); }

    int create(const lbm_srv_params &p);
    int submit(const lbm_srv_request &r, int64_t *ticket);
    int poll(lbm_srv_result *out, int max, int *n_out);
    int wait_idle(double timeout_s);
    int stats(lbm_srv_stats *out);
    double clock() const { return std::chrono::duration<double>(Clock::now() - t0_).count(); }

    // queue an already prepared sample (precision escalation from an FP32 server)
    void enqueue(Prepared &&pr);
    void prep_worker();
    bool fits(int64_t npad) const { return npad <= cap_; }

private:
    void shutdown();
    void loop();
    bool escalate_locked(Active &a, bool restart);
    bool make_fallback(int64_t min_nodes);
    int run_chunk(std::unique_lock<std::mutex> &lk);
    bool try_admit_locked();
    int rebuild_tables(std::string &msg);
    int64_t alloc_segment(int64_t len);
    void free_segment(int64_t start, int64_t len);
    template <int BLK>
    hipError_t launch_steps(int nsteps);
    hipError_t launch_steps_dispatch(int nsteps);
    template <int BLK>
    hipError_t launch_flux();
    template <int BLK>
    hipError_t launch_aa(bool update);
    int autotune();

    lbm_srv_params prm_{};
    Clock::time_point t0_ = Clock::now();
    int device_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    size_t st_ = 4;
    int64_t cap_ = 0;
    int blk_ = 256;
    void *pops_[2] = {nullptr, nullptr};
    int cur_ = 0;
    int16_t *d_nbr_ = nullptr;
    int16_t *d_nbr_hi_ = nullptr;   // high halves of wide samples' deltas
    // device preparation of submitted samples (loop thread only): pore flags, voxel -> node
    // numbering and the scan's tile counts, for samples of up to prep_cap_ voxels
    uint8_t *d_prep_pore_ = nullptr;
    int32_t *d_prep_idx_ = nullptr, *d_prep_tile_ = nullptr, *d_prep_flag_ = nullptr, *h_prep_flag_ = nullptr;
    int64_t prep_cap_ = 0;
    hipError_t prepare_on_device(Active &a);
    // Anderson acceleration: X, R_prev, G_prev, then m dR and m dG vectors, each 19 x CAP
    int aa_m_ = 0, aa_period_ = 0, aa_warm_ = 0;
    void *d_aa_vec_ = nullptr;
    AaDev *d_aa_ = nullptr, *h_aa_ = nullptr;   // h_aa_: two pinned halves, alternating by chunk
    int aa_half_ = 0;
    double *d_aa_partial_ = nullptr, *d_aa_sums_ = nullptr, *h_aa_sums_ = nullptr;
    SlotDev *d_slots_ = nullptr;
    TileDev *d_tiles_ = nullptr;
    int32_t *d_tlo_ = nullptr, *d_thi_ = nullptr;
    double *d_partial_ = nullptr, *d_sums_ = nullptr, *h_sums_ = nullptr;
    int64_t max_slots_ = 0, max_tiles_ = 0;
    int64_t ntiles_ = 0;
    int64_t ntiles_narrow_ = 0;   // tiles 0 .. ntiles_narrow_ - 1 belong to narrow samples
    int64_t global_step_ = 0;

    std::mutex mu_;
    std::condition_variable cv_, idle_cv_;
    std::thread thr_;
    bool stop_ = false;
    bool loop_error_ = false;
    std::string loop_msg_;
    int64_t next_ticket_ = 1;
    std::deque<Prepared> pending_;
    std::deque<PrepJob> prep_q_;   // submitted, geometry not built yet
    int prep_busy_ = 0;
    int nprep_ = 1;          // connectivity-test workers
    std::vector<std::thread> prep_thr_;
    std::condition_variable prep_cv_;
    std::vector<Active> active_;   // in tile order
    std::vector<Segment> free_;
    std::deque<lbm_srv_result> done_;
    int64_t n_done_ = 0, launches_ = 0, node_updates_ = 0;
    double busy_s_ = 0;
    // FP64 server for escalated samples, created at the first escalation
    std::unique_ptr<Server> fallback_;
    int64_t escalated_ = 0;
};

void Server::enqueue(Prepared &&pr)
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        pending_.push_back(std::move(pr));
    }
    cv_.notify_all();
}

// Read the sample's neighbour deltas back from the arena and hand the sample to the FP64
// server; false (the FP32 result stands) if that server cannot be made or cannot hold it.
// The internal FP64 server that escalated samples continue on.  It is made when the server is
// created, so that an escalation costs a state copy and not an arena allocation on the loop.
bool Server::make_fallback(int64_t min_nodes)
{
    lbm_srv_params fp = prm_;
    fp.precision = LBM_FP64;
    fp.escalate_k = 0.0;
    // small tiles: the FP64 server runs a few samples beside the FP32 arena's launches, and
    // 128-node workgroups find room between those launches' workgroups sooner (measured on the
    // 32-request stream: the escalated sample's 11,700 FP64 steps then end before the FP32 work)
    fp.block = 128;
    fp.capacity_nodes = std::max<int64_t>({int64_t(4) << 20, cap_ / 8, min_nodes});
    fp.mem_fraction = 0.0;
    auto f = std::unique_ptr<Server>(new (std::nothrow) Server());
    if (!f) return false;
    f->t0_ = t0_;   // one clock for submit, admit and done times
    if (f->create(fp) != LBM_OK) return false;
    fallback_ = std::move(f);
    return true;
}

// restart: the FP64 solve starts from rest and the FP32 steps are not counted, so that the
// result does not depend on how far the FP32 run had got (the hand-over of a sample found not to
// wrap comes from a preparation thread, at a time that depends on scheduling).
bool Server::escalate_locked(Active &a, bool restart)
{
    if (!fallback_ && !make_fallback(a.req.npad)) return false;
    if (!fallback_->fits(a.req.npad)) return false;
    a.req.nbr = std::shared_ptr<int16_t[]>(new int16_t[(size_t)18 * a.req.npad]);
    hipError_t e = hipMemcpy2DAsync(a.req.nbr.get(), (size_t)a.req.npad * sizeof(int16_t), d_nbr_ + a.base,
                                    (size_t)cap_ * sizeof(int16_t), (size_t)a.req.npad * sizeof(int16_t), 18,
                                    hipMemcpyDeviceToHost, stream_);
    if (a.req.wide && e == hipSuccess) {
        a.req.nbr_hi = std::shared_ptr<int16_t[]>(new int16_t[(size_t)18 * a.req.npad]);
        e = hipMemcpy2DAsync(a.req.nbr_hi.get(), (size_t)a.req.npad * sizeof(int16_t), d_nbr_hi_ + a.base,
                             (size_t)cap_ * sizeof(int16_t), (size_t)a.req.npad * sizeof(int16_t), 18,
                             hipMemcpyDeviceToHost, stream_);
    }
    // otherwise the FP64 solve continues from the FP32 state rather than from rest: the
    // transient so far (the pressure build-up of the slow mode) is kept
    if (!restart) {
        a.req.carried.resize((size_t)Q * a.req.npad);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(a.req.carried.data(), (size_t)a.req.npad * sizeof(float),
                                 (const float *)pops_[cur_] + a.base, (size_t)cap_ * sizeof(float),
                                 (size_t)a.req.npad * sizeof(float), Q, hipMemcpyDeviceToHost, stream_);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return false;
    if (!restart) {
        a.req.prior_steps += a.steps;
        a.req.prior_acc += a.aa_updates;
    }
    // the FP64 server is the answer whatever the connectivity test says
    a.req.wrap_pending = false;
    fallback_->enqueue(std::move(a.req));
    ++escalated_;
    return true;
}

int64_t Server::alloc_segment(int64_t len)
{
    for (size_t i = 0; i < free_.size(); ++i) {
        if (free_[i].len >= len) {
            const int64_t s = free_[i].start;
            free_[i].start += len;
            free_[i].len -= len;
            if (free_[i].len == 0) free_.erase(free_.begin() + (long)i);
            return s;
        }
    }
    return -1;
}

void Server::free_segment(int64_t start, int64_t len)
{
    free_.push_back({start, len});
    std::sort(free_.begin(), free_.end(), [](const Segment &a, const Segment &b) { return a.start < b.start; });
    std::vector<Segment> merged;
    for (const Segment &s : free_) {
        if (!merged.empty() && merged.back().start + merged.back().len == s.start)
            merged.back().len += s.len;
        else
            merged.push_back(s);
    }
    free_.swap(merged);
}

int Server::create(const lbm_srv_params &p)
{
    prm_ = p;
    if (prm_.check_interval <= 0) prm_.check_interval = 100;
    if (!(prm_.tol > 0)) prm_.tol = 1e-6;
    if (prm_.min_steps <= 0) prm_.min_steps = 2 * prm_.check_interval;
    if (prm_.max_steps <= 0) prm_.max_steps = 200000;
    // FP16 storage is not offered: on low-flux samples (g = 1e-6, porosity 0.15-0.2) its
    // rounding leaves k errors of 1e-4 to 1e-2 and the relative-change test never settles
    if (prm_.precision == LBM_FP16) return fail(LBM_ERR_ARG, "the server stores populations in FP32 or FP64");
    if (prm_.precision != LBM_FP64) prm_.precision = LBM_FP32;
    if (prm_.block != 0 && prm_.block != 128 && prm_.block != 256 && prm_.block != 512)
        return fail(LBM_ERR_ARG, "block must be 0, 128, 256 or 512");
    if (prm_.accel_depth < 0 || prm_.accel_depth > AA_MAX) return fail(LBM_ERR_ARG, "accel_depth must be 0..8");
    aa_m_ = prm_.accel_depth;
    aa_period_ = prm_.accel_period > 0 ? prm_.accel_period : 2 * prm_.check_interval;
    aa_warm_ = prm_.accel_warmup > 0 ? prm_.accel_warmup : 5 * prm_.check_interval;
    if (aa_m_ > 0 && (aa_period_ % prm_.check_interval != 0 || aa_warm_ % prm_.check_interval != 0))
        return fail(LBM_ERR_ARG, "accel_period and accel_warmup must be multiples of check_interval");
    st_ = prm_.precision == LBM_FP64 ? 8 : 4;
    device_ = p.device;
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    LBM_HIP_TRY(hipEventCreate(&ev0_));
    LBM_HIP_TRY(hipEventCreate(&ev1_));
    int64_t cap = p.capacity_nodes;
    if (cap <= 0) {
        size_t free_b = 0, total_b = 0;
        LBM_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        const double frac = p.mem_fraction > 0 && p.mem_fraction < 0.95 ? p.mem_fraction : 0.5;
        const double bytes = std::min(frac * (double)free_b, 64e9);
        cap = (int64_t)(bytes / (double)(2 * Q * st_ + 2 * 18 * sizeof(int16_t) + (aa_m_ ? (3 + 2 * aa_m_) * Q * st_ : 0)));
    }
    // the step kernel's 32-bit byte offsets reach (19 * CAP - 1) * sizeof(value)
    cap = std::min<int64_t>(cap, (int64_t)(0xffffffffu / (Q * st_)) - 511);
    cap_ = ((cap + 511) / 512) * 512;
    if (cap_ < 512) return fail(LBM_ERR_ARG, "arena too small");
    for (void *&b : pops_) LBM_HIP_TRY(hipMalloc(&b, (size_t)Q * cap_ * st_));
    LBM_HIP_TRY(hipMalloc(&d_nbr_, (size_t)18 * cap_ * sizeof(int16_t)));
    LBM_HIP_TRY(hipMalloc(&d_nbr_hi_, (size_t)18 * cap_ * sizeof(int16_t)));
    max_slots_ = 1 << 16;
    max_tiles_ = cap_ / 128 + 1;
    LBM_HIP_TRY(hipMalloc(&d_slots_, (size_t)max_slots_ * sizeof(SlotDev)));
    LBM_HIP_TRY(hipMalloc(&d_tiles_, (size_t)max_tiles_ * sizeof(TileDev)));
    LBM_HIP_TRY(hipMalloc(&d_tlo_, (size_t)max_slots_ * sizeof(int32_t)));
    LBM_HIP_TRY(hipMalloc(&d_thi_, (size_t)max_slots_ * sizeof(int32_t)));
    LBM_HIP_TRY(hipMalloc(&d_partial_, (size_t)max_tiles_ * 4 * sizeof(double)));
    LBM_HIP_TRY(hipMalloc(&d_sums_, (size_t)max_slots_ * 4 * sizeof(double)));
    LBM_HIP_TRY(hipHostMalloc((void **)&h_sums_, (size_t)max_slots_ * 4 * sizeof(double), hipHostMallocDefault));
    if (aa_m_ > 0) {
        LBM_HIP_TRY(hipMalloc(&d_aa_vec_, (size_t)(3 + 2 * aa_m_) * Q * cap_ * st_));
        LBM_HIP_TRY(hipMalloc(&d_aa_, (size_t)max_slots_ * sizeof(AaDev)));
        LBM_HIP_TRY(hipHostMalloc((void **)&h_aa_, (size_t)2 * max_slots_ * sizeof(AaDev), hipHostMallocDefault));
        LBM_HIP_TRY(hipMalloc(&d_aa_partial_, (size_t)max_tiles_ * AA_NV * sizeof(double)));
        LBM_HIP_TRY(hipMalloc(&d_aa_sums_, (size_t)max_slots_ * AA_NV * sizeof(double)));
        LBM_HIP_TRY(hipHostMalloc((void **)&h_aa_sums_, (size_t)max_slots_ * AA_NV * sizeof(double), hipHostMallocDefault));
    }
    free_.push_back({0, cap_});
    blk_ = prm_.block ? prm_.block : 256;
    if (prm_.block == 0) {
        int rc = autotune();
        if (rc != LBM_OK) return rc;
    }
    if (prm_.precision == LBM_FP32 && prm_.escalate_k > 0.0 && !make_fallback(0))
        return fail(LBM_ERR_NOMEM, "could not create the FP64 server for escalated samples");
    thr_ = std::thread([this] { loop(); });
    // geometry preparation off the caller's thread: each worker numbers a sample and builds
    // its neighbour deltas while the device steps the admitted ones
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    nprep_ = (int)std::min(8u, std::max(1u, hw / 2));
    for (int i = 0; i < nprep_; ++i) prep_thr_.emplace_back([this] { prep_worker(); });
    return LBM_OK;
}

void Server::shutdown()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    prep_cv_.notify_all();
    for (std::thread &t : prep_thr_)
        if (t.joinable()) t.join();
    prep_thr_.clear();
    if (thr_.joinable()) thr_.join();
    (void)hipSetDevice(device_);
    for (void *&b : pops_)
        if (b) (void)hipFree(b), b = nullptr;
    void **bufs[] = {(void **)&d_nbr_, (void **)&d_nbr_hi_, (void **)&d_slots_, (void **)&d_tiles_, (void **)&d_tlo_, (void **)&d_thi_,
                     (void **)&d_partial_, (void **)&d_sums_, &d_aa_vec_, (void **)&d_aa_, (void **)&d_aa_partial_,
                     (void **)&d_aa_sums_};
    for (void **b : bufs)
        if (*b) (void)hipFree(*b), *b = nullptr;
    for (void *b : {(void *)d_prep_pore_, (void *)d_prep_idx_, (void *)d_prep_tile_, (void *)d_prep_flag_})
        if (b) (void)hipFree(b);
    d_prep_pore_ = nullptr;
    d_prep_idx_ = d_prep_tile_ = d_prep_flag_ = nullptr;
    if (h_prep_flag_) (void)hipHostFree(h_prep_flag_), h_prep_flag_ = nullptr;
    if (h_sums_) (void)hipHostFree(h_sums_), h_sums_ = nullptr;
    if (h_aa_) (void)hipHostFree(h_aa_), h_aa_ = nullptr;
    if (h_aa_sums_) (void)hipHostFree(h_aa_sums_), h_aa_sums_ = nullptr;
    if (ev0_) (void)hipEventDestroy(ev0_), ev0_ = nullptr;
    if (ev1_) (void)hipEventDestroy(ev1_), ev1_ = nullptr;
    if (stream_) (void)hipStreamDestroy(stream_), stream_ = nullptr;
}

int Server::submit(const lbm_srv_request &r, int64_t *ticket)
{
    if (!r.solid_ijk || r.nx <= 0 || r.ny <= 0 || r.nz <= 0) return fail(LBM_ERR_ARG, "invalid request geometry");
    if (!(r.nu > 0)) return fail(LBM_ERR_ARG, "viscosity must be positive");
    if (r.perm_dir < 1 || r.perm_dir > 3) return fail(LBM_ERR_ARG, "perm_dir must be 1, 2 or 3");
    if (r.g[r.perm_dir - 1] == 0.0) return fail(LBM_ERR_ARG, "body force along perm_dir is zero");
    const int64_t V = (int64_t)r.nx * r.ny * r.nz;
    if (V >= ((int64_t)1 << 31)) return fail(LBM_ERR_ARG, "box has 2^31 or more voxels");
    // the checks that decide admissibility are O(voxels) and done here, on a byte copy of the
    // pore flags made in the same pass as the pore count; the sample is queued at once, and the
    // numbering and neighbour deltas are built on the device when it is admitted
    auto pore = std::make_shared<std::vector<uint8_t>>((size_t)V);
    int64_t N = 0;
    for (int64_t v = 0; v < V; ++v) {
        const uint8_t p = r.solid_ijk[v] == 0;
        (*pore)[(size_t)v] = p;
        N += p;
    }
    if (N == 0) return fail(LBM_ERR_ARG, "geometry has no pore voxels");
    const int64_t P = padded_stride(N);
    if (P > cap_) return fail(LBM_ERR_ARG, "sample larger than the whole arena");
    Prepared pr;
    pr.pore = pore;
    pr.nx = r.nx;
    pr.ny = r.ny;
    pr.nz = r.nz;
    pr.N = N;
    pr.npad = P;
    pr.nu = r.nu;
    for (int c = 0; c < 3; ++c) pr.g[c] = r.g[c];
    pr.dir = r.perm_dir;
    pr.volume = V;
    // the connectivity test (escalation on): a worker runs it beside the device work
    const bool test_wrap = prm_.precision == LBM_FP32 && prm_.escalate_k > 0.0 && fallback_ && fallback_->fits(P);
    {
        std::lock_guard<std::mutex> lk(mu_);
        pr.ticket = next_ticket_++;
        pr.t_submit = clock();
        if (ticket) *ticket = pr.ticket;
        pr.wrap_pending = test_wrap;
        if (test_wrap) {
            PrepJob job;
            job.ticket = pr.ticket;
            job.nx = r.nx;
            job.ny = r.ny;
            job.nz = r.nz;
            job.pore = pore;
            job.dir = r.perm_dir;
            prep_q_.push_back(std::move(job));
        }
        pending_.push_back(std::move(pr));
    }
    cv_.notify_all();
    if (test_wrap) prep_cv_.notify_one();
    return LBM_OK;
}

// A sample whose pores do not wrap round the box along the permeability direction has no
// throughflow: its k is the lattice's spurious-current residue, orders of magnitude below what
// FP32 deviations resolve.  With escalation on it goes to the FP64 server: at once if it is still
// queued, at its next check if it has been admitted meanwhile.  The test runs on these workers,
// beside the device work, so that it does not delay admission.
void Server::prep_worker()
{
    for (;;) {
        PrepJob job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            prep_cv_.wait(lk, [this] { return stop_ || !prep_q_.empty(); });
            if (stop_) return;
            job = std::move(prep_q_.front());
            prep_q_.pop_front();
            ++prep_busy_;
        }
        const bool wraps = pore_flags_wrap(job.nx, job.ny, job.nz, job.pore->data(), job.dir - 1);
        job.pore.reset();
        {
            std::lock_guard<std::mutex> lk(mu_);
            --prep_busy_;
            for (Prepared &p : pending_)
                if (p.ticket == job.ticket) p.wrap_pending = false;
            for (Active &a : active_)
                if (a.req.ticket == job.ticket) a.req.wrap_pending = false;
            if (!wraps) {
                bool moved = false;
                for (auto it = pending_.begin(); it != pending_.end(); ++it)
                    if (it->ticket == job.ticket) {
                        fallback_->enqueue(std::move(*it));
                        pending_.erase(it);
                        ++escalated_;
                        moved = true;
                        break;
                    }
                if (!moved)
                    for (Active &a : active_)
                        if (a.req.ticket == job.ticket) a.no_wrap = true;
            }
        }
        cv_.notify_all();
        idle_cv_.notify_all();
    }
}

int Server::poll(lbm_srv_result *out, int max, int *n_out)
{
    std::lock_guard<std::mutex> lk(mu_);
    int n = 0;
    while (n < max && !done_.empty()) {
        out[n++] = done_.front();
        done_.pop_front();
    }
    int rc = LBM_OK;
    if (fallback_ && n < max) {
        int m = 0;
        rc = fallback_->poll(out + n, max - n, &m);
        n += m;
        if (rc != LBM_OK) fail(rc, fallback_->error());
    }
    if (n_out) *n_out = n;
    if (loop_error_) return fail(LBM_ERR_HIP, loop_msg_);
    return rc;
}

int Server::wait_idle(double timeout_s)
{
    const auto t_end = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
    std::unique_lock<std::mutex> lk(mu_);
    auto idle = [this] {
        return loop_error_ || (prep_q_.empty() && prep_busy_ == 0 && pending_.empty() && active_.empty());
    };
    if (timeout_s > 0) {
        if (!idle_cv_.wait_until(lk, t_end, idle)) return LBM_ERR_BUSY;
    } else {
        idle_cv_.wait(lk, idle);
    }
    if (loop_error_) return fail(LBM_ERR_HIP, loop_msg_);
    // no sample is left here to escalate: wait for the FP64 server's share
    if (fallback_) {
        double rest = 0.0;
        if (timeout_s > 0) {
            rest = std::chrono::duration<double>(t_end - Clock::now()).count();
            if (rest <= 0) return LBM_ERR_BUSY;
        }
        const int rc = fallback_->wait_idle(rest);
        if (rc != LBM_OK) return rc == LBM_ERR_BUSY ? rc : fail(rc, fallback_->error());
    }
    return LBM_OK;
}

int Server::stats(lbm_srv_stats *out)
{
    std::lock_guard<std::mutex> lk(mu_);
    out->pending = (int64_t)(pending_.size() + prep_q_.size()) + prep_busy_;
    out->active = (int64_t)active_.size();
    out->done = n_done_;
    out->arena_nodes = cap_;
    int64_t used = 0;
    for (const Active &a : active_) used += a.req.npad;
    out->arena_used = used;
    out->launches = launches_;
    out->node_updates = node_updates_;
    out->block = blk_;
    out->pad = 0;
    out->busy_seconds = busy_s_;
    if (fallback_) {
        lbm_srv_stats f{};
        fallback_->stats(&f);
        out->pending += f.pending;
        out->active += f.active;
        out->done += f.done;
        out->launches += f.launches;
        out->node_updates += f.node_updates;
        out->busy_seconds += f.busy_seconds;
    }
    return LBM_OK;
}

// Admit queued samples in arrival order while they fit; a sample that does not fit waits and
// later, smaller ones may pass it (admission by HBM footprint).
bool Server::try_admit_locked()
{
    bool any = false;
    for (auto it = pending_.begin(); it != pending_.end();) {
        if ((int64_t)active_.size() >= max_slots_) break;
        const int64_t base = alloc_segment(it->npad);
        if (base < 0) {
            ++it;
            continue;
        }
        Active a;
        a.req = std::move(*it);
        a.base = base;
        a.first_step = global_step_;
        a.t_admit = clock();
        a.aa_next = aa_warm_;
        hipError_t e = hipSuccess;
        if (a.req.pore) {
            // a submitted sample: numbered and linked on the device, into the segment's rows
            e = prepare_on_device(a);
        } else {
            // a handed-over sample: its neighbour rows go straight into the arena rows
            e = hipMemcpy2DAsync(d_nbr_ + base, (size_t)cap_ * sizeof(int16_t), a.req.nbr.get(),
                                 (size_t)a.req.npad * sizeof(int16_t), (size_t)a.req.npad * sizeof(int16_t), 18,
                                 hipMemcpyHostToDevice, stream_);
            if (e == hipSuccess && a.req.wide)
                e = hipMemcpy2DAsync(d_nbr_hi_ + base, (size_t)cap_ * sizeof(int16_t), a.req.nbr_hi.get(),
                                     (size_t)a.req.npad * sizeof(int16_t), (size_t)a.req.npad * sizeof(int16_t), 18,
                                     hipMemcpyHostToDevice, stream_);
        }
        if (e == hipSuccess && !a.req.carried.empty()) {
            // an escalated sample resumes from its FP32 state: no equilibrium step
            std::vector<double> st(a.req.carried.begin(), a.req.carried.end());
            std::vector<float>().swap(a.req.carried);
            e = hipMemcpy2DAsync((char *)pops_[cur_] + (size_t)base * st_, (size_t)cap_ * st_, st.data(),
                                 (size_t)a.req.npad * sizeof(double), (size_t)a.req.npad * sizeof(double), Q,
                                 hipMemcpyHostToDevice, stream_);
            if (e == hipSuccess) e = hipStreamSynchronize(stream_);
            a.first_step = -1;
        }
        if (e != hipSuccess) {
            loop_error_ = true;
            loop_msg_ = std::string("neighbour upload failed: ") + hipGetErrorString(e);
            free_segment(base, a.req.npad);
            return any;
        }
        active_.push_back(std::move(a));
        it = pending_.erase(it);
        any = true;
    }
    if (any) {
        // the uploads read pageable host memory: finish them before the host copies go away
        const hipError_t e = hipStreamSynchronize(stream_);
        if (e != hipSuccess) {
            loop_error_ = true;
            loop_msg_ = std::string("neighbour upload failed: ") + hipGetErrorString(e);
        }
        for (Active &a : active_) {
            a.req.nbr.reset();
            a.req.nbr_hi.reset();
            a.req.pore.reset();
        }
    }
    return any;
}

// Called on the loop thread (and by autotune at create): errors are returned with their
// message in msg, not through fail(), so the loop thread never writes last_error.
// Number and link a submitted sample on the device (the preparation kernels above), into the
// arena rows of its segment a.base; sets a.req.wide.  Loop thread only.  The pore flags are read
// from pageable memory: the caller synchronises the stream before they go.
hipError_t Server::prepare_on_device(Active &a)
{
    const int64_t V = (int64_t)a.req.nx * a.req.ny * a.req.nz;
    hipError_t e = hipSuccess;
    if (V > prep_cap_) {
        for (void *b : {(void *)d_prep_pore_, (void *)d_prep_idx_, (void *)d_prep_tile_})
            if (b) (void)hipFree(b);
        d_prep_pore_ = nullptr;
        d_prep_idx_ = d_prep_tile_ = nullptr;
        prep_cap_ = 0;
        const int64_t cap = std::max<int64_t>(V, int64_t(1) << 22);
        e = hipMalloc(&d_prep_pore_, (size_t)cap);
        if (e == hipSuccess) e = hipMalloc(&d_prep_idx_, (size_t)cap * sizeof(int32_t));
        if (e == hipSuccess) e = hipMalloc(&d_prep_tile_, (size_t)(cap / PREP_TILE + 1) * sizeof(int32_t));
        if (e == hipSuccess && !d_prep_flag_) e = hipMalloc(&d_prep_flag_, sizeof(int32_t));
        if (e == hipSuccess && !h_prep_flag_) e = hipHostMalloc((void **)&h_prep_flag_, sizeof(int32_t), hipHostMallocDefault);
        if (e != hipSuccess) return e;
        prep_cap_ = cap;
    }
    const int64_t ntiles = (V + PREP_TILE - 1) / PREP_TILE;
    e = hipMemsetAsync(d_prep_flag_, 0, sizeof(int32_t), stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(d_prep_pore_, a.req.pore->data(), (size_t)V, hipMemcpyHostToDevice, stream_);
    if (e != hipSuccess) return e;
    srv_prep_count_kernel<<<(unsigned)ntiles, 256, 0, stream_>>>(d_prep_pore_, V, d_prep_tile_);
    srv_prep_scan_kernel<<<1, 256, 0, stream_>>>(d_prep_tile_, ntiles, a.req.N, d_prep_flag_);
    srv_prep_number_kernel<<<(unsigned)ntiles, 256, 0, stream_>>>(d_prep_pore_, V, d_prep_tile_, d_prep_idx_);
    const unsigned gl = (unsigned)((V + 255) / 256);
    srv_prep_links_kernel<false><<<gl, 256, 0, stream_>>>(d_prep_idx_, a.req.nx, a.req.ny, a.req.nz, a.req.N, d_nbr_,
                                                           d_nbr_hi_, cap_, a.base, d_prep_flag_);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h_prep_flag_, d_prep_flag_, sizeof(int32_t), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return e;
    if (*h_prep_flag_ & PREP_COUNT) return hipErrorInvalidValue;   // cannot happen: the counts are the host's
    a.req.wide = (*h_prep_flag_ & PREP_WIDE) != 0;
    if (a.req.wide) {
        srv_prep_links_kernel<true><<<gl, 256, 0, stream_>>>(d_prep_idx_, a.req.nx, a.req.ny, a.req.nz, a.req.N, d_nbr_,
                                                              d_nbr_hi_, cap_, a.base, d_prep_flag_);
        e = hipGetLastError();
    }
    return e;
}

int Server::rebuild_tables(std::string &msg)
{
    // order active samples by arena position (keeps neighbouring tiles adjacent in memory)
    std::sort(active_.begin(), active_.end(), [](const Active &a, const Active &b) { return a.base < b.base; });
    std::vector<SlotDev> slots(active_.size());
    std::vector<TileDev> tiles;
    std::vector<int32_t> tlo(active_.size()), thi(active_.size());
    const double c_s2 = 1.0 / 3.0;
    for (size_t s = 0; s < active_.size(); ++s) {
        const Active &a = active_[s];
        SlotDev d{};
        d.gx = a.req.g[0];
        d.gy = a.req.g[1];
        d.gz = a.req.g[2];
        const double tau = a.req.nu / (c_s2 * 1.0) + 0.5;   // SP:1302-1307
        d.sv = 1.0 / tau;
        d.sq = 8.0 * (2.0 - d.sv) / (8.0 - d.sv);
        d.rho0 = 1.0;
        d.first_step = a.first_step;
        d.count = a.req.N;
        d.wide = a.req.wide ? 1 : 0;
        slots[s] = d;
    }
    // the tiles of narrow samples first, then those of wide ones (each launched by its own
    // kernel); a sample's tiles stay consecutive and in node order
    for (int pass = 0; pass < 2; ++pass) {
        for (size_t s = 0; s < active_.size(); ++s) {
            const Active &a = active_[s];
            if (a.req.wide != (pass == 1)) continue;
            tlo[s] = (int32_t)tiles.size();
            for (int64_t b = 0; b < a.req.N; b += blk_) {
                TileDev t{};
                t.slot = (int32_t)s;
                t.begin = a.base + b;
                t.end = a.base + std::min<int64_t>(b + blk_, a.req.N);
                t.base = a.base;
                tiles.push_back(t);
            }
            thi[s] = (int32_t)tiles.size();
        }
        if (pass == 0) ntiles_narrow_ = (int64_t)tiles.size();
    }
    ntiles_ = (int64_t)tiles.size();
    if (ntiles_ > max_tiles_) {
        msg = "tile table overflow";
        return LBM_ERR_ARG;
    }
    if (!slots.empty()) {
        hipError_t e = hipMemcpyAsync(d_slots_, slots.data(), slots.size() * sizeof(SlotDev), hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_tiles_, tiles.data(), tiles.size() * sizeof(TileDev), hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_tlo_, tlo.data(), tlo.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_thi_, thi.data(), thi.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_);
        // pageable sources: make sure the copies have consumed them before the vectors go away
        if (e == hipSuccess) e = hipStreamSynchronize(stream_);
        if (e != hipSuccess) {
            msg = std::string("server table upload failed: ") + hipGetErrorString(e);
            return LBM_ERR_HIP;
        }
    }
    return LBM_OK;
}

template <int BLK>
hipError_t Server::launch_steps(int nsteps)
{
    const int64_t nn = ntiles_narrow_, nw = ntiles_ - ntiles_narrow_;
    for (int s = 0; s < nsteps; ++s) {
        if (prm_.precision == LBM_FP64) {
            const double *src = (const double *)pops_[cur_];
            double *dst = (double *)pops_[cur_ ^ 1];
            if (nn > 0)
                srv_step_kernel_wide<double, BLK, false><<<(unsigned)nn, BLK, 0, stream_>>>(
                    d_tiles_, d_slots_, src, dst, d_nbr_, d_nbr_hi_, cap_, global_step_, 0);
            if (nw > 0)
                srv_step_kernel_wide<double, BLK, true><<<(unsigned)nw, BLK, 0, stream_>>>(
                    d_tiles_, d_slots_, src, dst, d_nbr_, d_nbr_hi_, cap_, global_step_, nn);
        } else {
            const float *src = (const float *)pops_[cur_];
            float *dst = (float *)pops_[cur_ ^ 1];
            if (nn > 0)
                srv_step_kernel<float, BLK, false><<<(unsigned)nn, BLK, 0, stream_>>>(
                    d_tiles_, d_slots_, src, dst, d_nbr_, d_nbr_hi_, cap_, global_step_, 0);
            if (nw > 0)
                srv_step_kernel<float, BLK, true><<<(unsigned)nw, BLK, 0, stream_>>>(
                    d_tiles_, d_slots_, src, dst, d_nbr_, d_nbr_hi_, cap_, global_step_, nn);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        cur_ ^= 1;
        ++global_step_;
    }
    return hipSuccess;
}

hipError_t Server::launch_steps_dispatch(int nsteps)
{
    switch (blk_) {
    case 128: return launch_steps<128>(nsteps);
    case 512: return launch_steps<512>(nsteps);
    default: return launch_steps<256>(nsteps);
    }
}

template <int BLK>
hipError_t Server::launch_flux()
{
    if (prm_.precision == LBM_FP64)
        srv_flux_kernel<double, BLK><<<(unsigned)ntiles_, BLK, 0, stream_>>>(d_tiles_, d_slots_, (const double *)pops_[cur_],
                                                                             d_nbr_, d_nbr_hi_, cap_, d_partial_);
    else
        srv_flux_kernel<float, BLK><<<(unsigned)ntiles_, BLK, 0, stream_>>>(d_tiles_, d_slots_, (const float *)pops_[cur_],
                                                                            d_nbr_, d_nbr_hi_, cap_, d_partial_);
    return hipGetLastError();
}

// Chunk of check_interval steps for every active sample, then the permeability check.
// Called with the lock held; the lock is released while the device works.
template <int BLK>
hipError_t Server::launch_aa(bool update)
{
    const size_t V = (size_t)Q * cap_ * st_;
    char *base = (char *)d_aa_vec_;
    void *X = base, *RP = base + V, *GP = base + 2 * V, *DR = base + 3 * V, *DG = base + (3 + aa_m_) * V;
    if (prm_.precision == LBM_FP64) {
        if (update)
            srv_aa_update_kernel<double, BLK><<<(unsigned)ntiles_, BLK, 0, stream_>>>(
                d_tiles_, d_aa_, (double *)pops_[cur_], (double *)X, (const double *)DG, cap_);
        else
            srv_aa_resid_kernel<double, BLK><<<(unsigned)ntiles_, BLK, 0, stream_>>>(
                d_tiles_, d_aa_, (const double *)pops_[cur_], (double *)X, (double *)RP, (double *)GP, (double *)DR,
                (double *)DG, cap_, d_aa_partial_);
    } else {
        if (update)
            srv_aa_update_kernel<float, BLK><<<(unsigned)ntiles_, BLK, 0, stream_>>>(
                d_tiles_, d_aa_, (float *)pops_[cur_], (float *)X, (const float *)DG, cap_);
        else
            srv_aa_resid_kernel<float, BLK><<<(unsigned)ntiles_, BLK, 0, stream_>>>(
                d_tiles_, d_aa_, (const float *)pops_[cur_], (float *)X, (float *)RP, (float *)GP, (float *)DR,
                (float *)DG, cap_, d_aa_partial_);
    }
    return hipGetLastError();
}

// Chunk of check_interval steps for every active sample, then the permeability check and the
// acceleration events that fall on this chunk.  Called with the lock held; the lock is released
// while the device works.
int Server::run_chunk(std::unique_lock<std::mutex> &lk)
{
    const int nsteps = prm_.check_interval;
    const int64_t nodes = [&] {
        int64_t s = 0;
        for (const Active &a : active_) s += a.req.N;
        return s;
    }();
    const int nact = (int)active_.size();
    // acceleration events of this chunk; this chunk's half of the pinned table (the other half
    // may still be read by the previous chunk's update copy)
    AaDev *const hb = h_aa_ ? h_aa_ + (size_t)aa_half_ * max_slots_ : nullptr;
    aa_half_ ^= 1;
    bool any_aa = false;
    if (aa_m_ > 0) {
        for (int s = 0; s < nact; ++s) {
            const Active &a = active_[s];
            AaDev d{};
            d.mode = AA_NONE;
            if (a.steps + nsteps == a.aa_next) {
                d.mode = a.aa_stage == 0 ? AA_SAVE : (a.aa_stage == 1 ? AA_FIRST : AA_RESID);
                d.head = a.aa_head;
                d.count = std::min(a.aa_count + 1, aa_m_);
                any_aa = true;
            }
            hb[s] = d;
        }
    }
    lk.unlock();
    hipError_t e = hipEventRecord(ev0_, stream_);
    if (e == hipSuccess) e = launch_steps_dispatch(nsteps);
    if (e == hipSuccess) e = hipEventRecord(ev1_, stream_);
    if (e == hipSuccess) e = blk_ == 128 ? launch_flux<128>() : (blk_ == 512 ? launch_flux<512>() : launch_flux<256>());
    if (e == hipSuccess) {
        srv_slot_sum_kernel<4><<<(unsigned)nact, BLOCK, 0, stream_>>>(d_partial_, d_tlo_, d_thi_, d_sums_);
        e = hipGetLastError();
    }
    if (e == hipSuccess && any_aa) {
        e = hipMemcpyAsync(d_aa_, hb, (size_t)nact * sizeof(AaDev), hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess) e = blk_ == 128 ? launch_aa<128>(false) : (blk_ == 512 ? launch_aa<512>(false) : launch_aa<256>(false));
        if (e == hipSuccess) {
            srv_slot_sum_kernel<AA_NV><<<(unsigned)nact, BLOCK, 0, stream_>>>(d_aa_partial_, d_tlo_, d_thi_, d_aa_sums_);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_aa_sums_, d_aa_sums_, (size_t)nact * AA_NV * sizeof(double), hipMemcpyDeviceToHost, stream_);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h_sums_, d_sums_, (size_t)nact * 4 * sizeof(double), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0_, ev1_);
    lk.lock();
    if (e != hipSuccess) {
        loop_error_ = true;
        loop_msg_ = std::string("server chunk failed: ") + hipGetErrorString(e);
        return LBM_ERR_HIP;
    }
    launches_ += nsteps;
    node_updates_ += nodes * nsteps;
    busy_s_ += ms * 1e-3;
    // retire converged samples (Comput_Perm k = Q/V nu/g, SP:4675-4677; relative change test:
    // over one check interval when stepping plainly, between two acceleration events otherwise)
    std::vector<int> status(nact, 2);
    std::vector<double> rels(nact, INFINITY);
    bool any_update = false;
    for (int s = 0; s < nact; ++s) {
        Active &a = active_[s];
        a.steps += nsteps;
        const double *q = h_sums_ + 4 * s;
        double kd = 0.0;
        {
            const int c = a.req.dir - 1;
            kd = q[c] / (double)a.req.volume * a.req.nu / a.req.g[c];
        }
        const bool event = aa_m_ > 0 && hb[s].mode != AA_NONE;
        double rel = INFINITY;
        if (aa_m_ == 0) {
            rel = a.has_prev ? std::fabs(kd - a.k_prev) / std::max(std::fabs(kd), 1e-300) : INFINITY;
        } else if (event) {
            rel = a.has_k_ev ? std::fabs(kd - a.k_ev) / std::max(std::fabs(kd), 1e-300) : INFINITY;
            a.k_ev = kd;
            a.has_k_ev = true;
            a.small_events = rel <= prm_.tol ? a.small_events + 1 : 0;
        }
        rels[s] = rel;
        // accelerated samples retire after two consecutive small changes: Anderson mixing can
        // stall for one period on a slow mode before the next extrapolation moves k again
        const bool settled = aa_m_ == 0 ? (a.has_prev && rel <= prm_.tol) : (event && a.small_events >= 2);
        if (q[3] != 0.0 || !std::isfinite(kd))
            status[s] = -1;
        else if (a.steps >= prm_.min_steps && settled)
            status[s] = 1;
        else if (a.steps >= prm_.max_steps)
            status[s] = 0;
        // a sample whose connectivity test is still running keeps stepping (unless it blew up)
        if (a.req.wrap_pending && status[s] != -1) status[s] = 2;
        // precision escalation: a flux too small for FP32 deviations, or no convergence
        if (prm_.precision == LBM_FP32 && prm_.escalate_k > 0.0 &&
            (status[s] == 0 || (status[s] == 2 && a.no_wrap) ||
             (status[s] == 2 && a.steps >= aa_warm_ && a.steps >= prm_.min_steps && std::fabs(kd) < prm_.escalate_k)))
            status[s] = 3;
        a.k_prev = kd;
        a.has_prev = true;
        if (!event) continue;
        // advance the acceleration state; a sample that retires now needs no update
        AaDev &d = hb[s];
        if (d.mode == AA_RESID && status[s] == 2) {
            const double *v = h_aa_sums_ + (size_t)s * AA_NV;
            const int head = d.head, cnt = d.count;
            for (int i = 0; i < cnt; ++i) a.gram[i][head] = a.gram[head][i] = v[i];
            const double r2 = v[2 * AA_MAX];
            double gam[AA_MAX] = {0.0};
            bool ok = std::isfinite(r2) && r2 <= 1e6 * a.r2_min && solve_small(a.gram, v + AA_MAX, cnt, gam);
            if (std::isfinite(r2)) a.r2_min = std::min(a.r2_min, r2);
            if (!ok) {   // restart the history: this step stays a plain one
                for (int i = 0; i < AA_MAX; ++i) gam[i] = 0.0;
                a.aa_count = 0;
                a.aa_head = 0;
            } else {
                a.aa_count = cnt;
                a.aa_head = (head + 1) % aa_m_;
                ++a.aa_updates;
            }
            for (int i = 0; i < AA_MAX; ++i) d.gamma[i] = gam[i];
            any_update = true;
        } else {
            d.mode = AA_NONE;
        }
        a.aa_stage = std::min(a.aa_stage + 1, 2);
        a.aa_next += aa_period_;
    }
    if (any_update) {
        // x = g - dG gamma for the samples that accelerated, in stream order before the next chunk
        e = hipMemcpyAsync(d_aa_, hb, (size_t)nact * sizeof(AaDev), hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess) e = blk_ == 128 ? launch_aa<128>(true) : (blk_ == 512 ? launch_aa<512>(true) : launch_aa<256>(true));
        if (e != hipSuccess) {
            loop_error_ = true;
            loop_msg_ = std::string("acceleration update failed: ") + hipGetErrorString(e);
            return LBM_ERR_HIP;
        }
    }
    bool changed = false;
    std::vector<Active> keep;
    keep.reserve(active_.size());
    for (int s = 0; s < nact; ++s) {
        Active &a = active_[s];
        if (status[s] == 2) {
            keep.push_back(std::move(a));
            continue;
        }
        if (status[s] == 3) {
            const int64_t base = a.base, npad = a.req.npad;
            if (escalate_locked(a, a.no_wrap)) {
                free_segment(base, npad);
                changed = true;
                continue;
            }
            status[s] = 0;   // could not escalate: report the FP32 state as unsettled
        }
        const double *q = h_sums_ + 4 * s;
        lbm_srv_result r{};
        r.ticket = a.req.ticket;
        for (int c = 0; c < 3; ++c) r.k[c] = a.req.g[c] != 0.0 ? q[c] / (double)a.req.volume * a.req.nu / a.req.g[c] : 0.0;
        r.rel_change = rels[s];
        r.fluid_nodes = a.req.N;
        r.converged = status[s];
        r.accelerations = a.req.prior_acc + a.aa_updates;
        r.precision = prm_.precision;
        r.steps = a.req.prior_steps + a.steps;
        r.t_submit = a.req.t_submit;
        r.t_admit = a.t_admit;
        r.t_done = clock();
        done_.push_back(r);
        ++n_done_;
        free_segment(a.base, a.req.npad);
        changed = true;
    }
    active_.swap(keep);
    return changed ? 1 : 0;
}

void Server::loop()
{
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    bool tables_dirty = true;
    while (!stop_) {
        if (!pending_.empty() && try_admit_locked()) tables_dirty = true;
        if (loop_error_) break;
        if (active_.empty()) {
            idle_cv_.notify_all();
            cv_.wait(lk, [this] { return stop_ || !pending_.empty(); });
            continue;
        }
        if (tables_dirty) {
            std::string msg;
            if (rebuild_tables(msg) != LBM_OK) {
                loop_error_ = true;
                loop_msg_ = msg;
                break;
            }
            tables_dirty = false;
        }
        const int rc = run_chunk(lk);
        if (rc < 0) break;
        if (rc > 0) tables_dirty = true;
        if (pending_.empty() && active_.empty()) idle_cv_.notify_all();
    }
    idle_cv_.notify_all();
}

// Pick the tile size by timing the step kernel on a synthetic porous sample.
int Server::autotune()
{
    // a 96^3 random porous sample (hashed blobs, porosity ~0.35)
    const int L = 96;
    std::vector<uint8_t> pore((size_t)L * L * L);
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j)
            for (int k = 0; k < L; ++k) {
                const double v = std::sin(0.37 * i + 1.3) * std::sin(0.29 * j + 0.7) * std::sin(0.41 * k + 2.1) +
                                 0.5 * std::sin(0.13 * (i + 2 * j) + 0.4) * std::cos(0.17 * (k - j));
                pore[((size_t)i * L + j) * L + k] = v > -0.05 ? 0 : 1;
            }
    int64_t N = 0, P = 0;
    std::shared_ptr<int16_t[]> d16, d16_hi;
    if (!build_deltas(L, L, L, pore.data(), N, P, d16, d16_hi) || d16_hi || P > cap_) {
        blk_ = 256;
        return LBM_OK;
    }
    Active a;
    a.req.N = N;
    a.req.npad = P;
    a.req.nu = 1.0 / 6.0;
    a.req.g[0] = 1e-6;
    a.req.volume = (int64_t)L * L * L;
    a.base = 0;
    a.first_step = 0;
    LBM_HIP_TRY(hipMemcpy2DAsync(d_nbr_, (size_t)cap_ * sizeof(int16_t), d16.get(), (size_t)P * sizeof(int16_t),
                                 (size_t)P * sizeof(int16_t), 18, hipMemcpyHostToDevice, stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    active_.push_back(std::move(a));
    float best = 1e30f;
    int best_blk = 256;
    for (int blk : {128, 256, 512}) {
        blk_ = blk;
        global_step_ = 0;
        std::string msg;
        if (rebuild_tables(msg) != LBM_OK) break;
        LBM_HIP_TRY(launch_steps_dispatch(5));   // warm-up (first step is the equilibrium init)
        LBM_HIP_TRY(hipEventRecord(ev0_, stream_));
        LBM_HIP_TRY(launch_steps_dispatch(40));
        LBM_HIP_TRY(hipEventRecord(ev1_, stream_));
        LBM_HIP_TRY(hipEventSynchronize(ev1_));
        float ms = 0.f;
        LBM_HIP_TRY(hipEventElapsedTime(&ms, ev0_, ev1_));
        if (ms < best) {
            best = ms;
            best_blk = blk;
        }
    }
    active_.clear();
    global_step_ = 0;
    blk_ = best_blk;
    return LBM_OK;
}

}  // namespace lbm

// ---------------------------------------------------------------------------- C ABI
namespace {
lbm::Server *as_srv(lbm_handle *h) { return (h && h->kind == lbm_handle::SERVER) ? static_cast<lbm::Server *>(h) : nullptr; }
}  // namespace

extern "C" {

int lbm_srv_create(const lbm_srv_params *p, lbm_handle **out)
{
    if (!p || !out) return LBM_ERR_ARG;
    *out = nullptr;
    auto *s = new (std::nothrow) lbm::Server();
    if (!s) return LBM_ERR_NOMEM;
    const int rc = s->create(*p);
    if (rc != LBM_OK) {
        lbm::set_create_error(s->error());
        delete s;
        return rc;
    }
    *out = s;
    return LBM_OK;
}

int lbm_srv_submit(lbm_handle *h, const lbm_srv_request *req, int64_t *ticket)
{
    auto *s = as_srv(h);
    if (!s || !req) return LBM_ERR_ARG;
    return s->submit(*req, ticket);
}

int lbm_srv_poll(lbm_handle *h, lbm_srv_result *out, int32_t max, int32_t *n_out)
{
    auto *s = as_srv(h);
    if (!s || (!out && max > 0)) return LBM_ERR_ARG;
    int n = 0;
    const int rc = s->poll(out, max, &n);
    if (n_out) *n_out = n;
    return rc;
}

int lbm_srv_wait_idle(lbm_handle *h, double timeout_s)
{
    auto *s = as_srv(h);
    return s ? s->wait_idle(timeout_s) : LBM_ERR_ARG;
}

int lbm_srv_get_stats(lbm_handle *h, lbm_srv_stats *out)
{
    auto *s = as_srv(h);
    if (!s || !out) return LBM_ERR_ARG;
    return s->stats(out);
}

double lbm_srv_clock(const lbm_handle *h)
{
    auto *s = (h && h->kind == lbm_handle::SERVER) ? static_cast<const lbm::Server *>(h) : nullptr;
    return s ? s->clock() : -1.0;
}

}  // extern "C"
