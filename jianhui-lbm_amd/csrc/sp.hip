// sp.hip -- single-phase sparse D3Q19 MRT: fused pull-stream + collide, diagnostics, state I/O.
//
// One launch per time step does what the reference's collision() + comput_macro_variables()
// pair does per step (SP:1406-2013, SP:2702-2746): every thread owns one fluid node, pulls its
// 19 post-streaming populations from the upstream nodes (halfway bounce-back folded into the
// neighbour index), takes the macroscopic moments, collides in moment space and writes the 19
// post-collision populations, coalesced, to the other buffer.  The stored state is therefore
// post-collision; the reference's post-streaming state is the pull of it (get_state).
//
// Data layout in HBM (structure of arrays, row stride P = N rounded up to 256):
//   pops[2]  19 x P  storage type (double / float / half), deviations f - w*rho0
//   nbr      packed upstream node of direction q = 1..18 or bounce back (nbrpack.h):
//            18 x P uint16 residuals + per-wave 32-bit bases
//   u_aux     3 x P  compute type, u of the previous step (for the Error diagnostic)
//   rho_aux   1 x P  compute type, density of a loaded checkpoint (first step only)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "boundary.h"
#include "collide.h"
#include "common.h"
#include "handle.h"
#include "pull.h"
#include "storage.h"
#include "tune.h"

namespace lbm {

// A coupled passive solute advanced in the same launch (D3Q7 BGK, the solute step of st.hip):
// its six links are the first six rows of the pull table the flow step has just decoded, and its
// velocity is the one the flow step has just computed, so neither is read from memory again.
template <typename C>
struct StFused {
    const C *__restrict__ src;   // 7 x P post-collision populations of the previous step
    C *__restrict__ dst;
    C omega, w0, wq, inv_cs2;    // 1/tau_s, rest and axis weights, 1/c_s^2
};

template <typename ST>
struct SpArgs {
    using C = typename Store<ST>::C;
    const ST *__restrict__ src;
    ST *__restrict__ dst;
    NbrDev nbr;
    C *__restrict__ u_aux;
    const C *__restrict__ rho_aux;
    int64_t N, P;
    C gx, gy, gz, rho0, sv, sq;
    C rho_i, ux_i, uy_i, uz_i;
    float scale, inv_scale;
    BcDev bc;   // face boundaries (bc.wave == nullptr: none)
    C *__restrict__ nu_vn;   // power-law collision: per-node viscosity (nullptr: Newtonian)
    C pl_m, pl_n;            // nu = m |S|^(n-1)
    StFused<C> st;           // coupled solute (pull launches of the fused variant only)
    int32_t order;           // tile order of the step launch (TileOrder)
};

// Power-law collision (collision_nnf, SP:2399-2491): update the node's viscosity from its
// strain rate, estimated from the non-equilibrium populations with the viscosity of the
// previous step, and return the rates s_v, s_q of the new viscosity.
template <typename C>
__device__ __forceinline__ void power_law_rates(const C (&h)[Q], const Macro<C> &mc, C *__restrict__ nu_vn, int64_t n,
                                                C pl_m, C pl_n, C &sv, C &sq)
{
    const C cs2 = C(1.0 / 3.0);
    const C sv0 = C(1) / (nu_vn[n] / cs2 + C(0.5));
    C sxx = 0, syy = 0, szz = 0, sxy = 0, syz = 0, sxz = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const C d = h[q] - feq_dev(q, mc.rho, mc.drho, mc.ux, mc.uy, mc.uz);   // f - feq
        if (EX[q] != 0) sxx += d;
        if (EY[q] != 0) syy += d;
        if (EZ[q] != 0) szz += d;
        if (EX[q] * EY[q] != 0) sxy += C(EX[q] * EY[q]) * d;
        if (EY[q] * EZ[q] != 0) syz += C(EY[q] * EZ[q]) * d;
        if (EX[q] * EZ[q] != 0) sxz += C(EX[q] * EZ[q]) * d;
    }
    const C c = -sv0 / (C(2) * mc.rho * cs2);
    sxx *= c;
    syy *= c;
    szz *= c;
    sxy *= c;
    syz *= c;
    sxz *= c;
    const C s2 = sxx * sxx + syy * syy + szz * szz + C(2) * sxy * sxy + C(2) * syz * syz + C(2) * sxz * sxz;
    const C nu = pl_m * exp((pl_n - C(1)) * log(sqrt(s2)));
    nu_vn[n] = nu;
    sv = C(1) / (nu / cs2 + C(0.5));
    sq = C(8) * (C(2) - sv) / (C(8) - sv);
}

template <typename ST, int MODE>
__device__ __forceinline__ void gather(const SpArgs<ST> &a, int64_t n, typename Store<ST>::C (&h)[Q])
{
    uint32_t nb[Q - 1];
    gather_pops<ST, MODE>(a.src, a.nbr, a.P, n, a.inv_scale, h, nb);
}

template <typename ST, int MODE>
__device__ __forceinline__ Macro<typename Store<ST>::C> node_macro(const SpArgs<ST> &a, int64_t n,
                                                                   const typename Store<ST>::C (&m)[Q])
{
    using C = typename Store<ST>::C;
    Macro<C> mc;
    if (MODE == K_INIT) {
        mc.rho = a.rho_i;
        mc.drho = a.rho_i - a.rho0;
        mc.ux = a.ux_i;
        mc.uy = a.uy_i;
        mc.uz = a.uz_i;
    } else if (MODE == K_DIRECT) {   // a loaded checkpoint: rho and u as given (Backup_init)
        mc.rho = a.rho_aux[n];
        mc.drho = mc.rho - a.rho0;
        mc.ux = a.u_aux[n];
        mc.uy = a.u_aux[a.P + n];
        mc.uz = a.u_aux[2 * a.P + n];
    } else {
        mc = macro_from_moments(m, a.rho0, a.gx, a.gy, a.gz);
    }
    return mc;
}

// BC: face boundaries on.  A pull launch replaces the pulled populations of face nodes by
// their rule, and every launch saves the pre-collision state of the nodes the rules read.
// PL: power-law collision (pull launches of reference steps n > 10, SP:2397).
// FUSED: also the step of a coupled solute (pull launches).
// NB: nodes (threads) per workgroup; a.order: the tile order.
template <typename ST, int MODE, bool STORE_U, bool BC, bool PL = false, bool FUSED = false, int NB = BLOCK>
__global__ void __launch_bounds__(NB) sp_step_kernel(SpArgs<ST> a)
{
    using C = typename Store<ST>::C;
    const int64_t n = tile_of(a.order, blockIdx.x, gridDim.x) * NB + threadIdx.x;
    if (n >= a.N) return;
    C h[Q], m[Q];
    Macro<C> mc;
    if (MODE == K_INIT) {
        mc = node_macro<ST, MODE>(a, n, m);
        sp_eq_moments(mc, m);
        if (BC && bc_is_saved(a.bc, n)) {
#pragma unroll
            for (int q = 0; q < Q; ++q) h[q] = feq_dev(q, mc.rho, mc.drho, mc.ux, mc.uy, mc.uz);
            bc_save(a.bc, n, h, mc.drho, mc.ux, mc.uy, mc.uz);
        }
    } else {
        uint32_t nb[Q - 1];
        gather_pops<ST, MODE>(a.src, a.nbr, a.P, n, a.inv_scale, h, nb);
        if (BC && MODE == K_PULL) bc_patch(a.bc, n, a.rho0, h);
        moments_fwd(h, m);
        mc = node_macro<ST, MODE>(a, n, m);
        if (BC) bc_save(a.bc, n, h, mc.drho, mc.ux, mc.uy, mc.uz);
        if (FUSED && MODE == K_PULL) {
            C g[7];
            g[0] = a.st.src[n];
#pragma unroll
            for (int q = 1; q < 7; ++q) {
                const uint32_t v = nb[q - 1];
                const int64_t row = (v & BB_FLAG) ? (int64_t)opp(q) : (int64_t)q;
                g[q] = a.st.src[row * a.P + (int64_t)(v & ~BB_FLAG)];
            }
            C c = g[0];
#pragma unroll
            for (int q = 1; q < 7; ++q) c += g[q];
            const C wc = a.st.wq * c;
            __builtin_nontemporal_store(g[0] - a.st.omega * (g[0] - a.st.w0 * c), &a.st.dst[n]);
#pragma unroll
            for (int q = 1; q < 7; ++q) {
                const C eu = q == 1 ? mc.ux : q == 2 ? -mc.ux : q == 3 ? mc.uy : q == 4 ? -mc.uy : q == 5 ? mc.uz : -mc.uz;
                const C eq = wc * (C(1) + eu * a.st.inv_cs2);
                __builtin_nontemporal_store(g[q] - a.st.omega * (g[q] - eq), &a.st.dst[(int64_t)q * a.P + n]);
            }
        }
    }
    if (STORE_U && MODE != K_DIRECT) {
        a.u_aux[n] = mc.ux;
        a.u_aux[a.P + n] = mc.uy;
        a.u_aux[2 * a.P + n] = mc.uz;
    }
    C sv = a.sv, sq = a.sq;
    if (PL && MODE == K_PULL) power_law_rates(h, mc, a.nu_vn, n, a.pl_m, a.pl_n, sv, sq);
    sp_relax(m, mc, a.gx, a.gy, a.gz, sv, sq, h);
#pragma unroll
    for (int q = 0; q < Q; ++q) Store<ST>::store_nt(&a.dst[(int64_t)q * a.P + n], h[q], a.scale);
}

template <typename C>
__global__ void __launch_bounds__(BLOCK) fill_kernel(C *__restrict__ p, C v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- diagnostics
constexpr int NDIAG = 8;   // Q0 Q1 Q2 err_num err_den speed_sum umax2 nan  (+ box count below)

struct DiagBox {
    int32_t on, xn, xp, yn, yp, zn, zp;
    int32_t ny, nz;
};

// Sums over the nodes n_lo <= n < n_hi (all nodes, or one reference rank's share of them for
// Error over ranks); the Error terms skip the range's last node.  The launch covers the range
// from the 64-aligned node below n_lo, so that every wavefront stays within one 64-node group of
// the packed neighbour table (load_nbr); lanes before n_lo only add zeros.
template <typename ST, int MODE>
__global__ void __launch_bounds__(BLOCK) sp_diag_kernel(SpArgs<ST> a, const int32_t *__restrict__ supinv,
                                                        DiagBox box, int64_t n_lo, int64_t n_hi,
                                                        double *__restrict__ partial)
{
    using C = typename Store<ST>::C;
    const int64_t n = group_floor(n_lo) + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double v[NDIAG + 1];
#pragma unroll
    for (int k = 0; k < NDIAG + 1; ++k) v[k] = 0.0;
    if (n >= n_lo && n < n_hi) {
        C h[Q], m[Q];
        Macro<C> mc;
        if (MODE == K_INIT) {
            mc = node_macro<ST, MODE>(a, n, m);
        } else {
            gather<ST, MODE>(a, n, h);
            if (MODE == K_PULL && a.bc.wave) bc_patch(a.bc, n, a.rho0, h);
            moments_fwd(h, m);
            mc = node_macro<ST, MODE>(a, n, m);
        }
        const double ux = (double)mc.ux, uy = (double)mc.uy, uz = (double)mc.uz;
        bool inside = true;
        if (box.on) {
            const int32_t s = supinv[n];
            const int32_t nyz = box.ny * box.nz;
            const int32_t i = s / nyz, j = (s % nyz) / box.nz, k = s % box.nz;
            inside = i >= box.xn && i <= box.xp && j >= box.yn && j <= box.yp && k >= box.zn && k <= box.zp;
        }
        if (inside) {
            v[0] = ux;
            v[1] = uy;
            v[2] = uz;
            v[8] = 1.0;
        }
        if (n < n_hi - 1) {   // the reference's Error loop stops before a rank's last node (SP:3651)
            const double dx = ux - (double)a.u_aux[n], dy = uy - (double)a.u_aux[a.P + n],
                         dz = uz - (double)a.u_aux[2 * a.P + n];
            const double s2 = ux * ux + uy * uy + uz * uz;
            v[3] = dx * dx + dy * dy + dz * dz;
            v[4] = s2;
            v[5] = sqrt(s2);
            v[6] = s2 > 0.0 ? s2 : 0.0;   // a NaN never passes the reference's "> u_max" (SP:3660)
        }
        v[7] = (isfinite(ux) && isfinite(uy) && isfinite(uz)) ? 0.0 : 1.0;
    }
    const bool is_max[NDIAG + 1] = {false, false, false, false, false, false, true, false, false};
    block_reduce_store<NDIAG + 1>(v, is_max, partial + (int64_t)blockIdx.x * (NDIAG + 1));
}

// Sums of u and node counts per cube of a cube grid (Comput_Grop_Perm, SP:4886-5090), one
// workgroup per cube.  A cube's nodes are listed as entries (64-node group, lane mask); a wave
// takes one entry at a time, so its active lanes always lie in one group (as load_nbr and
// bc_patch require).  Each thread sums its nodes in a fixed order and the block reduction has a
// fixed tree: the sums are the same on every call.
template <typename ST, int MODE>
__global__ void __launch_bounds__(BLOCK) sp_cube_kernel(SpArgs<ST> a, const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ grp,
                                                        const uint64_t *__restrict__ mask, double *__restrict__ out)
{
    using C = typename Store<ST>::C;
    const int64_t c = blockIdx.x;
    const uint32_t e0 = off[c], e1 = off[c + 1];
    const uint32_t lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t e = e0 + wid; e < e1; e += BLOCK / WAVE) {
        const uint64_t mk = mask[e];
        const int64_t n = (int64_t)grp[e] * WAVE + lane;
        if (!((mk >> lane) & 1ull)) continue;
        C h[Q], m[Q];
        Macro<C> mc;
        if (MODE == K_INIT) {
            mc = node_macro<ST, MODE>(a, n, m);
        } else {
            gather<ST, MODE>(a, n, h);
            if (MODE == K_PULL && a.bc.wave) bc_patch(a.bc, n, a.rho0, h);
            moments_fwd(h, m);
            mc = node_macro<ST, MODE>(a, n, m);
        }
        v[0] += (double)mc.ux;
        v[1] += (double)mc.uy;
        v[2] += (double)mc.uz;
        v[3] += 1.0;
    }
    const bool is_max[4] = {false, false, false, false};
    block_reduce_store<4>(v, is_max, out + c * 4);
}

// ---------------------------------------------------------------- state I/O
// Post-streaming f, u and rho of the nodes n0 <= n < n1.  Without a voxel map the node's values
// land at local index n - n0 of the outputs; with one (dense output) at supinv[n] - v0.  The
// launch starts at the 64-aligned node below n0 (see sp_diag_kernel).
template <typename ST, int MODE>
__global__ void __launch_bounds__(BLOCK) sp_get_state_kernel(SpArgs<ST> a, int64_t n0, int64_t n1,
                                                             const int32_t *__restrict__ supinv, int64_t v0,
                                                             double *__restrict__ f_out, double *__restrict__ u_out,
                                                             double *__restrict__ rho_out)
{
    using C = typename Store<ST>::C;
    const int64_t n = group_floor(n0) + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (n < n0 || n >= n1) return;
    const int64_t l = supinv ? (int64_t)supinv[n] - v0 : n - n0;
    C h[Q], m[Q];
    Macro<C> mc;
    if (MODE == K_INIT) {
        mc = node_macro<ST, MODE>(a, n, m);
        // f = feq(rho_i, u_i) written the way the reference's feq evaluates it (SP:1386-1396)
        const double ux = (double)mc.ux, uy = (double)mc.uy, uz = (double)mc.uz;
        const double uv = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double eu = EX[q] * ux + EY[q] * uy + EZ[q] * uz;
            if (f_out) f_out[l * Q + q] = weight(q) * (double)mc.rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uv);
        }
    } else {
        gather<ST, MODE>(a, n, h);
        if (MODE == K_PULL && a.bc.wave) bc_patch(a.bc, n, a.rho0, h);
        moments_fwd(h, m);
        mc = node_macro<ST, MODE>(a, n, m);
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (f_out) f_out[l * Q + q] = (double)h[q] + weight(q) * (double)a.rho0;
    }
    if (u_out) {
        u_out[l * 3 + 0] = (double)mc.ux;
        u_out[l * 3 + 1] = (double)mc.uy;
        u_out[l * 3 + 2] = (double)mc.uz;
    }
    if (rho_out) rho_out[l] = (double)mc.rho;
}

// a chunk of a post-streaming state, nodes n0 <= n < n1 at local index n - n0 of the inputs
template <typename ST>
__global__ void __launch_bounds__(BLOCK) sp_load_state_kernel(SpArgs<ST> a, int64_t n0, int64_t n1,
                                                              const double *__restrict__ f, const double *__restrict__ u,
                                                              const double *__restrict__ rho,
                                                              typename Store<ST>::C *__restrict__ rho_aux)
{
    const int64_t n = n0 + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (n >= n1) return;
    const int64_t l = n - n0;
#pragma unroll
    for (int q = 0; q < Q; ++q)
        a.dst[(int64_t)q * a.P + n] = Store<ST>::put((typename Store<ST>::C)(f[l * Q + q] - weight(q) * (double)a.rho0),
                                                    a.scale);
    for (int c = 0; c < 3; ++c) a.u_aux[(int64_t)c * a.P + n] = (typename Store<ST>::C)u[l * 3 + c];
    rho_aux[n] = (typename Store<ST>::C)rho[l];
}

// ---------------------------------------------------------------- host side
namespace {

template <typename ST>
SpArgs<ST> make_args(const void *src, void *dst, const NbrDev &nbr, void *u_aux, const void *rho_aux, int64_t N,
                     int64_t P, const lbm_sp_params &p, double sv, double sq, double rho_i, const double *u_i,
                     float scale, float inv_scale)
{
    using C = typename Store<ST>::C;
    SpArgs<ST> a;
    a.src = (const ST *)src;
    a.dst = (ST *)dst;
    a.nbr = nbr;
    a.u_aux = (C *)u_aux;
    a.rho_aux = (const C *)rho_aux;
    a.N = N;
    a.P = P;
    a.gx = (C)p.g[0];
    a.gy = (C)p.g[1];
    a.gz = (C)p.g[2];
    a.rho0 = (C)p.rho0;
    a.sv = (C)sv;
    a.sq = (C)sq;
    a.rho_i = (C)rho_i;
    a.ux_i = (C)u_i[0];
    a.uy_i = (C)u_i[1];
    a.uz_i = (C)u_i[2];
    a.scale = scale;
    a.inv_scale = inv_scale;
    a.bc = BcDev{nullptr, nullptr, nullptr, nullptr, 0};
    a.nu_vn = nullptr;
    a.pl_m = a.pl_n = C(0);
    a.st = StFused<C>{nullptr, nullptr, C(0), C(0), C(0), C(0)};
    a.order = ORDER_RUNS8;
    return a;
}

// the pull step without faces or power law, at a workgroup size NB
template <typename ST, int NB>
void launch_pull_nb(const SpArgs<ST> &a, bool store_u, int64_t N, hipStream_t s)
{
    const dim3 g((unsigned)((N + NB - 1) / NB)), b(NB);
    if (a.st.src)
        store_u ? sp_step_kernel<ST, K_PULL, true, false, false, true, NB><<<g, b, 0, s>>>(a)
                : sp_step_kernel<ST, K_PULL, false, false, false, true, NB><<<g, b, 0, s>>>(a);
    else
        store_u ? sp_step_kernel<ST, K_PULL, true, false, false, false, NB><<<g, b, 0, s>>>(a)
                : sp_step_kernel<ST, K_PULL, false, false, false, false, NB><<<g, b, 0, s>>>(a);
}

template <typename ST, bool BC>
void launch_step_bc(const SpArgs<ST> &a, int mode, bool store_u, dim3 g, dim3 b, hipStream_t s)
{
    if (mode == K_PULL && a.st.src) {
        if (a.nu_vn)
            store_u ? sp_step_kernel<ST, K_PULL, true, BC, true, true><<<g, b, 0, s>>>(a)
                    : sp_step_kernel<ST, K_PULL, false, BC, true, true><<<g, b, 0, s>>>(a);
        else
            store_u ? sp_step_kernel<ST, K_PULL, true, BC, false, true><<<g, b, 0, s>>>(a)
                    : sp_step_kernel<ST, K_PULL, false, BC, false, true><<<g, b, 0, s>>>(a);
    } else if (mode == K_PULL && a.nu_vn)
        store_u ? sp_step_kernel<ST, K_PULL, true, BC, true><<<g, b, 0, s>>>(a)
                : sp_step_kernel<ST, K_PULL, false, BC, true><<<g, b, 0, s>>>(a);
    else if (mode == K_INIT)
        store_u ? sp_step_kernel<ST, K_INIT, true, BC><<<g, b, 0, s>>>(a)
                : sp_step_kernel<ST, K_INIT, false, BC><<<g, b, 0, s>>>(a);
    else if (mode == K_DIRECT)
        sp_step_kernel<ST, K_DIRECT, false, BC><<<g, b, 0, s>>>(a);
    else
        store_u ? sp_step_kernel<ST, K_PULL, true, BC><<<g, b, 0, s>>>(a)
                : sp_step_kernel<ST, K_PULL, false, BC><<<g, b, 0, s>>>(a);
}

template <typename ST>
hipError_t launch_step_t(SpArgs<ST> a, const BcDev &bc, void *nu_vn, double pl_m, double pl_n, int mode, bool store_u,
                         int64_t grid, hipStream_t s, const FusedSolute *fs, const lbm_launch &ls)
{
    using C = typename Store<ST>::C;
    dim3 g((unsigned)grid), b(BLOCK);
    a.bc = bc;
    a.order = ls.order;
    a.nu_vn = (C *)nu_vn;
    a.pl_m = (C)pl_m;
    a.pl_n = (C)pl_n;
    if (fs) {
        a.st.src = (const C *)fs->src;
        a.st.dst = (C *)fs->dst;
        a.st.omega = (C)fs->omega;
        a.st.w0 = (C)fs->w0;
        a.st.wq = (C)fs->wq;
        a.st.inv_cs2 = (C)fs->inv_cs2;
    }
    if (mode == K_PULL && !bc.wave && !nu_vn && ls.block != BLOCK) {
        if (ls.block == 128)
            launch_pull_nb<ST, 128>(a, store_u, a.N, s);
        else
            launch_pull_nb<ST, 512>(a, store_u, a.N, s);
    } else if (bc.wave) {
        launch_step_bc<ST, true>(a, mode, store_u, g, b, s);
    } else {
        launch_step_bc<ST, false>(a, mode, store_u, g, b, s);
    }
    return hipGetLastError();
}

template <typename ST>
hipError_t launch_diag_t(SpArgs<ST> a, const BcDev &bc, int mode, const int32_t *supinv, DiagBox box, int64_t n_lo,
                         int64_t n_hi, double *partial, int64_t grid, hipStream_t s)
{
    dim3 g((unsigned)grid), b(BLOCK);
    a.bc = bc;
    if (mode == K_INIT)
        sp_diag_kernel<ST, K_INIT><<<g, b, 0, s>>>(a, supinv, box, n_lo, n_hi, partial);
    else if (mode == K_DIRECT)
        sp_diag_kernel<ST, K_DIRECT><<<g, b, 0, s>>>(a, supinv, box, n_lo, n_hi, partial);
    else
        sp_diag_kernel<ST, K_PULL><<<g, b, 0, s>>>(a, supinv, box, n_lo, n_hi, partial);
    return hipGetLastError();
}

template <typename ST>
hipError_t launch_cube_t(SpArgs<ST> a, const BcDev &bc, int mode, const uint32_t *off, const uint32_t *grp,
                         const uint64_t *mask, int64_t ncubes, double *out, hipStream_t s)
{
    dim3 g((unsigned)ncubes), b(BLOCK);
    a.bc = bc;
    if (mode == K_INIT)
        sp_cube_kernel<ST, K_INIT><<<g, b, 0, s>>>(a, off, grp, mask, out);
    else if (mode == K_DIRECT)
        sp_cube_kernel<ST, K_DIRECT><<<g, b, 0, s>>>(a, off, grp, mask, out);
    else
        sp_cube_kernel<ST, K_PULL><<<g, b, 0, s>>>(a, off, grp, mask, out);
    return hipGetLastError();
}

template <typename ST>
hipError_t launch_get_t(SpArgs<ST> a, const BcDev &bc, int mode, int64_t n0, int64_t n1, const int32_t *supinv,
                        int64_t v0, double *f, double *u, double *rho, hipStream_t s)
{
    if (n1 <= n0) return hipSuccess;
    dim3 g((unsigned)((n1 - group_floor(n0) + BLOCK - 1) / BLOCK)), b(BLOCK);
    a.bc = bc;
    if (mode == K_INIT)
        sp_get_state_kernel<ST, K_INIT><<<g, b, 0, s>>>(a, n0, n1, supinv, v0, f, u, rho);
    else if (mode == K_DIRECT)
        sp_get_state_kernel<ST, K_DIRECT><<<g, b, 0, s>>>(a, n0, n1, supinv, v0, f, u, rho);
    else
        sp_get_state_kernel<ST, K_PULL><<<g, b, 0, s>>>(a, n0, n1, supinv, v0, f, u, rho);
    return hipGetLastError();
}

int kmode_of(StateMode m) { return m == STATE_INIT ? K_INIT : (m == STATE_DIRECT ? K_DIRECT : K_PULL); }

}  // namespace

SpSolver::~SpSolver() { release(); }

int SpSolver::release()
{
    if (device_ >= 0) (void)hipSetDevice(device_);
    for (void *&p : pops_)
        if (p) (void)hipFree(p), p = nullptr;
    nbr_.release();
    for (CubeGrid &g : cube_grids_) free_cube_grid(g);
    cube_grids_.clear();
    if (d_supinv_) (void)hipFree(d_supinv_), d_supinv_ = nullptr;
    if (d_u_) (void)hipFree(d_u_), d_u_ = nullptr;
    if (d_rho_) (void)hipFree(d_rho_), d_rho_ = nullptr;
    if (d_partial_) (void)hipFree(d_partial_), d_partial_ = nullptr;
    if (d_result_) (void)hipFree(d_result_), d_result_ = nullptr;
    if (h_result_) (void)hipHostFree(h_result_), h_result_ = nullptr;
    if (ev0_) (void)hipEventDestroy(ev0_), ev0_ = nullptr;
    if (ev1_) (void)hipEventDestroy(ev1_), ev1_ = nullptr;
    release_boundaries();
    if (d_nu_vn_) (void)hipFree(d_nu_vn_), d_nu_vn_ = nullptr;
    stage_.release();
    if (stream_) (void)hipStreamDestroy(stream_), stream_ = nullptr;
    return LBM_OK;
}

int SpSolver::create(const lbm_sp_params &p, const int32_t *solid)
{
    prm_ = p;
    if (p.precision < LBM_FP64 || p.precision > LBM_FP16) return fail(LBM_ERR_ARG, "unknown precision");
    if (!(p.nu > 0.0)) return fail(LBM_ERR_ARG, "viscosity must be positive");
    if (prm_.rho0 == 0.0) prm_.rho0 = 1.0;
    std::string err;
    int rc = build_sparse_geometry(p.nx, p.ny, p.nz, solid, geo_, err);
    if (rc != LBM_OK) return fail(rc, err);

    // relaxation rates as the reference derives them (SP:1291-1307)
    const double c_s2 = 1.0 / 3.0;
    const double tau = p.nu / (c_s2 * 1.0) + 0.5;
    sv_ = 1.0 / tau;
    sq_ = 8.0 * (2.0 - sv_) / (8.0 - sv_);
    if (p.precision == LBM_FP16) {
        const double s = p.fp16_scale > 0 ? p.fp16_scale : 4096.0;
        scale_ = (float)s;
        inv_scale_ = (float)(1.0 / s);
    }
    st_bytes_ = p.precision == LBM_FP64 ? 8 : (p.precision == LBM_FP32 ? 4 : 2);
    ct_bytes_ = p.precision == LBM_FP64 ? 8 : 4;

    device_ = p.device;
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    LBM_HIP_TRY(hipEventCreate(&ev0_));
    LBM_HIP_TRY(hipEventCreate(&ev1_));
    const int64_t P = geo_.P;
    grid_ = (geo_.N + BLOCK - 1) / BLOCK;
    for (void *&b : pops_) LBM_HIP_TRY(hipMalloc(&b, (size_t)Q * P * st_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_supinv_, (size_t)P * sizeof(int32_t)));
    LBM_HIP_TRY(hipMalloc(&d_u_, (size_t)3 * P * ct_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_rho_, (size_t)P * ct_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_partial_, (size_t)grid_ * (NDIAG + 1) * sizeof(double)));
    LBM_HIP_TRY(hipMalloc(&d_result_, 32 * sizeof(double)));
    LBM_HIP_TRY(hipHostMalloc((void **)&h_result_, 32 * sizeof(double), hipHostMallocDefault));
    PackedNbr packed;
    rc = pack_neighbours(geo_, packed, err);
    if (rc != LBM_OK) return fail(rc, err);
    LBM_HIP_TRY(nbr_.upload(packed, P, stream_));
    LBM_HIP_TRY(hipMemcpyAsync(d_supinv_, geo_.supinv.data(), geo_.supinv.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, stream_));
    LBM_HIP_TRY(hipMemsetAsync(d_u_, 0, (size_t)3 * P * ct_bytes_, stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    // the host copies of the neighbour table are no longer needed
    std::vector<uint32_t>().swap(geo_.nbr);
    const double zero[3] = {0, 0, 0};
    rc = init(1.0, zero);
    if (rc != LBM_OK) return rc;
    rc = autotune();
    if (rc != LBM_OK) return rc;
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    return LBM_OK;
}

// Time the pull step's block sizes and tile orders on this geometry (on the post-collision state
// of one init launch written to the spare buffer; the handle's state is left as it was).
int SpSolver::autotune()
{
    if (geo_.N < TUNE_MIN_NODES) return LBM_OK;
    lbm_launch c;
    if (tune_cache_get(LBM_KERNEL_SP_STEP, prm_.precision, geo_.N, c)) {
        launch_ = c;
        return LBM_OK;
    }
    void *state = pops_[cur_], *spare = pops_[cur_ ^ 1];
    LBM_HIP_TRY(launch_kernel(K_INIT, false, nullptr, launch_, state, spare, true));
    const auto cands = tune_candidates(launch_, {128, 256, 512}, {ORDER_HW, ORDER_RUNS8, ORDER_SLAB}, {0});
    lbm_launch best;
    LBM_HIP_TRY(tune_launch(stream_, cands,
                            [&](const lbm_launch &ls) { return launch_kernel(K_PULL, false, nullptr, ls, spare, state, true); },
                            best));
    launch_ = best;
    tune_cache_put(LBM_KERNEL_SP_STEP, prm_.precision, geo_.N, best);
    return LBM_OK;
}

int SpSolver::get_launch(int kernel, lbm_launch *out) const
{
    if (kernel != LBM_KERNEL_SP_STEP || !out) return LBM_ERR_ARG;
    *out = launch_;
    return LBM_OK;
}

int SpSolver::set_launch(int kernel, const lbm_launch *in)
{
    if (kernel != LBM_KERNEL_SP_STEP) return fail(LBM_ERR_ARG, "a single-phase handle launches LBM_KERNEL_SP_STEP only");
    if (!in || !launch_valid(*in, {128, 256, 512}, {0}))
        return fail(LBM_ERR_ARG, "launch settings: block 128, 256 or 512, order 0-2, variant 0");
    launch_ = *in;
    launch_.tuned = 0;
    launch_.time_us = 0.0;
    return LBM_OK;
}

int SpSolver::init(double rho, const double u0[3])
{
    if (!u0) return fail(LBM_ERR_ARG, "null u0");
    rho_i_ = rho;
    for (int c = 0; c < 3; ++c) u_i_[c] = u0[c];
    mode_ = STATE_INIT;
    steps_ = 0;
    return reset_viscosity();
}

#define SP_DISPATCH(EXPR_DOUBLE, EXPR_FLOAT, EXPR_HALF)                          \
    (prm_.precision == LBM_FP64 ? (EXPR_DOUBLE) : prm_.precision == LBM_FP32 ? (EXPR_FLOAT) : (EXPR_HALF))

int SpSolver::load_state(const double *f, const double *u, const double *rho)
{
    if (!f || !u || !rho) return fail(LBM_ERR_ARG, "null state array");
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(stage_.ensure());
    const int64_t N = geo_.N, chunk = stage_chunk(STAGE_BYTES, Q + 4);
    const int nxt = cur_;   // DIRECT state lives in the current buffer
    const double zero[3] = {0, 0, 0};
    hipError_t e = hipSuccess;
    for (int64_t n0 = 0; n0 < N && e == hipSuccess; n0 += chunk) {
        const int64_t c = std::min(chunk, N - n0);
        double *df = (double *)stage_.buf, *du = df + c * Q, *dr = du + c * 3;
        e = hipMemcpyAsync(df, f + n0 * Q, (size_t)c * Q * sizeof(double), hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess) e = hipMemcpyAsync(du, u + n0 * 3, (size_t)c * 3 * sizeof(double), hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess) e = hipMemcpyAsync(dr, rho + n0, (size_t)c * sizeof(double), hipMemcpyHostToDevice, stream_);
        if (e != hipSuccess) break;
        dim3 g((unsigned)((c + BLOCK - 1) / BLOCK)), b(BLOCK);
        if (prm_.precision == LBM_FP64) {
            auto a = make_args<double>(nullptr, pops_[nxt], nbr_.view(), d_u_, d_rho_, N, geo_.P, prm_, sv_, sq_, 1.0, zero,
                                       scale_, inv_scale_);
            sp_load_state_kernel<double><<<g, b, 0, stream_>>>(a, n0, n0 + c, df, du, dr, (double *)d_rho_);
        } else if (prm_.precision == LBM_FP32) {
            auto a = make_args<float>(nullptr, pops_[nxt], nbr_.view(), d_u_, d_rho_, N, geo_.P, prm_, sv_, sq_, 1.0, zero,
                                      scale_, inv_scale_);
            sp_load_state_kernel<float><<<g, b, 0, stream_>>>(a, n0, n0 + c, df, du, dr, (float *)d_rho_);
        } else {
            auto a = make_args<__half>(nullptr, pops_[nxt], nbr_.view(), d_u_, d_rho_, N, geo_.P, prm_, sv_, sq_, 1.0, zero,
                                       scale_, inv_scale_);
            sp_load_state_kernel<__half><<<g, b, 0, stream_>>>(a, n0, n0 + c, df, du, dr, (float *)d_rho_);
        }
        e = hipGetLastError();
    }
    // the host arrays are borrowed for this call only: wait for the copies either way
    const hipError_t es = hipStreamSynchronize(stream_);
    LBM_HIP_TRY(e);
    LBM_HIP_TRY(es);
    mode_ = STATE_DIRECT;
    steps_ = 0;
    // the reference's restart leaves niu_vn unset (it is set in init, SP:1355); here it starts
    // from the Newtonian viscosity as after init
    return reset_viscosity();
}

// one step launch from buffer src to buffer dst; faces and the power law as the handle has them
// unless bare (the autotuner's timing launches: the plain pull step, state untouched)
hipError_t SpSolver::launch_kernel(int mode, bool store_u, const FusedSolute *fs, const lbm_launch &ls, const void *src,
                                   void *dst, bool bare)
{
    // the power-law collision acts from the reference's loop step n = 11 on (SP:2397); the
    // launches since init or load_state are that loop's n
    void *pl_nu = (!bare && d_nu_vn_ && steps_ > 10) ? d_nu_vn_ : nullptr;
    const BcDev bc = bare ? BcDev{nullptr, nullptr, nullptr, nullptr, 0} : bc_view(true);
    if (prm_.precision == LBM_FP64)
        return launch_step_t(make_args<double>(src, dst, nbr_.view(), d_u_, d_rho_, geo_.N, geo_.P, prm_, sv_, sq_,
                                               rho_i_, u_i_, scale_, inv_scale_),
                             bc, pl_nu, pl_m_, pl_n_, mode, store_u, grid_, stream_, fs, ls);
    if (prm_.precision == LBM_FP32)
        return launch_step_t(make_args<float>(src, dst, nbr_.view(), d_u_, d_rho_, geo_.N, geo_.P, prm_, sv_, sq_, rho_i_,
                                              u_i_, scale_, inv_scale_),
                             bc, pl_nu, pl_m_, pl_n_, mode, store_u, grid_, stream_, fs, ls);
    return launch_step_t(make_args<__half>(src, dst, nbr_.view(), d_u_, d_rho_, geo_.N, geo_.P, prm_, sv_, sq_, rho_i_,
                                           u_i_, scale_, inv_scale_),
                         bc, pl_nu, pl_m_, pl_n_, mode, store_u, grid_, stream_, fs, ls);
}

int SpSolver::launch_step(bool store_u, const FusedSolute *fs)
{
    const int mode = kmode_of(mode_);
    const int nxt = cur_ ^ 1;
    LBM_HIP_TRY(launch_kernel(mode, store_u, fs, fs ? fs->ls : launch_, pops_[cur_], pops_[nxt], false));
    cur_ = nxt;
    if (d_bc_wave_) bc_cur_ ^= 1;
    mode_ = STATE_PULL;
    ++steps_;
    return LBM_OK;
}

BcDev SpSolver::bc_view(bool for_step) const
{
    if (!d_bc_wave_) return BcDev{nullptr, nullptr, nullptr, nullptr, 0};
    // a launch reads the buffer the previous one wrote; state reads use the same one
    return BcDev{d_bc_wave_, d_bc_entry_, d_bc_s_[bc_cur_], for_step ? d_bc_s_[bc_cur_ ^ 1] : nullptr, bc_sp_};
}

void SpSolver::release_boundaries()
{
    if (d_bc_wave_) (void)hipFree(d_bc_wave_), d_bc_wave_ = nullptr;
    if (d_bc_entry_) (void)hipFree(d_bc_entry_), d_bc_entry_ = nullptr;
    for (void *&b : d_bc_s_)
        if (b) (void)hipFree(b), b = nullptr;
    bc_sp_ = 0;
    bc_cur_ = 0;
}

int SpSolver::set_boundaries(const lbm_sp_bc *bc)
{
    if (mode_ == STATE_PULL)
        return fail(LBM_ERR_STATE, "set boundaries before stepping: after create, init or load_state");
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    release_boundaries();
    if (!bc) return LBM_OK;
    BoundaryTables t;
    std::string err;
    const int rc = build_boundaries(geo_, *bc, t, err);
    if (rc != LBM_OK) return fail(rc, err);
    if (t.entry.empty()) return LBM_OK;   // no fluid node on any active face
    bc_sp_ = std::max<int64_t>(1, (int64_t)((t.saved.size() + 63) / 64 * 64));
    LBM_HIP_TRY(hipMalloc((void **)&d_bc_wave_, t.wave.size() * sizeof(BcWave)));
    LBM_HIP_TRY(hipMalloc((void **)&d_bc_entry_, t.entry.size() * sizeof(BcEntry)));
    for (void *&b : d_bc_s_) {
        LBM_HIP_TRY(hipMalloc(&b, (size_t)BC_ROWS * bc_sp_ * ct_bytes_));
        LBM_HIP_TRY(hipMemsetAsync(b, 0, (size_t)BC_ROWS * bc_sp_ * ct_bytes_, stream_));
    }
    LBM_HIP_TRY(hipMemcpyAsync(d_bc_wave_, t.wave.data(), t.wave.size() * sizeof(BcWave), hipMemcpyHostToDevice, stream_));
    LBM_HIP_TRY(hipMemcpyAsync(d_bc_entry_, t.entry.data(), t.entry.size() * sizeof(BcEntry), hipMemcpyHostToDevice,
                               stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    return LBM_OK;
}

int SpSolver::reset_viscosity()
{
    if (!d_nu_vn_) return LBM_OK;
    const dim3 g((unsigned)grid_), b(BLOCK);
    if (ct_bytes_ == 8)
        fill_kernel<double><<<g, b, 0, stream_>>>((double *)d_nu_vn_, prm_.nu, geo_.P);
    else
        fill_kernel<float><<<g, b, 0, stream_>>>((float *)d_nu_vn_, (float)prm_.nu, geo_.P);
    LBM_HIP_TRY(hipGetLastError());
    return LBM_OK;
}

int SpSolver::set_power_law(double m, double n)
{
    if (mode_ == STATE_PULL)
        return fail(LBM_ERR_STATE, "set the power law before stepping: after create, init or load_state");
    LBM_HIP_TRY(hipSetDevice(device_));
    if (!(m > 0.0)) {   // back to the Newtonian collision
        LBM_HIP_TRY(hipStreamSynchronize(stream_));
        if (d_nu_vn_) (void)hipFree(d_nu_vn_), d_nu_vn_ = nullptr;
        pl_m_ = pl_n_ = 0.0;
        return LBM_OK;
    }
    if (!std::isfinite(n)) return fail(LBM_ERR_ARG, "power-law index must be finite");
    if (!d_nu_vn_) LBM_HIP_TRY(hipMalloc(&d_nu_vn_, (size_t)geo_.P * ct_bytes_));
    pl_m_ = m;
    pl_n_ = n;
    const int rc = reset_viscosity();   // niu_vn = niu (SP:1355)
    if (rc != LBM_OK) return rc;
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    return LBM_OK;
}

int SpSolver::get_viscosity(double *nu)
{
    if (!nu) return fail(LBM_ERR_ARG, "null output");
    LBM_HIP_TRY(hipSetDevice(device_));
    const int64_t N = geo_.N;
    if (!d_nu_vn_) {
        for (int64_t i = 0; i < N; ++i) nu[i] = prm_.nu;
        return LBM_OK;
    }
    if (ct_bytes_ == 8) {
        LBM_HIP_TRY(hipMemcpyAsync(nu, d_nu_vn_, (size_t)N * 8, hipMemcpyDeviceToHost, stream_));
        LBM_HIP_TRY(hipStreamSynchronize(stream_));
    } else {
        std::vector<float> t((size_t)N);
        LBM_HIP_TRY(hipMemcpyAsync(t.data(), d_nu_vn_, (size_t)N * 4, hipMemcpyDeviceToHost, stream_));
        LBM_HIP_TRY(hipStreamSynchronize(stream_));
        for (int64_t i = 0; i < N; ++i) nu[i] = t[(size_t)i];
    }
    return LBM_OK;
}

int SpSolver::step_storing_velocity()
{
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    LBM_HIP_TRY(hipSetDevice(device_));
    return launch_step(true);
}

int SpSolver::step_fused(const FusedSolute &fs, bool store_u)
{
    if (mode_ != STATE_PULL) return fail(LBM_ERR_STATE, "the fused step needs a flow state in pull form");
    LBM_HIP_TRY(hipSetDevice(device_));
    return launch_step(store_u, &fs);
}

int SpSolver::step(int nsteps, double *elapsed_ms)
{
    if (nsteps < 0) return fail(LBM_ERR_ARG, "negative step count");
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    LBM_HIP_TRY(hipSetDevice(device_));
    if (elapsed_ms) LBM_HIP_TRY(hipEventRecord(ev0_, stream_));
    for (int s = 0; s < nsteps; ++s) {
        // the last launch of a call keeps u of the step before it, which the Error
        // diagnostic compares against (u0 in SP:3625-3704)
        int rc = launch_step(s == nsteps - 1);
        if (rc != LBM_OK) return rc;
    }
    if (elapsed_ms) {
        LBM_HIP_TRY(hipEventRecord(ev1_, stream_));
        LBM_HIP_TRY(hipEventSynchronize(ev1_));
        float ms = 0.f;
        LBM_HIP_TRY(hipEventElapsedTime(&ms, ev0_, ev1_));
        *elapsed_ms = ms;
    }
    return LBM_OK;
}

// post-streaming f (AoS [.][19]), u ([.][3]) and rho of nodes n0 <= n < n1 into device buffers
// (any may be null), at local index n - n0, or at supinv[n] - v0 of a dense voxel range
hipError_t SpSolver::state_to_device(int64_t n0, int64_t n1, const int32_t *supinv, int64_t v0, double *df, double *du,
                                     double *dr)
{
    const int64_t N = geo_.N;
    const int mode = kmode_of(mode_);
    if (prm_.precision == LBM_FP64)
        return launch_get_t(make_args<double>(pops_[cur_], nullptr, nbr_.view(), d_u_, d_rho_, N, geo_.P, prm_, sv_, sq_,
                                              rho_i_, u_i_, scale_, inv_scale_),
                            bc_view(false), mode, n0, n1, supinv, v0, df, du, dr, stream_);
    if (prm_.precision == LBM_FP32)
        return launch_get_t(make_args<float>(pops_[cur_], nullptr, nbr_.view(), d_u_, d_rho_, N, geo_.P, prm_, sv_, sq_,
                                             rho_i_, u_i_, scale_, inv_scale_),
                            bc_view(false), mode, n0, n1, supinv, v0, df, du, dr, stream_);
    return launch_get_t(make_args<__half>(pops_[cur_], nullptr, nbr_.view(), d_u_, d_rho_, N, geo_.P, prm_, sv_, sq_,
                                          rho_i_, u_i_, scale_, inv_scale_),
                        bc_view(false), mode, n0, n1, supinv, v0, df, du, dr, stream_);
}

int SpSolver::get_state(double *f, double *u, double *rho)
{
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(stage_.ensure());
    return get_state_via(f, u, rho, stage_.buf, STAGE_BYTES);
}

// the state crosses through buf in chunks of whole blocks of nodes; each chunk's copies are
// ordered on the stream before the next chunk's kernel overwrites buf
int SpSolver::get_state_via(double *f, double *u, double *rho, void *buf, size_t bytes)
{
    const int per_node = (f ? Q : 0) + (u ? 3 : 0) + (rho ? 1 : 0);
    if (per_node == 0) return LBM_OK;
    if (!buf || bytes < (size_t)BLOCK * per_node * sizeof(double))
        return fail(LBM_ERR_ARG, "state transfer buffer smaller than one block of nodes");
    const int64_t N = geo_.N, chunk = stage_chunk(bytes, per_node);
    hipError_t e = hipSuccess;
    for (int64_t n0 = 0; n0 < N && e == hipSuccess; n0 += chunk) {
        const int64_t c = std::min(chunk, N - n0);
        double *df = f ? (double *)buf : nullptr;
        double *du = u ? (double *)buf + (f ? c * Q : 0) : nullptr;
        double *dr = rho ? (double *)buf + (f ? c * Q : 0) + (u ? c * 3 : 0) : nullptr;
        e = state_to_device(n0, n0 + c, nullptr, 0, df, du, dr);
        if (e == hipSuccess && f) e = hipMemcpyAsync(f + n0 * Q, df, (size_t)c * Q * sizeof(double), hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess && u) e = hipMemcpyAsync(u + n0 * 3, du, (size_t)c * 3 * sizeof(double), hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess && rho) e = hipMemcpyAsync(rho + n0, dr, (size_t)c * sizeof(double), hipMemcpyDeviceToHost, stream_);
    }
    const hipError_t es = hipStreamSynchronize(stream_);
    LBM_HIP_TRY(e);
    LBM_HIP_TRY(es);
    return LBM_OK;
}

int SpSolver::get_dense(double *u_ijk, double *rho_ijk)
{
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    if (!u_ijk && !rho_ijk) return LBM_OK;
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(stage_.ensure());
    // the box in chunks of consecutive voxels; the fluid nodes of a voxel range are a range of
    // node ids, since the numbering is i-major like the voxel index
    const int64_t N = geo_.N, V = (int64_t)geo_.nx * geo_.ny * geo_.nz;
    const int64_t chunk = stage_chunk(STAGE_BYTES, (u_ijk ? 3 : 0) + (rho_ijk ? 1 : 0));
    const std::vector<int32_t> &sup = geo_.supinv;
    hipError_t e = hipSuccess;
    for (int64_t v0 = 0; v0 < V && e == hipSuccess; v0 += chunk) {
        const int64_t c = std::min(chunk, V - v0);
        double *gu = u_ijk ? (double *)stage_.buf : nullptr;
        double *gr = rho_ijk ? (double *)stage_.buf + (u_ijk ? c * 3 : 0) : nullptr;
        // solid voxels keep the values the reference's writers put there: u = 0, rho = 1
        const dim3 b(BLOCK);
        if (gu) {
            fill_kernel<double><<<dim3((unsigned)((c * 3 + BLOCK - 1) / BLOCK)), b, 0, stream_>>>(gu, 0.0, c * 3);
            e = hipGetLastError();
        }
        if (e == hipSuccess && gr) {
            fill_kernel<double><<<dim3((unsigned)((c + BLOCK - 1) / BLOCK)), b, 0, stream_>>>(gr, 1.0, c);
            e = hipGetLastError();
        }
        const int64_t n0 = std::lower_bound(sup.begin(), sup.end(), v0) - sup.begin();
        const int64_t n1 = std::lower_bound(sup.begin(), sup.end(), (int64_t)(v0 + c)) - sup.begin();
        if (e == hipSuccess) e = state_to_device(n0, std::min(n1, N), d_supinv_, v0, nullptr, gu, gr);
        if (e == hipSuccess && gu)
            e = hipMemcpyAsync(u_ijk + v0 * 3, gu, (size_t)c * 3 * sizeof(double), hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess && gr)
            e = hipMemcpyAsync(rho_ijk + v0, gr, (size_t)c * sizeof(double), hipMemcpyDeviceToHost, stream_);
    }
    const hipError_t es = hipStreamSynchronize(stream_);
    LBM_HIP_TRY(e);
    LBM_HIP_TRY(es);
    return LBM_OK;
}

// the diagnostic sums over nodes [n_lo, n_hi) into h_result_ (NDIAG + 1 values)
int SpSolver::reduce_range(const DiagBox &db, int64_t n_lo, int64_t n_hi)
{
    const int mode = kmode_of(mode_);
    const int64_t grid = std::max<int64_t>(1, (n_hi - group_floor(n_lo) + BLOCK - 1) / BLOCK);
    hipError_t e;
    if (prm_.precision == LBM_FP64)
        e = launch_diag_t(make_args<double>(pops_[cur_], nullptr, nbr_.view(), d_u_, d_rho_, geo_.N, geo_.P, prm_, sv_, sq_,
                                            rho_i_, u_i_, scale_, inv_scale_),
                          bc_view(false), mode, d_supinv_, db, n_lo, n_hi, d_partial_, grid, stream_);
    else if (prm_.precision == LBM_FP32)
        e = launch_diag_t(make_args<float>(pops_[cur_], nullptr, nbr_.view(), d_u_, d_rho_, geo_.N, geo_.P, prm_, sv_, sq_,
                                           rho_i_, u_i_, scale_, inv_scale_),
                          bc_view(false), mode, d_supinv_, db, n_lo, n_hi, d_partial_, grid, stream_);
    else
        e = launch_diag_t(make_args<__half>(pops_[cur_], nullptr, nbr_.view(), d_u_, d_rho_, geo_.N, geo_.P, prm_, sv_, sq_,
                                            rho_i_, u_i_, scale_, inv_scale_),
                          bc_view(false), mode, d_supinv_, db, n_lo, n_hi, d_partial_, grid, stream_);
    LBM_HIP_TRY(e);
    reduce_partials_kernel<NDIAG + 1><<<1, BLOCK, 0, stream_>>>(d_partial_, grid, 1u << 6, d_result_);
    LBM_HIP_TRY(hipGetLastError());
    LBM_HIP_TRY(hipMemcpyAsync(h_result_, d_result_, (NDIAG + 1) * sizeof(double), hipMemcpyDeviceToHost, stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    return LBM_OK;
}

int SpSolver::diagnostics(const int32_t *box, lbm_sp_diag *out)
{
    if (!out) return fail(LBM_ERR_ARG, "null output");
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    LBM_HIP_TRY(hipSetDevice(device_));
    DiagBox db{};
    db.ny = geo_.ny;
    db.nz = geo_.nz;
    if (box) {
        db.on = 1;
        db.xn = box[0];
        db.xp = box[1];
        db.yn = box[2];
        db.yp = box[3];
        db.zn = box[4];
        db.zp = box[5];
    }
    const int rc = reduce_range(db, 0, geo_.N);
    if (rc != LBM_OK) return rc;
    const double *r = h_result_;
    out->Q[0] = r[0];
    out->Q[1] = r[1];
    out->Q[2] = r[2];
    out->err = std::sqrt(r[3]) / (std::sqrt(r[4]) + 1e-30);
    out->umax = std::sqrt(r[6]);
    out->sum_speed = r[5];
    out->nan_count = r[7];
    out->count_in_box = r[8];
    return LBM_OK;
}

int SpSolver::error_ranks(int nranks, double out[3])
{
    if (!out) return fail(LBM_ERR_ARG, "null output");
    if (nranks < 1) return fail(LBM_ERR_ARG, "rank count must be at least 1");
    if (nranks > geo_.nx) return fail(LBM_ERR_ARG, "more ranks than x-planes");
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    LBM_HIP_TRY(hipSetDevice(device_));
    std::vector<int64_t> ends;
    reference_rank_ends(geo_, nranks, ends);
    DiagBox db{};
    db.ny = geo_.ny;
    db.nz = geo_.nz;
    // per rank: its own err and u_max, then the largest of each and the sum of the speed sums
    // (SP:3666-3685)
    double err = 0, umax = 0, speed = 0;
    int64_t lo = 0;
    for (int r = 0; r < nranks; ++r) {
        const int64_t hi = ends[(size_t)r];
        const int rc = reduce_range(db, lo, hi);
        if (rc != LBM_OK) return rc;
        const double e = std::sqrt(h_result_[3]) / (std::sqrt(h_result_[4]) + 1e-30);
        const double um = std::sqrt(h_result_[6]);
        if (r == 0 || e > err) err = e;
        if (um > umax) umax = um;
        speed += h_result_[5];
        lo = hi;
    }
    out[0] = err;
    out[1] = umax;
    out[2] = speed;
    return LBM_OK;
}

void SpSolver::free_cube_grid(CubeGrid &g)
{
    for (void *p : {(void *)g.off, (void *)g.grp, (void *)g.mask, (void *)g.sums})
        if (p) (void)hipFree(p);
    g = CubeGrid{};
}

// The entries of every cube of the grid: for each (i, j) line of a cube (x, then y, ascending)
// the nodes with k in the cube's z range are one run of consecutive ids (found by binary search
// in supinv, which is ascending), so a cube's nodes come in ascending id order and a group shared
// by two of its runs is the last entry when the second run reaches it.
int SpSolver::build_cube_grid(const int32_t origin[3], int32_t size, int32_t m, CubeGrid &g)
{
    const int64_t nx = geo_.nx, ny = geo_.ny, nz = geo_.nz;
    const std::vector<int32_t> &sup = geo_.supinv;
    const int64_t ncubes = (int64_t)m * m * m;
    std::vector<uint32_t> off((size_t)ncubes + 1, 0u), grp;
    std::vector<uint64_t> mask;
    auto first_at_or_after = [&](int64_t voxel) {
        return (int64_t)(std::lower_bound(sup.begin(), sup.begin() + geo_.N, voxel) - sup.begin());
    };
    int64_t c = 0;
    for (int64_t ix = 0; ix < m; ++ix)
        for (int64_t iy = 0; iy < m; ++iy)
            for (int64_t iz = 0; iz < m; ++iz, ++c) {
                const int64_t x0 = origin[0] + ix * size, y0 = origin[1] + iy * size, z0 = origin[2] + iz * size;
                const int64_t zlo = std::max<int64_t>(z0, 0), zhi = std::min<int64_t>(z0 + size - 1, nz - 1);
                const size_t first = grp.size();
                if (zlo <= zhi)
                    for (int64_t i = std::max<int64_t>(x0, 0); i <= std::min<int64_t>(x0 + size - 1, nx - 1); ++i)
                        for (int64_t j = std::max<int64_t>(y0, 0); j <= std::min<int64_t>(y0 + size - 1, ny - 1); ++j) {
                            const int64_t base = (i * ny + j) * nz;
                            const int64_t lo = first_at_or_after(base + zlo), hi = first_at_or_after(base + zhi + 1);
                            for (int64_t n = lo; n < hi; ++n) {
                                const uint32_t w = (uint32_t)(n >> 6);
                                if (grp.size() == first || grp.back() != w) {
                                    grp.push_back(w);
                                    mask.push_back(0ull);
                                }
                                mask.back() |= 1ull << (n & 63);
                            }
                        }
                if (grp.size() > 0xFFFFFFF0ull) return fail(LBM_ERR_ARG, "cube grid: too many node groups");
                off[(size_t)c + 1] = (uint32_t)grp.size();
            }
    const size_t ne = std::max<size_t>(grp.size(), 1);
    grp.resize(ne, 0u);
    mask.resize(ne, 0ull);
    LBM_HIP_TRY(hipMalloc(&g.off, off.size() * sizeof(uint32_t)));
    LBM_HIP_TRY(hipMalloc(&g.grp, ne * sizeof(uint32_t)));
    LBM_HIP_TRY(hipMalloc(&g.mask, ne * sizeof(uint64_t)));
    LBM_HIP_TRY(hipMalloc(&g.sums, (size_t)ncubes * 4 * sizeof(double)));
    LBM_HIP_TRY(hipMemcpyAsync(g.off, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
    LBM_HIP_TRY(hipMemcpyAsync(g.grp, grp.data(), ne * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
    LBM_HIP_TRY(hipMemcpyAsync(g.mask, mask.data(), ne * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));   // the host vectors go out of scope
    g.ncubes = ncubes;
    return LBM_OK;
}

int SpSolver::cube_sums(const int32_t origin[3], int32_t size, int32_t m, double *out)
{
    if (!origin || !out) return fail(LBM_ERR_ARG, "null argument");
    if (size < 1 || m < 1 || m > 256) return fail(LBM_ERR_ARG, "cube grid: size >= 1 and 1 <= m <= 256");
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    LBM_HIP_TRY(hipSetDevice(device_));
    const int32_t key[5] = {origin[0], origin[1], origin[2], size, m};
    CubeGrid *g = nullptr;
    for (CubeGrid &cg : cube_grids_)
        if (std::equal(key, key + 5, cg.key)) g = &cg;
    if (!g) {
        if (cube_grids_.size() < 4) {
            cube_grids_.emplace_back();
            g = &cube_grids_.back();
        } else {
            g = &*std::min_element(cube_grids_.begin(), cube_grids_.end(),
                                   [](const CubeGrid &a, const CubeGrid &b) { return a.used < b.used; });
            free_cube_grid(*g);
        }
        const int rc = build_cube_grid(origin, size, m, *g);
        if (rc != LBM_OK) {
            free_cube_grid(*g);
            return rc;
        }
        std::copy(key, key + 5, g->key);
    }
    g->used = ++cube_clock_;
    const int mode = kmode_of(mode_);
    hipError_t e;
    if (prm_.precision == LBM_FP64)
        e = launch_cube_t(make_args<double>(pops_[cur_], nullptr, nbr_.view(), d_u_, d_rho_, geo_.N, geo_.P, prm_, sv_, sq_,
                                            rho_i_, u_i_, scale_, inv_scale_),
                          bc_view(false), mode, g->off, g->grp, g->mask, g->ncubes, g->sums, stream_);
    else if (prm_.precision == LBM_FP32)
        e = launch_cube_t(make_args<float>(pops_[cur_], nullptr, nbr_.view(), d_u_, d_rho_, geo_.N, geo_.P, prm_, sv_, sq_,
                                           rho_i_, u_i_, scale_, inv_scale_),
                          bc_view(false), mode, g->off, g->grp, g->mask, g->ncubes, g->sums, stream_);
    else
        e = launch_cube_t(make_args<__half>(pops_[cur_], nullptr, nbr_.view(), d_u_, d_rho_, geo_.N, geo_.P, prm_, sv_, sq_,
                                            rho_i_, u_i_, scale_, inv_scale_),
                          bc_view(false), mode, g->off, g->grp, g->mask, g->ncubes, g->sums, stream_);
    LBM_HIP_TRY(e);
    LBM_HIP_TRY(hipMemcpyAsync(out, g->sums, (size_t)g->ncubes * 4 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    return LBM_OK;
}

// ---------------------------------------------------------------- state read without allocation
// lbm_sp_get_state_noalloc stages the state through this fixed array of the code object instead
// of a buffer of the handle.  Every GPU the library runs on holds its own copy until the process
// ends.  In host code the array's name is not a device address: the address of each device's copy
// is taken once per device with hipGetSymbolAddress (with that device current) and only that
// address reaches kernels and copies.  Calls on one GPU share its copy: calls that overlap in
// time on one GPU race on it (a single-threaded caller, one call at a time per GPU, is safe).
}  // namespace lbm

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace lbm {
namespace {
constexpr int MAX_SCRATCH_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_ptr[MAX_SCRATCH_DEVICES] = {nullptr};
}  // namespace

int SpSolver::get_state_noalloc(double *f, double *u, double *rho)
{
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    if (device_ < 0 || device_ >= MAX_SCRATCH_DEVICES)
        return fail(LBM_ERR_ARG, "lbm_sp_get_state_noalloc: device ordinal beyond the 64 the library tracks");
    LBM_HIP_TRY(hipSetDevice(device_));
    void *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(scratch_mu);
        if (!scratch_ptr[device_]) {
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return fail(LBM_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
            scratch_ptr[device_] = p;
        }
        ptr = scratch_ptr[device_];
    }
    // the state streams through the array in chunks (91,136 nodes of f, u and rho per chunk), so
    // no node count is too large for it; nothing in it needs clearing, each chunk is written
    // whole before it is copied out
    return get_state_via(f, u, rho, ptr, NOALLOC_SCRATCH_BYTES);
}

}  // namespace lbm
