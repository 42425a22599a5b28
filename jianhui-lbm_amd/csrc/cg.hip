// cg.hip -- colour-gradient two-component sparse D3Q19 MRT (the reference's CG solver).
//
// One reference time step (collision CG:2687-3660 + comput_macro_variables CG:3698-4099)
// becomes two launches:
//
//  cg_collide_kernel  pull f from the upstream nodes, colour gradient C from psi at the 18
//                     neighbour sites (psi_solid at solid sites), MRT collision in the CG
//                     basis with the surface-tension equilibrium and the psi-interpolated
//                     viscosity, post-collision f written coalesced; plus an 8-value colour
//                     state per node (rho_r, rho_b, u, C/|C|) for the colour streaming.
//  cg_colour_kernel   colour streaming in pull form: rho_r(x) = sum_q g_r,q(x - e_q), or
//                     g_r,opp(q)(x) when x - e_q is solid, with each upstream g recomputed
//                     (equilibrium split + recolouring) from that node's colour state.  The
//                     reference pushes and accumulates (rhor[nbr] += g_r, CG:3513-3565); the
//                     pull form sums the same terms without atomics, in a fixed order.
//
// Data layout (row stride P):
//   pops[2] 19 x P  deviations f - w rho0, canonical (SP) velocity order
//   nbr     packed upstream node per direction or bounce back (nbrpack.h)
//   col      2 x P  rho_r, rho_b after colour streaming (the reference's rho_r/rho_b): doubles in
//                   the FP64 build, int32 fixed point (units of 2^-29) in the FP32 build
//   psi      1 x P  (rho_r - rho_b)/(rho_r + rho_b)
//   cs       8 x P  per node, contiguous: rho_r, rho_b (as in col), ux, uy, uz, nx, ny, nz
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "collide.h"
#include "colour_win.h"
#include "common.h"
#include "handle.h"
#include "pull.h"
#include "storage.h"
#include "tune.h"

namespace lbm {

// Stored colour densities.  The FP32 build keeps rho_r and rho_b in fixed point, int32 units of
// 2^-29 (values up to 4): the colour streaming then splits a node's colour into its 19 outgoing
// shares w_q rho exactly in integers (cg_colour_kernel_fx), so each colour's total is kept to
// the unit whatever the order of the sums, with FP32 arithmetic for the small terms.
template <typename C>
struct ColourStore {
    using T = double;
    static __host__ __device__ __forceinline__ double real(double v) { return v; }
    static __host__ __device__ __forceinline__ double put(double v) { return v; }
};
template <>
struct ColourStore<float> {
    using T = int32_t;
    static constexpr double ONE = 536870912.0;   // 2^29
    static __host__ __device__ __forceinline__ double real(int32_t v) { return (double)v * (1.0 / ONE); }
    static __host__ __device__ __forceinline__ int32_t put(double v) { return (int32_t)rint(v * ONE); }
};

template <typename ST>
struct CgArgs {
    using C = typename Store<ST>::C;
    using CT = typename ColourStore<C>::T;
    const ST *__restrict__ src;
    ST *__restrict__ dst;
    NbrDev nbr;
    CT *__restrict__ col;
    C *__restrict__ psi;
    C *__restrict__ cs;
    C *__restrict__ u_aux;
    const C *__restrict__ rho_aux;
    int64_t N, P;
    C gx, gy, gz, rho0, cap_a, psi_solid, rel1, rel2;
    C wl, wg, lg0, l1, l2, g1, g2, delta;
    C ux_i, uy_i, uz_i;
    int32_t bfa;
    float scale, inv_scale;
    int32_t order;                            // tile order of the collide launch (TileOrder)
    const ColourWin *__restrict__ wins;       // the staged variant's runs per CW_BLOCK-node block
};

// colour state of a node: its stored colour densities (raw) and six values of the compute type.
// FP32: two rows of 16-byte records, (rho_r, rho_b, ux, uy) at [n] and (uz, nx, ny, nz) at
// [P + n], so that a wavefront's loads of either half are 1 KB contiguous.
__device__ __forceinline__ void store_cs(float *cs, int64_t n, int64_t P, int32_t rr, int32_t rb, const float (&v)[6])
{
    float4 *p = reinterpret_cast<float4 *>(cs);
    p[n] = make_float4(__int_as_float(rr), __int_as_float(rb), v[0], v[1]);
    p[P + n] = make_float4(v[2], v[3], v[4], v[5]);
}
__device__ __forceinline__ void store_cs(double *cs, int64_t n, int64_t, double rr, double rb, const double (&w)[6])
{
    const double v[8] = {rr, rb, w[0], w[1], w[2], w[3], w[4], w[5]};
    double2 *p = reinterpret_cast<double2 *>(cs + 8 * n);
    p[0] = make_double2(v[0], v[1]);
    p[1] = make_double2(v[2], v[3]);
    p[2] = make_double2(v[4], v[5]);
    p[3] = make_double2(v[6], v[7]);
}
template <typename C>
__device__ __forceinline__ void load_cs(const C *cs, int64_t n, int64_t P, C (&v)[8]);
template <>
__device__ __forceinline__ void load_cs<float>(const float *cs, int64_t n, int64_t P, float (&v)[8])
{
    const float4 *p = reinterpret_cast<const float4 *>(cs);
    const float4 a = p[n], b = p[P + n];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
template <>
__device__ __forceinline__ void load_cs<double>(const double *cs, int64_t n, int64_t, double (&v)[8])
{
    const double2 *p = reinterpret_cast<const double2 *>(cs + 8 * n);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double2 a = p[k];
        v[2 * k] = a.x;
        v[2 * k + 1] = a.y;
    }
}

template <typename ST, int MODE>
__device__ __forceinline__ void cg_node_state(const CgArgs<ST> &a, int64_t n, typename Store<ST>::C (&h)[Q],
                                              typename Store<ST>::C (&m)[Q], uint32_t (&nb)[Q - 1],
                                              Macro<typename Store<ST>::C> &mc)
{
    using C = typename Store<ST>::C;
    load_nbr(a.nbr, a.P, n, nb);
    if (MODE == K_INIT) {
        mc.rho = C(1);
        mc.drho = C(1) - a.rho0;
        mc.ux = a.ux_i;
        mc.uy = a.uy_i;
        mc.uz = a.uz_i;
        sp_eq_moments(mc, m);   // f = M^-1 meq(1, u0) = feq(1, u0) (CG:2644-2681)
        return;
    }
    h[0] = Store<ST>::load(a.src[n], a.inv_scale);
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        if (MODE == K_DIRECT) {
            h[q] = Store<ST>::load(a.src[(int64_t)q * a.P + n], a.inv_scale);
        } else {
            const uint32_t v = nb[q - 1];
            const int64_t row = (v & BB_FLAG) ? (int64_t)opp(q) : (int64_t)q;
            h[q] = Store<ST>::load(a.src[row * a.P + (int64_t)(v & ~BB_FLAG)], a.inv_scale);
        }
    }
    moments_fwd(h, m);
    if (MODE == K_DIRECT) {
        mc.rho = a.rho_aux[n];
        mc.drho = mc.rho - a.rho0;
        mc.ux = a.u_aux[n];
        mc.uy = a.u_aux[a.P + n];
        mc.uz = a.u_aux[2 * a.P + n];
    } else {
        mc = macro_from_moments(m, a.rho0, a.gx, a.gy, a.gz);
    }
}

// One node of the collide kernel.  SLOTS > 0 and FAST: the phase field of the block's
// neighbourhood is in LDS (lpsi, at the slots of ColourWin w); otherwise it is read from global
// memory.
template <typename ST, int MODE, bool STORE_U, int SLOTS, bool FAST>
__device__ __forceinline__ void cg_collide_node(const CgArgs<ST> &a, int64_t n, const typename Store<ST>::C *lpsi,
                                                const ColourWin &w)
{
    using C = typename Store<ST>::C;
    const int32_t own = FAST ? (int32_t)n + w.own : 0;
    // psi at the site of table column j, whose entry is v (a solid site's value is not used;
    // its index is the node itself)
    auto psi_at = [&](int j, uint32_t v) -> C {
        if (SLOTS > 0 && FAST) {
            const int32_t up = (int32_t)v + w.d[j];
            return lpsi[(v & BB_FLAG) ? own : up];
        }
        return a.psi[v & ~BB_FLAG];
    };
    C h[Q], m[Q];
    uint32_t nb[Q - 1];
    Macro<C> mc;
    cg_node_state<ST, MODE>(a, n, h, m, nb, mc);
    using CS = ColourStore<C>;
    const typename CS::T rr_raw = a.col[n], rb_raw = a.col[a.P + n];
    const C rr = (C)CS::real(rr_raw), rb = (C)CS::real(rb_raw), ps = a.psi[n];

    // colour gradient C = sum_q 3 w_q e_q psi(x + e_q) (CG:2903-3012); x + e_q is the upstream
    // site of the opposite direction
    C cx = C(0), cy = C(0), cz = C(0);
    bool wall = false;
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        const uint32_t v = nb[opp(q) - 1];
        const bool solid = (v & BB_FLAG) != 0;
        wall |= solid;
        // unconditional load, then select
        const C pv = psi_at(opp(q) - 1, v);
        const C pq = solid ? a.psi_solid : pv;
        const C wq = C(3.0 * weight(q));
        if (EX[q] != 0) cx += C(EX[q]) * wq * pq;
        if (EY[q] != 0) cy += C(EY[q]) * wq * pq;
        if (EZ[q] != 0) cz += C(EZ[q]) * wq * pq;
    }
    // no colour gradient next to a wall inside a nearly pure phase (CG:3014-3015)
    const C dr = rr - rb;
    if (wall && (dr >= C(0.9) || dr <= C(-0.9))) cx = cy = cz = C(0);
    const C cc = sqrt(cx * cx + cy * cy + cz * cz);
    C nx = C(0), ny = C(0), nz = C(0);
    if (cc > C(0)) {
        const C inv = C(1) / cc;
        nx = cx * inv;
        ny = cy * inv;
        nz = cz * inv;
    }
    // body force gated by phase (CG:3040) -- the velocity keeps the full force (CG:3737-3739)
    const bool gate = (a.bfa == 1 && ps < a.rel1) || (a.bfa == -1 && ps > -a.rel2);
    const C gx = gate ? C(0) : a.gx, gy = gate ? C(0) : a.gy, gz = gate ? C(0) : a.gz;
    // viscosity interpolated in psi (CG:3163-3176)
    C sv;
    if (ps > C(0))
        sv = ps > a.delta ? a.wl : a.lg0 + a.l1 * ps + a.l2 * ps * ps;
    else
        sv = ps < -a.delta ? a.wg : a.lg0 + a.g1 * ps + a.g2 * ps * ps;

    if (STORE_U && MODE != K_DIRECT) {
        a.u_aux[n] = mc.ux;
        a.u_aux[a.P + n] = mc.uy;
        a.u_aux[2 * a.P + n] = mc.uz;
    }
    cg_relax(m, mc, gx, gy, gz, sv, a.cap_a * cc, nx, ny, nz, h);
#pragma unroll
    for (int q = 0; q < Q; ++q) Store<ST>::store_nt(&a.dst[(int64_t)q * a.P + n], h[q], a.scale);
    const C st[6] = {mc.ux, mc.uy, mc.uz, nx, ny, nz};
    store_cs(a.cs, n, a.P, rr_raw, rb_raw, st);
}

// SLOTS > 0: the staged variant (NB = CW_BLOCK).  The phase field of the block's neighbourhood -- the
// runs of node ids of the colour kernel's staging (ColourWin) -- is copied to LDS first, and the
// colour gradient reads its 18 neighbour values from there; a block with a read outside its runs
// reads them from global memory.
template <typename ST, int MODE, bool STORE_U, int NB = BLOCK, int SLOTS = 0>
__global__ void __launch_bounds__(NB) cg_collide_kernel(CgArgs<ST> a)
{
    using C = typename Store<ST>::C;
    const int64_t b = tile_of(a.order, blockIdx.x, gridDim.x);
    const int64_t n = b * NB + threadIdx.x;
    __shared__ C lpsi[SLOTS > 0 ? SLOTS : 1];
    ColourWin w{};
    if (SLOTS > 0) w = a.wins[b];
    if (SLOTS > 0 && w.fast == CW_ALL) {   // the same for the whole block
        // slots threadIdx.x + i NB: all loads in flight before the first LDS write
        constexpr int NS = SLOTS > 0 ? SLOTS / NB : 1;
        static_assert(SLOTS % NB == 0, "whole slots per lane");
        C v[NS];
#pragma unroll
        for (int i = 0; i < NS; ++i) v[i] = a.psi[cw_slot_node(w, (int32_t)threadIdx.x + i * NB)];
#pragma unroll
        for (int i = 0; i < NS; ++i) lpsi[threadIdx.x + i * NB] = v[i];
        __syncthreads();
    }
    if (n >= a.N) return;
    if (SLOTS > 0 && w.fast == CW_ALL)
        cg_collide_node<ST, MODE, STORE_U, SLOTS, true>(a, n, lpsi, w);
    else
        cg_collide_node<ST, MODE, STORE_U, SLOTS, false>(a, n, lpsi, w);
}

// e_q . (x, y, z) with the zero components left out (0 * v is not folded away in IEEE arithmetic)
// and the others added in the order x, y, z: the same value as the full sum
template <int q, typename T>
__device__ __forceinline__ T e_dot(T x, T y, T z)
{
    constexpr bool hx = EX[q] != 0, hy = EY[q] != 0, hz = EZ[q] != 0;
    T s = T(0);
    if constexpr (hx) s = EX[q] > 0 ? x : -x;
    if constexpr (hy) {
        const T t = EY[q] > 0 ? y : -y;
        s = hx ? s + t : t;
    }
    if constexpr (hz) {
        const T t = EZ[q] > 0 ? z : -z;
        s = (hx || hy) ? s + t : t;
    }
    return s;
}

// f(std::integral_constant<int, I>) for I = B .. E - 1, unrolled at compile time
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// g_r, g_b leaving a node with colour state v in direction e_q, or -e_q when bb (the opposite
// direction, used for the bounce-back term of the pull): the equilibrium split
// g_c = w rho_c phi_q, phi_q = 1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u (CG:3477-3485), then the
// recolouring g_r += d, g_b -= d, d = min(g_r,q, g_r,opp, g_b,q, g_b,opp) e.C/|C| (CG:3493-3508)
template <int q, typename C>
__device__ __forceinline__ void colour_pull_term(const C (&v)[8], bool bb, C &gr, C &gb)
{
    const C rr = v[0], rb = v[1], ux = v[2], uy = v[3], uz = v[4];
    const C uu = ux * ux + uy * uy + uz * uz;
    const C du = e_dot<q>(ux, uy, uz), eu = bb ? -du : du;
    const C base = C(1) - C(1.5) * uu + C(4.5) * eu * eu;
    const C phq = base + C(3) * eu, php = base - C(3) * eu;
    const C w = C(weight(q));
    const C wr = w * rr, wb = w * rb;
    gr = wr * phq;
    gb = wb * phq;
    const C mn = fmin(fmin(fmin(gr, wr * php), gb), wb * php);
    const C dn = e_dot<q>(v[5], v[6], v[7]), en = bb ? -dn : dn;
    gr += mn * en;
    gb -= mn * en;
}

// The FP64 build's colour sums of one node, direction by direction: begin() with the node's own
// colour state (the rest population stays on the node: e_0 = 0, no recolouring term), add() the
// term arriving along q from an upstream state (bb: BB_FLAG when the upstream site is solid and v
// is the node's own state, whose opposite term comes back), then store().  The gather and the
// staged kernel add the same terms in the same order, so they give the same bits.
struct ColourSums64 {
    double rr, rb;
    __device__ __forceinline__ void begin(const double (&own)[8])
    {
        const double uu = own[2] * own[2] + own[3] * own[3] + own[4] * own[4];
        const double w0 = weight(0) * (1.0 - 1.5 * uu);
        rr = w0 * own[0];
        rb = w0 * own[1];
    }
    template <int q>
    __device__ __forceinline__ void add(uint32_t bb, const double (&v)[8])
    {
        double gr, gb;
        colour_pull_term<q>(v, bb != 0, gr, gb);
        rr += gr;
        rb += gb;
    }
    __device__ __forceinline__ void store(int64_t n, int64_t P, double *__restrict__ col, double *__restrict__ psi) const
    {
        col[n] = rr;
        col[P + n] = rb;
        psi[n] = (rr - rb) / (rr + rb);
    }
};

// FP64 gathers of the 18 upstream colour states: all 18 source indices first, then the states in
// groups of three directions with no branch between them, so that each group's gathers are in
// flight together (measured best: 351 us at 3, 402 us at 6).  A solid upstream site sends back
// this node's own opposite term: same load address n, sign -1.
__device__ __forceinline__ void colour_gather64(const double *__restrict__ cs, const uint32_t (&src)[Q - 1], int64_t n,
                                                int64_t P, ColourSums64 &acc)
{
    constexpr int G = 3;
    static_for<0, (Q - 1) / G>([&](auto g) {
        constexpr int q0 = 1 + decltype(g)::value * G;
        double up[G][8];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const uint32_t v = src[q0 + j - 1];
            load_cs(cs, (v & BB_FLAG) ? n : (int64_t)v, P, up[j]);
        }
        static_for<0, G>([&](auto jj) {
            constexpr int j = decltype(jj)::value;
            acc.add<q0 + j>(src[q0 + j - 1] & BB_FLAG, up[j]);
        });
    });
}

__global__ void __launch_bounds__(BLOCK) cg_colour_kernel64(NbrDev nbr, const double *__restrict__ cs,
                                                            double *__restrict__ col, double *__restrict__ psi,
                                                            int64_t N, int64_t P, int32_t order)
{
    const int64_t n = tile_of(order, blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (n >= N) return;
    double own[8];
    load_cs(cs, n, P, own);
    uint32_t src[Q - 1];
    load_nbr(nbr, P, n, src);
    ColourSums64 acc;
    acc.begin(own);
    colour_gather64(cs, src, n, P, acc);
    acc.store(n, P, col, psi);
}

// The FP32 build's colour streaming.  Each node's stored colour R (fixed point, units of 2^-29)
// leaves as 19 shares: the bulk w_q R, in integers -- the nearest unit to R/18 along an axis, to
// R/36 along a diagonal, and the remainder R - 6 R/18 - 12 R/36 on the rest population -- and
// the small rest of g_q, w_q rho (phi_q - 1) plus the recolouring term, in FP32.  Every receiver
// computes the bulk share of an upstream node from the same integer, so the bulk shares of a
// node sum to its R exactly and each colour's total is conserved to the unit, in any summation
// order; the FP32 small terms sum to zero up to their rounding (~1e-12 of the colour per
// step).  The recolouring term of a pair of opposite directions comes out of the same
// operations with the sign of e_q flipped, so its two shares cancel exactly.  The received
// small terms are added to the integer sum rounded to the nearest unit.
//
// The bulk shares are floor((R + 9) / 18) and floor((R + 18) / 36), formed as (x * M) >> 34 and
// >> 35 with M = ceil(2^34 / 18), which is exact for 0 <= x < 2^31 (checked over that whole range
// on the host); a negative x gets the same deterministic form, so conservation holds for it too.
__device__ __forceinline__ int32_t colour_share(bool axis, int32_t R)
{
    const int32_t h = __mulhi(axis ? R + 9 : R + 18, 954437177);
    return axis ? h >> 2 : h >> 3;
}


// The colour sums of one node, direction by direction: begin() with the node's own colour state
// (the rest population), add() the share arriving along q from an upstream state (bb: the
// upstream site is solid and vs is the node's own state), then store().
struct ColourSums {
    static constexpr float UNIT = (float)(1.0 / ColourStore<float>::ONE);   // 2^-29, exact
    static constexpr float ONE = (float)ColourStore<float>::ONE;
    int32_t sr, sb;
    float tr, tb;
    __device__ __forceinline__ void begin(const float (&own)[8])
    {
        const int32_t R0 = __float_as_int(own[0]), B0 = __float_as_int(own[1]);
        // rest population: the remainder of the bulk, and w0 rho (phi_0 - 1) = -1.5 w0 rho u.u
        sr = R0 - 6 * colour_share(true, R0) - 12 * colour_share(false, R0);
        sb = B0 - 6 * colour_share(true, B0) - 12 * colour_share(false, B0);
        const float uu0 = own[2] * own[2] + own[3] * own[3] + own[4] * own[4];
        const float c0 = (float)weight(0) * -1.5f * uu0;
        tr = c0 * ((float)R0 * UNIT);
        tb = c0 * ((float)B0 * UNIT);
    }
    // bb: BB_FLAG (the sign bit) when the upstream site is solid, else 0
    template <int q>
    __device__ __forceinline__ void add(uint32_t bb, const float (&vs)[8])
    {
        constexpr bool axis = e2_of(q) == 1;
        const int32_t R = __float_as_int(vs[0]), B = __float_as_int(vs[1]);
        sr += colour_share(axis, R);
        sb += colour_share(axis, B);
        // w rho: (float)R * 2^-29 * w rounded once, as 2^-29 scales exactly
        constexpr float WU = (float)weight(q) * UNIT;
        const float wr = (float)R * WU, wb = (float)B * WU;
        const float ux = vs[2], uy = vs[3], uz = vs[4];
        const float uu = ux * ux + uy * uy + uz * uz;
        // a solid upstream site sends back this node's own opposite share: e_q negated, by
        // flipping the sign bit
        const float eu = __int_as_float(__float_as_int(e_dot<q>(ux, uy, uz)) ^ (int32_t)bb);
        // phi_q - 1 formed from the small terms alone: rounding 1 + small and taking 1 off again
        // would be biased, and would drift the colour totals
        const float s2 = 4.5f * eu * eu - 1.5f * uu;
        const float eps = s2 + 3.0f * eu;
        const float phq = 1.0f + eps, php = 1.0f + (s2 - 3.0f * eu);
        const float mn = fminf(fminf(fminf(wr * phq, wr * php), wb * phq), wb * php);
        const float en = __int_as_float(__float_as_int(e_dot<q>(vs[5], vs[6], vs[7])) ^ (int32_t)bb);
        tr += wr * eps + mn * en;
        tb += wb * eps - mn * en;
    }
    __device__ __forceinline__ void store(int64_t n, int64_t P, int32_t *__restrict__ col, float *__restrict__ psi) const
    {
        const int32_t Rn = sr + (int32_t)rintf(tr * ONE), Bn = sb + (int32_t)rintf(tb * ONE);
        col[n] = Rn;
        col[P + n] = Bn;
        psi[n] = (float)(Rn - Bn) / (float)(Rn + Bn);
    }
};

template <int q>
struct ColourStep {
    template <typename F>
    static __device__ __forceinline__ void run(ColourSums &acc, const uint32_t (&src)[Q - 1], F &&fetch)
    {
        float v[8];
        const uint32_t id = src[q - 1];
        fetch(q - 1, id, v);
        acc.add<q>(id & BB_FLAG, v);
        ColourStep<q + 1>::run(acc, src, fetch);
    }
};
template <>
struct ColourStep<Q> {
    template <typename F>
    static __device__ __forceinline__ void run(ColourSums &, const uint32_t (&)[Q - 1], F &&)
    {
    }
};

// The same over groups of CW_GROUP directions, software-pipelined: the states of the next group
// are read before the sums of this one, and the scheduler is kept from hoisting more reads (which
// would hold all 18 states in registers).
#ifndef CW_GROUP
#define CW_GROUP 3
#endif
static_assert((Q - 1) % CW_GROUP == 0, "groups cover the 18 directions");
// (T: the state's value type; Acc: ColourSums or ColourSums64)
template <int q, typename T, typename Acc, int i = 0>
struct ColourAdd {
    static __device__ __forceinline__ void run(Acc &acc, const uint32_t (&src)[Q - 1], const T (&v)[CW_GROUP][8])
    {
        acc.template add<q + i>(src[q + i - 1] & BB_FLAG, v[i]);
        if constexpr (i + 1 < CW_GROUP) ColourAdd<q, T, Acc, i + 1>::run(acc, src, v);
    }
};
template <int q, typename T, typename Acc>
struct ColourGroups {
    template <typename F>
    static __device__ __forceinline__ void run(Acc &acc, const uint32_t (&src)[Q - 1], const T (&cur)[CW_GROUP][8],
                                               F &&fetch)
    {
        constexpr int qn = q + CW_GROUP;
        T nxt[CW_GROUP][8];
        if constexpr (qn < Q) {
#pragma unroll
            for (int j = 0; j < CW_GROUP; ++j) fetch(qn + j - 1, src[qn + j - 1], nxt[j]);
        }
        __builtin_amdgcn_sched_barrier(0);
        ColourAdd<q, T, Acc>::run(acc, src, cur);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (qn < Q) ColourGroups<qn, T, Acc>::run(acc, src, nxt, fetch);
    }
};

// gathers of the 18 upstream colour states from global memory (through the caches)
__global__ void __launch_bounds__(BLOCK) cg_colour_kernel_fx(NbrDev nbr, const float *__restrict__ cs,
                                                             int32_t *__restrict__ col, float *__restrict__ psi,
                                                             int64_t N, int64_t P, int32_t order)
{
    const int64_t n = tile_of(order, blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (n >= N) return;
    float own[8];
    load_cs(cs, n, P, own);
    uint32_t src[Q - 1];
    load_nbr(nbr, P, n, src);
    ColourSums acc;
    acc.begin(own);
    ColourStep<1>::run(acc, src,
                       [&](int, uint32_t v, float (&out)[8]) { load_cs(cs, (v & BB_FLAG) ? n : (int64_t)v, P, out); });
    acc.store(n, P, col, psi);
}

// The same with the colour states of the block's upstream neighbourhood staged in LDS first.
// A block's CW_BLOCK = 512 nodes lie on five to seven z-lines of one x-plane; their upstream
// nodes are the nodes of a few neighbouring lines in the planes x-1, x and x+1, i.e. a few runs
// of consecutive node ids.  The host lists, per block, up to CW_RUNS such runs (ColourWin) that
// fit in SLOTS states; the block copies them with coalesced loads, then each of its 18 x 512 reads comes
// from LDS at slot id + d[j] -- one add per read, no search, no branch -- and the reads run a
// group of directions ahead of the sums.  Each state is copied by the few blocks whose
// neighbourhoods hold it, instead of being gathered by each of its 18 downstream nodes.  A block
// with a read outside its runs (across the periodic wrap of a plane) gathers all its states
// from global memory instead, as cg_colour_kernel_fx does.
template <int SLOTS>
__global__ void __launch_bounds__(CW_BLOCK) cg_colour_kernel_win(NbrDev nbr, const float *__restrict__ cs,
                                                              const ColourWin *__restrict__ wins,
                                                              int32_t *__restrict__ col, float *__restrict__ psi,
                                                              int64_t N, int64_t P, int32_t order)
{
    // a plain vector type, which the compiler keeps in registers (HIP's float4 is a struct whose
    // arrays it left in scratch memory here)
    using f4 = float __attribute__((ext_vector_type(4)));
    __shared__ f4 lds[2 * SLOTS];   // two rows of SLOTS records, as in global memory
    const int64_t b = tile_of(order, blockIdx.x, gridDim.x);
    const ColourWin w = wins[b];
    const f4 *__restrict__ cs4 = reinterpret_cast<const f4 *>(cs);
    const int64_t n = b * CW_BLOCK + threadIdx.x;
    // the neighbour table first, in flight during the staging (the table has P >= N rows, P a
    // multiple of 256, so a 512-node block can reach past its end: lanes past N load nothing)
    uint32_t src[Q - 1];
    if (n < N) load_nbr(nbr, P, n, src);
    if (w.fast == CW_ALL) {   // the same for the whole block
        // slots threadIdx.x + i CW_BLOCK: all loads in flight before the first LDS write
        constexpr int NS = SLOTS / CW_BLOCK;
        static_assert(SLOTS % CW_BLOCK == 0, "whole slots per lane");
        f4 ra[NS], rc[NS];
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int64_t g = cw_slot_node(w, (int32_t)threadIdx.x + i * CW_BLOCK);
            ra[i] = cs4[g];
            rc[i] = cs4[P + g];
        }
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            lds[threadIdx.x + i * CW_BLOCK] = ra[i];
            lds[SLOTS + threadIdx.x + i * CW_BLOCK] = rc[i];
        }
        __syncthreads();
    }
    if (n >= N) return;
    auto lds_state = [&](int32_t slot, float (&v)[8]) {
        const f4 a = lds[slot], c = lds[SLOTS + slot];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = c.x; v[5] = c.y; v[6] = c.z; v[7] = c.w;
    };
    ColourSums acc;
    if (w.fast == CW_ALL) {
        const int32_t own = (int32_t)n + w.own;
        float own_v[8];
        lds_state(own, own_v);
        // the reads select an address, not a value: hidden from the optimiser, `own` is not
        // matched with the read above, which would turn the select into a branch per lane group
        int32_t own_sel = own;
        asm volatile("" : "+v"(own_sel));
        auto fetch = [&](int j, uint32_t v, float (&out)[8]) {
            const int32_t up = (int32_t)v + w.d[j];
            lds_state((v & BB_FLAG) ? own_sel : up, out);
        };
        float first[CW_GROUP][8];
#pragma unroll
        for (int j = 0; j < CW_GROUP; ++j) fetch(j, src[j], first[j]);
        acc.begin(own_v);
        ColourGroups<1, float, ColourSums>::run(acc, src, first, fetch);
        acc.store(n, P, col, psi);
        return;
    }
    // the other blocks (a node or a direction of the block outside the runs): plain gathers
    float own[8];
    load_cs(cs, n, P, own);
    acc.begin(own);
    ColourStep<1>::run(acc, src,
                       [&](int, uint32_t v, float (&out)[8]) { load_cs(cs, (v & BB_FLAG) ? n : (int64_t)v, P, out); });
    acc.store(n, P, col, psi);
}

// The FP64 colour streaming with the same staging: a 64-byte state per LDS slot, in four rows of
// 16-byte quarters (so a wavefront's read of a quarter is 1 KB contiguous); 2560 slots fill the
// CU's 160 KB, one block per CU.
template <int SLOTS>
__global__ void __launch_bounds__(CW_BLOCK) cg_colour_kernel_win64(NbrDev nbr, const double *__restrict__ cs,
                                                                 const ColourWin *__restrict__ wins,
                                                                 double *__restrict__ col, double *__restrict__ psi,
                                                                 int64_t N, int64_t P, int32_t order)
{
    using d2 = double __attribute__((ext_vector_type(2)));
    __shared__ d2 lds[4 * SLOTS];
    const int64_t b = tile_of(order, blockIdx.x, gridDim.x);
    const ColourWin w = wins[b];
    const d2 *__restrict__ cs2 = reinterpret_cast<const d2 *>(cs);   // node m: cs2[4 m .. 4 m + 3]
    const int64_t n = b * CW_BLOCK + threadIdx.x;
    uint32_t src[Q - 1];
    if (n < N) load_nbr(nbr, P, n, src);   // lanes past N load nothing (see cg_colour_kernel_win)
    if (w.fast == CW_ALL) {   // the same for the whole block
        constexpr int NS = SLOTS / CW_BLOCK;
        static_assert(SLOTS % CW_BLOCK == 0, "whole slots per lane");
        d2 r[NS][4];
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int64_t g = cw_slot_node(w, (int32_t)threadIdx.x + i * CW_BLOCK);
#pragma unroll
            for (int c = 0; c < 4; ++c) r[i][c] = cs2[4 * g + c];
        }
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int c = 0; c < 4; ++c) lds[c * SLOTS + threadIdx.x + i * CW_BLOCK] = r[i][c];
        __syncthreads();
    }
    if (n >= N) return;
    auto lds_state = [&](int32_t slot, double (&v)[8]) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const d2 x = lds[c * SLOTS + slot];
            v[2 * c] = x.x;
            v[2 * c + 1] = x.y;
        }
    };
    ColourSums64 acc;
    if (w.fast == CW_ALL) {
        const int32_t own = (int32_t)n + w.own;
        double own_v[8];
        lds_state(own, own_v);
        int32_t own_sel = own;   // kept apart from `own`, as in cg_colour_kernel_win
        asm volatile("" : "+v"(own_sel));
        auto fetch = [&](int j, uint32_t v, double (&out)[8]) {
            const int32_t up = (int32_t)v + w.d[j];
            lds_state((v & BB_FLAG) ? own_sel : up, out);
        };
        double first[CW_GROUP][8];
#pragma unroll
        for (int j = 0; j < CW_GROUP; ++j) fetch(j, src[j], first[j]);
        acc.begin(own_v);
        ColourGroups<1, double, ColourSums64>::run(acc, src, first, fetch);
        acc.store(n, P, col, psi);
        return;
    }
    double own[8];
    load_cs(cs, n, P, own);
    acc.begin(own);
    colour_gather64(cs, src, n, P, acc);
    acc.store(n, P, col, psi);
}

// ---------------------------------------------------------------- diagnostics
constexpr int CG_NDIAG = 15;   // Ql[3] Qg[3] err_num err_den speed umax2 nan wet box Sd_l Sd_g
constexpr uint32_t CG_MAX_MASK = 1u << 9;

struct CgBox {
    int32_t on, xn, xp, yn, yp, zn, zp, ny, nz, perdir;
};

template <typename ST, int MODE>
__global__ void __launch_bounds__(BLOCK) cg_diag_kernel(CgArgs<ST> a, const int32_t *__restrict__ supinv, CgBox box,
                                                        double *__restrict__ partial)
{
    using C = typename Store<ST>::C;
    const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double v[CG_NDIAG];
#pragma unroll
    for (int k = 0; k < CG_NDIAG; ++k) v[k] = 0.0;
    if (n < a.N) {
        C h[Q], m[Q];
        uint32_t nb[Q - 1];
        Macro<C> mc;
        cg_node_state<ST, MODE>(a, n, h, m, nb, mc);
        const double ux = (double)mc.ux, uy = (double)mc.uy, uz = (double)mc.uz, ps = (double)a.psi[n];
        const int32_t s = supinv[n];
        const int32_t nyz = box.ny * box.nz;
        const int32_t i = s / nyz, j = (s % nyz) / box.nz, k = s % box.nz;
        bool inside = true;
        if (box.on) inside = i >= box.xn && i <= box.xp && j >= box.yn && j <= box.yp && k >= box.zn && k <= box.zp;
        if (inside) {
            // saturation "displacement": psi-weighted position along the flow axis (CG:6655-6667)
            const double sdir = box.perdir == 2 ? (double)j : (box.perdir == 3 ? (double)k : (double)i);
            if (ps >= 0.0)
                v[13] = ps * sdir;
            else
                v[14] = ps * sdir;
            if (ps >= (double)a.rel1) {
                v[0] = ux;
                v[1] = uy;
                v[2] = uz;
            }
            if (ps <= -(double)a.rel2) {
                v[3] = ux;
                v[4] = uy;
                v[5] = uz;
            }
            v[11] = ps >= 0.0 ? 1.0 : 0.0;
            v[12] = 1.0;
        }
        if (n < a.N - 1) {   // Error skips the last node (CG:5001)
            const double dx = ux - (double)a.u_aux[n], dy = uy - (double)a.u_aux[a.P + n],
                         dz = uz - (double)a.u_aux[2 * a.P + n];
            const double s2 = ux * ux + uy * uy + uz * uz;
            v[6] = dx * dx + dy * dy + dz * dz;
            v[7] = s2;
            v[8] = sqrt(s2);
            v[9] = s2 > 0.0 ? s2 : 0.0;   // a NaN never passes the reference's "> u_max" (CG:5005-5006)
        }
        v[10] = (isfinite(ux) && isfinite(uy) && isfinite(uz) && isfinite(ps)) ? 0.0 : 1.0;
    }
    bool is_max[CG_NDIAG];
#pragma unroll
    for (int k = 0; k < CG_NDIAG; ++k) is_max[k] = (CG_MAX_MASK >> k) & 1u;
    block_reduce_store<CG_NDIAG>(v, is_max, partial + (int64_t)blockIdx.x * CG_NDIAG);
}

// ---------------------------------------------------------------- state I/O
// state of the nodes n0 <= n < n1 at local index n - n0 of the outputs (any may be null)
template <typename ST, int MODE>
__global__ void __launch_bounds__(BLOCK) cg_get_state_kernel(CgArgs<ST> a, int64_t n0, int64_t n1,
                                                             double *__restrict__ f_out, double *__restrict__ u_out,
                                                             double *__restrict__ rho_out, double *__restrict__ psi_out,
                                                             double *__restrict__ rr_out, double *__restrict__ rb_out)
{
    using C = typename Store<ST>::C;
    const int64_t n = group_floor(n0) + (int64_t)blockIdx.x * BLOCK + threadIdx.x;   // see load_nbr
    if (n < n0 || n >= n1) return;
    const int64_t l = n - n0;
    C h[Q], m[Q];
    uint32_t nb[Q - 1];
    Macro<C> mc;
    cg_node_state<ST, MODE>(a, n, h, m, nb, mc);
    if (f_out) {
        double f[Q];
        if (MODE == K_INIT) {
            const double ux = (double)mc.ux, uy = (double)mc.uy, uz = (double)mc.uz;
            const double uv = ux * ux + uy * uy + uz * uz;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const double eu = EX[q] * ux + EY[q] * uy + EZ[q] * uz;
                f[q] = weight(q) * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uv);
            }
        } else {
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = (double)h[q] + weight(q) * (double)a.rho0;
        }
#pragma unroll
        for (int qc = 0; qc < Q; ++qc) f_out[l * Q + qc] = f[cg_to_canon(qc)];
    }
    if (u_out) {
        u_out[l * 3 + 0] = (double)mc.ux;
        u_out[l * 3 + 1] = (double)mc.uy;
        u_out[l * 3 + 2] = (double)mc.uz;
    }
    if (rho_out) rho_out[l] = (double)mc.rho;
    if (psi_out) psi_out[l] = (double)a.psi[n];
    if (rr_out) rr_out[l] = ColourStore<C>::real(a.col[n]);
    if (rb_out) rb_out[l] = ColourStore<C>::real(a.col[a.P + n]);
}

// a chunk of a post-streaming state: nodes n0 <= n < n1 at local index n - n0 of the inputs
template <typename ST>
__global__ void __launch_bounds__(BLOCK) cg_load_state_kernel(CgArgs<ST> a, int64_t n0, int64_t n1,
                                                              const double *__restrict__ f, const double *__restrict__ u,
                                                              const double *__restrict__ rho,
                                                              const double *__restrict__ psi, const double *__restrict__ rr,
                                                              const double *__restrict__ rb, typename Store<ST>::C *rho_aux)
{
    using C = typename Store<ST>::C;
    const int64_t n = n0 + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (n >= n1) return;
    const int64_t l = n - n0;
#pragma unroll
    for (int qc = 0; qc < Q; ++qc) {
        const int q = cg_to_canon(qc);
        a.dst[(int64_t)q * a.P + n] = Store<ST>::put((C)(f[l * Q + qc] - weight(q) * (double)a.rho0), a.scale);
    }
    for (int c = 0; c < 3; ++c) a.u_aux[(int64_t)c * a.P + n] = (C)u[l * 3 + c];
    rho_aux[n] = (C)rho[l];
    a.psi[n] = (C)psi[l];
    a.col[n] = ColourStore<C>::put(rr[l]);
    a.col[a.P + n] = ColourStore<C>::put(rb[l]);
}

// phase of the nodes n0 <= n < n1, psi_node at local index n - n0
template <typename C>
__global__ void __launch_bounds__(BLOCK) cg_init_colour_kernel(const double *__restrict__ psi_node,
                                                               typename ColourStore<C>::T *col, C *psi,
                                                               int64_t n0, int64_t n1, int64_t P)
{
    const int64_t n = n0 + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (n >= n1) return;
    double p = psi_node[n - n0];
    p = p > 1.0 ? 1.0 : (p < -1.0 ? -1.0 : p);
    const double rr = (p * 1.0 + 1.0) / 2;   // rho_r = (psi rho + rho)/2 with rho = 1 (CG:2531-2532)
    psi[n] = (C)p;
    col[n] = ColourStore<C>::put(rr);
    col[P + n] = ColourStore<C>::put(1.0 - rr);
}

// ---------------------------------------------------------------- host side
namespace {

template <typename ST>
CgArgs<ST> make_cg_args(const lbm_cg_params &p, const double *visc, const double *u_i, const void *src, void *dst,
                        const NbrDev &nbr, void *col, void *psi, void *cs, void *u_aux, const void *rho_aux,
                        int64_t N, int64_t P)
{
    using C = typename Store<ST>::C;
    CgArgs<ST> a;
    a.src = (const ST *)src;
    a.dst = (ST *)dst;
    a.nbr = nbr;
    a.col = (typename CgArgs<ST>::CT *)col;
    a.psi = (C *)psi;
    a.cs = (C *)cs;
    a.u_aux = (C *)u_aux;
    a.rho_aux = (const C *)rho_aux;
    a.N = N;
    a.P = P;
    a.gx = (C)p.g[0];
    a.gy = (C)p.g[1];
    a.gz = (C)p.g[2];
    a.rho0 = (C)p.rho0;
    a.cap_a = (C)p.cap_a;
    a.psi_solid = (C)p.cos_theta;
    a.rel1 = (C)p.rel_perm_psi;
    a.rel2 = (C)p.rel_perm_psi2;
    a.wl = (C)visc[0];
    a.wg = (C)visc[1];
    a.lg0 = (C)visc[2];
    a.l1 = (C)visc[3];
    a.l2 = (C)visc[4];
    a.g1 = (C)visc[5];
    a.g2 = (C)visc[6];
    a.delta = (C)visc[7];
    a.ux_i = (C)u_i[0];
    a.uy_i = (C)u_i[1];
    a.uz_i = (C)u_i[2];
    a.bfa = p.bodyforce_apply;
    a.scale = 1.0f;
    a.inv_scale = 1.0f;
    a.order = ORDER_RUNS8;
    a.wins = nullptr;
    return a;
}

int cg_kmode(StateMode m) { return m == STATE_INIT ? K_INIT : (m == STATE_DIRECT ? K_DIRECT : K_PULL); }

}  // namespace


CgSolver::~CgSolver() { release(); }

void CgSolver::release()
{
    (void)hipSetDevice(device_);
    stage_.release();
    for (void *&p : pops_)
        if (p) (void)hipFree(p), p = nullptr;
    nbr_.release();
    void **bufs[] = {(void **)&d_supinv_, &d_col_, &d_psi_, &d_cs_, &d_cwin_, &d_u_, &d_rho_,
                     (void **)&d_partial_, (void **)&d_result_};
    for (void **b : bufs)
        if (*b) (void)hipFree(*b), *b = nullptr;
    if (h_result_) (void)hipHostFree(h_result_), h_result_ = nullptr;
    if (ev0_) (void)hipEventDestroy(ev0_), ev0_ = nullptr;
    if (ev1_) (void)hipEventDestroy(ev1_), ev1_ = nullptr;
    if (stream_) (void)hipStreamDestroy(stream_), stream_ = nullptr;
}

int CgSolver::create(const lbm_cg_params &p, const int32_t *solid)
{
    prm_ = p;
    if (p.precision != LBM_FP64 && p.precision != LBM_FP32)
        return fail(LBM_ERR_ARG, "colour gradient supports LBM_FP64 and LBM_FP32");
    if (!(p.nu_l > 0.0) || !(p.nu_g > 0.0)) return fail(LBM_ERR_ARG, "viscosities must be positive");
    if (prm_.rho0 == 0.0) prm_.rho0 = 1.0;
    if (prm_.perm_dir < 1 || prm_.perm_dir > 3) prm_.perm_dir = 1;
    std::string err;
    int rc = build_sparse_geometry(p.nx, p.ny, p.nz, solid, geo_, err);
    if (rc != LBM_OK) return fail(rc, err);
    // viscosity interpolation constants, delta = 0.1 (CG:2368-2377)
    const double c_s2 = 1.0 / 3.0, delta = 0.1;
    const double wl = 1.0 / (p.nu_l / (c_s2 * 1.0) + 0.5), wg = 1.0 / (p.nu_g / (c_s2 * 1.0) + 0.5);
    const double lg0 = 2 * wl * wg / (wl + wg), l1 = 2 * (wl - lg0) / delta, l2 = -l1 / (2 * delta);
    const double g1 = 2 * (lg0 - wg) / delta, g2 = g1 / (2 * delta);
    const double v[8] = {wl, wg, lg0, l1, l2, g1, g2, delta};
    std::memcpy(visc_, v, sizeof(v));
    ct_bytes_ = p.precision == LBM_FP64 ? 8 : 4;
    const size_t st = ct_bytes_;

    device_ = p.device;
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    LBM_HIP_TRY(hipEventCreate(&ev0_));
    LBM_HIP_TRY(hipEventCreate(&ev1_));
    const int64_t P = geo_.P;
    grid_ = (geo_.N + BLOCK - 1) / BLOCK;
    cw_grid_ = (geo_.N + CW_BLOCK - 1) / CW_BLOCK;
    for (void *&b : pops_) LBM_HIP_TRY(hipMalloc(&b, (size_t)Q * P * st));
    LBM_HIP_TRY(hipMalloc(&d_supinv_, (size_t)P * sizeof(int32_t)));
    LBM_HIP_TRY(hipMalloc(&d_col_, (size_t)2 * P * ct_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_psi_, (size_t)P * ct_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_cs_, (size_t)8 * P * ct_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_u_, (size_t)3 * P * ct_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_rho_, (size_t)P * ct_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_partial_, (size_t)grid_ * CG_NDIAG * sizeof(double)));
    LBM_HIP_TRY(hipMalloc(&d_result_, 32 * sizeof(double)));
    LBM_HIP_TRY(hipHostMalloc((void **)&h_result_, 32 * sizeof(double), hipHostMallocDefault));
    PackedNbr packed;
    rc = pack_neighbours(geo_, packed, err);
    if (rc != LBM_OK) return fail(rc, err);
    LBM_HIP_TRY(nbr_.upload(packed, P, stream_));
    LBM_HIP_TRY(hipMemcpyAsync(d_supinv_, geo_.supinv.data(), geo_.supinv.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, stream_));
    LBM_HIP_TRY(hipMemsetAsync(d_u_, 0, (size_t)3 * P * ct_bytes_, stream_));
    {
        // LDS runs for the staged kernels: the smallest staging size in which 85 % of the blocks
        // read every state from LDS, else the size with the most such blocks.  A larger size
        // holds fewer blocks per CU: on a 192^3 sphere pack, 1536 states (91 % of the blocks)
        // gave a 259 us step, 2048 states (97 %) 297 us.
        std::vector<ColourWin> wins;
        int best = 0, last = 0;
        double best_fast = -1.0;
        for (int slots : {1024, 1536, 2560}) {
            const double fast = build_colour_windows(geo_, cw_grid_, slots, wins);
            last = slots;
            if (fast > best_fast) best = slots, best_fast = fast;
            if (fast >= 0.85) break;
        }
        if (last != best) build_colour_windows(geo_, cw_grid_, best, wins);
        cwin_slots_ = best;
        cwin_fast_ = best_fast;
        LBM_HIP_TRY(hipMalloc(&d_cwin_, wins.size() * sizeof(ColourWin)));
        LBM_HIP_TRY(hipMemcpyAsync(d_cwin_, wins.data(), wins.size() * sizeof(ColourWin), hipMemcpyHostToDevice, stream_));
        LBM_HIP_TRY(hipStreamSynchronize(stream_));
    }
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    std::vector<uint32_t>().swap(geo_.nbr);
    rc = autotune();
    if (rc != LBM_OK) return rc;
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    return LBM_OK;   // the state is set by init() or load_state()
}

int CgSolver::init(const double *psi_ijk, const double u0[3])
{
    if (!psi_ijk || !u0) return fail(LBM_ERR_ARG, "null phase field or velocity");
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(stage_.ensure());
    const int64_t N = geo_.N, chunk = stage_chunk(STAGE_BYTES, 1);
    std::vector<double> pn((size_t)std::min(chunk, N));
    hipError_t e = hipSuccess;
    for (int64_t n0 = 0; n0 < N && e == hipSuccess; n0 += chunk) {
        const int64_t c = std::min(chunk, N - n0);
        for (int64_t n = 0; n < c; ++n) pn[(size_t)n] = psi_ijk[geo_.supinv[(size_t)(n0 + n)]];
        double *d = (double *)stage_.buf;
        e = hipMemcpyAsync(d, pn.data(), (size_t)c * sizeof(double), hipMemcpyHostToDevice, stream_);
        if (e != hipSuccess) break;
        dim3 g((unsigned)((c + BLOCK - 1) / BLOCK)), b(BLOCK);
        if (prm_.precision == LBM_FP64)
            cg_init_colour_kernel<double><<<g, b, 0, stream_>>>(d, (double *)d_col_, (double *)d_psi_, n0, n0 + c, geo_.P);
        else
            cg_init_colour_kernel<float><<<g, b, 0, stream_>>>(d, (int32_t *)d_col_, (float *)d_psi_, n0, n0 + c, geo_.P);
        e = hipGetLastError();
        // the host chunk is refilled next round
        if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    }
    const hipError_t es = hipStreamSynchronize(stream_);
    LBM_HIP_TRY(e);
    LBM_HIP_TRY(es);
    for (int c = 0; c < 3; ++c) u_i_[c] = u0[c];
    mode_ = STATE_INIT;
    steps_ = 0;
    return LBM_OK;
}

int CgSolver::load_state(const double *f, const double *u, const double *rho, const double *psi, const double *rr,
                         const double *rb)
{
    if (!f || !u || !rho || !psi || !rr || !rb) return fail(LBM_ERR_ARG, "null state array");
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(stage_.ensure());
    const int64_t N = geo_.N, chunk = stage_chunk(STAGE_BYTES, Q + 7);
    hipError_t e = hipSuccess;
    for (int64_t n0 = 0; n0 < N && e == hipSuccess; n0 += chunk) {
        const int64_t c = std::min(chunk, N - n0);
        double *df = (double *)stage_.buf, *du = df + c * Q, *dr = du + 3 * c, *dp = dr + c, *drr = dp + c, *drb = drr + c;
        const double *srcs[6] = {f + n0 * Q, u + n0 * 3, rho + n0, psi + n0, rr + n0, rb + n0};
        double *dsts[6] = {df, du, dr, dp, drr, drb};
        const int64_t lens[6] = {c * Q, 3 * c, c, c, c, c};
        for (int k = 0; k < 6 && e == hipSuccess; ++k)
            e = hipMemcpyAsync(dsts[k], srcs[k], (size_t)lens[k] * sizeof(double), hipMemcpyHostToDevice, stream_);
        if (e != hipSuccess) break;
        dim3 g((unsigned)((c + BLOCK - 1) / BLOCK)), b(BLOCK);
        if (prm_.precision == LBM_FP64) {
            auto a = make_cg_args<double>(prm_, visc_, u_i_, nullptr, pops_[cur_], nbr_.view(), d_col_, d_psi_, d_cs_, d_u_,
                                          d_rho_, N, geo_.P);
            cg_load_state_kernel<double><<<g, b, 0, stream_>>>(a, n0, n0 + c, df, du, dr, dp, drr, drb, (double *)d_rho_);
        } else {
            auto a = make_cg_args<float>(prm_, visc_, u_i_, nullptr, pops_[cur_], nbr_.view(), d_col_, d_psi_, d_cs_, d_u_,
                                         d_rho_, N, geo_.P);
            cg_load_state_kernel<float><<<g, b, 0, stream_>>>(a, n0, n0 + c, df, du, dr, dp, drr, drb, (float *)d_rho_);
        }
        e = hipGetLastError();
    }
    const hipError_t es = hipStreamSynchronize(stream_);
    LBM_HIP_TRY(e);
    LBM_HIP_TRY(es);
    mode_ = STATE_DIRECT;
    steps_ = 0;
    return LBM_OK;
}

int CgSolver::set_body_force(const double g[3])
{
    if (!g) return fail(LBM_ERR_ARG, "null force");
    for (int c = 0; c < 3; ++c) prm_.g[c] = g[c];
    return LBM_OK;
}

template <typename ST, int NB>
static void cg_launch_pull_nb(const CgArgs<ST> &a, bool store_u, hipStream_t s)
{
    const dim3 g((unsigned)((a.N + NB - 1) / NB)), b(NB);
    store_u ? cg_collide_kernel<ST, K_PULL, true, NB><<<g, b, 0, s>>>(a)
            : cg_collide_kernel<ST, K_PULL, false, NB><<<g, b, 0, s>>>(a);
}

template <typename ST, int SLOTS>
static void cg_launch_pull_staged(const CgArgs<ST> &a, bool store_u, int64_t cw_grid, hipStream_t s)
{
    const dim3 g((unsigned)cw_grid), b(CW_BLOCK);
    store_u ? cg_collide_kernel<ST, K_PULL, true, CW_BLOCK, SLOTS><<<g, b, 0, s>>>(a)
            : cg_collide_kernel<ST, K_PULL, false, CW_BLOCK, SLOTS><<<g, b, 0, s>>>(a);
}

template <typename ST>
static hipError_t cg_launch_collide(CgArgs<ST> a, int mode, bool store_u, int64_t grid, int64_t cw_grid,
                                    const lbm_launch &ls, const void *wins, int slots, hipStream_t s)
{
    dim3 g((unsigned)grid), b(BLOCK);
    a.order = ls.order;
    a.wins = (const ColourWin *)wins;
    if (mode == K_PULL && ls.variant == LBM_VARIANT_STAGED && slots == 1024)
        cg_launch_pull_staged<ST, 1024>(a, store_u, cw_grid, s);
    else if (mode == K_PULL && ls.variant == LBM_VARIANT_STAGED && slots == 1536)
        cg_launch_pull_staged<ST, 1536>(a, store_u, cw_grid, s);
    else if (mode == K_PULL && ls.variant == LBM_VARIANT_STAGED && slots == 2560)
        cg_launch_pull_staged<ST, 2560>(a, store_u, cw_grid, s);
    else if (mode == K_INIT)
        store_u ? cg_collide_kernel<ST, K_INIT, true><<<g, b, 0, s>>>(a) : cg_collide_kernel<ST, K_INIT, false><<<g, b, 0, s>>>(a);
    else if (mode == K_DIRECT)
        cg_collide_kernel<ST, K_DIRECT, false><<<g, b, 0, s>>>(a);
    else if (ls.block == 128)
        cg_launch_pull_nb<ST, 128>(a, store_u, s);
    else if (ls.block == 512)
        cg_launch_pull_nb<ST, 512>(a, store_u, s);
    else
        store_u ? cg_collide_kernel<ST, K_PULL, true><<<g, b, 0, s>>>(a) : cg_collide_kernel<ST, K_PULL, false><<<g, b, 0, s>>>(a);
    return hipGetLastError();
}

// the collide launch from buffer src to dst
hipError_t CgSolver::launch_collide(int mode, bool store_u, const lbm_launch &ls, const void *src, void *dst)
{
    const int64_t N = geo_.N, P = geo_.P;
    if (prm_.precision == LBM_FP64)
        return cg_launch_collide(make_cg_args<double>(prm_, visc_, u_i_, src, dst, nbr_.view(), d_col_, d_psi_, d_cs_, d_u_,
                                                      d_rho_, N, P),
                                 mode, store_u, grid_, cw_grid_, ls, d_cwin_, cwin_slots_, stream_);
    return cg_launch_collide(make_cg_args<float>(prm_, visc_, u_i_, src, dst, nbr_.view(), d_col_, d_psi_, d_cs_, d_u_,
                                                 d_rho_, N, P),
                             mode, store_u, grid_, cw_grid_, ls, d_cwin_, cwin_slots_, stream_);
}

// the colour-streaming launch: gathers, or (FP32) the LDS-staged variant
hipError_t CgSolver::launch_colour(const lbm_launch &ls)
{
    const int64_t N = geo_.N, P = geo_.P;
    const dim3 g((unsigned)grid_), b(BLOCK);
    const dim3 gw((unsigned)cw_grid_), bw(CW_BLOCK);
    const ColourWin *w = (const ColourWin *)d_cwin_;
    if (prm_.precision == LBM_FP64) {
        const double *cs = (const double *)d_cs_;
        double *col = (double *)d_col_, *psi = (double *)d_psi_;
        if (ls.variant == LBM_VARIANT_STAGED && cwin_slots_ == 1024)
            cg_colour_kernel_win64<1024><<<gw, bw, 0, stream_>>>(nbr_.view(), cs, w, col, psi, N, P, ls.order);
        else if (ls.variant == LBM_VARIANT_STAGED && cwin_slots_ == 1536)
            cg_colour_kernel_win64<1536><<<gw, bw, 0, stream_>>>(nbr_.view(), cs, w, col, psi, N, P, ls.order);
        else if (ls.variant == LBM_VARIANT_STAGED && cwin_slots_ == 2560)
            cg_colour_kernel_win64<2560><<<gw, bw, 0, stream_>>>(nbr_.view(), cs, w, col, psi, N, P, ls.order);
        else
            cg_colour_kernel64<<<g, b, 0, stream_>>>(nbr_.view(), cs, col, psi, N, P, ls.order);
        return hipGetLastError();
    }
    const float *cs = (const float *)d_cs_;
    int32_t *col = (int32_t *)d_col_;
    float *psi = (float *)d_psi_;
    if (ls.variant == LBM_VARIANT_STAGED && cwin_slots_ == 1024)
        cg_colour_kernel_win<1024><<<gw, bw, 0, stream_>>>(nbr_.view(), cs, w, col, psi, N, P, ls.order);
    else if (ls.variant == LBM_VARIANT_STAGED && cwin_slots_ == 1536)
        cg_colour_kernel_win<1536><<<gw, bw, 0, stream_>>>(nbr_.view(), cs, w, col, psi, N, P, ls.order);
    else if (ls.variant == LBM_VARIANT_STAGED && cwin_slots_ == 2560)
        cg_colour_kernel_win<2560><<<gw, bw, 0, stream_>>>(nbr_.view(), cs, w, col, psi, N, P, ls.order);
    else
        cg_colour_kernel_fx<<<g, b, 0, stream_>>>(nbr_.view(), cs, col, psi, N, P, ls.order);
    return hipGetLastError();
}

int CgSolver::launch_step(bool store_u)
{
    const int mode = cg_kmode(mode_);
    const int nxt = cur_ ^ 1;
    LBM_HIP_TRY(launch_collide(mode, store_u, launch_collide_, pops_[cur_], pops_[nxt]));
    LBM_HIP_TRY(launch_colour(launch_colour_));
    cur_ = nxt;
    mode_ = STATE_PULL;
    ++steps_;
    return LBM_OK;
}

// Time the collide and colour launches' settings on this geometry.  The timing runs on a
// synthetic state (a random-looking +-1 phase field, one init launch); the handle has no state
// yet at create, and init or load_state sets every field the timing launches touch.
int CgSolver::autotune()
{
    if (geo_.N < TUNE_MIN_NODES) return LBM_OK;
    lbm_launch c1, c2;
    if (tune_cache_get(LBM_KERNEL_CG_COLLIDE, prm_.precision, geo_.N, c1) &&
        tune_cache_get(LBM_KERNEL_CG_COLOUR, prm_.precision, geo_.N, c2)) {
        launch_collide_ = c1;
        launch_colour_ = c2;
        return LBM_OK;
    }
    {
        const int64_t N = geo_.N, chunk = std::min<int64_t>(N, stage_chunk(STAGE_BYTES, 1));
        LBM_HIP_TRY(stage_.ensure());
        std::vector<double> pn((size_t)chunk);
        for (int64_t n0 = 0; n0 < N; n0 += chunk) {
            const int64_t c = std::min(chunk, N - n0);
            for (int64_t n = 0; n < c; ++n) pn[(size_t)n] = (((uint64_t)(n0 + n) * 2654435761u) >> 16) & 1 ? 1.0 : -1.0;
            LBM_HIP_TRY(hipMemcpyAsync(stage_.buf, pn.data(), (size_t)c * sizeof(double), hipMemcpyHostToDevice, stream_));
            dim3 g((unsigned)((c + BLOCK - 1) / BLOCK)), b(BLOCK);
            if (prm_.precision == LBM_FP64)
                cg_init_colour_kernel<double><<<g, b, 0, stream_>>>((const double *)stage_.buf, (double *)d_col_,
                                                                     (double *)d_psi_, n0, n0 + c, geo_.P);
            else
                cg_init_colour_kernel<float><<<g, b, 0, stream_>>>((const double *)stage_.buf, (int32_t *)d_col_,
                                                                    (float *)d_psi_, n0, n0 + c, geo_.P);
            LBM_HIP_TRY(hipGetLastError());
            LBM_HIP_TRY(hipStreamSynchronize(stream_));
        }
    }
    // each candidate is timed inside a step, collide then colour launch, as the two share the
    // caches in the run: timed alone, a kernel finds its inputs still cached from its own last
    // launch
    void *a = pops_[0], *b = pops_[1];
    LBM_HIP_TRY(launch_collide(K_INIT, false, launch_collide_, b, a));
    LBM_HIP_TRY(launch_colour(launch_colour_));
    lbm_launch best;
    const std::vector<int> staged{LBM_VARIANT_GATHER, LBM_VARIANT_STAGED};
    auto cc = tune_candidates(launch_collide_, {128, 256, 512}, {ORDER_HW, ORDER_RUNS8, ORDER_SLAB}, {0});
    for (const lbm_launch &l : tune_candidates(launch_collide_, {CW_BLOCK}, {ORDER_HW, ORDER_RUNS8, ORDER_SLAB}, staged))
        if (l.variant == LBM_VARIANT_STAGED) cc.push_back(l);
    LBM_HIP_TRY(tune_launch(stream_, cc,
                            [&](const lbm_launch &ls) {
                                hipError_t e = launch_collide(K_PULL, false, ls, a, b);
                                return e == hipSuccess ? launch_colour(launch_colour_) : e;
                            },
                            best));
    launch_collide_ = best;
    auto cv = tune_candidates(launch_colour_, {BLOCK}, {ORDER_HW, ORDER_RUNS8, ORDER_SLAB}, {LBM_VARIANT_GATHER});
    for (const lbm_launch &l : tune_candidates(launch_colour_, {CW_BLOCK}, {ORDER_HW, ORDER_RUNS8, ORDER_SLAB}, staged))
        if (l.variant == LBM_VARIANT_STAGED) cv.push_back(l);
    LBM_HIP_TRY(tune_launch(stream_, cv,
                            [&](const lbm_launch &ls) {
                                hipError_t e = launch_collide(K_PULL, false, launch_collide_, a, b);
                                return e == hipSuccess ? launch_colour(ls) : e;
                            },
                            best));
    launch_colour_ = best;
    tune_cache_put(LBM_KERNEL_CG_COLLIDE, prm_.precision, geo_.N, launch_collide_);
    tune_cache_put(LBM_KERNEL_CG_COLOUR, prm_.precision, geo_.N, launch_colour_);
    return LBM_OK;
}

int CgSolver::get_launch(int kernel, lbm_launch *out) const
{
    if (!out) return LBM_ERR_ARG;
    if (kernel == LBM_KERNEL_CG_COLLIDE) *out = launch_collide_;
    else if (kernel == LBM_KERNEL_CG_COLOUR) *out = launch_colour_;
    else return LBM_ERR_ARG;
    return LBM_OK;
}

int CgSolver::set_launch(int kernel, const lbm_launch *in)
{
    if (!in) return fail(LBM_ERR_ARG, "null launch settings");
    lbm_launch l = *in;
    l.tuned = 0;
    l.time_us = 0.0;
    if (kernel == LBM_KERNEL_CG_COLLIDE) {
        if (!(launch_valid(l, {128, 256, 512}, {0}) || launch_valid(l, {CW_BLOCK}, {1})))
            return fail(LBM_ERR_ARG, "collide launch: block 128, 256 or 512, order 0-2, variant 0; or block 512, variant 1");
        launch_collide_ = l;
    } else if (kernel == LBM_KERNEL_CG_COLOUR) {
        if (!(launch_valid(l, {BLOCK}, {0}) || launch_valid(l, {CW_BLOCK}, {1})))
            return fail(LBM_ERR_ARG, "colour launch: block 256, order 0-2, variant 0; or block 512, variant 1");
        launch_colour_ = l;
    } else {
        return fail(LBM_ERR_ARG, "a colour-gradient handle launches LBM_KERNEL_CG_COLLIDE and LBM_KERNEL_CG_COLOUR");
    }
    return LBM_OK;
}

int CgSolver::step(int nsteps, double *elapsed_ms)
{
    if (nsteps < 0) return fail(LBM_ERR_ARG, "negative step count");
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised: call lbm_cg_init or lbm_cg_load_state");
    LBM_HIP_TRY(hipSetDevice(device_));
    if (elapsed_ms) LBM_HIP_TRY(hipEventRecord(ev0_, stream_));
    for (int s = 0; s < nsteps; ++s) {
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

int CgSolver::get_state(double *f, double *u, double *rho, double *psi, double *rr, double *rb)
{
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    double *outs[6] = {f, u, rho, psi, rr, rb};
    const int width[6] = {Q, 3, 1, 1, 1, 1};
    int per_node = 0;
    for (int k = 0; k < 6; ++k) per_node += outs[k] ? width[k] : 0;
    if (per_node == 0) return LBM_OK;
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(stage_.ensure());
    const int64_t N = geo_.N, chunk = stage_chunk(STAGE_BYTES, per_node);
    const int mode = cg_kmode(mode_);
    hipError_t e = hipSuccess;
    for (int64_t n0 = 0; n0 < N && e == hipSuccess; n0 += chunk) {
        const int64_t c = std::min(chunk, N - n0), n1 = n0 + c;
        double *bufs[6] = {nullptr};
        double *p = (double *)stage_.buf;
        for (int k = 0; k < 6; ++k)
            if (outs[k]) bufs[k] = p, p += c * width[k];
        dim3 g((unsigned)((n1 - group_floor(n0) + BLOCK - 1) / BLOCK)), b(BLOCK);
        if (prm_.precision == LBM_FP64) {
            auto a = make_cg_args<double>(prm_, visc_, u_i_, pops_[cur_], nullptr, nbr_.view(), d_col_, d_psi_, d_cs_, d_u_,
                                          d_rho_, N, geo_.P);
            if (mode == K_INIT)
                cg_get_state_kernel<double, K_INIT><<<g, b, 0, stream_>>>(a, n0, n1, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4], bufs[5]);
            else if (mode == K_DIRECT)
                cg_get_state_kernel<double, K_DIRECT><<<g, b, 0, stream_>>>(a, n0, n1, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4], bufs[5]);
            else
                cg_get_state_kernel<double, K_PULL><<<g, b, 0, stream_>>>(a, n0, n1, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4], bufs[5]);
        } else {
            auto a = make_cg_args<float>(prm_, visc_, u_i_, pops_[cur_], nullptr, nbr_.view(), d_col_, d_psi_, d_cs_, d_u_,
                                         d_rho_, N, geo_.P);
            if (mode == K_INIT)
                cg_get_state_kernel<float, K_INIT><<<g, b, 0, stream_>>>(a, n0, n1, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4], bufs[5]);
            else if (mode == K_DIRECT)
                cg_get_state_kernel<float, K_DIRECT><<<g, b, 0, stream_>>>(a, n0, n1, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4], bufs[5]);
            else
                cg_get_state_kernel<float, K_PULL><<<g, b, 0, stream_>>>(a, n0, n1, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4], bufs[5]);
        }
        e = hipGetLastError();
        for (int k = 0; k < 6 && e == hipSuccess; ++k)
            if (outs[k])
                e = hipMemcpyAsync(outs[k] + n0 * width[k], bufs[k], (size_t)c * width[k] * sizeof(double),
                                   hipMemcpyDeviceToHost, stream_);
    }
    const hipError_t es = hipStreamSynchronize(stream_);
    LBM_HIP_TRY(e);
    LBM_HIP_TRY(es);
    return LBM_OK;
}

int CgSolver::diagnostics(const int32_t *box, lbm_cg_diag *out)
{
    if (!out) return fail(LBM_ERR_ARG, "null output");
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    LBM_HIP_TRY(hipSetDevice(device_));
    CgBox cb{};
    cb.ny = geo_.ny;
    cb.nz = geo_.nz;
    cb.perdir = prm_.perm_dir;
    if (box) {
        cb.on = 1;
        cb.xn = box[0];
        cb.xp = box[1];
        cb.yn = box[2];
        cb.yp = box[3];
        cb.zn = box[4];
        cb.zp = box[5];
    }
    const int mode = cg_kmode(mode_);
    dim3 g((unsigned)grid_), b(BLOCK);
    const int64_t N = geo_.N;
    if (prm_.precision == LBM_FP64) {
        auto a = make_cg_args<double>(prm_, visc_, u_i_, pops_[cur_], nullptr, nbr_.view(), d_col_, d_psi_, d_cs_, d_u_,
                                      d_rho_, N, geo_.P);
        if (mode == K_INIT)
            cg_diag_kernel<double, K_INIT><<<g, b, 0, stream_>>>(a, d_supinv_, cb, d_partial_);
        else if (mode == K_DIRECT)
            cg_diag_kernel<double, K_DIRECT><<<g, b, 0, stream_>>>(a, d_supinv_, cb, d_partial_);
        else
            cg_diag_kernel<double, K_PULL><<<g, b, 0, stream_>>>(a, d_supinv_, cb, d_partial_);
    } else {
        auto a = make_cg_args<float>(prm_, visc_, u_i_, pops_[cur_], nullptr, nbr_.view(), d_col_, d_psi_, d_cs_, d_u_,
                                     d_rho_, N, geo_.P);
        if (mode == K_INIT)
            cg_diag_kernel<float, K_INIT><<<g, b, 0, stream_>>>(a, d_supinv_, cb, d_partial_);
        else if (mode == K_DIRECT)
            cg_diag_kernel<float, K_DIRECT><<<g, b, 0, stream_>>>(a, d_supinv_, cb, d_partial_);
        else
            cg_diag_kernel<float, K_PULL><<<g, b, 0, stream_>>>(a, d_supinv_, cb, d_partial_);
    }
    LBM_HIP_TRY(hipGetLastError());
    reduce_partials_kernel<CG_NDIAG><<<1, BLOCK, 0, stream_>>>(d_partial_, grid_, CG_MAX_MASK, d_result_);
    LBM_HIP_TRY(hipGetLastError());
    LBM_HIP_TRY(hipMemcpyAsync(h_result_, d_result_, CG_NDIAG * sizeof(double), hipMemcpyDeviceToHost, stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    const double *r = h_result_;
    for (int c = 0; c < 3; ++c) {
        out->Ql[c] = r[c];
        out->Qg[c] = r[3 + c];
    }
    out->err = std::sqrt(r[6]) / (std::sqrt(r[7]) + 1e-30);
    out->sum_speed = r[8];
    out->umax = std::sqrt(r[9]);
    out->nan_count = r[10];
    out->wet_count = r[11];
    out->count_in_box = r[12];
    out->sd_l = r[13];
    out->sd_g = r[14];
    return LBM_OK;
}

}  // namespace lbm
