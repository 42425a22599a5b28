// boundary.h -- device side of the pressure / velocity faces (tables built by boundary.cpp).
//
// A face node's post-streaming populations are replaced by its rule (SP:2751-3620); rules
// that need the inward neighbour read that neighbour's pre-collision state of the same
// reference step, which on the device is the state the previous launch pulled for it and
// saved.  Both lookups are per-wave bit masks read with scalar loads: waves without face or
// saved nodes -- nearly all of them -- take one uniform branch.
#pragma once
#include "collide.h"
#include "handle.h"
#include "lattice.h"

namespace lbm {

// f_q - w_q rho0 of the equilibrium (SP:1386-1396) for density rho = rho0 + drho
template <typename C>
__device__ __forceinline__ C feq_dev(int q, C rho, C drho, C ux, C uy, C uz)
{
    const C eu = C(EX[q]) * ux + C(EY[q]) * uy + C(EZ[q]) * uz;
    const C uu = ux * ux + uy * uy + uz * uz;
    return C(weight(q)) * (drho + rho * (C(3) * eu + C(4.5) * eu * eu - C(1.5) * uu));
}

// The rule of entry e (deviations h on return).  Every rule fills its own array and one
// final loop copies it, so that h stays in registers across the divergent rules.
template <typename C>
__device__ __forceinline__ void bc_apply(const BcDev &bc, const BcEntry &e, C rho0, C (&h)[Q])
{
    const int rule = e.rule;
    const C rb = (C)e.val, drb = (C)(e.val - (double)rho0);
    const C *__restrict__ S = (const C *)bc.s_prev;
    const int64_t sp = bc.sp, s = rule == BC_EQ ? 0 : e.slot;
    C ux, uy, uz, uzl, dn = C(0);
    C hn[Q];
    if (rule == BC_EQ) {
        ux = (C)e.uf[0];
        uy = (C)e.uf[1];
        uz = uzl = (C)e.uf[2];
#pragma unroll
        for (int q = 0; q < Q; ++q) hn[q] = C(0);
    } else {
        ux = S[20 * sp + s];
        uy = S[21 * sp + s];
        uz = S[22 * sp + s];
        uzl = e.quirk ? uy : uz;   // the x- face's u_ls[2] = u[nb][1] (SP:3254, 3613)
        dn = S[19 * sp + s];
#pragma unroll
        for (int q = 0; q < Q; ++q) hn[q] = S[q * sp + s];
    }
    C out[Q];
    if (rule == BC_MOM) {
        // M^-1 (M f_nb with m0 = rho_b); in deviations only m0 = rho - rho0 changes
        C m[Q];
        moments_fwd(hn, m);
        m[0] = drb;
#pragma unroll
        for (int k = 0; k < Q; ++k) m[k] = m[k] / C(norm2(k));
        moments_inv(m, out);
    } else {
        // BC_EQ / BC_EQU: feq(rho_b, u_ls); BC_NEQ: + f_nb - feq(rho_nb, u_nb)
        const bool neq = rule == BC_NEQ;
        const C rn = rho0 + dn;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const C b = feq_dev(q, rb, drb, ux, uy, uzl);
            out[q] = neq ? b + hn[q] - feq_dev(q, rn, dn, ux, uy, uz) : b;
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) h[q] = out[q];
}

// Replace h of node n if it is a face node (post-streaming state of a pull launch).
template <typename C>
__device__ __forceinline__ void bc_patch(const BcDev &bc, int64_t n, C rho0, C (&h)[Q])
{
    const uint32_t w = __builtin_amdgcn_readfirstlane((uint32_t)(n >> 6));
    const uint64_t bm = bc.wave[w].b;
    if (bm == 0) return;
    const uint32_t lane = (uint32_t)(n & 63);
    if (!((bm >> lane) & 1ull)) return;
    const int64_t e = (int64_t)bc.wave[w].bbase + __popcll(bm & ((1ull << lane) - 1ull));
    bc_apply(bc, bc.entry[e], rho0, h);
}

// Save the pre-collision state of node n if a face rule reads it next step.
template <typename C>
__device__ __forceinline__ void bc_save(const BcDev &bc, int64_t n, const C (&h)[Q], C drho, C ux, C uy, C uz)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane((uint32_t)(n >> 6));
    const uint64_t sm = bc.wave[w].s;
    if (sm == 0) return;
    const uint32_t lane = (uint32_t)(n & 63);
    if (!((sm >> lane) & 1ull)) return;
    const int64_t s = (int64_t)bc.wave[w].sbase + __popcll(sm & ((1ull << lane) - 1ull));
    C *__restrict__ S = (C *)bc.s_cur;
    const int64_t sp = bc.sp;
#pragma unroll
    for (int q = 0; q < Q; ++q) S[q * sp + s] = h[q];
    S[19 * sp + s] = drho;
    S[20 * sp + s] = ux;
    S[21 * sp + s] = uy;
    S[22 * sp + s] = uz;
}

// whether node n's state is saved (for callers that must build h first)
__device__ __forceinline__ bool bc_is_saved(const BcDev &bc, int64_t n)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane((uint32_t)(n >> 6));
    const uint64_t sm = bc.wave[w].s;
    return sm != 0 && ((sm >> (uint32_t)(n & 63)) & 1ull);
}

}  // namespace lbm
