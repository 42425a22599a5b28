// collide.h -- per-node MRT collision arithmetic for the device kernels.
//
// Populations are handled as deviations h_q = f_q - w_q*rho0 from the rest state.  The
// moment transform is linear, so the collision is carried out on deviation moments
// (M h = M f - rho0 * M w); only the three rest-state moments (SP basis rows 0, 1, 2 =
// 1, -11, 3) differ, and they are handled in the equilibria below.
//
// Instead of the reference's dense 19x19 products (m = M f, F = M GuoF, meq = M feq and
// f* = M^-1 m*, SP:1642-1870), the transforms exploit the basis structure:
//   * the nine opposite pairs give s = h_a + h_b and d = h_a - h_b; even basis rows only see
//     s, odd rows only d, which halves the work of both M and M^-1 = M^T diag(1/|row|^2);
//   * M feq and M GuoF are replaced by their closed forms in (rho, u, g), which the
//     reference's per-direction feq (SP:1386-1396) and Guo term (SP:1519-1593) reduce to
//     for D3Q19 (the tests check the identities).
#pragma once
#include <hip/hip_runtime.h>

#include "lattice.h"

namespace lbm {

// m = M h (d'Humieres basis), zero coefficients skipped at compile time
template <typename T>
__device__ __forceinline__ void moments_fwd(const T (&h)[Q], T (&m)[Q])
{
    T s[NPAIR], d[NPAIR];
#pragma unroll
    for (int p = 0; p < NPAIR; ++p) {
        s[p] = h[PA[p]] + h[opp(PA[p])];
        d[p] = h[PA[p]] - h[opp(PA[p])];
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        T acc = T(-0.0);   // -0.0 is the additive identity the compiler may fold away
        if (even_row(k)) {
            if (basis(k, 0) != 0) acc += T(basis(k, 0)) * h[0];
#pragma unroll
            for (int p = 0; p < NPAIR; ++p)
                if (basis(k, PA[p]) != 0) acc += T(basis(k, PA[p])) * s[p];
        } else {
#pragma unroll
            for (int p = 0; p < NPAIR; ++p)
                if (basis(k, PA[p]) != 0) acc += T(basis(k, PA[p])) * d[p];
        }
        m[k] = acc;
    }
}

// h = M^-1 m given mt_k = m_k/|row_k|^2 (M^-1 = M^T diag(1/|row|^2))
template <typename T>
__device__ __forceinline__ void moments_inv(const T (&mt)[Q], T (&h)[Q])
{
    T h0 = T(-0.0);
#pragma unroll
    for (int k = 0; k < Q; ++k)
        if (even_row(k) && basis(k, 0) != 0) h0 += T(basis(k, 0)) * mt[k];
    h[0] = h0;
#pragma unroll
    for (int p = 0; p < NPAIR; ++p) {
        T ev = T(-0.0), od = T(-0.0);
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const int c = basis(k, PA[p]);
            if (c == 0) continue;
            if (even_row(k))
                ev += T(c) * mt[k];
            else
                od += T(c) * mt[k];
        }
        h[PA[p]] = ev + od;
        h[opp(PA[p])] = ev - od;
    }
}

template <typename T>
struct Macro {
    T drho, rho, ux, uy, uz;
};

// density deviation and velocity u = (sum e f + g/2)/rho from deviation moments
// (comput_macro_variables, SP:2713-2741)
template <typename T>
__device__ __forceinline__ Macro<T> macro_from_moments(const T (&m)[Q], T rho0, T gx, T gy, T gz)
{
    Macro<T> r;
    r.drho = m[0];
    r.rho = rho0 + m[0];
    const T inv = T(1) / r.rho;
    r.ux = (m[3] + T(0.5) * gx) * inv;
    r.uy = (m[5] + T(0.5) * gy) * inv;
    r.uz = (m[7] + T(0.5) * gz) * inv;
    return r;
}

// Single-phase MRT collision in moment space (SP:1642-1829).  On entry m holds the deviation
// moments M h of the pre-collision populations; on exit h holds the post-collision deviations.
template <typename T>
__device__ __forceinline__ void sp_relax(const T (&m)[Q], const Macro<T> &mc, T gx, T gy, T gz, T sv, T sq,
                                         T (&h)[Q])
{
    const T ux = mc.ux, uy = mc.uy, uz = mc.uz;
    const T rux = mc.rho * ux, ruy = mc.rho * uy, ruz = mc.rho * uz;
    const T ru2 = rux * ux + ruy * uy + ruz * uz;
    const T t9 = T(2) * rux * ux - ruy * uy - ruz * uz;
    const T t11 = ruy * uy - ruz * uz;
    // equilibrium moments of feq = w rho (1 + 3e.u + 4.5(e.u)^2 - 1.5u^2), as deviations
    T eq[Q];
    eq[0] = mc.drho;
    eq[1] = T(-11) * mc.drho + T(19) * ru2;
    eq[2] = T(3) * mc.drho - T(5.5) * ru2;
    eq[3] = rux;
    eq[4] = T(-2.0 / 3.0) * rux;
    eq[5] = ruy;
    eq[6] = T(-2.0 / 3.0) * ruy;
    eq[7] = ruz;
    eq[8] = T(-2.0 / 3.0) * ruz;
    eq[9] = t9;
    eq[10] = T(-0.5) * t9;
    eq[11] = t11;
    eq[12] = T(-0.5) * t11;
    eq[13] = rux * uy;
    eq[14] = ruy * uz;
    eq[15] = rux * uz;
    eq[16] = eq[17] = eq[18] = T(0);
    // Guo forcing moments M GuoF
    const T ug = ux * gx + uy * gy + uz * gz;
    const T f9 = T(2) * (T(2) * ux * gx - uy * gy - uz * gz);
    const T f11 = T(2) * (uy * gy - uz * gz);
    T F[Q];
    F[0] = T(0);
    F[1] = T(38) * ug;
    F[2] = T(-11) * ug;
    F[3] = gx;
    F[4] = T(-2.0 / 3.0) * gx;
    F[5] = gy;
    F[6] = T(-2.0 / 3.0) * gy;
    F[7] = gz;
    F[8] = T(-2.0 / 3.0) * gz;
    F[9] = f9;
    F[10] = T(-0.5) * f9;
    F[11] = f11;
    F[12] = T(-0.5) * f11;
    F[13] = ux * gy + uy * gx;
    F[14] = uy * gz + uz * gy;
    F[15] = ux * gz + uz * gx;
    F[16] = F[17] = F[18] = T(0);
    // m* = m - S (m - meq) + (1 - S/2) F, then scaled by 1/|row|^2 for the inverse
    const T av = T(1) - sv, cv = T(1) - T(0.5) * sv;
    const T aq = T(1) - sq, cq = T(1) - T(0.5) * sq;
    T mt[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        T ms;
        if (RCLASS[k] == 0)
            ms = m[k] + F[k];
        else if (RCLASS[k] == 1)
            ms = av * m[k] + sv * eq[k] + cv * F[k];
        else if (k >= 16)
            ms = aq * m[k];
        else
            ms = aq * m[k] + sq * eq[k] + cq * F[k];
        mt[k] = ms * T(1.0 / norm2(k));
    }
    moments_inv(mt, h);
}

// equilibrium deviation moments for the init step (f = feq(rho_i, u_i), SP:1343-1376)
template <typename T>
__device__ __forceinline__ void sp_eq_moments(const Macro<T> &mc, T (&m)[Q])
{
    const T ux = mc.ux, uy = mc.uy, uz = mc.uz;
    const T rux = mc.rho * ux, ruy = mc.rho * uy, ruz = mc.rho * uz;
    const T ru2 = rux * ux + ruy * uy + ruz * uz;
    const T t9 = T(2) * rux * ux - ruy * uy - ruz * uz;
    const T t11 = ruy * uy - ruz * uz;
    m[0] = mc.drho;
    m[1] = T(-11) * mc.drho + T(19) * ru2;
    m[2] = T(3) * mc.drho - T(5.5) * ru2;
    m[3] = rux;
    m[4] = T(-2.0 / 3.0) * rux;
    m[5] = ruy;
    m[6] = T(-2.0 / 3.0) * ruy;
    m[7] = ruz;
    m[8] = T(-2.0 / 3.0) * ruz;
    m[9] = t9;
    m[10] = T(-0.5) * t9;
    m[11] = t11;
    m[12] = T(-0.5) * t11;
    m[13] = rux * uy;
    m[14] = ruy * uz;
    m[15] = rux * uz;
    m[16] = m[17] = m[18] = T(0);
}

// ----------------------------------------------------------------------------- colour gradient
// The colour-gradient solver collides in its own moment basis (CG:80-98), which spans the same
// polynomial spaces as the SP basis: c = T s with s = M_SP h.  T is block diagonal with blocks
// on (0,1,2), (3,4), (5,6), (7,8), (9,10), (11,12) and the identity elsewhere.
template <typename T>
__device__ __forceinline__ void cg_from_sp(const T (&s)[Q], T (&c)[Q])
{
#pragma unroll
    for (int k = 0; k < Q; ++k) c[k] = s[k];
    c[1] = (T(11) * s[0] + s[1]) * T(1.0 / 19.0);
    c[2] = s[0] * T(1.0 / 19.0) + s[1] * T(11.0 / 133.0) + s[2] * T(2.0 / 7.0);
    c[4] = (T(2) * s[3] + T(3) * s[4]) * T(0.2);
    c[6] = (T(2) * s[5] + T(3) * s[6]) * T(0.2);
    c[8] = (T(2) * s[7] + T(3) * s[8]) * T(0.2);
    c[10] = (s[9] + T(2) * s[10]) * T(1.0 / 3.0);
    c[12] = (s[11] + T(2) * s[12]) * T(1.0 / 3.0);
}

template <typename T>
__device__ __forceinline__ void sp_from_cg(const T (&c)[Q], T (&s)[Q])
{
#pragma unroll
    for (int k = 0; k < Q; ++k) s[k] = c[k];
    s[1] = T(19) * c[1] - T(11) * c[0];
    s[2] = T(3.5) * (c[2] - c[0] * T(1.0 / 19.0) - s[1] * T(11.0 / 133.0));
    s[4] = (T(5) * c[4] - T(2) * c[3]) * T(1.0 / 3.0);
    s[6] = (T(5) * c[6] - T(2) * c[5]) * T(1.0 / 3.0);
    s[8] = (T(5) * c[8] - T(2) * c[7]) * T(1.0 / 3.0);
    s[10] = (T(3) * c[10] - c[9]) * T(0.5);
    s[12] = (T(3) * c[12] - c[11]) * T(0.5);
}

// Colour-gradient MRT collision (CG:3137-3400).  m: SP-basis deviation moments of the
// pre-collision populations; mc: rho and u; g: the Guo force after the per-phase gate
// (CG:3040); sv: the psi-interpolated shear rate (CG:3163-3176); capa_cc = CapA*|C| and n =
// C/|C| (zero when C = 0) give the surface-tension terms of the equilibrium (CG:3137-3143).
template <typename T>
__device__ __forceinline__ void cg_relax(const T (&m)[Q], const Macro<T> &mc, T gx, T gy, T gz, T sv, T capa_cc,
                                         T nx, T ny, T nz, T (&h)[Q])
{
    const T ux = mc.ux, uy = mc.uy, uz = mc.uz;
    const T sq = T(8) * (T(2) - sv) / (T(8) - sv);
    T c[Q];
    cg_from_sp(m, c);
    // equilibrium in the CG basis with rho_0 = 1 in the velocity terms
    const T hc = T(0.5) * capa_cc;
    T eq[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) eq[k] = T(0);
    eq[0] = mc.drho;
    eq[1] = ux * ux + uy * uy + uz * uz + capa_cc;
    eq[3] = ux;
    eq[5] = uy;
    eq[7] = uz;
    eq[9] = T(2) * ux * ux - uy * uy - uz * uz + hc * (T(2) * nx * nx - ny * ny - nz * nz);
    eq[11] = uy * uy - uz * uz + hc * (ny * ny - nz * nz);
    eq[13] = ux * uy + hc * (nx * ny);
    eq[14] = uy * uz + hc * (ny * nz);
    eq[15] = ux * uz + hc * (nx * nz);
    // Guo forcing in the CG basis (T applied to the SP closed forms; rows 2,4,6,8,10,12 vanish)
    T F[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) F[k] = T(0);
    F[1] = T(2) * (ux * gx + uy * gy + uz * gz);
    F[3] = gx;
    F[5] = gy;
    F[7] = gz;
    F[9] = T(2) * (T(2) * ux * gx - uy * gy - uz * gz);
    F[11] = T(2) * (uy * gy - uz * gz);
    F[13] = ux * gy + uy * gx;
    F[14] = uy * gz + uz * gy;
    F[15] = ux * gz + uz * gx;
    const T av = T(1) - sv, cv = T(1) - T(0.5) * sv;
    const T aq = T(1) - sq, cq = T(1) - T(0.5) * sq;
    T cs[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        if (RCLASS[k] == 0)
            cs[k] = c[k] + F[k];
        else if (RCLASS[k] == 1)
            cs[k] = av * c[k] + sv * eq[k] + cv * F[k];
        else
            cs[k] = aq * c[k] + sq * eq[k] + cq * F[k];
    }
    T s[Q], mt[Q];
    sp_from_cg(cs, s);
#pragma unroll
    for (int k = 0; k < Q; ++k) mt[k] = s[k] * T(1.0 / norm2(k));
    moments_inv(mt, h);
}

}  // namespace lbm
