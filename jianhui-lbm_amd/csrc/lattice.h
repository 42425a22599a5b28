// lattice.h -- D3Q19 tables shared by the host and the device code.
//
// The device works in one canonical velocity ordering, the single-phase reference's
// (SP:157-159): rest, +-x, +-y, +-z, then the xy, yz and xz diagonals.  Colour-gradient
// states are permuted into it on load and back on read (the CG reference orders its
// diagonals differently, CG:203-205).
//
// Moments use d'Humieres' orthogonal D3Q19 basis (the rows of SP:56-75), written here as
// polynomials of the velocity; its inverse is M^T diag(1/|row|^2).  The colour-gradient
// basis (CG:80-98) spans the same polynomial spaces and is reached from it by the small
// block matrix T (cg_from_sp / sp_from_cg in collide.h).
#pragma once
#include <cstdint>

namespace lbm {

constexpr int Q = 19;
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

constexpr int opp(int q)
{
    for (int r = 0; r < Q; ++r)
        if (EX[r] == -EX[q] && EY[r] == -EY[q] && EZ[r] == -EZ[q]) return r;
    return -1;
}

// the nine opposite pairs (PA[p], opp(PA[p]))
constexpr int NPAIR = 9;
constexpr int PA[NPAIR] = {1, 3, 5, 7, 8, 11, 12, 15, 16};

// lattice weight as a rational number: 1/3, 1/18, 1/36
constexpr int e2_of(int q) { return EX[q] * EX[q] + EY[q] * EY[q] + EZ[q] * EZ[q]; }
constexpr double weight(int q) { return e2_of(q) == 0 ? 1.0 / 3.0 : (e2_of(q) == 1 ? 1.0 / 18.0 : 1.0 / 36.0); }

// d'Humieres basis row k evaluated at velocity q (integer valued)
constexpr int basis(int k, int q)
{
    const int x = EX[q], y = EY[q], z = EZ[q], e2 = e2_of(q);
    const int q9 = 3 * x * x - e2, q11 = y * y - z * z;
    switch (k) {
    case 0: return 1;
    case 1: return 19 * e2 - 30;
    case 2: return (21 * e2 * e2 - 53 * e2 + 24) / 2;
    case 3: return x;
    case 4: return (5 * e2 - 9) * x;
    case 5: return y;
    case 6: return (5 * e2 - 9) * y;
    case 7: return z;
    case 8: return (5 * e2 - 9) * z;
    case 9: return q9;
    case 10: return (3 * e2 - 5) * q9;
    case 11: return q11;
    case 12: return (3 * e2 - 5) * q11;
    case 13: return x * y;
    case 14: return y * z;
    case 15: return x * z;
    case 16: return (y * y - z * z) * x;
    case 17: return (z * z - x * x) * y;
    default: return (x * x - y * y) * z;
    }
}

constexpr int norm2(int k)
{
    int s = 0;
    for (int q = 0; q < Q; ++q) s += basis(k, q) * basis(k, q);
    return s;
}

// rows that are even in e (basis(k,-e) == basis(k,e)); the others are odd
constexpr bool even_row(int k) { return !((k >= 3 && k <= 8) || k >= 16); }

// relaxation class of each moment: 0 conserved, 1 s_v (shear/energy), 2 s_q (fluxes/ghosts)
// -- the S vector of SP:1310-1328 and CG:3184-3186
constexpr int RCLASS[Q] = {0, 1, 1, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2};

// colour-gradient reference ordering (CG:203-205) -> canonical index
constexpr int CG_EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CG_EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CG_EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int cg_to_canon(int qc)
{
    for (int r = 0; r < Q; ++r)
        if (EX[r] == CG_EX[qc] && EY[r] == CG_EY[qc] && EZ[r] == CG_EZ[qc]) return r;
    return -1;
}

// upstream-index encoding: bit 31 set = the upstream site is solid, read the opposite
// population of this node instead (halfway bounce-back)
constexpr uint32_t BB_FLAG = 0x80000000u;

static_assert(norm2(0) == 19 && norm2(1) == 2394 && norm2(2) == 252, "basis norms");
static_assert(opp(7) == 10 && opp(8) == 9 && opp(11) == 14 && opp(15) == 18, "opposites");

}  // namespace lbm
