// connectivity.h -- does a cluster of pore nodes wrap round the periodic box along an axis?
//
// Union-find over the 18 lattice links with each node's displacement (along the axis) from its
// set's root.  Linking node n to its upstream node m = n - e_q fixes pos(n) - pos(m) = e_q; a
// link inside one set whose displacements disagree closes a loop that goes round the box along
// the axis (its displacement is a nonzero multiple of the box length).  Without such a loop the
// mean flux of a divergence-free field with no-penetration walls is zero along the axis.
#pragma once
#include <cstdint>
#include <vector>

#include "lattice.h"

namespace lbm {

// up(n, q) is node n's upstream node in direction q (1..18), or -1 for a solid site.
template <typename Up>
bool lattice_loop_wraps(int64_t N, int axis, Up up)
{
    const int *E = axis == 0 ? EX : (axis == 1 ? EY : EZ);
    // parent and offset of each node side by side (one cache line serves both); displacements
    // within a set are bounded by the node count, so 32 bits hold them
    std::vector<int32_t> pd((size_t)(2 * N));
    for (int64_t n = 0; n < N; ++n) {
        pd[(size_t)(2 * n)] = (int32_t)n;
        pd[(size_t)(2 * n + 1)] = 0;   // pos(n) - pos(parent)
    }
    // root of n and pos(n) - pos(root), halving the path on the way (each visited node is
    // re-pointed at its grandparent)
    auto find = [&pd](int32_t n, int32_t &d) {
        d = 0;
        for (;;) {
            const int32_t p = pd[(size_t)(2 * n)];
            if (p == n) return n;
            const int32_t gp = pd[(size_t)(2 * p)];
            const int32_t dp = pd[(size_t)(2 * n + 1)], dg = pd[(size_t)(2 * p + 1)];
            pd[(size_t)(2 * n)] = gp;
            pd[(size_t)(2 * n + 1)] = dp + dg;
            d += dp + dg;
            n = gp;
        }
    };
    int qs[9], nq = 0;
    for (int q = 1; q < Q; ++q)
        if (opp(q) > q) qs[nq++] = q;   // each link once: the reverse direction joins the same pair
    for (int64_t n = 0; n < N; ++n)
        for (int t = 0; t < nq; ++t) {
            const int q = qs[t];
            const int64_t m = up(n, q);
            if (m < 0) continue;
            int32_t dn, dm;
            const int32_t rn = find((int32_t)n, dn), rm = find((int32_t)m, dm);
            if (rn == rm) {
                if (dn - dm != E[q]) return true;
                continue;
            }
            // pos(rm) - pos(rn) = (pos(m) - dm) - (pos(n) - dn) = dn - dm - e_q; the later root
            // goes under the earlier one, which keeps the trees shallow in this scan order
            const int32_t d = dn - dm - E[q];
            if (rn < rm) {
                pd[(size_t)(2 * rm)] = rn;
                pd[(size_t)(2 * rm + 1)] = d;
            } else {
                pd[(size_t)(2 * rn)] = rm;
                pd[(size_t)(2 * rn + 1)] = -d;
            }
        }
    return false;
}

}  // namespace lbm
