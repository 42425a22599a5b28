// boundary.cpp -- host resolution of the pressure / velocity faces into per-node rules.
//
// The reference applies its faces after streaming, one face after the other in the order
// y+, y-, z+, z-, x+, x- (boundary_pressure, SP:3076-3620, then boundary_velocity,
// SP:2751-3075); a node on several faces keeps what the last one wrote.  Replaying those
// writes once on the host gives every face node one rule, so the device applies exactly one
// per node and step.
#include <algorithm>
#include <cstring>
#include <unordered_map>

#include "handle.h"

namespace lbm {

namespace {

// face index (x+, x-, y+, y-, z+, z-) in the order the reference visits them
constexpr int FACE_ORDER[6] = {2, 3, 4, 5, 0, 1};

struct Pending {
    BcEntry e;
    int64_t nb;   // inward neighbour node id, -1 if solid or unused
};

}  // namespace

int build_boundaries(const SparseGeometry &g, const lbm_sp_bc &bc, BoundaryTables &out, std::string &err)
{
    out = BoundaryTables();
    bool any_p = false, any_v = false;
    for (int f = 0; f < 6; ++f) {
        if ((bc.pre[f] != 0 && bc.pre[f] != 1) || (bc.vel[f] != 0 && bc.vel[f] != 1)) {
            err = "boundary face flags must be 0 or 1";
            return LBM_ERR_ARG;
        }
        any_p |= bc.pre[f] == 1;
        any_v |= bc.vel[f] == 1;
        if (bc.pre[f] == 1 && !(bc.p[f] > 0.0)) {
            err = "boundary density must be positive";
            return LBM_ERR_ARG;
        }
    }
    if (bc.sub_bc < 0 || bc.sub_bc > 3) {
        err = "Sub_BC must be 0, 1, 2 or 3";
        return LBM_ERR_ARG;
    }
    if (any_v && bc.sub_bc >= 2) {
        err = "velocity faces with Sub_BC 2 or 3 are not supported: the reference's x+ face indexes "
              "Solid[nx_l-1][k][j] (SP:2953, 3051) and Sub_BC 3 reads the density of solid neighbours";
        return LBM_ERR_ARG;
    }
    const int dim[3] = {g.nx, g.ny, g.nz};
    for (int f = 0; f < 6; ++f)
        if ((bc.pre[f] == 1 || bc.vel[f] == 1) && dim[f / 2] < 2) {
            err = "a boundary face needs at least 2 voxels along its axis";
            return LBM_ERR_ARG;
        }
    if (!any_p && !any_v) return LBM_OK;

    // dense voxel -> node map
    const int64_t vox = (int64_t)g.nx * g.ny * g.nz;
    std::vector<int32_t> node(vox, -1);
    for (int64_t n = 0; n < g.N; ++n) node[g.supinv[n]] = (int32_t)n;
    auto at = [&](int i, int j, int k) { return node[((int64_t)i * g.ny + j) * g.nz + k]; };

    std::unordered_map<int64_t, Pending> rule;
    rule.reserve((size_t)(2 * ((int64_t)g.nx * g.ny + (int64_t)g.ny * g.nz + (int64_t)g.nx * g.nz)));
    for (int pass = 0; pass < 2; ++pass) {            // pressure, then velocity
        const int32_t *on = pass == 0 ? bc.pre : bc.vel;
        for (int oi = 0; oi < 6; ++oi) {
            const int f = FACE_ORDER[oi];
            if (on[f] != 1) continue;
            const int ax = f / 2;
            const int plane = (f % 2 == 0) ? dim[ax] - 1 : 0;
            const int inward = (f % 2 == 0) ? plane - 1 : 1;
            const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
            for (int p1 = 0; p1 < dim[a1]; ++p1)
                for (int p2 = 0; p2 < dim[a2]; ++p2) {
                    int c[3], cn[3];
                    c[ax] = plane;
                    cn[ax] = inward;
                    c[a1] = cn[a1] = p1;
                    c[a2] = cn[a2] = p2;
                    const int32_t b = at(c[0], c[1], c[2]);
                    if (b < 0) continue;
                    const int32_t nb = at(cn[0], cn[1], cn[2]);
                    Pending pe{};
                    pe.nb = -1;
                    if (pass == 1) {   // velocity, Sub_BC 0 / 1: feq(1, v n)
                        pe.e.rule = BC_EQ;
                        pe.e.val = 1.0;
                        pe.e.uf[ax] = bc.v[f];
                    } else if (bc.sub_bc == 0 || nb < 0) {
                        // x faces of Sub_BC 2 with a solid neighbour: the reference writes the
                        // column's k = NZ node instead (SP:3451, 3475), which its own visit
                        // later overwrites, so this node keeps what it had (the k = NZ node
                        // itself does get the write)
                        if (bc.sub_bc == 2 && nb < 0 && ax == 0 && c[2] != g.nz - 1) continue;
                        pe.e.rule = BC_EQ;
                        pe.e.val = bc.p[f];
                    } else {
                        pe.e.rule = bc.sub_bc == 1 ? BC_EQU : (bc.sub_bc == 2 ? BC_MOM : BC_NEQ);
                        pe.e.val = bc.p[f];
                        pe.e.quirk = (f == 1 && bc.sub_bc != 2) ? 1 : 0;   // SP:3254, 3613
                        pe.nb = nb;
                    }
                    rule[b] = pe;
                }
        }
    }

    // entries in node order; slots for the neighbours that are read, in node order
    std::vector<int64_t> bnodes;
    bnodes.reserve(rule.size());
    for (const auto &kv : rule) bnodes.push_back(kv.first);
    std::sort(bnodes.begin(), bnodes.end());
    std::vector<int64_t> snodes;
    for (const auto &kv : rule)
        if (kv.second.nb >= 0) snodes.push_back(kv.second.nb);
    std::sort(snodes.begin(), snodes.end());
    snodes.erase(std::unique(snodes.begin(), snodes.end()), snodes.end());
    out.entry.resize(bnodes.size());
    for (size_t i = 0; i < bnodes.size(); ++i) {
        const Pending &pe = rule[bnodes[i]];
        out.entry[i] = pe.e;
        out.entry[i].slot = -1;
        if (pe.nb >= 0)
            out.entry[i].slot = (int32_t)(std::lower_bound(snodes.begin(), snodes.end(), pe.nb) - snodes.begin());
    }
    out.saved.assign(snodes.begin(), snodes.end());

    const int64_t nwave = g.P / 64;
    out.wave.assign((size_t)nwave, BcWave{0, 0, 0, 0});
    for (int64_t n : bnodes) out.wave[(size_t)(n >> 6)].b |= 1ull << (n & 63);
    for (int64_t n : snodes) out.wave[(size_t)(n >> 6)].s |= 1ull << (n & 63);
    uint32_t eb = 0, sb = 0;
    for (auto &w : out.wave) {
        w.bbase = eb;
        w.sbase = sb;
        eb += (uint32_t)__builtin_popcountll(w.b);
        sb += (uint32_t)__builtin_popcountll(w.s);
    }
    return LBM_OK;
}

}  // namespace lbm
