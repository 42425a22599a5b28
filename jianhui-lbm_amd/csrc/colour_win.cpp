// colour_win.cpp -- the LDS staging tables of the colour-gradient kernels' staged variant (host).
#include "colour_win.h"

#include <algorithm>

namespace lbm {

// The runs of node ids each block of the staged kernels copies to LDS: the block's own nodes and
// their upstream nodes, sorted, split where consecutive ids are more than RUN_GAP apart (ids in
// such a gap are copied too: a gap of 128 joins the far end of a z-line that a node at z = 0
// reads across the periodic wrap), and the runs with the most reads taken first while they fit
// in `slots` states and CW_RUNS runs.  Then the slot offsets of the blocks that read every state
// from these runs (ColourWin::fast).  Returns the fraction of such blocks.
double build_colour_windows(const SparseGeometry &g, int64_t grid, int slots, std::vector<ColourWin> &out)
{
    constexpr uint32_t RUN_GAP = 128;
    out.assign((size_t)grid, ColourWin{});
    int64_t nfast = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nfast)
    for (int64_t b = 0; b < grid; ++b) {
        std::vector<uint32_t> ids;
        ids.reserve((size_t)CW_BLOCK * Q);
        const int64_t n1 = std::min<int64_t>(g.N, (b + 1) * CW_BLOCK);
        for (int64_t n = b * CW_BLOCK; n < n1; ++n) {
            ids.push_back((uint32_t)n);
            for (int q = 1; q < Q; ++q) {
                const uint32_t v = g.nbr[(size_t)((q - 1) * g.P + n)];
                ids.push_back((v & BB_FLAG) ? (uint32_t)n : v);
            }
        }
        std::sort(ids.begin(), ids.end());
        struct Run {
            uint32_t lo, hi;
            int64_t reads;
        };
        std::vector<Run> runs;
        for (size_t i = 0; i < ids.size(); ++i) {
            if (runs.empty() || ids[i] > runs.back().hi - 1 + RUN_GAP)
                runs.push_back(Run{ids[i], ids[i] + 1, 0});
            runs.back().hi = ids[i] + 1;
            ++runs.back().reads;
        }
        std::sort(runs.begin(), runs.end(), [](const Run &x, const Run &y) { return x.reads > y.reads; });
        ColourWin w{};
        int used = 0, k = 0;
        for (const Run &r : runs) {
            const int len = (int)(r.hi - r.lo);
            if (k < CW_RUNS && used + len <= slots) {
                w.lo[k] = (int32_t)r.lo;
                w.len[k] = len;
                w.off[k] = used;
                used += len;
                ++k;
            }
        }
        for (int r = k; r < CW_RUNS; ++r) w.off[r] = used;
        // one run for all the block's nodes, and per table column one run for all the upstream
        // nodes that are not solid (a column of solid sites only takes run 0, unused)
        auto run_of = [&w](uint32_t id) {
            for (int r = 0; r < CW_RUNS; ++r)
                if (id - (uint32_t)w.lo[r] < (uint32_t)w.len[r]) return r;
            return -1;
        };
        int own = run_of((uint32_t)(b * CW_BLOCK));
        for (int64_t n = b * CW_BLOCK; n < n1 && own >= 0; ++n)
            if (run_of((uint32_t)n) != own) own = -1;
        if (own >= 0) {
            w.own = w.off[own] - w.lo[own];
            for (int j = 0; j < Q - 1; ++j) {
                int kj = -1;
                bool one = true;
                for (int64_t n = b * CW_BLOCK; n < n1 && one; ++n) {
                    const uint32_t v = g.nbr[(size_t)(j * g.P + n)];
                    if (v & BB_FLAG) continue;
                    const int r = run_of(v);
                    one = r >= 0 && (kj < 0 || r == kj);
                    kj = r;
                }
                if (!one) continue;
                w.fast |= 1u << j;
                w.d[j] = kj < 0 ? 0 : w.off[kj] - w.lo[kj];
            }
        }
        nfast += w.fast == CW_ALL;
        out[(size_t)b] = w;
    }
    return grid ? (double)nfast / (double)grid : 1.0;
}

}  // namespace lbm
