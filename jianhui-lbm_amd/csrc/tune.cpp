// tune.cpp -- the autotuner's cache and candidate lists (tune.h).
#include "tune.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

namespace lbm {

namespace {
std::mutex cache_mu;
std::map<std::tuple<int, int, int>, lbm_launch> cache;

// two size classes per octave of the node count
int size_class(int64_t N) { return (int)std::floor(2.0 * std::log2((double)std::max<int64_t>(N, 1))); }
}  // namespace

bool tune_cache_get(int kernel, int precision, int64_t N, lbm_launch &out)
{
    std::lock_guard<std::mutex> lk(cache_mu);
    auto it = cache.find(std::make_tuple(kernel, precision, size_class(N)));
    if (it == cache.end()) return false;
    out = it->second;
    return true;
}

void tune_cache_put(int kernel, int precision, int64_t N, const lbm_launch &ls)
{
    std::lock_guard<std::mutex> lk(cache_mu);
    cache[std::make_tuple(kernel, precision, size_class(N))] = ls;
}

std::vector<lbm_launch> tune_candidates(const lbm_launch &def, const std::vector<int> &blocks,
                                        const std::vector<int> &orders, const std::vector<int> &variants)
{
    std::vector<lbm_launch> c{def};
    for (int v : variants)
        for (int b : blocks)
            for (int o : orders) {
                if (b == def.block && o == def.order && v == def.variant) continue;
                lbm_launch l{};
                l.block = b;
                l.order = o;
                l.variant = v;
                c.push_back(l);
            }
    return c;
}

bool launch_valid(const lbm_launch &ls, const std::vector<int> &blocks, const std::vector<int> &variants)
{
    return std::find(blocks.begin(), blocks.end(), ls.block) != blocks.end() &&
           (ls.order == LBM_ORDER_HW || ls.order == LBM_ORDER_RUNS8 || ls.order == LBM_ORDER_SLAB) &&
           std::find(variants.begin(), variants.end(), ls.variant) != variants.end();
}

}  // namespace lbm
