// tune.h -- the launch autotuner: times a kernel's candidate launch settings on the handle's own
// geometry and keeps the fastest, with a process-wide cache of the choices.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "lbm_mi355x.h"

namespace lbm {

// boxes with fewer fluid nodes keep the default settings: their launches are too short for the
// choice to matter, and too short to time
constexpr int64_t TUNE_MIN_NODES = int64_t(1) << 20;
// an alternative replaces the default only when it is faster by this fraction
constexpr double TUNE_MARGIN = 0.01;

// cached choice for (kernel, precision, size class of N); false if none
bool tune_cache_get(int kernel, int precision, int64_t N, lbm_launch &out);
void tune_cache_put(int kernel, int precision, int64_t N, const lbm_launch &ls);

// The fastest of `cands` (cands[0] is the default): each candidate is launched `warm` times,
// then timed as the median of `rounds` rounds of `reps` launches between HIP events on s.
// launch(ls) queues one launch and returns its status.  On an error the default is returned
// with its status.  out.tuned = 1 and out.time_us = the median time of one launch.
template <typename F>
hipError_t tune_launch(hipStream_t s, const std::vector<lbm_launch> &cands, F &&launch, lbm_launch &out, int warm = 2,
                       int reps = 5, int rounds = 3)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    std::vector<double> t(cands.size(), 0.0);
    for (size_t c = 0; c < cands.size() && e == hipSuccess; ++c) {
        for (int w = 0; w < warm && e == hipSuccess; ++w) e = launch(cands[c]);
        std::vector<double> r;
        for (int k = 0; k < rounds && e == hipSuccess; ++k) {
            e = hipEventRecord(e0, s);
            for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch(cands[c]);
            if (e == hipSuccess) e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            r.push_back(1e3 * (double)ms / reps);
        }
        if (e == hipSuccess) {
            std::sort(r.begin(), r.end());
            t[c] = r[r.size() / 2];
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    out = cands[0];
    if (e != hipSuccess) return e;
    size_t best = 0;
    for (size_t c = 1; c < cands.size(); ++c)
        if (t[c] < t[best] * (best == 0 ? 1.0 - TUNE_MARGIN : 1.0)) best = c;
    out = cands[best];
    out.tuned = 1;
    out.time_us = t[best];
    return hipSuccess;
}

// the candidates of a kernel: block sizes x tile orders (x variants), the default first
std::vector<lbm_launch> tune_candidates(const lbm_launch &def, const std::vector<int> &blocks,
                                        const std::vector<int> &orders, const std::vector<int> &variants);

// argument checks shared by the handles' set_launch
bool launch_valid(const lbm_launch &ls, const std::vector<int> &blocks, const std::vector<int> &variants);

}  // namespace lbm
