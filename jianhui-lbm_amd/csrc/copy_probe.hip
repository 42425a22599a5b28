// copy_probe.hip -- device-to-device copies of known size.
//
// copy_kernel, a plain grid-stride float4 copy: its rocprofv3 FETCH_SIZE / WRITE_SIZE readings
// calibrate the counters (known bytes per dispatch / counter reading), so the step kernels'
// counter readings can be turned into HBM bytes.  copy_tuned, one tile of 4 x 256 float4 per
// workgroup with all four loads issued before the stores and nontemporal accesses, measures
// what a copy (one byte written per byte read) reaches on the device: the practical HBM rate
// the step kernels' traffic can be set against.  A measuring tool, built by `make copy_probe`,
// not by `make all`.
//   copy_probe [MiB per buffer, default 1024] [repetitions, default 20]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) copy_tuned(const f4 *__restrict__ src, f4 *__restrict__ dst, size_t n)
{
    constexpr int K = 4;
    const size_t base = (size_t)blockIdx.x * 256 * K + threadIdx.x;
    f4 v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const size_t i = base + (size_t)k * 256;
        if (i < n) v[k] = __builtin_nontemporal_load(&src[i]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const size_t i = base + (size_t)k * 256;
        if (i < n) __builtin_nontemporal_store(v[k], &dst[i]);
    }
}

#define CHECK(x)                                                                        \
    do {                                                                                \
        hipError_t e = (x);                                                             \
        if (e != hipSuccess) {                                                          \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));               \
            return 1;                                                                   \
        }                                                                               \
    } while (0)

int main(int argc, char **argv)
{
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    const size_t bytes = mib << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    CHECK(hipMemset(b, 0, bytes));
    const int grid = 256 * 8 * 4;   // 8 workgroups of 256 per CU, 4 passes per thread group
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    copy_kernel<<<grid, 256>>>(a, b, n);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) copy_kernel<<<grid, 256>>>(a, b, n);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / 1e3 / reps;
    std::printf("copy_probe: %zu MiB read + %zu MiB written per dispatch, %.2f us, %.1f GB/s\n", mib, mib,
                per * 1e6, 2.0 * bytes / per / 1e9);
    std::printf("bytes_read_per_dispatch: %zu\nbytes_written_per_dispatch: %zu\n", bytes, bytes);
    const unsigned tg = (unsigned)((n + 1023) / 1024);
    const f4 *ta = reinterpret_cast<const f4 *>(a);
    f4 *tb = reinterpret_cast<f4 *>(b);
    copy_tuned<<<tg, 256>>>(ta, tb, n);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) copy_tuned<<<tg, 256>>>(ta, tb, n);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double pt = ms / 1e3 / reps;
    std::printf("copy_tuned: %zu MiB read + %zu MiB written per dispatch, %.2f us, %.1f GB/s\n", mib, mib, pt * 1e6,
                2.0 * bytes / pt / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
