// storage.h -- population storage formats and their compute types.
//
//   double : FP64 storage, FP64 arithmetic (the parity build)
//   float  : FP32 storage of the deviation h = f - w*rho0, FP32 arithmetic
//   __half : FP16 storage of h*scale (scale a power of two), FP32 arithmetic
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace lbm {

template <typename ST>
struct Store;

template <>
struct Store<double> {
    using C = double;
    static __device__ __forceinline__ double load(double v, float) { return v; }
    static __device__ __forceinline__ double put(double v, float) { return v; }
};

template <>
struct Store<float> {
    using C = float;
    static __device__ __forceinline__ float load(float v, float) { return v; }
    static __device__ __forceinline__ float put(float v, float) { return v; }
};

template <>
struct Store<__half> {
    using C = float;
    static __device__ __forceinline__ float load(__half v, float inv_scale) { return __half2float(v) * inv_scale; }
    static __device__ __forceinline__ __half put(float v, float scale) { return __float2half_rn(v * scale); }
};

}  // namespace lbm
