// storage.h -- population storage formats and their compute types.
//
//   double : FP64 storage, FP64 arithmetic (the parity build)
//   float  : FP32 storage of the deviation h = f - w*rho0, FP32 arithmetic
//   __half : FP16 storage of h*scale (scale a power of two), FP32 arithmetic
//
// store_nt writes a post-collision value with a nontemporal hint: the next read of it is a
// whole time step away, so keeping it out of the caches leaves them to the pull gathers
// (measured 4 % faster on the 256^3 single-phase step).
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
    static __device__ __forceinline__ void store_nt(double *p, double v, float) { __builtin_nontemporal_store(v, p); }
};

template <>
struct Store<float> {
    using C = float;
    static __device__ __forceinline__ float load(float v, float) { return v; }
    static __device__ __forceinline__ float put(float v, float) { return v; }
    static __device__ __forceinline__ void store_nt(float *p, float v, float) { __builtin_nontemporal_store(v, p); }
};

template <>
struct Store<__half> {
    using C = float;
    static __device__ __forceinline__ float load(__half v, float inv_scale) { return __half2float(v) * inv_scale; }
    static __device__ __forceinline__ __half put(float v, float scale) { return __float2half_rn(v * scale); }
    static __device__ __forceinline__ void store_nt(__half *p, float v, float scale)
    {
        __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(v * scale)), reinterpret_cast<unsigned short *>(p));
    }
};

}  // namespace lbm
