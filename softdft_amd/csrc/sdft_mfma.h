// MFMA wrappers for CDNA4 (gfx950): one 16 x 16 x 4 product-accumulate per call.
// Lane layouts (verified on gfx950):
//   A: lane l holds A[l % 16][l / 16]      B: lane l holds B[l / 16][l % 16]
//   D f32: D[4 (l/16) + v][l % 16]         D f64: D[4 v + (l/16)][l % 16]
#pragma once
#include <hip/hip_runtime.h>

namespace sdft {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};

}  // namespace sdft
