// MFMA wrappers shared by the HIP kernels of libqumasun_hip (gfx950).
//
// 16x16x4 MFMA, one 16 x 16 result tile per wave.  Operand layout (lane l,
// group g = l / 16): A[i = l % 16][k = g], B[k = g][j = l % 16].  The result item
// r of lane l holds D[row(g, r)][l % 16]; the FP64 instruction interleaves the
// rows across lane groups (row = g + 4 r), the FP32 one does not (row = 4 g + r).
#pragma once

#include <hip/hip_runtime.h>

namespace qm {

template <typename T>
struct Mfma;

template <>
struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    typedef float vec_t __attribute__((ext_vector_type(4)));
    static constexpr int VEC = 4;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

template <>
struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    typedef double vec_t __attribute__((ext_vector_type(2)));
    static constexpr int VEC = 2;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};

}  // namespace qm
