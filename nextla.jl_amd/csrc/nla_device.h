// Device-side helpers shared by the gfx950 kernels of libnlahip.
//
// MFMA operand / result layouts (wave64, verified on gfx950 by a lane-layout
// probe; the f64 result layout differs from the f32 16x16 one):
//   v_mfma_f64_16x16x4_f64 : A[i][k] from lane i+16k, B[k][j] from lane j+16k,
//                            D[i][j] in lane j+16*(i%4), register i/4.
//   v_mfma_f32_32x32x2_f32 : A[i][k] from lane i+32k, B[k][j] from lane j+32k,
//                            D[i][j] in lane j+32*((i/4)%2), register 4*(i/8)+i%4.
//   v_mfma_f32_16x16x32_{bf16,f16}: A[i][k] from lane i+16*(k/8) element k%8,
//                            B likewise, D[i][j] in lane j+16*(i/4), register i%4.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nla {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef float f16_t __attribute__((ext_vector_type(16)));

template <typename T>
struct Mfma;

// One 16x16x4 f64 MFMA block: 2048 flops, 64 cycles on one SIMD.
template <>
struct Mfma<double> {
    static constexpr int BD = 16;  // block edge (rows and columns of D)
    static constexpr int KD = 4;   // k consumed per instruction
    static constexpr int NR = 4;   // result registers per lane
    typedef d4_t acc_t;
    __device__ static inline acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
    __device__ static inline acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // Coordinates (i, j) in D of lane l's result register r.
    __device__ static inline int di(int l, int r) { return 4 * r + (l >> 4); }
    __device__ static inline int dj(int l, int) { return l & 15; }
};

// One 32x32x2 f32 MFMA block: 4096 flops, 64 cycles on one SIMD.
template <>
struct Mfma<float> {
    static constexpr int BD = 32;
    static constexpr int KD = 2;
    static constexpr int NR = 16;
    typedef f16_t acc_t;
    __device__ static inline acc_t zero() {
        acc_t z;
#pragma unroll
        for (int r = 0; r < 16; ++r) z[r] = 0.0f;
        return z;
    }
    __device__ static inline acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int di(int l, int r) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }
    __device__ static inline int dj(int l, int) { return l & 31; }
};

}  // namespace nla
