// Split-precision 16-point DFT on the matrix cores (v_mfma_f32_32x32x8f16),
// shared by the unfused split passes (pf_split.hip) and the fused z stage
// (pf_fused.hip).
//
// The 16-point DFT is applied as the 32 x 32 real matrix [[C, -S], [S, C]]
// (rows Re/Im of the outputs, columns Re/Im of the inputs) to 32 GEMM columns
// of 16 complex inputs each.  Lane (g, h) = (lane % 32, lane / 32) holds, for
// GEMM column g, inputs m = 4 h + q and 8 + 4 h + q (q < 4); K step s of the
// B operand holds rows 8 s + 4 h + q = Re x[4h+q], Re x[8+4h+q], Im x[4h+q],
// Im x[8+4h+q] for s = 0..3, and the result rows give Re y[4h+q] = acc[q],
// Re y[8+4h+q] = acc[4+q], Im y[4h+q] = acc[8+q], Im y[8+4h+q] = acc[12+q].
// Each operand is split into FP16 hi + lo parts (inputs after an exact
// power-of-two scaling per GEMM column), Fh Xh + Fh Xl + Fl Xh is accumulated
// in FP32 (the Fl Xl term lies below FP32 rounding), and the scaling is undone
// exactly.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "pf_fft.h"

namespace pf {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Element (i, k) of the 32 x 32 real form of the 16-point DFT with sign DIR.
template <int DIR>
__device__ __forceinline__ float dft16_real(int i, int k) {
    const int m = ((i & 15) * (k & 15)) & 15;
    const double c = cospi((double)m / 8.0);
    const double d = (double)DIR * sinpi((double)m / 8.0);
    if (i < 16) return (float)(k < 16 ? c : -d);
    return (float)(k < 16 ? d : c);
}

__device__ __forceinline__ void split_half(float v, _Float16& hi, _Float16& lo) {
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

// The A operand (the DFT matrix, rows i = lane % 32) split into hi and lo.
template <int DIR>
__device__ __forceinline__ void dft16_operand(half4 (&fh)[4], half4 (&fl)[4]) {
    const int lane = threadIdx.x & 63;
    const int g = lane & 31, h = lane >> 5;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            _Float16 hi, lo;
            split_half(dft16_real<DIR>(g, 8 * s + 4 * h + q), hi, lo);
            fh[s][q] = hi;
            fl[s][q] = lo;
        }
}

// One 16-point DFT per GEMM column of the wavefront: x[q] = input 4 h + q,
// x[4 + q] = input 8 + 4 h + q of this lane's column; on return the same
// slots hold outputs 4 h + q and 8 + 4 h + q.  Every lane of the wavefront
// must call it (the MFMA and the cross-half maximum are wave-wide).
__device__ __forceinline__ void dft16_split_gemm(float2 (&x)[8], const half4 (&fh)[4],
                                                 const half4 (&fl)[4]) {
    float mx = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) mx = fmaxf(mx, fmaxf(fabsf(x[q].x), fabsf(x[q].y)));
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));   // the column's other half
    int e = 0;
    if (mx > 0.f && isfinite(mx)) (void)frexpf(mx, &e);
    const float sc = ldexpf(1.f, -e), unsc = ldexpf(1.f, e);
    half4 xh[4], xl[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        _Float16 hi, lo;
        split_half(x[q].x * sc, hi, lo);
        xh[0][q] = hi;
        xl[0][q] = lo;
        split_half(x[4 + q].x * sc, hi, lo);
        xh[1][q] = hi;
        xl[1][q] = lo;
        split_half(x[q].y * sc, hi, lo);
        xh[2][q] = hi;
        xl[2][q] = lo;
        split_half(x[4 + q].y * sc, hi, lo);
        xh[3][q] = hi;
        xl[3][q] = lo;
    }
    floatx16 acc = {};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        acc = __builtin_amdgcn_mfma_f32_32x32x8f16(fh[s], xh[s], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8f16(fh[s], xl[s], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8f16(fl[s], xh[s], acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        x[q] = c2(acc[q] * unsc, acc[8 + q] * unsc);
        x[4 + q] = c2(acc[4 + q] * unsc, acc[12 + q] * unsc);
    }
}

}  // namespace pf
