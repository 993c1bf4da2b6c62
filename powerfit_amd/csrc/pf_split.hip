// Split-precision DFT-as-GEMM first stage on the matrix cores (the optional
// precision mode of the unfused FFT, PF_FLAG_SPLIT_DFT).
//
// A strided-axis C2C transform of length n = 16 L is a Stockham plan whose
// first stage is a batch of 16-point DFTs over the stride-L elements:
//     y[16 j + k] = sum_m F16[k][m] x[j + m L],   j < L.
// Here that stage is a real GEMM on v_mfma_f32_32x32x8f16: the 32 x 32 real
// form of F16 ([[C, -S], [S, C]], rows Re/Im of the outputs, columns Re/Im of
// the inputs) times a 32 x 32 block of inputs (one GEMM column per tile column
// of the same j).  Each operand is split into FP16 hi + lo parts (inputs after
// an exact power-of-two scaling per GEMM column), and the product is
// Fh Xh + Fh Xl + Fl Xh accumulated in FP32 (the Fl Xl term is below FP32
// rounding); the scaling is undone exactly on the way out.  The remaining
// stages run in FP32 on the VALU with the LDS Stockham kernels.
#include <math.h>

#include "pf_internal.h"
#include "pf_mfma.h"

namespace pf {
namespace {

constexpr int kTq = 16;        // tile columns; a GEMM block is 2 stage rows j of 16 columns
constexpr int kThreads = 256;  // 4 waves

// Stages s0.. of plan p over LDS data (as lds_fft, starting at span Ns0).
template <int DIR>
__device__ float2* lds_fft_tail(float2* a, float2* b, const Plan1D& p, const float2* tw,
                                int nfft, int es, int s0, int Ns0) {
    int Ns = Ns0;
    for (int s = s0; s < p.nstages; ++s) {
        const int R = p.radix[s];
        const int nb = p.n / R;
        const int total = nfft * nb;
        for (int u = threadIdx.x; u < total; u += blockDim.x) {
            const int j = u / nfft;
            const int f = u - j * nfft;
            stage_dispatch<DIR>(a + f, b + f, R, j, Ns, p.n, es, tw);
        }
        __syncthreads();
        float2* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// In-place C2C along a strided axis (layout as k_c2c_axis), first radix-16
// stage on MFMA.  One block per (outer index, tile of 16 consecutive c); GEMM
// column g of block jb is tile column g % 16 at stage row j = 2 jb + g / 16.
template <int DIR>
__global__ __launch_bounds__(kThreads) void k_c2c_axis_split(float2* __restrict__ data,
                                                             int64_t stride, int64_t inner,
                                                             int64_t outer, int64_t outer_stride,
                                                             int64_t ntiles, Plan1D p,
                                                             const float2* __restrict__ tw) {
    extern __shared__ float2 lds[];
    const int n = p.n;
    const int L = n / 16;
    const int64_t blk = blockIdx.x;
    const int64_t o = blk / ntiles;
    const int64_t c0 = (blk - o * ntiles) * kTq;
    if (o >= outer || c0 >= inner) return;
    const int nc = (int)min((int64_t)kTq, inner - c0);
    float2* A = lds;
    float2* B = lds + (size_t)n * kTq;
    float2* base = data + o * outer_stride + c0;
    for (int u = threadIdx.x; u < n * kTq; u += blockDim.x) {
        const int k = u / kTq, c = u - k * kTq;
        A[u] = c < nc ? base[k * stride + c] : c2(0.f, 0.f);
    }
    __syncthreads();

    // ---- stage 0: y[16 j + k] = F16 x[j + m L], a 32 x 32 x 32 real GEMM per
    // pair of stage rows j
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane & 31, h = lane >> 5;
    const int c = g & 15;
    half4 fh[4], fl[4];
    dft16_operand<DIR>(fh, fl);
    for (int jb = wave; jb < (L + 1) / 2; jb += kThreads / 64) {
        const int j = 2 * jb + (g >> 4);
        const bool valid = j < L;
        const int jr = valid ? j : L - 1;   // an in-range row for the idle half
        // inputs m = 4 h + q and 8 + 4 h + q of GEMM column g
        float2 x[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            x[q] = A[(jr + (4 * h + q) * L) * kTq + c];
            x[4 + q] = A[(jr + (8 + 4 * h + q) * L) * kTq + c];
        }
        dft16_split_gemm(x, fh, fl);
        if (valid) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                B[(16 * j + 4 * h + q) * kTq + c] = x[q];
                B[(16 * j + 8 + 4 * h + q) * kTq + c] = x[4 + q];
            }
        }
    }
    __syncthreads();

    // ---- remaining stages in FP32
    const float2* Z = lds_fft_tail<DIR>(B, A, p, tw, kTq, kTq, 1, 16);
    for (int u = threadIdx.x; u < n * nc; u += blockDim.x) {
        const int k = u / nc, cc = u - k * nc;
        base[k * stride + cc] = Z[k * kTq + cc];
    }
}

}  // namespace

bool make_plan16(int n, Plan1D* p) {
    if (n % 16 != 0 || n > kSplitMaxLen) return false;
    Plan1D rest;
    if (n == 16) {
        rest.n = 1;
        rest.nstages = 0;
    } else if (!make_plan(n / 16, &rest)) {
        return false;
    }
    if (rest.nstages + 1 > kMaxStages) return false;
    p->n = n;
    p->nstages = rest.nstages + 1;
    p->radix[0] = 16;
    for (int s = 0; s < rest.nstages; ++s) p->radix[s + 1] = rest.radix[s];
    return true;
}

int launch_c2c_axis_split(const Plan1D& p, const float2* tw, float2* data, int64_t stride,
                          int64_t inner, int64_t outer, int64_t outer_stride, int dir,
                          hipStream_t s) {
    if (p.n <= 1 || inner <= 0 || outer <= 0) return kOk;
    if (p.radix[0] != 16 || p.n % 16 != 0 || p.n > kSplitMaxLen) {
        set_error("split-precision axis: plan must start with radix 16 (n <= %d)", kSplitMaxLen);
        return kErrUnsupported;
    }
    const int64_t ntiles = (inner + kTq - 1) / kTq;
    const int64_t blocks = ntiles * outer;
    if (blocks > 0x7fffffffLL) {
        set_error("split-precision axis: grid too large");
        return kErrUnsupported;
    }
    const size_t lds = 2 * sizeof(float2) * (size_t)kTq * p.n;   // <= 160 KB for n <= 320
    const void* fn = dir < 0 ? (const void*)k_c2c_axis_split<-1> : (const void*)k_c2c_axis_split<1>;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess) {
        set_error("split-precision axis: cannot reserve %zu bytes of LDS", lds);
        return kErrHip;
    }
    if (dir < 0)
        hipLaunchKernelGGL(k_c2c_axis_split<-1>, dim3((unsigned)blocks), dim3(kThreads), lds, s,
                           data, stride, inner, outer, outer_stride, ntiles, p, tw);
    else
        hipLaunchKernelGGL(k_c2c_axis_split<1>, dim3((unsigned)blocks), dim3(kThreads), lds, s,
                           data, stride, inner, outer, outer_stride, ntiles, p, tw);
    if (hipGetLastError() != hipSuccess) {
        set_error("k_c2c_axis_split launch failed");
        return kErrHip;
    }
    return kOk;
}

}  // namespace pf
