// z stage of the run-time planned fused pipeline with the z length fixed at
// compile time (one kernel per 2,3,5,7-smooth length from 16 to 256; other
// lengths use the LDS-planned k_zcorr_rt in pf_planes.hip).
//
// k_zcorr_ct: one block per (tile of TQ (ky,kx) columns, rotation chunk); the
// column's L points live in the registers of T threads (pf_ct_fft.h Flex).
// Per rotation: forward z FFT of the P and W (and W^2) xy-spectra, the
// products conj(X) F1 / conj(X) F2 with the F tiles held in registers for the
// whole chunk, inverse z FFT, written back in place.  The forward transform
// leaves each thread holding exactly the elements the inverse starts from, so
// product and inverse need no exchange; each transform exchanges through LDS
// once per stage boundary.
#include <algorithm>

#include "pf_ct_fft.h"
#include "pf_internal.h"

namespace pf {

namespace {

constexpr int kZTmax = 32;   // threads per transform
constexpr int kZTq = 16;     // columns per block

// power-of-two threads per column: a block is 16 T threads, whole wavefronts
// from T = 4 up.  EMAX: register slots per thread and transform (the default
// kFlexZSlots, or the alternative kFlexZSlotsAlt plan of more threads with
// fewer registers each, where one exists).
template <int L, int EMAX>
using ZF = Flex<L, EMAX, kZTmax, false, true>;
template <int L, int EMAX>
using ZB = Flex<L, EMAX, kZTmax, true, true>;

template <int L, bool BIN, int EMAX>
__global__ __launch_bounds__((kZTq * ZF<L, EMAX>::T)) void k_zcorr_ct(
    float2* __restrict__ X, const float2* __restrict__ F1, const float2* __restrict__ F2,
    int plane, int64_t H, int nrot, int chunk, int zr, const float2* __restrict__ twz) {
    using ZF = pf::ZF<L, EMAX>;
    using ZB = pf::ZB<L, EMAX>;
    constexpr int T = ZF::T, E = ZF::E;
    constexpr int EI = ZF::in_slots(), EO = ZF::out_slots();
    static_assert(ZB::T == T && ZB::E == E, "same threads and slots both ways");
    __shared__ float2 ex[L * kZTq];
    __shared__ float2 stw[L];
    for (int i = threadIdx.x; i < L; i += blockDim.x) stw[i] = twz[i];
    __syncthreads();
    const int c = threadIdx.x % kZTq, t = threadIdx.x / kZTq;
    const int q = blockIdx.x * kZTq + c;
    const bool active = q < plane;
    const int b0 = blockIdx.y * chunk;
    const int b1 = min(nrot, b0 + chunk);
    // target spectra at this thread's output-pattern elements
    float2 f1[EO], f2[EO];
#pragma unroll
    for (int s = 0; s < EO; ++s) {
        const int p = ZF::out_pos(t, s);
        const bool ok = active && p >= 0;
        f1[s] = ok ? F1[(int64_t)p * plane + q] : c2(0.f, 0.f);
        f2[s] = ok ? F2[(int64_t)p * plane + q] : c2(0.f, 0.f);
    }
    float2* const exc = ex + c;
    for (int b = b0; b < b1; ++b) {
        float2* const base = X + (int64_t)b * 3 * H + q;
#pragma unroll
        for (int g = 0; g < (BIN ? 2 : 3); ++g) {
            float2* const col = base + (int64_t)g * H;
            float2 v[E];
#pragma unroll
            for (int s = 0; s < E; ++s) v[s] = c2(0.f, 0.f);
#pragma unroll
            for (int s = 0; s < EI; ++s) {
                // planes beyond the support were not written: they are zero
                const int p = ZF::in_pos(t, s);
                const int sz = p <= L / 2 ? p : p - L;
                if (active && p >= 0 && sz >= -zr && sz <= zr) v[s] = col[(int64_t)p * plane];
            }
            flex_fft<L, EMAX, kZTmax, false, -1, true>(v, exc, kZTq, t, stw, active);
            if (BIN && g == 1) {
                // W's spectrum feeds both A (with F1) and B (with F2)
                float2 u[E];
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    u[s] = s < EO ? cmul(cconj(v[s]), f2[s]) : c2(0.f, 0.f);
                    v[s] = s < EO ? cmul(cconj(v[s]), f1[s]) : c2(0.f, 0.f);
                }
                flex_fft<L, EMAX, kZTmax, true, 1, true>(v, exc, kZTq, t, stw, active);
                flex_fft<L, EMAX, kZTmax, true, 1, true>(u, exc, kZTq, t, stw, active);
                if (active) {
#pragma unroll
                    for (int s = 0; s < EI; ++s) {
                        const int p = ZB::out_pos(t, s);
                        if (p >= 0) {
                            col[(int64_t)p * plane] = v[s];
                            col[H + (int64_t)p * plane] = u[s];
                        }
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < E; ++s)
                    v[s] = s < EO ? cmul(cconj(v[s]), g == 2 ? f2[s] : f1[s]) : c2(0.f, 0.f);
                flex_fft<L, EMAX, kZTmax, true, 1, true>(v, exc, kZTq, t, stw, active);
                if (active) {
#pragma unroll
                    for (int s = 0; s < EI; ++s) {
                        const int p = ZB::out_pos(t, s);
                        if (p >= 0) col[(int64_t)p * plane] = v[s];
                    }
                }
            }
        }
    }
}

template <int L, int EMAX>
int launch_zct(const Fft3D& f, const FusedArgs& a, int zr, hipStream_t s) {
    const int plane = f.ny * f.nxh;
    const int64_t H = (int64_t)f.nz * plane;
    const dim3 grid((plane + kZTq - 1) / kZTq, (a.nrot + a.chunk_z - 1) / a.chunk_z);
    const dim3 block(kZTq * ZF<L, EMAX>::T);
    if (a.binary_mask)
        hipLaunchKernelGGL((k_zcorr_ct<L, true, EMAX>), grid, block, 0, s, a.work, a.F1, a.F2,
                           plane, H, a.nrot, a.chunk_z, zr, f.twz);
    else
        hipLaunchKernelGGL((k_zcorr_ct<L, false, EMAX>), grid, block, 0, s, a.work, a.F1, a.F2,
                           plane, H, a.nrot, a.chunk_z, zr, f.twz);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of k_zcorr_ct failed: %s", hipGetErrorString(e));
        return kErrHip;
    }
    return kOk;
}

}  // namespace

bool zcorr_ct_supported(int nz) {
    switch (nz) {
#define PF_ZCASE(L) case L:
        PF_CT_COL_LENGTHS(PF_ZCASE)
#undef PF_ZCASE
        return true;
        default: return false;
    }
}

int launch_zcorr_ct(const Fft3D& f, const FusedArgs& a, int zr, hipStream_t s) {
    if (a.nrot <= 0) return kOk;
    if (a.nrot > 65535 || a.chunk_z < 1) {
        set_error("z stage: invalid batch or chunk size");
        return kErrArg;
    }
    // schedule bit 0: the alternative register plan where the length has one
    const bool alt = a.zsched & 1;
    switch (f.nz) {
#define PF_ZCASE(L)                                                          \
    case L:                                                                  \
        if constexpr (flex_zalt_ok(L))                                       \
            if (alt) return launch_zct<L, kFlexZSlotsAlt>(f, a, zr, s);      \
        return launch_zct<L, kFlexZSlots>(f, a, zr, s);
        PF_CT_COL_LENGTHS(PF_ZCASE)
#undef PF_ZCASE
        default:
            set_error("z stage: no compiled transform for length %d", f.nz);
            return kErrUnsupported;
    }
}

}  // namespace pf
