// Runtime-planned fused LCC pipeline for grids without a compiled fused size
// (non-cubic grids, odd x lengths, and cubic lengths whose radices the
// register FFTs cannot split evenly over threads, e.g. 120 = 8 * 3 * 5).
//
// The same three stages as the compiled pipeline (pf_fused.hip), with the FFT
// lengths and radix plans taken at run time: every transform is an LDS
// Stockham pass (lds_fft) over a whole z-plane or a tile of z-columns.
//
//   k_fwd_plane_rt  one block per (z-plane, rotation), one launch per grid:
//                   the P launch gathers the rotated template (trilinear) and
//                   mask (nearest) of the plane row by row and parks the mask
//                   values, the W (and W^2 for weighted masks) launches read
//                   them back; each then runs R2C along x (half-length complex
//                   FFT + post-processing), C2C along y, and writes its
//                   xy-spectrum in numpy's [z][ky][kx] layout.
//   k_zcorr_rt      one block per (tile of (ky,kx) columns, rotation chunk):
//                   forward FFT along z, conj(X) F1 / conj(X) F2 with the F
//                   tile in LDS, inverse FFT along z.
//   k_inv_plane_rt  one block per (z-plane, rotation chunk): inverse C2C along
//                   y and C2R along x for B, A then G, the LCC from the raw
//                   outputs, merged into the running (LCC, index) keys with
//                   atomicMax.
//
// Limits (plane_rt_supported): two planes of ny x (nx/2+1) complex (for odd
// nx, or its packed rows if larger) in LDS; at most kRtVox voxels of a plane
// per thread of the LCC stage.
#include "pf_planes_kern.h"

namespace pf {

namespace {

// z stage.  LDS: S0, S1, S2 (nz x TQ complex each, element (z, c) at z TQ + c)
// and the z twiddles; the F1, F2 tiles are read from global memory (L2) by
// each multiply.  BIN: W's spectrum gives both A (F1) and B (F2).
__global__ __launch_bounds__(kRtZThreads) void k_zcorr_rt(
    float2* __restrict__ X, const float2* __restrict__ F1, const float2* __restrict__ F2,
    RtGeo g, int nrot, int chunk, int bin, int zr, PlanRt pz, const float2* __restrict__ twz_g) {
    extern __shared__ float2 lds[];
    const int tile = g.nz * kRtTq;
    float2* const S0 = lds;
    float2* const S1 = lds + tile;
    float2* const S2 = lds + 2 * tile;
    float2* const twz = lds + 3 * tile;
    stage_tw(twz, twz_g, g.nz);
    const int q0 = blockIdx.x * kRtTq;
    const int nc = min(kRtTq, g.plane - q0);
    const int b0 = blockIdx.y * chunk, b1 = min(nrot, b0 + chunk);
    const int n = g.nz * nc;
    const float2* T1 = F1 + q0;
    const float2* T2 = F2 + q0;
    // a forward grid's column tile; planes beyond the support were not
    // written by the forward stage: they are zero
    auto load = [&](const float2* src, float2* d) {
        Step2 it(nc);
        for (int u = threadIdx.x; u < n; u += blockDim.x, it.next()) {
            const int z = it.r, c = it.c;
            const int sz = z <= g.nz / 2 ? z : z - g.nz;
            d[z * kRtTq + c] = (sz >= -zr && sz <= zr) ? src[(int64_t)z * g.plane + q0 + c]
                                                       : c2(0.f, 0.f);
        }
        __syncthreads();
    };
    auto store = [&](const float2* s, float2* d) {
        Step2 it(nc);
        for (int u = threadIdx.x; u < n; u += blockDim.x, it.next()) {
            const int z = it.r, c = it.c;
            d[(int64_t)z * g.plane + q0 + c] = s[z * kRtTq + c];
        }
    };
    // in place: s = conj(s) * t (and, with t2, d = conj(s) * t2 first); t, t2
    // are the target spectra at this tile's columns (row stride plane)
    auto mul = [&](float2* s, const float2* t, float2* d, const float2* t2) {
        Step2 it(nc);
        for (int u = threadIdx.x; u < n; u += blockDim.x, it.next()) {
            const int z = it.r, c = it.c;
            const int i = z * kRtTq + c;
            const int64_t gi = (int64_t)z * g.plane + c;
            const float2 x = cconj(s[i]);
            if (d) d[i] = cmul(x, t2[gi]);
            s[i] = cmul(x, t[gi]);
        }
        __syncthreads();
    };
    // Per rotation three outputs: G = IFFT(conj(FP) F1), A = IFFT(conj(FW) F1),
    // B = IFFT(conj(FW) F2) (FW^2 for weighted masks).  With a 0/1 mask FW
    // is transformed once and its F2 product parked in S2 for B.  One
    // forward and one inverse call site keep the kernel's code small.
    for (int b = b0; b < b1; ++b) {
        float2* base = X + (int64_t)b * 3 * g.H;
        for (int o = 0; o < 3; ++o) {
            float2* Z;
            float2* other;
            if (bin && o == 2) {
                Z = S2;
                other = S0;
            } else {
                load(base + (int64_t)(bin ? (o > 0 ? 1 : 0) : o) * g.H, S0);
                Z = tile_fft<-1, true>(S0, S1, pz, twz, nc, kRtTq);
                other = Z == S0 ? S1 : S0;
                const bool both = bin && o == 1;
                mul(Z, o == 2 ? T2 : T1, both ? S2 : nullptr, both ? T2 : nullptr);
            }
            const float2* Zi = tile_fft<1, true>(Z, other, pz, twz, nc, kRtTq);
            store(Zi, base + (int64_t)o * g.H);
            __syncthreads();
        }
    }
}

}  // namespace

bool plane_rt_supported(int nz, int ny, int nx) {
    if (nz < 2 || ny < 2 || nx < 3) return false;
    const RtGeo g = make_geo(nz, ny, nx);
    if (plane_lds_bytes(g) > kRtLds) return false;
    if ((int64_t)ny * nx > (int64_t)kRtVox * kRtThreads) return false;
    // z kernel: three column tiles and the z twiddles
    if ((size_t)(3 * kRtTq + 1) * nz * sizeof(float2) > kRtLds) return false;
    // radix plans that fit the packed kernel argument
    PlanRt p;
    return plan_rt(g.rl, &p) && plan_rt(ny, &p) && plan_rt(nz, &p);
}

int prepare_planes_rt() {
    PF_CHECK_STATUS(prepare_plane_kernels<false>());
    PF_CHECK_STATUS(prepare_plane_kernels_odd());
    PF_CHECK_STATUS(allow_lds((const void*)k_zcorr_rt, kRtLds));
    return kOk;
}

int launch_planes_rt(const Fft3D& f, const FusedArgs& a, int which, hipStream_t s) {
    if (a.nrot <= 0) return kOk;
    if (!plane_rt_supported(f.nz, f.ny, f.nx)) {
        set_error("runtime-planned fused pipeline not available for grid (%d, %d, %d)", f.nz,
                  f.ny, f.nx);
        return kErrUnsupported;
    }
    if (a.nrot > 65535 || a.chunk_z < 1 || a.chunk_lcc < 1) {
        set_error("fused: invalid batch or chunk sizes");
        return kErrArg;
    }
    // planes |z| <= zr only, the z stage reading the others as zero
    const int zr = std::max(0, std::min(a.zr, f.nz / 2));
    const bool skip = 2 * zr + 1 < f.nz;
    if (which == 1) {
        if (zcorr_ct_supported(f.nz)) return launch_zcorr_ct(f, a, skip ? zr : f.nz, s);
        const RtGeo g = make_geo(f.nz, f.ny, f.nx);
        PlanRt pz;
        if (!pack_plan(f.pz, &pz)) {
            set_error("z stage: no packed radix plan for length %d", f.nz);
            return kErrUnsupported;
        }
        const int tiles = (g.plane + kRtTq - 1) / kRtTq;
        const size_t lds = (size_t)(3 * kRtTq + 1) * f.nz * sizeof(float2);
        hipLaunchKernelGGL(k_zcorr_rt, dim3(tiles, (a.nrot + a.chunk_z - 1) / a.chunk_z),
                           dim3(kRtZThreads), lds, s, a.work, a.F1, a.F2, g, a.nrot, a.chunk_z,
                           a.binary_mask ? 1 : 0, skip ? zr : f.nz, pz, f.twz);
        return launched("k_zcorr_rt");
    }
    if (f.nx % 2) return launch_plane_kernel_odd(f, a, which, skip, zr, s);
    return launch_plane_kernel<false>(f, a, which, skip, zr, s);
}

}  // namespace pf
