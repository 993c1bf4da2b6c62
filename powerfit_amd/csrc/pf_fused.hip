// Fused LCC search pipeline for cubic N^3 grids (N even, compile-time).
//
// Per batch of rotations, three kernels replace rotate -> 3 x (R2C, conj-mul,
// C2R) -> calc_lcc -> take-best:
//
//   k_fwd_plane   one block per (z-plane, rotation): gathers the rotated
//                 template (trilinear) and mask (nearest) of the plane, runs
//                 the R2C along x and the C2C along y with the plane resident
//                 in LDS, and writes the xy-spectra of P, W (and W^2 for
//                 weighted masks) in numpy's [z][ky][kx] layout.
//   k_zcorr       one block per (tile of (ky,kx) columns, rotation chunk):
//                 forward FFT along z, multiply by the conjugate-free target
//                 spectrum (conj(X) F1 or conj(X) F2, F tiles kept in
//                 registers for the whole chunk), inverse FFT along z.
//   k_inv_plane   one block per (z-plane, rotation chunk): inverse C2C along
//                 y and C2R along x with the plane in LDS, then
//                 S = G / sqrt(N B - A^2) (variance in FP64) and a strict '>'
//                 running max/argmax over the chunk in registers, merged into
//                 the result with a 64-bit atomicMax on lcc_key.
//
// The DC and Nyquist columns of the x-transform are both real; they travel
// through the y transform packed as one complex column X0 + i XM and are
// separated with Hermitian symmetry on the way out (and packed on the way in
// for the inverse).
#include "pf_group_fft.h"
#include "pf_internal.h"
#include "pf_rotate.h"

namespace pf {

template <int N>
struct Geo {
    static constexpr int M = N / 2;          // complex points of the half-length x FFT
    static constexpr int NXH = M + 1;        // kx extent in global memory
    static constexpr int S = M + 1;          // LDS row stride (complex), padded
    static constexpr int PLANE = N * NXH;    // complex per z-plane in global memory
    static constexpr int64_t H = (int64_t)N * PLANE;
    static constexpr int64_t V = (int64_t)N * N * N;
};

// Column (y) FFT of the M packed columns of an LDS plane, NTHR threads.
template <int N, int NTHR, int DIR>
__device__ __forceinline__ void plane_columns(float2* plane, const float2* __restrict__ twN) {
    using G = Geo<N>;
    constexpr int TY = pick_T(N, 16, 32);
    constexpr int EY = N / TY;
    constexpr int CPI = NTHR / TY;
    constexpr int CITERS = (G::M + CPI - 1) / CPI;
    static_assert(TY > 0, "no column plan");
    const int gy = threadIdx.x / TY, ty = threadIdx.x % TY;
#pragma unroll
    for (int ci = 0; ci < CITERS; ++ci) {
        const int c = ci * CPI + gy;
        const bool active = (gy < CPI) && (c < G::M);
        float2 v[EY];
        if (active) {
#pragma unroll
            for (int e = 0; e < EY; ++e) v[e] = plane[(ty + TY * e) * G::S + c];
        }
        __syncthreads();
        group_fft<N, TY, DIR>(v, plane + (active ? c : 0), G::S, ty, twN, active);
        if (active) {
#pragma unroll
            for (int e = 0; e < EY; ++e) plane[(ty + TY * e) * G::S + c] = v[e];
        }
        __syncthreads();
    }
}

// Write the y-transformed plane to global [ky][kx], unpacking column 0.
template <int N, int NTHR>
__device__ __forceinline__ void store_plane(const float2* plane, float2* __restrict__ dst) {
    using G = Geo<N>;
    for (int f = threadIdx.x; f < G::PLANE; f += NTHR) {
        const int ky = f / G::NXH, kx = f - ky * G::NXH;
        float2 val;
        if (kx == 0 || kx == G::M) {
            const float2 D = plane[ky * G::S];
            const float2 Dm = cconj(plane[((N - ky) % N) * G::S]);
            val = (kx == 0) ? cscale(cadd(D, Dm), 0.5f) : mul_si<-1>(cscale(csub(D, Dm), 0.5f));
        } else {
            val = plane[ky * G::S + kx];
        }
        dst[f] = val;
    }
}

// Load a plane [ky][kx] from global into LDS, packing columns 0 and M.
template <int N, int NTHR>
__device__ __forceinline__ void load_plane(float2* plane, const float2* __restrict__ src) {
    using G = Geo<N>;
    for (int f = threadIdx.x; f < N * G::M; f += NTHR) {
        const int ky = f / G::M, kx = f - ky * G::M;
        float2 val;
        if (kx == 0) {
            const float2 y0 = src[ky * G::NXH], ym = src[ky * G::NXH + G::M];
            val = cadd(y0, mul_si<1>(ym));
        } else {
            val = src[ky * G::NXH + kx];
        }
        plane[ky * G::S + kx] = val;
    }
}

// ---------------------------------------------------------------------------
template <int N, int NTHR, bool BIN>
__global__ __launch_bounds__(NTHR) void k_fwd_plane(const float2* __restrict__ tm,
                                                    const double* __restrict__ rots, int r,
                                                    float2* __restrict__ out,
                                                    const float2* __restrict__ twM,
                                                    const float2* __restrict__ twN) {
    using G = Geo<N>;
    constexpr int M = G::M;
    constexpr int TX = pick_T(M, 16, 16);
    constexpr int EX = M / TX;
    constexpr int RPI = NTHR / TX;
    constexpr int ITERS = (N + RPI - 1) / RPI;
    static_assert(TX > 0 && NTHR % TX == 0, "row plan");
    __shared__ float2 plane[N * G::S];

    const int iz = blockIdx.x;
    const int b = blockIdx.y;
    const int gx = threadIdx.x / TX, tx = threadIdx.x % TX;
    double R[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = rots[9 * b + i];
    const float* tmpl = reinterpret_cast<const float*>(tm);

    float w[ITERS][EX][2];
    float2* dst = out + ((int64_t)b * 3) * G::H + (int64_t)iz * G::PLANE;

    // Row pass for grid g (0: P, 1: W, 2: W^2): fill v, x FFT, post-process.
#pragma unroll
    for (int g = 0; g < (BIN ? 2 : 3); ++g) {
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int row = it * RPI + gx;
            const bool active = row < N;
            float2 v[EX];
            if (active) {
#pragma unroll
                for (int e = 0; e < EX; ++e) {
                    const int m = tx + TX * e;
                    float pv[2];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (g == 0) {
                            int z, y, x;
                            float p = 0.f, wv = 0.f;
                            if (signed_triple(iz, row, 2 * m + h, N, N, N, r, &z, &y, &x)) {
                                const Coord c = source_coord(R, z, y, x);
                                p = (float)sample_trilinear<2>(tmpl, N, N, N, c);
                                wv = (float)sample_nearest<2>(tmpl + 1, N, N, N, c);
                            }
                            pv[h] = p;
                            w[it][e][h] = wv;
                        } else if (g == 1) {
                            pv[h] = w[it][e][h];
                        } else {
                            pv[h] = w[it][e][h] * w[it][e][h];
                        }
                    }
                    v[e] = c2(pv[0], pv[1]);
                }
            }
            float2* rowp = plane + row * G::S;
            group_fft<M, TX, -1>(v, active ? rowp : plane, 1, tx, twM, active);
            // R2C post-processing: X[k] = E[k] + w^k O[k]; X0, XM packed in column 0
            if (active) {
#pragma unroll
                for (int e = 0; e < EX; ++e) rowp[tx + TX * e] = v[e];
            }
            __syncthreads();
            float2 X[EX];
            if (active) {
#pragma unroll
                for (int e = 0; e < EX; ++e) {
                    const int k = tx + TX * e;
                    const float2 zk = rowp[k];
                    if (k == 0) {
                        X[e] = c2(zk.x + zk.y, zk.x - zk.y);
                    } else {
                        const float2 zm = cconj(rowp[M - k]);
                        const float2 Ev = cscale(cadd(zk, zm), 0.5f);
                        const float2 Od = mul_si<-1>(cscale(csub(zk, zm), 0.5f));
                        // exp(-2 pi i k / N) = twN[k]
                        X[e] = cadd(Ev, cmul(twN[k], Od));
                    }
                }
            }
            __syncthreads();
            if (active) {
#pragma unroll
                for (int e = 0; e < EX; ++e) rowp[tx + TX * e] = X[e];
            }
        }
        __syncthreads();
        plane_columns<N, NTHR, -1>(plane, twN);
        store_plane<N, NTHR>(plane, dst + (int64_t)g * G::H);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
template <int N, int TQ, int EMAX, bool BIN>
__global__ __launch_bounds__(TQ * pick_T(N, EMAX, 32)) void k_zcorr(
    float2* __restrict__ X, const float2* __restrict__ F1, const float2* __restrict__ F2,
    int nrot, int chunk, const float2* __restrict__ twN) {
    using G = Geo<N>;
    constexpr int TZ = pick_T(N, EMAX, 32);
    constexpr int EZ = N / TZ;
    __shared__ float2 ex[N * TQ];
    const int c = threadIdx.x % TQ, t = threadIdx.x / TQ;
    const int q = blockIdx.x * TQ + c;
    const bool active = q < G::PLANE;
    const int b0 = blockIdx.y * chunk;
    const int b1 = min(nrot, b0 + chunk);
    float2 f1[EZ], f2[EZ];
    if (active) {
#pragma unroll
        for (int e = 0; e < EZ; ++e) {
            f1[e] = F1[(int64_t)(t + TZ * e) * G::PLANE + q];
            f2[e] = F2[(int64_t)(t + TZ * e) * G::PLANE + q];
        }
    }
    float2* exc = ex + c;
    for (int b = b0; b < b1; ++b) {
        float2* base = X + (int64_t)b * 3 * G::H + q;
#pragma unroll
        for (int g = 0; g < (BIN ? 2 : 3); ++g) {
            float2* col = base + (int64_t)g * G::H;
            float2 v[EZ];
            if (active) {
#pragma unroll
                for (int e = 0; e < EZ; ++e) v[e] = col[(int64_t)(t + TZ * e) * G::PLANE];
            }
            group_fft<N, TZ, -1>(v, exc, TQ, t, twN, active);
            if (BIN && g == 1) {
                // W spectrum feeds both A (with F1) and B (with F2)
                float2 u[EZ];
#pragma unroll
                for (int e = 0; e < EZ; ++e) u[e] = cmul(cconj(v[e]), f2[e]);
#pragma unroll
                for (int e = 0; e < EZ; ++e) v[e] = cmul(cconj(v[e]), f1[e]);
                group_fft<N, TZ, 1>(v, exc, TQ, t, twN, active);
                group_fft<N, TZ, 1>(u, exc, TQ, t, twN, active);
                if (active) {
#pragma unroll
                    for (int e = 0; e < EZ; ++e) {
                        col[(int64_t)(t + TZ * e) * G::PLANE] = v[e];
                        col[G::H + (int64_t)(t + TZ * e) * G::PLANE] = u[e];
                    }
                }
            } else {
#pragma unroll
                for (int e = 0; e < EZ; ++e) v[e] = cmul(cconj(v[e]), g == 2 ? f2[e] : f1[e]);
                group_fft<N, TZ, 1>(v, exc, TQ, t, twN, active);
                if (active) {
#pragma unroll
                    for (int e = 0; e < EZ; ++e) col[(int64_t)(t + TZ * e) * G::PLANE] = v[e];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
template <int N, int NTHR>
__global__ __launch_bounds__(NTHR) void k_inv_plane(const float2* __restrict__ X,
                                                    const uint8_t* __restrict__ lmask,
                                                    unsigned long long* __restrict__ keys,
                                                    int nrot, int chunk, int64_t first,
                                                    double norm, double invV,
                                                    const float2* __restrict__ twM,
                                                    const float2* __restrict__ twN) {
    using G = Geo<N>;
    constexpr int M = G::M;
    constexpr int TX = pick_T(M, 16, 16);
    constexpr int EX = M / TX;
    constexpr int RPI = NTHR / TX;
    constexpr int ITERS = (N + RPI - 1) / RPI;
    static_assert(ITERS * EX * 2 <= 32, "owned voxels per thread must fit a 32-bit mask");
    __shared__ float2 plane[N * G::S];
    // chunk-local running max (strict '>') and its rotation offset, per voxel
    // of the plane; 0xff = no finite value seen yet.  Each entry is touched
    // only by the thread owning that voxel.
    __shared__ float sbest[N * N];
    __shared__ uint8_t sidx[N * N];

    const int iz = blockIdx.x;
    const int b0 = blockIdx.y * chunk;
    const int b1 = min(nrot, b0 + chunk);
    const int gx = threadIdx.x / TX, tx = threadIdx.x % TX;
    const float invVf = (float)invV;

    // owned voxels: (row = it*RPI + gx, x = 2 (tx + TX e) + h), bit 2(it*EX+e)+h
    uint32_t mbits = 0;
    float isd[ITERS][EX][2];
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int row = it * RPI + gx;
#pragma unroll
        for (int e = 0; e < EX; ++e) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (row < N) {
                    const int x = 2 * (tx + TX * e) + h;
                    const int64_t vox = ((int64_t)iz * N + row) * N + x;
                    if (lmask[vox]) mbits |= 1u << (2 * (it * EX + e) + h);
                    sidx[row * N + x] = 0xff;
                }
            }
        }
    }

    for (int b = b0; b < b1; ++b) {
#pragma unroll
        for (int gi = 0; gi < 3; ++gi) {
            const int g = 2 - gi;   // B (ave2), then A (ave), then G (gcc)
            load_plane<N, NTHR>(plane, X + ((int64_t)b * 3 + g) * G::H + (int64_t)iz * G::PLANE);
            __syncthreads();
            plane_columns<N, NTHR, 1>(plane, twN);
#pragma unroll
            for (int it = 0; it < ITERS; ++it) {
                const int row = it * RPI + gx;
                const bool active = row < N;
                float2* rowp = plane + (active ? row : 0) * G::S;
                float2 v[EX];
                if (active) {
                    const float2 c0 = rowp[0];   // (X0, XM), both real
#pragma unroll
                    for (int e = 0; e < EX; ++e) {
                        const int k = tx + TX * e;
                        const float2 xk = (k == 0) ? c2(c0.x, 0.f) : rowp[k];
                        const float2 xm = (k == 0) ? c2(c0.y, 0.f) : rowp[M - k];
                        const float2 xmc = cconj(xm);
                        const float2 d = cmul(csub(xk, xmc), cconj(twN[k]));
                        v[e] = cadd(cadd(xk, xmc), mul_si<1>(d));
                    }
                }
                __syncthreads();
                group_fft<M, TX, 1>(v, rowp, 1, tx, twM, active);
                if (active) {
#pragma unroll
                    for (int e = 0; e < EX; ++e) {
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const float val = h ? v[e].y : v[e].x;
                            if (g == 2) {
                                isd[it][e][h] = val;
                            } else if (g == 1) {
                                // N<W^2 T^2> - <W T>^2 in FP64, then 1/sqrt in FP32
                                const double a = (double)val * invV;
                                const double a2 = (double)isd[it][e][h] * invV * norm;
                                isd[it][e][h] = rsqrtf((float)(a2 - a * a));
                            } else if ((mbits >> (2 * (it * EX + e) + h)) & 1u) {
                                const float s = val * invVf * isd[it][e][h];
                                const int li = row * N + 2 * (tx + TX * e) + h;
                                const uint8_t bi = sidx[li];
                                if (bi == 0xff ? (s == s) : (s > sbest[li])) {
                                    sbest[li] = s;
                                    sidx[li] = (uint8_t)(b - b0);
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int row = it * RPI + gx;
#pragma unroll
        for (int e = 0; e < EX; ++e)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (row >= N) continue;
                const int x = 2 * (tx + TX * e) + h;
                const int li = row * N + x;
                if (sidx[li] != 0xff) {
                    const int64_t vox = ((int64_t)iz * N + row) * N + x;
                    atomicMax(keys + vox,
                              lcc_key(sbest[li], (uint32_t)(first + b0 + sidx[li])));
                }
            }
    }
}

__global__ void k_pack_tm(const float* __restrict__ t, const float* __restrict__ m,
                          float2* __restrict__ tm, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tm[i] = c2(t[i], m[i]);
}

// ---------------------------------------------------------------------------
// Host side: per-N configuration and dispatch.

namespace {

struct Tw {
    const float2 *M, *N;
};

int get_tw(int n, Tw* tw) {
    PF_CHECK_STATUS(twiddles(n / 2, &tw->M));
    PF_CHECK_STATUS(twiddles(n, &tw->N));
    return kOk;
}

int last(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return kErrHip;
    }
    return kOk;
}

template <int N, int NTHR1, int TQ, int EMAXZ, int NTHR3>
int run_fused(const FusedArgs& a, int which, hipStream_t s) {
    using G = Geo<N>;
    Tw tw;
    PF_CHECK_STATUS(get_tw(N, &tw));
    if (which < 0 || which == 0) {
        if (a.binary_mask)
            hipLaunchKernelGGL((k_fwd_plane<N, NTHR1, true>), dim3(N, a.nrot), dim3(NTHR1), 0, s,
                               a.tm, a.rots, a.rmax, a.work, tw.M, tw.N);
        else
            hipLaunchKernelGGL((k_fwd_plane<N, NTHR1, false>), dim3(N, a.nrot), dim3(NTHR1), 0, s,
                               a.tm, a.rots, a.rmax, a.work, tw.M, tw.N);
        PF_CHECK_STATUS(last("k_fwd_plane"));
    }
    if (which < 0 || which == 1) {
        const int ch = a.chunk_z;
        const dim3 grid((G::PLANE + TQ - 1) / TQ, (a.nrot + ch - 1) / ch);
        const int nthr = TQ * pick_T(N, EMAXZ, 32);
        if (a.binary_mask)
            hipLaunchKernelGGL((k_zcorr<N, TQ, EMAXZ, true>), grid, dim3(nthr), 0, s, a.work, a.F1,
                               a.F2, a.nrot, ch, tw.N);
        else
            hipLaunchKernelGGL((k_zcorr<N, TQ, EMAXZ, false>), grid, dim3(nthr), 0, s, a.work,
                               a.F1, a.F2, a.nrot, ch, tw.N);
        PF_CHECK_STATUS(last("k_zcorr"));
    }
    if (which < 0 || which == 2) {
        const int ch = a.chunk_lcc;
        hipLaunchKernelGGL((k_inv_plane<N, NTHR3>), dim3(N, (a.nrot + ch - 1) / ch),
                           dim3(NTHR3), 0, s, a.work, a.lcc_mask, a.keys, a.nrot, ch, a.first,
                           a.norm, a.invV, tw.M, tw.N);
        PF_CHECK_STATUS(last("k_inv_plane"));
    }
    return kOk;
}

}  // namespace

bool fused_supported(int nz, int ny, int nx) {
    if (nz != ny || ny != nx) return false;
    switch (nx) {
        case 64: case 96: case 128: return true;
        default: return false;
    }
}

int launch_fused(const FusedArgs& a, int which, hipStream_t s) {
    if (a.nrot <= 0) return kOk;
    if (a.nrot > 65535 || a.chunk_z < 1 || a.chunk_lcc < 1 || a.chunk_lcc > 255) {
        set_error("fused: invalid batch or chunk sizes");
        return kErrArg;
    }
    if (a.rmax < 0 || a.rmax > a.n - 2) {
        set_error("fused: invalid rmax");
        return kErrArg;
    }
    switch (a.n) {
        case 64: return run_fused<64, 256, 32, 8, 256>(a, which, s);
        case 96: return run_fused<96, 384, 32, 12, 384>(a, which, s);
        case 128: return run_fused<128, 512, 32, 8, 512>(a, which, s);
        default:
            set_error("fused pipeline not built for n = %d", a.n);
            return kErrUnsupported;
    }
}

int launch_pack_tm(const float* tmpl, const float* mask, float2* tm, int64_t n, hipStream_t s) {
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_pack_tm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tmpl, mask,
                       tm, n);
    return last("k_pack_tm");
}

}  // namespace pf
