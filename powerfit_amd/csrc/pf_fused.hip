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
//                 S = G / sqrt(N B - A^2) (variance in FP64), merged per
//                 rotation into the running max/argmax with a 64-bit
//                 atomicMax on lcc_key (masked voxels with S > 0 only), which
//                 equals the reference's strict '>' take-best in any order.
//
// k_fwd_plane2 / k_inv_plane2 are two-plane forms of the first and last
// kernel; each size's launch shape is a Shape below.
//
// The DC and Nyquist columns of the x-transform are both real; they travel
// through the y transform packed as one complex column X0 + i XM and are
// separated with Hermitian symmetry on the way out (and packed on the way in
// for the inverse).
#include <algorithm>
#include <type_traits>

#include "pf_group_fft.h"
#include "pf_mfma.h"
#include "pf_internal.h"
#include "pf_rotate.h"

namespace pf {

// LDS row stride (complex elements) for M packed columns: >= M and = 8 mod 32,
// so the rows one wave touches in the row FFTs fall on distinct bank groups.
constexpr int padded_stride(int M) {
    int s = M;
    while (s % 32 != 8) ++s;
    return s;
}

// Padding period of the row FFTs' LDS exchanges (0 = none) for M-point rows
// run by T threads: one free slot after every PAD elements spreads a stage's
// stride-R butterfly writes, and the rows a wavefront covers, over the banks
// (chosen with an LDS bank model of every stage's accesses; the padded row
// fits the row stride: pad_extent(M, PAD) <= padded_stride(M + 1)).  Measured
// per rotation: 128^3 K3 11.0 -> 10.4 us; 64^3 and 96^3 unchanged; 192^3 K3
// (M = 96, PAD 16) 76.4 -> 79.9 us, so that size keeps the plain layout.
constexpr int row_pad(int M, int T) {
    return (M == 32 && T == 4) ? 8
         : (M == 48 && T == 4) ? 4
         : (M == 64 && T == 8) ? 8
         : (M == 80 && T == 4) ? 8
         : 0;
}

// Row FFTs (groups of consecutive lanes, one wavefront) synchronise their
// stages per wavefront.  Column FFTs run lanes across columns (conflict-free
// LDS, coalesced loads) with workgroup barriers, or, with kWaveCols, each
// column's threads as consecutive lanes of one wavefront with wave-local
// stage synchronisation.
constexpr bool kWaveRows = true;
constexpr bool kWaveCols = false;

template <int N>
struct Geo {
    static constexpr int M = N / 2;          // complex points of the half-length x FFT
    static constexpr int NXH = M + 1;        // kx extent in global memory
    static constexpr int S = padded_stride(M + 1);  // LDS row stride (complex), >= M + 1
    static constexpr int NB = N / 8;         // 8^3 bricks per axis (occupancy map)
    static constexpr int PLANE = N * NXH;    // complex per z-plane in global memory
    static constexpr int64_t H = (int64_t)N * PLANE;
    static constexpr int64_t V = (int64_t)N * N * N;
};

// Copy a twiddle table into LDS (read there by the group FFTs: keeps the
// tables out of registers).
template <int L, int NTHR>
__device__ __forceinline__ void stage_twiddles(float2* dst, const float2* __restrict__ src) {
    for (int i = threadIdx.x; i < L; i += NTHR) dst[i] = src[i];
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

// Signed coordinate of index i on an axis of length N for rotation radius
// r <= N / 2: the value the reference's scatter loop writes last, i.e. i for
// i <= r (for i == r == N / 2 both +r and -r map here and +r comes later),
// otherwise i - N.
template <int N>
__device__ __forceinline__ int signed_coord(int i, int r) {
    return i <= r ? i : i - N;
}

// Rotated template (trilinear) and mask (nearest) at signed voxel (z, y, x),
// from the interleaved (template, mask) grid.  Coordinates, floor and round
// are FP64 in the reference's evaluation order (so corner and nearest indices
// are the reference's; the nearest voxel is always one of the eight trilinear
// corners, since round(s) is floor(s) or floor(s) + 1 with the same wrap); the
// interpolation itself runs in FP32 on the FP32 template.  Sources whose
// 2x2x2 corner cube lies in an all-zero region (occ bit clear for the 8^3
// brick of the lower corner, bricks dilated by one voxel towards +x/+y/+z
// with wrap) are exactly zero and skip the loads.
template <int N>
__device__ __forceinline__ void gather_pw(const float2* __restrict__ tm, const uint32_t* occ,
                                          const double* R, int z, int y, int x, float* p,
                                          float* w) {
    using G = Geo<N>;
    const Coord c = source_coord(R, z, y, x);   // FP64, contraction off inside
    const double fx = floor(c.xs), fy = floor(c.ys), fz = floor(c.zs);
    const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
    const int xi = x0 + (x0 < 0 ? N : 0), yi = y0 + (y0 < 0 ? N : 0), zi = z0 + (z0 < 0 ? N : 0);
    *p = 0.f;
    *w = 0.f;
    if ((unsigned)xi >= (unsigned)N || (unsigned)yi >= (unsigned)N || (unsigned)zi >= (unsigned)N)
        return;
    const int brick = ((zi >> 3) * G::NB + (yi >> 3)) * G::NB + (xi >> 3);
    if (!((occ[brick >> 5] >> (brick & 31)) & 1u)) return;
    const int xj = (xi + 1 == N) ? 0 : xi + 1;
    const int yj = (yi + 1 == N) ? 0 : yi + 1;
    const int zj = (zi + 1 == N) ? 0 : zi + 1;
    const float2* r00 = tm + (zi * N + yi) * N;
    const float2* r10 = tm + (zi * N + yj) * N;
    const float2* r01 = tm + (zj * N + yi) * N;
    const float2* r11 = tm + (zj * N + yj) * N;
    const float2 a000 = r00[xi], a001 = r00[xj], a010 = r10[xi], a011 = r10[xj];
    const float2 a100 = r01[xi], a101 = r01[xj], a110 = r11[xi], a111 = r11[xj];
    const float dx = (float)(c.xs - fx), dy = (float)(c.ys - fy), dz = (float)(c.zs - fz);
    const float c00 = a000.x + (a001.x - a000.x) * dx;
    const float c10 = a010.x + (a011.x - a010.x) * dx;
    const float c01 = a100.x + (a101.x - a100.x) * dx;
    const float c11 = a110.x + (a111.x - a110.x) * dx;
    const float c0 = c00 + (c10 - c00) * dy;
    const float c1 = c01 + (c11 - c01) * dy;
    *p = c0 + (c1 - c0) * dz;
    // round() is half away from zero: it moves up from floor when the
    // fraction exceeds one half, or equals it for non-negative values
    const double hx = c.xs - fx, hy = c.ys - fy, hz = c.zs - fz;
    const bool ux = hx > 0.5 || (hx == 0.5 && c.xs >= 0.0);
    const bool uy = hy > 0.5 || (hy == 0.5 && c.ys >= 0.0);
    const bool uz = hz > 0.5 || (hz == 0.5 && c.zs >= 0.0);
    const float2 n0 = uy ? (ux ? a011 : a010) : (ux ? a001 : a000);
    const float2 n1 = uy ? (ux ? a111 : a110) : (ux ? a101 : a100);
    *w = uz ? n1.y : n0.y;
}

// ---------------------------------------------------------------------------
// WSG: the mask values W of each thread's row voxels, gathered with P for grid 0
// and reused for grids 1 and 2, are kept in global memory instead of registers,
// in this plane's region of grid 2 (written last, by this block only, and
// only after every read of it; each thread reads back only what it wrote).
//
// EMAXC: element bound of the column pass.  It runs in one sweep over the M
// columns: the post-processing of column k reads columns k and M - k before
// the FFT overwrites them.
//
// GONLY >= 0: this launch transforms grid GONLY only (the grids then run as
// separate launches, grid 0 first; needs WSG).  A rolled loop over the grids
// keeps every loop-invariant address live across its large body.
template <int N, int NTHR, bool BIN, int EMAX, bool WSG = false, int EMAXC = EMAX,
          int GONLY = -1>
__global__ __launch_bounds__(NTHR) void k_fwd_plane(const float2* __restrict__ tm,
                                                    const double* __restrict__ rots, int r,
                                                    int zr, int zlo,
                                                    float2* __restrict__ out,
                                                    const float2* __restrict__ twM,
                                                    const float2* __restrict__ twN,
                                                    const uint32_t* __restrict__ occ_g) {
    using G = Geo<N>;
    constexpr int M = G::M;
    constexpr int TX = pick_T(M, EMAX, 16);
    constexpr int EX = M / TX;
    constexpr int RPI = NTHR / TX;
    constexpr int ITERS = (N + RPI - 1) / RPI;
    constexpr int OCCW = (G::NB * G::NB * G::NB + 31) / 32;
    constexpr int TY = pick_T_cols(N, M, NTHR, EMAXC);
    constexpr int EY = N / TY;
    constexpr int CPI = NTHR / TY;
    constexpr bool WROWS = kWaveRows && (64 % TX == 0);
    constexpr int XPAD = row_pad(M, TX);
    static_assert(pad_extent(M, XPAD) <= G::S, "padded row fits the row stride");
    constexpr bool WCOLS = kWaveCols && (64 % TY == 0);
    static_assert(TX > 0 && NTHR % TX == 0, "row plan");
    static_assert(TY > 0 && NTHR % TY == 0 && CPI >= M, "one column pass");
    __shared__ float2 plane[N * G::S];
    __shared__ uint32_t occ[OCCW];
    __shared__ float2 stwM[M], stwN[N];
    for (int i = threadIdx.x; i < OCCW; i += NTHR) occ[i] = occ_g[i];
    stage_twiddles<M, NTHR>(stwM, twM);
    stage_twiddles<N, NTHR>(stwN, twN);
    __syncthreads();

    // planes zlo, zlo + 1, ... (mod N); rows outside the support disc are zero
    const int iz = (blockIdx.x + zlo + N) % N;
    const int b = blockIdx.y;
    const int gx = threadIdx.x / TX, tx = threadIdx.x % TX;
    // column pass: lanes run across columns (conflict-free LDS)
    const int cg = WCOLS ? threadIdx.x / TY : threadIdx.x % CPI;
    const int ty = WCOLS ? threadIdx.x % TY : threadIdx.x / CPI;
    const bool cact = cg < M;
    const int sz = signed_coord<N>(iz, r);
    // gathered voxels: inside the rotation sphere and the support ball (|v| >
    // zr reads only zeros whatever the rotation)
    const int rg = r < zr ? r : zr, rg2 = rg * rg;
    double R[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = rots[9 * b + i];
    float w[WSG ? 1 : ITERS][WSG ? 1 : EX][2];
    float2* dst = out + ((int64_t)b * 3) * G::H + (int64_t)iz * G::PLANE;
    static_assert(N * M <= G::PLANE, "grid 2's region holds one float2 per voxel pair");

    // Grid g (0: P, 1: W, 2: W^2): gather the rows and run their half-length x
    // FFT, then the column pass (R2C post-processing, y FFT) and the store.
    static_assert(GONLY < 0 || WSG, "separate grid launches share W through global memory");
    constexpr int G0 = GONLY < 0 ? 0 : GONLY;
    constexpr int G1 = GONLY < 0 ? (BIN ? 2 : 3) : GONLY + 1;
#pragma unroll
    for (int g = G0; g < G1; ++g) {
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int row = it * RPI + gx;
            const int sy = signed_coord<N>(row, r);
            const bool in_plane = row < N;
            // rows whose every voxel lies outside the support are exactly zero
            const bool active = in_plane && sz * sz + sy * sy <= zr * zr;
            if (in_plane && !active) {
                float2* rowz = plane + row * G::S;
#pragma unroll
                for (int e = 0; e < EX; ++e) rowz[tx + TX * e] = c2(0.f, 0.f);
            }
            float2 v[EX];
            if (active && WSG) {
                float2* wst = dst + 2 * G::H + row * M + tx;
#pragma unroll
                for (int e = 0; e < EX; ++e) {
                    const int m = tx + TX * e;
                    if (g == 0) {
                        float pv[2], wv[2];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int x = signed_coord<N>(2 * m + h, r);
                            pv[h] = 0.f;
                            wv[h] = 0.f;
                            if (sz * sz + sy * sy + x * x <= rg2)
                                gather_pw<N>(tm, occ, R, sz, sy, x, &pv[h], &wv[h]);
                        }
                        v[e] = c2(pv[0], pv[1]);
                        wst[TX * e] = c2(wv[0], wv[1]);
                        // keep the scheduler from hoisting every element's
                        // gather loads at once (they exhaust the registers)
                        __builtin_amdgcn_sched_barrier(0);
                    } else {
                        const float2 wv = wst[TX * e];
                        v[e] = g == 1 ? wv : c2(wv.x * wv.x, wv.y * wv.y);
                    }
                }
            } else if (active) {
#pragma unroll
                for (int e = 0; e < EX; ++e) {
                    const int m = tx + TX * e;
                    float pv[2];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (g == 0) {
                            const int x = signed_coord<N>(2 * m + h, r);
                            float p = 0.f, wv = 0.f;
                            if (sz * sz + sy * sy + x * x <= rg2)
                                gather_pw<N>(tm, occ, R, sz, sy, x, &p, &wv);
                            pv[h] = p;
                            w[WSG ? 0 : it][WSG ? 0 : e][h] = wv;
                        } else if (g == 1) {
                            pv[h] = w[WSG ? 0 : it][WSG ? 0 : e][h];
                        } else {
                            pv[h] = w[WSG ? 0 : it][WSG ? 0 : e][h] * w[WSG ? 0 : it][WSG ? 0 : e][h];
                        }
                    }
                    v[e] = c2(pv[0], pv[1]);
                }
            }
            float2* rowp = plane + row * G::S;
            // row FFT; its result stays in the row (natural order)
            group_fft_ex<M, TX, -1, WROWS, XPAD>(v, active ? rowp : plane, 1, tx, stwM, active);
        }
        __syncthreads();
        {
            // column pass: R2C post-processing on the way in,
            // X[k] = E[k] + w^k O[k] from z[k] and conj(z[M - k]) of each row,
            // X0 and XM packed in column 0; then the y FFT, left in the plane
            float2 v[EY];
            if (cact) {
#pragma unroll
                for (int e = 0; e < EY; ++e) {
                    const float2* rowp = plane + (ty + TY * e) * G::S;
                    const float2 zk = rowp[cg];
                    if (cg == 0) {
                        v[e] = c2(zk.x + zk.y, zk.x - zk.y);
                    } else {
                        const float2 zm = cconj(rowp[M - cg]);
                        const float2 Ev = cscale(cadd(zk, zm), 0.5f);
                        const float2 Od = mul_si<-1>(cscale(csub(zk, zm), 0.5f));
                        // exp(-2 pi i k / N) = twN[k]
                        v[e] = cadd(Ev, cmul(stwN[cg], Od));
                    }
                }
            }
            __syncthreads();   // mirror reads are done before the FFT overwrites columns
            group_fft_ex<N, TY, -1, WCOLS>(v, plane + (cact ? cg : 0), G::S, ty, stwN, cact);
        }
        __syncthreads();
        store_plane<N, NTHR>(plane, dst + (int64_t)g * G::H);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// SPLIT: the forward z FFTs do their first stage, one radix-16 butterfly per
// thread, as a split-precision DFT-GEMM on the matrix cores (pf_mfma.h): a
// wavefront holds butterfly rows j0 = t & ~1 and j0 + 1 of its 32 columns,
// and each of the two 32 x 32 GEMMs takes one row, lane (c, h = t & 1)
// holding inputs 4 h + q and 8 + 4 h + q of that row's butterfly in column
// c.  The remaining stages, the products and the inverse transforms stay
// FP32 on the VALU.
template <int N, int TQ, int EMAX, bool BIN, bool SPLIT = false>
__global__ __launch_bounds__(TQ * pick_T(N, EMAX, 32)) void k_zcorr(
    float2* __restrict__ X, const float2* __restrict__ F1, const float2* __restrict__ F2,
    int nrot, int chunk, int zr, const float2* __restrict__ twN) {
    using G = Geo<N>;
    constexpr int TZ = pick_T(N, EMAX, 32);
    constexpr int EZ = N / TZ;
    static_assert(!SPLIT || (TQ == 32 && TZ % 2 == 0 && EZ == 16 &&
                             gplan_for(N, TZ).r[0] == 16),
                  "split first stage: one radix-16 butterfly per thread, 32-column tiles");
    // lanes run across columns (coalesced column loads); stages use
    // workgroup barriers (wave-local columns cost more in scattered loads)
    __shared__ float2 ex[N * TQ];
    __shared__ float2 stw[N];
    stage_twiddles<N, TQ * TZ>(stw, twN);
    __syncthreads();
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
    half4 fh[4], fl[4];
    if constexpr (SPLIT) dft16_operand<-1>(fh, fl);
    for (int b = b0; b < b1; ++b) {
        float2* base = X + (int64_t)b * 3 * G::H + q;
#pragma unroll
        for (int g = 0; g < (BIN ? 2 : 3); ++g) {
            float2* col = base + (int64_t)g * G::H;
            // planes beyond the support were not written: they are zero
            auto ld = [&](int z) -> float2 {
                const int sz = z <= N / 2 ? z : z - N;
                return (sz >= -zr && sz <= zr) ? col[(int64_t)z * G::PLANE] : c2(0.f, 0.f);
            };
            float2 v[EZ];
            if constexpr (SPLIT) {
                constexpr int NB = N / 16;   // stage-0 butterflies per column
                const int h = t & 1, j0 = t & ~1;
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const int j = j0 + jj;
                    float2 x[8];
#pragma unroll
                    for (int qq = 0; qq < 4; ++qq) {
                        x[qq] = active ? ld(j + NB * (4 * h + qq)) : c2(0.f, 0.f);
                        x[4 + qq] = active ? ld(j + NB * (8 + 4 * h + qq)) : c2(0.f, 0.f);
                    }
                    dft16_split_gemm(x, fh, fl);
                    if (active) {
#pragma unroll
                        for (int qq = 0; qq < 4; ++qq) {
                            exc[(16 * j + 4 * h + qq) * TQ] = x[qq];
                            exc[(16 * j + 8 + 4 * h + qq) * TQ] = x[4 + qq];
                        }
                    }
                }
                __syncthreads();
                if (active) GroupStage<N, TZ, -1, 0>::read_phase(v, exc, TQ, t);
                __syncthreads();
                GroupStage<N, TZ, -1, 1>::run(v, exc, TQ, t, stw, active);
            } else {
                if (active) {
#pragma unroll
                    for (int e = 0; e < EZ; ++e) v[e] = ld(t + TZ * e);
                }
                group_fft<N, TZ, -1>(v, exc, TQ, t, stw, active);
            }
            if (BIN && g == 1) {
                // W spectrum feeds both A (with F1) and B (with F2)
                float2 u[EZ];
#pragma unroll
                for (int e = 0; e < EZ; ++e) u[e] = cmul(cconj(v[e]), f2[e]);
#pragma unroll
                for (int e = 0; e < EZ; ++e) v[e] = cmul(cconj(v[e]), f1[e]);
                group_fft<N, TZ, 1>(v, exc, TQ, t, stw, active);
                group_fft<N, TZ, 1>(u, exc, TQ, t, stw, active);
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
                group_fft<N, TZ, 1>(v, exc, TQ, t, stw, active);
                if (active) {
#pragma unroll
                    for (int e = 0; e < EZ; ++e) col[(int64_t)(t + TZ * e) * G::PLANE] = v[e];
                }
            }
        }
    }
}

// 1 / sqrt of the local variance from the raw (unnormalised) inverse-FFT
// outputs A = V sum W T and B = V sum W^2 T^2 of one voxel: with nv = N V,
// N V B - A^2 = V^2 (N <W^2 T^2> - <W T>^2).  B nv (24 x 45 bits) and A^2
// (48 bits) are exact in FP64, so the difference has a single rounding; the
// LCC is then G_raw * rsq (the factors V cancel).  Raw v_rsq_f32: the
// argument is 0 (-> +inf), negative (-> NaN) or a normal float, never a
// denormal (FP32 FFT outputs are 0 or far above 1e-38).
__device__ __forceinline__ float rsq_var(float a, float b, double nv) {
    const double ad = (double)a;
    return __builtin_amdgcn_rsqf((float)fma(-ad, ad, (double)b * nv));
}

// ---------------------------------------------------------------------------
// XPF (one-sweep column plans only): the next plane is loaded while the current
// one is transformed (costs EY registers through the row phase); otherwise a
// plane's first sweep is loaded when the plane starts and each later sweep
// while the previous one is transformed.  ISDG: B, then
// 1/sqrt(N B - A^2), of each owned voxel is kept in global memory instead of
// registers, in the B plane's own region of X (this block is its only reader,
// and the plane has been read in full before the first value is written);
// each thread writes and reads back its own pairs, so no barrier is needed.
//
// RECOMP: the thread coordinates pass through an empty asm at each plane, so
// the compiler recomputes the addresses derived from them there instead of
// holding all of them in registers across the plane loop (fewer registers,
// more address arithmetic; faster at some sizes only).
template <int N, int NTHR, int EMAX, int EMAXC = EMAX, bool XPF = true, bool ISDG = false,
          bool RECOMP = false>
__global__ __launch_bounds__(NTHR) void k_inv_plane(typename std::conditional<ISDG, float2*, const float2*>::type __restrict__ X,
                                                    const uint8_t* __restrict__ lmask,
                                                    unsigned long long* __restrict__ keys,
                                                    int nrot, int chunk, int64_t first,
                                                    double norm,
                                                    const float2* __restrict__ twM,
                                                    const float2* __restrict__ twN) {
    using G = Geo<N>;
    constexpr int M = G::M;
    constexpr int TX = pick_T(M, EMAX, 16);
    constexpr int EX = M / TX;
    constexpr int RPI = NTHR / TX;
    constexpr int ITERS = (N + RPI - 1) / RPI;
    // the column pass runs in CSW sweeps of CPI columns (fewer registers per
    // thread than one sweep over all M columns)
    constexpr int TY = pick_T_cols(N, M, NTHR, EMAXC);
    constexpr int EY = N / TY;
    constexpr int CPI = NTHR / TY;
    constexpr int CSW = (M + CPI - 1) / CPI;
    constexpr bool WROWS = kWaveRows && (64 % TX == 0);
    constexpr int XPAD = row_pad(M, TX);
    static_assert(pad_extent(M, XPAD) <= G::S, "padded row fits the row stride");
    constexpr bool WCOLS = kWaveCols && (64 % TY == 0);
    static_assert(TY > 0 && NTHR % TY == 0, "column plan");
    static_assert(NTHR >= N, "one Nyquist element per thread");
    static_assert(ITERS * EX * 2 <= 64, "owned voxels per thread must fit a 64-bit mask");
    static_assert(N * M <= G::PLANE, "the B plane's region holds one float2 per voxel pair");
    __shared__ float2 plane[N * G::S];
    __shared__ float2 colM[N];
    __shared__ float2 stwM[M], stwN[N];
    stage_twiddles<M, NTHR>(stwM, twM);
    stage_twiddles<N, NTHR>(stwN, twN);

    const int iz = blockIdx.x;
    const int b0 = blockIdx.y * chunk;
    const int b1 = min(nrot, b0 + chunk);
    const int gx0 = threadIdx.x / TX, tx0 = threadIdx.x % TX;
    // column pass: lanes run across columns (coalesced loads, conflict-free LDS)
    const int cg0 = WCOLS ? threadIdx.x / TY : threadIdx.x % CPI;
    const int ty0 = WCOLS ? threadIdx.x % TY : threadIdx.x / CPI;
    const double nv = norm * (double)G::V;   // N V (exact)

    // owned voxels: (row = it*RPI + gx0, x = 2 (tx0 + TX e) + h), bit 2(it*EX+e)+h
    using MaskT = typename std::conditional<(ITERS * EX * 2 <= 32), uint32_t, uint64_t>::type;
    MaskT mbits = 0;
    constexpr int IR = ISDG ? 1 : ITERS, ER = ISDG ? 1 : EX;
    float isd[IR][ER][2];
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int row = it * RPI + gx0;
#pragma unroll
        for (int e = 0; e < EX; ++e)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (row < N) {
                    const int64_t vox = ((int64_t)iz * N + row) * N + 2 * (tx0 + TX * e) + h;
                    if (lmask[vox]) mbits |= MaskT(1) << (2 * (it * EX + e) + h);
                }
    }

    // Planes are visited in the order (b, B), (b, A), (b, G), (b + 1, B), ...
    // Each plane goes from global memory straight into the column registers
    // (sweep s: column cg0 + s CPI, rows ty0 + TY e; the Nyquist column kx = M,
    // one element per thread, through a small LDS buffer); loads are issued
    // ahead of the transform that needs them (see XPF above).
    const int nplanes = 3 * (b1 - b0);
    auto src_of = [&](int p) {
        const int b = b0 + p / 3, g = 2 - p % 3;
        return X + ((int64_t)b * 3 + g) * G::H + (int64_t)iz * G::PLANE;
    };
    constexpr bool ONE = CSW == 1 && XPF;   // one sweep, next plane prefetched
    float2 pf[EY];
    float2 pm = c2(0.f, 0.f);
    const bool cact0 = cg0 < M;
    auto prefetch = [&](const float2* src) {   // the ONE form
        if (cact0) {
#pragma unroll
            for (int e = 0; e < EY; ++e) pf[e] = src[(ty0 + TY * e) * G::NXH + cg0];
        }
        if (threadIdx.x < N) pm = src[threadIdx.x * G::NXH + M];
    };
    auto load_sweep = [&](const float2* src, int sw) {
        const int col = cg0 + sw * CPI;
        if (col < M) {
#pragma unroll
            for (int e = 0; e < EY; ++e) pf[e] = src[(ty0 + TY * e) * G::NXH + col];
        }
        if (sw == 0 && threadIdx.x < N) pm = src[threadIdx.x * G::NXH + M];
    };
    if (ONE && nplanes > 0) prefetch(src_of(0));
    for (int p = 0; p < nplanes; ++p) {
        const int b = b0 + p / 3, g = 2 - p % 3;
        int opq = 0;
        if constexpr (RECOMP) asm volatile("" : "+v"(opq));
        const int cg = cg0 + opq, ty = ty0 + opq, tx = tx0 + opq, gx = gx0 + opq;
        __syncthreads();   // previous plane's row phase is done with the LDS
        if (!ONE) load_sweep(src_of(p), 0);
        if (threadIdx.x < N) colM[threadIdx.x] = pm;
        __syncthreads();
        // y inverse of the M packed columns: column 0 carries X0 + i XM
        if constexpr (ONE) {
            float2 v[EY];
#pragma unroll
            for (int e = 0; e < EY; ++e) {
                v[e] = pf[e];
                if (cg == 0) v[e] = cadd(v[e], mul_si<1>(colM[ty + TY * e]));
            }
            if (p + 1 < nplanes) prefetch(src_of(p + 1));
            group_fft_ex<N, TY, 1, WCOLS>(v, plane + (cact0 ? cg : 0), G::S, ty, stwN, cact0);
        } else {
#pragma unroll
            for (int sw = 0; sw < CSW; ++sw) {
                const int col = cg + sw * CPI;
                const bool cact = col < M;
                float2 v[EY];
#pragma unroll
                for (int e = 0; e < EY; ++e) {
                    v[e] = pf[e];
                    if (col == 0) v[e] = cadd(v[e], mul_si<1>(colM[ty + TY * e]));
                }
                if (sw + 1 < CSW) load_sweep(src_of(p), sw + 1);
                group_fft_ex<N, TY, 1, WCOLS>(v, plane + (cact ? col : 0), G::S, ty, stwN,
                                              cact);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int row = it * RPI + gx;
            const bool active = row < N;
            float2* rowp = plane + (active ? row : 0) * G::S;
            float2 vr[EX];
            if (active) {
                const float2 c0 = rowp[0];   // (X0, XM), both real
#pragma unroll
                for (int e = 0; e < EX; ++e) {
                    const int k = tx + TX * e;
                    const float2 xk = (k == 0) ? c2(c0.x, 0.f) : rowp[k];
                    const float2 xm = (k == 0) ? c2(c0.y, 0.f) : rowp[M - k];
                    const float2 xmc = cconj(xm);
                    const float2 d = cmul(csub(xk, xmc), cconj(stwN[k]));
                    vr[e] = cadd(cadd(xk, xmc), mul_si<1>(d));
                }
            }
            // the row's own threads read it above and overwrite it below
            group_sync<WROWS>();
            group_fft<M, TX, 1, WROWS, XPAD>(vr, rowp, 1, tx, stwM, active);
            if constexpr (ISDG) {
                if (!active) continue;
                // B, then 1/sqrt(var), of voxels x = 2m, 2m + 1 in the B plane's region
                float2* slot = const_cast<float2*>(X) + ((int64_t)b * 3 + 2) * G::H +
                               (int64_t)iz * G::PLANE + row * M + tx;
                if (g == 2) {
#pragma unroll
                    for (int e = 0; e < EX; ++e) slot[TX * e] = vr[e];
                } else if (g == 1) {
#pragma unroll
                    for (int e = 0; e < EX; ++e) {
                        // N V B - A^2 in FP64 (exact products), then 1/sqrt in FP32
                        float2 held = slot[TX * e];
                        held.x = rsq_var(vr[e].x, held.x, nv);
                        held.y = rsq_var(vr[e].y, held.y, nv);
                        slot[TX * e] = held;
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < EX; ++e) {
                        const float2 held = slot[TX * e];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            if (!((mbits >> (2 * (it * EX + e) + h)) & MaskT(1))) continue;
                            // only S > 0 can beat the initial (0, 0); NaN never wins
                            const float sv = (h ? vr[e].y : vr[e].x) * (h ? held.y : held.x);
                            if (sv > 0.f) {
                                const int64_t vox =
                                    ((int64_t)iz * N + row) * N + 2 * (tx + TX * e) + h;
                                atomicMax(keys + vox, lcc_key(sv, (uint32_t)(first + b)));
                            }
                        }
                    }
                }
            } else if (active) {
                // one straight-line loop per grid (a grid test per element
                // makes the compiler copy held values between branches)
                if (g == 2) {
#pragma unroll
                    for (int e = 0; e < EX; ++e) {
                        isd[ISDG ? 0 : it][ISDG ? 0 : e][0] = vr[e].x;
                        isd[ISDG ? 0 : it][ISDG ? 0 : e][1] = vr[e].y;
                    }
                } else if (g == 1) {
#pragma unroll
                    for (int e = 0; e < EX; ++e) {
                        // N V B - A^2 in FP64 (exact products), then 1/sqrt in FP32
                        float* is = isd[ISDG ? 0 : it][ISDG ? 0 : e];
                        is[0] = rsq_var(vr[e].x, is[0], nv);
                        is[1] = rsq_var(vr[e].y, is[1], nv);
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < EX; ++e) {
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            if (!((mbits >> (2 * (it * EX + e) + h)) & MaskT(1))) continue;
                            // only S > 0 can beat the initial (0, 0); NaN never wins
                            const float sv = (h ? vr[e].y : vr[e].x) *
                                             isd[ISDG ? 0 : it][ISDG ? 0 : e][h];
                            if (sv > 0.f) {
                                const int64_t vox =
                                    ((int64_t)iz * N + row) * N + 2 * (tx + TX * e) + h;
                                atomicMax(keys + vox, lcc_key(sv, (uint32_t)(first + b)));
                            }
                        }
                    }
                }
            }
        }
    }
}


// ---------------------------------------------------------------------------
// Two planes per pass (the forward planes of P and W for 0/1 masks).  The
// column pass covers the 2M packed columns of both planes at once, so every
// thread runs a longer register FFT with fewer LDS exchange stages, and the
// pass needs half the workgroup barriers per plane; each thread transforms its
// row of both planes.  Used where one column pass and one row pass cover both
// planes and both fit in LDS.
template <int N, int NTHR, int EMAX>
struct PairPlan {
    static constexpr int M = N / 2;
    static constexpr int S = padded_stride(M + 1);
    static constexpr int TY = pick_T_cols(N, 2 * M, NTHR, EMAX);
    static constexpr int EY = TY > 0 ? N / TY : 1;
    static constexpr int CPI = TY > 0 ? NTHR / TY : 0;
    static constexpr int TX = pick_T(M, EMAX, 16);
    static constexpr int EX = TX > 0 ? M / TX : 1;
    static constexpr int RPI = TX > 0 ? NTHR / TX : 0;
    static constexpr int NB = N / 8;
    static constexpr int LDS = 8 * (2 * N * S + M + N) + 4 * ((NB * NB * NB + 31) / 32);
    static constexpr bool ok = TY > 0 && NTHR % TY == 0 && CPI >= 2 * M && TX > 0 &&
                               NTHR % TX == 0 && RPI >= N && NTHR >= N && EX * 2 <= 32 &&
                               LDS <= 160 * 1024 - 512;
};

// Forward planes of P and W for 0/1 masks (W^2 = W): one gather fills the
// rows of both, then one column pass covers both.
template <int N, int NTHR, int EMAX>
__global__ __launch_bounds__(NTHR) void k_fwd_plane2(const float2* __restrict__ tm,
                                                     const double* __restrict__ rots, int r,
                                                     int zr, int zlo, float2* __restrict__ out,
                                                     const float2* __restrict__ twM,
                                                     const float2* __restrict__ twN,
                                                     const uint32_t* __restrict__ occ_g) {
    using G = Geo<N>;
    using P = PairPlan<N, NTHR, EMAX>;
    constexpr int M = G::M, S = G::S, TY = P::TY, EY = P::EY, CPI = P::CPI, TX = P::TX,
                  EX = P::EX;
    constexpr int OCCW = (G::NB * G::NB * G::NB + 31) / 32;
    constexpr bool WROWS = kWaveRows && (64 % TX == 0);
    constexpr int XPAD = row_pad(M, TX);
    static_assert(pad_extent(M, XPAD) <= G::S, "padded row fits the row stride");
    static_assert(P::ok, "pair plan");
    __shared__ float2 planes[2 * N * S];
    __shared__ uint32_t occ[OCCW];
    __shared__ float2 stwM[M], stwN[N];
    for (int i = threadIdx.x; i < OCCW; i += NTHR) occ[i] = occ_g[i];
    stage_twiddles<M, NTHR>(stwM, twM);
    stage_twiddles<N, NTHR>(stwN, twN);
    __syncthreads();

    const int iz = (blockIdx.x + zlo + N) % N;
    const int b = blockIdx.y;
    const int sz = signed_coord<N>(iz, r);
    // gathered voxels: inside the rotation sphere and the support ball (|v| >
    // zr reads only zeros whatever the rotation)
    const int rg = r < zr ? r : zr, rg2 = rg * rg;
    double R[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = rots[9 * b + i];
    float2* dst = out + ((int64_t)b * 3) * G::H + (int64_t)iz * G::PLANE;

    {
        // rows: gather the row of P and W, half-length x FFT of each into its plane
        const int row = threadIdx.x / TX, tx = threadIdx.x % TX;
        const int sy = signed_coord<N>(row, r);
        const bool in_plane = row < N;
        // rows whose every voxel lies outside the support are exactly zero
        const bool active = in_plane && sz * sz + sy * sy <= zr * zr;
        if (in_plane && !active) {
#pragma unroll
            for (int e = 0; e < EX; ++e) {
                planes[row * S + tx + TX * e] = c2(0.f, 0.f);
                planes[N * S + row * S + tx + TX * e] = c2(0.f, 0.f);
            }
        }
        float2 vp[EX], vw[EX];
        if (active) {
#pragma unroll
            for (int e = 0; e < EX; ++e) {
                float p2[2], w2[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int x = signed_coord<N>(2 * (tx + TX * e) + h, r);
                    p2[h] = 0.f;
                    w2[h] = 0.f;
                    if (sz * sz + sy * sy + x * x <= rg2)
                        gather_pw<N>(tm, occ, R, sz, sy, x, &p2[h], &w2[h]);
                }
                vp[e] = c2(p2[0], p2[1]);
                vw[e] = c2(w2[0], w2[1]);
            }
        }
        float2* rowP = active ? planes + row * S : planes;
        float2* rowW = active ? planes + N * S + row * S : planes;
        group_fft_ex<M, TX, -1, WROWS, XPAD>(vp, rowP, 1, tx, stwM, active);
        group_fft_ex<M, TX, -1, WROWS, XPAD>(vw, rowW, 1, tx, stwM, active);
    }
    __syncthreads();
    {
        // columns of both planes: R2C post-processing on the way in
        // (X[k] = E[k] + w^k O[k], X0 and XM packed in column 0), y FFT left
        // in the plane
        const int cg = threadIdx.x % CPI, ty = threadIdx.x / CPI;
        const int cs = cg / M, cc = cg - cs * M;
        const bool cact = cs < 2;
        float2* pl = planes + (cact ? cs : 0) * (N * S);
        float2 v[EY];
        if (cact) {
#pragma unroll
            for (int e = 0; e < EY; ++e) {
                const float2* rowp = pl + (ty + TY * e) * S;
                const float2 zk = rowp[cc];
                if (cc == 0) {
                    v[e] = c2(zk.x + zk.y, zk.x - zk.y);
                } else {
                    const float2 zm = cconj(rowp[M - cc]);
                    const float2 Ev = cscale(cadd(zk, zm), 0.5f);
                    const float2 Od = mul_si<-1>(cscale(csub(zk, zm), 0.5f));
                    v[e] = cadd(Ev, cmul(stwN[cc], Od));
                }
            }
        }
        __syncthreads();   // mirror reads are done before the FFT overwrites columns
        group_fft_ex<N, TY, -1, false>(v, pl + (cact ? cc : 0), S, ty, stwN, cact);
    }
    __syncthreads();
    store_plane<N, NTHR>(planes, dst);
    store_plane<N, NTHR>(planes + N * S, dst + G::H);
}

// Copy a plane ([ky][kx] in global memory, NXH complex per row) into an LDS
// slot with row stride S by direct global->LDS loads: each wave moves 64
// consecutive floats of the slot per instruction (lane l -> float l of the
// chunk), nothing passes through VGPRs.  Padding floats of a row receive a
// copy of the row's first float.  The loads complete under vmcnt.
//
// The lane -> (row, float) pattern repeats every GR rows (GR * RF floats are a
// whole number CG of 64-float chunks), so each lane computes its CG byte
// offsets within such a group once per kernel (PlaneCopy holds them); the
// copies go through a buffer resource over the source plane, with the group's
// offset as the scalar offset, so issuing them takes no vector arithmetic (and
// a copy outside the plane's bytes would read zeros instead of faulting).
constexpr int dma_gcd(int a, int b) { return b == 0 ? a : dma_gcd(b, a % b); }

// HOIST: the offsets are computed once, when the PlaneCopy is made, and held
// in CG registers; otherwise per call, behind an empty asm so the compiler
// does not hoist them out of the caller's loop (for kernels short of
// registers).
template <int N, int NTHR, bool HOIST>
struct PlaneCopy {
    using G = Geo<N>;
    static constexpr int RF = 2 * G::S;     // floats per LDS row
    static constexpr int GF = 2 * G::NXH;   // floats per global row
    static constexpr int GR = 64 / dma_gcd(RF, 64);   // rows per group
    static constexpr int CG = GR * RF / 64;            // chunks per group
    static constexpr int NG = N / GR;                  // groups per plane
    static constexpr int NW = NTHR / 64;               // waves
    static_assert(N % GR == 0 && NTHR % 64 == 0, "whole groups of rows, whole waves");
    int off[HOIST ? CG : 1];

    // this lane's byte offset of chunk j within a group of rows
    static __device__ __forceinline__ void offsets(int lane, int (&o)[CG]) {
#pragma unroll
        for (int j = 0; j < CG; ++j) {
            const int f = j * 64 + lane;
            const int row = f / RF, fc = f - row * RF;
            o[j] = 4 * (row * GF + (fc < GF ? fc : 0));
        }
    }

    __device__ __forceinline__ PlaneCopy() {
        if constexpr (HOIST) offsets(threadIdx.x & 63, off);
    }

    __device__ __forceinline__ void operator()(float2* slot, const float2* __restrict__ src) const {
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float2*>(src), 0, (int)(sizeof(float) * N * GF), 0x00020000);
        float* d = reinterpret_cast<float*>(slot);
        int o[CG];
        if constexpr (HOIST) {
#pragma unroll
            for (int j = 0; j < CG; ++j) o[j] = off[j];
        } else {
            int lane = threadIdx.x & 63;
            asm volatile("" : "+v"(lane));
            offsets(lane, o);
        }
        // the wave index, as a wave-uniform (scalar) value
        const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        for (int g = wave; g < NG; g += NW) {
#pragma unroll
            for (int j = 0; j < CG; ++j)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rsrc, (__attribute__((address_space(3))) void*)(d + g * (GR * RF) + j * 64),
                    4, o[j], 4 * g * (GR * GF), 0, 0);
        }
    }
};

// Wait for this wave's outstanding global loads (LDS copies included) and
// stores, then for the whole workgroup.
__device__ __forceinline__ void wait_copies_and_sync() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// Inverse planes with the LCC, two LDS slots: per rotation one column pass
// over the pair (B, A) (16-point column FFTs), their row passes (giving
// 1/sqrt(N B - A^2) per voxel), then G with the one-plane column layout,
// transformed from slot 0 into slot 1.  Planes reach LDS by direct copies
// issued while the other slot is being transformed: G during the A rows, the
// next B during G's transform; only the next A's copy is waited for.
template <int N, int NTHR, int EMAX, bool HOISTC = false>
__global__ __launch_bounds__(NTHR) void k_inv_plane2(const float2* __restrict__ X,
                                                     const uint8_t* __restrict__ lmask,
                                                     unsigned long long* __restrict__ keys,
                                                     int nrot, int chunk, int64_t first,
                                                     double norm,
                                                     const float2* __restrict__ twM,
                                                     const float2* __restrict__ twN) {
    using G = Geo<N>;
    using P = PairPlan<N, NTHR, EMAX>;
    constexpr int M = G::M, S = G::S, TX = P::TX, EX = P::EX;
    constexpr int TYP = P::TY, EYP = P::EY, CPIP = P::CPI;         // pair layout
    constexpr int TYS = pick_T_cols(N, M, NTHR, EMAX);             // one-plane layout
    constexpr int EYS = N / TYS, CPIS = NTHR / TYS;
    constexpr bool WROWS = kWaveRows && (64 % TX == 0);
    constexpr int XPAD = row_pad(M, TX);
    static_assert(pad_extent(M, XPAD) <= G::S, "padded row fits the row stride");
    static_assert(P::ok, "pair plan");
    static_assert(TYS > 0 && NTHR % TYS == 0 && CPIS >= M, "one column pass");
    __shared__ float2 planes[2 * N * S];
    __shared__ float2 stwM[M], stwN[N];
    stage_twiddles<M, NTHR>(stwM, twM);
    stage_twiddles<N, NTHR>(stwN, twN);
    float2* const slot0 = planes;
    float2* const slot1 = planes + N * S;

    const int iz = blockIdx.x;
    const int b0 = blockIdx.y * chunk;
    const int b1 = min(nrot, b0 + chunk);
    const int row = threadIdx.x / TX, tx = threadIdx.x % TX;   // row pass: one row
    const bool ract = row < N;
    const int cgp = threadIdx.x % CPIP, typ = threadIdx.x / CPIP;
    const int csp = cgp / M, ccp = cgp - csp * M;               // pair: slot, column
    const bool cactp = csp < 2;
    const int ccs = threadIdx.x % CPIS, tys = threadIdx.x / CPIS;
    const bool cacts = ccs < M;
    const double nv = norm * (double)G::V;   // N V (exact)

    // owned voxels: (row, x = 2 (tx + TX e) + h), bit 2 e + h
    uint32_t mbits = 0;
    if (ract) {
#pragma unroll
        for (int e = 0; e < EX; ++e)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t vox = ((int64_t)iz * N + row) * N + 2 * (tx + TX * e) + h;
                if (lmask[vox]) mbits |= 1u << (2 * e + h);
            }
    }
    float isd[EX][2];

    auto plane_of = [&](int b, int g) {
        return X + ((int64_t)b * 3 + g) * G::H + (int64_t)iz * G::PLANE;
    };
    // C2R pre-processing of row `rowp` (packed (X0, XM) in column 0), then the
    // row's half-length inverse FFT; the values end up in vr
    auto row_inverse = [&](float2* rowp, float2 (&vr)[EX]) {
        if (ract) {
            const float2 c0 = rowp[0];
#pragma unroll
            for (int e = 0; e < EX; ++e) {
                const int k = tx + TX * e;
                const float2 xk = (k == 0) ? c2(c0.x, 0.f) : rowp[k];
                const float2 xm = (k == 0) ? c2(c0.y, 0.f) : rowp[M - k];
                const float2 xmc = cconj(xm);
                const float2 d = cmul(csub(xk, xmc), cconj(stwN[k]));
                vr[e] = cadd(cadd(xk, xmc), mul_si<1>(d));
            }
        }
        // the row's own threads read it above and overwrite it below
        group_sync<WROWS>();
        group_fft<M, TX, 1, WROWS, XPAD>(vr, rowp, 1, tx, stwM, ract);
    };

    const PlaneCopy<N, NTHR, HOISTC> copy;
    if (b0 < b1) {
        copy(slot0, plane_of(b0, 2));   // B
        copy(slot1, plane_of(b0, 1));   // A
    }
    for (int b = b0; b < b1; ++b) {
        wait_copies_and_sync();   // B and A of rotation b are in the slots
        {
            // y inverse of the packed columns of both slots: X0 + i XM in column 0
            float2 v[EYP];
            if (cactp) {
                const float2* col = planes + csp * (N * S) + ccp;
#pragma unroll
                for (int e = 0; e < EYP; ++e) {
                    v[e] = col[(typ + TYP * e) * S];
                    if (ccp == 0) v[e] = cadd(v[e], mul_si<1>(col[(typ + TYP * e) * S + M]));
                }
            }
            __syncthreads();   // every column is read before the FFT overwrites it
            group_fft_ex<N, TYP, 1, false>(v, planes + (cactp ? csp * (N * S) + ccp : 0), S, typ,
                                           stwN, cactp);
        }
        __syncthreads();
        {
            float2 vr[EX];
            row_inverse(slot0 + (ract ? row : 0) * S, vr);
            if (ract) {
#pragma unroll
                for (int e = 0; e < EX; ++e) {
                    isd[e][0] = vr[e].x;   // B
                    isd[e][1] = vr[e].y;
                }
            }
        }
        __syncthreads();   // slot 0 is free
        copy(slot0, plane_of(b, 0));   // G
        {
            float2 vr[EX];
            row_inverse(slot1 + (ract ? row : 0) * S, vr);
            if (ract) {
#pragma unroll
                for (int e = 0; e < EX; ++e)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        // N V B - A^2 in FP64 (exact products), then 1/sqrt in FP32
                        isd[e][h] = rsq_var(h ? vr[e].y : vr[e].x, isd[e][h], nv);
                    }
            }
        }
        wait_copies_and_sync();   // G is in slot 0, slot 1 is free
        {
            float2 v[EYS];
            if (cacts) {
                const float2* col = slot0 + ccs;
#pragma unroll
                for (int e = 0; e < EYS; ++e) {
                    v[e] = col[(tys + TYS * e) * S];
                    if (ccs == 0) v[e] = cadd(v[e], mul_si<1>(col[(tys + TYS * e) * S + M]));
                }
            }
            __syncthreads();   // slot 0 is read: the next B can come in
            if (b + 1 < b1) copy(slot0, plane_of(b + 1, 2));
            group_fft_ex<N, TYS, 1, false>(v, slot1 + (cacts ? ccs : 0), S, tys, stwN, cacts);
        }
        __syncthreads();
        {
            float2 vr[EX];
            row_inverse(slot1 + (ract ? row : 0) * S, vr);
            if (ract) {
#pragma unroll
                for (int e = 0; e < EX; ++e)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (!((mbits >> (2 * e + h)) & 1u)) continue;
                        // only S > 0 can beat the initial (0, 0); NaN never wins
                        const float sv = (h ? vr[e].y : vr[e].x) * isd[e][h];
                        if (sv > 0.f) {
                            const int64_t vox =
                                ((int64_t)iz * N + row) * N + 2 * (tx + TX * e) + h;
                            atomicMax(keys + vox, lcc_key(sv, (uint32_t)(first + b)));
                        }
                    }
            }
        }
        __syncthreads();   // slot 1 is free
        if (b + 1 < b1) copy(slot1, plane_of(b + 1, 1));
    }
}

// Brick occupancy for the gathers: bit b set when any voxel of 8^3 brick b
// (bricks ceil(n / 8) per axis, the last one partial), or of the one-voxel
// layer after it in +x/+y/+z (periodic), has a nonzero template or mask value.
// A corner cube whose lower corner lies in a clear brick is all zero.
__global__ void k_brick_occupancy(const float2* __restrict__ tm, int nz, int ny, int nx,
                                  uint32_t* __restrict__ occ) {
    const int nbz = (nz + 7) / 8, nby = (ny + 7) / 8, nbx = (nx + 7) / 8;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbz * nby * nbx) return;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    bool any = false;
    for (int dz = 0; dz <= 8 && !any; ++dz)
        for (int dy = 0; dy <= 8 && !any; ++dy)
            for (int dx = 0; dx <= 8; ++dx) {
                const int z = (bz * 8 + dz) % nz, y = (by * 8 + dy) % ny, x = (bx * 8 + dx) % nx;
                const float2 v = tm[((int64_t)z * ny + y) * nx + x];
                if (v.x != 0.f || v.y != 0.f) {
                    any = true;
                    break;
                }
            }
    if (any) atomicOr(occ + (b >> 5), 1u << (b & 31));
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

// K1 with one-plane passes: all grids in one launch, or (STASHW) one launch
// per grid, grid 0 first.
template <int N, int NTHR, bool BIN, int EMAX, bool STASHW, int EMAXC>
void launch_fwd_plane(dim3 grid, hipStream_t s, const float2* tm, const double* rots, int r,
                      int zr, int zlo, float2* out, const float2* twM, const float2* twN,
                      const uint32_t* occ) {
    if constexpr (STASHW) {
        hipLaunchKernelGGL((k_fwd_plane<N, NTHR, BIN, EMAX, true, EMAXC, 0>), grid, dim3(NTHR),
                           0, s, tm, rots, r, zr, zlo, out, twM, twN, occ);
        hipLaunchKernelGGL((k_fwd_plane<N, NTHR, BIN, EMAX, true, EMAXC, 1>), grid, dim3(NTHR),
                           0, s, tm, rots, r, zr, zlo, out, twM, twN, occ);
        if constexpr (!BIN)
            hipLaunchKernelGGL((k_fwd_plane<N, NTHR, BIN, EMAX, true, EMAXC, 2>), grid,
                               dim3(NTHR), 0, s, tm, rots, r, zr, zlo, out, twM, twN, occ);
    } else {
        hipLaunchKernelGGL((k_fwd_plane<N, NTHR, BIN, EMAX, false, EMAXC>), grid, dim3(NTHR), 0,
                           s, tm, rots, r, zr, zlo, out, twM, twN, occ);
    }
}

template <int N, int NTHR, int EMAX>
void launch_fwd_pair(dim3 grid, hipStream_t s, const float2* tm, const double* rots, int r, int zr,
                     int zlo, float2* out, const float2* twM, const float2* twN,
                     const uint32_t* occ) {
    if constexpr (PairPlan<N, NTHR, EMAX>::ok)
        hipLaunchKernelGGL((k_fwd_plane2<N, NTHR, EMAX>), grid, dim3(NTHR), 0, s, tm, rots, r, zr,
                           zlo, out, twM, twN, occ);
}

// z stage with element bound EZ (its radix schedule: pick_T(N, EZ, 32)).
template <int N, int TQ, int EZ, bool SPLIT = false>
void launch_zcorr(const FusedArgs& a, int ch, int zr_read, const float2* twN, hipStream_t s) {
    using G = Geo<N>;
    const dim3 grid((G::PLANE + TQ - 1) / TQ, (a.nrot + ch - 1) / ch);
    const int nthr = TQ * pick_T(N, EZ, 32);
    if (a.binary_mask)
        hipLaunchKernelGGL((k_zcorr<N, TQ, EZ, true, SPLIT>), grid, dim3(nthr), 0, s, a.work, a.F1,
                           a.F2, a.nrot, ch, zr_read, twN);
    else
        hipLaunchKernelGGL((k_zcorr<N, TQ, EZ, false, SPLIT>), grid, dim3(nthr), 0, s, a.work,
                           a.F1, a.F2, a.nrot, ch, zr_read, twN);
}

// The split z stage exists where the default schedule's z plan starts with
// radix 16 over one butterfly per thread in 32-column tiles.
template <int N, int TQ, int EZ>
constexpr bool zsplit_ok() {
    return TQ == 32 && pick_T(N, EZ, 32) % 2 == 0 && N / pick_T(N, EZ, 32) == 16 &&
           gplan_for(N, pick_T(N, EZ, 32)).r[0] == 16;
}

// Launch shape of one fused size.  K1: NTHR1 threads, element bounds EMAX (rows,
// and the two-plane kernels) and EMAXC1 (one-plane column pass), mask values
// kept in global memory and one launch per grid when STASHW (weighted masks)
// or STASHB (0/1 masks without the two-plane kernel).  K2: TQ columns per
// block (TQ2 the alternative width), element bounds EMAXZ / EMAXZ2 of the
// default and the alternative radix plan (FusedArgs::zsched: bit 0 the radix
// plan, bit 1 the tile width).  K3: NTHR3 threads, element bounds
// EMAX3 (rows) / EMAXC3 (columns), options XPF3, ISDG3 and RECOMP3 of the
// one-plane kernel.  Sizes that differ from the defaults derive from Shape and
// redefine members.
template <int NTHR1_, int TQ_, int EZ_, int NTHR3_, int E_ = 16, int EZ2_ = EZ_>
struct Shape {
    static constexpr int NTHR1 = NTHR1_, TQ = TQ_, EMAXZ = EZ_, NTHR3 = NTHR3_, EMAX = E_,
                         EMAXZ2 = EZ2_, EMAX3 = E_, EMAXC3 = E_, EMAXC1 = E_;
    static constexpr int TQ2 = TQ_ == 32 ? 16 : 32;   // alternative z tile width
    static constexpr bool XPF3 = true, ISDG3 = false, RECOMP3 = false, STASHW = false;
    static constexpr bool STASHB = STASHW;
    static constexpr bool HOIST3 = false;   // two-plane K3: copy offsets held in registers
    static constexpr bool PAIR1 = true, PAIR3 = true;   // two-plane K1 / K3 where they fit
};

// 72: two-plane K3 in 576-thread blocks (4.3 against 5.6 us per rotation);
// K1 in 768-thread blocks, two-plane for 0/1 masks (2.6 against 3.1 us)
struct Shape72 : Shape<768, 32, 12, 576, 16, 6> {
    static constexpr bool STASHW = true, STASHB = true;
};
struct Shape80 : Shape<320, 16, 16, 320, 12, 24> {
    static constexpr bool RECOMP3 = true, STASHW = true, STASHB = true;
};
// 48: one K1 launch per grid for weighted masks (the mask values parked; 0.77
// -> 0.64 us per rotation with a core-weighted mask); two-plane for 0/1 masks
struct Shape48 : Shape<192, 32, 12, 96, 16, 6> {
    static constexpr bool STASHW = true;
};
// 64: 0/1 masks one-plane, one launch per grid (0.60 -> 0.43 us per rotation
// against the two-plane K1); weighted masks one launch for all grids (faster
// there)
struct Shape64 : Shape<256, 32, 8, 256, 16, 4> {
    static constexpr bool PAIR1 = false, STASHB = true;
};
// 96: one-plane K1, one launch per grid for both mask kinds (per rotation 6.4
// -> 4.5 us core-weighted, 4.1 -> 3.3 us for 0/1 masks against the two-plane
// K1)
struct Shape96 : Shape<768, 32, 12, 384, 16, 24> {
    static constexpr bool STASHW = true, STASHB = true, PAIR1 = false;
};
// 128: one-plane K1 in 512-thread blocks, one launch per grid (7.3 against
// 8.4 us per rotation for the two-plane K1)
struct Shape128 : Shape<512, 32, 16, 1024, 16, 8> {
    static constexpr bool PAIR1 = false, STASHW = true, STASHB = true, HOIST3 = true;
};
// 144: the next plane is not prefetched during the row phase (K3 23.4 ->
// 22.1 us per rotation: the prefetch registers spill at 864 threads)
struct Shape144 : Shape<864, 16, 12, 864, 16, 6> {
    static constexpr bool RECOMP3 = true, XPF3 = false;
};
// 160 = 2^5 * 5 has no 16-element plan; K3 sweeps its 80 columns once with
// 20 elements per thread, holds the variance in registers and recomputes its
// addresses per plane (43.6 -> 33.0 us per rotation against the variance in
// global memory)
struct Shape160 : Shape<704, 16, 10, 640, 20, 20> {
    static constexpr int EMAX3 = 10, EMAXC3 = 20;
    static constexpr bool XPF3 = false, RECOMP3 = true, STASHW = true, STASHB = true;
};
// 192: K3 runs its 96 columns in two sweeps of 48 (12 elements per thread),
// variance in registers, addresses recomputed per plane (67.6 -> 49.4 us)
struct Shape192 : Shape<784, 16, 12, 768, 24, 24> {
    static constexpr int EMAX3 = 12, EMAXC3 = 12;
    static constexpr bool XPF3 = false, RECOMP3 = true, STASHW = true, STASHB = true;
};

template <int N, class SH>
int run_fused(const FusedArgs& a, int which, hipStream_t s) {
    constexpr int NTHR1 = SH::NTHR1, TQ = SH::TQ, NTHR3 = SH::NTHR3, EMAX = SH::EMAX;
    using G = Geo<N>;
    Tw tw;
    PF_CHECK_STATUS(get_tw(N, &tw));
    // planes |z| <= zr only when the z stage treats the others as zero
    const int zr = std::max(0, std::min(a.zr, N / 2));
    const bool skip = a.skip_planes && 2 * zr + 1 < N;
    const int nplanes = skip ? 2 * zr + 1 : N;
    const int zlo = skip ? -zr : 0;
    const int zr_read = skip ? zr : N;
    if (which < 0 || which == 0) {
        if (a.binary_mask && SH::PAIR1 && PairPlan<N, NTHR1, EMAX>::ok)
            launch_fwd_pair<N, NTHR1, EMAX>(dim3(nplanes, a.nrot), s, a.tm, a.rots, a.rmax, a.zr,
                                            zlo, a.work, tw.M, tw.N, a.occ);
        else if (a.binary_mask)
            launch_fwd_plane<N, NTHR1, true, EMAX, SH::STASHB, SH::EMAXC1>(
                dim3(nplanes, a.nrot), s, a.tm, a.rots, a.rmax, a.zr, zlo, a.work, tw.M, tw.N,
                a.occ);
        else
            launch_fwd_plane<N, NTHR1, false, EMAX, SH::STASHW, SH::EMAXC1>(
                dim3(nplanes, a.nrot), s, a.tm, a.rots, a.rmax, a.zr, zlo, a.work, tw.M, tw.N,
                a.occ);
        PF_CHECK_STATUS(last("k_fwd_plane"));
    }
    if (which < 0 || which == 1) {
        const int ch = a.chunk_z;
        const bool alt_plan = a.zsched & 1, alt_tile = a.zsched & 2;
        if (!a.split_z && (a.zsched & 4)) {
            // schedules 4 and 5: the register z stage of the run-time planned
            // pipeline (same work-buffer layout), bit 0 its register plan
            Fft3D fz{};
            fz.nz = N;
            fz.ny = N;
            fz.nx = N;
            fz.nxh = N / 2 + 1;
            fz.twz = tw.N;
            return launch_zcorr_ct(fz, a, zr_read, s);
        }
        if (a.split_z) {
            // split-precision first stage: the default schedule only
            if constexpr (zsplit_ok<N, TQ, SH::EMAXZ>()) {
                launch_zcorr<N, TQ, SH::EMAXZ, true>(a, ch, zr_read, tw.N, s);
            } else {
                set_error("fused z stage: no split-precision form for n = %d", N);
                return kErrUnsupported;
            }
        } else if (alt_tile && alt_plan)
            launch_zcorr<N, SH::TQ2, SH::EMAXZ2>(a, ch, zr_read, tw.N, s);
        else if (alt_tile)
            launch_zcorr<N, SH::TQ2, SH::EMAXZ>(a, ch, zr_read, tw.N, s);
        else if (alt_plan)
            launch_zcorr<N, TQ, SH::EMAXZ2>(a, ch, zr_read, tw.N, s);
        else
            launch_zcorr<N, TQ, SH::EMAXZ>(a, ch, zr_read, tw.N, s);
        PF_CHECK_STATUS(last("k_zcorr"));
    }
    if (which < 0 || which == 2) {
        const int ch = a.chunk_lcc;
        // two-slot inverse only for the large planes: it costs registers
        if constexpr (SH::PAIR3 && PairPlan<N, NTHR3, EMAX>::ok && NTHR3 >= 512)
            hipLaunchKernelGGL((k_inv_plane2<N, NTHR3, EMAX, SH::HOIST3>), dim3(N, (a.nrot + ch - 1) / ch),
                               dim3(NTHR3), 0, s, a.work, a.lcc_mask, a.keys, a.nrot, ch, a.first,
                               a.norm, tw.M, tw.N);
        else
            hipLaunchKernelGGL((k_inv_plane<N, NTHR3, SH::EMAX3, SH::EMAXC3, SH::XPF3, SH::ISDG3,
                                            SH::RECOMP3>),
                               dim3(N, (a.nrot + ch - 1) / ch),
                               dim3(NTHR3), 0, s, a.work, a.lcc_mask, a.keys, a.nrot, ch, a.first,
                               a.norm, tw.M, tw.N);
        PF_CHECK_STATUS(last("k_inv_plane"));
    }
    return kOk;
}

}  // namespace

bool fused_supported(int nz, int ny, int nx) {
    if (nz != ny || ny != nx) return false;
    switch (nx) {
        case 48: case 64: case 72: case 80: case 96: case 128: case 144: case 160: case 192:
            return true;
        default: return false;
    }
}

bool fused_split_supported(int n) {
    switch (n) {
        case 48: return zsplit_ok<48, 32, 12>();
        case 64: return zsplit_ok<64, 32, 8>();
        case 72: return zsplit_ok<72, Shape72::TQ, Shape72::EMAXZ>();
        case 80: return zsplit_ok<80, Shape80::TQ, Shape80::EMAXZ>();
        case 96: return zsplit_ok<96, 32, 12>();
        case 128: return zsplit_ok<128, Shape128::TQ, Shape128::EMAXZ>();
        case 144: return zsplit_ok<144, Shape144::TQ, Shape144::EMAXZ>();
        case 160: return zsplit_ok<160, Shape160::TQ, Shape160::EMAXZ>();
        case 192: return zsplit_ok<192, Shape192::TQ, Shape192::EMAXZ>();
        default: return false;
    }
}

int launch_fused(const FusedArgs& a, int which, hipStream_t s) {
    if (a.nrot <= 0) return kOk;
    if (a.nrot > 65535 || a.chunk_z < 1 || a.chunk_lcc < 1 || a.chunk_lcc > 255) {
        set_error("fused: invalid batch or chunk sizes");
        return kErrArg;
    }
    if (a.rmax < 0 || a.rmax > a.n / 2) {
        set_error("fused: rmax must be in [0, n/2]");
        return kErrArg;
    }
    switch (a.n) {
        case 48: return run_fused<48, Shape48>(a, which, s);
        case 64: return run_fused<64, Shape64>(a, which, s);
        case 72: return run_fused<72, Shape72>(a, which, s);
        case 80: return run_fused<80, Shape80>(a, which, s);
        case 96: return run_fused<96, Shape96>(a, which, s);
        case 128: return run_fused<128, Shape128>(a, which, s);
        case 144: return run_fused<144, Shape144>(a, which, s);
        case 160: return run_fused<160, Shape160>(a, which, s);
        case 192: return run_fused<192, Shape192>(a, which, s);
        default:
            set_error("fused pipeline not built for n = %d", a.n);
            return kErrUnsupported;
    }
}

int occupancy_words(int nz, int ny, int nx) {
    const int64_t nb = (int64_t)((nz + 7) / 8) * ((ny + 7) / 8) * ((nx + 7) / 8);
    return (int)((nb + 31) / 32);
}

int launch_brick_occupancy(const float2* tm, int nz, int ny, int nx, uint32_t* occ,
                           hipStream_t s) {
    const int nbricks = ((nz + 7) / 8) * ((ny + 7) / 8) * ((nx + 7) / 8);
    PF_HIP_CHECK(hipMemsetAsync(occ, 0, sizeof(uint32_t) * occupancy_words(nz, ny, nx), s));
    hipLaunchKernelGGL(k_brick_occupancy, dim3((nbricks + 255) / 256), dim3(256), 0, s, tm, nz,
                       ny, nx, occ);
    return last("k_brick_occupancy");
}

int launch_pack_tm(const float* tmpl, const float* mask, float2* tm, int64_t n, hipStream_t s) {
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_pack_tm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tmpl, mask,
                       tm, n);
    return last("k_pack_tm");
}

}  // namespace pf
