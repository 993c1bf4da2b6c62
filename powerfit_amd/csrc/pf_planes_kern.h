// Kernels and helpers of the run-time planned fused pipeline shared by its two
// translation units: pf_planes.hip (even x lengths, the z stage, the
// launcher) and pf_planes_odd.hip (the plane kernels for odd x lengths).
// Everything here has internal linkage in each unit that includes it.
#pragma once
#include <algorithm>

#include "pf_ct_fft.h"
#include "pf_internal.h"
#include "pf_rotate.h"

namespace pf {

namespace {


constexpr int kRtThreads = 1024;         // plane kernels: largest block
constexpr size_t kCuLds = 160 * 1024;    // LDS of one compute unit
constexpr int kRtVox = 16;               // voxels of a plane owned per thread (K3), at most
constexpr int kRtZThreads = 256;         // z kernel
constexpr int kRtTq = 16;                // z columns per block
constexpr size_t kRtLds = 150 * 1024;    // dynamic LDS a plane kernel may take

// Signed coordinate of index i on an axis of length n for rotation radius
// r <= n / 2: i for i <= r, else i - n (as the compiled pipeline's).
__device__ __forceinline__ int scoord(int i, int r, int n) { return i <= r ? i : i - n; }

// Rotated template (trilinear, FP32 blend) and mask (nearest) at signed voxel
// (z, y, x) from the interleaved (template, mask) grid, any shape.  FP64
// coordinates, floor and round in the reference's order (so corner and
// nearest indices are the reference's), one "+n" wrap for negatives, the
// upper neighbour wrapping only at n.  Sources whose corner cube starts in a
// clear 8^3 brick of occ (launch_brick_occupancy) are exactly zero and skip
// the loads.
__device__ __forceinline__ void gather_pw_rt(const float2* __restrict__ tm,
                                             const uint32_t* __restrict__ occ, int nz, int ny,
                                             int nx, const double* R, int z, int y, int x,
                                             float* p, float* w) {
    const Coord c = source_coord(R, z, y, x);
    const double fx = floor(c.xs), fy = floor(c.ys), fz = floor(c.zs);
    const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
    const int xi = x0 + (x0 < 0 ? nx : 0), yi = y0 + (y0 < 0 ? ny : 0), zi = z0 + (z0 < 0 ? nz : 0);
    *p = 0.f;
    *w = 0.f;
    if ((unsigned)xi >= (unsigned)nx || (unsigned)yi >= (unsigned)ny || (unsigned)zi >= (unsigned)nz)
        return;
    const int brick = ((zi >> 3) * ((ny + 7) >> 3) + (yi >> 3)) * ((nx + 7) >> 3) + (xi >> 3);
    if (!((occ[brick >> 5] >> (brick & 31)) & 1u)) return;
    const int xj = (xi + 1 == nx) ? 0 : xi + 1;
    const int yj = (yi + 1 == ny) ? 0 : yi + 1;
    const int zj = (zi + 1 == nz) ? 0 : zi + 1;
    const float2* r00 = tm + ((int64_t)zi * ny + yi) * nx;
    const float2* r10 = tm + ((int64_t)zi * ny + yj) * nx;
    const float2* r01 = tm + ((int64_t)zj * ny + yi) * nx;
    const float2* r11 = tm + ((int64_t)zj * ny + yj) * nx;
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
    // round() is half away from zero: up from floor when the fraction
    // exceeds one half, or equals it for non-negative values
    const double hx = c.xs - fx, hy = c.ys - fy, hz = c.zs - fz;
    const bool ux = hx > 0.5 || (hx == 0.5 && c.xs >= 0.0);
    const bool uy = hy > 0.5 || (hy == 0.5 && c.ys >= 0.0);
    const bool uz = hz > 0.5 || (hz == 0.5 && c.zs >= 0.0);
    const float2 n0 = uy ? (ux ? a011 : a010) : (ux ? a001 : a000);
    const float2 n1 = uy ? (ux ? a111 : a110) : (ux ? a101 : a100);
    *w = uz ? n1.y : n0.y;
}

// One Stockham butterfly of radix R (the stage of pf_fft.h) with its twiddle
// phase k = j mod Ns and twiddle step n / (Ns R) passed in, so a butterfly
// needs no integer division.
template <int R, int DIR>
__device__ __forceinline__ void bfly_k(const float2* sp, float2* dp, int j, int k, int Ns, int nb,
                                       int step, int es, const float2* tw) {
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = sp[(j + r * nb) * es];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<DIR>(tw, r * k * step));
    dft_fixed<R, DIR>(v);
    const int od = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dp[(od + r * Ns) * es] = v[r];
}

template <int DIR>
__device__ __noinline__ void bfly_k_any(const float2* sp, float2* dp, int R, int j, int k, int Ns,
                                        int nb, int step, int es, const float2* tw) {
    const int od = (j - k) * R + k;
    for (int q = 0; q < R; ++q) {
        float2 acc = c2(0.f, 0.f);
        for (int m = 0; m < R; ++m) {
            float2 x = sp[(j + m * nb) * es];
            if (m > 0) x = cmul(x, twid<DIR>(tw, m * k * step));
            const int e = (m * q) % R;
            acc = cadd(acc, e ? cmul(x, twid<DIR>(tw, e * nb)) : x);
        }
        dp[(od + q * Ns) * es] = acc;
    }
}

// One stage of lds_fft (pf_fft.h) for radix R (0: any radix, generic DFT),
// with each thread stepping through its butterflies by the block size:
// (transform f, butterfly j) and k = j mod Ns advance by adds and compares
// (nb is a multiple of Ns, so wrapping j past nb keeps k).
template <int R, int DIR>
__device__ __forceinline__ void rt_stage(const float2* a, float2* b, int Rr, int Ns, int n,
                                         const float2* tw, int nfft, int fs, int es) {
    const int T = blockDim.x, t = threadIdx.x;
    const int RR = R > 0 ? R : Rr;
    const int nb = n / RR, step = n / (Ns * RR), total = nfft * nb;
    int f, j, df, dj;
    if (es == 1) {   // u = f nb + j
        f = t / nb;
        j = t - f * nb;
        df = T / nb;
        dj = T - df * nb;
    } else {         // u = j nfft + f
        j = t / nfft;
        f = t - j * nfft;
        dj = T / nfft;
        df = T - dj * nfft;
    }
    int k = j % Ns;
    const int dk = dj % Ns;
    for (int u = t; u < total; u += T) {
        const float2* sp = a + f * fs;
        float2* dp = b + f * fs;
        if constexpr (R > 0) bfly_k<R, DIR>(sp, dp, j, k, Ns, nb, step, es, tw);
        else bfly_k_any<DIR>(sp, dp, RR, j, k, Ns, nb, step, es, tw);
        k += dk;
        if (k >= Ns) k -= Ns;
        if (es == 1) {
            j += dj;
            f += df;
            if (j >= nb) {
                j -= nb;
                ++f;
            }
        } else {
            f += df;
            j += dj;
            if (f >= nfft) {
                f -= nfft;
                ++j;
                if (++k >= Ns) k -= Ns;
            }
        }
    }
}

// A radix plan as a kernel argument: the length, the stage count and the
// radices in 8 bits each (stage s in bits 8 s .. 8 s + 7).  Four scalar
// registers, where Plan1D's radix array would hold eighteen for the whole
// kernel.
struct PlanRt {
    int n, nstages;
    unsigned long long radix;
};

// Pack a plan; false when it has more than 8 stages or a radix above 255.
inline bool pack_plan(const Plan1D& p, PlanRt* out) {
    if (p.nstages > 8) return false;
    out->n = p.n;
    out->nstages = p.nstages;
    out->radix = 0;
    for (int s = 0; s < p.nstages; ++s) {
        if (p.radix[s] < 2 || p.radix[s] > 255) return false;
        out->radix |= (unsigned long long)p.radix[s] << (8 * s);
    }
    return true;
}

// The packed plan of length n (make_plan's factorisation); false if none.
inline bool plan_rt(int n, PlanRt* out) {
    Plan1D p;
    return make_plan(n, &p) && pack_plan(p, out);
}

// lds_fft (pf_fft.h) over nfft transforms in LDS with the radix of each stage
// dispatched once per stage; twiddles are read from LDS.  Returns the buffer
// holding the result; ends with a barrier.
template <int DIR>
__device__ float2* lds_fft_rt(float2* a, float2* b, const PlanRt& p, const float2* tw, int nfft,
                              int fs, int es) {
    const int n = p.n;
    int Ns = 1;
    for (int s = 0; s < p.nstages; ++s) {
        const int R = (int)((p.radix >> (8 * s)) & 255ull);
        switch (R) {
            case 2: rt_stage<2, DIR>(a, b, R, Ns, n, tw, nfft, fs, es); break;
            case 3: rt_stage<3, DIR>(a, b, R, Ns, n, tw, nfft, fs, es); break;
            case 4: rt_stage<4, DIR>(a, b, R, Ns, n, tw, nfft, fs, es); break;
            case 5: rt_stage<5, DIR>(a, b, R, Ns, n, tw, nfft, fs, es); break;
            case 7: rt_stage<7, DIR>(a, b, R, Ns, n, tw, nfft, fs, es); break;
            case 8: rt_stage<8, DIR>(a, b, R, Ns, n, tw, nfft, fs, es); break;
            default: rt_stage<0, DIR>(a, b, R, Ns, n, tw, nfft, fs, es); break;
        }
        __syncthreads();
        float2* tmp = a;
        a = b;
        b = tmp;
        Ns *= R;
    }
    return a;
}

// Transforms of a plane or column tile: the compiled pass for the length
// when there is one (pf_ct_fft.h), else the run-time planned one.  COLS:
// transform f at a[f + e st]; otherwise at a[f st + e] (ODDROWS: the odd row
// lengths).
template <int DIR, bool COLS, bool ODDROWS = false>
__device__ __forceinline__ float2* tile_fft(float2* a, float2* b, const PlanRt& p,
                                            const float2* tw, int nfft, int st) {
    float2* out;
    if (ct_dispatch<DIR, COLS, ODDROWS>(p.n, a, b, nfft, st, tw, &out)) return out;
    return lds_fft_rt<DIR>(a, b, p, tw, nfft, COLS ? 1 : st, COLS ? st : 1);
}

// Row-major stepping over an (rows x cols) index space by the block size:
// (r, c) of u = threadIdx.x, advanced by adds and compares.
struct Step2 {
    int r, c, dr, dc, cols;
    __device__ __forceinline__ Step2(int cols_) : cols(cols_) {
        r = (int)threadIdx.x / cols;
        c = (int)threadIdx.x - r * cols;
        dr = (int)blockDim.x / cols;
        dc = (int)blockDim.x - dr * cols;
    }
    __device__ __forceinline__ void next() {
        r += dr;
        c += dc;
        if (c >= cols) {
            c -= cols;
            ++r;
        }
    }
};

// Copy a twiddle table into LDS.
__device__ __forceinline__ void stage_tw(float2* dst, const float2* __restrict__ src, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Plane geometry.  Even nx: each row's x transform is a half-length complex
// FFT (rl = M = nx / 2 points, nr = ny rows).  Odd nx: rows y = 2p and 2p + 1
// travel as one complex row (row 2p + i row 2p+1) through an nx-point FFT
// (rl = nx, nr = P = ceil(ny / 2) rows), separated by Hermitian symmetry on
// the way out and packed the same way on the way in.
struct RtGeo {
    int nz, ny, nx, M, nxh, plane;   // M = nx / 2, nxh = nx / 2 + 1, plane = ny * nxh
    int odd, rl, nr, buf;            // odd nx; row FFT length and count; LDS buffer size
    int64_t H;
};

__host__ __device__ inline RtGeo make_geo(int nz, int ny, int nx) {
    RtGeo g;
    g.nz = nz;
    g.ny = ny;
    g.nx = nx;
    g.M = nx / 2;
    g.nxh = nx / 2 + 1;
    g.plane = ny * g.nxh;
    g.odd = nx % 2;
    g.rl = g.odd ? nx : g.M;
    g.nr = g.odd ? (ny + 1) / 2 : ny;
    g.buf = g.odd ? (g.plane > g.nr * nx ? g.plane : g.nr * nx) : g.plane;
    g.H = (int64_t)nz * g.plane;
    return g;
}

// LDS of a plane kernel: two buffers (a plane of ny x nxh complex, or the
// packed rows of an odd-nx plane), then the twiddles of the x (rl), y (ny) and
// R2C/C2R (nx) tables.
__host__ __device__ inline size_t plane_lds_bytes(const RtGeo& g) {
    return sizeof(float2) * ((size_t)2 * g.buf + g.rl + g.ny + g.nx);
}

// Forward stage for grid GI (0: P, gathering and parking W; 1: W; 2: W^2),
// one launch per grid.  Planes zlo, zlo + 1, ... (mod nz) of the support.
template <bool ODD, int GI>
__global__ __launch_bounds__(kRtThreads) void k_fwd_plane_rt(
    const float2* __restrict__ tm, const uint32_t* __restrict__ occ,
    const double* __restrict__ rots, RtGeo g, int r, int zr,
    int zlo, float2* __restrict__ out, PlanRt px, PlanRt py,
    const float2* __restrict__ twx_g, const float2* __restrict__ twy_g,
    const float2* __restrict__ twf_g) {
    extern __shared__ float2 lds[];
    float2* const A = lds;
    float2* const B = lds + g.buf;
    float2* const twx = lds + 2 * g.buf;
    float2* const twy = twx + g.rl;
    float2* const twfull = twy + g.ny;
    stage_tw(twx, twx_g, g.rl);
    stage_tw(twy, twy_g, g.ny);
    stage_tw(twfull, twf_g, g.nx);
    const int iz = (blockIdx.x + zlo + g.nz) % g.nz, b = blockIdx.y;
    const int sz = scoord(iz, r, g.nz);
    // gathered voxels: inside the rotation sphere and the support ball
    const int rg = r < zr ? r : zr, rg2 = rg * rg;
    double R[9];
    if constexpr (GI == 0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) R[i] = rots[9 * b + i];
    }
    float2* dst = out + (int64_t)b * 3 * g.H + (int64_t)iz * g.plane;
    // The mask values gathered with P are parked in this plane's region of
    // grid 2 (read back by the W and W^2 launches; grid 2 is written last, by
    // the block of the same plane, after all its reads), voxel (y, x) at float
    // y nx + x.
    float2* const wst2 = dst + 2 * g.H;
    float* const wst = reinterpret_cast<float*>(wst2);
    const int rows = g.nr * g.rl;
    {
        constexpr int gi = GI;
        // packed rows of the x transform, gathered for grid 0 (P, with W
        // parked) and read back for W and W^2
        if constexpr (!ODD) {
            // row y, element m: (v(2m), v(2m + 1)), row stride M (the pair is
            // float2 y M + m of the parked values)
            const int M = g.M;
            Step2 it(M);
            for (int u = threadIdx.x; u < rows; u += blockDim.x, it.next()) {
                const int y = it.r, m = it.c;
                const int sy = scoord(y, r, g.ny);
                float2 v2 = c2(0.f, 0.f);
                if (sz * sz + sy * sy <= rg2) {
                    if (gi == 0) {
                        float2 w2 = c2(0.f, 0.f);
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int sx = scoord(2 * m + h, r, g.nx);
                            if (sz * sz + sy * sy + sx * sx <= rg2) {
                                float p, w;
                                gather_pw_rt(tm, occ, g.nz, g.ny, g.nx, R, sz, sy, sx, &p, &w);
                                if (h) {
                                    v2.y = p;
                                    w2.y = w;
                                } else {
                                    v2.x = p;
                                    w2.x = w;
                                }
                            }
                        }
                        wst2[u] = w2;
                    } else {
                        const float2 w2 = wst2[u];
                        v2 = gi == 1 ? w2 : c2(w2.x * w2.x, w2.y * w2.y);
                    }
                }
                A[u] = v2;
            }
        } else {
            // row p, element x: (v(2p, x), v(2p + 1, x))
            auto vox = [=](int y, int x) -> float {
                const int sy = scoord(y, r, g.ny), sx = scoord(x, r, g.nx);
                if (sz * sz + sy * sy + sx * sx > rg2) return 0.f;
                if (gi == 0) {
                    float p, w;
                    gather_pw_rt(tm, occ, g.nz, g.ny, g.nx, R, sz, sy, sx, &p, &w);
                    wst[y * g.nx + x] = w;
                    return p;
                }
                const float w = wst[y * g.nx + x];
                return gi == 1 ? w : w * w;
            };
            Step2 it(g.rl);
            for (int u = threadIdx.x; u < rows; u += blockDim.x, it.next()) {
                const int y0 = 2 * it.r, x = it.c;
                float2 v2 = c2(vox(y0, x), 0.f);
                if (y0 + 1 < g.ny) v2.y = vox(y0 + 1, x);
                A[u] = v2;
            }
        }
        __syncthreads();
        float2* Z = tile_fft<-1, false, ODD>(A, B, px, twx, g.nr, g.rl);
        float2* Xp = Z == A ? B : A;
        // R2C post-processing into rows of nxh.  Even nx: X[k] = E[k] + w^k O[k]
        // from the half-length row, X[0] and X[M] real.  Odd nx: rows 2p and
        // 2p + 1 from the packed row z: (z[k] + conj z[-k]) / 2 and
        // (z[k] - conj z[-k]) / 2i.
        {
            const int M = g.M, nx = g.nx;
            Step2 ip(g.nxh);
            for (int u = threadIdx.x; u < g.plane; u += blockDim.x, ip.next()) {
                const int k = ip.c, y = ip.r;
                float2 X;
                if constexpr (!ODD) {
                    const float2* z = Z + y * M;
                    if (k == 0 || k == M) {
                        const float2 z0 = z[0];
                        X = c2(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0.f);
                    } else {
                        const float2 zk = z[k], zm = cconj(z[M - k]);
                        const float2 E = cscale(cadd(zk, zm), 0.5f);
                        const float2 O = mul_si<-1>(cscale(csub(zk, zm), 0.5f));
                        X = cadd(E, cmul(twfull[k], O));
                    }
                } else {
                    const float2* z = Z + (y >> 1) * nx;
                    const float2 zk = z[k], zm = cconj(z[k ? nx - k : 0]);
                    X = (y & 1) ? mul_si<-1>(cscale(csub(zk, zm), 0.5f))
                                : cscale(cadd(zk, zm), 0.5f);
                }
                Xp[u] = X;
            }
        }
        __syncthreads();
        // y FFT of the nxh columns (element stride nxh)
        const float2* Y = tile_fft<-1, true>(Xp, Z, py, twy, g.nxh, g.nxh);
        float2* o = dst + (int64_t)gi * g.H;
        for (int u = threadIdx.x; u < g.plane; u += blockDim.x) o[u] = Y[u];
        __syncthreads();
    }
}

// Inverse stage and LCC.  Each thread owns VOX voxel slots j of the plane:
// even nx, voxel u = threadIdx.x + j blockDim.x (row-major (y, x)); odd nx,
// packed element e = threadIdx.x + (j / 2) blockDim.x of the x-transformed
// rows, voxel (2 (e / nx) + j % 2, e % nx).  It keeps B, then 1 / sqrt(N V B -
// A^2), of them in registers (VOX 12 where the plane allows: fewer registers).
template <bool ODD, int VOX>
__global__ __launch_bounds__(kRtThreads) void k_inv_plane_rt(
    const float2* __restrict__ X, const uint8_t* __restrict__ lmask,
    unsigned long long* __restrict__ keys, RtGeo g, int nrot, int chunk, int64_t first,
    double norm, PlanRt px, PlanRt py, const float2* __restrict__ twx_g,
    const float2* __restrict__ twy_g, const float2* __restrict__ twf_g) {
    extern __shared__ float2 lds[];
    float2* const A = lds;
    float2* const B = lds + g.buf;
    float2* const twx = lds + 2 * g.buf;
    float2* const twy = twx + g.rl;
    float2* const twfull = twy + g.ny;
    stage_tw(twx, twx_g, g.rl);
    stage_tw(twy, twy_g, g.ny);
    stage_tw(twfull, twf_g, g.nx);
    const int iz = blockIdx.x;
    const int b0 = blockIdx.y * chunk, b1 = min(nrot, b0 + chunk);
    const int nvox = g.ny * g.nx, npk = g.nr * g.nx;
    const double nv = norm * ((double)g.nz * g.ny * g.nx);   // N V (exact)
    const int64_t vz = (int64_t)iz * nvox;
    static_assert(VOX <= 32 && VOX % 2 == 0, "mask bits of the owned voxels fit a 32-bit word");
    // slot j: float index of its value in the x-transformed plane (-1: none)
    auto slot_fi = [&](int j) -> int {
        if constexpr (!ODD) {
            const int u = threadIdx.x + j * blockDim.x;
            return u < nvox ? u : -1;
        }
        const int e = threadIdx.x + (j >> 1) * blockDim.x;
        return e < npk ? 2 * e + (j & 1) : -1;
    };
    // slot j's voxel within the plane (row-major), for a slot with a value
    auto slot_vox = [&](int j) -> int {
        if constexpr (!ODD) return threadIdx.x + j * blockDim.x;
        const int e = threadIdx.x + (j >> 1) * blockDim.x;
        const int p = e / g.nx;
        return (2 * p + (j & 1)) * g.nx + (e - p * g.nx);
    };
    uint32_t mbits = 0;
#pragma unroll
    for (int j = 0; j < VOX; ++j) {
        const int fi = slot_fi(j);
        if (fi < 0) continue;
        const int v = slot_vox(j);
        if (v < nvox && lmask[vz + v]) mbits |= 1u << j;
    }
    float held[VOX];
    for (int b = b0; b < b1; ++b) {
        for (int gi = 2; gi >= 0; --gi) {
            const float2* src = X + ((int64_t)b * 3 + gi) * g.H + (int64_t)iz * g.plane;
            // inverse y over the nxh columns; the compiled pass reads the
            // plane from global memory in its first stage
            const int nxh0 = g.nxh;
            auto ldg = [=](int k, int y) -> float2 { return src[y * nxh0 + k]; };
            float2* Y;
            if (!ct_dispatch<1, true>(g.ny, B, A, g.nxh, g.nxh, twy, &Y, ldg)) {
                for (int u = threadIdx.x; u < g.plane; u += blockDim.x) A[u] = src[u];
                __syncthreads();
                Y = lds_fft_rt<1>(A, B, py, twy, g.nxh, 1, g.nxh);
            }
            float2* D = Y == A ? B : A;
            // C2R pre-processing of the rows of nxh into the x transform's
            // rows.  Even nx, rows of M: z[k] = (X[k] + conj X[M-k]) +
            // i w^-k (X[k] - conj X[M-k]).  Odd nx, packed rows of nx:
            // z[k] = X_2p[k] + i X_2p+1[k], each X Hermitian-extended
            // (X[k] = conj X[nx-k] for k > nx / 2) with the DC bin real.
            const int M = g.M, nxh = g.nxh, nx = g.nx, ny = g.ny;
            auto pre = [=](int f, int k) -> float2 {
                if constexpr (!ODD) {
                    const float2* xr = Y + f * nxh;
                    float2 xk = xr[k], xm = xr[M - k];
                    if (k == 0) {
                        xk.y = 0.f;
                        xm.y = 0.f;
                    }
                    const float2 xmc = cconj(xm);
                    const float2 d = cmul(csub(xk, xmc), cconj(twfull[k]));
                    return cadd(cadd(xk, xmc), mul_si<1>(d));
                }
                auto herm = [&](int y) -> float2 {
                    const float2* xr = Y + y * nxh;
                    if (k == 0) return c2(xr[0].x, 0.f);
                    return k < nxh ? xr[k] : cconj(xr[nx - k]);
                };
                const float2 a = herm(2 * f);
                const float2 bb = 2 * f + 1 < ny ? herm(2 * f + 1) : c2(0.f, 0.f);
                return c2(a.x - bb.y, a.y + bb.x);
            };
            // even nx: the compiled x pass computes the pre-processing while its
            // first stage reads; odd nx: the packing reads two rows per element,
            // which spills inside that stage, so it is a pass of its own
            float2* Z = nullptr;
            bool done = false;
            if constexpr (!ODD) done = ct_dispatch<1, false>(g.rl, Y, D, g.nr, g.rl, twx, &Z, pre);
            if (!done) {
                const int rows = g.nr * g.rl;
                Step2 it(g.rl);
                for (int u = threadIdx.x; u < rows; u += blockDim.x, it.next())
                    D[u] = pre(it.r, it.c);
                __syncthreads();
                if constexpr (ODD) Z = tile_fft<1, false, true>(D, Y, px, twx, g.nr, g.rl);
                else Z = lds_fft_rt<1>(D, Y, px, twx, g.nr, g.rl, 1);
            }
            // even nx: row y holds 2 M reals (voxel u at float u); odd nx: the
            // packed element e holds rows 2p and 2p + 1 as its real and
            // imaginary parts
            // One unrolled slot loop per grid, each straight-line in held[]
            // (a grid test inside the slot loop makes the compiler copy the
            // whole array at every slot).  Slots without a voxel read the
            // plane's first value instead and hold values nothing reads (their
            // mask bit is clear).
            const float* zr = reinterpret_cast<const float*>(Z);
            if (gi == 2) {
#pragma unroll
                for (int j = 0; j < VOX; ++j) held[j] = zr[max(slot_fi(j), 0)];
            } else if (gi == 1) {
#pragma unroll
                for (int j = 0; j < VOX; ++j) {
                    const double ad = (double)zr[max(slot_fi(j), 0)];
                    held[j] = __builtin_amdgcn_rsqf((float)fma(-ad, ad, (double)held[j] * nv));
                }
            } else {
#pragma unroll
                for (int j = 0; j < VOX; ++j) {
                    if (!((mbits >> j) & 1u)) continue;
                    // only S > 0 can beat the initial (0, 0); NaN never wins
                    const float sv = zr[slot_fi(j)] * held[j];
                    if (sv > 0.f)
                        atomicMax(keys + vz + slot_vox(j), lcc_key(sv, (uint32_t)(first + b)));
                }
            }
            __syncthreads();
        }
    }
}

int launched(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return kErrHip;
    }
    return kOk;
}

// Let a kernel take more than the default dynamic LDS (once per process and
// device; the attribute is per function).
int allow_lds(const void* fn, size_t bytes) {
    const hipError_t e =
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipFuncSetAttribute(MaxDynamicSharedMemorySize, %zu): %s", bytes,
                  hipGetErrorString(e));
        return kErrHip;
    }
    return kOk;
}


// Launch the forward (which = 0) or inverse (which = 2) plane kernel for x
// lengths of parity ODD.  Plane kernels run 512-thread blocks where two fit a
// compute unit's LDS (and the LCC stage's voxels per thread allow), so one
// block's barriers overlap the other's work; otherwise one 1024-thread block
// per unit.  Forward planes |z| <= zr only when skip (the z stage reads the
// others as zero).
template <bool ODD>
int launch_plane_kernel(const Fft3D& f, const FusedArgs& a, int which, bool skip, int zr,
                        hipStream_t s) {
    const RtGeo g = make_geo(f.nz, f.ny, f.nx);
    PlanRt px, py;
    if (!pack_plan(f.px, &px) || !pack_plan(f.py, &py)) {
        set_error("plane kernels: no packed radix plan for (%d, %d)", f.ny, f.nx);
        return kErrUnsupported;
    }
    const size_t lds_plane = plane_lds_bytes(g);
    const int nthr_plane =
        (2 * lds_plane <= kCuLds && (int64_t)f.ny * f.nx <= (int64_t)kRtVox * 512) ? 512
                                                                                   : kRtThreads;
    if (which == 0) {
        const int nplanes = skip ? 2 * zr + 1 : f.nz;
        const dim3 grid(nplanes, a.nrot);
        hipLaunchKernelGGL((k_fwd_plane_rt<ODD, 0>), grid, dim3(nthr_plane), lds_plane, s, a.tm,
                           a.occ, a.rots, g, a.rmax, a.zr, skip ? -zr : 0, a.work, px, py, f.twx,
                           f.twy, f.twfull);
        PF_CHECK_STATUS(launched("k_fwd_plane_rt"));
        hipLaunchKernelGGL((k_fwd_plane_rt<ODD, 1>), grid, dim3(nthr_plane), lds_plane, s, a.tm,
                           a.occ, a.rots, g, a.rmax, a.zr, skip ? -zr : 0, a.work, px, py, f.twx,
                           f.twy, f.twfull);
        PF_CHECK_STATUS(launched("k_fwd_plane_rt"));
        if (!a.binary_mask) {
            hipLaunchKernelGGL((k_fwd_plane_rt<ODD, 2>), grid, dim3(nthr_plane), lds_plane, s,
                               a.tm, a.occ, a.rots, g, a.rmax, a.zr, skip ? -zr : 0, a.work, px,
                               py, f.twx, f.twy, f.twfull);
            PF_CHECK_STATUS(launched("k_fwd_plane_rt"));
        }
        return kOk;
    }
    const dim3 grid3(f.nz, (a.nrot + a.chunk_lcc - 1) / a.chunk_lcc);
    if ((int64_t)f.ny * f.nx <= (int64_t)12 * nthr_plane)
        hipLaunchKernelGGL((k_inv_plane_rt<ODD, 12>), grid3, dim3(nthr_plane), lds_plane, s,
                           a.work, a.lcc_mask, a.keys, g, a.nrot, a.chunk_lcc, a.first, a.norm,
                           px, py, f.twx, f.twy, f.twfull);
    else
        hipLaunchKernelGGL((k_inv_plane_rt<ODD, kRtVox>), grid3, dim3(nthr_plane), lds_plane, s,
                           a.work, a.lcc_mask, a.keys, g, a.nrot, a.chunk_lcc, a.first, a.norm,
                           px, py, f.twx, f.twy, f.twfull);
    return launched("k_inv_plane_rt");
}

template <bool ODD>
int prepare_plane_kernels() {
    PF_CHECK_STATUS(allow_lds((const void*)k_fwd_plane_rt<ODD, 0>, kRtLds));
    PF_CHECK_STATUS(allow_lds((const void*)k_fwd_plane_rt<ODD, 1>, kRtLds));
    PF_CHECK_STATUS(allow_lds((const void*)k_fwd_plane_rt<ODD, 2>, kRtLds));
    PF_CHECK_STATUS(allow_lds((const void*)k_inv_plane_rt<ODD, 12>, kRtLds));
    PF_CHECK_STATUS(allow_lds((const void*)k_inv_plane_rt<ODD, kRtVox>, kRtLds));
    return kOk;
}

}  // namespace

// Odd x lengths (pf_planes_odd.hip).
int prepare_plane_kernels_odd();
int launch_plane_kernel_odd(const Fft3D& f, const FusedArgs& a, int which, bool skip, int zr,
                            hipStream_t s);

}  // namespace pf
