// Generic HIP kernels of the LCC search: rotation, 3-D R2C/C2R FFT passes,
// conj-multiply, calc_lcc + take-best.  Every kernel takes its extents as
// arguments and guards each global access, so any shape the host validates
// is safe to launch.
#include <math.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "pf_ct_fft.h"
#include "pf_internal.h"
#include "pf_rotate.h"

namespace pf {

// nfft transforms of p.n points in LDS, ping-ponging between a and b: the
// pass compiled for the length (pf_ct_fft.h) where there is one, else the
// run-time planned lds_fft.  COLS: transform f at a[f + e st]; otherwise at
// a[f st + e].  Returns the buffer holding the result; ends with a barrier.
template <int DIR, bool COLS>
__device__ __forceinline__ const float2* axis_fft(float2* a, float2* b, const Plan1D& p,
                                                  const float2* tw, int nfft, int st) {
    float2* out;
    if (ct_dispatch<DIR, COLS>(p.n, a, b, nfft, st, tw, &out)) return out;
    return lds_fft<DIR>(a, b, p, tw, nfft, COLS ? 1 : st, COLS ? st : 1);
}

// ---------------------------------------------------------------------------
// R2C along the contiguous x axis: rows of nx reals -> nx/2+1 complex.
__global__ void k_r2c_x(const float* __restrict__ in, float2* __restrict__ out, int nx,
                        int64_t rows, Plan1D p, const float2* __restrict__ tw,
                        const float2* __restrict__ twfull, int rpw) {
    extern __shared__ float2 lds[];
    const int M = p.n;
    const bool even = (nx % 2) == 0;
    const int nxh = nx / 2 + 1;
    const int64_t row0 = (int64_t)blockIdx.x * rpw;
    if (row0 >= rows) return;
    const int nrows = (int)min((int64_t)rpw, rows - row0);
    float2* A = lds;
    float2* B = lds + (size_t)rpw * M;
    for (int u = threadIdx.x; u < nrows * M; u += blockDim.x) {
        const int f = u / M, m = u - f * M;
        const float* src = in + (row0 + f) * nx;
        A[f * M + m] = even ? c2(src[2 * m], src[2 * m + 1]) : c2(src[m], 0.f);
    }
    __syncthreads();
    const float2* Z = axis_fft<-1, false>(A, B, p, tw, nrows, M);
    for (int u = threadIdx.x; u < nrows * nxh; u += blockDim.x) {
        const int f = u / nxh, k = u - f * nxh;
        const float2* z = Z + f * M;
        float2 X;
        if (even && (k == 0 || k == M)) {
            // DC and Nyquist are real: the sum and the difference of the even
            // and odd sums, with an imaginary part of exactly 0 as numpy's
            const float2 z0 = z[0];
            X = c2(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0.f);
        } else if (even) {
            const float2 zk = z[k];
            const float2 zm = cconj(z[M - k]);
            const float2 E = cscale(cadd(zk, zm), 0.5f);
            const float2 O = mul_si<-1>(cscale(csub(zk, zm), 0.5f));
            X = cadd(E, cmul(twfull[k], O));
        } else {
            X = z[k];
        }
        out[(row0 + f) * nxh + k] = X;
    }
}

// C2R along x: rows of nx/2+1 complex -> nx reals (unnormalised, times scale).
// Imaginary parts of the DC (and, for even nx, Nyquist) terms are ignored,
// as numpy's irfft does.
__global__ void k_c2r_x(const float2* __restrict__ in, float* __restrict__ out, int nx,
                        int64_t rows, Plan1D p, const float2* __restrict__ tw,
                        const float2* __restrict__ twfull, int rpw, float scale) {
    extern __shared__ float2 lds[];
    const int M = p.n;
    const bool even = (nx % 2) == 0;
    const int nxh = nx / 2 + 1;
    const int64_t row0 = (int64_t)blockIdx.x * rpw;
    if (row0 >= rows) return;
    const int nrows = (int)min((int64_t)rpw, rows - row0);
    float2* A = lds;
    float2* B = lds + (size_t)rpw * M;
    for (int u = threadIdx.x; u < nrows * M; u += blockDim.x) {
        const int f = u / M, k = u - f * M;
        const float2* X = in + (row0 + f) * nxh;
        float2 z;
        if (even) {
            float2 xk = X[k];
            float2 xm = X[M - k];
            if (k == 0) {
                xk.y = 0.f;
                xm.y = 0.f;
            }
            const float2 xmc = cconj(xm);
            const float2 d = cmul(csub(xk, xmc), cconj(twfull[k]));
            z = cadd(cadd(xk, xmc), mul_si<1>(d));
        } else {
            if (k == 0) z = c2(X[0].x, 0.f);
            else if (k <= nx / 2) z = X[k];
            else z = cconj(X[nx - k]);
        }
        A[f * M + k] = z;
    }
    __syncthreads();
    const float2* Z = axis_fft<1, false>(A, B, p, tw, nrows, M);
    for (int u = threadIdx.x; u < nrows * M; u += blockDim.x) {
        const int f = u / M, m = u - f * M;
        float* dst = out + (row0 + f) * nx;
        const float2 z = Z[f * M + m];
        if (even) {
            dst[2 * m] = z.x * scale;
            dst[2 * m + 1] = z.y * scale;
        } else {
            dst[m] = z.x * scale;
        }
    }
}

// In-place C2C along a strided axis: element (o, k, c) at
// data[o * outer_stride + k * stride + c], k < n, c < inner.  One block per
// (outer index, tile of tq consecutive c).
template <int DIR>
__global__ void k_c2c_axis(float2* __restrict__ data, int64_t stride, int64_t inner,
                           int64_t outer, int64_t outer_stride, int64_t ntiles, Plan1D p,
                           const float2* __restrict__ tw, int tq) {
    extern __shared__ float2 lds[];
    const int n = p.n;
    const int64_t blk = blockIdx.x;
    const int64_t o = blk / ntiles;
    const int64_t c0 = (blk - o * ntiles) * tq;
    if (o >= outer || c0 >= inner) return;
    const int nc = (int)min((int64_t)tq, inner - c0);
    float2* A = lds;
    float2* B = lds + (size_t)n * tq;
    float2* base = data + o * outer_stride + c0;
    for (int u = threadIdx.x; u < n * nc; u += blockDim.x) {
        const int k = u / nc, c = u - k * nc;
        A[k * tq + c] = base[k * stride + c];
    }
    __syncthreads();
    const float2* Z = axis_fft<DIR, true>(A, B, p, tw, nc, tq);
    for (int u = threadIdx.x; u < n * nc; u += blockDim.x) {
        const int k = u / nc, c = u - k * nc;
        base[k * stride + c] = Z[k * tq + c];
    }
}

// Template (trilinear) and mask (nearest) of `nrot` rotations; blockIdx.y = rotation.
__global__ void k_rotate_pw(const float* __restrict__ tmpl, const float* __restrict__ mask,
                            const double* __restrict__ rots, int nz, int ny, int nx, int r,
                            float* __restrict__ out) {
    const int64_t V = (int64_t)nz * ny * nx;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int b = blockIdx.y;
    const int ix = (int)(v % nx);
    const int iy = (int)((v / nx) % ny);
    const int iz = (int)(v / ((int64_t)nx * ny));
    float* P = out + (int64_t)(3 * b) * V;
    float* W = P + V;
    float* W2 = W + V;
    int z, y, x;
    if (!signed_triple(iz, iy, ix, nz, ny, nx, r, &z, &y, &x)) {
        P[v] = 0.f;
        W[v] = 0.f;
        W2[v] = 0.f;
        return;
    }
    double R[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = rots[9 * b + i];
    const Coord c = source_coord(R, z, y, x);
    const double p = sample_trilinear(tmpl, nz, ny, nx, c);
    const double w = sample_nearest(mask, nz, ny, nx, c);
    P[v] = (float)p;
    W[v] = (float)w;
    W2[v] = (float)(w * w);
}

// Drop-in rotate_grid3d on FP64 buffers: in-sphere voxels only.
__global__ void k_rotate_f64(const double* __restrict__ grid, int gz, int gy, int gx,
                             const double* __restrict__ rot, int r, double* __restrict__ out,
                             int oz, int oy, int ox, int nearest) {
    const int64_t V = (int64_t)oz * oy * ox;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int ix = (int)(v % ox);
    const int iy = (int)((v / ox) % oy);
    const int iz = (int)(v / ((int64_t)ox * oy));
    int z, y, x;
    if (!signed_triple(iz, iy, ix, oz, oy, ox, r, &z, &y, &x)) return;
    double R[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = rot[i];
    const Coord c = source_coord(R, z, y, x);
    out[v] = nearest ? sample_nearest(grid, gz, gy, gx, c)
                     : sample_trilinear(grid, gz, gy, gx, c);
}

// ---------------------------------------------------------------------------
__global__ void k_conj_mul_pw(float2* __restrict__ X, const float2* __restrict__ F1,
                              const float2* __restrict__ F2, int64_t H, int nrot) {
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= H) return;
    const float2 f1 = F1[h], f2 = F2[h];
    for (int b = 0; b < nrot; ++b) {
        float2* x = X + (int64_t)(3 * b) * H + h;
        x[0] = cmul(cconj(x[0]), f1);
        x[H] = cmul(cconj(x[H]), f1);
        x[2 * H] = cmul(cconj(x[2 * H]), f2);
    }
}

__global__ void k_conj_multiply(const float2* __restrict__ a, const float2* __restrict__ b,
                                float2* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cmul(cconj(a[i]), b[i]);
}

// calc_lcc (reference: _powerfit.pyx:56-72, with ave2 *= N from
// powerfitter.py:369) and take-best (powerfitter.py:327-330): strict '>'
// over the batch in rotation order, merged into the running result with
// atomicMax on lcc_key (ties keep the lower rotation index, NaN never wins).
// The local variance N<W^2 T^2> - <W T>^2 is formed in FP64.
__global__ void k_lcc_take_best(const float* __restrict__ real, const uint8_t* __restrict__ mask,
                                unsigned long long* __restrict__ keys, int64_t V, int nrot,
                                int64_t first, double norm, double invV) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || !mask[v]) return;
    float bv = 0.f;
    int32_t bi = -1;
    for (int b = 0; b < nrot; ++b) {
        const float* base = real + (int64_t)(3 * b) * V + v;
        const double g = (double)base[0] * invV;
        const double a = (double)base[V] * invV;
        const double a2 = (double)base[2 * V] * invV * norm;
        const float s = (float)(g / sqrt(a2 - a * a));
        if (bi < 0 ? (s == s) : (s > bv)) {
            bv = s;
            bi = (int32_t)(first + b);
        }
    }
    if (bi >= 0) atomicMax(keys + v, lcc_key(bv, (uint32_t)bi));
}

__global__ void k_fill_keys(unsigned long long* __restrict__ keys, int64_t n,
                            unsigned long long value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = value;
}

__global__ void k_calc_lcc(const float* __restrict__ gcc, const float* __restrict__ ave,
                           const float* __restrict__ ave2, const uint8_t* __restrict__ mask,
                           float* __restrict__ lcc, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !mask[i]) return;
    const double a = ave[i];
    lcc[i] = (float)((double)gcc[i] / sqrt((double)ave2[i] - a * a));
}

__global__ void k_convert(const double* __restrict__ in, float* __restrict__ out,
                          float* __restrict__ out_sq, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = in[i];
    out[i] = (float)x;
    if (out_sq) out_sq[i] = (float)(x * x);
}

// ---------------------------------------------------------------------------
// Host side.

static thread_local char g_err[1024];

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

}  // namespace pf

extern "C" const char* pf_last_error(void) { return pf::g_err; }

namespace pf {

static constexpr int kThreads = 256;
static constexpr size_t kLdsBudget = 64 * 1024;

static int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return kErrHip;
    }
    return kOk;
}

int twiddles(int n, const float2** out) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, float2*> cache;
    int dev = 0;
    PF_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_pair(dev, n);
    auto it = cache.find(key);
    if (it != cache.end()) {
        *out = it->second;
        return kOk;
    }
    // exp(-2 pi i m / n) = (-i)^q exp(-i (pi/2) r / n) with 4 m = q n + r,
    // 0 <= r < n: the quarter-turn points come out exactly (1, 0), (0, -1),
    // (-1, 0), (0, 1), as in numpy's FFT, instead of carrying the rounding of
    // cos and sin at multiples of pi / 2
    std::vector<float2> h(n);
    for (int m = 0; m < n; ++m) {
        const int64_t m4 = 4LL * m;
        const int q = (int)(m4 / n), r = (int)(m4 % n);
        const double th = 0.5 * M_PI * (double)r / (double)n;
        const double c = r == 0 ? 1.0 : cos(th), s = r == 0 ? 0.0 : sin(th);
        double re, im;   // (c, -s) times (-i)^q
        switch (q) {
            case 0: re = c; im = -s; break;
            case 1: re = -s; im = -c; break;
            case 2: re = -c; im = s; break;
            default: re = s; im = c; break;
        }
        h[m] = c2((float)(re + 0.0), (float)(im + 0.0));   // no negative zeros
    }
    float2* d = nullptr;
    PF_HIP_CHECK(hipMalloc(&d, sizeof(float2) * n));
    PF_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    cache[key] = d;
    *out = d;
    return kOk;
}

static int max_len_ok(int n) {
    // ping-pong LDS with at least one transform per block
    return (size_t)n * 2 * sizeof(float2) <= kLdsBudget;
}

int fft3d_init(Fft3D* f, int nz, int ny, int nx) {
    if (nz < 1 || ny < 1 || nx < 1) {
        set_error("fft3d: invalid shape (%d, %d, %d)", nz, ny, nx);
        return kErrArg;
    }
    f->nz = nz;
    f->ny = ny;
    f->nx = nx;
    f->nxh = nx / 2 + 1;
    const int mx = (nx % 2 == 0) ? nx / 2 : nx;
    if (!make_plan(mx, &f->px) || !make_plan(ny, &f->py) || !make_plan(nz, &f->pz)) {
        set_error("fft3d: cannot factor shape (%d, %d, %d)", nz, ny, nx);
        return kErrUnsupported;
    }
    if (!max_len_ok(mx) || !max_len_ok(ny) || !max_len_ok(nz)) {
        set_error("fft3d: axis length too large for LDS (%d, %d, %d)", nz, ny, nx);
        return kErrUnsupported;
    }
    PF_CHECK_STATUS(twiddles(mx, &f->twx));
    PF_CHECK_STATUS(twiddles(ny, &f->twy));
    PF_CHECK_STATUS(twiddles(nz, &f->twz));
    PF_CHECK_STATUS(twiddles(nx, &f->twfull));
    return kOk;
}

static int rows_per_block(int M) {
    int rpw = (int)(kLdsBudget / (2 * sizeof(float2) * (size_t)M));
    if (rpw > 32) rpw = 32;
    return rpw < 1 ? 1 : rpw;
}

int launch_r2c_x(const Fft3D& f, const float* in, float2* out, int64_t rows, hipStream_t s) {
    if (rows <= 0) return kOk;
    const int M = f.px.n;
    const int rpw = rows_per_block(M);
    const size_t lds = 2 * sizeof(float2) * (size_t)rpw * M;
    const int64_t blocks = (rows + rpw - 1) / rpw;
    hipLaunchKernelGGL(k_r2c_x, dim3((unsigned)blocks), dim3(kThreads), lds, s, in, out, f.nx,
                       rows, f.px, f.twx, f.twfull, rpw);
    return check_launch("k_r2c_x");
}

int launch_c2r_x(const Fft3D& f, const float2* in, float* out, int64_t rows, float scale,
                 hipStream_t s) {
    if (rows <= 0) return kOk;
    const int M = f.px.n;
    const int rpw = rows_per_block(M);
    const size_t lds = 2 * sizeof(float2) * (size_t)rpw * M;
    const int64_t blocks = (rows + rpw - 1) / rpw;
    hipLaunchKernelGGL(k_c2r_x, dim3((unsigned)blocks), dim3(kThreads), lds, s, in, out, f.nx,
                       rows, f.px, f.twx, f.twfull, rpw, scale);
    return check_launch("k_c2r_x");
}

int launch_c2c_axis(const Plan1D& p, const float2* tw, float2* data, int64_t stride,
                    int64_t inner, int64_t outer, int64_t outer_stride, int dir,
                    hipStream_t s) {
    if (p.n <= 1 || inner <= 0 || outer <= 0) return kOk;
    int tq = (int)(kLdsBudget / (2 * sizeof(float2) * (size_t)p.n));
    if (tq > 16) tq = 16;
    if (tq < 1) tq = 1;
    const int64_t ntiles = (inner + tq - 1) / tq;
    const int64_t blocks = ntiles * outer;
    if (blocks > 0x7fffffffLL) {
        set_error("c2c_axis: grid too large");
        return kErrUnsupported;
    }
    const size_t lds = 2 * sizeof(float2) * (size_t)tq * p.n;
    if (dir < 0)
        hipLaunchKernelGGL(k_c2c_axis<-1>, dim3((unsigned)blocks), dim3(kThreads), lds, s, data,
                           stride, inner, outer, outer_stride, ntiles, p, tw, tq);
    else
        hipLaunchKernelGGL(k_c2c_axis<1>, dim3((unsigned)blocks), dim3(kThreads), lds, s, data,
                           stride, inner, outer, outer_stride, ntiles, p, tw, tq);
    return check_launch("k_c2c_axis");
}

int fft3d_set_split(Fft3D* f, bool on) {
    if (!on) {
        f->split = false;
        return kOk;
    }
    if (!make_plan16(f->ny, &f->py16) || !make_plan16(f->nz, &f->pz16)) {
        set_error("split-precision FFT needs ny and nz divisible by 16 and at most %d "
                  "(got %d, %d)", kSplitMaxLen, f->ny, f->nz);
        return kErrUnsupported;
    }
    f->split = true;
    return kOk;
}

int launch_axis(const Fft3D& f, int axis, float2* data, int64_t stride, int64_t inner,
                int64_t outer, int64_t outer_stride, int dir, hipStream_t s) {
    const float2* tw = axis == 1 ? f.twy : f.twz;
    if (f.split)
        return launch_c2c_axis_split(axis == 1 ? f.py16 : f.pz16, tw, data, stride, inner, outer,
                                     outer_stride, dir, s);
    return launch_c2c_axis(axis == 1 ? f.py : f.pz, tw, data, stride, inner, outer,
                           outer_stride, dir, s);
}

int fft3d_r2c(const Fft3D& f, const float* in, float2* out, int batch, hipStream_t s) {
    const int64_t plane = (int64_t)f.ny * f.nxh;
    PF_CHECK_STATUS(launch_r2c_x(f, in, out, (int64_t)batch * f.nz * f.ny, s));
    PF_CHECK_STATUS(launch_axis(f, 1, out, f.nxh, f.nxh, (int64_t)batch * f.nz, plane, -1, s));
    PF_CHECK_STATUS(launch_axis(f, 2, out, plane, plane, batch, plane * f.nz, -1, s));
    return kOk;
}

int fft3d_c2r(const Fft3D& f, float2* in, float* out, int batch, float scale, hipStream_t s) {
    const int64_t plane = (int64_t)f.ny * f.nxh;
    PF_CHECK_STATUS(launch_axis(f, 2, in, plane, plane, batch, plane * f.nz, 1, s));
    PF_CHECK_STATUS(launch_axis(f, 1, in, f.nxh, f.nxh, (int64_t)batch * f.nz, plane, 1, s));
    PF_CHECK_STATUS(launch_c2r_x(f, in, out, (int64_t)batch * f.nz * f.ny, scale, s));
    return kOk;
}

static unsigned blocks_for(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

int launch_rotate_pw(const float* tmpl, const float* mask, const double* rots, int nrot,
                     int nz, int ny, int nx, int radius, float* out, hipStream_t s) {
    const int64_t V = (int64_t)nz * ny * nx;
    if (nrot <= 0) return kOk;
    if (nrot > 65535) {
        set_error("rotate: at most 65535 rotations per launch");
        return kErrArg;
    }
    hipLaunchKernelGGL(k_rotate_pw, dim3(blocks_for(V), nrot), dim3(kThreads), 0, s, tmpl, mask,
                       rots, nz, ny, nx, radius, out);
    return check_launch("k_rotate_pw");
}

int launch_rotate_f64(const double* grid, int gz, int gy, int gx, const double* rot,
                      int radius, double* out, int oz, int oy, int ox, int nearest,
                      hipStream_t s) {
    const int64_t V = (int64_t)oz * oy * ox;
    hipLaunchKernelGGL(k_rotate_f64, dim3(blocks_for(V)), dim3(kThreads), 0, s, grid, gz, gy, gx,
                       rot, radius, out, oz, oy, ox, nearest);
    return check_launch("k_rotate_f64");
}

int launch_conj_mul_pw(float2* X, const float2* F1, const float2* F2, int64_t H, int nrot,
                       hipStream_t s) {
    if (H <= 0 || nrot <= 0) return kOk;
    hipLaunchKernelGGL(k_conj_mul_pw, dim3(blocks_for(H)), dim3(kThreads), 0, s, X, F1, F2, H,
                       nrot);
    return check_launch("k_conj_mul_pw");
}

int launch_conj_multiply(const float2* a, const float2* b, float2* out, int64_t n,
                         hipStream_t s) {
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_conj_multiply, dim3(blocks_for(n)), dim3(kThreads), 0, s, a, b, out, n);
    return check_launch("k_conj_multiply");
}

int launch_lcc_take_best(const float* real, const uint8_t* mask, unsigned long long* keys,
                         int64_t V, int nrot, int64_t first, double norm, double invV,
                         hipStream_t s) {
    if (V <= 0 || nrot <= 0) return kOk;
    hipLaunchKernelGGL(k_lcc_take_best, dim3(blocks_for(V)), dim3(kThreads), 0, s, real, mask,
                       keys, V, nrot, first, norm, invV);
    return check_launch("k_lcc_take_best");
}

int launch_fill_keys(unsigned long long* keys, int64_t n, unsigned long long value,
                     hipStream_t s) {
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_fill_keys, dim3(blocks_for(n)), dim3(kThreads), 0, s, keys, n, value);
    return check_launch("k_fill_keys");
}

int launch_calc_lcc(const float* gcc, const float* ave, const float* ave2,
                    const uint8_t* mask, float* lcc, int64_t n, hipStream_t s) {
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_calc_lcc, dim3(blocks_for(n)), dim3(kThreads), 0, s, gcc, ave, ave2,
                       mask, lcc, n);
    return check_launch("k_calc_lcc");
}

int launch_convert(const double* in, float* out, float* out_sq, int64_t n, hipStream_t s) {
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_convert, dim3(blocks_for(n)), dim3(kThreads), 0, s, in, out, out_sq, n);
    return check_launch("k_convert");
}

}  // namespace pf
