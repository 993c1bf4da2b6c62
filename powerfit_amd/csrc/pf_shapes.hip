// Template and mask generation on the device: blur_points and dilate_points
// (src/powerfit_em/_powerfit.pyx:75-140 and :143-206), the step before the
// LCC search in the reference's CLI (volume.py:192-224, powerfit.py:245-255).
//
// The reference scatters every atom over its box, [ceil(p - e), floor(p + e)]
// per axis (e = 4 sigma or the atom's radius), clipped below at 0 (or at
// -dim + 1 with wraparound, a negative index addressing index + dim) and above
// at dim - 1, adding w exp(-d^2 / 2 sigma^2) (or writing 1) where d^2 <= e^2.
// Here each voxel gathers instead:
//   k_atom_boxes     the reference's clipped box of every atom;
//   k_brick_count    per 8^3 brick of the grid, the atoms whose box reaches it
//   k_scan_offsets   (directly or through the one wrap), counted, scanned and
//   k_brick_fill     listed in atom order (block-ordered append);
//   k_shape          one thread per voxel walks its brick's list in atom order
//                    and, per atom, the unwrapped coordinates {v - dim, v} that
//                    land on it in the reference's z, y, x loop order, so each
//                    voxel's sum has the reference's order of additions.
// d^2 is formed as in the reference with contraction off, so the mask is
// bit-identical; the template differs only by the rounding of exp.
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/powerfit_hip.h"
#include "pf_internal.h"

namespace pf {
namespace {

constexpr int kBrick = 8;
constexpr int kListThreads = 256;
constexpr int kScanThreads = 1024;

struct Grid {
    int n[3];      // x, y, z extents
    int nb[3];     // bricks per axis
};

__global__ void k_atom_boxes(const double* __restrict__ pts, const double* __restrict__ ext,
                             double ext_all, int64_t n, Grid g, int wrap,
                             int* __restrict__ box) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double e = ext ? ext[k] : ext_all;
    for (int a = 0; a < 3; ++a) {
        const double p = pts[(int64_t)a * n + k];
        double lo = ceil(p - e), hi = floor(p + e);
        int ilo = 1 << 30, ihi = -(1 << 30);   // empty (reaches no brick) unless finite
        if (lo == lo && hi == hi) {
            // clamp before the cast; the grid limits below are tighter
            lo = fmax(fmin(lo, 1e9), -1e9);
            hi = fmax(fmin(hi, 1e9), -1e9);
            ilo = max((int)lo, wrap ? -g.n[a] + 1 : 0);
            ihi = min((int)hi, g.n[a] - 1);
        }
        box[k * 6 + a] = ilo;
        box[k * 6 + 3 + a] = ihi;
    }
}

// Does the unwrapped range [lo, hi] (within [-n + 1, n - 1]) write an index
// in [b0, b1]?  Index i is written by u = i or u = i - n.
__device__ __forceinline__ bool axis_hits(int lo, int hi, int b0, int b1, int n) {
    return (lo <= b1 && hi >= b0) || (lo <= b1 - n && hi >= b0 - n);
}

__device__ __forceinline__ bool brick_hits(const int* __restrict__ bx, const int b0[3],
                                           const int b1[3], const Grid& g) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (!axis_hits(bx[a], bx[3 + a], b0[a], b1[a], g.n[a])) return false;
    return true;
}

__device__ __forceinline__ void brick_range(int64_t brick, const Grid& g, int b0[3], int b1[3]) {
    const int bx = (int)(brick % g.nb[0]);
    const int by = (int)((brick / g.nb[0]) % g.nb[1]);
    const int bz = (int)(brick / ((int64_t)g.nb[0] * g.nb[1]));
    const int c[3] = {bx, by, bz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b0[a] = c[a] * kBrick;
        b1[a] = min(b0[a] + kBrick, g.n[a]) - 1;
    }
}

__global__ __launch_bounds__(kListThreads) void k_brick_count(const int* __restrict__ box,
                                                              int64_t n, Grid g,
                                                              int64_t* __restrict__ counts) {
    __shared__ int64_t part[kListThreads / 64];
    int b0[3], b1[3];
    brick_range(blockIdx.x, g, b0, b1);
    int64_t c = 0;
    for (int64_t k = threadIdx.x; k < n; k += kListThreads) c += brick_hits(box + k * 6, b0, b1, g);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < kListThreads / 64; ++w) t += part[w];
        counts[blockIdx.x] = t;
    }
}

// Exclusive scan of counts[0..m) into offs[0..m], offs[m] = total (one block).
__global__ __launch_bounds__(kScanThreads) void k_scan_offsets(const int64_t* __restrict__ counts,
                                                               int64_t m,
                                                               int64_t* __restrict__ offs) {
    __shared__ int64_t buf[kScanThreads];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < m; base += kScanThreads) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < m ? counts[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kScanThreads; off <<= 1) {
            const int64_t add = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < m) offs[i] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += buf[kScanThreads - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) offs[m] = carry;
}

// Atoms reaching the brick, appended in atom order.
__global__ __launch_bounds__(kListThreads) void k_brick_fill(const int* __restrict__ box,
                                                             int64_t n, Grid g,
                                                             const int64_t* __restrict__ offs,
                                                             int32_t* __restrict__ list) {
    constexpr int NW = kListThreads / 64;
    __shared__ int wave_tot[NW];
    int b0[3], b1[3];
    brick_range(blockIdx.x, g, b0, b1);
    int64_t pos = offs[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t base = 0; base < n; base += kListThreads) {
        const int64_t k = base + threadIdx.x;
        const bool hit = k < n && brick_hits(box + k * 6, b0, b1, g);
        const uint64_t ball = __ballot(hit);
        const int before = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) wave_tot[wave] = __popcll(ball);
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w = 0; w < NW; ++w) {
            woff += (w < wave) ? wave_tot[w] : 0;
            tot += wave_tot[w];
        }
        if (hit) list[pos + woff + before] = (int32_t)k;
        pos += tot;
        __syncthreads();
    }
}

// Contribution of atom (p, box) at voxel index v: visits the unwrapped
// coordinates u in {v - n, v} inside the box, z outer, y, x inner, ascending.
// BLUR adds w exp(-d^2 / dsig2) to acc; otherwise sets hit.
template <bool BLUR>
__device__ __forceinline__ void visit_atom(const double p[3], const int* bx, const int v[3],
                                           const int n[3], double e2, double w, double dsig2,
                                           double& acc, bool& hit) {
#pragma clang fp contract(off)
#pragma unroll
    for (int iz = 0; iz < 2; ++iz) {
        const int uz = v[2] - (iz ? 0 : n[2]);
        if (uz < bx[2] || uz > bx[5]) continue;
        const double dz = (double)uz - p[2];
        const double z2 = dz * dz;
#pragma unroll
        for (int iy = 0; iy < 2; ++iy) {
            const int uy = v[1] - (iy ? 0 : n[1]);
            if (uy < bx[1] || uy > bx[4]) continue;
            const double dy = (double)uy - p[1];
            const double y2z2 = dy * dy + z2;
#pragma unroll
            for (int ix = 0; ix < 2; ++ix) {
                const int ux = v[0] - (ix ? 0 : n[0]);
                if (ux < bx[0] || ux > bx[3]) continue;
                const double dx = (double)ux - p[0];
                const double d2 = dx * dx + y2z2;
                if (d2 <= e2) {
                    if (BLUR) acc += w * exp(-d2 / dsig2);
                    else hit = true;
                }
            }
        }
    }
}

// One block per brick, one thread per voxel of it.  BLUR: extent 4 sigma for
// every atom; else extent = radius of the atom (ext[k]).
template <bool BLUR>
__global__ __launch_bounds__(512) void k_shape(const double* __restrict__ pts,
                                               const double* __restrict__ wr, int64_t n,
                                               double sigma, Grid g, const int* __restrict__ box,
                                               const int64_t* __restrict__ offs,
                                               const int32_t* __restrict__ list,
                                               double* __restrict__ out) {
    constexpr int CH = 512;
    __shared__ double sp[3][CH];
    __shared__ double sw[CH];
    __shared__ int sb[6][CH];
    int b0[3], b1[3];
    brick_range(blockIdx.x, g, b0, b1);
    const int v[3] = {b0[0] + (int)(threadIdx.x & 7), b0[1] + (int)((threadIdx.x >> 3) & 7),
                      b0[2] + (int)(threadIdx.x >> 6)};
    const bool inside = v[0] <= b1[0] && v[1] <= b1[1] && v[2] <= b1[2];
    const int64_t vox = ((int64_t)v[2] * g.n[1] + v[1]) * g.n[0] + v[0];
    double acc = inside ? out[vox] : 0.0;
    bool hit = false;
    // the reference's 4 sigma, (4 sigma)^2 and 2 sigma sigma, in its order
    const double extend = 4.0 * sigma;
    const double e2_blur = extend * extend;
    const double dsig2 = 2.0 * sigma * sigma;
    const int64_t l0 = offs[blockIdx.x], l1 = offs[blockIdx.x + 1];
    for (int64_t base = l0; base < l1; base += CH) {
        const int cnt = (int)min<int64_t>(CH, l1 - base);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const int64_t k = list[base + threadIdx.x];
            for (int a = 0; a < 3; ++a) sp[a][threadIdx.x] = pts[(int64_t)a * n + k];
            sw[threadIdx.x] = wr[k];
            for (int a = 0; a < 6; ++a) sb[a][threadIdx.x] = box[k * 6 + a];
        }
        __syncthreads();
        if (inside && (BLUR || !hit)) {
            for (int j = 0; j < cnt; ++j) {
                const double p[3] = {sp[0][j], sp[1][j], sp[2][j]};
                const int bx[6] = {sb[0][j], sb[1][j], sb[2][j], sb[3][j], sb[4][j], sb[5][j]};
                const double r = sw[j];
                const double e2 = BLUR ? e2_blur : r * r;
                visit_atom<BLUR>(p, bx, v, g.n, e2, r, dsig2, acc, hit);
                if (!BLUR && hit) break;
            }
        }
    }
    if (inside) out[vox] = BLUR ? acc : (hit ? 1.0 : acc);
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t count) {
        if (count == 0) count = 1;
        if (hipMalloc((void**)&p, sizeof(T) * count) != hipSuccess) {
            p = nullptr;
            set_error("shape generation: device allocation failed");
            return kErrHip;
        }
        return kOk;
    }
};

int run_shape(bool blur, int device, const double* points, const double* wr, int64_t n,
              double sigma, double* out, int nz, int ny, int nx, int wraparound) {
    if (!out || nz <= 0 || ny <= 0 || nx <= 0) {
        set_error("shape generation: output grid is null or has a dimension below 1");
        return kErrArg;
    }
    if (n < 0 || (n > 0 && (!points || !wr))) {
        set_error("shape generation: bad atom arrays");
        return kErrArg;
    }
    if (n > INT32_MAX) {
        set_error("shape generation: at most 2^31 - 1 atoms");
        return kErrArg;
    }
    if (n == 0) return kOk;
    if (hipSetDevice(device) != hipSuccess) {
        set_error("shape generation: hipSetDevice failed");
        return kErrHip;
    }
    Grid g;
    g.n[0] = nx;
    g.n[1] = ny;
    g.n[2] = nz;
    for (int a = 0; a < 3; ++a) g.nb[a] = (g.n[a] + kBrick - 1) / kBrick;
    const int64_t nbricks = (int64_t)g.nb[0] * g.nb[1] * g.nb[2];
    if (nbricks > INT32_MAX) {
        set_error("shape generation: grid too large");
        return kErrArg;
    }
    const int64_t V = (int64_t)nx * ny * nz;
    hipStream_t s;
    if (hipStreamCreate(&s) != hipSuccess) {
        set_error("shape generation: stream creation failed");
        return kErrHip;
    }
    DevBuf<double> d_pts, d_wr, d_out;
    DevBuf<int> d_box;
    DevBuf<int64_t> d_counts, d_offs;
    DevBuf<int32_t> d_list;
    int st = kOk;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (st == kOk && e != hipSuccess) {
            set_error(what);
            st = kErrHip;
        }
    };
    if (!st) st = d_pts.alloc(3 * n);
    if (!st) st = d_wr.alloc(n);
    if (!st) st = d_out.alloc(V);
    if (!st) st = d_box.alloc(6 * n);
    if (!st) st = d_counts.alloc(nbricks);
    if (!st) st = d_offs.alloc(nbricks + 1);
    if (!st) {
        hip_ok(hipMemcpyAsync(d_pts.p, points, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s),
               "shape generation: upload failed");
        hip_ok(hipMemcpyAsync(d_wr.p, wr, sizeof(double) * n, hipMemcpyHostToDevice, s),
               "shape generation: upload failed");
        hip_ok(hipMemcpyAsync(d_out.p, out, sizeof(double) * V, hipMemcpyHostToDevice, s),
               "shape generation: upload failed");
    }
    int64_t total = 0;
    if (!st) {
        hipLaunchKernelGGL(k_atom_boxes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           d_pts.p, blur ? nullptr : d_wr.p, 4.0 * sigma, n, g, wraparound,
                           d_box.p);
        hipLaunchKernelGGL(k_brick_count, dim3((unsigned)nbricks), dim3(kListThreads), 0, s,
                           d_box.p, n, g, d_counts.p);
        hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(kScanThreads), 0, s, d_counts.p,
                           nbricks, d_offs.p);
        hip_ok(hipGetLastError(), "shape generation: list launch failed");
        hip_ok(hipMemcpyAsync(&total, d_offs.p + nbricks, sizeof(int64_t), hipMemcpyDeviceToHost,
                              s),
               "shape generation: read-back failed");
        hip_ok(hipStreamSynchronize(s), "shape generation: list kernels failed");
    }
    if (!st) st = d_list.alloc((size_t)total);
    if (!st) {
        hipLaunchKernelGGL(k_brick_fill, dim3((unsigned)nbricks), dim3(kListThreads), 0, s,
                           d_box.p, n, g, d_offs.p, d_list.p);
        if (blur)
            hipLaunchKernelGGL(k_shape<true>, dim3((unsigned)nbricks), dim3(512), 0, s, d_pts.p,
                               d_wr.p, n, sigma, g, d_box.p, d_offs.p, d_list.p, d_out.p);
        else
            hipLaunchKernelGGL(k_shape<false>, dim3((unsigned)nbricks), dim3(512), 0, s, d_pts.p,
                               d_wr.p, n, sigma, g, d_box.p, d_offs.p, d_list.p, d_out.p);
        hip_ok(hipGetLastError(), "shape generation: launch failed");
        hip_ok(hipMemcpyAsync(out, d_out.p, sizeof(double) * V, hipMemcpyDeviceToHost, s),
               "shape generation: download failed");
        hip_ok(hipStreamSynchronize(s), "shape generation: kernels failed");
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    return st;
}

}  // namespace
}  // namespace pf

extern "C" int pf_blur_points(int device, const double* points, const double* weights, int64_t n,
                              double sigma, double* out, int nz, int ny, int nx, int wraparound) {
    return pf::run_shape(true, device, points, weights, n, sigma, out, nz, ny, nx, wraparound);
}

extern "C" int pf_dilate_points(int device, const double* points, const double* radii, int64_t n,
                                double* out, int nz, int ny, int nx, int wraparound) {
    return pf::run_shape(false, device, points, radii, n, 0.0, out, nz, ny, nx, wraparound);
}
