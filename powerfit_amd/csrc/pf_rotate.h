// Rotation sampling shared by the generic and fused kernels
// (reference: src/powerfit_em/_extensions.c:7-196).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace pf {

// Rotation (reference: _extensions.c:7-196).
//
// Gather form of the reference's scatter loop: output voxel (iz, iy, ix) takes
// the value written last by the reference, i.e. from the lexicographically
// largest signed triple (z, y, x), each congruent to its index and inside the
// sphere |v|^2 <= r^2.  Source coordinates are evaluated in FP64 in the
// reference's order without contraction; index arithmetic (floor/round, one
// "+dim" wrap for negatives, the upper trilinear neighbour wrapping only at 0)
// is the reference's.  Out-of-range source indices read 0.

__device__ __forceinline__ int first_cand(int i, int n, int r) {
    // largest s <= r with s == i (mod n)
    int q = r - i;
    int fl = q >= 0 ? q / n : -((-q + n - 1) / n);
    return i + fl * n;
}

// Returns false when no in-sphere triple maps to this voxel.
__device__ __forceinline__ bool signed_triple(int iz, int iy, int ix, int nz, int ny, int nx,
                                              int r, int* z, int* y, int* x) {
    const int r2 = r * r;
    for (int sz = first_cand(iz, nz, r); sz >= -r; sz -= nz) {
        const int d2z = sz * sz;
        if (d2z > r2) continue;
        for (int sy = first_cand(iy, ny, r); sy >= -r; sy -= ny) {
            const int d2zy = d2z + sy * sy;
            if (d2zy > r2) continue;
            for (int sx = first_cand(ix, nx, r); sx >= -r; sx -= nx) {
                if (d2zy + sx * sx <= r2) {
                    *z = sz;
                    *y = sy;
                    *x = sx;
                    return true;
                }
            }
        }
    }
    return false;
}

// Element i of a grid stored with element stride ST (ST = 2: one member of
// an interleaved pair); out-of-range indices read 0.
template <int ST = 1, typename T>
__device__ __forceinline__ double rd(const T* g, int64_t i, int64_t size) {
    return (i >= 0 && i < size) ? (double)g[i * ST] : 0.0;
}

struct Coord {
    double xs, ys, zs;
};

__device__ __forceinline__ Coord source_coord(const double* R, int z, int y, int x) {
#pragma clang fp contract(off)
    Coord c;
    const double xz = R[6] * z, yz = R[7] * z, zz = R[8] * z;
    const double xzy = xz + R[3] * y, yzy = yz + R[4] * y, zzy = zz + R[5] * y;
    c.xs = xzy + R[0] * x;
    c.ys = yzy + R[1] * x;
    c.zs = zzy + R[2] * x;
    return c;
}

template <int ST = 1, typename T>
__device__ __forceinline__ double sample_nearest(const T* g, int gz, int gy, int gx,
                                                 const Coord& c) {
    const int x0 = (int)round(c.xs), y0 = (int)round(c.ys), z0 = (int)round(c.zs);
    const int64_t gslice = (int64_t)gy * gx, gsize = gslice * gz;
    int64_t i = (int64_t)z0 * gslice + (int64_t)y0 * gx + x0;
    if (x0 < 0) i += gx;
    if (y0 < 0) i += gslice;
    if (z0 < 0) i += gsize;
    return rd<ST>(g, i, gsize);
}

template <int ST = 1, typename T>
__device__ __forceinline__ double sample_trilinear(const T* g, int gz, int gy, int gx,
                                                   const Coord& c) {
#pragma clang fp contract(off)
    const int x0 = (int)floor(c.xs), y0 = (int)floor(c.ys), z0 = (int)floor(c.zs);
    const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const int64_t gslice = (int64_t)gy * gx, gsize = gslice * gz;
    int64_t i = (int64_t)z0 * gslice + (int64_t)y0 * gx + x0;
    if (x0 < 0) i += gx;
    if (y0 < 0) i += gslice;
    if (z0 < 0) i += gsize;
    const double dx = c.xs - x0, dy = c.ys - y0, dz = c.zs - z0;
    const double dx1 = 1 - dx, dy1 = 1 - dy, dz1 = 1 - dz;
    const int64_t wx = (x1 == 0) ? -(int64_t)gx : 0;
    const int64_t wy = (y1 == 0) ? -gslice : 0;
    const int64_t wz = (z1 == 0) ? -gsize : 0;
    const double c00 = rd<ST>(g, i, gsize) * dx1 + rd<ST>(g, i + 1 + wx, gsize) * dx;
    const int64_t o10 = gx + wy;
    const double c10 = rd<ST>(g, i + o10, gsize) * dx1 + rd<ST>(g, i + o10 + 1 + wx, gsize) * dx;
    const int64_t o01 = gslice + wz;
    const double c01 = rd<ST>(g, i + o01, gsize) * dx1 + rd<ST>(g, i + o01 + 1 + wx, gsize) * dx;
    const int64_t o11 = gslice + gx + wz + wy;
    const double c11 = rd<ST>(g, i + o11, gsize) * dx1 + rd<ST>(g, i + o11 + 1 + wx, gsize) * dx;
    const double c0 = c00 * dy1 + c10 * dy;
    const double c1 = c01 * dy1 + c11 * dy;
    return c0 * dz1 + c1 * dz;
}


// Order-preserving 64-bit key of (LCC, rotation index): larger LCC first,
// then lower index.  atomicMax over keys reproduces the reference's
// sequential strict '>' take-best from (0, 0) in any merge order.
__host__ __device__ __forceinline__ unsigned long long lcc_key(float s, uint32_t idx) {
    uint32_t u = __builtin_bit_cast(uint32_t, s);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((unsigned long long)u << 32) | (unsigned long long)(0xFFFFFFFFu - idx);
}

__host__ __device__ __forceinline__ void lcc_unkey(unsigned long long k, float* s, int32_t* idx) {
    uint32_t u = (uint32_t)(k >> 32);
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    *s = __builtin_bit_cast(float, u);
    *idx = (int32_t)(0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFull));
}

}  // namespace pf
