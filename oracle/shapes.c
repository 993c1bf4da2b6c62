/*
 * ORACLE — test infrastructure only, never part of the product path.
 *
 * CPU restatement of PowerFit's template and mask generators
 * (src/powerfit_em/_powerfit.pyx):
 *
 *   blur_points   (:75-140)  out[z,y,x] += w_n exp(-d^2 / (2 sigma^2)) for
 *                            every atom n and voxel (x, y, z) of its box with
 *                            d^2 <= (4 sigma)^2;
 *   dilate_points (:143-206) out[z,y,x] = 1 for every voxel of the box of
 *                            atom n with d^2 <= radius_n^2.
 *
 * Box of an atom at p with extent e (4 sigma, or its radius):
 * [ceil(p - e), floor(p + e)] per axis, clipped below at 0 — or at -dim + 1
 * with wraparound, where a negative index then addresses index + dim (Python
 * indexing) — and clipped above at dim - 1 (no wrap above).
 * Atoms are visited in order, the box in z, y, x loop order; d^2 is formed
 * as ((x - px)^2 + ((y - py)^2 + (z - pz)^2)), FP64, no contraction.
 * points is (3, n) row-major: x row, y row, z row (grid units).
 */
#include <math.h>
#include <stdint.h>

#pragma STDC FP_CONTRACT OFF

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

typedef struct {
    int lo[3], hi[3];  /* inclusive, axis order x, y, z */
} Box;

static Box atom_box(const double p[3], double e, const int dim[3], int wraparound) {
    Box b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = imax((int)ceil(p[a] - e), wraparound ? -dim[a] + 1 : 0);
        b.hi[a] = imin((int)floor(p[a] + e), dim[a] - 1);
    }
    return b;
}

void oracle_blur_points(const double* points, const double* weights, int64_t n, double sigma,
                        double* out, int nz, int ny, int nx, int wraparound) {
    const double extend = 4.0 * sigma;
    const double extend2 = extend * extend;
    const double dsigma2 = 2.0 * sigma * sigma;
    const int dim[3] = {nx, ny, nz};
    for (int64_t k = 0; k < n; ++k) {
        const double p[3] = {points[k], points[n + k], points[2 * n + k]};
        const Box b = atom_box(p, extend, dim, wraparound);
        for (int z = b.lo[2]; z <= b.hi[2]; ++z) {
            const double dz = (double)z - p[2];
            const double z2 = dz * dz;
            const int64_t oz = (int64_t)(z < 0 ? z + nz : z) * ny;
            for (int y = b.lo[1]; y <= b.hi[1]; ++y) {
                const double dy = (double)y - p[1];
                const double y2z2 = dy * dy + z2;
                const int64_t oy = (oz + (y < 0 ? y + ny : y)) * nx;
                for (int x = b.lo[0]; x <= b.hi[0]; ++x) {
                    const double dx = (double)x - p[0];
                    const double d2 = dx * dx + y2z2;
                    if (d2 <= extend2) out[oy + (x < 0 ? x + nx : x)] += weights[k] * exp(-d2 / dsigma2);
                }
            }
        }
    }
}

void oracle_dilate_points(const double* points, const double* radii, int64_t n, double* out,
                          int nz, int ny, int nx, int wraparound) {
    const int dim[3] = {nx, ny, nz};
    for (int64_t k = 0; k < n; ++k) {
        const double p[3] = {points[k], points[n + k], points[2 * n + k]};
        const double radius = radii[k];
        const double radius2 = radius * radius;
        const Box b = atom_box(p, radius, dim, wraparound);
        for (int z = b.lo[2]; z <= b.hi[2]; ++z) {
            const double dz = (double)z - p[2];
            const double z2 = dz * dz;
            const int64_t oz = (int64_t)(z < 0 ? z + nz : z) * ny;
            for (int y = b.lo[1]; y <= b.hi[1]; ++y) {
                const double dy = (double)y - p[1];
                const double y2z2 = dy * dy + z2;
                const int64_t oy = (oz + (y < 0 ? y + ny : y)) * nx;
                for (int x = b.lo[0]; x <= b.hi[0]; ++x) {
                    const double dx = (double)x - p[0];
                    if (dx * dx + y2z2 <= radius2) out[oy + (x < 0 ? x + nx : x)] = 1.0;
                }
            }
        }
    }
}
