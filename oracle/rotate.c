/*
 * ORACLE — test infrastructure only, never part of the product path.
 *
 * CPU restatement of PowerFit's `rotate_grid3d`
 * (src/powerfit_em/_extensions.c:7-196): rotate a grid about voxel (0,0,0)
 * inside the sphere |v| <= radius, gathering grid(R^T v) for every integer
 * v = (x, y, z) in that sphere, with periodic output indices.
 *
 * Semantics restated, in the reference's own evaluation order:
 *   - loops z, y, x over [-radius, radius], skipped when the running integer
 *     sum of squares exceeds radius^2 (_extensions.c:55-80);
 *   - source coordinate  xs = (R[6] z + R[3] y) + R[0] x  (and the same with
 *     R[7],R[4],R[1] / R[8],R[5],R[2]) in double precision, no contraction
 *     (_extensions.c:60-62,74-76,87-89);
 *   - output index wraps negative coordinates by one "+dim" (:69-71,81-83,91-93);
 *   - nearest: (int) round() — half away from zero — with one "+dim" for
 *     negative indices (:95-110);
 *   - trilinear: floor corner, one "+dim" for negative corner indices, the
 *     upper neighbour wraps only when it equals 0 (:113-178);
 *   - voxels outside the sphere are never written (caller zero-fills).
 *
 * The source index arithmetic is kept exactly as the reference's (including
 * its wrap-once rules). Indices are additionally bounds-checked here so a
 * misuse reads 0 instead of faulting; for radius <= min(shape) - 2 no check
 * ever fires.
 */
#include <math.h>
#include <stdint.h>

#pragma STDC FP_CONTRACT OFF

static inline double rd(const double *g, int64_t i, int64_t size) {
    return (i >= 0 && i < size) ? g[i] : 0.0;
}

void oracle_rotate_grid3d(const double *grid, int gz, int gy, int gx,
                          const double *rotmat, int radius,
                          double *out, int oz, int oy, int ox, int nearest)
{
    const int64_t out_slice = (int64_t)oy * ox, out_size = out_slice * oz;
    const int64_t grid_slice = (int64_t)gy * gx, grid_size = grid_slice * gz;
    const int radius2 = radius * radius;
    for (int z = -radius; z <= radius; z++) {
        int d2z = z * z;
        if (d2z > radius2) continue;
        double xz = rotmat[6] * z, yz = rotmat[7] * z, zz = rotmat[8] * z;
        int64_t oz_ = (int64_t)z * out_slice;
        if (z < 0) oz_ += out_size;
        for (int y = -radius; y <= radius; y++) {
            int d2zy = d2z + y * y;
            if (d2zy > radius2) continue;
            double xzy = xz + rotmat[3] * y;
            double yzy = yz + rotmat[4] * y;
            double zzy = zz + rotmat[5] * y;
            int64_t ozy = oz_ + (int64_t)y * ox;
            if (y < 0) ozy += out_slice;
            for (int x = -radius; x <= radius; x++) {
                int d2 = d2zy + x * x;
                if (d2 > radius2) continue;
                double xs = xzy + rotmat[0] * x;
                double ys = yzy + rotmat[1] * x;
                double zs = zzy + rotmat[2] * x;
                int64_t o = ozy + x;
                if (x < 0) o += ox;
                if (o < 0 || o >= out_size) continue;
                if (nearest > 0) {
                    int x0 = (int)round(xs), y0 = (int)round(ys), z0 = (int)round(zs);
                    int64_t g = (int64_t)z0 * grid_slice + (int64_t)y0 * gx + x0;
                    if (x0 < 0) g += gx;
                    if (y0 < 0) g += grid_slice;
                    if (z0 < 0) g += grid_size;
                    out[o] = rd(grid, g, grid_size);
                } else {
                    int x0 = (int)floor(xs), y0 = (int)floor(ys), z0 = (int)floor(zs);
                    int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
                    int64_t g = (int64_t)z0 * grid_slice + (int64_t)y0 * gx + x0;
                    if (x0 < 0) g += gx;
                    if (y0 < 0) g += grid_slice;
                    if (z0 < 0) g += grid_size;
                    double dx = xs - x0, dy = ys - y0, dz = zs - z0;
                    double dx1 = 1 - dx, dy1 = 1 - dy, dz1 = 1 - dz;
                    int64_t off0, off1 = 1;
                    if (x1 == 0) off1 -= gx;
                    double c00 = rd(grid, g, grid_size) * dx1 + rd(grid, g + off1, grid_size) * dx;
                    off0 = gx;
                    if (y1 == 0) off0 -= grid_slice;
                    off1 = off0 + 1;
                    if (x1 == 0) off1 -= gx;
                    double c10 = rd(grid, g + off0, grid_size) * dx1 + rd(grid, g + off1, grid_size) * dx;
                    off0 = grid_slice;
                    if (z1 == 0) off0 -= grid_size;
                    off1 = off0 + 1;
                    if (x1 == 0) off1 -= gx;
                    double c01 = rd(grid, g + off0, grid_size) * dx1 + rd(grid, g + off1, grid_size) * dx;
                    off0 = grid_slice + gx;
                    if (z1 == 0) off0 -= grid_size;
                    if (y1 == 0) off0 -= grid_slice;
                    off1 = off0 + 1;
                    if (x1 == 0) off1 -= gx;
                    double c11 = rd(grid, g + off0, grid_size) * dx1 + rd(grid, g + off1, grid_size) * dx;
                    double c0 = c00 * dy1 + c10 * dy;
                    double c1 = c01 * dy1 + c11 * dy;
                    out[o] = c0 * dz1 + c1 * dz;
                }
            }
        }
    }
}
