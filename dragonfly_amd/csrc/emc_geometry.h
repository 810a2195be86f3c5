// Device geometry shared by the FP64 and mixed-precision kernels: rotated pixel position,
// trilinear cell and interpolated model value, as slice_gen3d (dragonfly/src/model.c:11-59).
#pragma once

#include <cfloat>

#include "emc_common.h"

namespace emc {

// ---- geometry ------------------------------------------------------------------------

// Rotated voxel position of a pixel and its trilinear cell, as slice_gen3d / slice_merge3d
// (dragonfly/src/model.c:19-45 and :184-207): p = R q + center, cell = trunc(p).
// Cell coordinates are 32-bit: truncation toward zero (v_cvt_i32_f64, which saturates far
// outside the grid, where the range test fails either way).
struct Cell {
    int x, y, z;
    double fx, fy, fz;
};

__device__ __forceinline__ bool locate(const double *rot, double q0, double q1, double q2,
                                       double center, long size, Cell &c) {
    double p0 = rot[0] * q0 + rot[1] * q1 + rot[2] * q2 + center;
    double p1 = rot[3] * q0 + rot[4] * q1 + rot[5] * q2 + center;
    double p2 = rot[6] * q0 + rot[7] * q1 + rot[8] * q2 + center;
    c.x = __double2int_rz(p0);
    c.y = __double2int_rz(p1);
    c.z = __double2int_rz(p2);
    const int hi = (int)size - 2;
    if (c.x < 0 || c.x > hi || c.y < 0 || c.y > hi || c.z < 0 || c.z > hi)
        return false;
    c.fx = p0 - c.x;
    c.fy = p1 - c.y;
    c.fz = p2 - c.z;
    return true;
}

// Interpolated model value times corr, DBL_MIN outside the grid (model.c:35-57).
__device__ __forceinline__ double slice_value(const double *__restrict__ model, long size,
                                              const double *rot, double q0, double q1, double q2,
                                              double corr, double center) {
    Cell c;
    if (!locate(rot, q0, q1, q2, center, size, c)) return DBL_MIN;
    const long s2 = size * size;
    const long x1 = (c.x + 1) % size, y1 = (c.y + 1) % size, z1 = (c.z + 1) % size;
    const long cx0 = c.x, cy0 = c.y, cz0 = c.z;
    const double cx = 1. - c.fx, cy = 1. - c.fy, cz = 1. - c.fz;
    const double fx = c.fx, fy = c.fy, fz = c.fz;
    double v = cx * cy * cz * model[cx0 * s2 + cy0 * size + cz0] +
               cx * cy * fz * model[cx0 * s2 + cy0 * size + z1] +
               cx * fy * cz * model[cx0 * s2 + y1 * size + cz0] +
               cx * fy * fz * model[cx0 * s2 + y1 * size + z1] +
               fx * cy * cz * model[x1 * s2 + cy0 * size + cz0] +
               fx * cy * fz * model[x1 * s2 + cy0 * size + z1] +
               fx * fy * cz * model[x1 * s2 + y1 * size + cz0] +
               fx * fy * fz * model[x1 * s2 + y1 * size + z1];
    return v * corr;
}

// slice_value without a branch around the loads: a pixel outside the grid reads the corners of
// cell (0, 0, 0) and returns DBL_MIN, so the eight loads of several calls can be in flight at
// once. The value is the same as slice_value's.
__device__ __forceinline__ double slice_value_flat(const double *__restrict__ model, long size,
                                                   const double *rot, double q0, double q1, double q2,
                                                   double corr, double center) {
    Cell c{0, 0, 0, 0., 0., 0.};
    const bool ok = locate(rot, q0, q1, q2, center, size, c);
    const long s2 = size * size;
    const long cx0 = ok ? c.x : 0, cy0 = ok ? c.y : 0, cz0 = ok ? c.z : 0;
    const long x1 = (cx0 + 1) % size, y1 = (cy0 + 1) % size, z1 = (cz0 + 1) % size;
    const double cx = 1. - c.fx, cy = 1. - c.fy, cz = 1. - c.fz;
    const double fx = c.fx, fy = c.fy, fz = c.fz;
    const double v = cx * cy * cz * model[cx0 * s2 + cy0 * size + cz0] +
                     cx * cy * fz * model[cx0 * s2 + cy0 * size + z1] +
                     cx * fy * cz * model[cx0 * s2 + y1 * size + cz0] +
                     cx * fy * fz * model[cx0 * s2 + y1 * size + z1] +
                     fx * cy * cz * model[x1 * s2 + cy0 * size + cz0] +
                     fx * cy * fz * model[x1 * s2 + cy0 * size + z1] +
                     fx * fy * cz * model[x1 * s2 + y1 * size + cz0] +
                     fx * fy * fz * model[x1 * s2 + y1 * size + z1];
    return ok ? v * corr : DBL_MIN;
}

// FP32 model copy in 4 x 4 x 4 voxel bricks: brick (X/4, Y/4, Z/4) of an nb^3 grid of bricks
// (nb = ceil(size / 4)) holds its 64 voxels contiguously, z fastest. A 128-byte cache line is
// then a 2 x 4 x 4 block of voxels instead of 32 voxels along one z row, so the eight corners of
// the cells a wavefront interpolates (a compact detector patch rotated into the volume, at any
// orientation) fall in fewer lines.
constexpr int kModelBrick = 4;
__host__ __device__ __forceinline__ int model_bricks(long size) { return (int)((size + kModelBrick - 1) / kModelBrick); }
__device__ __forceinline__ long brick_index(int nb, int x, int y, int z) {
    return ((((long)(x >> 2) * nb + (y >> 2)) * nb + (z >> 2)) << 6) + (((x & 3) << 4) | ((y & 3) << 2) | (z & 3));
}

// slice_value_flat on the FP32 brick copy: the rotated position, cell and fractions in FP64 (the
// same cell as the FP64 path), the eight corner weights and the interpolation in FP32. DBL_MIN
// outside the grid.
__device__ __forceinline__ double slice_value_f32b(const float *__restrict__ model, long size,
                                                   const double *rot, double q0, double q1, double q2,
                                                   double corr, double center) {
    Cell c{0, 0, 0, 0., 0., 0.};
    const bool ok = locate(rot, q0, q1, q2, center, size, c);
    const int nb = model_bricks(size);
    const int x0 = ok ? c.x : 0, y0 = ok ? c.y : 0, z0 = ok ? c.z : 0;
    // in the grid x0, y0, z0 <= size - 2, so the + 1 corners need no wrap
    const float fx = (float)c.fx, fy = (float)c.fy, fz = (float)c.fz;
    const float cx = 1.f - fx, cy = 1.f - fy, cz = 1.f - fz;
    const float m000 = model[brick_index(nb, x0, y0, z0)], m001 = model[brick_index(nb, x0, y0, z0 + 1)];
    const float m010 = model[brick_index(nb, x0, y0 + 1, z0)], m011 = model[brick_index(nb, x0, y0 + 1, z0 + 1)];
    const float m100 = model[brick_index(nb, x0 + 1, y0, z0)], m101 = model[brick_index(nb, x0 + 1, y0, z0 + 1)];
    const float m110 = model[brick_index(nb, x0 + 1, y0 + 1, z0)], m111 = model[brick_index(nb, x0 + 1, y0 + 1, z0 + 1)];
    const float v = cx * (cy * (cz * m000 + fz * m001) + fy * (cz * m010 + fz * m011)) +
                    fx * (cy * (cz * m100 + fz * m101) + fy * (cz * m110 + fz * m111));
    return ok ? (double)v * corr : DBL_MIN;
}

}  // namespace emc
