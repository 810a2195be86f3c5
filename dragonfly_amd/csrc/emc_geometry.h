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

// FP32 model copy in bricks of 4 x 4 x 4 trilinear cells: brick (X/4, Y/4, Z/4) of an nb^3
// grid (nb = (size - 2) / 4 + 1, the cells x <= size - 2 of the grid) holds the 5 x 5 x 5
// corners of its cells (shared faces are stored twice), padded to 128 floats, z fastest. The
// eight corners of any cell are then at fixed offsets {0, 1, 5, 6, 25, 26, 30, 31} from one
// address, all within the brick's 512 bytes: one index computation per interpolation, and a
// compact detector patch rotated into the volume at any orientation touches few cache lines.
constexpr int kModelBrick = 4;
constexpr int kHaloBrickFloats = 128;  // 5^3 = 125 corners, padded
__host__ __device__ __forceinline__ int model_bricks(long size) { return (int)((size - 2) / kModelBrick + 1); }
__device__ __forceinline__ long halo_index(int nb, int x, int y, int z) {
    return (((long)(x >> 2) * nb + (y >> 2)) * nb + (z >> 2)) * kHaloBrickFloats + (x & 3) * 25 + (y & 3) * 5 + (z & 3);
}

// Interpolated value of the FP32 brick copy at the rotated position of a pixel times the
// pixel's corr, all in FP32: the position R q + center, the truncated cell (toward zero, as
// model.c:31-33) and its fractions, then the trilinear value as three rounds of linear
// interpolation (z, y, x), each a fused multiply-add a + f (b - a). rot: the rotation as 9
// floats; q: the pixel's q-vector in FP32. -inf outside the grid (x, y or z outside
// [0, size - 2]), where the reference takes DBL_MIN.
__device__ __forceinline__ float slice_value_f32h(const float *__restrict__ model, int size, int nb, const float *rot,
                                                  float q0, float q1, float q2, float corr, float center) {
    const float p0 = __builtin_fmaf(rot[0], q0, __builtin_fmaf(rot[1], q1, __builtin_fmaf(rot[2], q2, center)));
    const float p1 = __builtin_fmaf(rot[3], q0, __builtin_fmaf(rot[4], q1, __builtin_fmaf(rot[5], q2, center)));
    const float p2 = __builtin_fmaf(rot[6], q0, __builtin_fmaf(rot[7], q1, __builtin_fmaf(rot[8], q2, center)));
    const int x = (int)p0, y = (int)p1, z = (int)p2;  // truncation toward zero
    const int hi = size - 2;
    const bool ok = x >= 0 && x <= hi && y >= 0 && y <= hi && z >= 0 && z <= hi;
    const float fx = p0 - (float)x, fy = p1 - (float)y, fz = p2 - (float)z;
    const float *b = model + (ok ? halo_index(nb, x, y, z) : 0);
    const float m000 = b[0], m001 = b[1], m010 = b[5], m011 = b[6];
    const float m100 = b[25], m101 = b[26], m110 = b[30], m111 = b[31];
    const float a00 = __builtin_fmaf(fz, m001 - m000, m000), a01 = __builtin_fmaf(fz, m011 - m010, m010);
    const float a10 = __builtin_fmaf(fz, m101 - m100, m100), a11 = __builtin_fmaf(fz, m111 - m110, m110);
    const float b0 = __builtin_fmaf(fy, a01 - a00, a00), b1 = __builtin_fmaf(fy, a11 - a10, a10);
    const float v = __builtin_fmaf(fx, b1 - b0, b0);
    return ok ? v * corr : -INFINITY;
}

}  // namespace emc
