// Expand: log W of one 64-orientation tile over one pixel group, plus the tile's partial u_r
// and "log W not finite on a mask < 1 pixel" flags. Shared by expand_kernel (one block per
// (tile, group)) and the persistent expand + likelihood kernel of the refinement subsets.
#pragma once

#include <cfloat>
#include <type_traits>

#include "emc_common.h"
#include "emc_geometry.h"

namespace emc {

// Lanes along pixels (neighbouring pixels hit neighbouring voxels), transpose through LDS,
// store with lanes along orientations (coalesced rows of logW). u_r (max_prob.c:36-38) sums W
// over mask < 1. The FP64 model is interpolated as slice_gen3d does (slice_value_flat); the
// FP32 copy (halo bricks, copy_f32_brick_kernel) in FP32 from an FP32 rotation
// (slice_value_f32h).
// elements of one mode's volume: size^3 for the FP64 model, the brick grid for the FP32 copy
__device__ __forceinline__ size_t mode_stride(const double *, long size) { return (size_t)(size * size * size); }
__device__ __forceinline__ size_t mode_stride(const float *, long size) {
    const size_t nb = (size_t)model_bricks(size);
    return nb * nb * nb * kHaloBrickFloats;
}
// log(DBL_MIN) = -1022 ln 2: the stored log W of a pixel outside the grid (FP32 model copy)
constexpr float kLogDblMin = (float)(-1022.0 * 0.693147180559945309417232121458);

// LDS of one expand block. With the FP32 model copy (MT float) the rotations are FP32 and the
// transposition tile holds W as FP32 (W = DBL_MIN outside the grid as -inf): 21 KB instead of
// 43 KB a block, so that twice as many blocks share a CU.
template <typename MT>
struct ExpandShared {
    static constexpr bool f32 = sizeof(MT) == sizeof(float);
    using V = typename std::conditional<f32, float, double>::type;
    V rot[kRTile * 9];
    V tile[64][kRTile + 1];
    double u[kWavesPerBlock][kRTile];
    int bad[kWavesPerBlock][kRTile];
    int pad[kRTile];
};

// Tile starting at chunk column c0, pixel group g; called by all kBlock threads of a block.
// MT: element type of the model read (double, or float for the FP32-gather path's copy)
template <typename T, int kW, typename MT>
__device__ __forceinline__ void expand_block(
    ExpandShared<MT> &sh, int c0, int g, const double *__restrict__ rots, const MT *__restrict__ model, long size,
    const double *__restrict__ qvals, const double *__restrict__ corr, const uint8_t *__restrict__ mask, int num_pix,
    int r0, int C, int pix_per_group, T *__restrict__ logw, double *__restrict__ upart, int *__restrict__ ubad,
    const int *__restrict__ slot_rot, int modes, int raw) {
    constexpr bool f32 = sizeof(MT) == sizeof(float);
    const int tid = threadIdx.x;
    for (int i = tid; i < kRTile * 9; i += kBlock) sh.rot[i] = (typename ExpandShared<MT>::V)rots[(size_t)(r0 + c0) * 9 + i];
    // padding slots of the subset order (slot_rot < 0) get W = 1 without interpolation
    if (tid < kRTile) sh.pad[tid] = slot_rot != nullptr && slot_rot[r0 + c0 + tid] < 0;
    __syncthreads();
    const double center = (double)(size / 2);
    const int nb = model_bricks(size);
    const int lane = tid & 63, wave = tid >> 6;
    const int t_begin = g * pix_per_group;
    const int t_end = min(num_pix, t_begin + pix_per_group);
    double uacc = 0.;
    int bad = 0;  // some mask < 1 pixel has W <= 0 (log W = -inf or NaN)
    for (int tb = t_begin; tb < t_end; tb += 64) {
        const int t = tb + lane;
        if (t < t_end) {
            const double q0 = qvals[3 * (size_t)t], q1 = qvals[3 * (size_t)t + 1],
                         q2 = qvals[3 * (size_t)t + 2], cr = corr[t];
            const float qf0 = (float)q0, qf1 = (float)q1, qf2 = (float)q2, crf = (float)cr;
#pragma unroll 4
            for (int k = 0; k < kRTile / kWavesPerBlock; ++k) {
                const int rr = wave + kWavesPerBlock * k;
                // combined orientation index r = rotation * modes + mode (max_prob.c:80-84)
                const MT *mdl = modes > 1 ? model + (size_t)((r0 + c0 + rr) % modes) * mode_stride(model, size) : model;
                if constexpr (f32) {
                    const float v =
                        slice_value_f32h(mdl, (int)size, nb, &sh.rot[rr * 9], qf0, qf1, qf2, crf, (float)center);
                    sh.tile[lane][rr] = sh.pad[rr] ? 1.f : v;
                } else {
                    const double v = slice_value_flat(mdl, size, &sh.rot[rr * 9], q0, q1, q2, cr, center);
                    sh.tile[lane][rr] = sh.pad[rr] ? 1. : v;
                }
            }
        }
        __syncthreads();
        for (int k = 0; k < 64 / kWavesPerBlock; ++k) {
            const int p = wave + kWavesPerBlock * k;
            const int tt = tb + p;
            if (tt < t_end) {
                // tile-major with kW-wide tiles: (c / kW) * num_pix * kW + t * kW + c % kW
                const int c = c0 + lane;
                T *dst = logw + (size_t)(c / kW) * ((size_t)num_pix * kW) + (size_t)tt * kW + (c % kW);
                if constexpr (f32) {
                    // FP32 tile: -inf outside the grid (W = DBL_MIN there: not bad, log W = log(DBL_MIN))
                    const float Wt = sh.tile[p][lane];
                    const bool outside = Wt == -INFINITY;
                    if (mask[tt] < 1) {
                        uacc += outside ? DBL_MIN : (double)Wt;
                        bad |= !outside && !(Wt > 0.f);
                    }
                    *dst = (T)(outside ? kLogDblMin : (Wt >= FLT_MIN ? __logf(Wt) : logf(Wt)));
                } else {
                    const double W = sh.tile[p][lane];
                    if (mask[tt] < 1) {
                        uacc += W;
                        bad |= !(W > 0.);
                    }
                    // raw: W itself (the background-aware likelihood takes log(phi W + b B))
                    *dst = raw ? (T)W : (T)log(W);
                }
            }
        }
        __syncthreads();
    }
    sh.u[wave][lane] = uacc;
    sh.bad[wave][lane] = bad;
    __syncthreads();
    if (tid < kRTile) {
        upart[(size_t)g * C + c0 + tid] = sh.u[0][tid] + sh.u[1][tid] + sh.u[2][tid] + sh.u[3][tid];
        ubad[(size_t)g * C + c0 + tid] = sh.bad[0][tid] | sh.bad[1][tid] | sh.bad[2][tid] | sh.bad[3][tid];
    }
    __syncthreads();  // the LDS is free again for the block's next piece of work
}

}  // namespace emc
