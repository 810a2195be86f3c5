// Brick merge: slice_merge3d of every orientation of a chunk (merge_tomogram,
// dragonfly/src/max_update.c:18-40; model.c:171-244) accumulated on chip.
//
// The volume is cut into bricks of kBrick^3 voxels. A work item is (orientation c of the
// chunk, patch p of 64 consecutive engine pixels); since engine pixels follow a Morton curve,
// a patch is a compact detector tile and its trilinear cells for one orientation fill a small
// box that meets 1-8 bricks. Which bricks an item meets depends only on the rotations and the
// detector, so the per-brick item lists are built once (brick_items_kernel, two passes:
// count, then fill) and reused by every iteration. Each list entry carries the 64-bit mask of
// the item's pixels that have a trilinear corner inside the brick: a patch's bounding box meets
// about four 16^3 bricks, its pixels' corners about 1.3 each, so a block loads and interpolates
// only the pixels it adds to. Each block then owns one brick (or a
// range of a crowded brick's items), adds every corner contribution that falls inside the
// brick into an LDS copy of it, and adds the brick to HBM once: one global atomic per touched
// voxel per block instead of one per (orientation, pixel, corner).
//
// Values are those of the reference's merge (same cell, same corner weights, w = view / corr
// for mask <= 1 pixels of orientations with psum_r > 0); only the summation order differs.
// The view is orientation-major: view[c][t], c the chunk-local orientation.
//
// With several modes (combined orientation r = rotation * modes + mode, max_prob.c:80-84) each
// mode's volume has its own bricks: brick id = mode * nbx^3 + brick, and a block adds into
// the mode's volume at mode * size^3.
#include <climits>

#include "emc_common.h"
#include "emc_geometry.h"
#include "emc_kernels.h"

namespace emc {

namespace {

__device__ __forceinline__ int wmin(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ int wmax(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

}  // namespace

// Pass 0 counts items per brick (counts[b] += 1); pass 1 writes item ids into
// list[off[b] + fill[b]++]. One wavefront per item (waves stride over the items); the wave's
// lanes are the patch's pixels.
__global__ __launch_bounds__(kBlock) void brick_items_kernel(
    const double *__restrict__ rots, int r0, int Cc, int num_rot, const double *__restrict__ qvals,
    const uint8_t *__restrict__ mask, int num_pix, long size, int npatch, int nbx, int pass,
    int *__restrict__ counts, const long *__restrict__ off, int *__restrict__ fill,
    int *__restrict__ list, unsigned long long *__restrict__ lmask, const int *__restrict__ slot_rot, int modes) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nb3 = nbx * nbx * nbx;
    const long items = (long)Cc * npatch, stride = (long)gridDim.x * kWavesPerBlock;
    for (long item = (long)blockIdx.x * kWavesPerBlock + wave; item < items; item += stride) {
        const int c = (int)(item / npatch), p = (int)(item % npatch);
        const int r = r0 + c;
        if (r >= num_rot) continue;
        if (slot_rot && slot_rot[r] < 0) continue;  // padding slot of the sparse-subset order
        const int t = p * 64 + lane;
        Cell cell;
        bool ok = false;
        if (t < num_pix && mask[t] <= 1)
            ok = locate(rots + 9 * (size_t)r, qvals[3 * (size_t)t], qvals[3 * (size_t)t + 1],
                        qvals[3 * (size_t)t + 2], (double)(size / 2), size, cell);
        const int xmax = wmax(ok ? (int)cell.x : -1);
        if (xmax < 0) continue;
        const int xmin = wmin(ok ? (int)cell.x : INT_MAX);
        const int ymin = wmin(ok ? (int)cell.y : INT_MAX), ymax = wmax(ok ? (int)cell.y : -1);
        const int zmin = wmin(ok ? (int)cell.z : INT_MAX), zmax = wmax(ok ? (int)cell.z : -1);
        // corners reach cell + 1 (never past size - 1: cells stop at size - 2)
        const int bx0 = xmin / kBrick, bx1 = (xmax + 1) / kBrick;
        const int by0 = ymin / kBrick, by1 = (ymax + 1) / kBrick;
        const int bz0 = zmin / kBrick, bz1 = (zmax + 1) / kBrick;
        const int ny = by1 - by0 + 1, nz = bz1 - bz0 + 1;
        const int nb = (bx1 - bx0 + 1) * ny * nz;
        const int mbase = modes > 1 ? (r % modes) * nb3 : 0;
        // the lane's corners span cells x .. x + 1 (likewise y, z): it meets the bricks from
        // x / kBrick to (x + 1) / kBrick along each axis
        const int cx0 = ok ? (int)cell.x / kBrick : INT_MAX, cx1 = ok ? ((int)cell.x + 1) / kBrick : -1;
        const int cy0 = ok ? (int)cell.y / kBrick : INT_MAX, cy1 = ok ? ((int)cell.y + 1) / kBrick : -1;
        const int cz0 = ok ? (int)cell.z / kBrick : INT_MAX, cz1 = ok ? ((int)cell.z + 1) / kBrick : -1;
        for (int i = 0; i < nb; ++i) {
            const int bx = bx0 + i / (ny * nz), by = by0 + (i / nz) % ny, bz = bz0 + i % nz;
            const bool in = bx >= cx0 && bx <= cx1 && by >= cy0 && by <= cy1 && bz >= cz0 && bz <= cz1;
            const unsigned long long m = __ballot(in);
            if (m == 0 || lane != 0) continue;
            const int b = mbase + (bx * nbx + by) * nbx + bz;
            if (pass == 0) {
                atomicAdd(&counts[b], 1);
            } else {
                const long at = off[b] + atomicAdd(&fill[b], 1);
                list[at] = (int)item;
                lmask[at] = m;
            }
        }
    }
}

// Adds the eight corner contributions of one (orientation, pixel) that fall inside the brick
// at (X0, Y0, Z0) into the LDS copies.
template <bool kModel, bool kWeight, typename A = double>
__device__ __forceinline__ void brick_add(const double *rot, double q0, double q1, double q2, double center,
                                          long size, int X0, int Y0, int Z0, double w, A *s_m, A *s_w) {
    Cell cell;
    if (!locate(rot, q0, q1, q2, center, size, cell)) return;
    const int lx = (int)cell.x - X0, ly = (int)cell.y - Y0, lz = (int)cell.z - Z0;
    if (lx < -1 || lx >= kBrick || ly < -1 || ly >= kBrick || lz < -1 || lz >= kBrick) return;
    const double cx = 1. - cell.fx, cy = 1. - cell.fy, cz = 1. - cell.fz;
    const double fx = cell.fx, fy = cell.fy, fz = cell.fz;
    const double f[8] = {cx * cy * cz, cx * cy * fz, cx * fy * cz, cx * fy * fz,
                         fx * cy * cz, fx * cy * fz, fx * fy * cz, fx * fy * fz};
    const int l0 = (lx * kBrick + ly) * kBrick + lz;
    if (lx >= 0 && lx < kBrick - 1 && ly >= 0 && ly < kBrick - 1 && lz >= 0 && lz < kBrick - 1) {
        // all eight corners inside the brick (the common case)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int li = l0 + ((q >> 2) * kBrick + ((q >> 1) & 1)) * kBrick + (q & 1);
            if (kWeight) unsafeAtomicAdd(&s_w[li], (A)f[q]);
            if (kModel) unsafeAtomicAdd(&s_m[li], (A)(f[q] * w));
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int ix = lx + (q >> 2), iy = ly + ((q >> 1) & 1), iz = lz + (q & 1);
        if (ix < 0 || ix >= kBrick || iy < 0 || iy >= kBrick || iz < 0 || iz >= kBrick) continue;
        const int li = l0 + ((q >> 2) * kBrick + ((q >> 1) & 1)) * kBrick + (q & 1);
        if (kWeight) unsafeAtomicAdd(&s_w[li], (A)f[q]);
        if (kModel) unsafeAtomicAdd(&s_m[li], (A)(f[q] * w));
    }
}

// Model-only merge (the weights are cached): most views of a sparse posterior are zero, so a
// wave compacts the nonzero (orientation, pixel) entries of its items into a 128-entry ring in
// LDS and interpolates and adds them 64 at a time, one entry per lane. A is the type of the
// brick's LDS sums: double for the FP64 path, float for the reduced-precision modes (the
// brick is added to the FP64 model once per work unit). VT: the view's type (double, or float
// where the reduced-precision pair tomogram writes it). nzmask (nullable): per item (chunk
// orientation c, patch p) at c * npatch + p, the pixels whose view is nonzero (the only view
// entries written); an item none of whose masked pixels is nonzero is passed over without a
// view read, and psum_r is not consulted (an orientation with psum_r = 0 has no nonzero view).
template <typename A, typename VT = double>
__global__ __launch_bounds__(kBlock) void merge_brick_model_kernel(
    const int4 *__restrict__ units, const int *__restrict__ list, const unsigned long long *__restrict__ lmask,
    const double *__restrict__ rots, int r0, int num_rot, const VT *__restrict__ view,
    const double *__restrict__ psum_r, int all_on, const double *__restrict__ qvals, const double *__restrict__ corr,
    const uint8_t *__restrict__ mask, int num_pix, long size, int npatch, int nbx, double *__restrict__ model2,
    const unsigned long long *__restrict__ nzmask) {
    constexpr int kVox = kBrick * kBrick * kBrick;
    constexpr int kRing = 128;
    __shared__ A s_m[kVox];
    __shared__ int q_r[kWavesPerBlock][kRing], q_t[kWavesPerBlock][kRing];
    __shared__ double q_w[kWavesPerBlock][kRing];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int4 u = units[blockIdx.x];
    const int nb3 = nbx * nbx * nbx, b = u.x % nb3;
    model2 += (size_t)(u.x / nb3) * (size_t)(size * size * size);  // the brick's mode volume
    const int X0 = (b / (nbx * nbx)) * kBrick, Y0 = ((b / nbx) % nbx) * kBrick, Z0 = (b % nbx) * kBrick;
    for (int i = tid; i < kVox; i += kBlock) s_m[i] = (A)0;
    __syncthreads();
    const double center = (double)(size / 2);
    int head = 0, tail = 0;  // wave-uniform ring positions
    auto drain = [&](int count) {
        if (lane < count) {
            const int at = (head + lane) & (kRing - 1);
            const int r = q_r[wave][at], t = q_t[wave][at];
            brick_add<true, false, A>(rots + 9 * (size_t)r, qvals[3 * (size_t)t], qvals[3 * (size_t)t + 1],
                                      qvals[3 * (size_t)t + 2], center, size, X0, Y0, Z0, q_w[wave][at] / corr[t],
                                      s_m, nullptr);
        }
        head += count;
    };
    // a wave reads the ids and pixel masks of 64 list entries at once (lane = entry), masks
    // them with the nonzero masks (or drops orientations with psum_r = 0), and then walks only
    // the live items, kItemsAhead at a time: their views are loaded before the first is
    // queued, so a wave has that many independent reads in flight
    constexpr int kItemsAhead = 4;
    for (int kb = u.y + wave * 64; kb < u.z; kb += kWavesPerBlock * 64) {
        const int k = kb + lane;
        int id = -1;
        unsigned long long m = 0ull;
        if (k < u.z) {
            id = list[k];
            m = lmask[k];
        }
        // item id = c * npatch + p (chunk orientation c, patch p), split once per lane
        const int ic = id >= 0 ? id / npatch : 0, ip = id >= 0 ? id - ic * npatch : 0;
        if (id >= 0) {
            if (nzmask)
                m &= nzmask[id];
            else if (!all_on && !(psum_r[r0 + ic] > 0.))
                m = 0ull;
        }
        unsigned long long live = __ballot(m != 0ull);
        while (live) {
            int cs[kItemsAhead], ps[kItemsAhead];
            unsigned long long ms[kItemsAhead];
#pragma unroll
            for (int a = 0; a < kItemsAhead; ++a) {
                cs[a] = -1;
                ps[a] = 0;
                ms[a] = 0ull;
                if (live) {
                    const int i = __ffsll((long long)live) - 1;
                    live &= live - 1ull;
                    cs[a] = __builtin_amdgcn_readlane(ic, i);
                    ps[a] = __builtin_amdgcn_readlane(ip, i);
                    ms[a] = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(m >> 32), i) << 32) |
                            (unsigned)__builtin_amdgcn_readlane((int)m, i);
                }
            }
            double wv[kItemsAhead];
#pragma unroll
            for (int a = 0; a < kItemsAhead; ++a) {
                wv[a] = 0.;  // the view; divided by corr when its entry is drained
                // a set bit: t < num_pix, mask[t] <= 1
                if (cs[a] >= 0 && ((ms[a] >> lane) & 1ull))
                    wv[a] = (double)view[(size_t)cs[a] * num_pix + ps[a] * 64 + lane];
            }
#pragma unroll
            for (int a = 0; a < kItemsAhead; ++a) {
                if (cs[a] < 0) continue;
                const bool nz = wv[a] != 0.;  // a zero view adds exactly nothing
                const unsigned long long bal = __ballot(nz);
                if (nz) {
                    const int at = (tail + __popcll(bal & ((1ull << lane) - 1))) & (kRing - 1);
                    q_r[wave][at] = r0 + cs[a];
                    q_t[wave][at] = ps[a] * 64 + lane;
                    q_w[wave][at] = wv[a];
                }
                tail += __popcll(bal);
                if (tail - head >= 64) drain(64);
            }
        }
    }
    if (tail > head) drain(tail - head);
    __syncthreads();
    const long s2 = size * size;
    for (int i = tid; i < kVox; i += kBlock) {
        const int x = X0 + i / (kBrick * kBrick), y = Y0 + (i / kBrick) % kBrick, z = Z0 + i % kBrick;
        if (x >= size || y >= size || z >= size) continue;
        if (s_m[i] != (A)0) unsafeAtomicAdd(&model2[x * s2 + (long)y * size + z], (double)s_m[i]);
    }
}

// One block per work unit (brick, item range). kModel adds w * f into model2, kWeight adds f
// into weight.
template <bool kModel, bool kWeight>
__global__ __launch_bounds__(kBlock) void merge_brick_kernel(
    const int4 *__restrict__ units, const int *__restrict__ list, const unsigned long long *__restrict__ lmask,
    const double *__restrict__ rots, int r0, int num_rot, const double *__restrict__ view,
    const double *__restrict__ psum_r, int all_on, const double *__restrict__ qvals, const double *__restrict__ corr,
    int num_pix, long size, int npatch, int nbx, double *__restrict__ model2, double *__restrict__ weight,
    const int *__restrict__ gate) {
    // gate: the number of orientations that merge nothing (psum_r = 0); none means the
    // weights come from the all-orientation cache instead and this pass has nothing to do
    if (gate && *gate == 0) return;
    constexpr int kVox = kBrick * kBrick * kBrick;
    __shared__ double s_m[kModel ? kVox : 1];
    __shared__ double s_w[kWeight ? kVox : 1];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int4 u = units[blockIdx.x];
    const int nb3 = nbx * nbx * nbx, b = u.x % nb3;
    const size_t moff = (size_t)(u.x / nb3) * (size_t)(size * size * size);  // the brick's mode volume
    if (kModel) model2 += moff;
    if (kWeight) weight += moff;
    const int X0 = (b / (nbx * nbx)) * kBrick, Y0 = ((b / nbx) % nbx) * kBrick, Z0 = (b % nbx) * kBrick;
    for (int i = tid; i < kVox; i += kBlock) {
        if (kModel) s_m[i] = 0.;
        if (kWeight) s_w[i] = 0.;
    }
    __syncthreads();
    const double center = (double)(size / 2);
    for (int k = u.y + wave; k < u.z; k += kWavesPerBlock) {
        // a set bit: the lane's pixel (t < num_pix, mask[t] <= 1) has a corner in this brick
        if (!((lmask[k] >> lane) & 1ull)) continue;
        const int item = list[k];
        const int c = item / npatch, p = item % npatch;
        const int r = r0 + c;
        if (!all_on && !(psum_r[r] > 0.)) continue;
        const int t = p * 64 + lane;
        double w = 0.;
        if (kModel) {
            w = view[(size_t)c * num_pix + t] / corr[t];
            // model-only merge: a zero view adds exactly nothing
            if (!kWeight && w == 0.) continue;
        }
        Cell cell;
        if (!locate(rots + 9 * (size_t)r, qvals[3 * (size_t)t], qvals[3 * (size_t)t + 1],
                    qvals[3 * (size_t)t + 2], center, size, cell))
            continue;
        const int lx = (int)cell.x - X0, ly = (int)cell.y - Y0, lz = (int)cell.z - Z0;
        if (lx < -1 || lx >= kBrick || ly < -1 || ly >= kBrick || lz < -1 || lz >= kBrick) continue;
        const double cx = 1. - cell.fx, cy = 1. - cell.fy, cz = 1. - cell.fz;
        const double fx = cell.fx, fy = cell.fy, fz = cell.fz;
        const double f[8] = {cx * cy * cz, cx * cy * fz, cx * fy * cz, cx * fy * fz,
                             fx * cy * cz, fx * cy * fz, fx * fy * cz, fx * fy * fz};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int ix = lx + (q >> 2), iy = ly + ((q >> 1) & 1), iz = lz + (q & 1);
            if (ix < 0 || ix >= kBrick || iy < 0 || iy >= kBrick || iz < 0 || iz >= kBrick) continue;
            const int li = (ix * kBrick + iy) * kBrick + iz;
            if (kWeight) unsafeAtomicAdd(&s_w[li], f[q]);
            if (kModel) unsafeAtomicAdd(&s_m[li], f[q] * w);
        }
    }
    __syncthreads();
    const long s2 = size * size;
    for (int i = tid; i < kVox; i += kBlock) {
        const int x = X0 + i / (kBrick * kBrick), y = Y0 + (i / kBrick) % kBrick, z = Z0 + i % kBrick;
        if (x >= size || y >= size || z >= size) continue;
        const long g = x * s2 + (long)y * size + z;
        if (kModel && s_m[i] != 0.) unsafeAtomicAdd(&model2[g], s_m[i]);
        if (kWeight && s_w[i] != 0.) unsafeAtomicAdd(&weight[g], s_w[i]);
    }
}

void launch_brick_items(hipStream_t s, const double *rots, int r0, int Cc, int num_rot,
                        const double *qvals, const uint8_t *mask, int num_pix, long size, int npatch,
                        int nbx, int pass, int *counts, const long *off, int *fill, int *list,
                        unsigned long long *lmask, const int *slot_rot, int modes) {
    const long items = (long)Cc * npatch;
    const unsigned grid = stride_grid((items + kWavesPerBlock - 1) / kWavesPerBlock);
    brick_items_kernel<<<grid, kBlock, 0, s>>>(rots, r0, Cc, num_rot, qvals, mask, num_pix, size, npatch,
                                               nbx, pass, counts, off, fill, list, lmask, slot_rot, modes);
}

void launch_merge_bricks(hipStream_t s, const int4 *units, int num_units, const int *list,
                         const unsigned long long *lmask, const double *rots, int r0, int num_rot,
                         const double *view, const double *psum_r, int all_on, const double *qvals,
                         const double *corr, const uint8_t *mask, int num_pix, long size, int npatch, int nbx,
                         double *model2, double *weight, int do_model, int do_weight, int f32_sums,
                         const int *gate, const unsigned long long *nzmask) {
    if (num_units <= 0) return;
#define EMC_MB(M, W)                                                                                        \
    merge_brick_kernel<M, W><<<num_units, kBlock, 0, s>>>(units, list, lmask, rots, r0, num_rot, view, psum_r, \
                                                          all_on, qvals, corr, num_pix, size, npatch, nbx,     \
                                                          model2, weight, gate)
    if (do_model && do_weight) EMC_MB(true, true);
    else if (do_model && f32_sums)
        merge_brick_model_kernel<float><<<num_units, kBlock, 0, s>>>(units, list, lmask, rots, r0, num_rot, view,
                                                                     psum_r, all_on, qvals, corr, mask, num_pix,
                                                                     size, npatch, nbx, model2, nzmask);
    else if (do_model)
        merge_brick_model_kernel<double><<<num_units, kBlock, 0, s>>>(units, list, lmask, rots, r0, num_rot, view,
                                                                      psum_r, all_on, qvals, corr, mask, num_pix,
                                                                      size, npatch, nbx, model2, nzmask);
    else if (do_weight) EMC_MB(false, true);
#undef EMC_MB
}

void launch_merge_model_f32(hipStream_t s, const int4 *units, int num_units, const int *list,
                            const unsigned long long *lmask, const double *rots, int r0, int num_rot,
                            const float *view, const double *psum_r, int all_on, const double *qvals,
                            const double *corr, const uint8_t *mask, int num_pix, long size, int npatch, int nbx,
                            double *model2, const unsigned long long *nzmask) {
    if (num_units <= 0) return;
    merge_brick_model_kernel<float, float><<<num_units, kBlock, 0, s>>>(units, list, lmask, rots, r0, num_rot, view,
                                                                        psum_r, all_on, qvals, corr, mask, num_pix,
                                                                        size, npatch, nbx, model2, nzmask);
}

}  // namespace emc
