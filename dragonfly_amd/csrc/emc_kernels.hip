// HIP kernels of the plain FP64 EMC iteration for gfx950 (MI355X).
//
// Layout in HBM (all row-major):
//   logW / view   [C/64][num_pix][64]   one orientation chunk of C (multiple of 64) columns,
//                                       tile-major: the 64-orientation slab of every pixel is one
//                                       contiguous block and a photon's row is a 512-byte
//                                       coalesced read by one wavefront
//   R / P         [Rp/64][num_data][64] per-frame log-likelihood, then posterior, in place;
//                                       Rp = num_rot rounded up to 64, same tile-major form
// Element (row i, orientation r) of a tile-major buffer with tile stride ts is at
// (r / 64) * ts + i * 64 + r % 64.
//   model         [size][size][size] x-major (index x*size*size + y*size + z)
// Every kernel keeps the reference's operation order for the values it computes
// (the library is compiled with -ffp-contract=off, so no multiply-add is fused).
#include <cfloat>
#include <climits>

#include "emc_common.h"
#include "emc_expand.h"
#include "emc_geometry.h"
#include "emc_kernels.h"

namespace emc {

// Scatter one pixel value into model2/weight (model.c:209-242); w = view / corr.
__device__ __forceinline__ void merge_value(double *__restrict__ model2, double *__restrict__ weight,
                                            long size, const Cell &c, double w) {
    const long s2 = size * size;
    const long x1 = (c.x + 1) % size, y1 = (c.y + 1) % size, z1 = (c.z + 1) % size;
    const double cx = 1. - c.fx, cy = 1. - c.fy, cz = 1. - c.fz;
    const double fx = c.fx, fy = c.fy, fz = c.fz;
    const long idx[8] = {c.x * s2 + c.y * size + c.z, c.x * s2 + c.y * size + z1,
                         c.x * s2 + y1 * size + c.z, c.x * s2 + y1 * size + z1,
                         x1 * s2 + c.y * size + c.z, x1 * s2 + c.y * size + z1,
                         x1 * s2 + y1 * size + c.z, x1 * s2 + y1 * size + z1};
    const double f[8] = {cx * cy * cz, cx * cy * fz, cx * fy * cz, cx * fy * fz,
                         fx * cy * cz, fx * cy * fz, fx * fy * cz, fx * fy * fz};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        unsafeAtomicAdd(&weight[idx[k]], f[k]);
        unsafeAtomicAdd(&model2[idx[k]], f[k] * w);
    }
}

// ---- expand: log W for one orientation chunk, plus partial u_r ---------------------------
// Block = 64 orientations x one pixel group (expand_block, emc_expand.h).
template <typename T, int kW, typename MT = double>
__global__ __launch_bounds__(kBlock) void expand_kernel(
    const double *__restrict__ rots, const MT *__restrict__ model, long size,
    const double *__restrict__ qvals, const double *__restrict__ corr,
    const uint8_t *__restrict__ mask, int num_pix, int r0, int C, int pix_per_group,
    T *__restrict__ logw, double *__restrict__ upart, int *__restrict__ ubad,
    const int *__restrict__ slot_rot, int modes, int raw) {
    __shared__ ExpandShared<MT> sh;
    // XCD-aware order: workgroups go round-robin to the kNumXcd XCDs, so the b-th block of XCD
    // x takes tile x * per + b % per of its own contiguous range of tiles (neighbouring
    // orientations: the slot order follows a chain of neighbouring cells) and pixel group
    // b / per; the blocks resident on an XCD then read one region of the model through its L2
    const int ntiles = C / kRTile, per = (ntiles + kNumXcd - 1) / kNumXcd;
    const int xcd = blockIdx.x % kNumXcd, b = blockIdx.x / kNumXcd;
    const int tile = xcd * per + b % per, g = b / per;
    if (tile >= ntiles || (long)g * pix_per_group >= num_pix) return;
    expand_block<T, kW, MT>(sh, tile * kRTile, g, rots, model, size, qvals, corr, mask, num_pix, r0, C, pix_per_group,
                            logw, upart, ubad, slot_rot, modes, raw);
}

// grid of expand_kernel: kNumXcd ranges of per tiles, each over every pixel group
static unsigned expand_grid(int C, int num_pix, int pix_per_group) {
    const long per = (C / kRTile + kNumXcd - 1) / kNumXcd;
    return (unsigned)(kNumXcd * per * ((num_pix + pix_per_group - 1) / pix_per_group));
}

// u[r0 + c] = sum over pixel groups, fixed order.
__global__ void u_reduce_kernel(const double *__restrict__ upart, const int *__restrict__ ubad,
                                int ngroups, int C, int r0, int num_rot, double *__restrict__ u,
                                int *__restrict__ rbad) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C || r0 + c >= num_rot) return;
    double acc = 0.;
    int bad = 0;
    for (int g = 0; g < ngroups; ++g) {
        acc += upart[(size_t)g * C + c];
        bad |= ubad[(size_t)g * C + c];
    }
    u[r0 + c] = acc;
    rbad[r0 + c] = bad;
}

// rescale = mean_count / sum_r w_r u_r (max_prob.c:40, :59); one block, fixed order.
__global__ __launch_bounds__(kBlock) void rescale_kernel(const double *__restrict__ quats,
                                                         const double *__restrict__ u, int num_rot,
                                                         double mean_count, double modes,
                                                         double *__restrict__ scalars) {
    __shared__ double s[kBlock];
    double acc = 0.;
    for (int r = threadIdx.x; r < num_rot; r += kBlock) acc += quats[5 * (size_t)r + 4] * u[r];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    // rescale = mean_count / total * num_modes (max_prob.c:59)
    if (threadIdx.x == 0) scalars[0] = mean_count / s[0] * modes;
}

// ---- likelihood: G_dr = sum_t K_dt log W_rt over one frame's photons (mask < 1) ------------
// One wavefront = one frame x NT*64 orientations; the frame's (pixel, count) list is read
// with scalar loads (wave-uniform) and each photon costs NT 512-byte row reads of logW.
// Row loads of several photons are issued before the additions, which stay in list order.
template <int NT, int U, typename T>
__global__ __launch_bounds__(kBlock) void likelihood_kernel(
    const T *__restrict__ logw, int C, size_t tsW, const long *__restrict__ off,
    const int *__restrict__ pix, const double *__restrict__ cnt,
    const uint8_t *__restrict__ blacklist, int num_data, double *__restrict__ G, size_t tsG,
    int r0, const int *__restrict__ fmap, int xcd_order) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    int fg = blockIdx.x, tb = blockIdx.y;
    if (xcd_order) {
        // 1-D grid, XCD-aware: workgroups go round-robin to the kNumXcd XCDs. The XCDs form
        // groups of G = xcd_order (1, 2, 4 or 8); group g takes orientation tiles
        // g, g + kNumXcd / G, ... one at a time, its G members splitting the frames, so each
        // L2 sees the rows of one log W tile and kNumXcd / G tiles are live chip-wide
        const int G = xcd_order, ngrp = kNumXcd / G;
        const int nfg = (num_data + kWavesPerBlock - 1) / kWavesPerBlock, nfgG = (nfg + G - 1) / G;
        const int x = blockIdx.x % kNumXcd, j = blockIdx.x / kNumXcd;
        tb = x / G + ngrp * (j / nfgG);
        fg = (j % nfgG) * G + x % G;
        if (tb * (kRTile * NT) >= C || fg >= nfg) return;
    }
    const int d = fg * kWavesPerBlock + wave;
    if (d >= num_data) return;
    const int fd = fmap ? fmap[d] : d;  // frame whose photons row d holds
    if (blacklist[fd]) return;
    const int c0 = tb * (kRTile * NT);
    const int nt = min(NT, (C - c0) / kRTile);
    const T *__restrict__ col = logw + (size_t)(c0 >> 6) * tsW + lane;
    const long k1 = off[fd + 1];
    long k = off[fd];
    double acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = 0.;
    for (; k + U <= k1; k += U) {
        double v[U][NT];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t row = (size_t)pix[k + u] * kRTile;
#pragma unroll
            for (int j = 0; j < NT; ++j) v[u][j] = (j < nt) ? (double)col[row + tsW * j] : 0.;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int j = 0; j < NT; ++j) acc[j] += cnt[k + u] * v[u][j];
    }
    for (; k < k1; ++k) {
        const size_t row = (size_t)pix[k] * kRTile;
#pragma unroll
        for (int j = 0; j < NT; ++j)
            if (j < nt) acc[j] += cnt[k] * (double)col[row + tsW * j];
    }
#pragma unroll
    for (int j = 0; j < NT; ++j)
        if (j < nt) G[(size_t)((r0 + c0) / kRTile + j) * tsG + (size_t)d * kRTile + lane] = acc[j];
}

// Likelihood over pixel ranges: block (range x, frame group, tile) sums the frame's photons of
// range x (lists split by build_lik_split) against one log W tile into its own partial G_x.
// Workgroups go round-robin to the XCDs and the range varies fastest, so with X = 8 XCD x only
// ever reads the rows of range x: about 1/8 of a tile's rows, which its L2 can hold. The
// partials are added in range order by likelihood_split_reduce_kernel.
template <int U>
__global__ __launch_bounds__(kBlock) void likelihood_split_kernel(
    const double *__restrict__ logw, size_t tsW, const long *__restrict__ off8, int X,
    const int *__restrict__ pix, const double *__restrict__ cnt, const uint8_t *__restrict__ blacklist,
    int num_data, int nfg, double *__restrict__ part, size_t tsPart) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int x = blockIdx.x % X, fg = blockIdx.x / X;
    if (fg >= nfg) return;
    const int d = fg * kWavesPerBlock + wave;
    if (d >= num_data || blacklist[d]) return;
    const int tile = blockIdx.y;
    const double *__restrict__ col = logw + (size_t)tile * tsW + lane;
    long k = off8[(size_t)d * (X + 1) + x];
    const long k1 = off8[(size_t)d * (X + 1) + x + 1];
    double acc = 0.;
    for (; k + U <= k1; k += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = col[(size_t)pix[k + u] * kRTile];
#pragma unroll
        for (int u = 0; u < U; ++u) acc += cnt[k + u] * v[u];
    }
    for (; k < k1; ++k) acc += cnt[k] * col[(size_t)pix[k] * kRTile];
    part[(size_t)x * tsPart + (size_t)tile * num_data * kRTile + (size_t)d * kRTile + lane] = acc;
}

// G[i] = sum over ranges x (in order) of part[x][i], for the chunk's tiles.
__global__ void likelihood_split_reduce_kernel(const double *__restrict__ part, int X, size_t n,
                                               double *__restrict__ G) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double acc = 0.;
        for (int x = 0; x < X; ++x) acc += part[(size_t)x * n + i];
        G[i] = acc;
    }
}

// ---- block helpers -----------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double *s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    double r = s[0];
    __syncthreads();
    return r;
}

// ---- per-frame maximum: R = log w_r - phi_d u_r + G, first arg-max (max_prob.c:122-127,
// :199-203, normalize_prob :223-241) ----------------------------------------------------------
__global__ __launch_bounds__(kBlock) void frame_max_kernel(
    double *__restrict__ R, size_t tsR, int num_rot, const double *__restrict__ logwr,
    const double *__restrict__ u, const double *__restrict__ scale, int use_scale,
    const double *__restrict__ scalars, const uint8_t *__restrict__ blacklist,
    const uint8_t *__restrict__ dense, const int *__restrict__ rbad,
    const int *__restrict__ qmap, const uint32_t *__restrict__ relbits, int words,
    double *__restrict__ max_exp, int *__restrict__ rmax) {
    __shared__ double s_v[kBlock];
    __shared__ int s_i[kBlock];
    const int d = blockIdx.x;
    // refinement (legacy src/max_emc.c:264-274): frame d only takes orientations whose coarse
    // cell qmap[r] is set in its bit row; the others are not evaluated (R = -inf: never the
    // maximum, never kept)
    const uint32_t *bits = relbits ? relbits + (size_t)d * words : nullptr;
    // dense frames sum K log W over every mask < 1 pixel (max_prob.c:173-177): an orientation
    // with log W = -inf or NaN on such a pixel gives 0 * -inf = NaN and is never kept
    const bool dense_frame = dense != nullptr && dense[d];
    if (blacklist[d]) {
        if (threadIdx.x == 0) {
            max_exp[d] = kMaxExpStart;
            rmax[d] = -1;
        }
        return;
    }
    const double phi = use_scale ? scale[d] : scalars[0];
    double *row = R + (size_t)d * kRTile;
    double best = kMaxExpStart;
    int bi = -1;
    for (int r = threadIdx.x; r < num_rot; r += kBlock) {
        double &x = row[(size_t)(r >> 6) * tsR + (r & 63)];
        bool skip = dense_frame && rbad[r];
        if (bits) {
            const int c = qmap[r];
            skip |= !((bits[c >> 5] >> (c & 31)) & 1u);
        }
        const double v = skip ? -INFINITY : (logwr[r] - u[r] * phi) + x;
        x = v;
        if (v > best) {
            best = v;
            bi = r;
        }
    }
    s_v[threadIdx.x] = best;
    s_i[threadIdx.x] = bi;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double ov = s_v[threadIdx.x + w];
            const int oi = s_i[threadIdx.x + w];
            if (ov > s_v[threadIdx.x] || (ov == s_v[threadIdx.x] && oi >= 0 && oi < s_i[threadIdx.x])) {
                s_v[threadIdx.x] = ov;
                s_i[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        max_exp[d] = s_v[0];
        rmax[d] = (s_v[0] == kMaxExpStart) ? -1 : s_i[0];
    }
}

// Kept set of the reference's single-thread keep list (max_prob.c:185-203): the first
// arg-max rf; r > rf kept if R + 14/beta > M; r < rf kept if R > M - 14/beta.
__device__ __forceinline__ bool kept(double v, int r, int rf, double M, double c) {
    if (r == rf) return true;
    return r > rf ? (v + c > M) : (v > M - c);
}

// p_norm_d = sum_kept exp(beta (R - M)) (normalize_prob, max_prob.c:245-250).
__global__ __launch_bounds__(kBlock) void frame_norm_kernel(
    const double *__restrict__ R, size_t tsR, int num_rot, const double *__restrict__ beta,
    const uint8_t *__restrict__ blacklist, const double *__restrict__ max_exp,
    const int *__restrict__ rmax, double *__restrict__ p_norm) {
    __shared__ double s[kBlock];
    const int d = blockIdx.x;
    const int rf = rmax[d];
    if (blacklist[d] || rf < 0) {
        if (threadIdx.x == 0) p_norm[d] = 0.;
        return;
    }
    const double b = beta[d], M = max_exp[d], c = kPdiffThresh / b;
    const double *row = R + (size_t)d * kRTile;
    double acc = 0.;
    for (int r = threadIdx.x; r < num_rot; r += kBlock) {
        const double v = row[(size_t)(r >> 6) * tsR + (r & 63)];
        if (kept(v, r, rf, M, c)) acc += exp(b * (v - M));
    }
    const double tot = block_sum(acc, s);
    if (threadIdx.x == 0) p_norm[d] = tot;
}

// Posterior and per-frame statistics (calc_psum_r, max_update.c:95-130):
// P = exp(beta (R - M)) / p_norm on the kept set, 0 elsewhere (written in place);
// likelihood += P (R - sum_fact); psum_d += P u_r rescale; info += P log(P / w_r * num_modes)
// for P >= 1e-6.
__global__ __launch_bounds__(kBlock) void posterior_kernel(
    double *__restrict__ R, size_t tsR, int num_rot, const double *__restrict__ logwr,
    const double *__restrict__ u, const double *__restrict__ beta,
    const double *__restrict__ sum_fact, const uint8_t *__restrict__ blacklist,
    const double *__restrict__ max_exp, const int *__restrict__ rmax,
    const double *__restrict__ p_norm, const double *__restrict__ scalars, int want_psum_d,
    double log_modes, double *__restrict__ likelihood, double *__restrict__ info, double *__restrict__ psum_d) {
    __shared__ double s[kBlock];
    const int d = blockIdx.x;
    double *row = R + (size_t)d * kRTile;
    const int rf = rmax[d];
    if (blacklist[d] || rf < 0) {
        for (int r = threadIdx.x; r < num_rot; r += kBlock) row[(size_t)(r >> 6) * tsR + (r & 63)] = 0.;
        if (threadIdx.x == 0) {
            likelihood[d] = 0.;
            info[d] = 0.;
            if (want_psum_d) psum_d[d] = 0.;
        }
        return;
    }
    const double b = beta[d], M = max_exp[d], c = kPdiffThresh / b, pn = p_norm[d];
    const double sf = sum_fact[d], rescale = scalars[0], log_pn = log(pn);
    double lk = 0., inf = 0., ps = 0.;
    for (int r = threadIdx.x; r < num_rot; r += kBlock) {
        double &x = row[(size_t)(r >> 6) * tsR + (r & 63)];
        const double v = x;
        double prob = 0.;
        if (kept(v, r, rf, M, c)) {
            const double e = b * (v - M);
            prob = exp(e) / pn;
            lk += prob * (v - sf);
            ps += prob * u[r] * rescale;
            // log(P / w_r * num_modes) (max_update.c:127-130) taken as e - log p_norm - log w_r
            // + log num_modes: the same value to rounding, without a logarithm per pair (one
            // mode adds exactly 0)
            if (!(prob < kProbMin)) inf += prob * ((e - log_pn) - logwr[r] + log_modes);
        }
        x = prob;
    }
    const double tl = block_sum(lk, s);
    const double ti = block_sum(inf, s);
    const double tp = block_sum(ps, s);
    if (threadIdx.x == 0) {
        likelihood[d] = tl;
        info[d] = ti;
        if (want_psum_d) psum_d[d] = tp;
    }
}

// Sparse posterior (the saved probabilities, max_prob.c:295-326: entries P > 1e-6 of every
// frame, resparsify's cut). Count pass: one block per frame.
// kRows: P is row-major, P[d * num_rot + r], instead of tile-major.
template <bool kRows>
__device__ __forceinline__ double post_at(const double *__restrict__ P, size_t tsR, int num_rot, int d, int r) {
    return kRows ? P[(size_t)d * num_rot + r] : P[(size_t)d * kRTile + (size_t)(r >> 6) * tsR + (r & 63)];
}

template <bool kRows>
__global__ __launch_bounds__(kBlock) void sparse_count_kernel(const double *__restrict__ P, size_t tsR,
                                                              int num_rot, int *__restrict__ counts) {
    __shared__ int s_n[kWavesPerBlock];
    const int d = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int n = 0;
    for (int r = threadIdx.x; r < num_rot; r += kBlock) n += post_at<kRows>(P, tsR, num_rot, d, r) > kProbMin;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if (lane == 0) s_n[wave] = n;
    __syncthreads();
    if (threadIdx.x == 0) counts[d] = s_n[0] + s_n[1] + s_n[2] + s_n[3];
}

// Fill pass: frame d0 + blockIdx.x writes its entries, orientations ascending, at off[blockIdx.x]
// (offsets relative to the batch's first entry): place = orientation, prob = P.
template <bool kRows>
__global__ __launch_bounds__(kBlock) void sparse_fill_kernel(const double *__restrict__ P, size_t tsR,
                                                             int num_rot, int d0, const long *__restrict__ off,
                                                             int *__restrict__ place, double *__restrict__ prob) {
    __shared__ int s_n[kWavesPerBlock];
    const int d = d0 + blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    long at = off[blockIdx.x];
    for (int base = 0; base < num_rot; base += kBlock) {
        const int r = base + threadIdx.x;
        const double v = r < num_rot ? post_at<kRows>(P, tsR, num_rot, d, r) : 0.;
        const bool keep = v > kProbMin;
        const unsigned long long bal = __ballot(keep);
        if (lane == 0) s_n[wave] = __popcll(bal);
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < kWavesPerBlock; ++w) {
            before += w < wave ? s_n[w] : 0;
            total += s_n[w];
        }
        if (keep) {
            const long k = at + before + __popcll(bal & ((1ull << lane) - 1));
            place[k] = r;
            prob[k] = v;
        }
        at += total;
        __syncthreads();  // s_n is rewritten by the next round
    }
}

// Occupancy of every frame in every mode (max_update.c:122-124): quat_norm[d][m] = sum of the
// posterior P_dr >= 1e-6 over the combined orientations r = rotation * modes + m. One block
// per frame; P is the tile-major posterior (0 where not kept).
__global__ __launch_bounds__(kBlock) void occupancy_kernel(const double *__restrict__ P, size_t tsR,
                                                           int num_rot, int modes,
                                                           double *__restrict__ quat_norm) {
    __shared__ double s[kBlock];
    const int d = blockIdx.x;
    const double *row = P + (size_t)d * kRTile;
    for (int m = 0; m < modes; ++m) {
        double acc = 0.;
        for (int r = m + modes * (int)threadIdx.x; r < num_rot; r += modes * kBlock) {
            const double v = row[(size_t)(r >> 6) * tsR + (r & 63)];
            if (!(v < kProbMin)) acc += v;
        }
        const double tot = block_sum(acc, s);
        if (threadIdx.x == 0) quat_norm[(size_t)d * modes + m] = tot;
    }
}

// psum_r = sum_d P_dr phi'_d, phi' = scale if need_scaling else 1 (max_update.c:108-116).
__global__ __launch_bounds__(kBlock) void psum_r_partial_kernel(
    const double *__restrict__ P, size_t tsP, int num_rot, int num_data, int frames_per_part,
    const double *__restrict__ scale, int need_scaling, const uint8_t *__restrict__ blacklist,
    double *__restrict__ part) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    const int part_id = blockIdx.y;
    if (r >= num_rot) return;
    const int d0 = part_id * frames_per_part, d1 = min(num_data, d0 + frames_per_part);
    double acc = 0.;
    for (int d = d0; d < d1; ++d) {
        if (blacklist[d]) continue;
        const double p = P[(size_t)(r >> 6) * tsP + (size_t)d * kRTile + (r & 63)];
        acc += need_scaling ? p * scale[d] : p;
    }
    part[(size_t)part_id * num_rot + r] = acc;
}

__global__ void psum_r_reduce_kernel(const double *__restrict__ part, int nparts, int num_rot,
                                     double *__restrict__ psum_r) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= num_rot) return;
    double acc = 0.;
    for (int i = 0; i < nparts; ++i) acc += part[(size_t)i * num_rot + r];
    psum_r[r] = acc;
}

// ---- tomogram: view_rt = sum_{d: P > 1e-6} P_dr K_dt / psum_r (update_tomogram_nobg,
// max_update.c:142-230), gathered per pixel over its frames (d ascending, as the
// reference accumulates). One wavefront = one pixel x NT*64 orientations. P rows of
// blacklisted frames are zero (posterior_kernel), so they add nothing here.
template <int NT>
__global__ __launch_bounds__(kBlock) void tomogram_kernel(
    const double *__restrict__ P, size_t tsP, int r0, int C, size_t tsW,
    const long *__restrict__ toff, const int *__restrict__ tframe, const double *__restrict__ tcnt,
    const double *__restrict__ psum_r, int num_rot, int num_pix, double *__restrict__ view) {
    constexpr int U = NT >= 4 ? 2 : (NT == 2 ? 4 : 8);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * kWavesPerBlock + wave;
    if (t >= num_pix) return;
    const int c0 = blockIdx.y * (kRTile * NT);
    const int nt = min(NT, (C - c0) / kRTile);
    const double *__restrict__ col = P + (size_t)((r0 + c0) >> 6) * tsP + lane;
    const long k1 = toff[t + 1];
    long k = toff[t];
    double acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = 0.;
    for (; k + U <= k1; k += U) {
        double v[U][NT];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t row = (size_t)tframe[k + u] * kRTile;
#pragma unroll
            for (int j = 0; j < NT; ++j) v[u][j] = (j < nt) ? col[row + tsP * j] : 0.;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int j = 0; j < NT; ++j)
                if (v[u][j] > kProbMin) acc[j] += tcnt[k + u] * v[u][j];
    }
    for (; k < k1; ++k) {
        const size_t row = (size_t)tframe[k] * kRTile;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const double p = (j < nt) ? col[row + tsP * j] : 0.;
            if (p > kProbMin) acc[j] += tcnt[k] * p;
        }
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        if (j >= nt) continue;
        const int r = r0 + c0 + kRTile * j + lane;
        const double ps = (r < num_rot) ? psum_r[r] : 0.;
        view[(size_t)((c0 >> 6) + j) * tsW + (size_t)t * kRTile + lane] = (ps > 0.) ? acc[j] / ps : acc[j];
    }
}

// Frame-partitioned form of the tomogram. Frames are split into nparts contiguous ranges and
// block b works on range b % nparts: with blocks dispatched round-robin over the 8 XCDs
// (nparts a multiple of 8), each XCD's L2 holds only its own frames' P slab of the
// orientation tile. Each range writes a partial view; tomogram_reduce_kernel adds the
// partials in range order and divides by psum_r. toff2[t * nparts + x] is the first photon of
// pixel t with a frame in range x (the per-pixel list is frame-ascending).
//
// kSeq orders the blocks so that each XCD walks its ranges one after another (all pixels of
// range x, then of range x + 8, ...) instead of interleaving them, which keeps one range's
// P slab in that XCD's L2 at a time; nparts is then a multiple of 8.
template <int U, bool kSeq, typename T>
__global__ __launch_bounds__(kBlock) void tomogram_part_kernel(
    const T *__restrict__ P, size_t tsP, int r0, int C, const long *__restrict__ toff2,
    int nparts, const int *__restrict__ tframe, const double *__restrict__ tcnt, int num_pix,
    int ngroups, double *__restrict__ partial) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    int part, group;
    if (kSeq) {
        const int seq = blockIdx.x / kNumXcd;
        part = (seq / ngroups) * kNumXcd + blockIdx.x % kNumXcd;
        group = seq % ngroups;
    } else {
        part = blockIdx.x % nparts;
        group = blockIdx.x / nparts;
    }
    const int t = group * kWavesPerBlock + wave;
    if (t >= num_pix) return;
    const int c0 = blockIdx.y * kRTile;
    const T *__restrict__ col = P + (size_t)((r0 + c0) >> 6) * tsP + lane;
    const size_t at = (size_t)t * nparts + part;
    long k = toff2[at];
    const long k1 = toff2[at + 1];
    double acc = 0.;
    for (; k + U <= k1; k += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = (double)col[(size_t)tframe[k + u] * kRTile];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (v[u] > kProbMin) acc += tcnt[k + u] * v[u];
    }
    for (; k < k1; ++k) {
        const double p = (double)col[(size_t)tframe[k] * kRTile];
        if (p > kProbMin) acc += tcnt[k] * p;
    }
    const size_t tsW = (size_t)num_pix * kRTile, tsPart = tsW * (C / kRTile);
    partial[part * tsPart + (size_t)(c0 >> 6) * tsW + (size_t)t * kRTile + lane] = acc;
}

// Same as tomogram_part_kernel with a wavefront covering 32 orientations: lanes 0-31 take the
// even photons of the pixel's list and lanes 32-63 the odd ones (two 256-byte P segments per
// load instruction); the halves are added at the end. The L2 slab of one frame range for one
// 32-orientation block is half as large.
__global__ __launch_bounds__(kBlock) void tomogram_part32_kernel(
    const double *__restrict__ P, size_t tsP, int r0, int C, const long *__restrict__ toff2,
    int nparts, const int *__restrict__ tframe, const double *__restrict__ tcnt, int num_pix,
    double *__restrict__ partial) {
    constexpr int U = 8;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, half = lane >> 5;
    const int part = blockIdx.x % nparts;
    const int t = (blockIdx.x / nparts) * kWavesPerBlock + wave;
    if (t >= num_pix) return;
    const int c0 = blockIdx.y * (kRTile / 2);          // chunk-local first orientation (multiple of 32)
    const int rg = r0 + c0;                             // global first orientation
    const double *__restrict__ col = P + (size_t)(rg >> 6) * tsP + (rg & 63) + (lane & 31);
    const size_t at = (size_t)t * nparts + part;
    long k = toff2[at];
    const long k1 = toff2[at + 1];
    double acc = 0.;
    for (; k + 2 * U <= k1; k += 2 * U) {
        double v[U], c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long kk = k + 2 * u + half;
            v[u] = col[(size_t)tframe[kk] * kRTile];
            c[u] = tcnt[kk];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (v[u] > kProbMin) acc += c[u] * v[u];
    }
    for (; k < k1; k += 2) {
        const long kk = k + half;
        if (kk < k1) {
            const double p = col[(size_t)tframe[kk] * kRTile];
            if (p > kProbMin) acc += tcnt[kk] * p;
        }
    }
    acc += __shfl_down(acc, 32);
    if (half == 0) {
        const size_t tsW = (size_t)num_pix * kRTile, tsPart = tsW * (C / kRTile);
        partial[part * tsPart + (size_t)(c0 >> 6) * tsW + (size_t)t * kRTile + (c0 & 63) + lane] = acc;
    }
}

// Same sum, written orientation-major (view[c][t], c chunk-local) through an LDS transpose:
// block = 64 pixels x 64 orientations.
__global__ __launch_bounds__(kBlock) void tomogram_reduce_rmajor_kernel(
    const double *__restrict__ partial, int nparts, int r0, int C, int num_pix,
    const double *__restrict__ psum_r, int num_rot, double *__restrict__ view) {
    __shared__ double s[64][kRTile + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t0 = blockIdx.x * 64, c0 = blockIdx.y * kRTile;
    const size_t tsW = (size_t)num_pix * kRTile, tsPart = tsW * (C / kRTile);
    const int r = r0 + c0 + lane;
    const double ps = (psum_r && r < num_rot) ? psum_r[r] : 0.;  // no psum: sums only
    for (int k = wave; k < 64; k += kWavesPerBlock) {
        const int t = t0 + k;
        if (t >= num_pix) break;
        const size_t i = (size_t)(c0 >> 6) * tsW + (size_t)t * kRTile + lane;
        double acc = 0.;
        for (int x = 0; x < nparts; ++x) acc += partial[x * tsPart + i];
        s[k][lane] = (ps > 0.) ? acc / ps : acc;
    }
    __syncthreads();
    const int t = t0 + lane;
    if (t < num_pix)
        for (int k = wave; k < kRTile; k += kWavesPerBlock)
            view[(size_t)(c0 + k) * num_pix + t] = s[lane][k];
}

__global__ void tomogram_reduce_kernel(const double *__restrict__ partial, int nparts, int r0,
                                       int C, int num_pix, const double *__restrict__ psum_r,
                                       int num_rot, double *__restrict__ view) {
    const size_t tsW = (size_t)num_pix * kRTile, n = tsW * (C / kRTile);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        double acc = 0.;
        for (int x = 0; x < nparts; ++x) acc += partial[x * n + i];
        const int r = r0 + (int)(i / tsW) * kRTile + (int)(i % kRTile);
        const double ps = (r < num_rot) ? psum_r[r] : 0.;
        view[i] = (ps > 0.) ? acc / ps : acc;
    }
}

// Tile form of the tomogram. Engine pixels are grouped in tiles of kTomoTile consecutive
// pixels (compact detector patches, Morton order). For a tile the photons are listed frame by
// frame: ft_off[tile]..ft_off[tile+1] index the tile's distinct frames (ft_frame), and
// fe_off[f]..fe_off[f+1] that frame's photons in the tile (fe_code = count << 8 | local pixel).
// A block = one tile x 64 orientations: each frame's P row segment is loaded once and added
// into LDS accumulators for every photon of that frame in the tile, so bright tiles reuse
// each 512-byte P load several times. Accumulation order across waves is not fixed.
constexpr int kTomoTile = 128;

__global__ __launch_bounds__(kBlock) void tomogram_tile_kernel(
    const double *__restrict__ P, size_t tsP, int r0, int C, size_t tsW,
    const long *__restrict__ ft_off, const int *__restrict__ ft_frame, const long *__restrict__ fe_off,
    const int *__restrict__ fe_code, const double *__restrict__ psum_r, int num_rot, int num_pix,
    double *__restrict__ view) {
    __shared__ double s_acc[kTomoTile][kRTile];
    const int tile = blockIdx.x;
    const int c0 = blockIdx.y * kRTile;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int i = threadIdx.x; i < kTomoTile * kRTile; i += kBlock) (&s_acc[0][0])[i] = 0.;
    __syncthreads();
    const double *__restrict__ col = P + (size_t)((r0 + c0) >> 6) * tsP + lane;
    const long f1 = ft_off[tile + 1];
    long f = ft_off[tile] + wave;
    // four frames per step (one per wave of the block), two frames' rows in flight per wave
    for (; f + kWavesPerBlock < f1; f += 2 * kWavesPerBlock) {
        const long g = f + kWavesPerBlock;
        const double pa = col[(size_t)ft_frame[f] * kRTile];
        const double pb = col[(size_t)ft_frame[g] * kRTile];
        if (pa > kProbMin)
            for (long e = fe_off[f]; e < fe_off[f + 1]; ++e) {
                const int code = fe_code[e];
                unsafeAtomicAdd(&s_acc[code & 255][lane], (double)(code >> 8) * pa);
            }
        if (pb > kProbMin)
            for (long e = fe_off[g]; e < fe_off[g + 1]; ++e) {
                const int code = fe_code[e];
                unsafeAtomicAdd(&s_acc[code & 255][lane], (double)(code >> 8) * pb);
            }
    }
    for (; f < f1; f += kWavesPerBlock) {
        const double pa = col[(size_t)ft_frame[f] * kRTile];
        if (pa > kProbMin)
            for (long e = fe_off[f]; e < fe_off[f + 1]; ++e) {
                const int code = fe_code[e];
                unsafeAtomicAdd(&s_acc[code & 255][lane], (double)(code >> 8) * pa);
            }
    }
    __syncthreads();
    const int r = r0 + c0 + lane;
    const double ps = (r < num_rot) ? psum_r[r] : 0.;
    for (int lp = wave; lp < kTomoTile; lp += kWavesPerBlock) {
        const int t = tile * kTomoTile + lp;
        if (t >= num_pix) break;
        const double acc = s_acc[lp][lane];
        view[(size_t)(c0 >> 6) * tsW + (size_t)t * kRTile + lane] = (ps > 0.) ? acc / ps : acc;
    }
}

// ---- merge: scatter views into model2 / inter_weight (merge_tomogram, max_update.c:33-36;
// slice_merge3d model.c:171-244), only orientations with psum_r > 0 and pixels mask <= 1 ----
//
// Pixels are stored in 8x8 q-space tiles (Morton order), so the 64 lanes of a wavefront hold
// one compact detector patch. For one orientation the patch lands in a small voxel box: the
// wavefront accumulates its 8-corner contributions in a private LDS box with LDS atomics and
// then adds each touched voxel to HBM once. A box larger than kBoxCap falls back to direct
// HBM atomics. The voxel sums are the same as the reference's up to summation order.
constexpr int kBoxCap = 1024;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ void corner_weights(const Cell &c, double *f) {
    const double cx = 1. - c.fx, cy = 1. - c.fy, cz = 1. - c.fz;
    const double fx = c.fx, fy = c.fy, fz = c.fz;
    f[0] = cx * cy * cz; f[1] = cx * cy * fz; f[2] = cx * fy * cz; f[3] = cx * fy * fz;
    f[4] = fx * cy * cz; f[5] = fx * cy * fz; f[6] = fx * fy * cz; f[7] = fx * fy * fz;
}

template <bool kModel, bool kWeight>
__global__ __launch_bounds__(kBlock) void merge_box_kernel(
    const double *__restrict__ rots, int r0, int C, int num_rot, const double *__restrict__ view,
    const double *__restrict__ psum_r, int all_on, const double *__restrict__ qvals,
    const double *__restrict__ corr, const uint8_t *__restrict__ mask, int num_pix, long size,
    double *__restrict__ model2_all, double *__restrict__ weight_all, int modes) {
    constexpr int kPer = (kModel ? kBoxCap : 0) + (kWeight ? kBoxCap : 0);
    __shared__ double s_rot[kRTile * 9];
    __shared__ int s_on[kRTile];
    __shared__ double s_box[kWavesPerBlock][kPer];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = blockIdx.x * kRTile;
    for (int i = tid; i < kRTile * 9; i += kBlock) s_rot[i] = rots[(size_t)(r0 + c0) * 9 + i];
    if (tid < kRTile) {
        const int r = r0 + c0 + tid;
        s_on[tid] = (r < num_rot) && (all_on || psum_r[r] > 0.);
    }
    __syncthreads();
    const int t = (blockIdx.y * kWavesPerBlock + wave) * 64 + lane;
    const bool pix_ok = t < num_pix && mask[t] <= 1;
    double q0 = 0., q1 = 0., q2 = 0., cr = 1.;
    if (pix_ok) {
        q0 = qvals[3 * (size_t)t];
        q1 = qvals[3 * (size_t)t + 1];
        q2 = qvals[3 * (size_t)t + 2];
        cr = corr[t];
    }
    if (!__any(pix_ok)) return;
    double *bm = &s_box[wave][0];
    double *bw = bm + (kModel ? kBoxCap : 0);
    const double center = (double)(size / 2);
    const long s2 = size * size;
    for (int cc = 0; cc < kRTile; ++cc) {
        if (!s_on[cc]) continue;
        // the volume of this orientation's mode (slice_merge3d, model.c:175-176)
        const size_t moff = modes > 1 ? (size_t)((r0 + c0 + cc) % modes) * (size_t)(size * s2) : 0;
        double *model2 = kModel ? model2_all + moff : nullptr;
        double *weight = kWeight ? weight_all + moff : nullptr;
        Cell cell;
        const bool ok = pix_ok && locate(&s_rot[cc * 9], q0, q1, q2, center, size, cell);
        double w = 0.;
        if (kModel && ok) w = view[(size_t)(c0 >> 6) * ((size_t)num_pix * kRTile) + (size_t)t * kRTile + cc] / cr;
        const int xmax = wave_max(ok ? (int)cell.x : -1);
        if (xmax < 0) continue;
        const int xmin = wave_min(ok ? (int)cell.x : INT_MAX);
        const int ymin = wave_min(ok ? (int)cell.y : INT_MAX), ymax = wave_max(ok ? (int)cell.y : -1);
        const int zmin = wave_min(ok ? (int)cell.z : INT_MAX), zmax = wave_max(ok ? (int)cell.z : -1);
        const int dx = xmax - xmin + 2, dy = ymax - ymin + 2, dz = zmax - zmin + 2;
        const int bvol = dx * dy * dz;
        double f[8];
        if (ok) corner_weights(cell, f);
        if (bvol <= kBoxCap) {
            for (int i = lane; i < bvol; i += 64) {
                if (kModel) bm[i] = 0.;
                if (kWeight) bw[i] = 0.;
            }
            wave_sync();
            if (ok) {
                const int bx = (int)cell.x - xmin, by = (int)cell.y - ymin, bz = (int)cell.z - zmin;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int li = ((bx + (k >> 2)) * dy + (by + ((k >> 1) & 1))) * dz + bz + (k & 1);
                    if (kWeight) unsafeAtomicAdd(&bw[li], f[k]);
                    if (kModel) unsafeAtomicAdd(&bm[li], f[k] * w);
                }
            }
            wave_sync();
            for (int i = lane; i < bvol; i += 64) {
                const int iz = i % dz, iy = (i / dz) % dy, ix = i / (dz * dy);
                const long g = (long)(xmin + ix) * s2 + (long)(ymin + iy) * size + (zmin + iz);
                if (kModel) {
                    const double v = bm[i];
                    if (v != 0.) unsafeAtomicAdd(&model2[g], v);
                }
                if (kWeight) {
                    const double v = bw[i];
                    if (v != 0.) unsafeAtomicAdd(&weight[g], v);
                }
            }
            wave_sync();
        } else if (ok) {
            const long x1 = cell.x + 1, y1 = cell.y + 1, z1 = cell.z + 1;
            const long idx[8] = {cell.x * s2 + cell.y * size + cell.z, cell.x * s2 + cell.y * size + z1,
                                 cell.x * s2 + y1 * size + cell.z, cell.x * s2 + y1 * size + z1,
                                 x1 * s2 + cell.y * size + cell.z, x1 * s2 + cell.y * size + z1,
                                 x1 * s2 + y1 * size + cell.z, x1 * s2 + y1 * size + z1};
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (kWeight) unsafeAtomicAdd(&weight[idx[k]], f[k]);
                if (kModel) unsafeAtomicAdd(&model2[idx[k]], f[k] * w);
            }
        }
    }
}

// count of orientations with psum_r <= 0 (they are not merged, max_update.c:34-36); one block
__global__ __launch_bounds__(kBlock) void count_unmerged_kernel(const double *__restrict__ psum_r,
                                                                int num_rot, int *__restrict__ count) {
    __shared__ int s[kBlock];
    int local = 0;
    for (int r = threadIdx.x; r < num_rot; r += kBlock) local += !(psum_r[r] > 0.);
    s[threadIdx.x] = local;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = s[0];
}

// ---- single-view geometry (Model.slice_gen / slice_merge); perm maps the engine's pixel
// order to the caller's --------------------------------------------------------------------
__global__ void slice_gen_kernel(const double *__restrict__ rot, const double *__restrict__ model,
                                 long size, const double *__restrict__ qvals,
                                 const double *__restrict__ corr, const int *__restrict__ perm,
                                 int num_pix, double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_pix) return;
    out[perm[t]] = slice_value(model, size, rot, qvals[3 * (size_t)t], qvals[3 * (size_t)t + 1],
                               qvals[3 * (size_t)t + 2], corr[t], (double)(size / 2));
}

__global__ void slice_merge_kernel(const double *__restrict__ rot, const double *__restrict__ view,
                                   const double *__restrict__ qvals,
                                   const double *__restrict__ corr,
                                   const uint8_t *__restrict__ mask, const int *__restrict__ perm,
                                   int num_pix, long size, double *__restrict__ model2,
                                   double *__restrict__ weight) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_pix || mask[t] > 1) return;
    Cell cell;
    if (!locate(rot, qvals[3 * (size_t)t], qvals[3 * (size_t)t + 1], qvals[3 * (size_t)t + 2],
                (double)(size / 2), size, cell))
        return;
    merge_value(model2, weight, size, cell, view[perm[t]] / corr[t]);
}

// ---- per-frame statistics and scale ------------------------------------------------------
// sums of likelihood and info over non-blacklisted frames (max_prob.c:367-374).
__global__ __launch_bounds__(kBlock) void frame_stats_kernel(
    const double *__restrict__ likelihood, const double *__restrict__ info,
    const uint8_t *__restrict__ blacklist, int num_data, int denom, double *__restrict__ scalars) {
    __shared__ double s[kBlock];
    double a = 0., b = 0.;
    for (int d = threadIdx.x; d < num_data; d += kBlock)
        if (!blacklist[d]) {
            a += likelihood[d];
            b += info[d];
        }
    const double ta = block_sum(a, s);
    const double tb = block_sum(b, s);
    if (threadIdx.x == 0) {
        scalars[1] = ta / denom;
        scalars[2] = tb / denom;
    }
}

// update_scale: scale_d = fcounts_d / psum_d (max_scale.c:13-15).
__global__ void scale_kernel(const int *__restrict__ fcounts, const double *__restrict__ psum_d,
                             const uint8_t *__restrict__ blacklist, int num_data,
                             double *__restrict__ scale) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= num_data || blacklist[d]) return;
    scale[d] = fcounts[d] / psum_d[d];
}

__global__ void fill_kernel(double *__restrict__ a, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// ---- model update (EMCRecon.update_model, recon.pyx:112-149) ------------------------------
// Friedel symmetrisation, model.c:470-490. Voxel pairs with z < center are each visited once
// by the reference loop; pairs on the z == center plane are visited twice (once from each
// end, in x-major order), which is reproduced here so the result is the same bit for bit.
__global__ void friedel_offplane_kernel(double *__restrict__ a, const double *__restrict__ w,
                                        int size) {
    const int lo = (size % 2 == 0) ? 1 : 0, c = size / 2;
    const int n = size - lo, nz = c - lo;
    const long total = (long)n * n * nz;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x) {
        const int z = lo + (int)(i % nz);
        const int y = lo + (int)((i / nz) % n);
        const int x = lo + (int)(i / ((long)nz * n));
        const long v1 = (long)x * size * size + (long)y * size + z;
        const long v2 = (long)(2 * c - x) * size * size + (long)(2 * c - y) * size + (2 * c - z);
        double ai = 0.5 * (a[v1] * w[v1] + a[v2] * w[v2]);
        const double aw = 0.5 * (w[v1] + w[v2]);
        ai = (aw > 0) ? ai / aw : 0.;
        a[v1] = ai;
        a[v2] = ai;
    }
}

__global__ void friedel_plane_kernel(double *__restrict__ a, const double *__restrict__ w, int size) {
    const int lo = (size % 2 == 0) ? 1 : 0, c = size / 2;
    const int n = size - lo;
    const long total = (long)n * n;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x) {
        const int y = lo + (int)(i % n), x = lo + (int)(i / n);
        const int xm = 2 * c - x, ym = 2 * c - y;
        // this thread owns the pair if (x, y) is visited first
        if (xm < x || (xm == x && ym < y)) continue;
        const long v1 = (long)x * size * size + (long)y * size + c;
        const long v2 = (long)xm * size * size + (long)ym * size + c;
        double ai = 0.5 * (a[v1] * w[v1] + a[v2] * w[v2]);
        double aw = 0.5 * (w[v1] + w[v2]);
        ai = (aw > 0) ? ai / aw : 0.;
        a[v1] = ai;
        a[v2] = ai;
        if (v1 == v2) continue;
        ai = 0.5 * (a[v2] * w[v2] + a[v1] * w[v1]);
        aw = 0.5 * (w[v2] + w[v1]);
        ai = (aw > 0) ? ai / aw : 0.;
        a[v2] = ai;
        a[v1] = ai;
    }
}

// model2 /= weight where weight > 0; partial sums of (model2 - model1)^2; blend into model1.
__global__ __launch_bounds__(kBlock) void model_divide_blend_kernel(
    double *__restrict__ model1, double *__restrict__ model2, const double *__restrict__ w,
    long vol, double alpha, double *__restrict__ part) {
    __shared__ double s[kBlock];
    double acc = 0.;
    for (long x = blockIdx.x * (long)kBlock + threadIdx.x; x < vol; x += (long)gridDim.x * kBlock) {
        double m2 = model2[x];
        if (w[x] > 0.) m2 /= w[x];
        model2[x] = m2;
        const double diff = m2 - model1[x];
        acc += diff * diff;
        model1[x] = (alpha > 0.) ? alpha * model1[x] + (1 - alpha) * m2 : m2;
    }
    const double t = block_sum(acc, s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void sum_parts_kernel(const double *__restrict__ part, int n,
                                                           double *__restrict__ out) {
    __shared__ double s[kBlock];
    double acc = 0.;
    for (int i = threadIdx.x; i < n; i += kBlock) acc += part[i];
    const double t = block_sum(acc, s);
    if (threadIdx.x == 0) *out = t;
}

// Iterate.normalize_scale (iterate.pyx:361-384)
__global__ __launch_bounds__(kBlock) void scale_mean_kernel(const double *__restrict__ scale,
                                                            const uint8_t *__restrict__ blacklist,
                                                            int num_data, double *__restrict__ out) {
    __shared__ double s[kBlock];
    __shared__ double sc[kBlock];
    double acc = 0., cnt = 0.;
    for (int d = threadIdx.x; d < num_data; d += kBlock)
        if (!blacklist[d]) {
            acc += scale[d];
            cnt += 1.;
        }
    const double t = block_sum(acc, s);
    const double n = block_sum(cnt, sc);
    if (threadIdx.x == 0) *out = t / n;
}

__global__ void scale_divide_kernel(double *__restrict__ scale, const uint8_t *__restrict__ blacklist,
                                    int num_data, const double *__restrict__ mean) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < num_data && !blacklist[d]) scale[d] /= *mean;
}

__global__ void volume_scale_kernel(double *__restrict__ a, long n, const double *__restrict__ f) {
    const double v = *f;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        a[i] *= v;
}

}  // namespace emc

// ---- launchers (host) --------------------------------------------------------------------
namespace emc {

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }
static inline unsigned grid_cap(size_t blocks) { return (unsigned)(blocks < 65536 ? blocks : 65536); }
static inline int round_up_int(int a, int b) { return (a + b - 1) / b * b; }

void launch_expand(hipStream_t s, const double *rots, const double *model, long size,
                   const double *qvals, const double *corr, const uint8_t *mask, int num_pix,
                   int r0, int C, int pix_per_group, double *logw, double *upart, int *ubad,
                   const int *slot_rot, int modes, int raw) {
    expand_kernel<double, kRTile><<<expand_grid(C, num_pix, pix_per_group), kBlock, 0, s>>>(rots, model, size, qvals, corr, mask, num_pix, r0, C,
                                                          pix_per_group, logw, upart, ubad, slot_rot, modes, raw);
}

void launch_expand(hipStream_t s, const double *rots, const float *model, long size,
                   const double *qvals, const double *corr, const uint8_t *mask, int num_pix,
                   int r0, int C, int pix_per_group, float *logw, double *upart, int *ubad,
                   const int *slot_rot, int modes) {
    expand_kernel<float, 2 * kRTile, float><<<expand_grid(C, num_pix, pix_per_group), kBlock, 0, s>>>(rots, model, size, qvals, corr, mask, num_pix,
                                                                    r0, C, pix_per_group, logw, upart, ubad,
                                                                    slot_rot, modes, 0);
}

__global__ void copy_if_zero_kernel(double *__restrict__ dst, const double *__restrict__ src, size_t n,
                                    const int *__restrict__ gate) {
    if (*gate != 0) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy_if_zero(hipStream_t s, double *dst, const double *src, size_t n, const int *gate) {
    if (n == 0) return;
    copy_if_zero_kernel<<<grid_cap(cdiv(n, 256)), 256, 0, s>>>(dst, src, n, gate);
}

__global__ void copy_f32_kernel(const double *__restrict__ in, size_t n, float *__restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

void launch_copy_f32(hipStream_t s, const double *in, size_t n, float *out) {
    if (n == 0) return;
    copy_f32_kernel<<<grid_cap(cdiv(n, 256)), 256, 0, s>>>(in, n, out);
}

// FP32 brick copy of `modes` volumes of size^3 (halo_index layout: every brick holds the 5^3
// corners of its 4^3 cells; corners past size - 1 and the padding are 0 and never read): one
// thread per output element.
__global__ void copy_f32_brick_kernel(const double *__restrict__ in, long size, int modes, float *__restrict__ out) {
    const int nb = model_bricks(size);
    const size_t per_mode = (size_t)nb * nb * nb * kHaloBrickFloats, n = per_mode * modes;
    const size_t vol = (size_t)size * size * size;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t m = i / per_mode, j = i % per_mode, b = j / kHaloBrickFloats;
        const int l = (int)(j % kHaloBrickFloats);
        float v = 0.f;
        if (l < 125) {
            const long x = (long)(b / ((size_t)nb * nb)) * 4 + l / 25;
            const long y = (long)((b / nb) % nb) * 4 + (l / 5) % 5;
            const long z = (long)(b % nb) * 4 + l % 5;
            if (x < size && y < size && z < size) v = (float)in[m * vol + (size_t)(x * size + y) * size + z];
        }
        out[i] = v;
    }
}

size_t model_brick_elems(long size, int modes) {
    const size_t nb = (size_t)model_bricks(size);
    return nb * nb * nb * kHaloBrickFloats * (size_t)modes;
}

void launch_copy_f32_brick(hipStream_t s, const double *in, long size, int modes, float *out) {
    const size_t n = model_brick_elems(size, modes);
    if (n == 0) return;
    copy_f32_brick_kernel<<<grid_cap(cdiv(n, 256)), 256, 0, s>>>(in, size, modes, out);
}

// FP32 copy of one orientation chunk of the tile-major posterior in 128-wide tiles:
// out[(c / 128) * N * 128 + d * 128 + c % 128] = P[((r0 + c) / 64) * N * 64 + d * 64 + c % 64]
// for chunk orientations c < C (r0 a multiple of 64; columns of the last tile past C are 0).
__global__ void to_f32_w128_kernel(const double *__restrict__ P, int num_data, int r0, int C,
                                   float *__restrict__ out) {
    const size_t n = (size_t)num_data * ((C + 127) / 128) * 128;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int cw = (int)(i % 128);
        const size_t rest = i / 128;
        const int d = (int)(rest % num_data);
        const int c = (int)(rest / num_data) * 128 + cw;
        out[i] = c < C ? (float)P[(size_t)((r0 + c) / kRTile) * num_data * kRTile + (size_t)d * kRTile + (c % kRTile)]
                       : 0.f;
    }
}

void launch_to_f32(hipStream_t s, const double *P, int num_data, int r0, int C, float *out) {
    const size_t n = (size_t)num_data * ((C + 127) / 128) * 128;
    to_f32_w128_kernel<<<grid_cap(cdiv(n, 256)), 256, 0, s>>>(P, num_data, r0, C, out);
}

// Likelihood with FP32 log W in 128-wide tiles: one wavefront = one frame x 128 orientations,
// every lane reads two neighbouring orientations (one 8-byte load, 512 bytes per row) and
// keeps two FP64 sums; writes the 64-wide tile-major G of orientations < C.
template <int U>
__global__ __launch_bounds__(kBlock) void likelihood_w128_kernel(
    const float *__restrict__ logw, int C, size_t tsW, const long *__restrict__ off,
    const int *__restrict__ pix, const double *__restrict__ cnt, const uint8_t *__restrict__ blacklist,
    int num_data, double *__restrict__ G, size_t tsG, int r0, const int *__restrict__ fmap) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int d = blockIdx.x * kWavesPerBlock + wave;
    if (d >= num_data) return;
    const int fd = fmap ? fmap[d] : d;  // frame whose photons row d holds
    if (blacklist[fd]) return;
    const int tile = blockIdx.y;
    const float2 *__restrict__ col = reinterpret_cast<const float2 *>(logw + (size_t)tile * tsW) + lane;
    const long k1 = off[fd + 1];
    long k = off[fd];
    double a0 = 0., a1 = 0.;
    for (; k + U <= k1; k += U) {
        float2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = col[(size_t)pix[k + u] * 64];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a0 += cnt[k + u] * (double)v[u].x;
            a1 += cnt[k + u] * (double)v[u].y;
        }
    }
    for (; k < k1; ++k) {
        const float2 v = col[(size_t)pix[k] * 64];
        a0 += cnt[k] * (double)v.x;
        a1 += cnt[k] * (double)v.y;
    }
    const int c = tile * 128 + 2 * lane;
    if (c < C) {
        double2 *dst = reinterpret_cast<double2 *>(G + (size_t)((r0 + c) / kRTile) * tsG + (size_t)d * kRTile + (c % kRTile));
        *dst = make_double2(a0, a1);
    }
}

// Frame-partitioned tomogram over an FP32 posterior in 128-wide tiles (see to_f32_w128_kernel):
// each lane gathers two orientations per photon; partial sums in FP64, written in the 64-wide
// tile-major partial layout of tomogram_part_kernel.
template <int U, typename OT>
__global__ __launch_bounds__(kBlock) void tomogram_part_w128_kernel(
    const float *__restrict__ P32, size_t tsP, int C, const long *__restrict__ toff2, int nparts,
    const int *__restrict__ tframe, const double *__restrict__ tcnt, int num_pix, OT *__restrict__ partial) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int part = blockIdx.x % nparts;
    const int t = (blockIdx.x / nparts) * kWavesPerBlock + wave;
    if (t >= num_pix) return;
    const int tile = blockIdx.y;
    const float2 *__restrict__ col = reinterpret_cast<const float2 *>(P32 + (size_t)tile * tsP) + lane;
    const size_t at = (size_t)t * nparts + part;
    long k = toff2[at];
    const long k1 = toff2[at + 1];
    double a0 = 0., a1 = 0.;
    for (; k + U <= k1; k += U) {
        float2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = col[(size_t)tframe[k + u] * 64];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double c = tcnt[k + u];
            if ((double)v[u].x > kProbMin) a0 += c * (double)v[u].x;
            if ((double)v[u].y > kProbMin) a1 += c * (double)v[u].y;
        }
    }
    for (; k < k1; ++k) {
        const float2 v = col[(size_t)tframe[k] * 64];
        if ((double)v.x > kProbMin) a0 += tcnt[k] * (double)v.x;
        if ((double)v.y > kProbMin) a1 += tcnt[k] * (double)v.y;
    }
    const int c = tile * 128 + 2 * lane;
    if (c < C) {
        const size_t tsW = (size_t)num_pix * kRTile, tsPart = tsW * (C / kRTile);
        OT *dst = partial + part * tsPart + (size_t)(c / kRTile) * tsW + (size_t)t * kRTile + (c % kRTile);
        dst[0] = (OT)a0;
        dst[1] = (OT)a1;
    }
}

void launch_u_reduce(hipStream_t s, const double *upart, const int *ubad, int ngroups, int C,
                     int r0, int num_rot, double *u, int *rbad) {
    u_reduce_kernel<<<cdiv(C, 256), 256, 0, s>>>(upart, ubad, ngroups, C, r0, num_rot, u, rbad);
}

void launch_rescale(hipStream_t s, const double *quats, const double *u, int num_rot,
                    double mean_count, double *scalars, int modes) {
    rescale_kernel<<<1, kBlock, 0, s>>>(quats, u, num_rot, mean_count, (double)modes, scalars);
}

template <typename T>
static void likelihood_launch(hipStream_t s, const T *logw, int C, int num_pix, const long *off,
                              const int *pix, const double *cnt, const uint8_t *blacklist, int num_data,
                              double *G, int r0, int tiles, int unroll, const int *fmap = nullptr,
                              int xcd_order = 0) {
    const unsigned gx = cdiv(num_data, kWavesPerBlock);
    const size_t tsW = (size_t)num_pix * kRTile, tsG = (size_t)num_data * kRTile;
    /* XCDs sharing a tile in the 1-D order: 1, 2, 4 or 8 (0: 2-D grid) */
    const int xg = xcd_order >= 8 ? 8 : xcd_order >= 4 ? 4 : xcd_order >= 2 ? 2 : (xcd_order > 0 ? 1 : 0);
#define EMC_LIK(NT, U)                                                                           \
    do {                                                                                         \
        const size_t gy = cdiv(C, kRTile * NT);                                                  \
        const size_t nblk = xg ? (size_t)kNumXcd * cdiv(gx, xg) * cdiv(gy, kNumXcd / xg) : 0;    \
        /* the 1-D order needs its work-item count to fit the 32-bit grid size */                \
        const int xo = xg && nblk * kBlock < ((size_t)1 << 32);                                  \
        const dim3 grid = xo ? dim3((unsigned)nblk) : dim3(gx, (unsigned)gy);                    \
        likelihood_kernel<NT, U, T><<<grid, kBlock, 0, s>>>(logw, C, tsW, off, pix, cnt, blacklist, num_data, G, \
                                                            tsG, r0, fmap, xo ? xg : 0);        \
    } while (0)
    if (tiles >= 4) EMC_LIK(4, 2);
    else if (tiles == 2) EMC_LIK(2, 4);
    else if (unroll >= 16) EMC_LIK(1, 16);
    else if (unroll == 4) EMC_LIK(1, 4);
    else EMC_LIK(1, 8);
#undef EMC_LIK
}

void launch_likelihood(hipStream_t s, const double *logw, int C, int num_pix, const long *off,
                       const int *pix, const double *cnt, const uint8_t *blacklist, int num_data,
                       double *G, int r0, int tiles, int unroll, int xcd_order) {
    likelihood_launch(s, logw, C, num_pix, off, pix, cnt, blacklist, num_data, G, r0, tiles, unroll, nullptr,
                      xcd_order);
}

void launch_likelihood_split(hipStream_t s, const double *logw, int C, int num_pix, const long *off8, int X,
                             const int *pix, const double *cnt, const uint8_t *blacklist, int num_data,
                             double *part, double *G, int r0) {
    if (num_data <= 0 || C <= 0) return;
    const int nfg = (int)cdiv(num_data, kWavesPerBlock);
    // the range index varies fastest and the row of blocks is a multiple of the XCD count, so
    // block (x, fg) lands on an XCD of group x (x = XCD index when X = kNumXcd)
    const unsigned gx = (unsigned)(X * round_up_int(nfg, kNumXcd));
    const int ntiles = C / kRTile;
    const size_t tsW = (size_t)num_pix * kRTile, n = (size_t)ntiles * num_data * kRTile;
    likelihood_split_kernel<8><<<dim3(gx, ntiles), kBlock, 0, s>>>(logw, tsW, off8, X, pix, cnt, blacklist, num_data,
                                                                  nfg, part, n);
    likelihood_split_reduce_kernel<<<grid_cap(cdiv(n, 256)), 256, 0, s>>>(
        part, X, n, G + (size_t)(r0 / kRTile) * num_data * kRTile);
}

void launch_likelihood_mapped(hipStream_t s, const double *logw, int C, int num_pix, const long *off,
                              const int *pix, const double *cnt, const uint8_t *blacklist, const int *fmap,
                              int num_rows, double *G, int r0) {
    if (num_rows > 0) likelihood_launch(s, logw, C, num_pix, off, pix, cnt, blacklist, num_rows, G, r0, 1, 8, fmap);
}

static void likelihood_w128_launch(hipStream_t s, const float *logw, int C, int num_pix, const long *off,
                                   const int *pix, const double *cnt, const uint8_t *blacklist, int num_data,
                                   double *G, int r0, int unroll, const int *fmap) {
    const unsigned gx = cdiv(num_data, kWavesPerBlock);
    const size_t tsW = (size_t)num_pix * 128, tsG = (size_t)num_data * kRTile;
    const dim3 grid(gx, cdiv(C, 128));
    if (unroll >= 16)
        likelihood_w128_kernel<16><<<grid, kBlock, 0, s>>>(logw, C, tsW, off, pix, cnt, blacklist, num_data, G, tsG, r0,
                                                           fmap);
    else
        likelihood_w128_kernel<8><<<grid, kBlock, 0, s>>>(logw, C, tsW, off, pix, cnt, blacklist, num_data, G, tsG, r0,
                                                          fmap);
}

void launch_likelihood(hipStream_t s, const float *logw, int C, int num_pix, const long *off,
                       const int *pix, const double *cnt, const uint8_t *blacklist, int num_data,
                       double *G, int r0, int tiles, int unroll) {
    (void)tiles;
    likelihood_w128_launch(s, logw, C, num_pix, off, pix, cnt, blacklist, num_data, G, r0, unroll, nullptr);
}

void launch_likelihood_mapped(hipStream_t s, const float *logw, int C, int num_pix, const long *off,
                              const int *pix, const double *cnt, const uint8_t *blacklist, const int *fmap,
                              int num_rows, double *G, int r0) {
    if (num_rows > 0) likelihood_w128_launch(s, logw, C, num_pix, off, pix, cnt, blacklist, num_rows, G, r0, 16, fmap);
}

void launch_frame_max(hipStream_t s, double *R, int num_rot, int num_data,
                      const double *logwr, const double *u, const double *scale, int use_scale,
                      const double *scalars, const uint8_t *blacklist, const uint8_t *dense,
                      const int *rbad, const int *qmap, const uint32_t *relbits, int words,
                      double *max_exp, int *rmax) {
    frame_max_kernel<<<num_data, kBlock, 0, s>>>(R, (size_t)num_data * kRTile, num_rot, logwr, u, scale, use_scale,
                                                 scalars, blacklist, dense, rbad, qmap, relbits, words,
                                                 max_exp, rmax);
}

void launch_frame_norm(hipStream_t s, const double *R, int num_rot, int num_data,
                       const double *beta, const uint8_t *blacklist, const double *max_exp,
                       const int *rmax, double *p_norm) {
    frame_norm_kernel<<<num_data, kBlock, 0, s>>>(R, (size_t)num_data * kRTile, num_rot, beta, blacklist, max_exp, rmax,
                                                  p_norm);
}

__global__ void subset_mass_kernel(const double *__restrict__ R, int num_data, const long *__restrict__ off,
                                   const int *__restrict__ idx, double *__restrict__ mass) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= num_data) return;
    double acc = 0.;
    for (long k = off[d]; k < off[d + 1]; ++k) {
        const int r = idx[k];
        acc += R[(size_t)(r / kRTile) * ((size_t)num_data * kRTile) + (size_t)d * kRTile + r % kRTile];
    }
    mass[d] = acc;
}

void launch_subset_mass(hipStream_t s, const double *R, int num_data, const long *off, const int *idx,
                        double *mass) {
    if (num_data <= 0) return;
    subset_mass_kernel<<<cdiv(num_data, 256), 256, 0, s>>>(R, num_data, off, idx, mass);
}

// Posterior mass a frame's fine search would put outside its cells, estimated from a coarse
// full search (every cell centre) against the fine maximum inside the cells. One wavefront
// per frame; lanes stride over the coarse orientations. With the coarse posterior
// Pc = exp(beta (Rc - Mc)) / pnc, a cell outside the frame's set is counted as n_c fine
// orientations at the centre's log-likelihood with the centre's weight w_c replaced by the
// cell's mean fine weight wbar_c:
//   Z_out = sum_c n_c (wbar_c / w_c)^beta exp(beta (Rc - M)) = pnc exp(beta (Mc - M)) sum_c Pc g_c
// with g_c = exp(log_n[c] + beta * log_wr[c]); the fine search inside holds Z_in = p_norm
// (relative to its maximum M), and the estimate is Z_out / (Z_in + Z_out). Coarse centres
// outside the coarse keep set (Pc = 0, more than 14 / beta below the coarse maximum) count 0.
// Frames with skip[d] (full frames, blacklisted frames) get 0. With prob_floor > 0 only the
// cells whose estimated posterior per fine orientation (the centre's term over Z_in + Z_out)
// exceeds prob_floor are counted in the numerator: the mass outside that is spread thinner
// than prob_floor per orientation is mass the tomogram drops anyway (P <= 1e-6,
// max_update.c:119-120, 196).
__global__ __launch_bounds__(kBlock) void judge_outside_kernel(
    const double *__restrict__ Pc, int num_data, int num_coarse, const double *__restrict__ Mc,
    const double *__restrict__ pnc, const double *__restrict__ Mf, const double *__restrict__ pnf,
    const double *__restrict__ beta, const uint8_t *__restrict__ skip, const uint32_t *__restrict__ bits,
    int words, const double *__restrict__ log_n, const double *__restrict__ log_wr, double prob_floor,
    double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int d = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (d >= num_data) return;
    if (skip[d]) {
        if (lane == 0) out[d] = 0.;
        return;
    }
    const double b = beta[d];
    const uint32_t *row = bits + (size_t)d * words;
    auto pc = [&](int c) {
        return Pc[(size_t)(c / kRTile) * ((size_t)num_data * kRTile) + (size_t)d * kRTile + c % kRTile];
    };
    double acc = 0.;
    for (int c = lane; c < num_coarse; c += 64) {
        if ((row[c >> 5] >> (c & 31)) & 1u) continue;
        const double p = pc(c);
        if (p > 0.) acc += p * exp(log_n[c] + b * log_wr[c]);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (!(acc > 0.)) {
        if (lane == 0) out[d] = 0.;
        return;
    }
    // Z_out = exp(lz) with the coarse sum acc; the share outside is Z_out / (Z_in + Z_out)
    const double lz0 = log(pnc[d]) + b * (Mc[d] - Mf[d]);  // log of Z_out / acc
    const double lz = lz0 + log(acc);
    const double lr = log(pnf[d]) - lz;  // log Z_in / Z_out
    if (lr > 700.) {
        if (lane == 0) out[d] = 0.;
        return;
    }
    const double share = 1. / (1. + exp(lr));
    double v = share;
    if (prob_floor > 0.) {
        // log (Z_in + Z_out): the normaliser of the per-orientation estimates
        const double lzt = lz + log1p(exp(lr));
        double sig = 0.;
        for (int c = lane; c < num_coarse; c += 64) {
            if ((row[c >> 5] >> (c & 31)) & 1u) continue;
            const double p = pc(c);
            if (!(p > 0.)) continue;
            const double term = p * exp(log_n[c] + b * log_wr[c]);    // the cell's share of acc
            const double per = exp(lz0 + log(term) - log_n[c] - lzt);  // per fine orientation
            if (per > prob_floor) sig += term;
        }
        for (int o = 32; o > 0; o >>= 1) sig += __shfl_xor(sig, o);
        v = share * (sig / acc);
    }
    if (lane == 0) out[d] = v;
}

// For the frames judge_outside_kernel (prob_floor 0) puts above outside_max: the fewest coarse
// cells outside the frame's set after whose addition the estimate, still with the fine
// maximum and normaliser of the set alone (both only grow with more cells, so the estimate is
// an upper bound), is at most target -- the cells of largest share of Z_out, found by
// bisection on the share. One wavefront per frame, a lane holding cells lane + 64 j
// (num_coarse <= 64 * kWidenPerLane). out: the estimate; num_add: 0 (kept, skipped, or no mass
// outside), k (add[d * max_add + i], i < k, cells ascending), -1 (more than max_add cells, or
// a non-finite normaliser: the full search).
constexpr int kWidenPerLane = 64;
static_assert(64 * kWidenPerLane == kWidenCells, "a wavefront's cells");
__global__ __launch_bounds__(kBlock) void judge_widen_kernel(
    const double *__restrict__ Pc, int num_data, int num_coarse, const double *__restrict__ Mc,
    const double *__restrict__ pnc, const double *__restrict__ Mf, const double *__restrict__ pnf,
    const double *__restrict__ beta, const uint8_t *__restrict__ skip, const uint32_t *__restrict__ bits,
    int words, const double *__restrict__ log_n, const double *__restrict__ log_wr, double outside_max, double target,
    int max_add, double *__restrict__ out, int *__restrict__ num_add, int *__restrict__ add) {
    const int lane = threadIdx.x & 63;
    const int d = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (d >= num_data) return;
    if (skip[d]) {
        if (lane == 0) out[d] = 0., num_add[d] = 0;
        return;
    }
    const double b = beta[d];
    const uint32_t *row = bits + (size_t)d * words;
    // each outside cell's share of Z_out (up to the common factor), as judge_outside_kernel sums it
    double acc = 0.;
    float lt[kWidenPerLane];  // its log (-inf: inside the set or no posterior)
#pragma unroll
    for (int j = 0; j < kWidenPerLane; ++j) {
        const int c = lane + 64 * j;
        lt[j] = -INFINITY;
        if (c < num_coarse && !((row[c >> 5] >> (c & 31)) & 1u)) {
            const double p = Pc[(size_t)(c / kRTile) * ((size_t)num_data * kRTile) + (size_t)d * kRTile + c % kRTile];
            if (p > 0.) {
                const double term = p * exp(log_n[c] + b * log_wr[c]);
                acc += term;
                lt[j] = (float)log(term);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (!(acc > 0.)) {
        if (lane == 0) out[d] = 0., num_add[d] = 0;
        return;
    }
    const double lz0 = log(pnc[d]) + b * (Mc[d] - Mf[d]);  // log of Z_out / acc
    const double lr = log(pnf[d]) - (lz0 + log(acc));    // log Z_in / Z_out
    const double share = lr > 700. ? 0. : 1. / (1. + exp(lr));
    if (lane == 0) out[d] = share;
    if (!(share > outside_max)) {
        if (lane == 0) num_add[d] = 0;
        return;
    }
    // the remaining sum allowed: lz0 + log(rem) <= log(target / (1 - target)) + log Z_in
    const double lallow = log(target / (1. - target)) + log(pnf[d]) - lz0;
    float lmax = -INFINITY, lmin = INFINITY;
#pragma unroll
    for (int j = 0; j < kWidenPerLane; ++j)
        if (lt[j] > -INFINITY) lmax = fmaxf(lmax, lt[j]), lmin = fminf(lmin, lt[j]);
    for (int o = 32; o > 0; o >>= 1) {
        lmax = fmaxf(lmax, __shfl_xor(lmax, o));
        lmin = fminf(lmin, __shfl_xor(lmin, o));
    }
    if (!isfinite(lallow) || !isfinite(lmax)) {
        if (lane == 0) num_add[d] = -1;
        return;
    }
    // shares relative to the largest: s = lt - lmax <= 0; find the largest cut t with
    // sum over s < t of exp(s) <= exp(lallow - lmax); cells with s >= t are added
    const float g = (float)(lallow - lmax);
    float lo = lmin - lmax - 1.f, hi = 0.5f;  // sum below lo: 0 (fits); below hi: everything (does not)
    for (int it = 0; it < 28; ++it) {
        const float mid = 0.5f * (lo + hi);
        float f = 0.f;
#pragma unroll
        for (int j = 0; j < kWidenPerLane; ++j) {
            const float sj = lt[j] - lmax;
            if (sj < mid) f += __expf(sj);  // -inf entries add 0
        }
        for (int o = 32; o > 0; o >>= 1) f += __shfl_xor(f, o);
        if (f > 0.f && __logf(f) > g)
            hi = mid;
        else
            lo = mid;
    }
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < kWidenPerLane; ++j) cnt += (lt[j] > -INFINITY && lt[j] - lmax >= lo) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if (cnt > max_add) {
        if (lane == 0) num_add[d] = -1;
        return;
    }
    int base = 0;
#pragma unroll
    for (int j = 0; j < kWidenPerLane; ++j) {
        const bool take = lt[j] > -INFINITY && lt[j] - lmax >= lo;
        const unsigned long long m = __ballot(take);
        if (take) add[(size_t)d * max_add + base + __popcll(m & ((1ull << lane) - 1ull))] = lane + 64 * j;
        base += __popcll(m);
    }
    if (lane == 0) num_add[d] = cnt;
}

// out[off[d] + i] = add[d * max_add + i] for i < num_add[d] (> 0)
__global__ void widen_compact_kernel(const int *__restrict__ add, const int *__restrict__ num_add,
                                     const long *__restrict__ off, int num_data, int max_add, int *__restrict__ out) {
    const int d = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (d >= num_data) return;
    const int n = num_add[d];
    for (int i = lane; i < n; i += 64) out[off[d] + i] = add[(size_t)d * max_add + i];
}

void launch_judge_widen(hipStream_t s, const double *Pc, int num_data, int num_coarse, const double *Mc,
                        const double *pnc, const double *Mf, const double *pnf, const double *beta, const uint8_t *skip,
                        const uint32_t *bits, int words, const double *log_n, const double *log_wr, double outside_max,
                        double target, int max_add, double *out, int *num_add, int *add) {
    if (num_data <= 0) return;
    judge_widen_kernel<<<cdiv(num_data, kWavesPerBlock), kBlock, 0, s>>>(Pc, num_data, num_coarse, Mc, pnc, Mf, pnf,
                                                                         beta, skip, bits, words, log_n, log_wr,
                                                                         outside_max, target, max_add, out, num_add,
                                                                         add);
}

void launch_widen_compact(hipStream_t s, const int *add, const int *num_add, const long *off, int num_data,
                          int max_add, int *out) {
    if (num_data <= 0) return;
    widen_compact_kernel<<<cdiv(num_data, kWavesPerBlock), kBlock, 0, s>>>(add, num_add, off, num_data, max_add, out);
}

void launch_judge_outside(hipStream_t s, const double *Pc, int num_data, int num_coarse, const double *Mc,
                          const double *pnc, const double *Mf, const double *pnf, const double *beta,
                          const uint8_t *skip, const uint32_t *bits, int words, const double *log_n,
                          const double *log_wr, double prob_floor, double *out) {
    if (num_data <= 0) return;
    judge_outside_kernel<<<cdiv(num_data, kWavesPerBlock), kBlock, 0, s>>>(
        Pc, num_data, num_coarse, Mc, pnc, Mf, pnf, beta, skip, bits, words, log_n, log_wr, prob_floor, out);
}

void launch_posterior(hipStream_t s, double *R, int num_rot, int num_data,
                      const double *logwr, const double *u, const double *beta,
                      const double *sum_fact, const uint8_t *blacklist, const double *max_exp,
                      const int *rmax, const double *p_norm, const double *scalars,
                      int want_psum_d, double *likelihood, double *info, double *psum_d, int modes) {
    posterior_kernel<<<num_data, kBlock, 0, s>>>(R, (size_t)num_data * kRTile, num_rot, logwr, u, beta, sum_fact,
                                                 blacklist, max_exp, rmax, p_norm, scalars,
                                                 want_psum_d, log((double)modes), likelihood, info, psum_d);
}

void launch_occupancy(hipStream_t s, const double *P, int num_rot, int num_data, int modes,
                      double *quat_norm) {
    occupancy_kernel<<<num_data, kBlock, 0, s>>>(P, (size_t)num_data * kRTile, num_rot, modes, quat_norm);
}

void launch_sparse_count(hipStream_t s, const double *P, int num_rot, int num_data, int *counts) {
    sparse_count_kernel<false><<<num_data, kBlock, 0, s>>>(P, (size_t)num_data * kRTile, num_rot, counts);
}

void launch_sparse_fill(hipStream_t s, const double *P, int num_rot, int num_data, int d0, int nd,
                        const long *off, int *place, double *prob) {
    if (nd <= 0) return;
    sparse_fill_kernel<false><<<nd, kBlock, 0, s>>>(P, (size_t)num_data * kRTile, num_rot, d0, off, place, prob);
}

void launch_sparse_count_rows(hipStream_t s, const double *P, int num_rot, int nrows, int *counts) {
    if (nrows > 0) sparse_count_kernel<true><<<nrows, kBlock, 0, s>>>(P, 0, num_rot, counts);
}

void launch_sparse_fill_rows(hipStream_t s, const double *P, int num_rot, int nrows, const long *off, int *place,
                             double *prob) {
    if (nrows > 0) sparse_fill_kernel<true><<<nrows, kBlock, 0, s>>>(P, 0, num_rot, 0, off, place, prob);
}

void launch_psum_r(hipStream_t s, const double *P, int num_rot, int num_data,
                   int frames_per_part, const double *scale, int need_scaling,
                   const uint8_t *blacklist, double *part, double *psum_r) {
    const int nparts = (int)cdiv(num_data, frames_per_part);
    dim3 grid(cdiv(num_rot, kBlock), nparts);
    psum_r_partial_kernel<<<grid, kBlock, 0, s>>>(P, (size_t)num_data * kRTile, num_rot, num_data, frames_per_part,
                                                  scale, need_scaling, blacklist, part);
    psum_r_reduce_kernel<<<cdiv(num_rot, 256), 256, 0, s>>>(part, nparts, num_rot, psum_r);
}

void launch_tomogram(hipStream_t s, const double *P, int num_data, int r0, int C,
                     const long *toff, const int *tframe, const double *tcnt, const double *psum_r,
                     int num_rot, int num_pix, double *view, int tiles) {
    const unsigned gx = cdiv(num_pix, kWavesPerBlock);
    const size_t tsP = (size_t)num_data * kRTile, tsW = (size_t)num_pix * kRTile;
#define EMC_TOMO(NT)                                                                             \
    tomogram_kernel<NT><<<dim3(gx, cdiv(C, kRTile * NT)), kBlock, 0, s>>>(P, tsP, r0, C, tsW, toff, tframe, \
                                                                        tcnt, psum_r, num_rot, num_pix, view)
    if (tiles >= 4) EMC_TOMO(4);
    else if (tiles == 2) EMC_TOMO(2);
    else EMC_TOMO(1);
#undef EMC_TOMO
}

void launch_tomogram_tiles(hipStream_t s, const double *P, int num_data, int r0, int C,
                           const long *ft_off, const int *ft_frame, const long *fe_off,
                           const int *fe_code, const double *psum_r, int num_rot, int num_pix,
                           double *view) {
    dim3 grid(cdiv(num_pix, kTomoTile), C / kRTile);
    tomogram_tile_kernel<<<grid, kBlock, 0, s>>>(P, (size_t)num_data * kRTile, r0, C,
                                                 (size_t)num_pix * kRTile, ft_off, ft_frame, fe_off,
                                                 fe_code, psum_r, num_rot, num_pix, view);
}

void launch_tomogram_parts(hipStream_t s, const double *P, int num_data, int r0, int C,
                           const long *toff2, int nparts, const int *tframe, const double *tcnt,
                           const double *psum_r, int num_rot, int num_pix, double *partial,
                           double *view, int width32, int unroll, int seq) {
    if (width32) {
        dim3 grid(cdiv(num_pix, kWavesPerBlock) * nparts, C / (kRTile / 2));
        tomogram_part32_kernel<<<grid, kBlock, 0, s>>>(P, (size_t)num_data * kRTile, r0, C, toff2,
                                                       nparts, tframe, tcnt, num_pix, partial);
    } else {
        const int ngroups = (int)cdiv(num_pix, kWavesPerBlock);
        dim3 grid(ngroups * nparts, C / kRTile);
        const bool sq = seq && nparts % kNumXcd == 0;
#define EMC_TP(U, SQ)                                                                            \
    tomogram_part_kernel<U, SQ, double><<<grid, kBlock, 0, s>>>(P, (size_t)num_data * kRTile, r0, C, toff2, \
                                                                nparts, tframe, tcnt, num_pix, ngroups, partial)
        if (sq) {
            if (unroll >= 16) EMC_TP(16, true); else EMC_TP(8, true);
        } else {
            if (unroll >= 16) EMC_TP(16, false); else EMC_TP(8, false);
        }
#undef EMC_TP
    }
}

// FP32 posterior of one orientation chunk (128-wide tiles, to_f32_w128_kernel), FP64 sums.
void launch_tomogram_parts_f32(hipStream_t s, const float *P32, int num_data, int C, const long *toff2,
                               int nparts, const int *tframe, const double *tcnt, int num_pix,
                               double *partial) {
    const int ngroups = (int)cdiv(num_pix, kWavesPerBlock);
    dim3 grid(ngroups * nparts, cdiv(C, 128));
    tomogram_part_w128_kernel<16, double><<<grid, kBlock, 0, s>>>(P32, (size_t)num_data * 128, C, toff2, nparts,
                                                                  tframe, tcnt, num_pix, partial);
}

void launch_tomogram_parts_f32(hipStream_t s, const float *P32, int num_data, int C, const long *toff2,
                               int nparts, const int *tframe, const double *tcnt, int num_pix, float *partial) {
    const int ngroups = (int)cdiv(num_pix, kWavesPerBlock);
    dim3 grid(ngroups * nparts, cdiv(C, 128));
    tomogram_part_w128_kernel<16, float><<<grid, kBlock, 0, s>>>(P32, (size_t)num_data * 128, C, toff2, nparts,
                                                                 tframe, tcnt, num_pix, partial);
}

void launch_tomogram_reduce_rmajor(hipStream_t s, const double *partial, int nparts, int r0, int C,
                                   const double *psum_r, int num_rot, int num_pix, double *view) {
    dim3 grid(cdiv(num_pix, 64), C / kRTile);
    tomogram_reduce_rmajor_kernel<<<grid, kBlock, 0, s>>>(partial, nparts, r0, C, num_pix, psum_r, num_rot,
                                                          view);
}

void launch_tomogram_reduce(hipStream_t s, const double *partial, int nparts, int r0, int C,
                            const double *psum_r, int num_rot, int num_pix, double *view) {
    const size_t n = (size_t)num_pix * C;
    tomogram_reduce_kernel<<<grid_cap(cdiv(n, 256)), 256, 0, s>>>(partial, nparts, r0, C, num_pix,
                                                                  psum_r, num_rot, view);
}

void launch_merge(hipStream_t s, const double *rots, int r0, int C, int num_rot, const double *view,
                  const double *psum_r, int all_on, const double *qvals, const double *corr,
                  const uint8_t *mask, int num_pix, long size, double *model2, double *weight,
                  int do_model, int do_weight, int modes) {
    dim3 grid(C / kRTile, cdiv(num_pix, 64 * kWavesPerBlock));
#define EMC_MERGE(M, W)                                                                          \
    merge_box_kernel<M, W><<<grid, kBlock, 0, s>>>(rots, r0, C, num_rot, view, psum_r, all_on, qvals, \
                                                   corr, mask, num_pix, size, model2, weight, modes)
    if (do_model && do_weight) EMC_MERGE(true, true);
    else if (do_model) EMC_MERGE(true, false);
    else if (do_weight) EMC_MERGE(false, true);
#undef EMC_MERGE
}

void launch_count_unmerged(hipStream_t s, const double *psum_r, int num_rot, int *count) {
    count_unmerged_kernel<<<1, kBlock, 0, s>>>(psum_r, num_rot, count);
}

void launch_slice_gen(hipStream_t s, const double *rot, const double *model, long size,
                      const double *qvals, const double *corr, const int *perm, int num_pix,
                      double *out) {
    slice_gen_kernel<<<cdiv(num_pix, 256), 256, 0, s>>>(rot, model, size, qvals, corr, perm, num_pix, out);
}

void launch_slice_merge(hipStream_t s, const double *rot, const double *view, const double *qvals,
                        const double *corr, const uint8_t *mask, const int *perm, int num_pix,
                        long size, double *model2, double *weight) {
    slice_merge_kernel<<<cdiv(num_pix, 256), 256, 0, s>>>(rot, view, qvals, corr, mask, perm, num_pix,
                                                          size, model2, weight);
}

void launch_frame_stats(hipStream_t s, const double *likelihood, const double *info,
                        const uint8_t *blacklist, int num_data, int denom, double *scalars) {
    frame_stats_kernel<<<1, kBlock, 0, s>>>(likelihood, info, blacklist, num_data, denom, scalars);
}

void launch_scale(hipStream_t s, const int *fcounts, const double *psum_d, const uint8_t *blacklist,
                  int num_data, double *scale) {
    scale_kernel<<<cdiv(num_data, 256), 256, 0, s>>>(fcounts, psum_d, blacklist, num_data, scale);
}

void launch_fill(hipStream_t s, double *a, size_t n, double v) {
    fill_kernel<<<grid_cap(cdiv(n, 256)), 256, 0, s>>>(a, n, v);
}

void launch_update_model(hipStream_t s, double *model1, double *model2, const double *w, long size,
                         double alpha, double *part, int nparts, double *out_change, int modes) {
    const long vol = size * size * size;
    // Friedel symmetry per mode, then divide, change and blend over every mode's volume
    // (recon.pyx:125-149)
    for (int m = 0; m < modes; ++m) {
        double *m2 = model2 + (size_t)m * vol;
        const double *mw = w + (size_t)m * vol;
        friedel_offplane_kernel<<<grid_cap(cdiv(vol / 2 + 1, 256)), 256, 0, s>>>(m2, mw, (int)size);
        friedel_plane_kernel<<<cdiv((size_t)size * size, 256), 256, 0, s>>>(m2, mw, (int)size);
    }
    model_divide_blend_kernel<<<nparts, kBlock, 0, s>>>(model1, model2, w, vol * modes, alpha, part);
    sum_parts_kernel<<<1, kBlock, 0, s>>>(part, nparts, out_change);
}

void launch_normalize_scale(hipStream_t s, double *scale, const uint8_t *blacklist, int num_data,
                            double *model1, long vol, double *mean_out) {
    scale_mean_kernel<<<1, kBlock, 0, s>>>(scale, blacklist, num_data, mean_out);
    volume_scale_kernel<<<grid_cap(cdiv(vol, 256)), 256, 0, s>>>(model1, vol, mean_out);
    scale_divide_kernel<<<cdiv(num_data, 256), 256, 0, s>>>(scale, blacklist, num_data, mean_out);
}

}  // namespace emc
