// Per-frame orientation subsets evaluated sparsely (refinement, legacy src/max_emc.c:264-274).
//
// With refinement every frame is evaluated only at the orientations of its coarse cells. The
// engine then orders orientations by coarse cell and pads every cell to whole tiles of 64
// ("slots": slot s holds orientation slot_rot[s], or -1 for padding), so the orientations a
// frame may take are a union of whole slot tiles. A (frame, tile) pair is one row of 64
// values; the pairs of a frame are contiguous (pair_off[d] .. pair_off[d + 1]) and every
// kernel below works on pairs only, so the work and memory scale with the subsets' size, not
// with frames x orientations.
//
// The values are the plain path's: the same likelihood sums in the same photon order, the
// same first arg-max (ties broken by the original orientation index), the same keep rule
// (ordered by the original index), posterior, sums and tomogram. Only the order of additions
// across frames in psum_r and the tomogram differs from the reference.
#include <cfloat>

#include "emc_common.h"
#include "emc_kernels.h"

namespace emc {

// G[k][lane] = sum over the frame's photons of count * logW[tile - t0][pix][lane] for pairs
// whose tile lies in [t0, t1) (one orientation chunk); other pairs are left alone.
template <int U>
__global__ __launch_bounds__(kBlock) void pair_likelihood_kernel(
    const double *__restrict__ logw, size_t tsW, const long *__restrict__ off,
    const int *__restrict__ pix, const double *__restrict__ cnt, const int *__restrict__ pair_frame,
    const int *__restrict__ pair_tile, long num_pairs, int t0, int t1, double *__restrict__ G) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const long stride = (long)gridDim.x * kWavesPerBlock;
    for (long k = (long)blockIdx.x * kWavesPerBlock + wave; k < num_pairs; k += stride) {
        const int T = pair_tile[k];
        if (T < t0 || T >= t1) continue;
        const int d = pair_frame[k];
        const double *__restrict__ col = logw + (size_t)(T - t0) * tsW + lane;
        long j = off[d];
        const long j1 = off[d + 1];
        double acc = 0.;
        for (; j + U <= j1; j += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = col[(size_t)pix[j + u] * kRTile];
#pragma unroll
            for (int u = 0; u < U; ++u) acc += cnt[j + u] * v[u];
        }
        for (; j < j1; ++j) acc += cnt[j] * col[(size_t)pix[j] * kRTile];
        G[(size_t)k * kRTile + lane] = acc;
    }
}

// Per frame over its pairs: R = log w - u phi + G for real, allowed orientations (-inf for
// padding slots and, for dense frames, orientations with non-finite log W), maximum and first
// arg-max by original orientation index.
__global__ __launch_bounds__(kBlock) void pair_max_kernel(
    double *__restrict__ R, const long *__restrict__ pair_off, const int *__restrict__ pair_tile,
    const int *__restrict__ slot_rot, const double *__restrict__ logwr_slot,
    const double *__restrict__ u_slot, const int *__restrict__ rbad_slot,
    const double *__restrict__ scale, int use_scale, const double *__restrict__ scalars,
    const uint8_t *__restrict__ blacklist, const uint8_t *__restrict__ dense,
    double *__restrict__ max_exp, int *__restrict__ rmax) {
    __shared__ double s_v[kBlock];
    __shared__ int s_i[kBlock];
    const int d = blockIdx.x;
    if (blacklist[d]) {
        if (threadIdx.x == 0) {
            max_exp[d] = kMaxExpStart;
            rmax[d] = -1;
        }
        return;
    }
    const bool dense_frame = dense != nullptr && dense[d];
    const double phi = use_scale ? scale[d] : scalars[0];
    const long k0 = pair_off[d], k1 = pair_off[d + 1];
    double best = kMaxExpStart;
    int bi = -1;
    for (long e = k0 * kRTile + threadIdx.x; e < k1 * kRTile; e += kBlock) {
        const long k = e / kRTile;
        const int s = pair_tile[k] * kRTile + (int)(e % kRTile);
        const int r = slot_rot[s];
        double &x = R[e];
        double v = -INFINITY;
        if (r >= 0 && !(dense_frame && rbad_slot[s])) v = (logwr_slot[s] - u_slot[s] * phi) + x;
        x = v;
        if (v > best || (v == best && r >= 0 && bi >= 0 && r < bi)) {
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
            const double mv = s_v[threadIdx.x];
            const int mi = s_i[threadIdx.x];
            if (ov > mv || (ov == mv && oi >= 0 && (mi < 0 || oi < mi))) {
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

__device__ __forceinline__ bool pair_kept(double v, int r, int rf, double M, double c) {
    if (r == rf) return true;
    return r > rf ? (v + c > M) : (v > M - c);
}

__device__ __forceinline__ double block_sum_pairs(double v, double *s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

// p_norm over the frame's kept pairs (normalize_prob, max_prob.c:245-250).
__global__ __launch_bounds__(kBlock) void pair_norm_kernel(
    const double *__restrict__ R, const long *__restrict__ pair_off, const int *__restrict__ pair_tile,
    const int *__restrict__ slot_rot, const double *__restrict__ beta,
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
    double acc = 0.;
    for (long e = pair_off[d] * kRTile + threadIdx.x; e < pair_off[d + 1] * kRTile; e += kBlock) {
        const int r = slot_rot[pair_tile[e / kRTile] * kRTile + (int)(e % kRTile)];
        const double v = R[e];
        if (r >= 0 && pair_kept(v, r, rf, M, c)) acc += exp(b * (v - M));
    }
    const double t = block_sum_pairs(acc, s);
    if (threadIdx.x == 0) p_norm[d] = t;
}

// P in place, and the frame's likelihood, information and psum_d (calc_psum_r,
// max_update.c:42-140).
__global__ __launch_bounds__(kBlock) void pair_posterior_kernel(
    double *__restrict__ R, const long *__restrict__ pair_off, const int *__restrict__ pair_tile,
    const int *__restrict__ slot_rot, const double *__restrict__ logwr_slot,
    const double *__restrict__ u_slot, const double *__restrict__ beta,
    const double *__restrict__ sum_fact, const uint8_t *__restrict__ blacklist,
    const double *__restrict__ max_exp, const int *__restrict__ rmax, const double *__restrict__ p_norm,
    const double *__restrict__ scalars, int want_psum_d, double *__restrict__ likelihood,
    double *__restrict__ info, double *__restrict__ psum_d) {
    __shared__ double s[kBlock];
    const int d = blockIdx.x;
    const int rf = rmax[d];
    const long e0 = pair_off[d] * kRTile, e1 = pair_off[d + 1] * kRTile;
    if (blacklist[d] || rf < 0) {
        for (long e = e0 + threadIdx.x; e < e1; e += kBlock) R[e] = 0.;
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
    for (long e = e0 + threadIdx.x; e < e1; e += kBlock) {
        const int sl = pair_tile[e / kRTile] * kRTile + (int)(e % kRTile);
        const int r = slot_rot[sl];
        const double v = R[e];
        double prob = 0.;
        if (r >= 0 && pair_kept(v, r, rf, M, c)) {
            const double ex = b * (v - M);
            prob = exp(ex) / pn;
            lk += prob * (v - sf);
            ps += prob * u_slot[sl] * rescale;
            if (!(prob < kProbMin)) inf += prob * ((ex - log_pn) - logwr_slot[sl]);
        }
        R[e] = prob;
    }
    const double tl = block_sum_pairs(lk, s);
    const double ti = block_sum_pairs(inf, s);
    const double tp = block_sum_pairs(ps, s);
    if (threadIdx.x == 0) {
        likelihood[d] = tl;
        info[d] = ti;
        if (want_psum_d) psum_d[d] = tp;
    }
}

// psum of every slot of a tile: sum over the tile's pairs (frames ascending) of P * phi'
// (phi' = scale if need_scaling else 1; blacklisted frames have P = 0). One wave per tile.
__global__ __launch_bounds__(kBlock) void pair_psum_kernel(const double *__restrict__ P,
                                                           const long *__restrict__ tile_off,
                                                           const int *__restrict__ tile_pair,
                                                           const int *__restrict__ pair_frame, int num_tiles,
                                                           const double *__restrict__ scale, int need_scaling,
                                                           double *__restrict__ psum_slot) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int T = blockIdx.x * kWavesPerBlock + wave;
    if (T >= num_tiles) return;
    double acc = 0.;
    for (long j = tile_off[T]; j < tile_off[T + 1]; ++j) {
        const int k = tile_pair[j];
        const double p = P[(size_t)k * kRTile + lane];
        acc += need_scaling ? p * scale[pair_frame[k]] : p;
    }
    psum_slot[(size_t)T * kRTile + lane] = acc;
}

// Tomogram of the chunk's tiles [t0, t0 + ntiles) over pixel ranges of kPairRange pixels:
// block (tile, range) accumulates K * P (P > 1e-6) of the tile's pairs into LDS rows of the
// range's pixels (frame photon lists sorted by pixel, range offsets roff[d][range]), divides
// by psum and writes the orientation-major view[(T - t0) * 64 + j][t].
__global__ __launch_bounds__(kBlock) void pair_tomogram_kernel(
    const double *__restrict__ P, const long *__restrict__ tile_off, const int *__restrict__ tile_pair,
    const int *__restrict__ pair_frame, const long *__restrict__ foff, const int *__restrict__ roff,
    int nranges, const int *__restrict__ fpix, const double *__restrict__ fcnt,
    const double *__restrict__ psum_slot, int t0, int num_pix, const uint8_t *__restrict__ tile_dense,
    double *__restrict__ view) {
    __shared__ double acc[kPairRange][kRTile + 1];
    const int T = t0 + blockIdx.y, rg = blockIdx.x;
    const long j0 = tile_off[T], j1 = tile_off[T + 1];
    if (j0 == j1) return;  // no frame takes this tile: psum 0, the merge skips it
    if (tile_dense && tile_dense[T]) return;  // gathered per pixel (pair_tomogram_dense_kernel)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < kPairRange * (kRTile + 1); i += kBlock) (&acc[0][0])[i] = 0.;
    __syncthreads();
    const int p0 = rg * kPairRange;
    for (long j = j0 + wave; j < j1; j += kWavesPerBlock) {
        const int k = tile_pair[j];
        const double p = P[(size_t)k * kRTile + lane];
        if (!(p > kProbMin)) continue;
        const int d = pair_frame[k];
        const int *__restrict__ rr = roff + (size_t)d * (nranges + 1);
        const long base = foff[d];
        for (long q = base + rr[rg]; q < base + rr[rg + 1]; ++q)
            unsafeAtomicAdd(&acc[fpix[q] - p0][lane], fcnt[q] * p);
    }
    __syncthreads();
    const double ps = psum_slot[(size_t)T * kRTile + lane];
    for (int i = wave; i < kPairRange; i += kWavesPerBlock)
        acc[i][lane] = ps > 0. ? acc[i][lane] / ps : acc[i][lane];
    __syncthreads();
    // transpose: lanes along pixels
    for (int jj = wave; jj < kRTile; jj += kWavesPerBlock)
        for (int i = lane; i < kPairRange; i += 64) {
            const int t = p0 + i;
            if (t < num_pix) view[((size_t)blockIdx.y * kRTile + jj) * num_pix + t] = acc[i][jj];
        }
}

// Tomogram of tiles that most frames take (dense tiles), gathered per pixel like the plain
// path: block = 64 pixels x one dense tile; for each pixel, the frames with a photon there
// (tomogram photon list, frames ascending) that have the tile find their pair through
// pair_of[d][cell] (first pair of the frame's cell, -1 if none) + the tile's index in its cell.
__global__ __launch_bounds__(kBlock) void pair_tomogram_dense_kernel(
    const double *__restrict__ P, const int *__restrict__ dense_tiles, const int *__restrict__ tile_cell,
    const int *__restrict__ tile_sub, const int *__restrict__ pair_of, int num_cells,
    const long *__restrict__ toff, const int *__restrict__ tframe, const double *__restrict__ tcnt,
    const double *__restrict__ psum_slot, int t0, int num_pix, double *__restrict__ view) {
    __shared__ double s[64][kRTile + 1];
    const int T = dense_tiles[blockIdx.y];
    const int cell = tile_cell[T], sub = tile_sub[T];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p0 = blockIdx.x * 64;
    for (int i = wave; i < 64; i += kWavesPerBlock) {
        const int t = p0 + i;
        double acc = 0.;
        if (t < num_pix) {
            for (long q = toff[t]; q < toff[t + 1]; ++q) {
                const int d = tframe[q];
                const int k0 = pair_of[(size_t)d * num_cells + cell];
                if (k0 < 0) continue;
                const double p = P[((size_t)k0 + sub) * kRTile + lane];
                if (p > kProbMin) acc += tcnt[q] * p;
            }
        }
        s[i][lane] = acc;
    }
    __syncthreads();
    const double ps = psum_slot[(size_t)T * kRTile + lane];
    for (int i = wave; i < 64; i += kWavesPerBlock) s[i][lane] = ps > 0. ? s[i][lane] / ps : s[i][lane];
    __syncthreads();
    const int t = p0 + lane;
    if (t < num_pix)
        for (int j = wave; j < kRTile; j += kWavesPerBlock)
            view[((size_t)(T - t0) * kRTile + j) * num_pix + t] = s[lane][j];
}

// Scatter of pair values to a dense (num_data x num_rot) matrix (zeroed by the caller).
__global__ __launch_bounds__(kBlock) void pair_scatter_kernel(const double *__restrict__ P,
                                                              const int *__restrict__ pair_frame,
                                                              const int *__restrict__ pair_tile,
                                                              const int *__restrict__ slot_rot, long num_pairs,
                                                              int num_rot, double *__restrict__ out) {
    const long n = num_pairs * kRTile;
    for (long e = blockIdx.x * (long)kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock) {
        const long k = e / kRTile;
        const int r = slot_rot[pair_tile[k] * kRTile + (int)(e % kRTile)];
        if (r >= 0) out[(size_t)pair_frame[k] * num_rot + r] = P[e];
    }
}

// Gather of slot values into the original orientation order: out[r] = in[rot_slot[r]].
__global__ void slot_gather_kernel(const double *__restrict__ in, const int *__restrict__ rot_slot,
                                   int num_rot, double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < num_rot) out[r] = in[rot_slot[r]];
}

// ---- launchers -----------------------------------------------------------------------------

void launch_pair_likelihood(hipStream_t s, const double *logw, int num_pix, const long *off, const int *pix,
                            const double *cnt, const int *pair_frame, const int *pair_tile, long num_pairs,
                            int t0, int t1, double *G) {
    const unsigned grid = stride_grid((num_pairs + kWavesPerBlock - 1) / kWavesPerBlock);
    if (grid == 0) return;
    pair_likelihood_kernel<8><<<grid, kBlock, 0, s>>>(logw, (size_t)num_pix * kRTile, off, pix, cnt, pair_frame,
                                                      pair_tile, num_pairs, t0, t1, G);
}

void launch_pair_max(hipStream_t s, double *R, const long *pair_off, const int *pair_tile, const int *slot_rot,
                     const double *logwr_slot, const double *u_slot, const int *rbad_slot, const double *scale,
                     int use_scale, const double *scalars, const uint8_t *blacklist, const uint8_t *dense,
                     int num_data, double *max_exp, int *rmax) {
    pair_max_kernel<<<num_data, kBlock, 0, s>>>(R, pair_off, pair_tile, slot_rot, logwr_slot, u_slot, rbad_slot,
                                                scale, use_scale, scalars, blacklist, dense, max_exp, rmax);
}

void launch_pair_norm(hipStream_t s, const double *R, const long *pair_off, const int *pair_tile,
                      const int *slot_rot, const double *beta, const uint8_t *blacklist, const double *max_exp,
                      const int *rmax, int num_data, double *p_norm) {
    pair_norm_kernel<<<num_data, kBlock, 0, s>>>(R, pair_off, pair_tile, slot_rot, beta, blacklist, max_exp, rmax,
                                                 p_norm);
}

void launch_pair_posterior(hipStream_t s, double *R, const long *pair_off, const int *pair_tile,
                           const int *slot_rot, const double *logwr_slot, const double *u_slot,
                           const double *beta, const double *sum_fact, const uint8_t *blacklist,
                           const double *max_exp, const int *rmax, const double *p_norm, const double *scalars,
                           int want_psum_d, int num_data, double *likelihood, double *info, double *psum_d) {
    pair_posterior_kernel<<<num_data, kBlock, 0, s>>>(R, pair_off, pair_tile, slot_rot, logwr_slot, u_slot, beta,
                                                      sum_fact, blacklist, max_exp, rmax, p_norm, scalars,
                                                      want_psum_d, likelihood, info, psum_d);
}

void launch_pair_psum(hipStream_t s, const double *P, const long *tile_off, const int *tile_pair,
                      const int *pair_frame, int num_tiles, const double *scale, int need_scaling,
                      double *psum_slot) {
    pair_psum_kernel<<<(num_tiles + kWavesPerBlock - 1) / kWavesPerBlock, kBlock, 0, s>>>(
        P, tile_off, tile_pair, pair_frame, num_tiles, scale, need_scaling, psum_slot);
}

void launch_pair_tomogram(hipStream_t s, const double *P, const long *tile_off, const int *tile_pair,
                          const int *pair_frame, const long *foff, const int *roff, int nranges, const int *fpix,
                          const double *fcnt, const double *psum_slot, int t0, int ntiles, int num_pix,
                          const uint8_t *tile_dense, double *view) {
    if (ntiles <= 0) return;
    dim3 grid(nranges, ntiles);
    pair_tomogram_kernel<<<grid, kBlock, 0, s>>>(P, tile_off, tile_pair, pair_frame, foff, roff, nranges, fpix,
                                                 fcnt, psum_slot, t0, num_pix, tile_dense, view);
}

void launch_pair_tomogram_dense(hipStream_t s, const double *P, const int *dense_tiles, int ndense,
                                const int *tile_cell, const int *tile_sub, const int *pair_of, int num_cells,
                                const long *toff, const int *tframe, const double *tcnt, const double *psum_slot,
                                int t0, int num_pix, double *view) {
    if (ndense <= 0) return;
    dim3 grid((num_pix + 63) / 64, ndense);
    pair_tomogram_dense_kernel<<<grid, kBlock, 0, s>>>(P, dense_tiles, tile_cell, tile_sub, pair_of, num_cells,
                                                       toff, tframe, tcnt, psum_slot, t0, num_pix, view);
}

void launch_pair_scatter(hipStream_t s, const double *P, const int *pair_frame, const int *pair_tile,
                         const int *slot_rot, long num_pairs, int num_rot, double *out) {
    const long n = num_pairs * kRTile;
    if (n == 0) return;
    pair_scatter_kernel<<<stride_grid((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(P, pair_frame, pair_tile, slot_rot,
                                                                                num_pairs, num_rot, out);
}

void launch_slot_gather(hipStream_t s, const double *in, const int *rot_slot, int num_rot, double *out) {
    slot_gather_kernel<<<(num_rot + 255) / 256, 256, 0, s>>>(in, rot_slot, num_rot, out);
}

}  // namespace emc
