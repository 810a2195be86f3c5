// Per-frame orientation subsets evaluated sparsely (refinement, legacy src/max_emc.c:264-274).
//
// With refinement every frame is evaluated only at the orientations of its coarse cells. The
// engine orders orientations by coarse cell and pads every cell to whole tiles of 64
// ("slots": slot s holds orientation slot_rot[s], or -1 for padding), so the orientations a
// frame may take are a union of whole slot tiles.
//
// Two kinds of frames:
//  * partial frames (some cells): a (frame, tile) pair is one row of 64 values; the pairs of a
//    frame are contiguous in R (pair_off[d] .. pair_off[d + 1]) and a tile-major copy of the
//    pair list (tile_off / tile_pair, frames ascending) drives the per-tile passes;
//  * full frames (every cell: the fall-back to the plain search): their rows are kept
//    tile-major like the plain path, RF[(T * nF + f) * 64 + j] with f = ffull[d], so the plain
//    path's likelihood, psum and partitioned tomogram kernels run on them unchanged.
// Work and memory scale with the pairs of partial frames plus nF x tiles.
//
// The values are the plain path's: the same likelihood sums in the same photon order, the
// same first arg-max (ties broken by the original orientation index), the same keep rule
// (ordered by the original index), posterior, sums and tomogram. Only the order of additions
// across frames in psum_r and the tomogram differs from the reference.
#include <algorithm>
#include <cfloat>

#include "emc_common.h"
#include "emc_expand.h"
#include "emc_kernels.h"

namespace emc {

namespace {

// The rows of frame d: element e (e0 <= e < e1) has address at(e) and slot `slot`.
struct FrameRows {
    double *base;
    long e0, e1;
    int f, nF;
    const int *pair_tile;

    __device__ __forceinline__ double *at(long e, int &slot) const {
        if (f >= 0) {
            slot = (int)e;
            return base + ((size_t)(e / kRTile) * nF + f) * kRTile + (e % kRTile);
        }
        slot = pair_tile[e / kRTile] * kRTile + (int)(e % kRTile);
        return base + e;
    }
};

__device__ __forceinline__ FrameRows frame_rows(double *R, const long *pair_off, const int *pair_tile,
                                                const FullRows &full, int d) {
    FrameRows fr;
    fr.f = full.ffull ? full.ffull[d] : -1;
    fr.nF = full.nF;
    fr.pair_tile = pair_tile;
    if (fr.f >= 0) {
        fr.base = full.RF;
        fr.e0 = 0;
        fr.e1 = (long)full.num_tiles * kRTile;
    } else {
        fr.base = R;
        fr.e0 = pair_off[d] * kRTile;
        fr.e1 = pair_off[d + 1] * kRTile;
    }
    return fr;
}

}  // namespace

// G[k][lane] = sum over the frame's photons of count * logW[tile - t0][pix][lane] for the
// partial-frame pairs of tiles [t0, t1). Pairs are walked tile-major (tile_off / tile_pair), so
// the waves resident at one time read rows of the same log W tile. log W is chunk-local and
// tile-major with kW-wide tiles of type TW (64 x f64, or 128 x f32: two 64-slot tiles per row).
template <int U, typename TW, int kW>
__global__ __launch_bounds__(kBlock) void pair_likelihood_kernel(
    const TW *__restrict__ logw, size_t tsW, const long *__restrict__ off,
    const int *__restrict__ pix, const double *__restrict__ cnt, const int *__restrict__ pair_frame,
    const int *__restrict__ pair_tile, const long *__restrict__ tile_off, const int *__restrict__ tile_pair,
    int t0, int t1, double *__restrict__ G) {
    constexpr int kSub = kW / kRTile;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const long j1 = tile_off[t1];
    const long stride = (long)gridDim.x * kWavesPerBlock;
    for (long j = tile_off[t0] + (long)blockIdx.x * kWavesPerBlock + wave; j < j1; j += stride) {
        const int k = tile_pair[j];
        const int d = pair_frame[k];
        const int lt = pair_tile[k] - t0;
        const TW *__restrict__ col = logw + (size_t)(lt / kSub) * tsW + (lt % kSub) * kRTile + lane;
        long q = off[d];
        const long q1 = off[d + 1];
        double acc = 0.;
        for (; q + U <= q1; q += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = (double)col[(size_t)pix[q + u] * kW];
#pragma unroll
            for (int u = 0; u < U; ++u) acc += cnt[q + u] * v[u];
        }
        for (; q < q1; ++q) acc += cnt[q] * (double)col[(size_t)pix[q] * kW];
        G[(size_t)k * kRTile + lane] = acc;
    }
}

// Per frame over its rows: R = log w - u phi + G for real, allowed orientations (-inf for
// padding slots and, for dense frames, orientations with non-finite log W), maximum and first
// arg-max by original orientation index.
__global__ __launch_bounds__(kBlock) void pair_max_kernel(
    double *__restrict__ R, const long *__restrict__ pair_off, const int *__restrict__ pair_tile,
    FullRows full, const int *__restrict__ slot_rot, const int *__restrict__ slot_cell,
    const uint32_t *__restrict__ relbits, int words, const double *__restrict__ logwr_slot,
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
    const FrameRows fr = frame_rows(R, pair_off, pair_tile, full, d);
    // packed slots: a partial frame's tiles may hold orientations of cells outside its set
    const uint32_t *__restrict__ bits = (slot_cell && fr.f < 0) ? relbits + (size_t)d * words : nullptr;
    double best = kMaxExpStart;
    int bi = -1;
    for (long e = fr.e0 + threadIdx.x; e < fr.e1; e += kBlock) {
        int s;
        double *x = fr.at(e, s);
        const int r = slot_rot[s];
        bool on = r >= 0 && !(dense_frame && rbad_slot[s]);
        if (on && bits) {
            const int c = slot_cell[s];
            on = (bits[c >> 5] >> (c & 31)) & 1u;
        }
        double v = -INFINITY;
        if (on) v = (logwr_slot[s] - u_slot[s] * phi) + *x;
        *x = v;
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

// p_norm over the frame's kept orientations (normalize_prob, max_prob.c:245-250).
__global__ __launch_bounds__(kBlock) void pair_norm_kernel(
    double *__restrict__ R, const long *__restrict__ pair_off, const int *__restrict__ pair_tile,
    FullRows full, const int *__restrict__ slot_rot, const double *__restrict__ beta,
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
    const FrameRows fr = frame_rows(R, pair_off, pair_tile, full, d);
    double acc = 0.;
    for (long e = fr.e0 + threadIdx.x; e < fr.e1; e += kBlock) {
        int sl;
        const double v = *fr.at(e, sl);
        const int r = slot_rot[sl];
        if (r >= 0 && pair_kept(v, r, rf, M, c)) acc += exp(b * (v - M));
    }
    const double t = block_sum_pairs(acc, s);
    if (threadIdx.x == 0) p_norm[d] = t;
}

// P in place, and the frame's likelihood, information and psum_d (calc_psum_r,
// max_update.c:42-140).
__global__ __launch_bounds__(kBlock) void pair_posterior_kernel(
    double *__restrict__ R, const long *__restrict__ pair_off, const int *__restrict__ pair_tile,
    FullRows full, const int *__restrict__ slot_rot, const double *__restrict__ logwr_slot,
    const double *__restrict__ u_slot, const double *__restrict__ beta,
    const double *__restrict__ sum_fact, const uint8_t *__restrict__ blacklist,
    const double *__restrict__ max_exp, const int *__restrict__ rmax, const double *__restrict__ p_norm,
    const double *__restrict__ scalars, int want_psum_d, double *__restrict__ likelihood,
    double *__restrict__ info, double *__restrict__ psum_d) {
    __shared__ double s[kBlock];
    const int d = blockIdx.x;
    const int rf = rmax[d];
    const FrameRows fr = frame_rows(R, pair_off, pair_tile, full, d);
    if (blacklist[d] || rf < 0) {
        for (long e = fr.e0 + threadIdx.x; e < fr.e1; e += kBlock) {
            int sl;
            *fr.at(e, sl) = 0.;
        }
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
    for (long e = fr.e0 + threadIdx.x; e < fr.e1; e += kBlock) {
        int sl;
        double *x = fr.at(e, sl);
        const int r = slot_rot[sl];
        const double v = *x;
        double prob = 0.;
        if (r >= 0 && pair_kept(v, r, rf, M, c)) {
            const double ex = b * (v - M);
            prob = exp(ex) / pn;
            lk += prob * (v - sf);
            ps += prob * u_slot[sl] * rescale;
            if (!(prob < kProbMin)) inf += prob * ((ex - log_pn) - logwr_slot[sl]);
        }
        *x = prob;
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

// psum of every slot of a tile: the full frames' sum (psum_full, or 0) plus the sum over the
// tile's partial-frame pairs of P * phi' (phi' = scale if need_scaling else 1; blacklisted
// frames have P = 0). One block per tile: wave w takes the batches of 64 pairs w, w + 4, ...,
// its lanes load a batch's pair indices and factors together and pass them round by shuffles,
// so the P-row loads of a batch are independent; the waves' partial sums are added in wave
// order (a fixed order: the result does not depend on timing).
__global__ __launch_bounds__(kBlock) void pair_psum_kernel(const double *__restrict__ P,
                                                           const long *__restrict__ tile_off,
                                                           const int *__restrict__ tile_pair,
                                                           const int *__restrict__ pair_frame, int num_tiles,
                                                           const double *__restrict__ scale, int need_scaling,
                                                           const double *__restrict__ psum_full,
                                                           double *__restrict__ psum_slot) {
    __shared__ double part[kWavesPerBlock][kRTile];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int T = blockIdx.x;
    const long j0 = tile_off[T], j1 = tile_off[T + 1];
    double acc = 0.;
    for (long b = j0 + (long)wave * 64; b < j1; b += (long)kWavesPerBlock * 64) {
        const long j = b + lane;
        int k = 0;
        double f = 1.;
        if (j < j1) {
            k = tile_pair[j];
            if (need_scaling) f = scale[pair_frame[k]];
        }
        const int n = (int)min((long)64, j1 - b);
#pragma unroll 8
        for (int i = 0; i < n; ++i) {
            const int ki = __shfl(k, i);
            const double fi = __shfl(f, i);
            acc += P[(size_t)ki * kRTile + lane] * fi;
        }
    }
    part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0) {
        double sum = psum_full ? psum_full[(size_t)T * kRTile + lane] : 0.;
        for (int w = 0; w < kWavesPerBlock; ++w) sum += part[w][lane];
        psum_slot[(size_t)T * kRTile + lane] = sum;
    }
}

// Pairs of every tile whose row holds some P > 1e-6 (the tomogram's condition, on the values
// the tomogram reads), in the tile's order (frames ascending): sig_pair / sig_frame
// [tile_off[T] + i] for i < sig_cnt[T]. At a late iteration a frame's posterior sits in one
// or two of its proposed cells, so most pairs add nothing to any view.
template <typename TP>
__global__ __launch_bounds__(kBlock) void pair_sig_kernel(const TP *__restrict__ P, long num_pairs,
                                                          uint8_t *__restrict__ flag) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long stride = (long)gridDim.x * kWavesPerBlock;
    for (long k = (long)blockIdx.x * kWavesPerBlock + wave; k < num_pairs; k += stride) {
        const unsigned long long b = __ballot((double)P[(size_t)k * kRTile + lane] > kProbMin);
        if (lane == 0) flag[k] = b != 0ull;
    }
}

__global__ __launch_bounds__(kBlock) void sig_compact_kernel(const long *__restrict__ tile_off,
                                                             const int *__restrict__ tile_pair,
                                                             const int *__restrict__ tile_frame,
                                                             const uint8_t *__restrict__ flag, int *__restrict__ sig_pair,
                                                             int *__restrict__ sig_frame, int *__restrict__ sig_cnt) {
    __shared__ int s_w[kWavesPerBlock];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int T = blockIdx.x;
    const long j0 = tile_off[T], j1 = tile_off[T + 1];
    int base = 0;  // block-uniform count so far
    for (long jb = j0; jb < j1; jb += kBlock) {
        const long j = jb + threadIdx.x;
        int k = 0, d = 0;
        bool sig = false;
        if (j < j1) {
            k = tile_pair[j];
            d = tile_frame[j];
            sig = flag[k] != 0;
        }
        const unsigned long long b = __ballot(sig);
        if (lane == 0) s_w[wave] = __popcll(b);
        __syncthreads();
        int before = base, total = base;
        for (int w = 0; w < kWavesPerBlock; ++w) {
            if (w < wave) before += s_w[w];
            total += s_w[w];
        }
        if (sig) {
            const long at = j0 + before + __popcll(b & ((1ull << lane) - 1ull));
            sig_pair[at] = k;
            sig_frame[at] = d;
        }
        base = total;
        __syncthreads();  // s_w is rewritten by the next round
    }
    if (threadIdx.x == 0) sig_cnt[T] = base;
}

// Full frames' rows RF[T][f][64] with some P > 1e-6, per tile T, f ascending:
// sigF_row / sigF_frame[T * nF + i] = f / fidx[f] for i < sigF_cnt[T] (flag: one byte per row,
// pair_sig_kernel over the T * nF rows).
__global__ __launch_bounds__(kBlock) void full_sig_compact_kernel(const uint8_t *__restrict__ flag, int nF,
                                                                  const int *__restrict__ fidx,
                                                                  int *__restrict__ sigF_row,
                                                                  int *__restrict__ sigF_frame,
                                                                  int *__restrict__ sigF_cnt) {
    __shared__ int s_w[kWavesPerBlock];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t base_row = (size_t)blockIdx.x * nF;
    int base = 0;
    for (int fb = 0; fb < nF; fb += kBlock) {
        const int f = fb + threadIdx.x;
        const bool sig = f < nF && flag[base_row + f] != 0;
        const unsigned long long b = __ballot(sig);
        if (lane == 0) s_w[wave] = __popcll(b);
        __syncthreads();
        int before = base, total = base;
        for (int w = 0; w < kWavesPerBlock; ++w) {
            if (w < wave) before += s_w[w];
            total += s_w[w];
        }
        if (sig) {
            const size_t at = base_row + before + __popcll(b & ((1ull << lane) - 1ull));
            sigF_row[at] = f;
            sigF_frame[at] = fidx[f];
        }
        base = total;
        __syncthreads();
    }
    if (threadIdx.x == 0) sigF_cnt[blockIdx.x] = base;
}

// *total += set bits of m[0 .. n) (the caller zeroes *total)
__global__ __launch_bounds__(kBlock) void popcount_kernel(const unsigned long long *__restrict__ m, size_t n,
                                                          unsigned long long *__restrict__ total) {
    __shared__ unsigned long long s[kBlock];
    unsigned long long a = 0;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        a += (unsigned long long)__popcll(m[i]);
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && s[0]) atomicAdd(total, s[0]);
}

__global__ void zero_u64_kernel(unsigned long long *x) { *x = 0ull; }

void launch_zero_u64(hipStream_t s, unsigned long long *x) { zero_u64_kernel<<<1, 1, 0, s>>>(x); }

void launch_popcount(hipStream_t s, const unsigned long long *m, size_t n, unsigned long long *total) {
    if (n == 0) return;
    popcount_kernel<<<(unsigned)std::min<size_t>((n + kBlock - 1) / kBlock, 2048), kBlock, 0, s>>>(m, n, total);
}

// *total = sum of cnt[0 .. n) (one block)
__global__ __launch_bounds__(kBlock) void sum_counts_kernel(const int *__restrict__ cnt, int n, long *__restrict__ total) {
    __shared__ long s[kBlock];
    long a = 0;
    for (int i = threadIdx.x; i < n; i += kBlock) a += cnt[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s[0];
}

// Per-(tile, pixel range) buckets of the significant pairs' photons, so that a pair-tomogram
// block reads only the entries of its own range instead of every pair of its tile: an entry
// is (pair, frame range f), the frame's photons [fr_q0[f], fr_q1[f]) in range fr_rg[f] (the
// frame's distinct ranges, fsort order). Count: one block per tile, LDS counters per range
// (nr + 1 of them, dynamic LDS), then their exclusive scan: loc[T * (nr + 1) + rg] and the
// tile's total.
__global__ __launch_bounds__(kBlock) void bucket_count_kernel(const long *__restrict__ tile_off,
                                                              const int *__restrict__ sig_frame,
                                                              const int *__restrict__ sig_cnt,
                                                              const long *__restrict__ fr_off,
                                                              const int *__restrict__ fr_rg, int nr,
                                                              int *__restrict__ loc, long *__restrict__ tile_tot) {
    extern __shared__ int cnt[];
    __shared__ int part[kBlock];
    const int T = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = tid; i <= nr; i += kBlock) cnt[i] = 0;
    __syncthreads();
    const long j0 = tile_off[T];
    const int n = sig_cnt[T];
    for (int i = wave; i < n; i += kWavesPerBlock) {
        const int d = sig_frame[j0 + i];
        for (long f = fr_off[d] + lane; f < fr_off[d + 1]; f += 64) atomicAdd(&cnt[fr_rg[f]], 1);
    }
    __syncthreads();
    // exclusive scan of cnt[0 .. nr): each thread a chunk, then the chunks' totals
    const int per = (nr + kBlock - 1) / kBlock, a = tid * per, b = min(nr, a + per);
    int sum = 0;
    for (int i = a; i < b; ++i) sum += cnt[i];
    part[tid] = sum;
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int w = 0; w < kBlock; ++w) {
            const int v = part[w];
            part[w] = run;
            run += v;
        }
        tile_tot[T] = run;
        loc[(size_t)T * (nr + 1) + nr] = run;
    }
    __syncthreads();
    int run = part[tid];
    for (int i = a; i < b; ++i) {
        const int v = cnt[i];
        loc[(size_t)T * (nr + 1) + i] = run;
        run += v;
    }
}

// base[0] = 0, base[t + 1] = base[t] + tot[t] (one block)
__global__ __launch_bounds__(kBlock) void tile_scan_kernel(const long *__restrict__ tot, int n, long *__restrict__ base) {
    __shared__ long part[kBlock];
    const int tid = threadIdx.x, per = (n + kBlock - 1) / kBlock, a = tid * per, b = min(n, a + per);
    long sum = 0;
    for (int i = a; i < b; ++i) sum += tot[i];
    part[tid] = sum;
    __syncthreads();
    if (tid == 0) {
        long run = 0;
        for (int w = 0; w < kBlock; ++w) {
            const long v = part[w];
            part[w] = run;
            run += v;
        }
        base[n] = run;
    }
    __syncthreads();
    long run = part[tid];
    for (int i = a; i < b; ++i) {
        base[i] = run;
        run += tot[i];
    }
}

// Fill: the same walk, each entry at base[T] + loc[T][rg] + (its rank, from LDS counters)
__global__ __launch_bounds__(kBlock) void bucket_fill_kernel(const long *__restrict__ tile_off,
                                                             const int *__restrict__ sig_pair,
                                                             const int *__restrict__ sig_frame,
                                                             const int *__restrict__ sig_cnt,
                                                             const long *__restrict__ fr_off,
                                                             const int *__restrict__ fr_rg, int nr,
                                                             const int *__restrict__ loc, const long *__restrict__ base,
                                                             int *__restrict__ bk_pair, int *__restrict__ bk_fr) {
    extern __shared__ int fill[];
    const int T = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = tid; i < nr; i += kBlock) fill[i] = loc[(size_t)T * (nr + 1) + i];
    __syncthreads();
    const long j0 = tile_off[T], b0 = base[T];
    const int n = sig_cnt[T];
    for (int i = wave; i < n; i += kWavesPerBlock) {
        const int k = sig_pair[j0 + i], d = sig_frame[j0 + i];
        for (long f = fr_off[d] + lane; f < fr_off[d + 1]; f += 64) {
            const long at = b0 + atomicAdd(&fill[fr_rg[f]], 1);
            bk_pair[at] = k;
            bk_fr[at] = (int)f;
        }
    }
}

void launch_pair_buckets(hipStream_t s, int num_tiles, const long *tile_off, const int *sig_pair,
                         const int *sig_frame, const int *sig_cnt, const long *fr_off, const int *fr_rg, int nr,
                         int *loc, long *tile_tot, long *base, int *bk_pair, int *bk_fr, int pass) {
    if (num_tiles <= 0) return;
    const size_t lds = (size_t)(nr + 1) * sizeof(int);
    if (pass == 0) {
        bucket_count_kernel<<<num_tiles, kBlock, lds, s>>>(tile_off, sig_frame, sig_cnt, fr_off, fr_rg, nr, loc,
                                                          tile_tot);
        tile_scan_kernel<<<1, kBlock, 0, s>>>(tile_tot, num_tiles, base);
    } else {
        bucket_fill_kernel<<<num_tiles, kBlock, lds, s>>>(tile_off, sig_pair, sig_frame, sig_cnt, fr_off, fr_rg, nr,
                                                         loc, base, bk_pair, bk_fr);
    }
}

// Tomogram of the chunk's tiles [t0, t0 + ntiles) over pixel ranges of kPairRange pixels:
// block (tile, range) starts from the full frames' partial sums (tomogram_part_kernel
// layout partial[x][tile][t][64], added over the nparts frame ranges in order) or from zero,
// adds K * P (P > 1e-6) of the tile's significant partial-frame pairs (sig_pair, sig_cnt)
// and, with RF, of its significant full-frame rows (sigF_*: rows RF[T][f][64], in place of the
// partial sums) into LDS rows of the range's pixels (frame photon lists sorted by pixel, range
// offsets roff[d][range]), divides by psum and writes the orientation-major
// view[(T - t0) * 64 + j][t] (VT: double, or float for the FP32-sum merge). The tile's rows
// are examined kBlock at a time, one per thread, and only those with photons in the range are
// queued for the waves, so a row costs one read per range it has photons in.
template <typename TP, typename VT, typename PT>
__global__ __launch_bounds__(kBlock) void pair_tomogram_kernel(
    const TP *__restrict__ P, const long *__restrict__ tile_off, const int *__restrict__ tile_pair,
    const int *__restrict__ tile_frame, const int *__restrict__ sig_cnt, const long *__restrict__ foff,
    const int *__restrict__ roff, int nranges, const int *__restrict__ fpix, const double *__restrict__ fcnt,
    const double *__restrict__ psum_slot, int t0, int ntiles, int num_pix,
    const PT *__restrict__ partial, int nparts, VT *__restrict__ view,
    unsigned long long *__restrict__ nzmask, const double *__restrict__ RF, int nF,
    const int *__restrict__ sigF_row, const int *__restrict__ sigF_frame, const int *__restrict__ sigF_cnt,
    const PairBuckets bk) {
    static_assert(kPairRange == 64, "a pixel range is one 64-pixel merge patch");
    // LDS sums in the view's type: FP64 for an FP64 view, FP32 for an FP32 one (half the LDS,
    // more resident blocks; each sum has a few terms and is stored in FP32 anyway)
    using A = VT;
    __shared__ A acc[kPairRange][kRTile + 1];
    __shared__ long s_row[kBlock];  // element offset of the row (P or RF)
    __shared__ bool s_full[kBlock];
    __shared__ long s_q0[kBlock], s_q1[kBlock];
    __shared__ int s_n;
    // XCD-aware order: workgroups go round-robin to the kNumXcd XCDs, so XCD x takes tiles
    // x, x + kNumXcd, ... one at a time, all ranges of a tile together, and its L2 keeps the
    // tile's P rows while the ranges read them
    const int xcd = blockIdx.x % kNumXcd, j = blockIdx.x / kNumXcd;
    const int lt = xcd + kNumXcd * (j / nranges), rg = j % nranges;
    if (lt >= ntiles) return;
    const int T = t0 + lt;
    const long j0 = tile_off[T];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int npatch = (num_pix + 63) / 64;
    const bool rows_full = RF != nullptr;
    if (j0 == tile_off[T + 1] && !partial && !rows_full) {
        // no frame takes this tile: psum 0, views 0 (unwritten), the merge skips the items
        if (nzmask && tid < kRTile) nzmask[((size_t)lt * kRTile + tid) * npatch + rg] = 0ull;
        return;
    }
    // with buckets: this (tile, range)'s entries [e0, e0 + nRows); else every significant row
    // of the tile, filtered below by its photons in the range
    const bool buckets = bk.base != nullptr;
    long e0 = 0, nP, nRows;
    if (buckets) {
        const int *loc = bk.loc + (size_t)T * (nranges + 1);
        e0 = bk.base[T] + loc[rg];
        nP = nRows = loc[rg + 1] - loc[rg];
    } else {
        nP = sig_cnt[T];
        nRows = nP + (rows_full ? sigF_cnt[T] : 0);
    }
    const size_t frow0 = (size_t)T * nF;
    const int p0 = rg * kPairRange;
    if (partial) {
        const size_t tsW = (size_t)num_pix * kRTile, tsPart = tsW * ntiles;
        for (int i = wave; i < kPairRange; i += kWavesPerBlock) {
            const int t = p0 + i;
            double a = 0.;
            if (t < num_pix) {
                const size_t at = (size_t)lt * tsW + (size_t)t * kRTile + lane;
                for (int x = 0; x < nparts; ++x) a += (double)partial[x * tsPart + at];
            }
            acc[i][lane] = (A)a;
        }
    } else {
        for (int i = tid; i < kPairRange * (kRTile + 1); i += kBlock) (&acc[0][0])[i] = (A)0;
    }
    for (long jb = 0; jb < nRows; jb += kBlock) {
        if (tid == 0) s_n = 0;
        __syncthreads();
        const long j = jb + tid;
        if (buckets) {
            if (j < nRows) {
                // every entry has photons in the range: its queue position is its own
                const int f = bk.fr[e0 + j];
                s_row[tid] = (long)bk.pair[e0 + j] * kRTile;
                s_full[tid] = false;
                s_q0[tid] = bk.q0[f];
                s_q1[tid] = bk.q1[f];
            }
            if (tid == 0) s_n = (int)min((long)kBlock, nRows - jb);
        } else if (j < nRows) {
            long row;
            int d;
            const bool full = j >= nP;
            if (!full) {
                row = (long)tile_pair[j0 + j] * kRTile;
                d = tile_frame[j0 + j];
            } else {
                const size_t i = frow0 + (size_t)(j - nP);
                row = ((long)frow0 + sigF_row[i]) * kRTile;
                d = sigF_frame[i];
            }
            const int *__restrict__ rr = roff + (size_t)d * (nranges + 1);
            const int q0 = rr[rg], q1 = rr[rg + 1];
            if (q0 < q1) {
                const int at = atomicAdd(&s_n, 1);
                s_row[at] = row;
                s_full[at] = full;
                s_q0[at] = foff[d] + q0;
                s_q1[at] = foff[d] + q1;
            }
        }
        __syncthreads();
        const int n = s_n;
        // the P rows and first photons (a frame has about one photon in a 128-pixel range) of
        // kQueueRows queue entries are loaded before any is added, so a wave has that many
        // reads in flight instead of one chain of dependent ones
        constexpr int kQueueRows = 8;
        for (int i0 = wave; i0 < n; i0 += kWavesPerBlock * kQueueRows) {
            double p[kQueueRows], c0[kQueueRows];
            int px0[kQueueRows];
            long qa[kQueueRows], qb[kQueueRows];
#pragma unroll
            for (int u = 0; u < kQueueRows; ++u) {
                const int i = i0 + u * kWavesPerBlock;
                qa[u] = i < n ? s_q0[i] : 0;
                qb[u] = i < n ? s_q1[i] : 0;
                p[u] = i < n ? (s_full[i] ? RF[s_row[i] + lane] : (double)P[s_row[i] + lane]) : 0.;
                px0[u] = i < n ? fpix[qa[u]] : 0;
                c0[u] = i < n ? fcnt[qa[u]] : 0.;
            }
#pragma unroll
            for (int u = 0; u < kQueueRows; ++u) {
                if (!(p[u] > kProbMin)) continue;  // also the entries past the queue (p = 0)
                unsafeAtomicAdd(&acc[px0[u] - p0][lane], (A)(c0[u] * p[u]));
                for (long q = qa[u] + 1; q < qb[u]; ++q)
                    unsafeAtomicAdd(&acc[fpix[q] - p0][lane], (A)(fcnt[q] * p[u]));
            }
        }
        __syncthreads();  // every wave has read the queue before it is refilled
    }
    // psum of the tile's slots (FP32 views: its FP32 reciprocal, one division per slot)
    __shared__ double s_ps[kRTile];
    __shared__ float s_inv[kRTile];
    if (tid < kRTile) {
        const double ps = psum_slot[(size_t)T * kRTile + tid];
        s_ps[tid] = ps;
        s_inv[tid] = ps > 0. ? (float)(1. / ps) : 1.f;
    }
    __syncthreads();
    // divide by psum while transposing (lanes along pixels). With nzmask, item (slot, patch rg)
    // records which of its pixels have a nonzero view and only those are stored: the merge
    // reads no other.
    for (int jj = wave; jj < kRTile; jj += kWavesPerBlock) {
        const int t = p0 + lane;
        double v = 0.;
        if (t < num_pix) {
            if constexpr (sizeof(VT) == sizeof(float)) {
                v = (double)(acc[lane][jj] * s_inv[jj]);
            } else {
                const double ps = s_ps[jj];
                v = ps > 0. ? (double)acc[lane][jj] / ps : (double)acc[lane][jj];
            }
        }
        const size_t at = ((size_t)lt * kRTile + jj) * num_pix + t;
        if (nzmask) {
            const unsigned long long m = __ballot(v != 0.);
            if (lane == 0) nzmask[((size_t)lt * kRTile + jj) * npatch + rg] = m;
            if (v != 0.) view[at] = (VT)v;
        } else if (t < num_pix) {
            view[at] = (VT)v;
        }
    }
}

// Scatter of partial-frame pair values to a dense (num_data x num_rot) matrix (zeroed by the
// caller).
// (pairs k0 .. k0 + num_pairs; rows of out counted from frame d0)
__global__ __launch_bounds__(kBlock) void pair_scatter_kernel(const double *__restrict__ P,
                                                              const int *__restrict__ pair_frame,
                                                              const int *__restrict__ pair_tile,
                                                              const int *__restrict__ slot_rot, long k0,
                                                              long num_pairs, int d0, int num_rot,
                                                              double *__restrict__ out) {
    const long n = num_pairs * kRTile;
    for (long e = blockIdx.x * (long)kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock) {
        const long k = k0 + e / kRTile;
        const int r = slot_rot[pair_tile[k] * kRTile + (int)(e % kRTile)];
        if (r >= 0) out[(size_t)(pair_frame[k] - d0) * num_rot + r] = P[k * kRTile + e % kRTile];
    }
}

// The same for full frames f0 .. f1: RF[(T * nF + f) * 64 + j] -> out[fidx[f] - d0][slot_rot[T * 64 + j]].
__global__ __launch_bounds__(kBlock) void full_scatter_kernel(const double *__restrict__ RF, int nF, int f0,
                                                              int f1, int num_tiles, const int *__restrict__ fidx,
                                                              const int *__restrict__ slot_rot, int d0,
                                                              int num_rot, double *__restrict__ out) {
    const int nb = f1 - f0;
    const long n = (long)nb * num_tiles * kRTile;
    for (long e = blockIdx.x * (long)kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock) {
        const long row = e / kRTile;
        const int T = (int)(row / nb), f = f0 + (int)(row % nb), j = (int)(e % kRTile);
        const int r = slot_rot[T * kRTile + j];
        if (r >= 0) out[(size_t)(fidx[f] - d0) * num_rot + r] = RF[((size_t)T * nF + f) * kRTile + j];
    }
}

// Gather of slot values into the original orientation order: out[r] = in[rot_slot[r]].
__global__ void slot_gather_kernel(const double *__restrict__ in, const int *__restrict__ rot_slot,
                                   int num_rot, double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < num_rot) out[r] = in[rot_slot[r]];
}

// Per-frame values of the full frames: out[f] = in[fidx[f]].
template <typename V>
__global__ void frame_gather_kernel(const V *__restrict__ in, const int *__restrict__ fidx, int nF,
                                    V *__restrict__ out) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nF) out[f] = in[fidx[f]];
}

// out[d - d0][c] = sum of frame d's posterior over the orientations of coarse cell c (f32), from
// the plain path's rows (plain: R[T][num_data][64], orientation T * 64 + j) or from the
// subset rows (partial-frame pairs and RF). Sums in LDS (num_coarse doubles).
__global__ __launch_bounds__(kBlock) void cell_posterior_kernel(
    double *__restrict__ R, int plain, int num_data, int num_rot, const long *__restrict__ pair_off,
    const int *__restrict__ pair_tile, FullRows full, const int *__restrict__ slot_rot,
    const int *__restrict__ qmap, int num_coarse, int d0, float *__restrict__ out) {
    extern __shared__ double s_cell[];
    const int d = d0 + blockIdx.x;
    for (int c = threadIdx.x; c < num_coarse; c += kBlock) s_cell[c] = 0.;
    __syncthreads();
    if (plain) {
        const size_t tsR = (size_t)num_data * kRTile;
        for (int r = threadIdx.x; r < num_rot; r += kBlock) {
            const double p = R[(size_t)(r >> 6) * tsR + (size_t)d * kRTile + (r & 63)];
            if (p != 0.) unsafeAtomicAdd(&s_cell[qmap[r]], p);
        }
    } else {
        const FrameRows fr = frame_rows(R, pair_off, pair_tile, full, d);
        for (long e = fr.e0 + threadIdx.x; e < fr.e1; e += kBlock) {
            int sl;
            const double p = *fr.at(e, sl);
            const int r = slot_rot[sl];
            if (r >= 0 && p != 0.) unsafeAtomicAdd(&s_cell[qmap[r]], p);
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < num_coarse; c += kBlock) out[(size_t)blockIdx.x * num_coarse + c] = (float)s_cell[c];
}

// ---- full-frame photon lists ----------------------------------------------------------------
// Per engine pixel, the photons (mask < 2) of the full frames in frame order, split into
// nparts ranges of full-frame index f (range x holds f with nF * x / nparts <= f). One wave
// per pixel. Pass 0 counts the photons of every (pixel, range); pass 1 writes them from the
// pixel's first output slot off2[t * nparts], keeping the pixel's frame order.
__device__ __forceinline__ int full_part_of(int f, int nF, int nparts) {
    int x = (int)(((long)f * nparts) / nF);  // then the largest x with nF * x / nparts <= f
    while (x + 1 < nparts && (long)nF * (x + 1) / nparts <= f) ++x;
    while (x > 0 && (long)nF * x / nparts > f) --x;
    return x;
}

__global__ __launch_bounds__(kBlock) void full_lists_kernel(
    int pass, const long *__restrict__ tom_off, const int *__restrict__ tom_frame,
    const double *__restrict__ tom_cnt, const int *__restrict__ fof, int num_pix, int nF, int nparts,
    int *__restrict__ counts, const long *__restrict__ off2, int *__restrict__ out_frame,
    double *__restrict__ out_cnt) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (t >= num_pix) return;
    const long k0 = tom_off[t], k1 = tom_off[t + 1];
    long at = pass ? off2[(size_t)t * nparts] : 0;
    int cnt_lane = 0;  // pass 0: lane x < nparts counts range x
    for (long kb = k0; kb < k1; kb += 64) {
        const long k = kb + lane;
        const int f = k < k1 ? fof[tom_frame[k]] : -1;
        const unsigned long long keep = __ballot(f >= 0);
        if (pass == 0) {
            const int x = f >= 0 ? full_part_of(f, nF, nparts) : -1;
            for (int y = 0; y < nparts; ++y) {
                const int n = __popcll(__ballot(x == y));
                if (lane == y) cnt_lane += n;
            }
        } else {
            if (f >= 0) {
                const long pos = at + __popcll(keep & ((1ull << lane) - 1ull));
                out_frame[pos] = f;
                out_cnt[pos] = tom_cnt[k];
            }
            at += __popcll(keep);
        }
    }
    if (pass == 0 && lane < nparts) counts[(size_t)t * nparts + lane] = cnt_lane;
}

// ---- launchers -----------------------------------------------------------------------------

void launch_full_lists(hipStream_t s, int pass, const long *tom_off, const int *tom_frame, const double *tom_cnt,
                       const int *fof, int num_pix, int nF, int nparts, int *counts, const long *off2,
                       int *out_frame, double *out_cnt) {
    if (num_pix <= 0) return;
    full_lists_kernel<<<(num_pix + kWavesPerBlock - 1) / kWavesPerBlock, kBlock, 0, s>>>(
        pass, tom_off, tom_frame, tom_cnt, fof, num_pix, nF, nparts, counts, off2, out_frame, out_cnt);
}

void launch_cell_posterior(hipStream_t s, double *R, int plain, int num_data, int num_rot, const long *pair_off,
                           const int *pair_tile, const FullRows &full, const int *slot_rot, const int *qmap,
                           int num_coarse, int d0, int count, float *out) {
    if (count <= 0) return;
    cell_posterior_kernel<<<count, kBlock, num_coarse * sizeof(double), s>>>(
        R, plain, num_data, num_rot, pair_off, pair_tile, full, slot_rot, qmap, num_coarse, d0, out);
}

// ---- persistent expand + likelihood ------------------------------------------------------
// One launch per orientation chunk does the chunk's expand and its (frame, tile) pair
// likelihoods. Each block takes whole 64-orientation tiles from a counter, one at a time:
// it expands the tile over every pixel group into log W (expand_block) and then, with its
// waves, evaluates every pair of that tile from the log W it has just written, while that is
// still in L2 / the Infinity Cache. A tile's pairs can be many (frames of up to 20,000 photons
// with subsets that favour some tiles) or few; blocks that finish early take the next tile, so
// the uneven work of a chunk stays balanced without any block waiting on another. The
// arithmetic is expand_block's and pair_likelihood_kernel's, in the same order.
__global__ void fused_reset_kernel(unsigned *__restrict__ counter) { *counter = 0u; }

template <int U, typename T, int kW, typename MT>
__global__ __launch_bounds__(kBlock) void fused_expand_likelihood_kernel(
    const double *__restrict__ rots, const MT *__restrict__ model, long size, const double *__restrict__ qvals,
    const double *__restrict__ corr, const uint8_t *__restrict__ mask, int num_pix, int r0, int C, int pix_per_group,
    int ngroups, T *__restrict__ logw, double *__restrict__ upart, int *__restrict__ ubad,
    const int *__restrict__ slot_rot, const long *__restrict__ off, const int *__restrict__ pix,
    const double *__restrict__ cnt, const int *__restrict__ pair_frame, const int *__restrict__ pair_tile,
    const long *__restrict__ tile_off, const int *__restrict__ tile_pair, int t0, int ntiles,
    unsigned *__restrict__ counter, double *__restrict__ G) {
    constexpr int kSub = kW / kRTile;
    __shared__ ExpandShared<MT> sh;
    __shared__ int s_tile;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const size_t tsW = (size_t)num_pix * kW;
    for (;;) {
        if (tid == 0) s_tile = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int lt = s_tile;  // local tile of the chunk
        __syncthreads();
        if (lt >= ntiles) break;  // every block leaves here once the tiles run out
        for (int g = 0; g < ngroups; ++g)
            expand_block<T, kW, MT>(sh, lt * kRTile, g, rots, model, size, qvals, corr, mask, num_pix, r0, C,
                                    pix_per_group, logw, upart, ubad, slot_rot, 1, 0);
        __threadfence();  // the tile's log W, written by every wave, is read below by every wave
        __syncthreads();
        const long j1 = tile_off[t0 + lt + 1];
        for (long j = tile_off[t0 + lt] + wave; j < j1; j += kWavesPerBlock) {
            const int kp = tile_pair[j];
            const int d = pair_frame[kp];
            const T *__restrict__ col = logw + (size_t)(lt / kSub) * tsW + (lt % kSub) * kRTile + lane;
            long q = off[d];
            const long q1 = off[d + 1];
            double acc = 0.;
            for (; q + U <= q1; q += U) {
                double v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) v[u] = (double)col[(size_t)pix[q + u] * kW];
#pragma unroll
                for (int u = 0; u < U; ++u) acc += cnt[q + u] * v[u];
            }
            for (; q < q1; ++q) acc += cnt[q] * (double)col[(size_t)pix[q] * kW];
            G[(size_t)kp * kRTile + lane] = acc;
        }
    }
}

template <typename T, int kW, typename MT>
static void fused_launch(hipStream_t s, const double *rots, const MT *model, long size, const double *qvals,
                         const double *corr, const uint8_t *mask, int num_pix, int r0, int C, int pix_per_group,
                         T *logw, double *upart, int *ubad, const int *slot_rot, const long *off, const int *pix,
                         const double *cnt, const int *pair_frame, const int *pair_tile, const long *tile_off,
                         const int *tile_pair, int t0, unsigned *counter, double *G) {
    const int ntiles = C / kRTile;
    if (ntiles <= 0) return;
    fused_reset_kernel<<<1, 1, 0, s>>>(counter);
    static int resident = 0;  // blocks of this kernel that fit the device at once
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fused_expand_likelihood_kernel<8, T, kW, MT>,
                                                           kBlock, 0);
        resident = std::max(1, cus * std::max(per_cu, 1));
    }
    const unsigned grid = (unsigned)std::min(ntiles, resident);
    const int ngroups = (num_pix + pix_per_group - 1) / pix_per_group;
    fused_expand_likelihood_kernel<8, T, kW, MT><<<grid, kBlock, 0, s>>>(
        rots, model, size, qvals, corr, mask, num_pix, r0, C, pix_per_group, ngroups, logw, upart, ubad, slot_rot, off,
        pix, cnt, pair_frame, pair_tile, tile_off, tile_pair, t0, ntiles, counter, G);
}

void launch_fused_expand_likelihood(hipStream_t s, const double *rots, const double *model, long size,
                                    const double *qvals, const double *corr, const uint8_t *mask, int num_pix, int r0,
                                    int C, int pix_per_group, double *logw, double *upart, int *ubad,
                                    const int *slot_rot, const long *off, const int *pix, const double *cnt,
                                    const int *pair_frame, const int *pair_tile, const long *tile_off,
                                    const int *tile_pair, int t0, unsigned *counter, double *G) {
    fused_launch<double, kRTile, double>(s, rots, model, size, qvals, corr, mask, num_pix, r0, C, pix_per_group, logw,
                                         upart, ubad, slot_rot, off, pix, cnt, pair_frame, pair_tile, tile_off,
                                         tile_pair, t0, counter, G);
}

void launch_fused_expand_likelihood(hipStream_t s, const double *rots, const float *model, long size,
                                    const double *qvals, const double *corr, const uint8_t *mask, int num_pix, int r0,
                                    int C, int pix_per_group, float *logw, double *upart, int *ubad,
                                    const int *slot_rot, const long *off, const int *pix, const double *cnt,
                                    const int *pair_frame, const int *pair_tile, const long *tile_off,
                                    const int *tile_pair, int t0, unsigned *counter, double *G) {
    fused_launch<float, 2 * kRTile, float>(s, rots, model, size, qvals, corr, mask, num_pix, r0, C, pix_per_group,
                                           logw, upart, ubad, slot_rot, off, pix, cnt, pair_frame, pair_tile, tile_off,
                                           tile_pair, t0, counter, G);
}

void launch_pair_likelihood(hipStream_t s, const double *logw, int num_pix, const long *off, const int *pix,
                            const double *cnt, const int *pair_frame, const int *pair_tile, const long *tile_off,
                            const int *tile_pair, long num_pairs, int t0, int t1, double *G) {
    const unsigned grid = stride_grid((num_pairs + kWavesPerBlock - 1) / kWavesPerBlock);
    if (grid == 0) return;
    pair_likelihood_kernel<8, double, kRTile><<<grid, kBlock, 0, s>>>(logw, (size_t)num_pix * kRTile, off, pix, cnt,
                                                                      pair_frame, pair_tile, tile_off, tile_pair, t0,
                                                                      t1, G);
}

void launch_pair_likelihood(hipStream_t s, const float *logw, int num_pix, const long *off, const int *pix,
                            const double *cnt, const int *pair_frame, const int *pair_tile, const long *tile_off,
                            const int *tile_pair, long num_pairs, int t0, int t1, double *G) {
    const unsigned grid = stride_grid((num_pairs + kWavesPerBlock - 1) / kWavesPerBlock);
    if (grid == 0) return;
    pair_likelihood_kernel<8, float, 2 * kRTile><<<grid, kBlock, 0, s>>>(logw, (size_t)num_pix * 2 * kRTile, off, pix,
                                                                         cnt, pair_frame, pair_tile, tile_off,
                                                                         tile_pair, t0, t1, G);
}

void launch_pair_max(hipStream_t s, double *R, const long *pair_off, const int *pair_tile, const FullRows &full,
                     const int *slot_rot, const int *slot_cell, const uint32_t *relbits, int words,
                     const double *logwr_slot, const double *u_slot, const int *rbad_slot, const double *scale,
                     int use_scale, const double *scalars, const uint8_t *blacklist, const uint8_t *dense,
                     int num_data, double *max_exp, int *rmax) {
    pair_max_kernel<<<num_data, kBlock, 0, s>>>(R, pair_off, pair_tile, full, slot_rot, slot_cell, relbits, words,
                                                logwr_slot, u_slot, rbad_slot, scale, use_scale, scalars, blacklist,
                                                dense, max_exp, rmax);
}

void launch_pair_norm(hipStream_t s, double *R, const long *pair_off, const int *pair_tile, const FullRows &full,
                      const int *slot_rot, const double *beta, const uint8_t *blacklist, const double *max_exp,
                      const int *rmax, int num_data, double *p_norm) {
    pair_norm_kernel<<<num_data, kBlock, 0, s>>>(R, pair_off, pair_tile, full, slot_rot, beta, blacklist, max_exp,
                                                 rmax, p_norm);
}

void launch_pair_posterior(hipStream_t s, double *R, const long *pair_off, const int *pair_tile,
                           const FullRows &full, const int *slot_rot, const double *logwr_slot,
                           const double *u_slot, const double *beta, const double *sum_fact,
                           const uint8_t *blacklist, const double *max_exp, const int *rmax, const double *p_norm,
                           const double *scalars, int want_psum_d, int num_data, double *likelihood, double *info,
                           double *psum_d) {
    pair_posterior_kernel<<<num_data, kBlock, 0, s>>>(R, pair_off, pair_tile, full, slot_rot, logwr_slot, u_slot,
                                                      beta, sum_fact, blacklist, max_exp, rmax, p_norm, scalars,
                                                      want_psum_d, likelihood, info, psum_d);
}

void launch_pair_psum(hipStream_t s, const double *P, const long *tile_off, const int *tile_pair,
                      const int *pair_frame, int num_tiles, const double *scale, int need_scaling,
                      const double *psum_full, double *psum_slot) {
    if (num_tiles <= 0) return;
    pair_psum_kernel<<<num_tiles, kBlock, 0, s>>>(
        P, tile_off, tile_pair, pair_frame, num_tiles, scale, need_scaling, psum_full, psum_slot);
}

template <typename TP, typename VT, typename PT>
static void pair_tomogram_launch(hipStream_t s, const TP *P, const long *tile_off, const int *sig_pair,
                                 const int *sig_frame, const int *sig_cnt, const long *foff, const int *roff,
                                 int nranges, const int *fpix, const double *fcnt, const double *psum_slot, int t0,
                                 int ntiles, int num_pix, const PT *partial, int nparts, VT *view,
                                 unsigned long long *nzmask, const SigFullRows &fr, const PairBuckets &bk) {
    if (ntiles <= 0) return;
    const unsigned grid = (unsigned)((ntiles + kNumXcd - 1) / kNumXcd * kNumXcd) * (unsigned)nranges;
    pair_tomogram_kernel<TP, VT, PT><<<grid, kBlock, 0, s>>>(P, tile_off, sig_pair, sig_frame, sig_cnt, foff, roff,
                                                             nranges, fpix, fcnt, psum_slot, t0, ntiles, num_pix,
                                                             partial, nparts, view, nzmask, fr.RF, fr.nF, fr.row,
                                                             fr.frame, fr.cnt, bk);
}

void launch_pair_tomogram(hipStream_t s, const double *P, const long *tile_off, const int *sig_pair,
                          const int *sig_frame, const int *sig_cnt, const long *foff, const int *roff, int nranges,
                          const int *fpix, const double *fcnt, const double *psum_slot, int t0, int ntiles,
                          int num_pix, const double *partial, int nparts, double *view,
                          unsigned long long *nzmask, const SigFullRows &fr, const PairBuckets &bk) {
    pair_tomogram_launch(s, P, tile_off, sig_pair, sig_frame, sig_cnt, foff, roff, nranges, fpix, fcnt, psum_slot,
                         t0, ntiles, num_pix, partial, nparts, view, nzmask, fr, bk);
}

void launch_pair_tomogram(hipStream_t s, const float *P, const long *tile_off, const int *sig_pair,
                          const int *sig_frame, const int *sig_cnt, const long *foff, const int *roff, int nranges,
                          const int *fpix, const double *fcnt, const double *psum_slot, int t0, int ntiles,
                          int num_pix, const float *partial, int nparts, float *view,
                          unsigned long long *nzmask, const SigFullRows &fr, const PairBuckets &bk) {
    pair_tomogram_launch(s, P, tile_off, sig_pair, sig_frame, sig_cnt, foff, roff, nranges, fpix, fcnt, psum_slot,
                         t0, ntiles, num_pix, partial, nparts, view, nzmask, fr, bk);
}

template <typename TP>
static void sig_pairs_launch(hipStream_t s, const TP *P, long num_pairs, const long *tile_off, const int *tile_pair,
                             const int *tile_frame, int num_tiles, uint8_t *flag, int *sig_pair, int *sig_frame,
                             int *sig_cnt) {
    if (num_tiles <= 0) return;
    if (num_pairs > 0)
        pair_sig_kernel<TP><<<stride_grid((num_pairs + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, s>>>(
            P, num_pairs, flag);
    sig_compact_kernel<<<num_tiles, kBlock, 0, s>>>(tile_off, tile_pair, tile_frame, flag, sig_pair, sig_frame,
                                                    sig_cnt);
}

void launch_sig_pairs(hipStream_t s, const double *P, long num_pairs, const long *tile_off, const int *tile_pair,
                      const int *tile_frame, int num_tiles, uint8_t *flag, int *sig_pair, int *sig_frame,
                      int *sig_cnt) {
    sig_pairs_launch(s, P, num_pairs, tile_off, tile_pair, tile_frame, num_tiles, flag, sig_pair, sig_frame, sig_cnt);
}

void launch_sig_full_rows(hipStream_t s, const double *RF, int nF, int num_tiles, const int *fidx, uint8_t *flag,
                          int *sigF_row, int *sigF_frame, int *sigF_cnt, long *total) {
    if (num_tiles <= 0 || nF <= 0) return;
    const long rows = (long)num_tiles * nF;
    pair_sig_kernel<double><<<stride_grid((rows + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, s>>>(RF, rows,
                                                                                                         flag);
    full_sig_compact_kernel<<<num_tiles, kBlock, 0, s>>>(flag, nF, fidx, sigF_row, sigF_frame, sigF_cnt);
    sum_counts_kernel<<<1, kBlock, 0, s>>>(sigF_cnt, num_tiles, total);
}

void launch_sig_pairs(hipStream_t s, const float *P, long num_pairs, const long *tile_off, const int *tile_pair,
                      const int *tile_frame, int num_tiles, uint8_t *flag, int *sig_pair, int *sig_frame,
                      int *sig_cnt) {
    sig_pairs_launch(s, P, num_pairs, tile_off, tile_pair, tile_frame, num_tiles, flag, sig_pair, sig_frame, sig_cnt);
}

// FP32 copy of the pair rows (the posterior read by the FP32 pair tomogram).
__global__ void pair_to_f32_kernel(const double *__restrict__ P, long n, float *__restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (float)P[i];
}

void launch_pair_to_f32(hipStream_t s, const double *P, long n, float *out) {
    if (n > 0) pair_to_f32_kernel<<<stride_grid((n + 255) / 256), 256, 0, s>>>(P, n, out);
}

void launch_pair_scatter(hipStream_t s, const double *P, const int *pair_frame, const int *pair_tile,
                         const int *slot_rot, long num_pairs, int num_rot, double *out, long k0, int d0) {
    const long n = num_pairs * kRTile;
    if (n == 0) return;
    pair_scatter_kernel<<<stride_grid((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(P, pair_frame, pair_tile, slot_rot,
                                                                                  k0, num_pairs, d0, num_rot, out);
}

void launch_full_scatter(hipStream_t s, const double *RF, int nF, int num_tiles, const int *fidx,
                         const int *slot_rot, int num_rot, double *out, int f0, int f1, int d0) {
    if (f1 < 0) f1 = nF;
    const long n = (long)(f1 - f0) * num_tiles * kRTile;
    if (n <= 0) return;
    full_scatter_kernel<<<stride_grid((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(RF, nF, f0, f1, num_tiles, fidx,
                                                                                  slot_rot, d0, num_rot, out);
}

void launch_slot_gather(hipStream_t s, const double *in, const int *rot_slot, int num_rot, double *out) {
    slot_gather_kernel<<<(num_rot + 255) / 256, 256, 0, s>>>(in, rot_slot, num_rot, out);
}

void launch_frame_gather(hipStream_t s, const double *in, const int *fidx, int nF, double *out) {
    if (nF > 0) frame_gather_kernel<double><<<(nF + 255) / 256, 256, 0, s>>>(in, fidx, nF, out);
}

void launch_frame_gather(hipStream_t s, const uint8_t *in, const int *fidx, int nF, uint8_t *out) {
    if (nF > 0) frame_gather_kernel<uint8_t><<<(nF + 255) / 256, 256, 0, s>>>(in, fidx, nF, out);
}

}  // namespace emc
