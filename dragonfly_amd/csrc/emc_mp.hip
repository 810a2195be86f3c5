// Mixed-precision path for dense integer frames on gfx950 matrix cores.
//
// For dense frames the two photon contractions of an iteration are dense products:
//   likelihood  G[d][r]    = sum_t K[d][t] log W[r][t]         (mask < 1 pixels)
//   tomogram    view[r][t] = sum_d P[d][r] K[d][t]              (mask < 2, P > 1e-6)
// They run here as FP16 x FP16 products on MFMA (v_mfma_f32_32x32x16_f16) with FP32
// accumulation; everything else of the iteration (per-frame maximum, keep rule, p_norm,
// posterior, likelihood and information sums, psum_r / psum_d, merge, model update) stays in
// FP64 and is shared with the plain path.
//
// Precision:
//   * photon counts are integers 0..2048 and exact in FP16;
//   * log W is centred per pixel, x_rt = log W_rt - c_t with c_t the mean of log W over a
//     strided sample of 64 orientations, so the FP16 operand and the FP32 partial sums are
//     small; sum_t K_dt c_t is added back per frame in FP64. x is stored as hi = fp16(x) and
//     (two-term mode) lo = fp16(x - hi): hi + lo carries ~22 bits, and both products go into
//     the same FP32 accumulators;
//   * P is stored as fp16(P * 2^15) (P in (1e-6, 1] is then a normal FP16 number) and the
//     tomogram epilogue multiplies by 2^-15 / psum_r in FP64.
//
// Layouts (row-major, K-contiguous operands, all padded with zeros):
//   klik  [Mp_d][ldp]   FP16 counts, frames x engine pixels (mask < 1, else 0)
//   ktom  [Mp_t][ldd]   FP16 counts, engine pixels x frames (mask < 2, else 0)
//   lw_hi, lw_lo [Np_c][ldp]  centred log W of one orientation chunk
//   p16   [Np_c][ldd]   scaled posterior of one orientation chunk
// with ldp = num_pix rounded up to kMpGemmBK and ldd = num_data rounded up to kMpGemmBK; row
// counts are rounded up to the GEMM tile. Outputs are the FP64 buffers of the plain path: R
// tile-major (frame_max_kernel and the later per-frame passes read it unchanged) and the view
// orientation-major, as the brick merge reads it.
#include <cfloat>

#include "emc_common.h"
#include "emc_gemm.h"
#include "emc_geometry.h"
#include "emc_kernels.h"

namespace emc {


// ---- product kernel ----------------------------------------------------------------------
// C[m][n] = sum_k A[m][k] * (B0[n][k] + B1[n][k]) (B1 only when NB == 2) with the shared
// matrix-core tile (emc_gemm.h). Epilogue, for elements (m, n) with m < M_valid, n < N_valid:
//   EPI 0: out[(n / 64) * ts + m * 64 + n % 64] = acc + row_add[m]   (tile-major; likelihood
//          of frame m, orientation n, plus sum_t K_mt c_t)
//   EPI 1: out[m * ts + n] = acc * scale / row_div[m] if m < div_valid and row_div[m] > 0,
//          else acc * scale   (row-major; tomogram of orientation m at pixel n; rows at or past
//          div_valid are padding orientations)
template <typename T, int NB, int EPI>
__global__ __launch_bounds__(kGemmThreads) void mp_gemm_kernel(
    const T *__restrict__ A, size_t lda, const T *__restrict__ B0,
    const T *__restrict__ B1, size_t ldb, int K, int m_tiles, int n_tiles, int M_valid,
    int N_valid, double *__restrict__ out, size_t ts, const double *__restrict__ row_add,
    const double *__restrict__ row_div, int div_valid, double scale) {
    extern __shared__ char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    const GemmTile tile = gemm_block_tile(m_tiles, n_tiles);
    if (!tile.valid) return;
    floatx16 acc[2][4];
    // one-term products double-buffer LDS (2 x 72 KiB) and their fragment reads (the two-term
    // ones fit neither: LDS, registers)
    gemm_mainloop<T, NB, kMpGemmBK, NB == 1, NB == 1>(A, lda, B0, B1, ldb, K, tile.m0, tile.n0, smem, acc);
    gemm_for_each(tile.m0, tile.n0, acc, [&](size_t m, size_t n, float a) {
        if (m >= (size_t)M_valid || n >= (size_t)N_valid) return;
        const double x = (double)a;
        if (EPI == 0) {
            out[(n >> 6) * ts + m * kRTile + (n & 63)] = x + row_add[m];
        } else {
            const double ps = (m < (size_t)div_valid) ? row_div[m] : 0.;
            out[m * ts + n] = ps > 0. ? x * scale / ps : x * scale;
        }
    });
}

// ---- operand builders ---------------------------------------------------------------------

// Round a double to the 16-bit operand type: FP16 directly, BF16 through FP32.
template <typename T>
__device__ __forceinline__ T to_op(double x);
template <>
__device__ __forceinline__ _Float16 to_op<_Float16>(double x) { return (_Float16)x; }
template <>
__device__ __forceinline__ __bf16 to_op<__bf16>(double x) { return (__bf16)(float)x; }
template <typename T>
__device__ __forceinline__ double from_op(T x) { return (double)(float)x; }
template <>
__device__ __forceinline__ double from_op<_Float16>(_Float16 x) { return (double)x; }

// klik[d][pix] = count for the frame's mask < 1 photons (likelihood photon list, engine order),
// as T (integers are exact to 2048 in FP16, to 256 in BF16).
template <typename T>
__global__ __launch_bounds__(kBlock) void build_klik_kernel(const long *__restrict__ off,
                                                            const int *__restrict__ pix,
                                                            const double *__restrict__ cnt,
                                                            int num_data, size_t ldp,
                                                            T *__restrict__ klik) {
    const int d = blockIdx.x;
    if (d >= num_data) return;
    for (long k = off[d] + threadIdx.x; k < off[d + 1]; k += blockDim.x)
        klik[(size_t)d * ldp + pix[k]] = to_op<T>(cnt[k]);
}

// ktom[t][frame] = count for the pixel's mask < 2 photons (tomogram photon list).
__global__ __launch_bounds__(kBlock) void build_ktom_kernel(const long *__restrict__ toff,
                                                            const int *__restrict__ tframe,
                                                            const double *__restrict__ tcnt,
                                                            int num_pix, size_t ldd,
                                                            _Float16 *__restrict__ ktom) {
    const int t = blockIdx.x;
    if (t >= num_pix) return;
    for (long k = toff[t] + threadIdx.x; k < toff[t + 1]; k += blockDim.x)
        ktom[(size_t)t * ldd + tframe[k]] = (_Float16)tcnt[k];
}

// c_t = mean of log W_rt over the finite values of a strided sample of orientations
// (mask < 1 pixels; 0 elsewhere).
__global__ __launch_bounds__(kBlock) void center_kernel(
    const double *__restrict__ rots, int num_rot, int nsample, const double *__restrict__ model,
    long size, const double *__restrict__ qvals, const double *__restrict__ corr,
    const uint8_t *__restrict__ mask, int num_pix, double *__restrict__ center_t) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_pix) return;
    double c = 0.;
    if (mask[t] < 1) {
        const double center = (double)(size / 2);
        const double q0 = qvals[3 * (size_t)t], q1 = qvals[3 * (size_t)t + 1], q2 = qvals[3 * (size_t)t + 2];
        double s = 0.;
        int n = 0;
        for (int k = 0; k < nsample; ++k) {
            const int r = (int)((long)k * num_rot / nsample);
            const double w = slice_value(model, size, rots + 9 * (size_t)r, q0, q1, q2, corr[t], center);
            if (w > 0.) {
                s += log(w);
                ++n;
            }
        }
        c = n ? s / n : 0.;
    }
    center_t[t] = c;
}

// base_d = sum over the frame's mask < 1 photons of count * c_t (one wave per frame).
__global__ __launch_bounds__(kBlock) void center_base_kernel(const long *__restrict__ off,
                                                             const int *__restrict__ pix,
                                                             const double *__restrict__ cnt,
                                                             const double *__restrict__ center_t,
                                                             int num_data, double *__restrict__ base) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int d = blockIdx.x * kWavesPerBlock + wave;
    if (d >= num_data) return;
    double s = 0.;
    for (long k = off[d] + lane; k < off[d + 1]; k += 64) s += cnt[k] * center_t[pix[k]];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if (lane == 0) base[d] = s;
}

// Expand for the mixed path: W as slice_gen3d, u_r and the non-finite flag as expand_kernel,
// and the centred log W split into 16-bit (T: FP16 or BF16) hi / lo rows lw[c][t] (lanes
// along pixels, so the stores are contiguous without a transpose). mask >= 1 and non-finite
// entries store 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void expand_mp_kernel(
    const double *__restrict__ rots, const double *__restrict__ model, long size,
    const double *__restrict__ qvals, const double *__restrict__ corr,
    const uint8_t *__restrict__ mask, const double *__restrict__ center_t, int num_pix, int r0,
    int C, int pix_per_group, size_t ldp, int two_terms, T *__restrict__ lw_hi,
    T *__restrict__ lw_lo, double *__restrict__ upart, int *__restrict__ ubad) {
    __shared__ double s_rot[kRTile * 9];
    __shared__ double s_tile[64][kRTile + 1];
    __shared__ double s_u[kWavesPerBlock][kRTile];
    __shared__ int s_bad[kWavesPerBlock][kRTile];
    const int tid = threadIdx.x;
    const int c0 = blockIdx.x * kRTile;
    const int g = blockIdx.y;
    for (int i = tid; i < kRTile * 9; i += kBlock) s_rot[i] = rots[(size_t)(r0 + c0) * 9 + i];
    __syncthreads();
    const double center = (double)(size / 2);
    const int lane = tid & 63, wave = tid >> 6;
    const int t_begin = g * pix_per_group;
    const int t_end = min(num_pix, t_begin + pix_per_group);
    double uacc = 0.;
    int bad = 0;
    for (int tb = t_begin; tb < t_end; tb += 64) {
        const int t = tb + lane;
        if (t < t_end) {
            const double q0 = qvals[3 * (size_t)t], q1 = qvals[3 * (size_t)t + 1],
                         q2 = qvals[3 * (size_t)t + 2], cr = corr[t], ct = center_t[t];
            const bool good = mask[t] < 1;
            for (int k = 0; k < kRTile / kWavesPerBlock; ++k) {
                const int rr = wave + kWavesPerBlock * k;
                const double W = slice_value(model, size, &s_rot[rr * 9], q0, q1, q2, cr, center);
                s_tile[lane][rr] = W;
                double x = (good && W > 0.) ? log(W) - ct : 0.;
                const T hi = to_op<T>(x);
                const size_t at = (size_t)(c0 + rr) * ldp + t;
                lw_hi[at] = hi;
                if (two_terms) lw_lo[at] = to_op<T>(x - from_op<T>(hi));
            }
        }
        __syncthreads();
        for (int k = 0; k < 64 / kWavesPerBlock; ++k) {
            const int p = wave + kWavesPerBlock * k;
            const int tt = tb + p;
            if (tt < t_end && mask[tt] < 1) {
                const double W = s_tile[p][lane];
                uacc += W;
                bad |= !(W > 0.);
            }
        }
        __syncthreads();
    }
    s_u[wave][lane] = uacc;
    s_bad[wave][lane] = bad;
    __syncthreads();
    if (tid < kRTile) {
        upart[(size_t)g * C + c0 + tid] = s_u[0][tid] + s_u[1][tid] + s_u[2][tid] + s_u[3][tid];
        ubad[(size_t)g * C + c0 + tid] = s_bad[0][tid] | s_bad[1][tid] | s_bad[2][tid] | s_bad[3][tid];
    }
}

// p16[c][d] = fp16(P[d][r0 + c] * 2^15) where P > 1e-6, else 0; P is the tile-major FP64
// posterior. Block = 64 orientations x 64 frames, transposed through LDS.
__global__ __launch_bounds__(kBlock) void posterior_fp16_kernel(const double *__restrict__ P,
                                                                size_t tsP, int r0, int num_data,
                                                                size_t ldd, double pscale,
                                                                _Float16 *__restrict__ p16) {
    __shared__ _Float16 s[64][64 + 2];
    const int c0 = blockIdx.x * kRTile, d0 = blockIdx.y * 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double *__restrict__ tileP = P + (size_t)((r0 + c0) >> 6) * tsP;
    for (int k = wave; k < 64; k += kWavesPerBlock) {
        const int d = d0 + k;
        double p = (d < num_data) ? tileP[(size_t)d * kRTile + lane] : 0.;
        s[lane][k] = (p > kProbMin) ? (_Float16)(p * pscale) : (_Float16)0.f;
    }
    __syncthreads();
    for (int k = wave; k < 64; k += kWavesPerBlock) {
        const int d = d0 + lane;
        if (d < num_data) p16[(size_t)(c0 + k) * ldd + d] = s[k][lane];
    }
}

// ---- launchers -----------------------------------------------------------------------------

size_t gemm_lds_bytes(int nb) {
    return gemm_lds_elems<_Float16>(nb, kMpGemmBK, nb == 1 ? 2 : 1) * sizeof(_Float16);
}

int launch_mp_gemm(hipStream_t s, int nb, int epi, const _Float16 *A, size_t lda, const _Float16 *B0,
                   const _Float16 *B1, size_t ldb, int K, int Mp, int Np, int M_valid, int N_valid,
                   double *out, size_t ts, const double *row_add, const double *row_div,
                   int div_valid, double scale, int bf16) {
    if (Mp % kGemmBM || Np % kGemmBN || K % kMpGemmBK || lda % 8 || ldb % 8 || (size_t)K > lda ||
        (size_t)K > ldb || M_valid > Mp || N_valid > Np)
        return -1;
    const int m_tiles = Mp / kGemmBM, n_tiles = Np / kGemmBN;
    const dim3 grid(gemm_grid(m_tiles, n_tiles));
    const size_t lds = gemm_lds_bytes(nb);
#define EMC_GEMM(T, NB, EPI)                                                                      \
    mp_gemm_kernel<T, NB, EPI><<<grid, kGemmThreads, lds, s>>>(                                   \
        (const T *)A, lda, (const T *)B0, (const T *)B1, ldb, K, m_tiles, n_tiles, M_valid, N_valid, out, ts, \
        row_add, row_div, div_valid, scale)
    if (bf16 && nb == 2 && epi == 0) EMC_GEMM(__bf16, 2, 0);
    else if (bf16 && nb == 1 && epi == 0) EMC_GEMM(__bf16, 1, 0);
    else if (bf16) return -1;
    else if (nb == 2 && epi == 0) EMC_GEMM(_Float16, 2, 0);
    else if (nb == 1 && epi == 0) EMC_GEMM(_Float16, 1, 0);
    else if (nb == 1 && epi == 1) EMC_GEMM(_Float16, 1, 1);
    else return -1;
#undef EMC_GEMM
    return 0;
}

void launch_build_klik(hipStream_t s, const long *off, const int *pix, const double *cnt, int num_data,
                       size_t ldp, _Float16 *klik, int bf16) {
    if (bf16)
        build_klik_kernel<__bf16><<<num_data, kBlock, 0, s>>>(off, pix, cnt, num_data, ldp, (__bf16 *)klik);
    else
        build_klik_kernel<_Float16><<<num_data, kBlock, 0, s>>>(off, pix, cnt, num_data, ldp, klik);
}

void launch_build_ktom(hipStream_t s, const long *toff, const int *tframe, const double *tcnt,
                       int num_pix, size_t ldd, _Float16 *ktom) {
    build_ktom_kernel<<<num_pix, kBlock, 0, s>>>(toff, tframe, tcnt, num_pix, ldd, ktom);
}

void launch_center(hipStream_t s, const double *rots, int num_rot, const double *model, long size,
                   const double *qvals, const double *corr, const uint8_t *mask, int num_pix,
                   double *center_t) {
    const int nsample = num_rot < 64 ? num_rot : 64;
    center_kernel<<<(num_pix + kBlock - 1) / kBlock, kBlock, 0, s>>>(rots, num_rot, nsample, model, size,
                                                                     qvals, corr, mask, num_pix, center_t);
}

void launch_center_base(hipStream_t s, const long *off, const int *pix, const double *cnt,
                        const double *center_t, int num_data, double *base) {
    center_base_kernel<<<(num_data + kWavesPerBlock - 1) / kWavesPerBlock, kBlock, 0, s>>>(
        off, pix, cnt, center_t, num_data, base);
}

void launch_expand_mp(hipStream_t s, const double *rots, const double *model, long size,
                      const double *qvals, const double *corr, const uint8_t *mask,
                      const double *center_t, int num_pix, int r0, int C, int pix_per_group,
                      size_t ldp, int two_terms, _Float16 *lw_hi, _Float16 *lw_lo, double *upart,
                      int *ubad, int bf16) {
    dim3 grid(C / kRTile, (num_pix + pix_per_group - 1) / pix_per_group);
    if (bf16)
        expand_mp_kernel<__bf16><<<grid, kBlock, 0, s>>>(rots, model, size, qvals, corr, mask, center_t,
                                                         num_pix, r0, C, pix_per_group, ldp, two_terms,
                                                         (__bf16 *)lw_hi, (__bf16 *)lw_lo, upart, ubad);
    else
        expand_mp_kernel<_Float16><<<grid, kBlock, 0, s>>>(rots, model, size, qvals, corr, mask, center_t,
                                                           num_pix, r0, C, pix_per_group, ldp, two_terms, lw_hi,
                                                           lw_lo, upart, ubad);
}

void launch_posterior_fp16(hipStream_t s, const double *P, int num_data, int r0, int C, size_t ldd,
                           double pscale, _Float16 *p16) {
    dim3 grid(C / kRTile, (num_data + 63) / 64);
    posterior_fp16_kernel<<<grid, kBlock, 0, s>>>(P, (size_t)num_data * kRTile, r0, num_data, ldd,
                                                  pscale, p16);
}

}  // namespace emc
