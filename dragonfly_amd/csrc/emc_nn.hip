// Orientation-proposal network kernels: a BF16 multilayer perceptron trained and evaluated on
// the GPU (forward, backward, optimiser), plus the polar photon histogram it reads and the
// top-k selection of its output.
//
// Layout: activations are row-major [batch][features] BF16; weights W[out][in] BF16 with an
// FP32 master copy and a transposed BF16 copy Wt[in][out] for the data gradient; all leading
// dimensions are padded to the product tile (emc_gemm.h), padding entries are zero.
//   forward      Y  = relu(X W^T + b)            product A = X,    B = W
//   data grad    dX = (dY W) * (H > 0)           product A = dY,   B = Wt
//   weight grad  dW = dY^T X, db = sum_rows dY   product A = dY^T, B = X^T
// Products accumulate in FP32 on the matrix cores (v_mfma_f32_32x32x16_bf16).
#include <algorithm>
#include <cfloat>
#include <climits>

#include "emc_common.h"
#include "emc_gemm.h"
#include "emc_kernels.h"

namespace emc {

typedef __bf16 bf16;

// ---- products ------------------------------------------------------------------------------

// Y[m][n] (m < M_valid, n < N_valid) = act(sum_k X[m][k] W[n][k] + bias[n]); out_f32 selects
// FP32 output (logits) instead of BF16; relu applies max(0, .).
template <bool kF32Out>
__global__ __launch_bounds__(kGemmThreads) void nn_fwd_kernel(
    const bf16 *__restrict__ X, size_t ldx, const bf16 *__restrict__ W, size_t ldw, int K,
    int m_tiles, int n_tiles, int M_valid, int N_valid, const float *__restrict__ bias, int relu,
    void *__restrict__ out, size_t ldo) {
    extern __shared__ bf16 nn_smem[];
    const GemmTile tile = gemm_block_tile(m_tiles, n_tiles);
    if (!tile.valid) return;
    floatx16 acc[2][4];
    gemm_mainloop<bf16, 1>(X, ldx, W, nullptr, ldw, K, tile.m0, tile.n0, nn_smem, acc);
    gemm_for_each(tile.m0, tile.n0, acc, [&](size_t m, size_t n, float a) {
        if (m >= (size_t)M_valid || n >= (size_t)N_valid) return;
        float y = a + bias[n];
        if (relu) y = y > 0.f ? y : 0.f;
        if (kF32Out) static_cast<float *>(out)[m * ldo + n] = y;
        else static_cast<bf16 *>(out)[m * ldo + n] = (bf16)y;
    });
}

// dX[m][n] = (sum_k dY[m][k] Wt[n][k]) * (H[m][n] > 0) (H null: no mask), BF16.
__global__ __launch_bounds__(kGemmThreads) void nn_bwd_data_kernel(
    const bf16 *__restrict__ dY, size_t ldy, const bf16 *__restrict__ Wt, size_t ldwt, int K,
    int m_tiles, int n_tiles, int M_valid, int N_valid, const bf16 *__restrict__ H, size_t ldh,
    bf16 *__restrict__ dX, size_t ldx) {
    extern __shared__ bf16 nn_smem[];
    const GemmTile tile = gemm_block_tile(m_tiles, n_tiles);
    if (!tile.valid) return;
    floatx16 acc[2][4];
    gemm_mainloop<bf16, 1>(dY, ldy, Wt, nullptr, ldwt, K, tile.m0, tile.n0, nn_smem, acc);
    gemm_for_each(tile.m0, tile.n0, acc, [&](size_t m, size_t n, float a) {
        if (m >= (size_t)M_valid || n >= (size_t)N_valid) return;
        if (H && !((float)H[m * ldh + n] > 0.f)) a = 0.f;
        dX[m * ldx + n] = (bf16)a;
    });
}

// dW[m][n] = sum_k dYt[m][k] Xt[n][k] (FP32).
__global__ __launch_bounds__(kGemmThreads) void nn_bwd_weight_kernel(
    const bf16 *__restrict__ dYt, size_t lda, const bf16 *__restrict__ Xt, size_t ldb, int K,
    int m_tiles, int n_tiles, int M_valid, int N_valid, float *__restrict__ dW, size_t ldo) {
    extern __shared__ bf16 nn_smem[];
    const GemmTile tile = gemm_block_tile(m_tiles, n_tiles);
    if (!tile.valid) return;
    floatx16 acc[2][4];
    gemm_mainloop<bf16, 1>(dYt, lda, Xt, nullptr, ldb, K, tile.m0, tile.n0, nn_smem, acc);
    gemm_for_each(tile.m0, tile.n0, acc, [&](size_t m, size_t n, float a) {
        if (m >= (size_t)M_valid || n >= (size_t)N_valid) return;
        dW[m * ldo + n] = a;
    });
}

// ---- elementwise and reductions --------------------------------------------------------------

// out[c][r] = in[r][c] for r < rows, c < cols (BF16), 64 x 64 tiles through LDS.
__global__ __launch_bounds__(kBlock) void nn_transpose_kernel(const bf16 *__restrict__ in, size_t ldi,
                                                              int rows, int cols, bf16 *__restrict__ out,
                                                              size_t ldo) {
    __shared__ bf16 s[64][66];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = wave; k < 64; k += kWavesPerBlock) {
        const int r = r0 + k, c = c0 + lane;
        s[k][lane] = (r < rows && c < cols) ? in[(size_t)r * ldi + c] : (bf16)0.f;
    }
    __syncthreads();
    for (int k = wave; k < 64; k += kWavesPerBlock) {
        const int c = c0 + k, r = r0 + lane;
        if (c < cols && r < rows) out[(size_t)c * ldo + r] = s[lane][k];
    }
}

// Soft-target cross entropy per row: p = softmax(z[0..n)), loss = -sum_j y_j log p_j,
// dZ = (p - y) * grad_scale (BF16; columns n..ldz stay 0). One block per row.
__global__ __launch_bounds__(kBlock) void nn_softmax_xent_kernel(
    const float *__restrict__ Z, size_t ldz, const float *__restrict__ Y, size_t ldy, int n,
    float grad_scale, float *__restrict__ loss, bf16 *__restrict__ dZ, size_t ldd) {
    __shared__ float s[kBlock];
    const int row = blockIdx.x, tid = threadIdx.x;
    const float *z = Z + (size_t)row * ldz;
    const float *y = Y ? Y + (size_t)row * ldy : nullptr;
    float mx = -FLT_MAX;
    for (int j = tid; j < n; j += kBlock) mx = fmaxf(mx, z[j]);
    s[tid] = mx;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) s[tid] = fmaxf(s[tid], s[tid + w]);
        __syncthreads();
    }
    mx = s[0];
    __syncthreads();
    float se = 0.f;
    for (int j = tid; j < n; j += kBlock) se += __expf(z[j] - mx);
    s[tid] = se;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) s[tid] += s[tid + w];
        __syncthreads();
    }
    const float lse = mx + __logf(s[0]);
    __syncthreads();
    if (!y) return;
    float l = 0.f;
    for (int j = tid; j < n; j += kBlock) {
        const float logp = z[j] - lse;
        l -= y[j] * logp;
        if (dZ) dZ[(size_t)row * ldd + j] = (bf16)((__expf(logp) - y[j]) * grad_scale);
    }
    s[tid] = l;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) s[tid] += s[tid + w];
        __syncthreads();
    }
    if (tid == 0 && loss) loss[row] = s[0];
}

// db[n] = sum over rows of dY[row][n] (BF16 in, FP32 out). Block = 64 columns x 16 waves: wave
// w sums rows w, w + 16, ... of its lane's column (eight loads in flight), then the 16 partial
// sums are added in wave order.
constexpr int kColsumWaves = 16;
__global__ __launch_bounds__(64 * kColsumWaves) void nn_colsum_kernel(const bf16 *__restrict__ dY, size_t ldy,
                                                                      int rows, int cols, float *__restrict__ db) {
    __shared__ float part[kColsumWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + lane;
    float acc = 0.f;
    if (c < cols) {
        int r = wave;
        for (; r + 7 * kColsumWaves < rows; r += 8 * kColsumWaves) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = (float)dY[(size_t)(r + u * kColsumWaves) * ldy + c];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u];
        }
        for (; r < rows; r += kColsumWaves) acc += (float)dY[(size_t)r * ldy + c];
    }
    part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && c < cols) {
        float t = 0.f;
        for (int w = 0; w < kColsumWaves; ++w) t += part[w][lane];
        db[c] = t;
    }
}

// Adam step on an FP32 master tensor [rows][cols] (leading dimension ld) with FP32 gradient;
// refreshes the BF16 copy (same layout) and, when wt is given, the transposed BF16 copy
// wt[c][r] (leading dimension ldt). step = 1-based step count.
__global__ void nn_adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                               float *__restrict__ v, int rows, int cols, size_t ld, float lr, float b1,
                               float b2, float eps, float wd, int step, bf16 *__restrict__ wb,
                               bf16 *__restrict__ wt, size_t ldt) {
    const size_t n = (size_t)rows * cols;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / cols, c = i % cols, at = r * ld + c;
        const float gi = g[at] + wd * w[at];
        const float mi = b1 * m[at] + (1.f - b1) * gi;
        const float vi = b2 * v[at] + (1.f - b2) * gi * gi;
        m[at] = mi;
        v[at] = vi;
        const float wi = w[at] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[at] = wi;
        if (wb) wb[at] = (bf16)wi;
        if (wt) wt[c * ldt + r] = (bf16)wi;
    }
}

// Polar photon histogram of each frame: counts of the frame's photons (likelihood list,
// mask < 1) per bin (bin_of[pixel], -1 = none), scaled to unit sum, then log(1 + 1000 x);
// one block per frame, histogram in LDS. out [frame][ldo] BF16.
__global__ __launch_bounds__(kBlock) void nn_polar_hist_kernel(
    const long *__restrict__ off, const int *__restrict__ pix, const double *__restrict__ cnt,
    const int *__restrict__ bin_of, int nbins, int num_frames, bf16 *__restrict__ out, size_t ldo) {
    extern __shared__ float hist[];
    const int d = blockIdx.x, tid = threadIdx.x;
    if (d >= num_frames) return;
    for (int i = tid; i < nbins; i += kBlock) hist[i] = 0.f;
    __syncthreads();
    float tot = 0.f;
    for (long k = off[d] + tid; k < off[d + 1]; k += kBlock) {
        const int b = bin_of[pix[k]];
        if (b < 0) continue;
        atomicAdd(&hist[b], (float)cnt[k]);
        tot += (float)cnt[k];
    }
    __shared__ float s[kBlock];
    s[tid] = tot;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) s[tid] += s[tid + w];
        __syncthreads();
    }
    const float inv = s[0] > 0.f ? 1.f / s[0] : 0.f;
    for (int i = tid; i < nbins; i += kBlock)
        out[(size_t)d * ldo + i] = (bf16)__logf(1.f + 1000.f * hist[i] * inv);
}

// Indices of the k largest of z[row][0..n) in descending order (k <= 64), one wave per row:
// k rounds of a wave-wide arg-max over the lane-strided row (ties: lower index). When mass is
// given, mass[row] = the softmax probability of the selected entries (kProb: z are
// non-negative masses, and mass[row] = their sum over the selected entries).
template <bool kProb>
__global__ __launch_bounds__(kBlock) void nn_topk_kernel(const float *__restrict__ Z, size_t ldz, int rows,
                                                         int n, int k, int *__restrict__ idx,
                                                         float *__restrict__ mass) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * kWavesPerBlock + wave;
    if (row >= rows) return;
    const float *z = Z + (size_t)row * ldz;
    float zmax = -FLT_MAX;
    for (int j = lane; j < n; j += 64) zmax = fmaxf(zmax, z[j]);
    for (int o = 32; o > 0; o >>= 1) zmax = fmaxf(zmax, __shfl_xor(zmax, o));
    float zsum = 0.f;
    if (!kProb) {
        for (int j = lane; j < n; j += 64) zsum += __expf(z[j] - zmax);
        for (int o = 32; o > 0; o >>= 1) zsum += __shfl_xor(zsum, o);
    }
    float captured = 0.f;
    unsigned long long taken_lo = 0;  // lane-local bitmap over this lane's (up to 64) columns
    unsigned long long taken_hi = 0;
    for (int round = 0; round < k; ++round) {
        float best = -FLT_MAX;
        int bi = INT_MAX;
        int slot = 0;
        for (int j = lane, s = 0; j < n; j += 64, ++s) {
            const bool taken = s < 64 ? (taken_lo >> s) & 1ull : (taken_hi >> (s - 64)) & 1ull;
            if (!taken && (z[j] > best || (z[j] == best && j < bi))) {
                best = z[j];
                bi = j;
                slot = s;
            }
        }
        float wb = best;
        int wi = bi;
        for (int o = 32; o > 0; o >>= 1) {
            const float ob = __shfl_xor(wb, o);
            const int oi = __shfl_xor(wi, o);
            if (ob > wb || (ob == wb && oi < wi)) {
                wb = ob;
                wi = oi;
            }
        }
        if (wi == bi && bi != INT_MAX) {
            if (slot < 64) taken_lo |= 1ull << slot;
            else taken_hi |= 1ull << (slot - 64);
        }
        if (lane == 0) idx[(size_t)row * k + round] = (wi == INT_MAX) ? -1 : wi;
        if (wi != INT_MAX) captured += kProb ? wb : __expf(wb - zmax);
    }
    if (mass && lane == 0) mass[row] = kProb ? captured : captured / zsum;
}

// ---- launchers --------------------------------------------------------------------------------

namespace {
bool tiles_ok(int M, int N, int K, size_t lda, size_t ldb) {
    return M % kGemmBM == 0 && N % kGemmBN == 0 && K % kGemmBK == 0 && lda % 8 == 0 && ldb % 8 == 0 &&
           (size_t)K <= lda && (size_t)K <= ldb;
}
}  // namespace

int launch_nn_fwd(hipStream_t s, const bf16 *X, size_t ldx, const bf16 *W, size_t ldw, int M, int N, int K,
                  int M_valid, int N_valid, const float *bias, int relu, int out_f32, void *out, size_t ldo) {
    if (!tiles_ok(M, N, K, ldx, ldw)) return -1;
    const int mt = M / kGemmBM, nt = N / kGemmBN;
    const size_t lds = gemm_lds_elems<bf16>(1) * sizeof(bf16);
    if (out_f32)
        nn_fwd_kernel<true><<<gemm_grid(mt, nt), kGemmThreads, lds, s>>>(X, ldx, W, ldw, K, mt, nt, M_valid,
                                                                         N_valid, bias, relu, out, ldo);
    else
        nn_fwd_kernel<false><<<gemm_grid(mt, nt), kGemmThreads, lds, s>>>(X, ldx, W, ldw, K, mt, nt, M_valid,
                                                                          N_valid, bias, relu, out, ldo);
    return 0;
}

int launch_nn_bwd_data(hipStream_t s, const bf16 *dY, size_t ldy, const bf16 *Wt, size_t ldwt, int M, int N,
                       int K, int M_valid, int N_valid, const bf16 *H, size_t ldh, bf16 *dX, size_t ldx) {
    if (!tiles_ok(M, N, K, ldy, ldwt)) return -1;
    const int mt = M / kGemmBM, nt = N / kGemmBN;
    nn_bwd_data_kernel<<<gemm_grid(mt, nt), kGemmThreads, gemm_lds_elems<bf16>(1) * sizeof(bf16), s>>>(
        dY, ldy, Wt, ldwt, K, mt, nt, M_valid, N_valid, H, ldh, dX, ldx);
    return 0;
}

int launch_nn_bwd_weight(hipStream_t s, const bf16 *dYt, size_t lda, const bf16 *Xt, size_t ldb, int M, int N,
                         int K, int M_valid, int N_valid, float *dW, size_t ldo) {
    if (!tiles_ok(M, N, K, lda, ldb)) return -1;
    const int mt = M / kGemmBM, nt = N / kGemmBN;
    nn_bwd_weight_kernel<<<gemm_grid(mt, nt), kGemmThreads, gemm_lds_elems<bf16>(1) * sizeof(bf16), s>>>(
        dYt, lda, Xt, ldb, K, mt, nt, M_valid, N_valid, dW, ldo);
    return 0;
}

void launch_nn_transpose(hipStream_t s, const bf16 *in, size_t ldi, int rows, int cols, bf16 *out, size_t ldo) {
    dim3 grid((cols + 63) / 64, (rows + 63) / 64);
    nn_transpose_kernel<<<grid, kBlock, 0, s>>>(in, ldi, rows, cols, out, ldo);
}

void launch_nn_softmax_xent(hipStream_t s, const float *Z, size_t ldz, const float *Y, size_t ldy, int rows,
                            int n, float grad_scale, float *loss, bf16 *dZ, size_t ldd) {
    nn_softmax_xent_kernel<<<rows, kBlock, 0, s>>>(Z, ldz, Y, ldy, n, grad_scale, loss, dZ, ldd);
}

void launch_nn_colsum(hipStream_t s, const bf16 *dY, size_t ldy, int rows, int cols, float *db) {
    nn_colsum_kernel<<<(cols + 63) / 64, 64 * kColsumWaves, 0, s>>>(dY, ldy, rows, cols, db);
}

void launch_nn_adam(hipStream_t s, float *w, const float *g, float *m, float *v, int rows, int cols, size_t ld,
                    float lr, float b1, float b2, float eps, float wd, int step, bf16 *wb, bf16 *wt, size_t ldt) {
    const size_t n = (size_t)rows * cols;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    nn_adam_kernel<<<grid, 256, 0, s>>>(w, g, m, v, rows, cols, ld, lr, b1, b2, eps, wd, step, wb, wt, ldt);
}

void launch_nn_polar_hist(hipStream_t s, const long *off, const int *pix, const double *cnt, const int *bin_of,
                          int nbins, int num_frames, bf16 *out, size_t ldo) {
    nn_polar_hist_kernel<<<num_frames, kBlock, nbins * sizeof(float), s>>>(off, pix, cnt, bin_of, nbins,
                                                                           num_frames, out, ldo);
}

void launch_nn_topk(hipStream_t s, const float *Z, size_t ldz, int rows, int n, int k, int *idx, float *mass) {
    nn_topk_kernel<false><<<(rows + kWavesPerBlock - 1) / kWavesPerBlock, kBlock, 0, s>>>(Z, ldz, rows, n, k, idx,
                                                                                          mass);
}

void launch_prob_topk(hipStream_t s, const float *Z, size_t ldz, int rows, int n, int k, int *idx, float *mass) {
    if (rows <= 0) return;
    nn_topk_kernel<true><<<(rows + kWavesPerBlock - 1) / kWavesPerBlock, kBlock, 0, s>>>(Z, ldz, rows, n, k, idx,
                                                                                         mass);
}

}  // namespace emc

// ---- C ABI (device pointers, work ordered on the given HIP stream) ----------------------------
extern "C" {

typedef __bf16 emc_bf16;

int emc_nn_forward(void *stream, const void *X, size_t ldx, const void *W, size_t ldw, int M, int N, int K,
                   int M_valid, int N_valid, const float *bias, int relu, int out_f32, void *out, size_t ldo) {
    if (!X || !W || !bias || !out) return -1;
    return emc::launch_nn_fwd((hipStream_t)stream, (const emc_bf16 *)X, ldx, (const emc_bf16 *)W, ldw, M, N, K,
                              M_valid, N_valid, bias, relu, out_f32, out, ldo);
}

int emc_nn_backward_data(void *stream, const void *dY, size_t ldy, const void *Wt, size_t ldwt, int M, int N,
                         int K, int M_valid, int N_valid, const void *H, size_t ldh, void *dX, size_t ldx) {
    if (!dY || !Wt || !dX) return -1;
    return emc::launch_nn_bwd_data((hipStream_t)stream, (const emc_bf16 *)dY, ldy, (const emc_bf16 *)Wt, ldwt,
                                   M, N, K, M_valid, N_valid, (const emc_bf16 *)H, ldh, (emc_bf16 *)dX, ldx);
}

int emc_nn_backward_weight(void *stream, const void *dYt, size_t lda, const void *Xt, size_t ldb, int M, int N,
                           int K, int M_valid, int N_valid, float *dW, size_t ldo) {
    if (!dYt || !Xt || !dW) return -1;
    return emc::launch_nn_bwd_weight((hipStream_t)stream, (const emc_bf16 *)dYt, lda, (const emc_bf16 *)Xt, ldb,
                                     M, N, K, M_valid, N_valid, dW, ldo);
}

int emc_nn_transpose(void *stream, const void *in, size_t ldi, int rows, int cols, void *out, size_t ldo) {
    if (!in || !out || rows < 0 || cols < 0) return -1;
    emc::launch_nn_transpose((hipStream_t)stream, (const emc_bf16 *)in, ldi, rows, cols, (emc_bf16 *)out, ldo);
    return 0;
}

int emc_nn_softmax_xent(void *stream, const float *Z, size_t ldz, const float *Y, size_t ldy, int rows, int n,
                        float grad_scale, float *loss, void *dZ, size_t ldd) {
    if (!Z || rows < 0 || n <= 0) return -1;
    emc::launch_nn_softmax_xent((hipStream_t)stream, Z, ldz, Y, ldy, rows, n, grad_scale, loss, (emc_bf16 *)dZ,
                                ldd);
    return 0;
}

int emc_nn_colsum(void *stream, const void *dY, size_t ldy, int rows, int cols, float *db) {
    if (!dY || !db) return -1;
    emc::launch_nn_colsum((hipStream_t)stream, (const emc_bf16 *)dY, ldy, rows, cols, db);
    return 0;
}

int emc_nn_adam(void *stream, float *w, const float *g, float *m, float *v, int rows, int cols, size_t ld,
                float lr, float b1, float b2, float eps, float wd, int step, void *wb, void *wt, size_t ldt) {
    if (!w || !g || !m || !v || step < 1) return -1;
    emc::launch_nn_adam((hipStream_t)stream, w, g, m, v, rows, cols, ld, lr, b1, b2, eps, wd, step,
                        (emc_bf16 *)wb, (emc_bf16 *)wt, ldt);
    return 0;
}

int emc_nn_polar_hist(void *stream, const long *off, const int *pix, const double *cnt, const int *bin_of,
                      int nbins, int num_frames, void *out, size_t ldo) {
    if (!off || !pix || !cnt || !bin_of || !out || nbins <= 0 || nbins > 16384 || (size_t)nbins > ldo) return -1;
    emc::launch_nn_polar_hist((hipStream_t)stream, off, pix, cnt, bin_of, nbins, num_frames, (emc_bf16 *)out, ldo);
    return 0;
}

int emc_nn_topk(void *stream, const float *Z, size_t ldz, int rows, int n, int k, int *idx, float *mass) {
    if (!Z || !idx || k < 1 || k > 64 || k > n || n > 128 * 64) return -1;
    emc::launch_nn_topk((hipStream_t)stream, Z, ldz, rows, n, k, idx, mass);
    return 0;
}

}  // extern "C"
