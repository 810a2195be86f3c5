// Matrix-core product tile shared by the mixed-precision EMC path and the orientation
// proposal network: C[m][n] = sum_k A[m][k] * (B0[n][k] (+ B1[n][k])) with 16-bit operands
// (FP16 or BF16), FP32 accumulation, on gfx950 v_mfma_f32_32x32x16_{f16,bf16}.
//
// Block tile kGemmBM x kGemmBN, 8 waves as 4 (m) x 2 (n), wave tile 64 x 128 = 2 x 4 MFMA
// tiles of 32 x 32. K advances kGemmBK elements per step; the next step's operands are
// loaded into registers while the current step's MFMAs read LDS. LDS rows are padded by 8
// elements (80-byte stride) so the 16-byte fragment reads of 32 consecutive rows hit
// distinct banks.
//
// MFMA operand layout (checked on the GPU): lane l supplies A[l % 32][8 (l / 32) + i] and
// B[l % 32][8 (l / 32) + i], i = 0..7; accumulator element v of lane l is
// C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
//
// Shapes: rows of A and B padded to the block tile, K to kGemmBK, leading dimensions
// multiples of 8 elements; the padding holds finite values (zeros).
#pragma once

#include "emc_common.h"

namespace emc {

typedef float floatx16 __attribute__((ext_vector_type(16)));

template <typename T>
struct Mfma;

template <>
struct Mfma<_Float16> {
    typedef _Float16 vec8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ floatx16 run(vec8 a, vec8 b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<__bf16> {
    typedef __bf16 vec8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ floatx16 run(vec8 a, vec8 b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

constexpr int kGemmLdsPad = 8;
constexpr int kGemmRow = kGemmBK + kGemmLdsPad;

template <typename T>
constexpr size_t gemm_lds_elems(int nb, int bk = kGemmBK, int stages = 1) {
    return (size_t)stages * (kGemmBM + nb * kGemmBN) * (bk + kGemmLdsPad);
}

// Block -> output tile. Consecutive blocks go round-robin to the 8 XCDs; each XCD gets a
// contiguous range of tiles, walked in panels of kGemmGroupM m-tiles per n-tile, so the
// blocks resident on one XCD share A and B k-slices in its L2. Launch with
// gemm_grid(m_tiles, n_tiles) blocks.
struct GemmTile {
    size_t m0, n0;
    bool valid;
};

__host__ __device__ inline int gemm_grid(int m_tiles, int n_tiles) {
    const int tiles = m_tiles * n_tiles;
    return (tiles + kNumXcd - 1) / kNumXcd * kNumXcd;
}

__device__ __forceinline__ GemmTile gemm_block_tile(int m_tiles, int n_tiles) {
    const int tiles = m_tiles * n_tiles;
    const int per_xcd = (tiles + kNumXcd - 1) / kNumXcd;
    const int xcd = blockIdx.x % kNumXcd, seq = blockIdx.x / kNumXcd;
    const int tile = xcd * per_xcd + seq;
    GemmTile t{0, 0, seq < per_xcd && tile < tiles};
    if (!t.valid) return t;
    const int panel_tiles = kGemmGroupM * n_tiles;
    const int panel = tile / panel_tiles, in_panel = tile % panel_tiles;
    const int group_m = min(kGemmGroupM, m_tiles - panel * kGemmGroupM);
    t.m0 = (size_t)(panel * kGemmGroupM + in_panel % group_m) * kGemmBM;
    t.n0 = (size_t)(in_panel / group_m) * kGemmBN;
    return t;
}

// Accumulates the block's tile into acc (zeroed here). smem holds
// gemm_lds_elems<T>(NB, BK, DB ? 2 : 1); K advances BK elements per step. With DB the steps
// alternate between two LDS stages: step k's fragments are read from one while step k + 1's
// operands are stored to the other, one barrier per step. With PF the fragments of the next
// 16 elements of K are read from LDS while the current ones feed the MFMAs, the reads spread
// between the MFMAs (scheduling groups: 0x100 LDS read, 0x008 MFMA).
template <typename T, int NB, int BK = kGemmBK, bool DB = false, bool PF = false>
__device__ __forceinline__ void gemm_mainloop(const T *__restrict__ A, size_t lda,
                                              const T *__restrict__ B0, const T *__restrict__ B1,
                                              size_t ldb, int K, size_t m0, size_t n0, T *smem,
                                              floatx16 (&acc)[2][4]) {
    typedef typename Mfma<T>::vec8 vec8;
    constexpr int kRow = BK + kGemmLdsPad;
    // one stage: the A tile [kGemmBM][kRow], then NB consecutive [kGemmBN][kRow] B tiles
    constexpr int kStage = (kGemmBM + NB * kGemmBN) * kRow;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 3, wn = wave >> 2;
    // global -> register staging: each thread moves kLoads 16-byte chunks per operand per
    // step (rows tid / kChunks + (kGemmThreads / kChunks) j, chunk tid % kChunks of the row)
    constexpr int kChunks = BK / 8;
    constexpr int kLoads = kGemmBM * BK / 8 / kGemmThreads;
    static_assert(kGemmBM == kGemmBN, "square block tile");
    static_assert(BK % 16 == 0 && kLoads >= 1, "K step");
    const int lrow = tid / kChunks, lchunk = (tid % kChunks) * 8;
    vec8 ra[kLoads], rb[NB][kLoads];
    auto load = [&](int k0) {
#pragma unroll
        for (int j = 0; j < kLoads; ++j) {
            const size_t row = lrow + (kGemmThreads / kChunks) * j;
            ra[j] = *reinterpret_cast<const vec8 *>(A + (m0 + row) * lda + k0 + lchunk);
            rb[0][j] = *reinterpret_cast<const vec8 *>(B0 + (n0 + row) * ldb + k0 + lchunk);
            if (NB == 2)
                rb[NB - 1][j] = *reinterpret_cast<const vec8 *>(B1 + (n0 + row) * ldb + k0 + lchunk);
        }
    };
    auto store = [&](T *sA) {
        T *sB = sA + kGemmBM * kRow;
#pragma unroll
        for (int j = 0; j < kLoads; ++j) {
            const int row = lrow + (kGemmThreads / kChunks) * j;
            *reinterpret_cast<vec8 *>(sA + row * kRow + lchunk) = ra[j];
#pragma unroll
            for (int b = 0; b < NB; ++b)
                *reinterpret_cast<vec8 *>(sB + (b * kGemmBN + row) * kRow + lchunk) = rb[b][j];
        }
    };
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

    const int frag_row = lane & 31, frag_k = (lane >> 5) * 8;
    constexpr int kReads = 2 + 4 * NB, kMfmas = 8 * NB, kPerRead = kMfmas / kReads;
    auto compute_pf = [&](const T *sA) {
        const T *sB = sA + kGemmBM * kRow;
        vec8 fa[2][2], fb[2][NB][4];
        auto read = [&](int kk, int buf) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
                fa[buf][i] = *reinterpret_cast<const vec8 *>(sA + (wm * 64 + i * 32 + frag_row) * kRow + kk + frag_k);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    fb[buf][b][j] = *reinterpret_cast<const vec8 *>(
                        sB + (b * kGemmBN + wn * 128 + j * 32 + frag_row) * kRow + kk + frag_k);
        };
        read(0, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, kReads, 0);
#pragma unroll
        for (int st = 0; st < BK / 16; ++st) {
            const int cur = st & 1;
            const bool next = st + 1 < BK / 16;
            if (next) read((st + 1) * 16, cur ^ 1);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::run(fa[cur][i], fb[cur][b][j], acc[i][j]);
            if (next) {
#pragma unroll
                for (int r = 0; r < kReads; ++r) {
                    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x008, kPerRead, 0);
                }
                if (kMfmas > kPerRead * kReads)
                    __builtin_amdgcn_sched_group_barrier(0x008, kMfmas - kPerRead * kReads, 0);
            } else {
                __builtin_amdgcn_sched_group_barrier(0x008, kMfmas, 0);
            }
        }
    };
    auto compute = [&](const T *sA) {
        if constexpr (PF) {
            compute_pf(sA);
            return;
        }
        const T *sB = sA + kGemmBM * kRow;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            vec8 fa[2], fb[NB][4];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                fa[i] = *reinterpret_cast<const vec8 *>(sA + (wm * 64 + i * 32 + frag_row) * kRow + kk + frag_k);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    fb[b][j] = *reinterpret_cast<const vec8 *>(
                        sB + (b * kGemmBN + wn * 128 + j * 32 + frag_row) * kRow + kk + frag_k);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::run(fa[i], fb[b][j], acc[i][j]);
        }
    };
    load(0);
    if (!DB) {
        for (int k0 = 0; k0 < K; k0 += BK) {
            __syncthreads();  // previous step's fragment reads are done
            store(smem);
            __syncthreads();
            if (k0 + BK < K) load(k0 + BK);
            compute(smem);
        }
    } else {
        store(smem);
        __syncthreads();
        int stage = 0;
        for (int k0 = 0; k0 < K; k0 += BK) {
            const bool more = k0 + BK < K;
            if (more) load(k0 + BK);
            compute(smem + stage * kStage);
            // the other stage was last read by the previous step, which every wave finished
            // before the previous barrier
            if (more) store(smem + (stage ^ 1) * kStage);
            __syncthreads();
            stage ^= 1;
        }
    }
}

// Calls f(m, n, value) for every accumulator element of the block's tile.
template <typename F>
__device__ __forceinline__ void gemm_for_each(size_t m0, size_t n0, const floatx16 (&acc)[2][4], F f) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 3, wn = wave >> 2;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t n = n0 + wn * 128 + j * 32 + (lane & 31);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const size_t m = m0 + wm * 64 + i * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                f(m, n, acc[i][j][v]);
            }
        }
}

}  // namespace emc
