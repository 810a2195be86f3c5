// Host launchers of the EMC kernels (defined in emc_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace emc {

void launch_expand(hipStream_t s, const double *rots, const double *model, long size,
                   const double *qvals, const double *corr, const uint8_t *mask, int num_pix,
                   int r0, int C, int pix_per_group, double *logw, double *upart, int *ubad,
                   const int *slot_rot = nullptr, int modes = 1, int raw = 0);
// FP32 log W from an FP32 copy of the model (FP32-gather path)
void launch_expand(hipStream_t s, const double *rots, const float *model, long size,
                   const double *qvals, const double *corr, const uint8_t *mask, int num_pix,
                   int r0, int C, int pix_per_group, float *logw, double *upart, int *ubad,
                   const int *slot_rot = nullptr, int modes = 1);
// out[i] = (float)in[i], i < n
void launch_copy_f32(hipStream_t s, const double *in, size_t n, float *out);
// FP32 copy of `modes` volumes in 4^3-voxel bricks (the FP32-gather expand's model)
size_t model_brick_elems(long size, int modes);
void launch_copy_f32_brick(hipStream_t s, const double *in, long size, int modes, float *out);
void launch_to_f32(hipStream_t s, const double *P, int num_data, int r0, int C, float *out);
void launch_likelihood(hipStream_t s, const float *logw, int C, int num_pix, const long *off,
                       const int *pix, const double *cnt, const uint8_t *blacklist, int num_data,
                       double *G, int r0, int tiles, int unroll);
// Likelihood over X pixel ranges (lists split per range, off8[d * (X + 1) + x]) into partials
// part[X][C / 64][num_data][64], then added in range order into the chunk's tiles of G.
void launch_likelihood_split(hipStream_t s, const double *logw, int C, int num_pix, const long *off8, int X,
                             const int *pix, const double *cnt, const uint8_t *blacklist, int num_data,
                             double *part, double *G, int r0);
// fmap: row d of G holds frame fmap[d] (photon lists and blacklist indexed by fmap[d]).
void launch_likelihood_mapped(hipStream_t s, const double *logw, int C, int num_pix, const long *off,
                              const int *pix, const double *cnt, const uint8_t *blacklist, const int *fmap,
                              int num_rows, double *G, int r0);
void launch_likelihood_mapped(hipStream_t s, const float *logw, int C, int num_pix, const long *off,
                              const int *pix, const double *cnt, const uint8_t *blacklist, const int *fmap,
                              int num_rows, double *G, int r0);
void launch_tomogram_parts_f32(hipStream_t s, const float *P32, int num_data, int C, const long *toff2,
                               int nparts, const int *tframe, const double *tcnt, int num_pix,
                               double *partial);
// the same with FP32 partial sums (each the FP64 sum rounded once)
void launch_tomogram_parts_f32(hipStream_t s, const float *P32, int num_data, int C, const long *toff2,
                               int nparts, const int *tframe, const double *tcnt, int num_pix,
                               float *partial);
void launch_u_reduce(hipStream_t s, const double *upart, const int *ubad, int ngroups, int C,
                     int r0, int num_rot, double *u, int *rbad);
void launch_rescale(hipStream_t s, const double *quats, const double *u, int num_rot,
                    double mean_count, double *scalars, int modes = 1);
void launch_likelihood(hipStream_t s, const double *logw, int C, int num_pix, const long *off,
                       const int *pix, const double *cnt, const uint8_t *blacklist, int num_data,
                       double *G, int r0, int tiles, int unroll, int xcd_order = 0);
void launch_frame_max(hipStream_t s, double *R, int num_rot, int num_data,
                      const double *logwr, const double *u, const double *scale, int use_scale,
                      const double *scalars, const uint8_t *blacklist, const uint8_t *dense,
                      const int *rbad, const int *qmap, const uint32_t *relbits, int words,
                      double *max_exp, int *rmax);
void launch_frame_norm(hipStream_t s, const double *R, int num_rot, int num_data,
                       const double *beta, const uint8_t *blacklist, const double *max_exp,
                       const int *rmax, double *p_norm);
// mass[d] = sum of the posterior P_dr (tile-major R) over r in idx[off[d] .. off[d + 1])
void launch_subset_mass(hipStream_t s, const double *R, int num_data, const long *off, const int *idx,
                        double *mass);
// Estimated posterior mass outside each frame's cells (coarse full search Pc [num_coarse/64][N][64]
// against the fine maximum Mf / normaliser pnf inside the cells); 0 where skip[d].
// Top-k columns (descending, ties to the lower index; k <= 64, n <= 8192) of every row of a
// non-negative FP32 matrix, with the sum of the selected entries (emc_nn.hip).
void launch_prob_topk(hipStream_t s, const float *Z, size_t ldz, int rows, int n, int k, int *idx, float *mass);
// Persistent expand + pair likelihood of one orientation chunk of C columns (a multiple of
// 64) starting at slot r0 = 64 t0 (emc_subset.hip): blocks take whole tiles from a counter
// (device scratch, reset by the launch), expand them and evaluate their pairs.
void launch_fused_expand_likelihood(hipStream_t s, const double *rots, const double *model, long size,
                                    const double *qvals, const double *corr, const uint8_t *mask, int num_pix, int r0,
                                    int C, int pix_per_group, double *logw, double *upart, int *ubad,
                                    const int *slot_rot, const long *off, const int *pix, const double *cnt,
                                    const int *pair_frame, const int *pair_tile, const long *tile_off,
                                    const int *tile_pair, int t0, unsigned *counter, double *G);
void launch_fused_expand_likelihood(hipStream_t s, const double *rots, const float *model, long size,
                                    const double *qvals, const double *corr, const uint8_t *mask, int num_pix, int r0,
                                    int C, int pix_per_group, float *logw, double *upart, int *ubad,
                                    const int *slot_rot, const long *off, const int *pix, const double *cnt,
                                    const int *pair_frame, const int *pair_tile, const long *tile_off,
                                    const int *tile_pair, int t0, unsigned *counter, double *G);
// the fewest cells to add to each frame the judge rejects (judge_widen_kernel); num_coarse
// <= kWidenCells; add: num_data * max_add; then compacted to the CSR offsets off
constexpr int kWidenCells = 4096;
void launch_judge_widen(hipStream_t s, const double *Pc, int num_data, int num_coarse, const double *Mc,
                        const double *pnc, const double *Mf, const double *pnf, const double *beta, const uint8_t *skip,
                        const uint32_t *bits, int words, const double *log_n, const double *log_wr, double outside_max,
                        double target, int max_add, double *out, int *num_add, int *add);
void launch_widen_compact(hipStream_t s, const int *add, const int *num_add, const long *off, int num_data,
                          int max_add, int *out);
void launch_judge_outside(hipStream_t s, const double *Pc, int num_data, int num_coarse, const double *Mc,
                          const double *pnc, const double *Mf, const double *pnf, const double *beta,
                          const uint8_t *skip, const uint32_t *bits, int words, const double *log_n,
                          const double *log_wr, double prob_floor, double *out);
void launch_posterior(hipStream_t s, double *R, int num_rot, int num_data,
                      const double *logwr, const double *u, const double *beta,
                      const double *sum_fact, const uint8_t *blacklist, const double *max_exp,
                      const int *rmax, const double *p_norm, const double *scalars,
                      int want_psum_d, double *likelihood, double *info, double *psum_d, int modes = 1);
void launch_occupancy(hipStream_t s, const double *P, int num_rot, int num_data, int modes,
                      double *quat_norm);
// background-aware path (emc_background.hip)
void launch_likelihood_bg(hipStream_t s, const double *W, int C, int num_pix, const long *off, const int *pix,
                          const double *cnt, const uint8_t *blacklist, int num_data, const double *scale,
                          const double *bgscale, const double *bg, double *G, int r0);
void launch_scalemax(hipStream_t s, const double *P, int num_data, int Rp, const double *scale, double *part,
                     int nseg, double *out);
void launch_tomogram_bg(hipStream_t s, int first, const double *P, int num_data, int r0, int C, int num_rot,
                        const long *tom_off, const int *tom_frame, const double *tom_cnt, const double *scale,
                        const double *bgscale, const double *bg, const double *powder, const uint8_t *raw_mask,
                        int num_pix, double mean_count, const double *psum_r, const double *scalemax,
                        int *last_round, double *view, int rmajor);
void launch_gradient_d(hipStream_t s, const double *W, int num_pix, const double *P, int num_data, int r0, int C,
                       int num_rot, const long *off, const int *pix, const double *cnt, const uint8_t *dmask,
                       const double *sc, const double *bgscale, const double *bg, double rescale, double *part,
                       double *gacc);
void launch_sparse_count(hipStream_t s, const double *P, int num_rot, int num_data, int *counts);
void launch_sparse_fill(hipStream_t s, const double *P, int num_rot, int num_data, int d0, int nd,
                        const long *off, int *place, double *prob);
// the same over row-major rows P[row * num_rot + r]
void launch_sparse_count_rows(hipStream_t s, const double *P, int num_rot, int nrows, int *counts);
void launch_sparse_fill_rows(hipStream_t s, const double *P, int num_rot, int nrows, const long *off, int *place,
                             double *prob);
void launch_psum_r(hipStream_t s, const double *P, int num_rot, int num_data,
                   int frames_per_part, const double *scale, int need_scaling,
                   const uint8_t *blacklist, double *part, double *psum_r);
void launch_tomogram(hipStream_t s, const double *P, int num_data, int r0, int C,
                     const long *toff, const int *tframe, const double *tcnt, const double *psum_r,
                     int num_rot, int num_pix, double *view, int tiles);
void launch_tomogram_tiles(hipStream_t s, const double *P, int num_data, int r0, int C,
                           const long *ft_off, const int *ft_frame, const long *fe_off,
                           const int *fe_code, const double *psum_r, int num_rot, int num_pix,
                           double *view);
void launch_tomogram_parts(hipStream_t s, const double *P, int num_data, int r0, int C,
                           const long *toff2, int nparts, const int *tframe, const double *tcnt,
                           const double *psum_r, int num_rot, int num_pix, double *partial,
                           double *view, int width32, int unroll, int seq);
void launch_tomogram_reduce(hipStream_t s, const double *partial, int nparts, int r0, int C,
                            const double *psum_r, int num_rot, int num_pix, double *view);
void launch_tomogram_reduce_rmajor(hipStream_t s, const double *partial, int nparts, int r0, int C,
                                   const double *psum_r, int num_rot, int num_pix, double *view);
// ---- brick merge (emc_merge.hip) ----
// per-brick item lists: pass 0 counts, pass 1 writes the item ids and, beside each, the mask of
// the item's pixels with a trilinear corner inside the brick
void launch_brick_items(hipStream_t s, const double *rots, int r0, int Cc, int num_rot,
                        const double *qvals, const uint8_t *mask, int num_pix, long size, int npatch,
                        int nbx, int pass, int *counts, const long *off, int *fill, int *list,
                        unsigned long long *lmask, const int *slot_rot, int modes = 1);
void launch_merge_bricks(hipStream_t s, const int4 *units, int num_units, const int *list,
                         const unsigned long long *lmask, const double *rots, int r0, int num_rot,
                         const double *view, const double *psum_r, int all_on, const double *qvals,
                         const double *corr, const uint8_t *mask, int num_pix, long size, int npatch, int nbx,
                         double *model2, double *weight, int do_model, int do_weight, int f32_sums = 0,
                         const int *gate = nullptr, const unsigned long long *nzmask = nullptr);
// model-only merge with FP32 brick sums from an FP32 view (nzmask: as launch_merge_bricks')
void launch_merge_model_f32(hipStream_t s, const int4 *units, int num_units, const int *list,
                            const unsigned long long *lmask, const double *rots, int r0, int num_rot,
                            const float *view, const double *psum_r, int all_on, const double *qvals,
                            const double *corr, const uint8_t *mask, int num_pix, long size, int npatch, int nbx,
                            double *model2, const unsigned long long *nzmask = nullptr);
// dst = src (n doubles) when *gate == 0
void launch_copy_if_zero(hipStream_t s, double *dst, const double *src, size_t n, const int *gate);
void launch_merge(hipStream_t s, const double *rots, int r0, int C, int num_rot, const double *view,
                  const double *psum_r, int all_on, const double *qvals, const double *corr,
                  const uint8_t *mask, int num_pix, long size, double *model2, double *weight,
                  int do_model, int do_weight, int modes = 1);
void launch_count_unmerged(hipStream_t s, const double *psum_r, int num_rot, int *count);
void launch_slice_gen(hipStream_t s, const double *rot, const double *model, long size,
                      const double *qvals, const double *corr, const int *perm, int num_pix,
                      double *out);
void launch_slice_merge(hipStream_t s, const double *rot, const double *view, const double *qvals,
                        const double *corr, const uint8_t *mask, const int *perm, int num_pix,
                        long size, double *model2, double *weight);
void launch_frame_stats(hipStream_t s, const double *likelihood, const double *info,
                        const uint8_t *blacklist, int num_data, int denom, double *scalars);
void launch_scale(hipStream_t s, const int *fcounts, const double *psum_d, const uint8_t *blacklist,
                  int num_data, double *scale);
void launch_fill(hipStream_t s, double *a, size_t n, double v);
void launch_update_model(hipStream_t s, double *model1, double *model2, const double *w, long size,
                         double alpha, double *part, int nparts, double *out_change, int modes = 1);
void launch_normalize_scale(hipStream_t s, double *scale, const uint8_t *blacklist, int num_data,
                            double *model1, long vol, double *mean_out);

// ---- mixed-precision path (emc_mp.hip) ----
size_t gemm_lds_bytes(int nb);
// bf16: the operands hold BF16 values (the buffers are 16-bit either way), else FP16
int launch_mp_gemm(hipStream_t s, int nb, int epi, const _Float16 *A, size_t lda, const _Float16 *B0,
                   const _Float16 *B1, size_t ldb, int K, int Mp, int Np, int M_valid, int N_valid,
                   double *out, size_t ts, const double *row_add, const double *row_div,
                   int div_valid, double scale, int bf16 = 0);
void launch_build_klik(hipStream_t s, const long *off, const int *pix, const double *cnt, int num_data,
                       size_t ldp, _Float16 *klik, int bf16 = 0);
void launch_build_ktom(hipStream_t s, const long *toff, const int *tframe, const double *tcnt,
                       int num_pix, size_t ldd, _Float16 *ktom);
void launch_center(hipStream_t s, const double *rots, int num_rot, const double *model, long size,
                   const double *qvals, const double *corr, const uint8_t *mask, int num_pix,
                   double *center_t);
void launch_center_base(hipStream_t s, const long *off, const int *pix, const double *cnt,
                        const double *center_t, int num_data, double *base);
void launch_expand_mp(hipStream_t s, const double *rots, const double *model, long size,
                      const double *qvals, const double *corr, const uint8_t *mask,
                      const double *center_t, int num_pix, int r0, int C, int pix_per_group,
                      size_t ldp, int two_terms, _Float16 *lw_hi, _Float16 *lw_lo, double *upart,
                      int *ubad, int bf16 = 0);
void launch_posterior_fp16(hipStream_t s, const double *P, int num_data, int r0, int C, size_t ldd,
                           double pscale, _Float16 *p16);

// ---- sparse orientation subsets (emc_subset.hip) ----
// Rows of full frames (every coarse cell), tile-major like the plain path:
// RF[(T * nF + f) * 64 + j], f = ffull[d] (-1 for partial frames).
struct FullRows {
    double *RF;
    const int *ffull;
    int nF;
    int num_tiles;
};
void launch_pair_likelihood(hipStream_t s, const double *logw, int num_pix, const long *off, const int *pix,
                            const double *cnt, const int *pair_frame, const int *pair_tile, const long *tile_off,
                            const int *tile_pair, long num_pairs, int t0, int t1, double *G);
// FP32 log W in 128-wide tiles (expand_kernel<float, 128>)
void launch_pair_likelihood(hipStream_t s, const float *logw, int num_pix, const long *off, const int *pix,
                            const double *cnt, const int *pair_frame, const int *pair_tile, const long *tile_off,
                            const int *tile_pair, long num_pairs, int t0, int t1, double *G);
// slot_cell (packed slot order, else null) with the frames' cell bit rows (relbits, words per
// frame) masks the orientations of a partial frame's tiles outside its cells.
void launch_pair_max(hipStream_t s, double *R, const long *pair_off, const int *pair_tile, const FullRows &full,
                     const int *slot_rot, const int *slot_cell, const uint32_t *relbits, int words,
                     const double *logwr_slot, const double *u_slot, const int *rbad_slot, const double *scale,
                     int use_scale, const double *scalars, const uint8_t *blacklist, const uint8_t *dense,
                     int num_data, double *max_exp, int *rmax);
void launch_pair_norm(hipStream_t s, double *R, const long *pair_off, const int *pair_tile, const FullRows &full,
                      const int *slot_rot, const double *beta, const uint8_t *blacklist, const double *max_exp,
                      const int *rmax, int num_data, double *p_norm);
void launch_pair_posterior(hipStream_t s, double *R, const long *pair_off, const int *pair_tile,
                           const FullRows &full, const int *slot_rot, const double *logwr_slot,
                           const double *u_slot, const double *beta, const double *sum_fact,
                           const uint8_t *blacklist, const double *max_exp, const int *rmax, const double *p_norm,
                           const double *scalars, int want_psum_d, int num_data, double *likelihood, double *info,
                           double *psum_d);
void launch_pair_psum(hipStream_t s, const double *P, const long *tile_off, const int *tile_pair,
                      const int *pair_frame, int num_tiles, const double *scale, int need_scaling,
                      const double *psum_full, double *psum_slot);
// Per-(tile, range) buckets of the significant pairs' frame ranges (launch_pair_buckets), read
// by the pair tomogram instead of walking each tile's pairs; base null: none.
struct PairBuckets {
    const long *base = nullptr;  // [num_tiles + 1] first entry of each tile
    const int *loc = nullptr;    // [num_tiles][nranges + 1] offsets of the ranges in their tile
    const int *pair = nullptr, *fr = nullptr;  // entries: pair, frame range
    const long *q0 = nullptr, *q1 = nullptr;   // frame range f: photons [q0[f], q1[f])
};
// pass 0: counts into loc (and each tile's total), then base; pass 1: the entries
// (bk_pair / bk_fr sized base[num_tiles]). Dynamic LDS of (nr + 1) ints per block.
void launch_pair_buckets(hipStream_t s, int num_tiles, const long *tile_off, const int *sig_pair,
                         const int *sig_frame, const int *sig_cnt, const long *fr_off, const int *fr_rg, int nr,
                         int *loc, long *tile_tot, long *base, int *bk_pair, int *bk_fr, int pass);
// *x = 0; *total += set bits of m[0 .. n)
void launch_zero_u64(hipStream_t s, unsigned long long *x);
void launch_popcount(hipStream_t s, const unsigned long long *m, size_t n, unsigned long long *total);
// Significant full-frame rows per tile (launch_sig_full_rows), read by the pair tomogram from
// RF in place of partial sums; RF null: none.
struct SigFullRows {
    const double *RF = nullptr;
    int nF = 0;
    const int *row = nullptr, *frame = nullptr, *cnt = nullptr;
};
// flag: num_tiles * nF bytes; sigF_row / sigF_frame: num_tiles * nF; sigF_cnt: num_tiles;
// total: the sum of sigF_cnt
void launch_sig_full_rows(hipStream_t s, const double *RF, int nF, int num_tiles, const int *fidx, uint8_t *flag,
                          int *sigF_row, int *sigF_frame, int *sigF_cnt, long *total);
// sig_pair / sig_frame / sig_cnt: the tiles' significant pairs (launch_sig_pairs). nzmask
// (nullable): per (chunk slot, 64-pixel patch) the mask of pixels with a nonzero view; with it
// only those view entries are written.
void launch_pair_tomogram(hipStream_t s, const double *P, const long *tile_off, const int *sig_pair,
                          const int *sig_frame, const int *sig_cnt, const long *foff, const int *roff, int nranges,
                          const int *fpix, const double *fcnt, const double *psum_slot, int t0, int ntiles,
                          int num_pix, const double *partial, int nparts, double *view,
                          unsigned long long *nzmask, const SigFullRows &fr = SigFullRows(),
                          const PairBuckets &bk = PairBuckets());
void launch_pair_tomogram(hipStream_t s, const float *P, const long *tile_off, const int *sig_pair,
                          const int *sig_frame, const int *sig_cnt, const long *foff, const int *roff, int nranges,
                          const int *fpix, const double *fcnt, const double *psum_slot, int t0, int ntiles,
                          int num_pix, const float *partial, int nparts, float *view, unsigned long long *nzmask,
                          const SigFullRows &fr = SigFullRows(), const PairBuckets &bk = PairBuckets());
// pairs with some P > 1e-6 in their row, per tile in the tile's order (flag: num_pairs bytes;
// sig_pair / sig_frame: tile_off layout; sig_cnt: per tile)
void launch_sig_pairs(hipStream_t s, const double *P, long num_pairs, const long *tile_off, const int *tile_pair,
                      const int *tile_frame, int num_tiles, uint8_t *flag, int *sig_pair, int *sig_frame,
                      int *sig_cnt);
void launch_sig_pairs(hipStream_t s, const float *P, long num_pairs, const long *tile_off, const int *tile_pair,
                      const int *tile_frame, int num_tiles, uint8_t *flag, int *sig_pair, int *sig_frame,
                      int *sig_cnt);
void launch_pair_to_f32(hipStream_t s, const double *P, long n, float *out);
// pairs k0 .. k0 + num_pairs / full frames f0 .. f1 (-1: all), output rows from frame d0
void launch_pair_scatter(hipStream_t s, const double *P, const int *pair_frame, const int *pair_tile,
                         const int *slot_rot, long num_pairs, int num_rot, double *out, long k0 = 0, int d0 = 0);
void launch_full_scatter(hipStream_t s, const double *RF, int nF, int num_tiles, const int *fidx,
                         const int *slot_rot, int num_rot, double *out, int f0 = 0, int f1 = -1, int d0 = 0);
void launch_slot_gather(hipStream_t s, const double *in, const int *rot_slot, int num_rot, double *out);
// Full-frame photon lists per engine pixel in nparts ranges of full-frame index (nparts <= 64):
// pass 0 writes counts[t * nparts + x]; pass 1 fills out_frame / out_cnt from off2 (the
// exclusive scan of the counts). fof[d] is frame d's full index or -1.
void launch_full_lists(hipStream_t s, int pass, const long *tom_off, const int *tom_frame, const double *tom_cnt,
                       const int *fof, int num_pix, int nF, int nparts, int *counts, const long *off2,
                       int *out_frame, double *out_cnt);
void launch_cell_posterior(hipStream_t s, double *R, int plain, int num_data, int num_rot, const long *pair_off,
                           const int *pair_tile, const FullRows &full, const int *slot_rot, const int *qmap,
                           int num_coarse, int d0, int count, float *out);
void launch_frame_gather(hipStream_t s, const double *in, const int *fidx, int nF, double *out);
void launch_frame_gather(hipStream_t s, const uint8_t *in, const int *fidx, int nF, uint8_t *out);

}  // namespace emc
