/*
 * emc_hip.h -- C ABI of the MI355X EMC engine (libemc_hip.so).
 *
 * One engine handle holds, in HBM, everything one EMC iteration reads and writes:
 * detector geometry, orientation samples, the 3D intensity model, the sparse photon
 * frames and the per-frame state. emc_maximize() runs one full E+M step on the GPU
 * (expand -> likelihood -> normalise -> posterior -> tomogram -> merge -> scale) and
 * replaces the reference's C entry point
 *     double maximize(struct max_data*)            dragonfly/src/maximize.h:45, maximize.c:8-56
 * The setters replace the fields the reference's Cython objects fill before calling it
 * (struct iterate, dragonfly/src/iterate.h:11-38, filled by dragonfly/iterate.pyx:40-182);
 * the getters replace the numpy-view properties of EMCRecon (dragonfly/recon.pyx:294-415).
 *
 * Conventions: plain pointers and sizes; all input pointers are HOST pointers that are
 * copied into engine-owned device memory; getters copy device results into caller host
 * buffers. Every call returns 0 on success or a negative EMC_E* code; emc_last_error()
 * gives the message of the last failure on the calling thread. A handle is used by one
 * caller at a time (not re-entrant); calls on it are ordered on the engine's stream and
 * every getter synchronises that stream.
 */
#ifndef EMC_HIP_H
#define EMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMC_OK 0
#define EMC_E_ARG (-1)      /* bad argument (null pointer, size mismatch, out of range) */
#define EMC_E_STATE (-2)    /* a required setter has not been called */
#define EMC_E_HIP (-3)      /* HIP runtime error (allocation, launch, copy) */
#define EMC_E_NOMEM (-4)    /* device memory budget exceeded */

#define EMC_MAX_MODES 64    /* largest num_modes emc_set_modes accepts */

typedef struct emc_engine emc_engine;

/* Options of one iteration; the fields of struct params that maximize() reads
 * (dragonfly/src/params.h, set from config by dragonfly/params.pyx:67-128). */
typedef struct emc_iter_opts {
    int iteration;      /* params.iteration, 1-based (max_prob.c:124) */
    int need_scaling;   /* params.need_scaling (max_prob.c:124, max_update.c:108) */
    int update_scale;   /* params.update_scale (maximize.c:151) */
    int known_scale;    /* params.known_scale (max_prob.c:124) */
} emc_iter_opts;

#define EMC_NUM_STAGES 9

/* Results of one iteration that the reference returns or stores in struct iterate. */
typedef struct emc_iter_stats {
    double avg_likelihood; /* return value of maximize() (max_prob.c:370-374) */
    double mutual_info;    /* iter->mutual_info (max_prob.c:369-373) */
    double rescale;        /* iter->rescale[0] (max_prob.c:59) */
    double stage_ms[EMC_NUM_STAGES]; /* GPU ms per stage (HIP events on the engine stream):
                              expand, likelihood, rescale, max, norm, posterior, tomogram
                              (gather), tomogram reduce, merge */
    int num_chunks;        /* orientation chunks the iteration was split into (launches per stage) */
    double view_nonzero;   /* nonzero view entries (orientation x pixel) the merge read, on the
                              refinement-subset path when the views are stored sparsely; else 0 */
} emc_iter_stats;

/* Per-frame / per-orientation / volume fields readable with emc_get_field(). */
enum emc_field {
    EMC_FIELD_U = 0,          /* f64[num_rot]   u[0][r]      (max_prob.c:38)  recon.pyx:371 */
    EMC_FIELD_MAX_EXP,        /* f64[num_data]  max_exp      (max_prob.c:234) recon.pyx:363 */
    EMC_FIELD_P_NORM,         /* f64[num_data]  p_norm       (max_prob.c:245) recon.pyx:367 */
    EMC_FIELD_RMAX,           /* i32[num_data]  rmax         (max_prob.c:237) recon.pyx:397 */
    EMC_FIELD_LIKELIHOOD,     /* f64[num_data]  likelihood   (max_update.c:105) recon.pyx:393 */
    EMC_FIELD_INFO,           /* f64[num_data]  info         (max_update.c:128) recon.pyx:389 */
    EMC_FIELD_PSUM_D,         /* f64[num_data]  psum_d       (max_update.c:113) recon.pyx:358 */
    EMC_FIELD_SCALE,          /* f64[num_data]  iter->scale  (max_scale.c:15) */
    EMC_FIELD_MODEL1,         /* f64[num_modes][size^3] mod->model1 */
    EMC_FIELD_MODEL2,         /* f64[num_modes][size^3] mod->model2 (max_prob.c:281) */
    EMC_FIELD_INTER_WEIGHT,   /* f64[num_modes][size^3] mod->inter_weight (max_prob.c:282) */
    EMC_FIELD_PSUM_R,         /* f64[num_rot]   psum_r per orientation (max_update.c:109,116) */
    EMC_FIELD_PROB,           /* f64[num_data][num_rot] posterior P_dr, 0 where not kept
                                 (dense form of max_data.prob/place_prob, recon.pyx:409-415) */
    EMC_FIELD_QUAT_NORM,      /* f64[num_data][num_modes] occupancy of each frame in each mode,
                                 num_modes > 1 only (max_update.c:122-124, recon.pyx:259-260) */
    EMC_FIELD_COUNT
};

/* ---- lifetime ------------------------------------------------------------------------ */
int emc_create(int device, emc_engine **out);
void emc_destroy(emc_engine *e);
const char *emc_last_error(void);
int emc_synchronize(emc_engine *e);

/* ---- orientation sampling (host; no GPU needed) --------------------------------------
 * quat_gen(num_div, quat)   dragonfly/src/quaternion.c:811-842, bit-exact.
 * quats_out receives emc_num_rot(num_div) rows of (q0, q1, q2, q3, weight).            */
int emc_num_rot(int num_div);
int emc_quat_gen(int num_div, double *quats_out);
/* voronoi_subset(qcoarse, qfine, nearest)  dragonfly/src/quaternion.c:844-862: the coarse
 * orientation nearest to every fine one (used for refinement).                          */
int emc_voronoi_subset(int num_coarse, const double *qcoarse, int num_fine,
                       const double *qfine, int *nearest);

/* ---- inputs ----------------------------------------------------------------------------
 * detector: struct detector qvals/corr/raw_mask  (dragonfly/src/detector.h:6-16,
 *           filled by dragonfly/detector.pyx:31-74 / setters :161-206)               */
int emc_set_detector(emc_engine *e, int num_pix, const double *qvals, const double *corr,
                     const uint8_t *raw_mask);
/* modes: params.num_modes (dragonfly/params.pyx:104) and the model's mode count
 *        (Model(size, num_modes), iterate.pyx:164). The engine's orientations become the
 *        combined ones, r = rotation * num_modes + mode (max_prob.c:80-84): num_rot,
 *        rmax, u, psum_r and the posterior's columns count num_rot * num_modes. Default 1.
 *        Setting a new count re-expands the quaternions, switches refinement off and drops
 *        the model (set it again with num_modes volumes). Non-rotating modes
 *        (num_nonrot_modes) are not taken: the Cython engine's model holds num_modes
 *        volumes only. The mixed-precision and refinement paths take one mode.          */
int emc_set_modes(emc_engine *e, int num_modes);
/* detector background: det->background[num_pix] (caller's pixel order, detector.h:11-12,
 *        CDetector.parse_background detector.pyx:101-104) and iter->bgscale[num_data]
 *        (iterate.h:21, parse_scale(bg=True) iterate.pyx:284-315); the powder sum
 *        (calc_powder, iterate.c:119-179) is computed from the frames. Call after the
 *        detector and frames (setting either switches the background off again); NULL
 *        switches it off. With need_scaling the iteration then runs the background-aware
 *        likelihood, tomogram and scale updates (max_prob.c:134-156, max_update.c:232-543,
 *        max_scale.c:21-288), in FP64 for sparse frames with one mode and no refinement. */
int emc_set_background(emc_engine *e, const double *background, const double *bgscale);
/* orientations: struct quaternion quats[num_rot][5] (dragonfly/src/quaternion.h:11-15) */
int emc_set_quats(emc_engine *e, int num_rot, const double *quats);
/* model: struct model model1[num_modes][size^3] (dragonfly/src/model.h:15-20,
 *        Model.allocate dragonfly/model.pyx:43-103)                                   */
int emc_set_model(emc_engine *e, long size, const double *model1);
/* emc_set_model with src's model1 (same device and mode count), copied on the device: a
 * second engine over the same frames (the coarse search that judges orientation proposals)
 * takes the current model without a round trip through the host. No counterpart in the
 * reference.                                                                            */
int emc_copy_model(emc_engine *dst, const emc_engine *src);
/* sparse photon frames: struct dataset ones, multi, place_ones, place_multi, count_multi
 *        (dragonfly/src/emcfile.h:12-32, parse_binarydataset emcfile.c:3-57)          */
int emc_set_frames_sparse(emc_engine *e, int num_data, const int *ones, const int *multi,
                          const long *ones_accum, const long *multi_accum,
                          const int *place_ones, const int *place_multi,
                          const int *count_multi);
/* dense integer frames (ftype DENSE_INT, emcfile.c:39-43): frames[num_data][num_pix].
 * The likelihood then sums K log W over every mask < 1 pixel (max_prob.c:173-177), so an
 * orientation with log W = -inf on any such pixel is dropped for these frames; the tomogram
 * adds every pixel (max_update.c:211-214).                                             */
int emc_set_frames_dense(emc_engine *e, int num_data, int num_pix, const int *frames);
/* refinement: orientation subsets per frame (legacy engine: quat_mapping from
 * voronoi_subset, rel_quat/num_rel_quat from parse_rel_quat, src/iterate.c:103-130,
 * :364-525; applied in src/max_emc.c:264-274). quat_mapping[r] is the coarse cell of
 * orientation r (num_rot entries, values < num_coarse); frame d is evaluated only at the
 * orientations whose cell is listed in rel_quat[rel_off[d] .. rel_off[d + 1]). A null
 * quat_mapping switches refinement off; setting quaternions or frames also does. A frame
 * whose list is the single entry -1 takes every coarse cell (the full search) without
 * listing them.                                                                          */
int emc_set_refine(emc_engine *e, int num_coarse, const int *quat_mapping, const long *rel_off,
                   const int *rel_quat);
/* arithmetic of the two photon contractions. No counterpart in the reference (which
 * computes everything in FP64): 0 = plain FP64 (default, reproduces the reference);
 * 1 = dense frames only: likelihood and tomogram as FP16 products on matrix cores with
 * FP32 accumulation, log W centred per pixel and rounded once to FP16; 2 = as 1 with
 * log W carried as an FP16 hi + lo pair; 3 = any frames: log W and the posterior read by
 * the photon gathers in FP32 (half the bytes), sums in FP64; 4, 5 = as 1, 2 with the
 * likelihood product's log W and counts in BF16 (the tomogram product stays FP16). Every
 * other step stays FP64. Modes 1, 2, 4 and 5 need frames set with emc_set_frames_dense and
 * counts <= 2048 (FP16) or <= 256 (BF16).                                                */
int emc_set_precision(emc_engine *e, int mode);
/* kernel-variant options, no counterpart in the reference (results are the same to the
 * parity tolerance; only speed changes): lik_tiles (1-4 orientation tiles per likelihood
 * wave), lik_unroll (4/8/16 rows in flight), lik_xcd (XCD-aware tile order of the
 * likelihood launch: 1, 2, 4 or 8 XCDs share a log W tile; 0 off), lik_split (2-8: photon lists split into pixel ranges, one per XCD
 * group, partial sums added in range order), tomo_tiles, tomo_tile_mode (0 per-pixel gather,
 * 1 LDS tiles, 2 frame-partitioned, default), tomo_parts (frame ranges, 1-64), tomo_width32,
 * tomo_unroll (8/16), tomo_seq, max_chunk_cols (0 = by memory), merge_mode (0 box, 1 brick),
 * subset (1 sparse refinement subsets, 0 masked), slot_pack (1: refinement cells packed
 * back to back along a nearest-neighbour chain, 0: each cell padded to whole 64-orientation
 * tiles; takes effect at the next emc_set_refine), sparse_rows (frames per batch of
 * emc_get_sparse_prob after a subset iteration, 0: 256 MB of rows), full_part_frames
 * (frames per range of the subset path's full-frame tomogram, default 12500), estep_only (1:
 * emc_maximize stops after the posterior and its statistics -- no tomogram, merge or scale
 * update), mstep_only (1: on the subset path, the M-step of the last estep_only iteration
 * when nothing changed since), graph (1: the subset iteration captured and replayed as a HIP
 * graph), fused (1: expand and pair likelihood in one persistent launch), sig_full_rows (mean
 * significant full-frame rows per tile up to which the subset path's pair tomogram reads them
 * directly instead of partial views; default 1024), unit_items (brick merge: list entries per
 * work unit, >= 64), tomo_buckets (1, default: the subset path's pair tomogram reads per-(tile,
 * pixel range) buckets of its significant pairs; 0: each block walks its tile's pairs). Start
 * values come from EMC_<NAME>.                                                               */
int emc_set_option(emc_engine *e, const char *name, int value);
int emc_get_option(emc_engine *e, const char *name, int *value);
/* per-frame state of struct iterate: scale, beta, blacklist, fcounts, sum_fact,
 *        mean_count[0], num_blacklist (iterate.h:17-24,33; iterate.c:3-117)          */
int emc_set_frame_params(emc_engine *e, const double *scale, const double *beta,
                         const uint8_t *blacklist, const int *fcounts,
                         const double *sum_fact, double mean_count, int num_blacklist);
/* per-iteration beta written by EMCRecon.update_beta (recon.pyx:151-174)             */
int emc_set_beta(emc_engine *e, const double *beta);
/* scale factors (Iterate.parse_scale, iterate.pyx:284-315)                             */
int emc_set_scale(emc_engine *e, const double *scale);

/* ---- the iteration --------------------------------------------------------------------
 * maximize(struct max_data*)  dragonfly/src/maximize.c:8-56. Writes model2,
 * inter_weight, the per-frame fields, and (need_scaling && update_scale) scale.         */
int emc_maximize(emc_engine *e, const emc_iter_opts *opts, emc_iter_stats *stats);

/* EMCRecon.update_model, 3D point group 1 (dragonfly/recon.pyx:112-149 with
 * symmetrize_friedel dragonfly/src/model.c:470-490): Friedel-symmetrise model2 with
 * inter_weight, divide, rms change, model1 = alpha*model1 + (1-alpha)*model2.          */
int emc_update_model(emc_engine *e, double alpha, double *rms_change);
/* Iterate.normalize_scale (dragonfly/iterate.pyx:361-384): mean of scale over
 * non-blacklisted frames; model1 *= mean; scale /= mean. Returns the mean.             */
int emc_normalize_scale(emc_engine *e, double *mean_scale);

/* ---- geometry on the device ----------------------------------------------------------
 * Model.slice_gen (dragonfly/model.pyx:105-131 -> slice_gen3d model.c:11-59): view of
 * model1 at one quaternion.                                                             */
int emc_slice_gen(emc_engine *e, const double *quat, double *view_out);
/* emc_slice_gen without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * or per-orientation loops, harnesses that forbid hipMalloc) once the engine's detector and
 * model are set. The rotation and the device copy of the view are staged in a fixed
 * library array of EMC_STATIC_MAX_PIX + 16 doubles per GPU (8 MiB, held until the process
 * ends) instead of engine buffers; num_pix above EMC_STATIC_MAX_PIX returns EMC_E_ARG.
 * Same arguments, result and errors as emc_slice_gen otherwise. Every call on one GPU
 * shares that array: calls that overlap on one GPU (several threads, several engines on
 * the device) race and must be serialised by the caller.                                */
#define EMC_STATIC_MAX_PIX 1048576
int emc_slice_gen_static(emc_engine *e, const double *quat, double *view_out);
/* Model.slice_merge (dragonfly/model.pyx:133-150 -> slice_merge3d model.c:171-244):
 * merge one view into model2 / inter_weight.                                            */
int emc_slice_merge(emc_engine *e, const double *quat, const double *view);

/* ---- outputs ------------------------------------------------------------------------- */
/* Sparse posterior of the last iteration: for every frame the orientations with P > 1e-6 and
 * their P, orientations ascending (common->prob / place_prob after resparsify with PROB_MIN,
 * max_prob.c:295-326; saved as probabilities/place and probabilities/prob, recon.pyx:269-279,
 * the single-thread order). num_prob[num_data] receives the counts and *total their sum;
 * with place and prob non-NULL (capacity >= total entries) the entries follow, frame after
 * frame. Counted and compacted on the GPU (after an iteration on refinement subsets, frame
 * batches are first scattered from the subset rows into orientation order). */
int emc_get_sparse_prob(emc_engine *e, int *num_prob, long capacity, int *place, double *prob,
                        long *total);
/* Copy field `field` (enum emc_field) into dst; count is the element count of dst and
 * must equal the field's size.                                                          */
int emc_get_field(emc_engine *e, int field, void *dst, size_t count);
/* Posterior mass of every frame in every coarse cell after an iteration:
 * out[d][c] = sum over orientations r with quat_mapping[r] == c of P_dr (f32,
 * num_data x num_coarse, host memory; num_coarse <= 6144). No counterpart in the reference:
 * it gives an orientation-proposal model its training targets (the cell posteriors an
 * iteration produced) without moving the dense num_data x num_rot posterior.            */
int emc_cell_posterior(emc_engine *e, int num_coarse, const int *quat_mapping, float *out);
/* The k most likely coarse cells of every frame after an iteration: idx[d][0..k) the cells
 * in descending order of posterior mass (ties to the lower cell), mass[d] the mass they hold
 * (host memory; k <= 64, num_coarse <= 6144). No counterpart in the reference: the
 * per-frame proposals of an earlier iteration's posterior, without moving the cell
 * posteriors (emc_cell_posterior) to the host.                                          */
int emc_cell_topk(emc_engine *e, int num_coarse, const int *quat_mapping, int k, int *idx, float *mass);
/* Posterior mass of listed orientations, per frame, after a plain (not refinement-subset)
 * iteration: mass[d] = sum of P_dr over r in rel_quat[rel_off[d] .. rel_off[d + 1]) (rel_off:
 * num_data + 1 ascending offsets; orientations < num_rot, each listed once). No counterpart
 * in the reference: a full search at a coarse sampling, run with the option estep_only,
 * judges orientation proposals by the likelihood -- the mass its posterior puts inside each
 * frame's proposed cells.                                                                */
int emc_subset_mass(emc_engine *e, const long *rel_off, const int *rel_quat, double *mass);
/* Likelihood judge of orientation proposals, after an iteration of e on refinement subsets
 * run with the option estep_only (the frames' cells are the proposals) and a plain estep_only
 * iteration of `coarse` on the same frames, scale factors and beta over one orientation per
 * coarse cell (quat_gen of the coarse sampling, the cells of e's quat_mapping). Per frame,
 * outside[d] (host, num_data) estimates the posterior mass the full fine search would put
 * outside the frame's cells: each outside cell counts as its fine orientations at the centre's
 * log-likelihood (the centre's weight replaced by the cell's mean fine weight), against the
 * fine search's maximum and normaliser inside the cells. Full and blacklisted frames get 0.
 * With prob_floor > 0 (< 1) only cells whose estimated posterior per fine orientation
 * exceeds prob_floor count (mass spread thinner is dropped by the tomogram's P > 1e-6 cut).
 * No counterpart in the reference: proposals are accepted or sent to the full search by this
 * estimate (refinement in the reference, src/max_emc.c:264-274, takes its cells as given).  */
int emc_judge_outside(emc_engine *e, emc_engine *coarse, double prob_floor, double *outside);
/* The same judge (prob_floor 0), and for every frame it puts above outside_max (in (0, 1)) the
 * fewest coarse cells outside the frame's set whose addition brings the estimate, still taken
 * with the fine maximum and normaliser of the set alone (an upper bound, since both only grow
 * with more cells), to at most target (in (0, outside_max]): the cells of largest share of
 * the outside mass.
 * outside[d] as emc_judge_outside; num_add[d] (host, num_data) = 0 (at most outside_max, or a
 * full or blacklisted frame), k > 0 (k cells to add), -1 (more than max_add cells: the full
 * search); *total = the cells to add over all frames, kept in the engine for
 * emc_judge_widen_cells. At most 4096 coarse cells. No counterpart in the reference.        */
int emc_judge_widen(emc_engine *e, emc_engine *coarse, double outside_max, double target, int max_add,
                    double *outside, int *num_add, long *total);
/* The cells of the last emc_judge_widen, frame after frame, ascending within a frame
 * (capacity >= its *total).                                                               */
int emc_judge_widen_cells(emc_engine *e, int *cells, long capacity);
/* Zero model2 and inter_weight (max_data allocate_memory, maximize.c:93-94).            */
int emc_clear_merge(emc_engine *e);

/* ---- orientation proposal network --------------------------------------------------------
 * No counterpart in the reference: kernels of the learned orientation proposal (a BF16
 * multilayer perceptron trained on (pattern, posterior) pairs whose top-k coarse cells feed
 * emc_set_refine). All pointers are device pointers; work is ordered on `stream` (a
 * hipStream_t). Products need M, N multiples of 256, K of 32, leading dimensions (in
 * elements) multiples of 8 and >= K; BF16 operands, FP32 accumulation. Return 0, or -1 on a
 * bad shape or pointer (nothing launched).                                                */
/* out[m][n] = act(X[m] . W[n] + bias[n]), m < M_valid, n < N_valid; relu: max(0, .);
 * out_f32: FP32 output, else BF16                                                          */
int emc_nn_forward(void *stream, const void *X, size_t ldx, const void *W, size_t ldw, int M, int N,
                   int K, int M_valid, int N_valid, const float *bias, int relu, int out_f32, void *out,
                   size_t ldo);
/* dX[m][n] = (dY[m] . Wt[n]) * (H[m][n] > 0) (H may be null), BF16                         */
int emc_nn_backward_data(void *stream, const void *dY, size_t ldy, const void *Wt, size_t ldwt, int M,
                         int N, int K, int M_valid, int N_valid, const void *H, size_t ldh, void *dX,
                         size_t ldx);
/* dW[m][n] = dYt[m] . Xt[n], FP32                                                           */
int emc_nn_backward_weight(void *stream, const void *dYt, size_t lda, const void *Xt, size_t ldb, int M,
                           int N, int K, int M_valid, int N_valid, float *dW, size_t ldo);
/* out[c][r] = in[r][c], BF16                                                                */
int emc_nn_transpose(void *stream, const void *in, size_t ldi, int rows, int cols, void *out, size_t ldo);
/* per row: p = softmax(Z[0..n)), loss = -sum Y log p, dZ = (p - Y) * grad_scale (BF16)      */
int emc_nn_softmax_xent(void *stream, const float *Z, size_t ldz, const float *Y, size_t ldy, int rows,
                        int n, float grad_scale, float *loss, void *dZ, size_t ldd);
/* db[c] = sum over rows of dY[r][c] (BF16 in, FP32 out)                                     */
int emc_nn_colsum(void *stream, const void *dY, size_t ldy, int rows, int cols, float *db);
/* Adam step on FP32 master weights; refreshes the BF16 copy wb and transposed copy wt       */
int emc_nn_adam(void *stream, float *w, const float *g, float *m, float *v, int rows, int cols, size_t ld,
                float lr, float b1, float b2, float eps, float wd, int step, void *wb, void *wt,
                size_t ldt);
/* polar photon histogram per frame: photon lists (off, pix, cnt), bin_of[pixel] (-1 none),
 * written as log(1 + 1000 * fraction) to out[frame][0..nbins), BF16                        */
int emc_nn_polar_hist(void *stream, const long *off, const int *pix, const double *cnt,
                      const int *bin_of, int nbins, int num_frames, void *out, size_t ldo);
/* indices of the k (<= 64) largest of Z[row][0..n), descending, ties to the lower index;
 * mass (may be null) receives the softmax probability those k entries hold                */
int emc_nn_topk(void *stream, const float *Z, size_t ldz, int rows, int n, int k, int *idx, float *mass);

#ifdef __cplusplus
}
#endif

#endif /* EMC_HIP_H */
