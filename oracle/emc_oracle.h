/*
 * CPU oracle for one EMC iteration -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference engine's algorithm (duaneloh/Dragonfly,
 * Cython engine, dragonfly/src/), used by tests/, by the smoke check and by the
 * CPU-baseline leg of bench.py as the checker. Nothing in the product links it.
 *
 * Each function cites the reference file:line it follows. Parity status of every
 * function is listed in the header comment of emc_oracle.c.
 */
#ifndef EMC_ORACLE_H
#define EMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem {
    /* detector (dragonfly/src/detector.h:6-16) */
    int num_pix;
    const double *qvals;      /* [num_pix][3] voxel units */
    const double *corr;       /* [num_pix] */
    const uint8_t *mask;      /* [num_pix] 0 good, 1 likelihood-excluded, 2 bad */
    /* orientations (dragonfly/src/quaternion.h:11-15), [num_rot][5] = q0..q3, weight */
    int num_rot;
    const double *quats;
    /* model (dragonfly/src/model.h:15-20): num_modes volumes of size^3 */
    long size;
    const double *model1;
    double *model2;           /* out: merged numerator */
    double *inter_weight;     /* out: merged weights */
    /* frames (dragonfly/src/emcfile.h:12-32): ftype 0 sparse, 1 dense int */
    int num_data;
    int ftype;
    const int *ones, *multi;
    const long *ones_accum, *multi_accum;
    const int *place_ones, *place_multi, *count_multi;
    const int *int_frames;    /* [num_data][num_pix] when ftype == 1 */
    /* per-frame state (dragonfly/src/iterate.h:11-38) */
    const double *beta;
    double *scale;            /* in/out when need_scaling && update_scale */
    const uint8_t *blacklist;
    const int *fcounts;
    const double *sum_fact;
    double mean_count;
    int num_blacklist;
    /* refinement (legacy src/iterate.c:103-130, src/max_emc.c:264-274): when quat_mapping is
     * set, frame d is evaluated only at orientations r whose coarse cell quat_mapping[r] is
     * one of rel_quat[rel_off[d] .. rel_off[d + 1]) */
    const int *quat_mapping;  /* [num_rot] or NULL */
    const long *rel_off;      /* [num_data + 1] */
    const int *rel_quat;
    /* flags (dragonfly/src/params.h) */
    int need_scaling, update_scale, iteration, known_scale;
    int num_threads;          /* OpenMP threads; 1 gives the canonical single-thread keep rule */
    /* params.num_modes (0 is taken as 1). Orientations are then the combined
     * r = rotation * num_modes + mode, r < num_rot * num_modes (max_prob.c:80-84); u,
     * prob_dense and rmax count combined orientations. */
    int num_modes;
    /* background (dragonfly/src/detector.h:11-12, iterate.h:21): with_bg switches on the
     * background-aware likelihood, tomogram and scale updates when need_scaling is set
     * (max_prob.c:94,134-156, max_update.c:11, max_scale.c:9). background and powder are
     * [num_pix] (powder from orc_calc_powder), bgscale [num_data]. */
    int with_bg;
    const double *background;
    const double *powder;
    const double *bgscale;
} orc_problem;

typedef struct orc_result {
    double *u;                /* [num_rot] */
    double *max_exp;          /* [num_data] */
    double *p_norm;           /* [num_data] */
    int *rmax;                /* [num_data] */
    double *likelihood;       /* [num_data] */
    double *info;             /* [num_data] */
    double *psum_d;           /* [num_data], only when need_scaling && update_scale */
    double *prob_dense;       /* optional [num_data][num_rot] posterior dump (0 where not kept) */
    double rescale;
    double mutual_info;
    double avg_likelihood;
    double seconds[7];        /* wall: rescale, prob, norm, update, total; then thread-seconds
                                 in the orientation loops spent on per-orientation work (slice
                                 generation, log, merge) and in those loops altogether */
    double *quat_norm;        /* optional [num_data][num_modes] occupancies (max_update.c:122-124) */
} orc_result;

int orc_num_rot(int num_div);
/* voronoi_subset (dragonfly/src/quaternion.c:844-862): nearest coarse orientation of every
 * fine one by 1 - (q_f . q_c)^2, first minimum kept */
void orc_voronoi_subset(int num_coarse, const double *qcoarse, int num_fine, const double *qfine,
                        int *nearest);
int orc_quat_gen(int num_div, double *quats);
void orc_make_rot_quat(const double *q, double *rot9);
void orc_slice_gen3d(const double *quat, const double *model, long size,
                     int num_pix, const double *qvals, const double *corr, double *slice);
void orc_slice_merge3d(const double *quat, const double *slice, double *model2,
                       double *weight, long size, int num_pix, const double *qvals,
                       const double *corr, const uint8_t *mask);
void orc_symmetrize_friedel(double *array, const double *weights, int size);
/* calc_powder (dragonfly/src/iterate.c:119-179), sparse frames: mean photon count per mask < 1
 * pixel over all frames */
void orc_calc_powder(int num_pix, const uint8_t *mask, int num_data, const int *ones, const int *multi,
                     const long *ones_accum, const long *multi_accum, const int *place_ones,
                     const int *place_multi, const int *count_multi, double *powder);
double orc_maximize(const orc_problem *p, orc_result *res);
/* update_tomogram_bg's view for one orientation (max_update.c:343-543), exposed for tests */
void orc_tomogram_bg_view(const orc_problem *p, const double *pr, double psum_r, double scalemax, double *view);
double orc_update_model(double *model1, double *model2, const double *inter_weight,
                        long size, double alpha);
/* update_model over num_modes volumes: Friedel symmetry per mode, then divide, rms change and
 * blend over all of them, rms = sqrt(change / num_modes / size^3) (recon.pyx:112-149) */
double orc_update_model_modes(double *model1, double *model2, const double *inter_weight, long size,
                              double alpha, int num_modes);

#ifdef __cplusplus
}
#endif

#endif
