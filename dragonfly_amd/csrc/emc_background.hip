// Background-aware iteration (a detector background B_t with per-frame scales phi_d and
// background scales b_d, when the scale factors are on): the likelihood of a photon is
// log(phi_d W_rt + b_d B_t) (max_prob.c:134-156), the tomogram is the root in W of the
// derivative of the log-likelihood per (pixel, orientation) (update_tomogram_bg /
// gradient_rt, max_update.c:232-543), and the scale of a frame the root in phi of its
// derivative (update_scale_bg / gradient_d, max_scale.c:21-288). Roots are bracketed and then
// found by Ridder's method with the reference's state machine and tolerances.
//
// Tomogram: one wavefront per (pixel, 64 orientations), one lane per orientation; every
// gradient evaluation walks the pixel's photon list (frames ascending, the order of
// gradient_rt) and gathers the 64 posterior values of each frame from its 512-byte row.
// The (pixel, orientation) solves are independent except in one place: a pixel that is
// bracketed in round i of the bracket search keeps its gradient only if that round was the
// orientation's last; a later round, run for the sake of other pixels, re-evaluates it to 0.
// The first pass finds each orientation's last round (the round in which every pixel was
// settled), the second pass solves with it.
#include <cfloat>

#include "emc_common.h"
#include "emc_kernels.h"

namespace emc {

// G[tile][d][lane] = sum over the frame's photons (mask < 1, ones then multi) of
// count * log(max(phi_d W + b_d B_t, DBL_MIN if negative)); W is the raw (not logged) chunk.
__global__ __launch_bounds__(kBlock) void likelihood_bg_kernel(
    const double *__restrict__ W, size_t tsW, const long *__restrict__ off, const int *__restrict__ pix,
    const double *__restrict__ cnt, const uint8_t *__restrict__ blacklist, int num_data,
    const double *__restrict__ scale, const double *__restrict__ bgscale, const double *__restrict__ bg,
    double *__restrict__ G, size_t tsG, int r0) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int d = blockIdx.x * kWavesPerBlock + wave;
    if (d >= num_data || blacklist[d]) return;
    const int tile = blockIdx.y;
    const double *__restrict__ col = W + (size_t)tile * tsW + lane;
    const double sc = scale[d], bs = bgscale[d];
    double acc = 0.;
    for (long k = off[d]; k < off[d + 1]; ++k) {
        const int t = pix[k];
        double ip = col[(size_t)t * kRTile] * sc + bs * bg[t];
        if (ip < 0) ip = DBL_MIN;
        acc += cnt[k] * log(ip);
    }
    G[(size_t)(r0 / kRTile + tile) * tsG + (size_t)d * kRTile + lane] = acc;
}

// scalemax of calc_psum_r (max_update.c:50,120-121): per orientation, the largest scale of the
// frames with P >= 1e-6, -DBL_MAX if none. Segments of frames give partial maxima.
__global__ __launch_bounds__(64) void scalemax_part_kernel(const double *__restrict__ P, int num_data,
                                                           const double *__restrict__ scale, int seg_len,
                                                           double *__restrict__ part, int Rp) {
    const int tile = blockIdx.x, seg = blockIdx.y, lane = threadIdx.x;
    const double *row = P + (size_t)tile * num_data * kRTile + lane;
    const int d0 = seg * seg_len, d1 = min(num_data, d0 + seg_len);
    double m = -DBL_MAX;
    for (int d = d0; d < d1; ++d) {
        const double v = row[(size_t)d * kRTile];
        if (!(v < kProbMin) && scale[d] > m) m = scale[d];
    }
    part[(size_t)seg * Rp + tile * kRTile + lane] = m;
}

__global__ void scalemax_reduce_kernel(const double *__restrict__ part, int nseg, int Rp,
                                       double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Rp) return;
    double m = -DBL_MAX;
    for (int s = 0; s < nseg; ++s) m = fmax(m, part[(size_t)s * Rp + r]);
    out[r] = m;
}

namespace {

// Pixel states of update_tomogram_bg.
constexpr int kOpt = 0, kNoPhoton = 1, kNegligible = 128, kLowBound = 160, kBracketed = 192, kDone = 255;

struct PixelPhotons {
    const int *frame;
    const double *cnt;
    long k0, k1;
};

// gradient_rt for one (pixel, orientation): the state update of its first loop and the sum
// over the frames with P > 1e-6 (blacklisted frames have P = 0).
__device__ __forceinline__ double gradient_rt(const PixelPhotons &ph, const double *__restrict__ prow,
                                              const double *__restrict__ scale,
                                              const double *__restrict__ bgscale, double bgt, double psum,
                                              double w, int &mask) {
    double g;
    if (mask < kNegligible) {
        g = -psum;
        mask = kNoPhoton;
    } else if (mask == kLowBound) {
        g = -1e100;
    } else {
        g = 0.;
    }
    for (long k = ph.k0; k < ph.k1; ++k) {
        const int d = ph.frame[k];
        const double pv = prow[(size_t)d * kRTile];
        if (!(pv > kProbMin)) continue;
        const double c = ph.cnt[k], sc = scale[d], bs = bgscale[d];
        if (mask < kNegligible) {
            g += pv * c * sc / (w * sc + bs * bgt);
            mask = kOpt;
        } else if (mask == kNegligible) {
            g += pv * c;
        } else if (mask == kLowBound) {
            g = fmax(g, sc / bs);
        }
    }
    return g;
}

}  // namespace

// kFirst: run the bracket search until the pixel is settled and record, per orientation, the
// last round the reference runs (atomicMax of each pixel's first settled round into last_round).
// Otherwise: solve with the recorded rounds and write the view, orientation-major
// (view[c * num_pix + t]) when rmajor, else in 64-wide tiles.
template <bool kFirst>
__global__ __launch_bounds__(kBlock) void tomogram_bg_kernel(
    const double *__restrict__ P, int num_data, int r0, int C, int num_rot, const long *__restrict__ tom_off,
    const int *__restrict__ tom_frame, const double *__restrict__ tom_cnt, const double *__restrict__ scale,
    const double *__restrict__ bgscale, const double *__restrict__ bg, const double *__restrict__ powder,
    const uint8_t *__restrict__ raw_mask, int num_pix, double mean_count, const double *__restrict__ psum_r,
    const double *__restrict__ scalemax, int *__restrict__ last_round, double *__restrict__ view_out,
    int rmajor) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t = blockIdx.x * kWavesPerBlock + wave;
    if (t >= num_pix) return;
    const int tile = blockIdx.y, c = tile * kRTile + lane, r = r0 + c;
    if (c >= C) return;
    const bool on = r < num_rot;
    const double grad_tol = 1.e-4;
    const PixelPhotons ph{tom_frame, tom_cnt, tom_off[t], tom_off[t + 1]};
    const double *__restrict__ prow = P + (size_t)(r / kRTile) * num_data * kRTile + (r % kRTile);
    const double bgt = bg[t], pw = powder[t];
    const double psum = on ? psum_r[r] : 0., smax = on ? scalemax[r] : -DBL_MAX;
    double view = 0., w_old = 0., w_new = 0., w_mid = 0., w_lat = 0.;
    double g_old = 0., g_new = 0., g_mid = 0., g_lat = 0.;
    int mask;
    if (raw_mask[t] < 2) {
        if (bgt / smax < 1.e-2 * pw) {
            mask = kNegligible;
        } else {
            w_new = 1.e-8 - bgt / smax;
            mask = kOpt;
        }
    } else {
        mask = kDone;
    }
    if (!on) mask = kDone;
    if (on) g_old = gradient_rt(ph, prow, scale, bgscale, bgt, psum, w_old, mask);
    if (on) {
        if (fabs(g_old) < grad_tol) {
            view = 0.;
            mask = kDone;
        } else if (g_old > 0.) {
            const double val = mean_count / num_pix;
            w_new = pw > val ? pw : val;
            w_new *= 2.;
        }
    }
    // bracket search: rounds 0 .. last (the reference stops after the round in which every
    // pixel of the orientation was settled, or after 10)
    const int last = kFirst ? 9 : (on ? min(last_round[r], 9) : -1);
    // first round in which this pixel counts as settled; states other than 0 (and 160, not
    // reachable yet) count from round 0. A settled pixel's later rounds change nothing it
    // uses again except a bracketed pixel's gradient, which they reset to 0 (below), so
    // they are skipped.
    int settled = (mask == kOpt) ? 10 : 0;
    for (int i = 0; i < 10; ++i) {
        const bool act = on && i <= last && settled == 10;
        if (!__any(act)) break;
        if (!act) continue;
        g_new = gradient_rt(ph, prow, scale, bgscale, bgt, psum, w_new, mask);
        bool counted = false;
        if (mask == kOpt) {
            if (g_old < 0. && g_new < 0.) {
                mask = kLowBound;
            } else if (g_old > 0. && g_new > 0.) {
                w_new *= 2;
            } else {
                mask = kBracketed;
                counted = true;
            }
        } else if (mask == kLowBound) {
            w_new = 1.e-8 - bgt / g_new;
            mask = kOpt;
        } else {
            counted = true;
        }
        if (counted && settled == 10) settled = i;
    }
    if (kFirst) {
        if (on) atomicMax(&last_round[r], settled);
        return;
    }
    // gradient_rt gives 0 for a bracketed pixel: a round of its orientation after the one that
    // bracketed it left that in g_new
    if (on && mask == kBracketed && settled < last) g_new = 0.;
    // Ridder's method
    for (int i = 0; i < 50; ++i) {
        if (!__any(on && mask != kDone && mask != kLowBound)) break;
        if (mask == kNoPhoton) {
            mask = kDone;
            view = 0.;
        } else if (mask == kNegligible) {
            view = g_old / psum;
            mask = kDone;
        } else if (mask == kBracketed) {
            mask = kOpt;
            w_mid = 0.5 * (w_old + w_new);
        } else if (mask == kOpt) {
            w_mid = 0.5 * (w_old + w_new);
        }
        const bool live = on && mask != kDone;
        if (!__any(live)) break;
        if (live) g_mid = gradient_rt(ph, prow, scale, bgscale, bgt, psum, w_mid, mask);
        if (mask == kNoPhoton) {
            mask = kDone;
            view = 0.;
        } else if (mask == kOpt) {
            if (fabs(g_mid) < grad_tol) {
                view = w_mid;
                mask = kDone;
            } else {
                const double sq = sqrt(g_mid * g_mid - g_old * g_new);
                if (sq < 1.e-15) {
                    w_lat = w_mid;
                } else {
                    w_lat = w_mid + (w_mid - w_old) * copysign(1.0, g_old) * g_mid / sq;
                    const double lo = fmin(w_old, w_new), hi = fmax(w_old, w_new);
                    if (w_lat <= lo || w_lat >= hi) w_lat = w_mid;
                }
            }
        }
        const bool live2 = on && mask != kDone;
        if (!__any(live2)) break;
        if (live2) g_lat = gradient_rt(ph, prow, scale, bgscale, bgt, psum, w_lat, mask);
        if (mask == kNoPhoton) {
            mask = kDone;
            view = 0.;
        } else if (mask == kOpt) {
            if (fabs(g_lat) < grad_tol) {
                view = w_lat;
                mask = kDone;
            } else if (g_mid * g_lat < 0.) {
                w_old = w_mid;
                g_old = g_mid;
                w_new = w_lat;
                g_new = g_lat;
            } else if (g_old * g_lat < 0.) {
                w_new = w_lat;
                g_new = g_lat;
            } else {
                w_old = w_lat;
                g_old = g_lat;
            }
        }
    }
    if (rmajor)
        view_out[(size_t)c * num_pix + t] = view;
    else
        view_out[(size_t)tile * num_pix * kRTile + (size_t)t * kRTile + lane] = view;
}

// gradient_d (max_scale.c:21-132), one orientation chunk: part[tile][d] = sum over the tile's
// orientations with P > 1e-6 of P times the sum over the frame's photons (mask < 1) of
// count * v / (v s_d + b_d B_t), v = W * rescale; frames with mask != 0 write 0.
__global__ __launch_bounds__(kBlock) void gradient_d_kernel(
    const double *__restrict__ W, size_t tsW, const double *__restrict__ P, int num_data, int r0, int C,
    int num_rot, const long *__restrict__ off, const int *__restrict__ pix, const double *__restrict__ cnt,
    const uint8_t *__restrict__ dmask, const double *__restrict__ s, const double *__restrict__ bgscale,
    const double *__restrict__ bg, double rescale, double *__restrict__ part) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int d = blockIdx.x * kWavesPerBlock + wave;
    if (d >= num_data) return;
    const int tile = blockIdx.y, c = tile * kRTile + lane, r = r0 + c;
    double acc = 0.;
    if (dmask[d] == 0 && c < C && r < num_rot) {
        const double pv = P[(size_t)(r / kRTile) * num_data * kRTile + (size_t)d * kRTile + (r % kRTile)];
        if (pv > kProbMin) {
            const double *__restrict__ col = W + (size_t)tile * tsW + lane;
            const double sd = s[d], bs = bgscale[d];
            for (long k = off[d]; k < off[d + 1]; ++k) {
                const int t = pix[k];
                const double v = col[(size_t)t * kRTile] * rescale;
                acc += pv * cnt[k] * v / (v * sd + bs * bg[t]);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) part[(size_t)tile * num_data + d] = acc;
}

// gacc[d] += sum over the chunk's tiles (in order) of part[tile][d]
__global__ void gradient_d_accumulate_kernel(const double *__restrict__ part, int ntiles, int num_data,
                                             double *__restrict__ gacc) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= num_data) return;
    double a = 0.;
    for (int j = 0; j < ntiles; ++j) a += part[(size_t)j * num_data + d];
    gacc[d] += a;
}

void launch_likelihood_bg(hipStream_t s, const double *W, int C, int num_pix, const long *off, const int *pix,
                          const double *cnt, const uint8_t *blacklist, int num_data, const double *scale,
                          const double *bgscale, const double *bg, double *G, int r0) {
    dim3 grid((num_data + kWavesPerBlock - 1) / kWavesPerBlock, C / kRTile);
    likelihood_bg_kernel<<<grid, kBlock, 0, s>>>(W, (size_t)num_pix * kRTile, off, pix, cnt, blacklist, num_data,
                                                 scale, bgscale, bg, G, (size_t)num_data * kRTile, r0);
}

void launch_scalemax(hipStream_t s, const double *P, int num_data, int Rp, const double *scale, double *part,
                     int nseg, double *out) {
    const int seg_len = (num_data + nseg - 1) / nseg;
    scalemax_part_kernel<<<dim3(Rp / kRTile, nseg), 64, 0, s>>>(P, num_data, scale, seg_len, part, Rp);
    scalemax_reduce_kernel<<<(Rp + 255) / 256, 256, 0, s>>>(part, nseg, Rp, out);
}

void launch_tomogram_bg(hipStream_t s, int first, const double *P, int num_data, int r0, int C, int num_rot,
                        const long *tom_off, const int *tom_frame, const double *tom_cnt, const double *scale,
                        const double *bgscale, const double *bg, const double *powder, const uint8_t *raw_mask,
                        int num_pix, double mean_count, const double *psum_r, const double *scalemax,
                        int *last_round, double *view, int rmajor) {
    dim3 grid((num_pix + kWavesPerBlock - 1) / kWavesPerBlock, (C + kRTile - 1) / kRTile);
    if (first)
        tomogram_bg_kernel<true><<<grid, kBlock, 0, s>>>(P, num_data, r0, C, num_rot, tom_off, tom_frame, tom_cnt,
                                                         scale, bgscale, bg, powder, raw_mask, num_pix, mean_count,
                                                         psum_r, scalemax, last_round, view, rmajor);
    else
        tomogram_bg_kernel<false><<<grid, kBlock, 0, s>>>(P, num_data, r0, C, num_rot, tom_off, tom_frame, tom_cnt,
                                                          scale, bgscale, bg, powder, raw_mask, num_pix, mean_count,
                                                          psum_r, scalemax, last_round, view, rmajor);
}

void launch_gradient_d(hipStream_t s, const double *W, int num_pix, const double *P, int num_data, int r0, int C,
                       int num_rot, const long *off, const int *pix, const double *cnt, const uint8_t *dmask,
                       const double *sc, const double *bgscale, const double *bg, double rescale, double *part,
                       double *gacc) {
    dim3 grid((num_data + kWavesPerBlock - 1) / kWavesPerBlock, C / kRTile);
    gradient_d_kernel<<<grid, kBlock, 0, s>>>(W, (size_t)num_pix * kRTile, P, num_data, r0, C, num_rot, off, pix,
                                              cnt, dmask, sc, bgscale, bg, rescale, part);
    gradient_d_accumulate_kernel<<<(num_data + 255) / 256, 256, 0, s>>>(part, C / kRTile, num_data, gacc);
}

}  // namespace emc
