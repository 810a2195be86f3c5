// C ABI of the MI355X EMC engine: device-resident state and the one-iteration driver.
// The ABI is declared (with the reference interfaces each entry point replaces) in
// include/emc_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cctype>
#include <climits>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/emc_hip.h"
#include "emc_common.h"
#include "emc_kernels.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// Counts device allocations and frees of the engine buffers (every engine): a captured graph
// holds buffer addresses, so it is captured again once any buffer may have moved.
std::atomic<unsigned long> g_alloc_gen{0};

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t err_ = (expr);                                                            \
        if (err_ != hipSuccess)                                                              \
            return fail(EMC_E_HIP, "%s failed: %s", #expr, hipGetErrorString(err_));        \
    } while (0)

// Device buffer with an element count.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;    // elements in use
    size_t cap = 0;  // elements allocated: a smaller request reuses the allocation
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), cap(o.cap) {
        o.p = nullptr;
        o.n = o.cap = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            cap = o.cap;
            o.p = nullptr;
            o.n = o.cap = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int alloc(size_t count) {
        if (p && count <= cap && count > 0) {
            n = count;
            return 0;
        }
        release();
        if (count == 0) return 0;
        ++g_alloc_gen;
        hipError_t err = hipMalloc((void **)&p, count * sizeof(T));
        if (err != hipSuccess) {
            p = nullptr;
            return fail(EMC_E_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                        hipGetErrorString(err));
        }
        n = cap = count;
        return 0;
    }
    int upload(const T *src, size_t count, hipStream_t s) {
        int rc = alloc(count);
        if (rc) return rc;
        if (count) HIP_TRY(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
        return 0;
    }
    void release() {
        if (p) {
            ++g_alloc_gen;
            (void)hipFree(p);
        }
        p = nullptr;
        n = cap = 0;
    }
};

// Page-locked host memory (grow only) that the per-iteration lists are built in, so that
// their uploads run at the host link's speed without a staging copy.
template <typename T>
struct PinnedBuf {
    T *p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
    int reserve(size_t count) {
        count = std::max<size_t>(count, 1);
        if (p && count <= cap) return 0;
        release();
        const size_t want = count + count / 4;  // room for lists that grow a little
        if (hipHostMalloc((void **)&p, want * sizeof(T), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return fail(EMC_E_NOMEM, "hipHostMalloc of %zu bytes failed", want * sizeof(T));
        }
        cap = want;
        return 0;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Host threads for a loop over n items, at least per items each, at most 16.
int host_threads(long n, long per) {
    const long hw = std::max(1u, std::thread::hardware_concurrency());
    return (int)std::max(1L, std::min(std::min(16L, hw), n / std::max(per, 1L)));
}

// fn(part, i0, i1) over nparts contiguous ranges of [0, n), part 0 on the calling thread.
// Returns EMC_E_NOMEM (with the message) when a part ran out of host memory, else 0: an
// exception must not leave a host thread.
template <typename F>
int parallel_ranges(long n, int nparts, F &&fn) {
    std::atomic<bool> failed{false};
    auto run = [&fn, &failed, n, nparts](int pi) {
        try {
            fn(pi, n * pi / nparts, n * (pi + 1) / nparts);
        } catch (const std::exception &) {
            failed = true;
        }
    };
    std::vector<std::thread> pool;
    for (int pi = 1; pi < nparts; ++pi) pool.emplace_back(run, pi);
    run(0);
    for (auto &th : pool) th.join();
    return failed ? fail(EMC_E_NOMEM, "host memory ran out while building lists") : 0;
}

void make_rot(const double *q, double *rot) {
    // make_rot_quat, dragonfly/src/model.c:759-786
    double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    double q01 = q0 * q1, q02 = q0 * q2, q03 = q0 * q3, q11 = q1 * q1, q12 = q1 * q2;
    double q13 = q1 * q3, q22 = q2 * q2, q23 = q2 * q3, q33 = q3 * q3;
    rot[0] = (1. - 2. * (q22 + q33));
    rot[1] = 2. * (q12 + q03);
    rot[2] = 2. * (q13 - q02);
    rot[3] = 2. * (q12 - q03);
    rot[4] = (1. - 2. * (q11 + q33));
    rot[5] = 2. * (q01 + q23);
    rot[6] = 2. * (q02 + q13);
    rot[7] = 2. * (q23 - q01);
    rot[8] = (1. - 2. * (q11 + q22));
}

// Engine pixel order: pixels sorted along a Morton (Z-order) curve of their transverse
// q-coordinates (qx, qy) quantised to the mean pixel spacing, so every 64 consecutive pixels
// form a compact detector patch (an 8x8 tile on a regular grid). perm[new] = old.
std::vector<int> morton_order(int num_pix, const double *qvals) {
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    for (int t = 0; t < num_pix; ++t) {
        x0 = std::min(x0, qvals[3 * t]);
        x1 = std::max(x1, qvals[3 * t]);
        y0 = std::min(y0, qvals[3 * t + 1]);
        y1 = std::max(y1, qvals[3 * t + 1]);
    }
    double area = (x1 - x0) * (y1 - y0);
    double step = area > 0 ? std::sqrt(area / num_pix) : 1.;
    if (!(step > 0)) step = 1.;
    std::vector<std::pair<uint64_t, int>> key(num_pix);
    for (int t = 0; t < num_pix; ++t) {
        uint32_t ix = (uint32_t)std::min(65535.0, std::max(0.0, std::floor((qvals[3 * t] - x0) / step + 0.5)));
        uint32_t iy = (uint32_t)std::min(65535.0, std::max(0.0, std::floor((qvals[3 * t + 1] - y0) / step + 0.5)));
        uint64_t code = 0;
        for (int b = 0; b < 16; ++b) code |= (uint64_t)((ix >> b) & 1) << (2 * b + 1) | (uint64_t)((iy >> b) & 1) << (2 * b);
        key[t] = {code, t};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> perm(num_pix);
    for (int t = 0; t < num_pix; ++t) perm[t] = key[t].second;
    return perm;
}

constexpr int kPixPerGroup = 1024;
constexpr int kFramesPerPsumPart = 512;
constexpr int kFramesPerTomoPart = 12500;  // default of the full_part_frames option
constexpr int kSigFullRowsPerTile = 1024;  // full-frame rows read by the pair tomogram, mean per tile
constexpr int kMaxBucketRanges = 8192;     // pixel ranges of the pair tomogram's buckets (LDS counters)
constexpr int kModelParts = 1024;

// Staging memory of emc_slice_gen_static: the rotation matrix at [0, 9) and the device copy
// of one view from kStaticViewOffset on. Each GPU holds its own instance; its address on a
// device comes from hipGetSymbolAddress with that device current (scratch_ptr below), never
// from the host-side name.
constexpr long kStaticViewOffset = 16;
static __device__ __align__(256) double scratch[kStaticViewOffset + EMC_STATIC_MAX_PIX];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, resolved once per device (the caller has made the
// device current). Returns 0, or an error code with the message set.
int resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(EMC_E_ARG, "device %d beyond %d devices", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void *ptr = nullptr;
    hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !ptr)
        return fail(EMC_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(err));
    g_scratch_ptr[device] = (double *)ptr;
    return 0;
}

}  // namespace

struct emc_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[10] = {};

    // detector
    int num_pix = 0;
    std::vector<uint8_t> h_mask;       // caller's pixel order
    std::vector<int> perm, inv;        // perm[engine] = caller, inv[caller] = engine
    DevBuf<double> qvals, corr;        // engine pixel order
    DevBuf<uint8_t> mask;
    DevBuf<int> d_perm;
    // orientations (rots padded to Rp with identity). With num_modes > 1 the engine's
    // orientations are the combined ones, r = rotation * num_modes + mode (max_prob.c:80-84):
    // num_rot = num_rot_user * num_modes, and quats / rots / logwr hold rotation r / num_modes.
    int num_rot = 0, Rp = 0, num_rot_user = 0, num_modes = 1;
    std::vector<double> h_quats;  // the caller's [num_rot_user][5]
    DevBuf<double> quats, rots, logwr;
    // model: num_modes volumes of size^3 each
    long size = 0;
    DevBuf<double> model1, model2, weight;
    DevBuf<double> quat_norm;  // [num_data][num_modes] occupancies when num_modes > 1
    // detector background (background-aware updates when the scale factors are on)
    bool with_bg = false;
    DevBuf<double> bg, powder, bgscale;  // engine pixel order; per frame
    DevBuf<double> bg_smax, bg_smax_part, bg_gpart, bg_gacc, bg_s;
    DevBuf<uint8_t> bg_dmask;
    DevBuf<int> bg_last;
    // frames
    int num_data = 0;
    std::vector<int> h_ones, h_multi, h_place_ones, h_place_multi, h_count_multi;
    std::vector<long> h_ones_accum, h_multi_accum;
    std::vector<uint8_t> h_dense;   // 1 for frames given as dense integer frames
    bool any_dense = false;
    DevBuf<uint8_t> dense;
    bool lists_dirty = true;
    DevBuf<long> lik_off, tom_off;
    DevBuf<int> lik_pix, tom_frame;
    DevBuf<double> lik_cnt, tom_cnt;
    // per (pixel, frame range) offsets into the tomogram list for the frame-partitioned form
    DevBuf<long> tom_off2;
    // host copies of the per-pixel photon lists (tom_off / tom_frame / tom_cnt)
    std::vector<long> h_tom_off;
    std::vector<int> h_tom_frame;
    std::vector<double> h_tom_cnt;
    // host copies of the per-frame likelihood lists, and the same lists split into lik_split
    // pixel ranges of about equal photon count (lik_split > 1: one range per XCD group)
    std::vector<long> h_lik_off;
    std::vector<int> h_lik_pix;
    std::vector<double> h_lik_cnt;
    int lik_split = 1, split_built = 0;
    DevBuf<long> lik8_off;
    DevBuf<int> lik8_pix;
    DevBuf<double> lik8_cnt, lik_part;
    int tom_parts = 8, tomo_width32 = 0;
    // upper bound on the orientation chunk width (0: memory decides)
    int max_chunk_cols = 0;
    DevBuf<double> tom_partial;
    DevBuf<float> tom_partial32;  // FP32 partial views of the full frames (subset path, FP32 gathers)
    // log W of every slot (one orientation chunk), left by an expectation step of the subset
    // path with the current model and geometry and not written over since: a later expectation
    // step (a second pass over new lists) skips the expand. Cleared by every change of model,
    // detector, orientations, modes, precision or options, by the M-step (the view takes the
    // buffer), and by the plain path.
    bool logw_resident = false;
    const void *res_logw = nullptr, *res_u = nullptr, *res_rbad = nullptr;
    int res_S = 0;
    // tile form of the tomogram list: tiles of 128 engine pixels -> frames -> photons
    DevBuf<long> ft_off, fe_off;
    DevBuf<int> ft_frame, fe_code;
    long tile_frames = 0, tile_entries = 0;
    // per-frame state
    DevBuf<double> scale, beta, sum_fact;
    DevBuf<uint8_t> blacklist;
    DevBuf<int> fcounts;
    double mean_count = 0.;
    int num_blacklist = 0;
    bool have_params = false;
    // work and results
    int C = 0;
    DevBuf<double> logw, upart, u, R, psum_r, psum_part, scalars, model_part;
    DevBuf<int> ubad, rbad;   // orientation has log W = -inf / NaN on a likelihood pixel
    DevBuf<double> max_exp, p_norm, likelihood, info, psum_d;
    DevBuf<int> rmax;
    DevBuf<double> one_rot, one_view;
    DevBuf<int> d_count;
    // inter_weight summed over every orientation: data-independent, reused whenever all
    // orientations have psum_r > 0
    DevBuf<double> weight_all;
    bool weight_all_valid = false;

    bool have_results = false;
    int last_unmerged = -1;
    // orientation tiles (of 64) per wavefront in the two photon-gather kernels; tomogram
    // form (2: frame-partitioned gather, 1: pixel-tile LDS accumulation, 0: per-pixel gather)
    int lik_tiles = 1, tomo_tiles = 1, tomo_tile_mode = 2;
    int lik_unroll = 8, tomo_unroll = 16, tomo_seq = 0, lik_xcd = 1;

    // mixed-precision dense-frame path (0: plain FP64; 1: FP16 log W; 2: FP16 hi + lo log W)
    int precision = 0;
    int max_count = 0;                  // largest photon count in the frames
    size_t mp_ldp = 0, mp_ldd = 0;      // padded pixel / frame counts (K of the two products)
    size_t mp_rows_d = 0, mp_rows_t = 0, mp_rows_c = 0;
    bool mp_built = false;
    DevBuf<_Float16> klik, ktom, lw_hi, lw_lo, p16;
    // FP32 gathers (precision 3): log W of one chunk and the chunk's posterior in FP32
    DevBuf<float> logw32, p32;
    DevBuf<double> center_t, center_base;

    // brick merge (1, default) or per-wave voxel boxes (0); the brick form needs an
    // orientation-major view, which the frame-partitioned tomogram and the mixed path write
    int merge_mode = 1;
    struct BrickChunk {
        int r0 = -1, Cc = 0, num_units = 0;
        DevBuf<int> list;
        DevBuf<unsigned long long> lmask;  // per list entry: the item's pixels with a corner in the brick
        DevBuf<int4> units;
    };
    std::vector<BrickChunk> bricks;
    bool bricks_valid = false;
    DevBuf<int> brick_counts, brick_fill;
    DevBuf<long> brick_off;
    // refinement: coarse cell of every orientation and per-frame bit rows of allowed cells
    bool refine = false;
    int rel_words = 0;
    DevBuf<int> qmap;
    DevBuf<uint32_t> relbits;
    // sparse evaluation of the subsets (default with refinement; EMC_SUBSET=0 evaluates every
    // pair and masks instead): orientations ordered by coarse cell in slots, cells padded to
    // tiles of 64, (frame, tile) pairs
    int subset_mode = 1;
    bool subset_ready = false, subset_ran = false;
    int num_slots = 0, num_tiles = 0, slot_C = 0;
    // slot order of the last coarse mapping (kept while the mapping does not change)
    std::vector<int> h_slot_qmap, h_tile0;  // h_tile0: first tile of each cell (per-cell padding)
    int slot_num_coarse = 0;
    // slot_pack = 1: cells placed back to back along a nearest-neighbour chain of their mean
    // orientations, no padding per cell; a tile may then hold orientations of two or three
    // cells, and slots outside a partial frame's cells are masked (slot_cell, relbits)
    int slot_pack = 1, slot_pack_built = -1;
    int sparse_rows = 0;  // frames per batch of the subset sparse posterior (0: 256 MB of rows)
    int full_part_frames = kFramesPerTomoPart;  // frames per range of the full-frame tomogram
    int sig_full_rows = kSigFullRowsPerTile;  // full-frame rows per tile the pair tomogram takes
    int unit_items = emc::kUnitItems;         // brick merge: list entries per work unit (block)
    int tomo_buckets = 1;                     // pair tomogram over per-(tile, range) buckets
    int estep_only = 0;   // stop after the posterior: no tomogram, merge or scale update
    int mstep_only = 0;   // subset path: resume after an estep_only iteration with the M-step
    int graph_mode = 0;   // subset path: run the iteration as a captured HIP graph
    int fused = 0;        // subset path: persistent expand + pair likelihood kernel
    DevBuf<unsigned> fused_counter;  // the next tile of the persistent kernel
    hipGraphExec_t graph_exec = nullptr;
    std::vector<long> graph_key;  // what the graph was captured for (buffers, sizes, options)
    unsigned long layout_gen = 0;  // bumped by every setter that changes sizes, lists or options
    // state generation: every setter (and model / scale update) increments it; an estep_only
    // subset iteration records it, and mstep_only runs only while it is unchanged
    unsigned long gen = 0, estep_gen = ~0ul;
    DevBuf<float> model32;        // FP32 brick copy of model1 for the FP32-gather expand
    bool model32_valid = false;   // cleared whenever model1 changes
    DevBuf<long> mass_off;
    DevBuf<int> mass_idx;
    DevBuf<double> mass_out;
    // emc_judge_outside: per-cell constants, per-frame skip flags and estimates
    DevBuf<double> judge_logn, judge_lwr, judge_out;
    DevBuf<uint8_t> judge_skip;
    std::vector<uint8_t> h_blacklist;  // host copy of the blacklist (emc_set_frame_params)
    std::vector<int> h_ctile0, h_ctile1;  // packed: first and last tile of each cell
    DevBuf<int> slot_cell;
    // full frames (every coarse cell): rows tile-major in RF; photon lists of mask < 2 pixels
    // of the full frames per pixel (frame index f), split into tom_parts frame ranges
    int num_full = 0, full_parts = 1;
    std::vector<int> h_fidx;
    DevBuf<int> fidx, ffull, ftom_frame;
    DevBuf<long> ftom_off2;
    DevBuf<double> ftom_cnt, RF, scale_full, psum_full, psum_full_part;
    DevBuf<int> d_fof, d_fcounts;   // full index of every frame (-1: partial) and list counts
    DevBuf<uint8_t> blacklist_full;
    bool full_lists_valid = false;
    // host scratch of the per-iteration subset set-up, kept so that its pages stay mapped
    PinnedBuf<uint32_t> h_relbits;
    std::vector<long> h_poff, h_toff, h_ftom_off2;
    std::vector<int> h_ffull, h_fof;
    PinnedBuf<int> h_pframe, h_ptile, h_tpair, h_tframe;
    long num_pairs = 0;
    DevBuf<int> slot_rot, rot_slot, pair_frame, pair_tile, tile_pair, tile_frame, rbad_slot;
    DevBuf<long> pair_off, tile_off;
    DevBuf<double> rots_slot, quats_slot, logwr_slot, u_slot, psum_slot, Rsub;
    DevBuf<float> Rsub32;  // FP32 copy of the pair rows' posterior (FP32 gathers)
    // pairs with some P > 1e-6, per tile (tile_off layout, sig_cnt per tile): the tomogram's
    DevBuf<int> sig_pair, sig_frame, sig_cnt;
    DevBuf<uint8_t> pair_flag;
    // per (chunk slot, 64-pixel patch): pixels with a nonzero view (pair tomogram -> merge)
    DevBuf<unsigned long long> nzmask;
    // full frames' rows with some P > 1e-6, per tile (T * nF layout), their count per tile and
    // in total: the pair tomogram reads them from RF when they are few
    DevBuf<uint8_t> full_flag;
    DevBuf<int> sigF_row, sigF_frame, sigF_cnt;
    DevBuf<long> sigF_total;
    DevBuf<unsigned long long> nz_count;  // nonzero view entries of the last subset iteration
    // cells the judge adds to rejected frames (emc_judge_widen): per frame (stride max_add),
    // their counts, and compacted frame after frame
    DevBuf<int> judge_nadd, judge_add, judge_cells;
    DevBuf<long> judge_off;
    long judge_cells_n = -1;
    bool nz_counted = false;
    std::vector<BrickChunk> slot_bricks;
    bool slot_bricks_valid = false;
    // frame photon lists of mask < 2 pixels sorted by engine pixel, with per-range offsets
    bool fsort_valid = false;
    int nranges = 0;
    DevBuf<long> fsort_off;
    DevBuf<int> fsort_pix, fsort_roff;
    DevBuf<double> fsort_cnt;
    // each frame's distinct pixel ranges (fsort order): fr_off per frame, the range and its
    // photons [fr_q0, fr_q1) per entry; and the pair tomogram's per-(tile, range) buckets
    DevBuf<long> fr_off, fr_q0, fr_q1, bk_tot, bk_base;
    DevBuf<int> fr_rg, bk_loc, bk_pair, bk_fr;

};

namespace {

// FP32 copy of model1 (every mode) for the FP32-gather expand, refreshed after model1 changed.
int ensure_model32(emc_engine *e) {
    if (e->model32_valid) return 0;
    int rc;
    if ((rc = e->model32.alloc(emc::model_brick_elems(e->size, e->num_modes)))) return rc;
    emc::launch_copy_f32_brick(e->stream, e->model1.p, e->size, e->num_modes, e->model32.p);
    HIP_TRY(hipGetLastError());
    e->model32_valid = true;
    return 0;
}

// precision modes of the dense-frame products: 1, 2 FP16 log W (one term, hi + lo); 4, 5 the
// same in BF16 (the tomogram product stays FP16)
inline bool mp_mode(const emc_engine *e) {
    return e->precision == 1 || e->precision == 2 || e->precision == 4 || e->precision == 5;
}
inline bool mp_two(const emc_engine *e) { return e->precision == 2 || e->precision == 5; }
inline bool mp_bf16(const emc_engine *e) { return e->precision == 4 || e->precision == 5; }

int check_ready(emc_engine *e) {
    if (!e) return fail(EMC_E_ARG, "null engine");
    if (!e->num_pix) return fail(EMC_E_STATE, "detector not set");
    if (!e->num_rot) return fail(EMC_E_STATE, "quaternions not set");
    if (!e->size) return fail(EMC_E_STATE, "model not set");
    if (!e->num_data) return fail(EMC_E_STATE, "frames not set");
    if (!e->have_params) return fail(EMC_E_STATE, "frame parameters not set");
    return 0;
}

// Photon lists of every frame in two device layouts:
//  * likelihood list: per frame, (pixel, count) for mask < 1, ones then multi
//    (the order calculate_prob sums them, max_prob.c:158-170)
//  * tomogram list: per pixel with mask < 2, (frame, count) with frames ascending
//    (the order update_tomogram_nobg accumulates them, max_update.c:196-209)
int build_lists(emc_engine *e) {
    const int N = e->num_data, P = e->num_pix;
    const uint8_t *mask = e->h_mask.data();
    std::vector<long> loff(N + 1, 0);
    std::vector<long> tcount(P + 1, 0);
    for (int d = 0; d < N; ++d) {
        long n = 0;
        for (int k = 0; k < e->h_ones[d]; ++k) {
            int p = e->h_place_ones[e->h_ones_accum[d] + k];
            if (p < 0 || p >= P) return fail(EMC_E_ARG, "frame %d: pixel %d out of range", d, p);
            if (mask[p] < 1) ++n;
            if (mask[p] < 2) ++tcount[e->inv[p] + 1];
        }
        for (int k = 0; k < e->h_multi[d]; ++k) {
            int p = e->h_place_multi[e->h_multi_accum[d] + k];
            if (p < 0 || p >= P) return fail(EMC_E_ARG, "frame %d: pixel %d out of range", d, p);
            if (mask[p] < 1) ++n;
            if (mask[p] < 2) ++tcount[e->inv[p] + 1];
        }
        loff[d + 1] = loff[d] + n;
    }
    for (int p = 0; p < P; ++p) tcount[p + 1] += tcount[p];
    std::vector<int> lpix(loff[N]);
    std::vector<double> lcnt(loff[N]);
    std::vector<int> tframe(tcount[P]);
    std::vector<double> tcnt(tcount[P]);
    std::vector<long> tfill(tcount.begin(), tcount.end() - 1);
    int max_count = 1;
    for (int d = 0; d < N; ++d) {
        long at = loff[d];
        for (int k = 0; k < e->h_ones[d]; ++k) {
            int p = e->h_place_ones[e->h_ones_accum[d] + k], q = e->inv[p];
            if (mask[p] < 1) { lpix[at] = q; lcnt[at] = 1.; ++at; }
            if (mask[p] < 2) { tframe[tfill[q]] = d; tcnt[tfill[q]] = 1.; ++tfill[q]; }
        }
        for (int k = 0; k < e->h_multi[d]; ++k) {
            long j = e->h_multi_accum[d] + k;
            int p = e->h_place_multi[j], q = e->inv[p];
            double c = (double)e->h_count_multi[j];
            max_count = std::max(max_count, e->h_count_multi[j]);
            if (mask[p] < 1) { lpix[at] = q; lcnt[at] = c; ++at; }
            if (mask[p] < 2) { tframe[tfill[q]] = d; tcnt[tfill[q]] = c; ++tfill[q]; }
        }
    }
    // frame-range offsets: toff2[t * nparts + x] = first photon of pixel t in frame range x
    const int nparts = e->tom_parts;
    std::vector<long> toff2((size_t)P * nparts + 1);
    for (int t = 0; t < P; ++t) {
        long k = tcount[t];
        for (int x = 0; x < nparts; ++x) {
            const long lo = (long)N * x / nparts;
            while (k < tcount[t + 1] && tframe[k] < lo) ++k;
            toff2[(size_t)t * nparts + x] = k;
        }
    }
    toff2[(size_t)P * nparts] = tcount[P];

    // tile form: photons of mask < 2 pixels grouped by 128-pixel tile, then frame
    const int ntiles = (P + 127) / 128;
    std::vector<long> tf_count(ntiles + 1, 0), te_count(ntiles + 1, 0);
    std::vector<int> last(ntiles, -1);
    auto visit = [&](auto &&fn) {
        for (int d = 0; d < N; ++d) {
            for (int k = 0; k < e->h_ones[d]; ++k) {
                int p = e->h_place_ones[e->h_ones_accum[d] + k];
                if (mask[p] < 2) fn(d, e->inv[p], 1);
            }
            for (int k = 0; k < e->h_multi[d]; ++k) {
                long j = e->h_multi_accum[d] + k;
                int p = e->h_place_multi[j];
                if (mask[p] < 2) fn(d, e->inv[p], e->h_count_multi[j]);
            }
        }
    };
    bool big_count = false;
    visit([&](int d, int q, int c) {
        const int tl = q / 128;
        if (last[tl] != d) { last[tl] = d; ++tf_count[tl + 1]; }
        ++te_count[tl + 1];
        if (c >= (1 << 23) || c < 0) big_count = true;
    });
    if (big_count) return fail(EMC_E_ARG, "photon count too large for the tile tomogram list");
    for (int tl = 0; tl < ntiles; ++tl) {
        tf_count[tl + 1] += tf_count[tl];
        te_count[tl + 1] += te_count[tl];
    }
    const long nf = tf_count[ntiles], ne = te_count[ntiles];
    std::vector<int> ft_frame(nf > 0 ? nf : 1), fe_code(ne > 0 ? ne : 1);
    std::vector<long> fe_off(nf + 1, 0);
    std::vector<long> fpos(tf_count.begin(), tf_count.end() - 1), epos(te_count.begin(), te_count.end() - 1);
    std::fill(last.begin(), last.end(), -1);
    // entries of one (tile, frame) are contiguous because frames are visited in order
    visit([&](int d, int q, int c) {
        const int tl = q / 128;
        if (last[tl] != d) {
            last[tl] = d;
            ft_frame[fpos[tl]] = d;
            fe_off[fpos[tl]] = epos[tl];
            ++fpos[tl];
        }
        fe_code[epos[tl]++] = (c << 8) | (q & 127);
    });
    fe_off[nf] = ne;
    // fe_off of the first frame of each tile equals te_count[tile]; fill ends: the frame
    // after the last of a tile starts the next tile's entries, so fe_off is monotone
    e->tile_frames = nf;
    e->tile_entries = ne;

    int rc;
    if ((rc = e->tom_off2.upload(toff2.data(), toff2.size(), e->stream))) return rc;
    e->any_dense = false;
    for (uint8_t v : e->h_dense) e->any_dense |= (v != 0);
    if ((rc = e->dense.upload(e->h_dense.data(), e->h_dense.size(), e->stream))) return rc;
    if ((rc = e->ft_off.upload(tf_count.data(), tf_count.size(), e->stream))) return rc;
    if ((rc = e->ft_frame.upload(ft_frame.data(), ft_frame.size(), e->stream))) return rc;
    if ((rc = e->fe_off.upload(fe_off.data(), fe_off.size(), e->stream))) return rc;
    if ((rc = e->fe_code.upload(fe_code.data(), fe_code.size(), e->stream))) return rc;
    if ((rc = e->lik_off.upload(loff.data(), loff.size(), e->stream))) return rc;
    if ((rc = e->lik_pix.upload(lpix.data(), lpix.size() ? lpix.size() : 0, e->stream))) return rc;
    if ((rc = e->lik_cnt.upload(lcnt.data(), lcnt.size(), e->stream))) return rc;
    if ((rc = e->tom_off.upload(tcount.data(), tcount.size(), e->stream))) return rc;
    if ((rc = e->tom_frame.upload(tframe.data(), tframe.size(), e->stream))) return rc;
    if ((rc = e->tom_cnt.upload(tcnt.data(), tcnt.size(), e->stream))) return rc;
    // empty lists still need a valid pointer for the kernels
    if (!e->lik_pix.p && (rc = e->lik_pix.alloc(1))) return rc;
    if (!e->lik_cnt.p && (rc = e->lik_cnt.alloc(1))) return rc;
    if (!e->tom_frame.p && (rc = e->tom_frame.alloc(1))) return rc;
    if (!e->tom_cnt.p && (rc = e->tom_cnt.alloc(1))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->h_lik_off = std::move(loff);
    e->h_lik_pix = std::move(lpix);
    e->h_lik_cnt = std::move(lcnt);
    e->split_built = 0;
    e->h_tom_off = std::move(tcount);
    e->h_tom_frame = std::move(tframe);
    e->h_tom_cnt = std::move(tcnt);
    e->max_count = max_count;
    e->lists_dirty = false;
    e->mp_built = false;
    e->fsort_valid = false;
    e->full_lists_valid = false;
    return 0;
}

size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// Likelihood lists split into X pixel ranges of about equal photon count: frame d's photons
// of range x are lik8_pix[lik8_off[d * (X + 1) + x] .. lik8_off[d * (X + 1) + x + 1]), in
// their list order. Engine pixels follow a Morton curve, so a range is a compact detector
// region and its rows of a log W tile are a compact slab.
int build_lik_split(emc_engine *e, int X) {
    const int N = e->num_data, P = e->num_pix;
    const std::vector<long> &off = e->h_lik_off;
    const std::vector<int> &pix = e->h_lik_pix;
    std::vector<long> hist(P + 1, 0);
    for (int q : pix) ++hist[q + 1];
    for (int t = 0; t < P; ++t) hist[t + 1] += hist[t];
    const long total = hist[P];
    std::vector<uint8_t> range_of(P);
    int x = 0;
    for (int t = 0; t < P; ++t) {
        while (x < X - 1 && hist[t] >= (long)((double)total * (x + 1) / X)) ++x;
        range_of[t] = (uint8_t)x;
    }
    std::vector<long> off8((size_t)N * (X + 1));
    std::vector<int> pix8(pix.size() > 0 ? pix.size() : 1);
    std::vector<double> cnt8(pix.size() > 0 ? pix.size() : 1);
    std::vector<long> fill(X);
    for (int d = 0; d < N; ++d) {
        std::fill(fill.begin(), fill.end(), 0);
        for (long k = off[d]; k < off[d + 1]; ++k) ++fill[range_of[pix[k]]];
        long at = off[d];
        for (int r = 0; r < X; ++r) {
            off8[(size_t)d * (X + 1) + r] = at;
            const long n = fill[r];
            fill[r] = at;
            at += n;
        }
        off8[(size_t)d * (X + 1) + X] = at;
        for (long k = off[d]; k < off[d + 1]; ++k) {
            const long to = fill[range_of[pix[k]]]++;
            pix8[to] = pix[k];
            cnt8[to] = e->h_lik_cnt[k];
        }
    }
    int rc;
    if ((rc = e->lik8_off.upload(off8.data(), off8.size(), e->stream))) return rc;
    if ((rc = e->lik8_pix.upload(pix8.data(), pix8.size(), e->stream))) return rc;
    if ((rc = e->lik8_cnt.upload(cnt8.data(), cnt8.size(), e->stream))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->split_built = X;
    return 0;
}

// Per-chunk brick item lists and work units of the brick merge (geometry only). `rots` holds
// `total` rotations (a multiple of 64) walked in chunks of C; rotations at or past num_valid,
// and slots with slot_rot < 0 when slot_rot is given, make no items.
int build_bricks(emc_engine *e, std::vector<emc_engine::BrickChunk> &out, const double *rots, int total,
                 int C, int num_valid, const int *slot_rot, int modes) {
    hipStream_t s = e->stream;
    const int P = e->num_pix;
    const int nbx = (int)((e->size + emc::kBrick - 1) / emc::kBrick);
    const size_t nb = (size_t)nbx * nbx * nbx * modes;  // bricks of every mode's volume
    const int npatch = (P + 63) / 64;
    int rc;
    if ((rc = e->brick_counts.alloc(nb))) return rc;
    if ((rc = e->brick_fill.alloc(nb))) return rc;
    if ((rc = e->brick_off.alloc(nb))) return rc;
    std::vector<int> counts(nb);
    std::vector<long> off(nb);
    out.clear();
    for (int r0 = 0; r0 < total; r0 += C) {
        const int Cc = (total - r0 < C) ? (total - r0) : C;
        out.emplace_back();
        emc_engine::BrickChunk &bc = out.back();
        bc.r0 = r0;
        bc.Cc = Cc;
        HIP_TRY(hipMemsetAsync(e->brick_counts.p, 0, nb * sizeof(int), s));
        emc::launch_brick_items(s, rots, r0, Cc, num_valid, e->qvals.p, e->mask.p, P, e->size, npatch, nbx, 0,
                                e->brick_counts.p, nullptr, nullptr, nullptr, nullptr, slot_rot, modes);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(counts.data(), e->brick_counts.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        long tot = 0;
        std::vector<int4> units;
        for (size_t b = 0; b < nb; ++b) {
            off[b] = tot;
            for (int k = 0; k < counts[b]; k += e->unit_items)
                units.push_back(make_int4((int)b, (int)(tot + k),
                                          (int)(tot + std::min(counts[b], k + e->unit_items)), 0));
            tot += counts[b];
        }
        if (tot >= INT32_MAX) return fail(EMC_E_ARG, "brick merge: too many work items");
        if ((rc = bc.list.alloc(tot > 0 ? tot : 1))) return rc;
        if ((rc = bc.lmask.alloc(tot > 0 ? tot : 1))) return rc;
        if ((rc = e->brick_off.upload(off.data(), nb, s))) return rc;
        HIP_TRY(hipMemsetAsync(e->brick_fill.p, 0, nb * sizeof(int), s));
        emc::launch_brick_items(s, rots, r0, Cc, num_valid, e->qvals.p, e->mask.p, P, e->size, npatch, nbx, 1,
                                nullptr, e->brick_off.p, e->brick_fill.p, bc.list.p, bc.lmask.p, slot_rot, modes);
        HIP_TRY(hipGetLastError());
        bc.num_units = (int)units.size();
        if (!units.empty() && (rc = bc.units.upload(units.data(), units.size(), s))) return rc;
        HIP_TRY(hipStreamSynchronize(s));
    }
    return 0;
}

int build_bricks(emc_engine *e) {
    int rc = build_bricks(e, e->bricks, e->rots.p, e->Rp, e->C, e->num_rot, nullptr, e->num_modes);
    if (!rc) e->bricks_valid = true;
    return rc;
}

int merge_chunk(const emc_engine *e, const emc_engine::BrickChunk &bc, const double *rots, int num_valid,
                const double *view, const double *psum, bool all_on, bool do_model, bool do_weight,
                const int *gate = nullptr, const float *view32 = nullptr,
                const unsigned long long *nzmask = nullptr) {
    const int nbx = (int)((e->size + emc::kBrick - 1) / emc::kBrick);
    if ((view32 || nzmask) && (!do_model || do_weight))
        return fail(EMC_E_STATE, "internal: an FP32 or masked view merges the model only");
    if (view32) {
        // an FP32 view: FP32 brick sums
        emc::launch_merge_model_f32(e->stream, bc.units.p, bc.num_units, bc.list.p, bc.lmask.p, rots, bc.r0, num_valid,
                                    view32, psum, all_on, e->qvals.p, e->corr.p, e->mask.p, e->num_pix, e->size,
                                    (e->num_pix + 63) / 64, nbx, e->model2.p, nzmask);
        return 0;
    }
    emc::launch_merge_bricks(e->stream, bc.units.p, bc.num_units, bc.list.p, bc.lmask.p, rots, bc.r0, num_valid,
                             view, psum,
                             all_on, e->qvals.p, e->corr.p, e->mask.p, e->num_pix, e->size,
                             (e->num_pix + 63) / 64, nbx, e->model2.p, e->weight.p, do_model, do_weight,
                             e->precision != 0, gate, nzmask);  // reduced-precision modes sum bricks in FP32
    return 0;
}

// Slot order and (frame, tile) pairs of the refinement subsets: orientations grouped by coarse
// cell (ascending index inside a cell), every cell padded to whole tiles of 64. Frames whose
// set holds every cell are full frames (rows in RF, tile-major); the pairs of the others are
// the tiles of their (distinct) cells in ascending tile order, and the tile-major copy of
// that pair list (tile_off / tile_pair) lists each tile's pairs by ascending frame.
int build_subset(emc_engine *e, int num_coarse, const int *qmap, const long *rel_off, const int *rel_quat) {
    const int Rn = e->num_rot, N = e->num_data;
    hipStream_t s = e->stream;
    int rc;
    const bool pack = e->slot_pack != 0;
    const bool same_slots = e->slot_num_coarse == num_coarse && (int)e->h_slot_qmap.size() == Rn &&
                            e->slot_pack_built == (int)pack &&
                            std::equal(e->h_slot_qmap.begin(), e->h_slot_qmap.end(), qmap);
    if (!same_slots) {
        e->logw_resident = false;  // a new slot order
        std::vector<int> cnt(num_coarse, 0), tile0(num_coarse + 1, 0);
        for (int r = 0; r < Rn; ++r) ++cnt[qmap[r]];
        int T = 0;
        std::vector<int> slot_rot, rot_slot(Rn), cell_slot;
        std::vector<int> ctile0(num_coarse, 0), ctile1(num_coarse, -1);
        if (pack) {
            // cells along a greedy nearest-neighbour chain of their mean orientations (from the
            // first cell, always on to the closest cell not yet placed), so that the cells
            // around an orientation tend to sit next to each other in slot order; members in
            // ascending orientation index
            std::vector<double> mean((size_t)num_coarse * 4, 0.), first((size_t)num_coarse * 4, 0.);
            std::vector<uint8_t> has(num_coarse, 0);
            for (int r = 0; r < Rn; ++r) {
                const int c = qmap[r];
                const double *q = &e->h_quats[5 * (size_t)r];
                if (!has[c]) {
                    has[c] = 1;
                    for (int k = 0; k < 4; ++k) first[4 * (size_t)c + k] = q[k];
                }
                double dot = 0.;
                for (int k = 0; k < 4; ++k) dot += q[k] * first[4 * (size_t)c + k];
                const double sg = dot < 0. ? -1. : 1.;
                for (int k = 0; k < 4; ++k) mean[4 * (size_t)c + k] += sg * q[k];
            }
            for (int c = 0; c < num_coarse; ++c) {
                double *m = &mean[4 * (size_t)c];
                double nrm = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2] + m[3] * m[3]);
                if (nrm == 0.) nrm = 1.;
                for (int k = 0; k < 4; ++k) m[k] /= nrm;
            }
            std::vector<int> chain;
            chain.reserve(num_coarse);
            std::vector<uint8_t> placed(num_coarse, 0);
            int cur = 0;
            for (int n = 0; n < num_coarse; ++n) {
                chain.push_back(cur);
                placed[cur] = 1;
                const double *a = &mean[4 * (size_t)cur];
                int next = -1;
                double best = -1.;
                for (int c = 0; c < num_coarse; ++c) {
                    if (placed[c]) continue;
                    const double *b = &mean[4 * (size_t)c];
                    const double dd = std::fabs(a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
                    if (dd > best) {
                        best = dd;
                        next = c;
                    }
                }
                if (next < 0) break;
                cur = next;
            }
            std::vector<int> start(num_coarse, 0), fill(num_coarse, 0);
            int at = 0;
            for (int c : chain) {
                start[c] = at;
                at += cnt[c];
            }
            const int S = (Rn + emc::kRTile - 1) / emc::kRTile * emc::kRTile;
            T = S / emc::kRTile;
            slot_rot.assign(S, -1);
            cell_slot.assign(S, -1);
            for (int r = 0; r < Rn; ++r) {
                const int c = qmap[r];
                const int sl = start[c] + fill[c]++;
                slot_rot[sl] = r;
                rot_slot[r] = sl;
                cell_slot[sl] = c;
            }
            for (int c = 0; c < num_coarse; ++c)
                if (cnt[c] > 0) {
                    ctile0[c] = start[c] / emc::kRTile;
                    ctile1[c] = (start[c] + cnt[c] - 1) / emc::kRTile;
                }
        } else {
            for (int c = 0; c < num_coarse; ++c) tile0[c + 1] = tile0[c] + (cnt[c] + emc::kRTile - 1) / emc::kRTile;
            T = tile0[num_coarse];
            slot_rot.assign((size_t)T * emc::kRTile, -1);
            std::vector<int> fill(num_coarse, 0);
            for (int r = 0; r < Rn; ++r) {
                const int c = qmap[r];
                const int sl = tile0[c] * emc::kRTile + fill[c]++;
                slot_rot[sl] = r;
                rot_slot[r] = sl;
            }
            for (int c = 0; c < num_coarse; ++c) {
                ctile0[c] = tile0[c];
                ctile1[c] = tile0[c + 1] - 1;
            }
        }
        const int S = T * emc::kRTile;
        if (S <= 0) return fail(EMC_E_ARG, "emc_set_refine: no orientations");
        std::vector<double> rots((size_t)S * 9, 0.), quats((size_t)S * 5, 0.), lw(S, 0.);
        for (int sl = 0; sl < S; ++sl) {
            const int r = slot_rot[sl];
            if (r >= 0) {
                make_rot(&e->h_quats[5 * (size_t)r], &rots[9 * (size_t)sl]);
                for (int k = 0; k < 5; ++k) quats[5 * (size_t)sl + k] = e->h_quats[5 * (size_t)r + k];
                lw[sl] = std::log(e->h_quats[5 * (size_t)r + 4]);
            } else {
                rots[9 * (size_t)sl] = rots[9 * (size_t)sl + 4] = rots[9 * (size_t)sl + 8] = 1.;
            }
        }
        if ((rc = e->slot_rot.upload(slot_rot.data(), S, s))) return rc;
        if ((rc = e->rot_slot.upload(rot_slot.data(), Rn, s))) return rc;
        if ((rc = e->rots_slot.upload(rots.data(), rots.size(), s))) return rc;
        if ((rc = e->quats_slot.upload(quats.data(), quats.size(), s))) return rc;
        if ((rc = e->logwr_slot.upload(lw.data(), S, s))) return rc;
        if (pack) {
            if ((rc = e->slot_cell.upload(cell_slot.data(), S, s))) return rc;
        } else {
            e->slot_cell.release();
        }
        HIP_TRY(hipStreamSynchronize(s));
        e->h_slot_qmap.assign(qmap, qmap + Rn);
        e->h_tile0 = tile0;
        e->h_ctile0 = ctile0;
        e->h_ctile1 = ctile1;
        e->slot_pack_built = (int)pack;
        e->slot_num_coarse = num_coarse;
        e->num_slots = S;
        e->num_tiles = T;
        e->slot_bricks_valid = false;
    }
    const std::vector<int> &ct0 = e->h_ctile0, &ct1 = e->h_ctile1;
    const int T = e->num_tiles;
    // (frame, tile) pairs from the frames' bit rows (built by emc_set_refine): host threads list
    // ranges of frames (a frame's distinct tiles, ascending), the frame offsets follow, then
    // the threads write the pairs and, from per-thread tile counts, the tile-major copy in
    // frame order
    const int words = e->rel_words;
    const uint32_t *bits = e->h_relbits.p;
    const int nth = host_threads(N, 2048);
    struct Part {
        std::vector<int> tiles;   // the range's frames' tiles, frame after frame
        std::vector<long> hist;  // pairs per tile
    };
    std::vector<Part> parts(nth);
    std::vector<int> ntl(N);  // tiles per frame, -1 for a full frame
    if ((rc = parallel_ranges(N, nth, [&](int pi, long d0, long d1) {
        Part &pt = parts[pi];
        std::vector<int> tl;
        for (long d = d0; d < d1; ++d) {
            const uint32_t *row = bits + (size_t)d * words;
            int n = 0;
            for (int w = 0; w < words; ++w) n += __builtin_popcount(row[w]);
            if (n == num_coarse) {
                ntl[d] = -1;
                continue;
            }
            // the tiles holding the frame's cells, ascending and distinct (with packed cells
            // neighbouring cells share tiles)
            tl.clear();
            for (int w = 0; w < words; ++w)
                for (uint32_t x = row[w]; x; x &= x - 1) {
                    const int c = w * 32 + __builtin_ctz(x);
                    for (int t = ct0[c]; t <= ct1[c]; ++t) tl.push_back(t);
                }
            std::sort(tl.begin(), tl.end());
            tl.erase(std::unique(tl.begin(), tl.end()), tl.end());
            ntl[d] = (int)tl.size();
            pt.tiles.insert(pt.tiles.end(), tl.begin(), tl.end());
        }
    })))
        return rc;
    std::vector<long> &poff = e->h_poff;
    std::vector<int> &ffull = e->h_ffull;
    std::vector<int> fidx;
    poff.assign(N + 1, 0);
    ffull.assign(N, -1);
    for (int d = 0; d < N; ++d) {
        poff[d + 1] = poff[d] + std::max(ntl[d], 0);
        if (ntl[d] < 0) {
            ffull[d] = (int)fidx.size();
            fidx.push_back(d);
        }
    }
    const long K = poff[N];
    if (K >= INT32_MAX) return fail(EMC_E_ARG, "emc_set_refine: too many (frame, tile) pairs");
    if ((rc = e->h_pframe.reserve(K)) || (rc = e->h_ptile.reserve(K)) || (rc = e->h_tpair.reserve(K)) ||
        (rc = e->h_tframe.reserve(K)))
        return rc;
    int *pframe = e->h_pframe.p, *ptile = e->h_ptile.p, *tpair = e->h_tpair.p, *tframe = e->h_tframe.p;
    pframe[0] = ptile[0] = tpair[0] = tframe[0] = 0;  // one element at least is uploaded
    if ((rc = parallel_ranges(N, nth, [&](int pi, long d0, long d1) {
        Part &pt = parts[pi];
        pt.hist.assign(T, 0);
        size_t i = 0;
        for (long d = d0; d < d1; ++d)
            for (long k = poff[d]; k < poff[d + 1]; ++k) {
                const int t = pt.tiles[i++];
                pframe[k] = (int)d;
                ptile[k] = t;
                ++pt.hist[t];
            }
    })))
        return rc;
    std::vector<long> &toff = e->h_toff;
    toff.assign(T + 1, 0);
    for (int t = 0; t < T; ++t) {
        long at = toff[t];
        for (Part &pt : parts) {
            const long n = pt.hist[t];
            pt.hist[t] = at;  // the part's first position in tile t
            at += n;
        }
        toff[t + 1] = at;
    }
    if ((rc = parallel_ranges(N, nth, [&](int pi, long d0, long d1) {
        std::vector<long> &at = parts[pi].hist;
        for (long k = poff[d0]; k < poff[d1]; ++k) {
            const long j = at[ptile[k]]++;
            tpair[j] = (int)k;
            tframe[j] = pframe[k];  // the frame beside the pair: the pair tomogram reads it with it
        }
    })))
        return rc;
    const size_t Ku = (size_t)std::max(K, 1L);
    if ((rc = e->pair_off.upload(poff.data(), poff.size(), s))) return rc;
    if ((rc = e->pair_frame.upload(pframe, Ku, s))) return rc;
    if ((rc = e->pair_tile.upload(ptile, Ku, s))) return rc;
    if ((rc = e->tile_off.upload(toff.data(), toff.size(), s))) return rc;
    if ((rc = e->tile_pair.upload(tpair, Ku, s))) return rc;
    if ((rc = e->tile_frame.upload(tframe, Ku, s))) return rc;
    if ((rc = e->ffull.upload(ffull.data(), ffull.size(), s))) return rc;
    if (!fidx.empty() && (rc = e->fidx.upload(fidx.data(), fidx.size(), s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    e->num_pairs = K;
    e->num_full = (int)fidx.size();
    e->h_fidx = std::move(fidx);
    e->full_lists_valid = false;
    e->subset_ready = true;
    return 0;
}

// Per-pixel photon lists (mask < 2) restricted to the full frames, frame index f, with the
// offsets of tom_parts ranges of f: the plain path's partitioned tomogram lists for RF.
int build_full_lists(emc_engine *e) {
    // frame ranges of about kFramesPerTomoPart frames (at most tom_parts): fewer full frames
    // share fewer partial views, whose writes and reads would otherwise dominate
    const int P = e->num_pix, nF = e->num_full, N = e->num_data;
    const int per = e->full_part_frames;
    const int nparts = std::max(1, std::min(std::min(e->tom_parts, 64), (nF + per - 1) / per));
    e->full_parts = nparts;
    hipStream_t s = e->stream;
    std::vector<int> &fof = e->h_fof;
    fof.assign(N, -1);
    for (int f = 0; f < nF; ++f) fof[e->h_fidx[f]] = f;
    int rc;
    const size_t ncell = (size_t)P * nparts;
    if ((rc = e->d_fof.upload(fof.data(), fof.size(), s))) return rc;
    if ((rc = e->d_fcounts.alloc(ncell))) return rc;
    // counts per (pixel, range) on the GPU, the scan here, then the fill on the GPU
    emc::launch_full_lists(s, 0, e->tom_off.p, e->tom_frame.p, e->tom_cnt.p, e->d_fof.p, P, std::max(nF, 1),
                           nparts, e->d_fcounts.p, nullptr, nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    std::vector<int> cnt(ncell);
    HIP_TRY(hipMemcpyAsync(cnt.data(), e->d_fcounts.p, ncell * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<long> &off2 = e->h_ftom_off2;
    off2.resize(ncell + 1);
    long total = 0;
    for (size_t i = 0; i < ncell; ++i) {
        off2[i] = total;
        total += cnt[i];
    }
    off2[ncell] = total;
    if ((rc = e->ftom_off2.upload(off2.data(), off2.size(), s))) return rc;
    if ((rc = e->ftom_frame.alloc((size_t)std::max(total, 1L)))) return rc;
    if ((rc = e->ftom_cnt.alloc((size_t)std::max(total, 1L)))) return rc;
    emc::launch_full_lists(s, 1, e->tom_off.p, e->tom_frame.p, e->tom_cnt.p, e->d_fof.p, P, std::max(nF, 1), nparts,
                           nullptr, e->ftom_off2.p, e->ftom_frame.p, e->ftom_cnt.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    e->full_lists_valid = true;
    return 0;
}

// Photons of mask < 2 pixels per frame, sorted by engine pixel, with the offset (relative to
// the frame's first photon) of the first photon at or past every range start p * kPairRange.
int build_fsort(emc_engine *e) {
    const int N = e->num_data, P = e->num_pix;
    const int nr = (P + emc::kPairRange - 1) / emc::kPairRange;
    const uint8_t *mask = e->h_mask.data();
    std::vector<long> off(N + 1, 0), r_off(N + 1, 0), r_q0, r_q1;
    std::vector<int> pix, roff((size_t)N * (nr + 1)), r_rg;
    std::vector<double> cnt;
    std::vector<std::pair<int, double>> fr;
    for (int d = 0; d < N; ++d) {
        fr.clear();
        for (int k = 0; k < e->h_ones[d]; ++k) {
            const int p = e->h_place_ones[e->h_ones_accum[d] + k];
            if (mask[p] < 2) fr.emplace_back(e->inv[p], 1.);
        }
        for (int k = 0; k < e->h_multi[d]; ++k) {
            const long j = e->h_multi_accum[d] + k;
            const int p = e->h_place_multi[j];
            if (mask[p] < 2) fr.emplace_back(e->inv[p], (double)e->h_count_multi[j]);
        }
        std::stable_sort(fr.begin(), fr.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        int q = 0;
        for (int rg = 0; rg <= nr; ++rg) {
            while (q < (int)fr.size() && fr[q].first < rg * emc::kPairRange) ++q;
            roff[(size_t)d * (nr + 1) + rg] = q;
        }
        // the frame's distinct ranges and their photons
        const long base = (long)pix.size();
        for (size_t a = 0; a < fr.size();) {
            const int rg = fr[a].first / emc::kPairRange;
            size_t b = a + 1;
            while (b < fr.size() && fr[b].first / emc::kPairRange == rg) ++b;
            r_rg.push_back(rg);
            r_q0.push_back(base + (long)a);
            r_q1.push_back(base + (long)b);
            a = b;
        }
        r_off[d + 1] = (long)r_rg.size();
        for (const auto &x : fr) {
            pix.push_back(x.first);
            cnt.push_back(x.second);
        }
        off[d + 1] = (long)pix.size();
    }
    if (r_rg.empty()) {
        r_rg.push_back(0);
        r_q0.push_back(0);
        r_q1.push_back(0);
    }
    if (pix.empty()) {
        pix.push_back(0);
        cnt.push_back(0.);
    }
    int rc;
    hipStream_t s = e->stream;
    if ((rc = e->fsort_off.upload(off.data(), off.size(), s))) return rc;
    if ((rc = e->fsort_pix.upload(pix.data(), pix.size(), s))) return rc;
    if ((rc = e->fsort_cnt.upload(cnt.data(), cnt.size(), s))) return rc;
    if ((rc = e->fsort_roff.upload(roff.data(), roff.size(), s))) return rc;
    if ((rc = e->fr_off.upload(r_off.data(), r_off.size(), s)) || (rc = e->fr_rg.upload(r_rg.data(), r_rg.size(), s)) ||
        (rc = e->fr_q0.upload(r_q0.data(), r_q0.size(), s)) || (rc = e->fr_q1.upload(r_q1.data(), r_q1.size(), s)))
        return rc;
    HIP_TRY(hipStreamSynchronize(s));
    e->nranges = nr;
    e->fsort_valid = true;
    return 0;
}

int merge_chunk(emc_engine *e, int chunk, const double *view, bool all_on, bool do_model, bool do_weight) {
    return merge_chunk(e, e->bricks[chunk], e->rots.p, e->num_rot, view, e->psum_r.p, all_on, do_model,
                       do_weight);
}

// Work buffers. The posterior matrix R/P holds all frames x all orientations (the keep
// rule needs each frame's global maximum before any normalisation); log W / view hold one
// orientation chunk of C columns, C chosen to fit the remaining memory.
// Device bytes the merge's per-brick item lists will take beyond what they hold now, for
// `cols` orientation columns: about 2.8 list entries (int item id + 64-bit pixel mask) per
// (orientation, 64-pixel patch) at the benchmark geometries, budgeted at 4. The chunk buffers
// are sized on what is left.
size_t brick_list_extra(const std::vector<emc_engine::BrickChunk> &chunks, size_t cols, int num_pix) {
    const size_t want = cols * (size_t)((num_pix + 63) / 64) * 4 * (sizeof(int) + sizeof(unsigned long long));
    size_t held = 0;
    for (const auto &bc : chunks) held += bc.list.cap * sizeof(int) + bc.lmask.cap * sizeof(unsigned long long);
    return want > held ? want - held : 0;
}

int ensure_work(emc_engine *e) {
    int rc;
    const size_t N = e->num_data, Rp = e->Rp, P = e->num_pix;
    // the sparse-subset rows are not needed by the plain path
    e->Rsub.release();
    e->RF.release();
    if ((rc = e->R.alloc(N * Rp))) return rc;
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    const size_t reserve = ((size_t)4 << 30) + brick_list_extra(e->bricks, Rp, P);
    size_t budget = freeb > reserve ? freeb - reserve : 0;
    const size_t cap = (size_t)120 << 30;
    budget += e->logw.cap * sizeof(double);
    budget += e->tom_partial.cap * sizeof(double);
    budget += (e->lw_hi.cap + e->lw_lo.cap + e->p16.cap) * sizeof(_Float16);
    const bool mp = mp_mode(e);
    const bool g32 = e->precision == 3;
    budget += (e->logw32.cap + e->p32.cap) * sizeof(float);
    if (mp) {
        e->mp_ldp = round_up(P, emc::kMpGemmBK);
        e->mp_ldd = round_up(N, emc::kMpGemmBK);
        e->mp_rows_d = round_up(N, emc::kGemmBM);
        e->mp_rows_t = round_up(P, emc::kGemmBM);
        const size_t kbytes = (e->mp_rows_d * e->mp_ldp + e->mp_rows_t * e->mp_ldd) * sizeof(_Float16);
        const size_t have = (e->klik.cap + e->ktom.cap) * sizeof(_Float16);
        budget = budget + have > kbytes ? budget + have - kbytes : 0;
    }
    if (budget > cap) budget = cap;
    // per orientation column: log W / view, plus the partial views of the partitioned
    // tomogram (plain path) or the FP16 log W rows and posterior column (mixed path)
    const bool split = !mp && !g32 && e->lik_split > 1;
    budget += e->lik_part.cap * sizeof(double);
    if (!split) e->lik_part.release();  // before the chunk buffers grow into its room
    const size_t per_col = mp ? P * sizeof(double) + (mp_two(e) ? 2 : 1) * e->mp_ldp * sizeof(_Float16) +
                                    e->mp_ldd * sizeof(_Float16)
                              : P * sizeof(double) * (1 + (e->tomo_tile_mode == 2 ? e->tom_parts : 0)) +
                                    (g32 ? P * sizeof(float) + N * sizeof(float) : 0) +
                                    (split ? (size_t)e->lik_split * N * sizeof(double) : 0);
    size_t cmax = budget / per_col;
    if (e->max_chunk_cols > 0 && cmax > (size_t)e->max_chunk_cols) cmax = e->max_chunk_cols;
    cmax = cmax / emc::kRTile * emc::kRTile;
    if (cmax < (size_t)emc::kRTile) return fail(EMC_E_NOMEM, "no device memory for an orientation chunk");
    const int C = (int)(cmax < Rp ? cmax : Rp);
    if (C != e->C || !e->logw.p) {
        // exact sizes: the budget above counted these buffers as free
        e->bricks_valid = false;
        e->tom_partial.release();
        e->logw.release();
        e->lik_part.release();
        e->logw32.release();
        e->p32.release();
        if ((rc = e->logw.alloc(P * (size_t)C))) return rc;
        e->C = C;
    }
    if (mp) {
        e->tom_partial.release();
        const size_t rows_c = round_up((size_t)e->C, emc::kGemmBN);
        if (rows_c != e->mp_rows_c || !e->lw_hi.p || mp_two(e) != (e->lw_lo.p != nullptr)) {
            e->lw_hi.release();
            e->lw_lo.release();
            e->p16.release();
            if ((rc = e->lw_hi.alloc(rows_c * e->mp_ldp))) return rc;
            if (mp_two(e) && (rc = e->lw_lo.alloc(rows_c * e->mp_ldp))) return rc;
            if ((rc = e->p16.alloc(rows_c * e->mp_ldd))) return rc;
            // padding rows and columns stay zero: the kernels only write valid entries
            HIP_TRY(hipMemsetAsync(e->lw_hi.p, 0, e->lw_hi.n * sizeof(_Float16), e->stream));
            if (e->lw_lo.p) HIP_TRY(hipMemsetAsync(e->lw_lo.p, 0, e->lw_lo.n * sizeof(_Float16), e->stream));
            HIP_TRY(hipMemsetAsync(e->p16.p, 0, e->p16.n * sizeof(_Float16), e->stream));
            e->mp_rows_c = rows_c;
        }
        if (!e->mp_built) {
            if ((rc = e->klik.alloc(e->mp_rows_d * e->mp_ldp))) return rc;
            if ((rc = e->ktom.alloc(e->mp_rows_t * e->mp_ldd))) return rc;
            HIP_TRY(hipMemsetAsync(e->klik.p, 0, e->klik.n * sizeof(_Float16), e->stream));
            HIP_TRY(hipMemsetAsync(e->ktom.p, 0, e->ktom.n * sizeof(_Float16), e->stream));
            emc::launch_build_klik(e->stream, e->lik_off.p, e->lik_pix.p, e->lik_cnt.p, (int)N, e->mp_ldp, e->klik.p,
                                   mp_bf16(e));
            emc::launch_build_ktom(e->stream, e->tom_off.p, e->tom_frame.p, e->tom_cnt.p, (int)P, e->mp_ldd, e->ktom.p);
            HIP_TRY(hipGetLastError());
            e->mp_built = true;
        }
        if ((rc = e->center_t.alloc(P))) return rc;
        if ((rc = e->center_base.alloc(N))) return rc;
    } else {
        e->lw_hi.release();
        e->lw_lo.release();
        e->p16.release();
        e->klik.release();
        e->ktom.release();
        e->mp_built = false;
        e->mp_rows_c = 0;
        if (e->tomo_tile_mode == 2 && (rc = e->tom_partial.alloc((size_t)e->tom_parts * P * (size_t)C))) return rc;
    }
    if (split) {
        if (e->split_built != e->lik_split && (rc = build_lik_split(e, e->lik_split))) return rc;
        if ((rc = e->lik_part.alloc((size_t)e->lik_split * N * e->C))) return rc;
    }
    if (g32) {
        // 128-wide tiles: room for the chunk rounded up to 128 orientations
        const size_t C128 = round_up((size_t)e->C, 128);
        if ((rc = e->logw32.alloc(P * C128))) return rc;
        if ((rc = e->p32.alloc(N * C128))) return rc;
    } else {
        e->logw32.release();
        e->p32.release();
    }
    const size_t ngroups = (P + kPixPerGroup - 1) / kPixPerGroup;
    if ((rc = e->upart.alloc(ngroups * (size_t)e->C))) return rc;
    if ((rc = e->ubad.alloc(ngroups * (size_t)e->C))) return rc;
    if ((rc = e->rbad.alloc(Rp))) return rc;
    if ((rc = e->u.alloc(Rp))) return rc;
    if ((rc = e->psum_r.alloc(Rp))) return rc;
    const size_t nparts = (N + kFramesPerPsumPart - 1) / kFramesPerPsumPart;
    if ((rc = e->psum_part.alloc(nparts * (size_t)e->num_rot))) return rc;
    if ((rc = e->scalars.alloc(8))) return rc;
    if ((rc = e->model_part.alloc(kModelParts))) return rc;
    if ((rc = e->max_exp.alloc(N))) return rc;
    if ((rc = e->p_norm.alloc(N))) return rc;
    if ((rc = e->likelihood.alloc(N))) return rc;
    if ((rc = e->info.alloc(N))) return rc;
    if ((rc = e->psum_d.alloc(N))) return rc;
    if ((rc = e->rmax.alloc(N))) return rc;
    return 0;
}

// One iteration over the refinement subsets only (pairs of partial frames and slot tiles, and
// every slot tile of the full frames). Same outputs as emc_maximize's plain path;
// per-orientation outputs are mapped back to the caller's orientation order.
int maximize_subset(emc_engine *e, const emc_iter_opts *opts, emc_iter_stats *stats) {
    int rc;
    hipStream_t s = e->stream;
    const int N = e->num_data, Rn = e->num_rot, P = e->num_pix, S = e->num_slots, T = e->num_tiles;
    const int nF = e->num_full;
    const size_t vol = (size_t)e->size * e->size * e->size;
    if (!e->fsort_valid && (rc = build_fsort(e))) return rc;
    if (nF > 0 && !e->full_lists_valid && (rc = build_full_lists(e))) return rc;
    const int nparts = e->full_parts;
    // memory: the plain path's rows are not needed here
    e->R.release();
    const size_t rsub_need = (size_t)std::max<long>(e->num_pairs, 1) * emc::kRTile;
    const size_t rf_need = nF > 0 ? (size_t)nF * T * emc::kRTile : 0;
    if (e->Rsub.cap < rsub_need || e->RF.cap < rf_need) {
        // the rows grow: hand the orientation-chunk buffers' room back first, so that the
        // chunk width below is sized on what the new rows leave free; a resident log W (the
        // second pass of a judged iteration) stays when the rows fit beside it
        size_t freeb = 0, totalb = 0;
        HIP_TRY(hipMemGetInfo(&freeb, &totalb));
        const size_t grow = (e->Rsub.cap < rsub_need ? rsub_need : 0) * sizeof(double) +
                            (e->RF.cap < rf_need ? rf_need : 0) * sizeof(double);
        const bool keep_logw = e->logw_resident && freeb > grow + ((size_t)8 << 30);
        if (!keep_logw) {
            e->logw.release();
            e->C = 0;
            e->bricks_valid = false;
            e->logw32.release();
            e->logw_resident = false;
        }
        e->tom_partial.release();
        e->tom_partial32.release();
        e->p32.release();
    }
    if ((rc = e->Rsub.alloc((size_t)std::max<long>(e->num_pairs, 1) * emc::kRTile))) return rc;
    if (nF > 0) {
        if ((rc = e->RF.alloc((size_t)nF * T * emc::kRTile))) return rc;
        const size_t pp = (nF + kFramesPerPsumPart - 1) / kFramesPerPsumPart;
        if ((rc = e->psum_full_part.alloc(pp * (size_t)S))) return rc;
        if ((rc = e->psum_full.alloc(S))) return rc;
        if ((rc = e->scale_full.alloc(nF))) return rc;
        if ((rc = e->blacklist_full.alloc(nF))) return rc;
    }  // with no full frames RF keeps its room: the next set of lists may have them again
    const bool g32 = e->precision == 3;
    if (g32 && (rc = e->Rsub32.alloc((size_t)std::max<long>(e->num_pairs, 1) * emc::kRTile))) return rc;
    if ((rc = e->sig_pair.alloc((size_t)std::max<long>(e->num_pairs, 1)))) return rc;
    if ((rc = e->sig_frame.alloc((size_t)std::max<long>(e->num_pairs, 1)))) return rc;
    if ((rc = e->pair_flag.alloc((size_t)std::max<long>(e->num_pairs, 1)))) return rc;
    if ((rc = e->sig_cnt.alloc((size_t)std::max(T, 1)))) return rc;
    if (nF > 0) {
        const size_t rows = (size_t)nF * T;
        if ((rc = e->full_flag.alloc(rows))) return rc;
        if ((rc = e->sigF_row.alloc(rows))) return rc;
        if ((rc = e->sigF_frame.alloc(rows))) return rc;
        if ((rc = e->sigF_cnt.alloc((size_t)T))) return rc;
        if ((rc = e->sigF_total.alloc(1))) return rc;
    }
    if (g32 && (rc = ensure_model32(e))) return rc;
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    const size_t reserve = ((size_t)4 << 30) + brick_list_extra(e->slot_bricks, S, P);
    // free memory plus the chunk buffers sized below (the full frames' ones only when there are
    // full frames: else they keep their room for the next lists that have some)
    size_t budget = (freeb > reserve ? freeb - reserve : 0) + e->logw.cap * sizeof(double) +
                    e->logw32.cap * sizeof(float) +
                    (nF > 0 ? e->tom_partial.cap * sizeof(double) + (e->p32.cap + e->tom_partial32.cap) * sizeof(float)
                            : 0);
    budget = std::min(budget, (size_t)120 << 30);
    // per slot column: log W / view and the full frames' partial views (FP64; with FP32 gathers
    // all FP32, log W and the view sharing one buffer, plus the full frames' FP32 posterior
    // column), and the view's nonzero masks
    const size_t parts_col = nF > 0 ? (size_t)nparts : 0;
    const size_t per_col = (g32 ? (size_t)P * sizeof(float) * (1 + parts_col) + (size_t)nF * sizeof(float)
                                : (size_t)P * sizeof(double) * (1 + parts_col)) +
                           (size_t)((P + 63) / 64) * sizeof(unsigned long long);
    size_t cmax = budget / per_col;
    if (e->max_chunk_cols > 0) cmax = std::min(cmax, (size_t)e->max_chunk_cols);
    cmax = cmax / emc::kRTile * emc::kRTile;
    if (cmax < (size_t)emc::kRTile) return fail(EMC_E_NOMEM, "no device memory for an orientation chunk");
    int C = (int)std::min(cmax, (size_t)S);
    // keep the chunk width of the previous iteration while it still fits: the merge's brick
    // lists are built per chunk, and lists that alternate between fewer and more full frames
    // (proposals judged in two passes) would otherwise rebuild them every time
    if (e->slot_C > 0 && e->slot_C < C && e->slot_bricks_valid) C = e->slot_C;
    if (C != e->slot_C) e->slot_bricks_valid = false;
    e->slot_C = C;
    // the chunk buffers at the sizes the budget assumed: every buffer larger than that (a wider
    // chunk before) is released before any is allocated, since the budget counted all of it as
    // free; a size of 0 (no full frames) leaves the buffer as it is
    auto shrink = [](auto &buf, size_t n) {
        if (n > 0 && buf.cap > n) buf.release();
    };
    auto fit = [](auto &buf, size_t n) -> int { return n > 0 ? buf.alloc(n) : 0; };
    const size_t C128 = round_up((size_t)C, 128);  // FP32 log W: 128-wide tiles
    if (g32 ? e->logw.cap > 0 : e->logw.cap != (size_t)P * C) {
        // the FP64 log W / view buffer: none with FP32 gathers (log W and the view are FP32)
        e->logw.release();
        e->C = 0;
        e->bricks_valid = false;
        if (!g32 && (rc = e->logw.alloc((size_t)P * C))) return rc;
    }
    if (g32 && e->logw32.cap > (size_t)P * C128) e->logw_resident = false;  // released below
    if (g32) {
        const size_t n_part = nF > 0 ? (size_t)nparts * P * (size_t)C : 0, n_logw = (size_t)P * C128,
                     n_p32 = nF > 0 ? (size_t)nF * C128 : 0;
        e->tom_partial.release();
        shrink(e->logw32, n_logw);
        shrink(e->tom_partial32, n_part);
        shrink(e->p32, n_p32);
        if ((rc = fit(e->logw32, n_logw)) || (rc = fit(e->tom_partial32, n_part)) || (rc = fit(e->p32, n_p32)))
            return rc;
    } else {
        e->tom_partial32.release();
        e->logw32.release();
        e->p32.release();
        const size_t n_part = nF > 0 ? (size_t)nparts * P * (size_t)C : 0;
        shrink(e->tom_partial, n_part);
        if ((rc = fit(e->tom_partial, n_part))) return rc;
    }
    const int ngroups = (P + kPixPerGroup - 1) / kPixPerGroup;
    if ((rc = e->upart.alloc((size_t)ngroups * C))) return rc;
    if ((rc = e->ubad.alloc((size_t)ngroups * C))) return rc;
    if ((rc = e->nzmask.alloc((size_t)C * ((P + 63) / 64)))) return rc;
    if ((rc = e->nz_count.alloc(1))) return rc;
    if ((rc = e->u_slot.alloc(S))) return rc;
    if ((rc = e->rbad_slot.alloc(S))) return rc;
    if ((rc = e->psum_slot.alloc(S))) return rc;
    if ((rc = e->u.alloc(e->Rp))) return rc;
    if ((rc = e->rbad.alloc(e->Rp))) return rc;
    if ((rc = e->psum_r.alloc(e->Rp))) return rc;
    if ((rc = e->scalars.alloc(8))) return rc;
    if ((rc = e->model_part.alloc(kModelParts))) return rc;
    if ((rc = e->max_exp.alloc(N))) return rc;
    if ((rc = e->p_norm.alloc(N))) return rc;
    if ((rc = e->likelihood.alloc(N))) return rc;
    if ((rc = e->info.alloc(N))) return rc;
    if ((rc = e->psum_d.alloc(N))) return rc;
    if ((rc = e->rmax.alloc(N))) return rc;
    if ((rc = e->d_count.alloc(1))) return rc;
    const emc::FullRows full{nF > 0 ? e->RF.p : nullptr, nF > 0 ? e->ffull.p : nullptr, nF, T};

    const int use_scale = opts->need_scaling && (opts->iteration > 1 || opts->known_scale);
    const int want_psum_d = opts->need_scaling && opts->update_scale;
    const bool graph = e->graph_mode != 0;
    if (graph && (e->estep_only || e->mstep_only))
        return fail(EMC_E_STATE, "emc_maximize: the graph option runs whole iterations only");
    float ms[EMC_NUM_STAGES] = {};
    bool capturing = false;  // enqueue into a graph: no events, no host synchronisation
    auto timed = [&](int stage, auto &&fn) -> int {
        if (capturing) return fn();
        HIP_TRY(hipEventRecord(e->ev[0], s));
        int r = fn();
        if (r) return r;
        HIP_TRY(hipEventRecord(e->ev[1], s));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventSynchronize(e->ev[1]));
        float a = 0.f;
        HIP_TRY(hipEventElapsedTime(&a, e->ev[0], e->ev[1]));
        ms[stage] += a;
        return 0;
    };
    const uint8_t *dense = e->any_dense ? e->dense.p : nullptr;
    // expectation step: expand, likelihood, per-frame maximum, normaliser and posterior, psum
    const bool fused = e->fused != 0;
    if (fused && (rc = e->fused_counter.alloc(1))) return rc;
    // log W of every slot still in place from the previous expectation step (one chunk)
    const void *logw_buf = g32 ? (const void *)e->logw32.p : (const void *)e->logw.p;
    const bool one_chunk = C >= S;
    const bool resident = one_chunk && !fused && !graph && e->logw_resident && e->res_logw == logw_buf &&
                          e->res_u == e->u_slot.p && e->res_rbad == e->rbad_slot.p && e->res_S == S;
    e->logw_resident = false;  // set again below once the expand has run
    auto estep_work = [&]() -> int {
            for (int s0 = 0; s0 < S; s0 += C) {
                const int Cc = std::min(C, S - s0);
                if (fused) {
                    // expand and pair likelihood in one persistent launch (timed as likelihood)
                    if ((rc = timed(1, [&] {
                             const int t0 = s0 / emc::kRTile;
                             if (g32)
                                 emc::launch_fused_expand_likelihood(
                                     s, e->rots_slot.p, e->model32.p, e->size, e->qvals.p, e->corr.p, e->mask.p, P, s0,
                                     Cc, kPixPerGroup, e->logw32.p, e->upart.p, e->ubad.p, e->slot_rot.p, e->lik_off.p,
                                     e->lik_pix.p, e->lik_cnt.p, e->pair_frame.p, e->pair_tile.p, e->tile_off.p,
                                     e->tile_pair.p, t0, e->fused_counter.p, e->Rsub.p);
                             else
                                 emc::launch_fused_expand_likelihood(
                                     s, e->rots_slot.p, e->model1.p, e->size, e->qvals.p, e->corr.p, e->mask.p, P, s0,
                                     Cc, kPixPerGroup, e->logw.p, e->upart.p, e->ubad.p, e->slot_rot.p, e->lik_off.p,
                                     e->lik_pix.p, e->lik_cnt.p, e->pair_frame.p, e->pair_tile.p, e->tile_off.p,
                                     e->tile_pair.p, t0, e->fused_counter.p, e->Rsub.p);
                             emc::launch_u_reduce(s, e->upart.p, e->ubad.p, ngroups, Cc, s0, S, e->u_slot.p,
                                                  e->rbad_slot.p);
                             if (nF > 0) {
                                 if (g32)
                                     emc::launch_likelihood_mapped(s, e->logw32.p, Cc, P, e->lik_off.p, e->lik_pix.p,
                                                                   e->lik_cnt.p, e->blacklist.p, e->fidx.p, nF, e->RF.p,
                                                                   s0);
                                 else
                                     emc::launch_likelihood_mapped(s, e->logw.p, Cc, P, e->lik_off.p, e->lik_pix.p,
                                                                   e->lik_cnt.p, e->blacklist.p, e->fidx.p, nF, e->RF.p,
                                                                   s0);
                             }
                             return 0;
                         })))
                        return rc;
                    continue;
                }
                if (!resident && (rc = timed(0, [&] {
                                      if (g32)
                                          emc::launch_expand(s, e->rots_slot.p, e->model32.p, e->size, e->qvals.p,
                                                             e->corr.p, e->mask.p, P, s0, Cc, kPixPerGroup,
                                                             e->logw32.p, e->upart.p, e->ubad.p, e->slot_rot.p);
                                      else
                                          emc::launch_expand(s, e->rots_slot.p, e->model1.p, e->size, e->qvals.p,
                                                             e->corr.p, e->mask.p, P, s0, Cc, kPixPerGroup,
                                                             e->logw.p, e->upart.p, e->ubad.p, e->slot_rot.p);
                                      emc::launch_u_reduce(s, e->upart.p, e->ubad.p, ngroups, Cc, s0, S, e->u_slot.p,
                                                           e->rbad_slot.p);
                                      return 0;
                                  })))
                    return rc;
                if (one_chunk && !capturing) {
                    e->logw_resident = true;
                    e->res_logw = logw_buf;
                    e->res_u = e->u_slot.p;
                    e->res_rbad = e->rbad_slot.p;
                    e->res_S = S;
                }
                if ((rc = timed(1, [&] {
                         const int t0 = s0 / emc::kRTile, t1 = (s0 + Cc) / emc::kRTile;
                         if (g32) {
                             emc::launch_pair_likelihood(s, e->logw32.p, P, e->lik_off.p, e->lik_pix.p, e->lik_cnt.p,
                                                         e->pair_frame.p, e->pair_tile.p, e->tile_off.p, e->tile_pair.p,
                                                         e->num_pairs, t0, t1, e->Rsub.p);
                             if (nF > 0)
                                 emc::launch_likelihood_mapped(s, e->logw32.p, Cc, P, e->lik_off.p, e->lik_pix.p,
                                                               e->lik_cnt.p, e->blacklist.p, e->fidx.p, nF, e->RF.p, s0);
                         } else {
                             emc::launch_pair_likelihood(s, e->logw.p, P, e->lik_off.p, e->lik_pix.p, e->lik_cnt.p,
                                                         e->pair_frame.p, e->pair_tile.p, e->tile_off.p, e->tile_pair.p,
                                                         e->num_pairs, t0, t1, e->Rsub.p);
                             if (nF > 0)
                                 emc::launch_likelihood_mapped(s, e->logw.p, Cc, P, e->lik_off.p, e->lik_pix.p,
                                                               e->lik_cnt.p, e->blacklist.p, e->fidx.p, nF, e->RF.p, s0);
                         }
                         return 0;
                     })))
                    return rc;
            }
            if ((rc = timed(2, [&] {
                     emc::launch_rescale(s, e->quats_slot.p, e->u_slot.p, S, e->mean_count, e->scalars.p);
                     return 0;
                 })))
                return rc;
            if ((rc = timed(3, [&] {
                     emc::launch_pair_max(s, e->Rsub.p, e->pair_off.p, e->pair_tile.p, full, e->slot_rot.p,
                                          e->slot_pack_built == 1 ? e->slot_cell.p : nullptr, e->relbits.p, e->rel_words,
                                          e->logwr_slot.p, e->u_slot.p, e->rbad_slot.p, e->scale.p, use_scale, e->scalars.p,
                                          e->blacklist.p, dense, N, e->max_exp.p, e->rmax.p);
                     return 0;
                 })))
                return rc;
            if ((rc = timed(4, [&] {
                     emc::launch_pair_norm(s, e->Rsub.p, e->pair_off.p, e->pair_tile.p, full, e->slot_rot.p, e->beta.p,
                                           e->blacklist.p, e->max_exp.p, e->rmax.p, N, e->p_norm.p);
                     return 0;
                 })))
                return rc;
            if ((rc = timed(5, [&] {
                     emc::launch_pair_posterior(s, e->Rsub.p, e->pair_off.p, e->pair_tile.p, full, e->slot_rot.p,
                                                e->logwr_slot.p, e->u_slot.p, e->beta.p, e->sum_fact.p, e->blacklist.p,
                                                e->max_exp.p, e->rmax.p, e->p_norm.p, e->scalars.p, want_psum_d, N,
                                                e->likelihood.p, e->info.p, e->psum_d.p);
                     if (nF > 0) {
                         emc::launch_frame_gather(s, e->scale.p, e->fidx.p, nF, e->scale_full.p);
                         emc::launch_frame_gather(s, e->blacklist.p, e->fidx.p, nF, e->blacklist_full.p);
                         emc::launch_psum_r(s, e->RF.p, S, nF, kFramesPerPsumPart, e->scale_full.p, opts->need_scaling,
                                            e->blacklist_full.p, e->psum_full_part.p, e->psum_full.p);
                     }
                     emc::launch_pair_psum(s, e->Rsub.p, e->tile_off.p, e->tile_pair.p, e->pair_frame.p, T, e->scale.p,
                                           opts->need_scaling, nF > 0 ? e->psum_full.p : nullptr, e->psum_slot.p);
                     emc::launch_slot_gather(s, e->psum_slot.p, e->rot_slot.p, Rn, e->psum_r.p);
                     emc::launch_slot_gather(s, e->u_slot.p, e->rot_slot.p, Rn, e->u.p);
                     if (nF > 0)
                         emc::launch_sig_full_rows(s, e->RF.p, nF, T, e->fidx.p, e->full_flag.p, e->sigF_row.p,
                                                   e->sigF_frame.p, e->sigF_cnt.p, e->sigF_total.p);
                     if (g32) {
                         emc::launch_pair_to_f32(s, e->Rsub.p, e->num_pairs * emc::kRTile, e->Rsub32.p);
                         emc::launch_sig_pairs(s, e->Rsub32.p, e->num_pairs, e->tile_off.p, e->tile_pair.p,
                                               e->tile_frame.p, T, e->pair_flag.p, e->sig_pair.p, e->sig_frame.p,
                                               e->sig_cnt.p);
                     } else {
                         emc::launch_sig_pairs(s, e->Rsub.p, e->num_pairs, e->tile_off.p, e->tile_pair.p,
                                               e->tile_frame.p, T, e->pair_flag.p, e->sig_pair.p, e->sig_frame.p,
                                               e->sig_cnt.p);
                     }
                     return 0;
                 })))
                return rc;
        return 0;
    };
    // all-orientation weights (they depend on the geometry only), computed once
    auto cache_weights = [&]() -> int {
        if (e->weight_all_valid) return 0;
        int r2;
        if ((r2 = e->weight_all.alloc(vol))) return r2;
        emc::launch_fill(s, e->weight.p, vol, 0.);
        for (const auto &bc : e->slot_bricks)
            merge_chunk(e, bc, e->rots_slot.p, S, nullptr, e->psum_slot.p, true, false, true);
        HIP_TRY(hipMemcpyAsync(e->weight_all.p, e->weight.p, vol * sizeof(double), hipMemcpyDeviceToDevice, s));
        e->weight_all_valid = true;
        return 0;
    };
    // M-step: tomogram and merge chunk by chunk. device_weights: the choice between the cached
    // weights and merging them is made on the device (no read-back), as a graph needs
    auto mstep_work = [&](bool device_weights) -> int {
        e->logw_resident = false;  // the views go into the log W buffer
        bool cached_weight = false;
        if (device_weights) {
            emc::launch_count_unmerged(s, e->psum_r.p, Rn, e->d_count.p);
            emc::launch_copy_if_zero(s, e->weight.p, e->weight_all.p, vol, e->d_count.p);
            e->last_unmerged = -1;
        } else {
            // merge: bricks over the slot order (padding slots make no items)
            if (!e->slot_bricks_valid) {
                if ((rc = timed(8, [&] {
                         return build_bricks(e, e->slot_bricks, e->rots_slot.p, S, C, S, e->slot_rot.p, 1);
                     })))
                    return rc;
                e->slot_bricks_valid = true;
            }
            int unmerged = 0;
            emc::launch_count_unmerged(s, e->psum_r.p, Rn, e->d_count.p);
            HIP_TRY(hipMemcpyAsync(&unmerged, e->d_count.p, sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            e->last_unmerged = unmerged;
            cached_weight = unmerged == 0;
            if (cached_weight && (rc = timed(8, [&] {
                                      if (!e->weight_all_valid) return cache_weights();  // leaves them in weight too
                                      HIP_TRY(hipMemcpyAsync(e->weight.p, e->weight_all.p, vol * sizeof(double),
                                                             hipMemcpyDeviceToDevice, s));
                                      return 0;
                                  })))
                return rc;
        }
        // the full frames' significant rows go to the pair tomogram directly when they average
        // at most sig_full_rows a tile (its blocks walk every row of their tile); else
        // their tomogram is the plain path's, into partial views (also inside a graph, which
        // cannot read the count back)
        bool sig_full = false;
        if (nF > 0 && !device_weights) {
            long tot = 0;
            HIP_TRY(hipMemcpyAsync(&tot, e->sigF_total.p, sizeof(long), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            sig_full = tot <= (long)e->sig_full_rows * T;
        }
        const emc::SigFullRows frows =
            sig_full ? emc::SigFullRows{e->RF.p, nF, e->sigF_row.p, e->sigF_frame.p, e->sigF_cnt.p}
                     : emc::SigFullRows{};
        const bool parts = nF > 0 && !sig_full;
        e->nz_counted = false;
        emc::launch_zero_u64(s, e->nz_count.p);
        // per-(tile, range) buckets of the significant pairs' photons (eager only: their size is
        // read back; without room for them the blocks walk their tiles' pairs instead)
        emc::PairBuckets bkv;
        if (e->tomo_buckets && !device_weights && !sig_full && e->nranges <= kMaxBucketRanges) {
            const int nr = e->nranges;
            if ((rc = timed(7, [&]() -> int {
                     int r2;
                     if ((r2 = e->bk_loc.alloc((size_t)T * (nr + 1))) || (r2 = e->bk_tot.alloc((size_t)T)) ||
                         (r2 = e->bk_base.alloc((size_t)T + 1)))
                         return r2;
                     emc::launch_pair_buckets(s, T, e->tile_off.p, e->sig_pair.p, e->sig_frame.p, e->sig_cnt.p,
                                              e->fr_off.p, e->fr_rg.p, nr, e->bk_loc.p, e->bk_tot.p, e->bk_base.p,
                                              nullptr, nullptr, 0);
                     HIP_TRY(hipGetLastError());
                     long total = 0;
                     HIP_TRY(hipMemcpyAsync(&total, e->bk_base.p + T, sizeof(long), hipMemcpyDeviceToHost, s));
                     HIP_TRY(hipStreamSynchronize(s));
                     const size_t n = (size_t)std::max(total, 1L);
                     if (e->bk_pair.alloc(n) || e->bk_fr.alloc(n)) {
                         e->bk_pair.release();  // no room: the blocks walk their tiles' pairs
                         e->bk_fr.release();
                         return 0;
                     }
                     emc::launch_pair_buckets(s, T, e->tile_off.p, e->sig_pair.p, e->sig_frame.p, e->sig_cnt.p,
                                              e->fr_off.p, e->fr_rg.p, nr, e->bk_loc.p, e->bk_tot.p, e->bk_base.p,
                                              e->bk_pair.p, e->bk_fr.p, 1);
                     HIP_TRY(hipGetLastError());
                     bkv = emc::PairBuckets{e->bk_base.p, e->bk_loc.p, e->bk_pair.p, e->bk_fr.p, e->fr_q0.p, e->fr_q1.p};
                     return 0;
                 })))
                return rc;
        }
        for (size_t k = 0; k < e->slot_bricks.size(); ++k) {
            const auto &bc = e->slot_bricks[k];
            const int t0 = bc.r0 / emc::kRTile, t1 = (bc.r0 + bc.Cc) / emc::kRTile;
            if (parts && (rc = timed(6, [&] {
                               if (g32) {
                                   emc::launch_to_f32(s, e->RF.p, nF, bc.r0, bc.Cc, e->p32.p);
                                   emc::launch_tomogram_parts_f32(s, e->p32.p, nF, bc.Cc, e->ftom_off2.p, nparts,
                                                                  e->ftom_frame.p, e->ftom_cnt.p, P,
                                                                  e->tom_partial32.p);
                               } else {
                                   emc::launch_tomogram_parts(s, e->RF.p, nF, bc.r0, bc.Cc, e->ftom_off2.p, nparts,
                                                              e->ftom_frame.p, e->ftom_cnt.p, nullptr, S, P,
                                                              e->tom_partial.p, e->logw.p, 0, 16, 0);
                               }
                               return 0;
                           })))
                return rc;
            // partial frames' pairs added to the full frames' partial sums, divided by psum;
            // with nonzero masks when the model merges alone (only nonzero views are stored)
            unsigned long long *nz = g32 || device_weights || cached_weight ? e->nzmask.p : nullptr;
            if ((rc = timed(7, [&] {
                     // FP32 gathers: an FP32 view (in the FP32 log W buffer, free in the M-step)
                     if (g32)
                         emc::launch_pair_tomogram(s, e->Rsub32.p, e->tile_off.p, e->sig_pair.p, e->sig_frame.p,
                                                   e->sig_cnt.p, e->fsort_off.p, e->fsort_roff.p, e->nranges,
                                                   e->fsort_pix.p, e->fsort_cnt.p, e->psum_slot.p, t0, t1 - t0, P,
                                                   parts ? e->tom_partial32.p : nullptr, nparts, e->logw32.p, nz,
                                                   frows, bkv);
                     else
                         emc::launch_pair_tomogram(s, e->Rsub.p, e->tile_off.p, e->sig_pair.p, e->sig_frame.p,
                                                   e->sig_cnt.p, e->fsort_off.p, e->fsort_roff.p, e->nranges,
                                                   e->fsort_pix.p, e->fsort_cnt.p, e->psum_slot.p, t0, t1 - t0, P,
                                                   parts ? e->tom_partial.p : nullptr, nparts, e->logw.p, nz,
                                                   frows, bkv);
                     return 0;
                 })))
                return rc;
            if (nz) {
                emc::launch_popcount(s, nz, (size_t)bc.Cc * ((P + 63) / 64), e->nz_count.p);
                e->nz_counted = true;
            }
            if ((rc = timed(8, [&] {
                     const float *v32 = g32 ? e->logw32.p : nullptr;
                     if (!device_weights && !g32)
                         return merge_chunk(e, bc, e->rots_slot.p, S, e->logw.p, e->psum_slot.p, false, true,
                                            !cached_weight, nullptr, nullptr, nz);
                     // weights of the orientations that merge, only when some orientation does not
                     // (device: gate = the count of those that do not), then the model from every view
                     if (device_weights)
                         merge_chunk(e, bc, e->rots_slot.p, S, nullptr, e->psum_slot.p, false, false, true,
                                     e->d_count.p);
                     else if (!cached_weight)
                         merge_chunk(e, bc, e->rots_slot.p, S, nullptr, e->psum_slot.p, false, false, true);
                     return merge_chunk(e, bc, e->rots_slot.p, S, e->logw.p, e->psum_slot.p, false, true, false,
                                        nullptr, v32, nz);
                 })))
                return rc;
        }
        emc::launch_frame_stats(s, e->likelihood.p, e->info.p, e->blacklist.p, N, N - e->num_blacklist, e->scalars.p);
        if (want_psum_d) emc::launch_scale(s, e->fcounts.p, e->psum_d.p, e->blacklist.p, N, e->scale.p);
        HIP_TRY(hipGetLastError());
        return 0;
    };
    auto finish = [&]() -> int {
        double host_scalars[3];
        HIP_TRY(hipMemcpyAsync(host_scalars, e->scalars.p, sizeof host_scalars, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        e->have_results = true;
        e->subset_ran = true;
        if (stats) {
            stats->rescale = host_scalars[0];
            stats->avg_likelihood = host_scalars[1];
            stats->mutual_info = host_scalars[2];
            for (int i = 0; i < EMC_NUM_STAGES; ++i) stats->stage_ms[i] = ms[i];
            stats->num_chunks = (int)e->slot_bricks.size();
            stats->view_nonzero = 0.;
            if (e->nz_counted) {
                unsigned long long nnz = 0;
                HIP_TRY(hipMemcpyAsync(&nnz, e->nz_count.p, sizeof nnz, hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
                stats->view_nonzero = (double)nnz;
            }
        }
        return 0;
    };

    if (graph) {
        // the set-up that depends on the geometry only stays outside the graph
        if (!e->slot_bricks_valid) {
            if ((rc = build_bricks(e, e->slot_bricks, e->rots_slot.p, S, C, S, e->slot_rot.p, 1))) return rc;
            e->slot_bricks_valid = true;
        }
        if ((rc = cache_weights())) return rc;
        // the graph holds the kernels' arguments: capture again when a buffer moved or a size,
        // list or option changed
        const std::vector<long> key = {(long)g_alloc_gen.load(), (long)e->layout_gen, use_scale, want_psum_d,
                                       opts->need_scaling, C, (long)e->slot_bricks.size(), (long)(intptr_t)e};
        if (!e->graph_exec || key != e->graph_key) {
            if (e->graph_exec) {
                (void)hipGraphExecDestroy(e->graph_exec);
                e->graph_exec = nullptr;
            }
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            capturing = true;
            emc::launch_fill(s, e->model2.p, vol, 0.);
            emc::launch_fill(s, e->weight.p, vol, 0.);
            rc = estep_work();
            if (!rc) rc = mstep_work(true);
            capturing = false;
            hipGraph_t g = nullptr;
            const hipError_t end = hipStreamEndCapture(s, &g);
            if (rc || end != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc ? rc : fail(EMC_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(end));
            }
            const hipError_t inst = hipGraphInstantiate(&e->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (inst != hipSuccess) {
                e->graph_exec = nullptr;
                return fail(EMC_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(inst));
            }
            e->graph_key = key;
        }
        HIP_TRY(hipGraphLaunch(e->graph_exec, s));
        return finish();
    }

    emc::launch_fill(s, e->model2.p, vol, 0.);
    emc::launch_fill(s, e->weight.p, vol, 0.);
    // mstep_only: the posterior, psum and per-frame sums of the estep_only iteration just run
    // on this state are still in place; only the M-step below remains
    const bool resume = e->mstep_only != 0;
    if (resume && (e->estep_only || e->estep_gen != e->gen))
        return fail(EMC_E_STATE, "emc_maximize: mstep_only needs an estep_only iteration of the current state");
    if (!resume && (rc = estep_work())) return rc;
    if (e->estep_only) {
        // expectation step only: per-frame statistics of the posterior over the subsets;
        // model2, inter_weight and the scale factors are left as they are
        emc::launch_frame_stats(s, e->likelihood.p, e->info.p, e->blacklist.p, N, N - e->num_blacklist,
                                e->scalars.p);
        HIP_TRY(hipGetLastError());
        double host_scalars[3];
        HIP_TRY(hipMemcpyAsync(host_scalars, e->scalars.p, sizeof host_scalars, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        e->have_results = true;
        e->subset_ran = true;
        e->estep_gen = e->gen;
        if (stats) {
            stats->rescale = host_scalars[0];
            stats->avg_likelihood = host_scalars[1];
            stats->mutual_info = host_scalars[2];
            for (int i = 0; i < EMC_NUM_STAGES; ++i) stats->stage_ms[i] = ms[i];
            stats->num_chunks = (int)((S + C - 1) / C);
            stats->view_nonzero = 0.;
        }
        return 0;
    }
    e->estep_gen = ~0ul;  // the M-step below consumes the expectation step
    if ((rc = mstep_work(false))) return rc;
    return finish();
}

// Kernel-variant options (emc_set_option / EMC_<NAME>): the field, its allowed range, and
// whether a change invalidates the per-pixel lists (built for a number of frame ranges).
struct EngineOption {
    const char *name;
    int emc_engine::*field;
    int lo, hi;
    bool lists;
};
const EngineOption kOptions[] = {
    {"lik_tiles", &emc_engine::lik_tiles, 1, 4, false},
    {"lik_unroll", &emc_engine::lik_unroll, 4, 16, false},
    {"lik_xcd", &emc_engine::lik_xcd, 0, 8, false},
    {"lik_split", &emc_engine::lik_split, 1, 8, false},
    {"tomo_tiles", &emc_engine::tomo_tiles, 1, 4, false},
    {"tomo_tile_mode", &emc_engine::tomo_tile_mode, 0, 2, false},
    {"tomo_parts", &emc_engine::tom_parts, 1, 64, true},
    {"tomo_width32", &emc_engine::tomo_width32, 0, 1, false},
    {"tomo_unroll", &emc_engine::tomo_unroll, 8, 16, false},
    {"tomo_seq", &emc_engine::tomo_seq, 0, 1, false},
    {"max_chunk_cols", &emc_engine::max_chunk_cols, 0, INT_MAX, false},
    {"merge_mode", &emc_engine::merge_mode, 0, 1, false},
    {"subset", &emc_engine::subset_mode, 0, 1, false},
    {"slot_pack", &emc_engine::slot_pack, 0, 1, false},
    {"sparse_rows", &emc_engine::sparse_rows, 0, 1 << 30, false},
    {"full_part_frames", &emc_engine::full_part_frames, 1, 1 << 30, true},
    {"estep_only", &emc_engine::estep_only, 0, 1, false},
    {"mstep_only", &emc_engine::mstep_only, 0, 1, false},
    {"graph", &emc_engine::graph_mode, 0, 1, false},
    {"fused", &emc_engine::fused, 0, 1, false},
    {"sig_full_rows", &emc_engine::sig_full_rows, 0, 1 << 20, false},
    {"unit_items", &emc_engine::unit_items, 64, 1 << 20, false},
    {"tomo_buckets", &emc_engine::tomo_buckets, 0, 1, false},
};

int set_option(emc_engine *e, const EngineOption &o, int value) {
    if (value < o.lo || value > o.hi)
        return fail(EMC_E_ARG, "option %s = %d outside [%d, %d]", o.name, value, o.lo, o.hi);
    if (e->*(o.field) != value && o.lists) {
        e->lists_dirty = true;
        e->full_lists_valid = false;
    }
    if (e->*(o.field) != value && o.field == &emc_engine::unit_items)
        e->bricks_valid = e->slot_bricks_valid = false;  // the merge's work units are cut again
    e->*(o.field) = value;
    return 0;
}

}  // namespace

extern "C" {

const char *emc_last_error(void) { return g_last_error.c_str(); }

int emc_create(int device, emc_engine **out) {
    if (!out) return fail(EMC_E_ARG, "null output pointer");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(EMC_E_ARG, "device %d not present (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    emc_engine *e = new emc_engine();
    e->device = device;
    hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (err != hipSuccess) {
        delete e;
        return fail(EMC_E_HIP, "hipStreamCreate failed: %s", hipGetErrorString(err));
    }
    for (auto &ev : e->ev) (void)hipEventCreate(&ev);
    // staging memory of emc_slice_gen_static on this device (resolved once per device)
    if (int rc = resolve_scratch(device)) {
        for (auto &ev : e->ev) (void)hipEventDestroy(ev);
        (void)hipStreamDestroy(e->stream);
        delete e;
        return rc;
    }
    // start values of the kernel-variant options from EMC_<NAME> (upper case)
    for (const EngineOption &o : kOptions) {
        std::string env = "EMC_";
        for (const char *c = o.name; *c; ++c) env += (char)toupper(*c);
        if (const char *v = getenv(env.c_str())) (void)set_option(e, o, atoi(v));
    }

    *out = e;
    return 0;
}

int emc_set_option(emc_engine *e, const char *name, int value) {
    if (!e || !name) return fail(EMC_E_ARG, "emc_set_option: null argument");
    // the two pass switches leave the iteration's state as it is; every other option may not
    if (strcmp(name, "estep_only") && strcmp(name, "mstep_only")) ++e->gen, ++e->layout_gen, e->logw_resident = false;
    for (const EngineOption &o : kOptions)
        if (!strcmp(o.name, name)) return set_option(e, o, value);
    return fail(EMC_E_ARG, "emc_set_option: unknown option '%s'", name);
}

int emc_get_option(emc_engine *e, const char *name, int *value) {
    if (!e || !name || !value) return fail(EMC_E_ARG, "emc_get_option: null argument");
    for (const EngineOption &o : kOptions)
        if (!strcmp(o.name, name)) {
            *value = e->*(o.field);
            return 0;
        }
    return fail(EMC_E_ARG, "emc_get_option: unknown option '%s'", name);
}

void emc_destroy(emc_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    for (DevBuf<double> *b : {&e->qvals, &e->corr, &e->quats, &e->rots, &e->logwr, &e->model1,
                              &e->model2, &e->weight, &e->lik_cnt, &e->tom_cnt, &e->scale,
                              &e->beta, &e->sum_fact, &e->logw, &e->upart, &e->u, &e->R,
                              &e->psum_r, &e->psum_part, &e->scalars, &e->model_part,
                              &e->max_exp, &e->p_norm, &e->likelihood, &e->info, &e->psum_d,
                              &e->one_rot, &e->one_view})
        b->release();
    e->mask.release();
    e->dense.release();
    e->ubad.release();
    e->rbad.release();
    e->d_perm.release();
    e->d_count.release();
    e->weight_all.release();

    e->blacklist.release();
    e->lik_off.release();
    e->tom_off.release();
    e->tom_off2.release();
    e->tom_partial.release();
    e->ft_off.release();
    e->fe_off.release();
    e->ft_frame.release();
    e->fe_code.release();
    e->lik_pix.release();
    e->tom_frame.release();
    e->fcounts.release();
    e->rmax.release();
    e->mass_off.release();
    e->mass_idx.release();
    e->mass_out.release();
    for (auto &ev : e->ev) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(e->stream);
    delete e;
}

int emc_synchronize(emc_engine *e) {
    if (!e) return fail(EMC_E_ARG, "null engine");
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int emc_set_detector(emc_engine *e, int num_pix, const double *qvals, const double *corr,
                     const uint8_t *raw_mask) {
    if (e) ++e->gen, ++e->layout_gen;
    if (e) e->logw_resident = false;
    if (!e || num_pix <= 0 || !qvals || !corr || !raw_mask) return fail(EMC_E_ARG, "emc_set_detector: bad argument");
    HIP_TRY(hipSetDevice(e->device));
    std::vector<int> perm = morton_order(num_pix, qvals);
    std::vector<int> inv(num_pix);
    std::vector<double> pq(3 * (size_t)num_pix), pc(num_pix);
    std::vector<uint8_t> pm(num_pix);
    for (int t = 0; t < num_pix; ++t) {
        const int o = perm[t];
        inv[o] = t;
        pq[3 * (size_t)t] = qvals[3 * (size_t)o];
        pq[3 * (size_t)t + 1] = qvals[3 * (size_t)o + 1];
        pq[3 * (size_t)t + 2] = qvals[3 * (size_t)o + 2];
        pc[t] = corr[o];
        pm[t] = raw_mask[o];
    }
    int rc;
    if ((rc = e->qvals.upload(pq.data(), pq.size(), e->stream))) return rc;
    if ((rc = e->corr.upload(pc.data(), pc.size(), e->stream))) return rc;
    if ((rc = e->mask.upload(pm.data(), pm.size(), e->stream))) return rc;
    if ((rc = e->d_perm.upload(perm.data(), perm.size(), e->stream))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->h_mask.assign(raw_mask, raw_mask + num_pix);
    e->with_bg = false;  // the background and powder follow the detector and frames
    e->perm.swap(perm);
    e->inv.swap(inv);
    e->num_pix = num_pix;
    e->lists_dirty = true;
    e->have_results = false;
    e->weight_all_valid = false;
    e->bricks_valid = false;
    e->slot_bricks_valid = false;  // the refinement subsets' brick lists hold pixels of the old detector
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

// Rotation matrices, log weights and quaternions of the combined orientations
// r = rotation * num_modes + mode from the caller's quaternions (h_quats).
static int upload_quats(emc_engine *e) {
    const int M = e->num_modes, nu = e->num_rot_user;
    if ((long)nu * M > INT32_MAX - emc::kRTile) return fail(EMC_E_ARG, "emc_set_quats: too many orientations");
    const int num_rot = nu * M;
    const int Rp = (num_rot + emc::kRTile - 1) / emc::kRTile * emc::kRTile;
    std::vector<double> rot((size_t)Rp * 9, 0.), lw(Rp, 0.), q((size_t)num_rot * 5);
    for (int r = 0; r < Rp; ++r) {
        if (r < num_rot) {
            const double *qu = &e->h_quats[5 * (size_t)(r / M)];
            make_rot(qu, &rot[9 * (size_t)r]);
            lw[r] = std::log(qu[4]);  // log(w_r) of the rotation, max_prob.c:125
            std::copy(qu, qu + 5, &q[5 * (size_t)r]);
        } else {
            rot[9 * (size_t)r] = rot[9 * (size_t)r + 4] = rot[9 * (size_t)r + 8] = 1.;
        }
    }
    int rc;
    if ((rc = e->quats.upload(q.data(), q.size(), e->stream))) return rc;
    if ((rc = e->rots.upload(rot.data(), rot.size(), e->stream))) return rc;
    if ((rc = e->logwr.upload(lw.data(), lw.size(), e->stream))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->num_rot = num_rot;
    e->Rp = Rp;
    return 0;
}

int emc_set_modes(emc_engine *e, int num_modes) {
    if (e) ++e->gen, ++e->layout_gen;
    if (e) e->model32_valid = false, e->logw_resident = false;  // model1 changes below
    if (!e || num_modes < 1 || num_modes > EMC_MAX_MODES)
        return fail(EMC_E_ARG, "emc_set_modes: num_modes must be 1..%d", EMC_MAX_MODES);
    if (num_modes == e->num_modes) return 0;
    HIP_TRY(hipSetDevice(e->device));
    e->num_modes = num_modes;
    e->size = 0;  // the model holds num_modes volumes: set it again
    e->have_results = false;
    e->weight_all_valid = false;
    e->bricks_valid = false;
    e->refine = false;
    e->h_slot_qmap.clear();
    e->slot_num_coarse = 0;
    if (e->num_rot_user > 0) return upload_quats(e);
    return 0;
}

int emc_set_quats(emc_engine *e, int num_rot, const double *quats) {
    if (e) ++e->gen, ++e->layout_gen;
    if (e) e->logw_resident = false;
    if (!e || num_rot <= 0 || !quats) return fail(EMC_E_ARG, "emc_set_quats: bad argument");
    HIP_TRY(hipSetDevice(e->device));
    e->h_quats.assign(quats, quats + 5 * (size_t)num_rot);
    e->num_rot_user = num_rot;
    int rc;
    if ((rc = upload_quats(e))) return rc;
    e->have_results = false;
    e->weight_all_valid = false;
    e->bricks_valid = false;
    e->refine = false;  // the orientation -> coarse cell map no longer applies
    e->h_slot_qmap.clear();
    e->slot_num_coarse = 0;
    return 0;
}

int emc_set_model(emc_engine *e, long size, const double *model1) {
    if (e) ++e->gen;  // the same buffer unless the size changes (then it is allocated anew)
    if (e) e->model32_valid = false, e->logw_resident = false;  // model1 changes below
    if (!e || size < 3 || !model1) return fail(EMC_E_ARG, "emc_set_model: bad argument");
    HIP_TRY(hipSetDevice(e->device));
    const size_t vol = (size_t)size * size * size * e->num_modes;  // every mode's volume
    int rc;
    if ((rc = e->model1.upload(model1, vol, e->stream))) return rc;
    if ((rc = e->model2.alloc(vol))) return rc;
    if ((rc = e->weight.alloc(vol))) return rc;
    emc::launch_fill(e->stream, e->model2.p, vol, 0.);
    emc::launch_fill(e->stream, e->weight.p, vol, 0.);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (size != e->size) e->weight_all_valid = false, e->bricks_valid = e->slot_bricks_valid = false;
    e->size = size;
    return 0;
}

int emc_copy_model(emc_engine *dst, const emc_engine *src) {
    if (dst) ++dst->gen;
    if (dst) dst->model32_valid = false, dst->logw_resident = false;  // model1 changes below
    if (!dst || !src) return fail(EMC_E_ARG, "emc_copy_model: null engine");
    if (!src->size || !src->model1.p) return fail(EMC_E_STATE, "emc_copy_model: the source holds no model");
    if (src->device != dst->device || src->num_modes != dst->num_modes)
        return fail(EMC_E_ARG, "emc_copy_model: engines on different devices or with different mode counts");
    HIP_TRY(hipSetDevice(dst->device));
    const long size = src->size;
    const size_t vol = (size_t)size * size * size * dst->num_modes;
    int rc;
    if ((rc = dst->model1.alloc(vol)) || (rc = dst->model2.alloc(vol)) || (rc = dst->weight.alloc(vol))) return rc;
    // the source's model is complete on its own stream before the copy on the destination's
    HIP_TRY(hipStreamSynchronize(src->stream));
    HIP_TRY(hipMemcpyAsync(dst->model1.p, src->model1.p, vol * sizeof(double), hipMemcpyDeviceToDevice, dst->stream));
    emc::launch_fill(dst->stream, dst->model2.p, vol, 0.);
    emc::launch_fill(dst->stream, dst->weight.p, vol, 0.);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(dst->stream));
    if (size != dst->size) dst->weight_all_valid = false, dst->bricks_valid = dst->slot_bricks_valid = false;
    dst->size = size;
    return 0;
}

int emc_set_background(emc_engine *e, const double *background, const double *bgscale) {
    if (e) ++e->gen, ++e->layout_gen;
    if (e) e->logw_resident = false;
    if (!e) return fail(EMC_E_ARG, "null engine");
    if (!background) {
        e->with_bg = false;
        return 0;
    }
    if (!bgscale) return fail(EMC_E_ARG, "emc_set_background: bgscale is NULL");
    if (!e->num_pix || !e->num_data) return fail(EMC_E_STATE, "emc_set_background: set the detector and frames first");
    for (uint8_t v : e->h_dense)
        if (v) return fail(EMC_E_STATE, "emc_set_background: the background-aware path takes sparse frames");
    HIP_TRY(hipSetDevice(e->device));
    const int P = e->num_pix, N = e->num_data;
    // calc_powder (iterate.c:119-179): mean photons per mask < 1 pixel over all frames
    std::vector<double> pw(P, 0.);
    for (int d = 0; d < N; ++d) {
        for (int k = 0; k < e->h_ones[d]; ++k) {
            const int t = e->h_place_ones[e->h_ones_accum[d] + k];
            if (e->h_mask[t] < 1) pw[t]++;
        }
        for (int k = 0; k < e->h_multi[d]; ++k) {
            const long at = e->h_multi_accum[d] + k;
            const int t = e->h_place_multi[at];
            if (e->h_mask[t] < 1) pw[t] += e->h_count_multi[at];
        }
    }
    std::vector<double> b(P), pe(P);
    for (int t = 0; t < P; ++t) {
        b[t] = background[e->perm[t]];
        pe[t] = pw[e->perm[t]] / N;
    }
    int rc;
    if ((rc = e->bg.upload(b.data(), P, e->stream))) return rc;
    if ((rc = e->powder.upload(pe.data(), P, e->stream))) return rc;
    if ((rc = e->bgscale.upload(bgscale, N, e->stream))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->with_bg = true;
    return 0;
}

int emc_set_frames_sparse(emc_engine *e, int num_data, const int *ones, const int *multi,
                          const long *ones_accum, const long *multi_accum, const int *place_ones,
                          const int *place_multi, const int *count_multi) {
    if (e) ++e->gen, ++e->layout_gen;
    if (!e || num_data <= 0 || !ones || !multi || !ones_accum || !multi_accum)
        return fail(EMC_E_ARG, "emc_set_frames_sparse: bad argument");
    long ones_total = ones_accum[num_data - 1] + ones[num_data - 1];
    long multi_total = multi_accum[num_data - 1] + multi[num_data - 1];
    if ((ones_total && !place_ones) || (multi_total && (!place_multi || !count_multi)))
        return fail(EMC_E_ARG, "emc_set_frames_sparse: missing photon arrays");
    for (int d = 0; d < num_data; ++d)
        if (ones[d] < 0 || multi[d] < 0 || ones_accum[d] < 0 || multi_accum[d] < 0 ||
            ones_accum[d] + ones[d] > ones_total || multi_accum[d] + multi[d] > multi_total)
            return fail(EMC_E_ARG, "emc_set_frames_sparse: inconsistent offsets at frame %d", d);
    e->h_ones.assign(ones, ones + num_data);
    e->h_multi.assign(multi, multi + num_data);
    e->h_ones_accum.assign(ones_accum, ones_accum + num_data);
    e->h_multi_accum.assign(multi_accum, multi_accum + num_data);
    e->h_place_ones.assign(place_ones, place_ones + ones_total);
    e->h_place_multi.assign(place_multi, place_multi + multi_total);
    e->h_count_multi.assign(count_multi, count_multi + multi_total);
    e->h_dense.assign(num_data, 0);
    e->num_data = num_data;
    e->lists_dirty = true;
    e->have_params = false;
    e->with_bg = false;  // the powder and per-frame background scales belong to the old frames
    e->have_results = false;
    e->refine = false;  // per-frame orientation sets belong to the previous frames
    return 0;
}

int emc_set_frames_dense(emc_engine *e, int num_data, int num_pix, const int *frames) {
    if (e) ++e->gen, ++e->layout_gen;
    if (!e || num_data <= 0 || num_pix <= 0 || !frames)
        return fail(EMC_E_ARG, "emc_set_frames_dense: bad argument");
    if (e->num_pix && num_pix != e->num_pix)
        return fail(EMC_E_ARG, "emc_set_frames_dense: %d pixels, detector has %d", num_pix, e->num_pix);
    // photon lists of the non-zero counts; the dense rule for zero counts is applied per
    // orientation in the likelihood (frame_max_kernel)
    std::vector<int> ones(num_data, 0), multi(num_data, 0), po, pm, cm;
    std::vector<long> oa(num_data, 0), ma(num_data, 0);
    for (int d = 0; d < num_data; ++d) {
        oa[d] = (long)po.size();
        ma[d] = (long)pm.size();
        const int *f = frames + (size_t)d * num_pix;
        for (int t = 0; t < num_pix; ++t) {
            if (f[t] < 0) return fail(EMC_E_ARG, "emc_set_frames_dense: negative count");
            if (f[t] == 1) { po.push_back(t); ++ones[d]; }
            else if (f[t] > 1) { pm.push_back(t); cm.push_back(f[t]); ++multi[d]; }
        }
    }
    if (po.empty()) po.push_back(0);
    if (pm.empty()) { pm.push_back(0); cm.push_back(0); }
    int rc = emc_set_frames_sparse(e, num_data, ones.data(), multi.data(), oa.data(), ma.data(),
                                   po.data(), pm.data(), cm.data());
    if (rc) return rc;
    e->h_dense.assign(num_data, 1);
    return 0;
}

int emc_set_refine(emc_engine *e, int num_coarse, const int *quat_mapping, const long *rel_off,
                   const int *rel_quat) {
    if (e) ++e->gen, ++e->layout_gen;
    if (!e) return fail(EMC_E_ARG, "null engine");
    if (!quat_mapping) {
        e->refine = false;
        return 0;
    }
    if (!e->num_rot || !e->num_data)
        return fail(EMC_E_STATE, "emc_set_refine: set quaternions and frames first");
    if (num_coarse <= 0 || !rel_off || !rel_quat) return fail(EMC_E_ARG, "emc_set_refine: bad argument");
    // the reference indexes quat_mapping (one entry per rotation) by the combined orientation
    // index (max_prob.c:111), which runs past it for several modes
    if (e->num_modes > 1) return fail(EMC_E_STATE, "emc_set_refine: refinement takes one mode");
    HIP_TRY(hipSetDevice(e->device));
    const int N = e->num_data, words = (num_coarse + 31) / 32;
    std::vector<int> qm(e->Rp, 0);
    for (int r = 0; r < e->num_rot; ++r) {
        if (quat_mapping[r] < 0 || quat_mapping[r] >= num_coarse)
            return fail(EMC_E_ARG, "emc_set_refine: quat_mapping[%d] = %d out of range", r, quat_mapping[r]);
        qm[r] = quat_mapping[r];
    }
    if (rel_off[0] != 0) return fail(EMC_E_ARG, "emc_set_refine: rel_off[0] must be 0");
    for (int d = 0; d < N; ++d)
        if (rel_off[d + 1] < rel_off[d]) return fail(EMC_E_ARG, "emc_set_refine: rel_off not ascending");
    int rc;
    // one bit row per frame, built by host threads over ranges of frames
    if ((rc = e->h_relbits.reserve((size_t)N * words))) return rc;
    uint32_t *bits = e->h_relbits.p;
    const int nth = host_threads(N, 2048);
    std::vector<long> bad(nth, -1);  // first entry out of range in each range of frames
    if ((rc = parallel_ranges(N, nth, [&](int pi, long d0, long d1) {
        std::memset(bits + (size_t)d0 * words, 0, (size_t)(d1 - d0) * words * sizeof(uint32_t));
        for (long d = d0; d < d1; ++d) {
            uint32_t *row = bits + (size_t)d * words;
            if (rel_off[d + 1] - rel_off[d] == 1 && rel_quat[rel_off[d]] == -1) {
                // the single entry -1: every coarse cell (the full search)
                for (int w = 0; w < words; ++w) row[w] = ~0u;
                if (num_coarse & 31) row[words - 1] = (1u << (num_coarse & 31)) - 1u;
                continue;
            }
            for (long k = rel_off[d]; k < rel_off[d + 1]; ++k) {
                const int c = rel_quat[k];
                if (c < 0 || c >= num_coarse) {
                    if (bad[pi] < 0) bad[pi] = k;
                    continue;
                }
                row[c >> 5] |= 1u << (c & 31);
            }
        }
    })))
        return rc;
    for (long k : bad)
        if (k >= 0) return fail(EMC_E_ARG, "emc_set_refine: rel_quat %d out of range", rel_quat[k]);
    if ((rc = e->qmap.upload(qm.data(), qm.size(), e->stream))) return rc;
    if ((rc = e->relbits.upload(bits, (size_t)N * words, e->stream))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->rel_words = words;
    e->refine = true;
    return build_subset(e, num_coarse, quat_mapping, rel_off, rel_quat);
}

int emc_cell_posterior(emc_engine *e, int num_coarse, const int *quat_mapping, float *out) {
    if (!e || !quat_mapping || !out) return fail(EMC_E_ARG, "emc_cell_posterior: null argument");
    if (num_coarse <= 0 || num_coarse > 6144)
        return fail(EMC_E_ARG, "emc_cell_posterior: num_coarse %d (1 to 6144)", num_coarse);
    if (!e->have_results) return fail(EMC_E_STATE, "emc_cell_posterior: no iteration has run");
    HIP_TRY(hipSetDevice(e->device));
    const int N = e->num_data, Rn = e->num_rot;
    for (int r = 0; r < Rn; ++r)
        if (quat_mapping[r] < 0 || quat_mapping[r] >= num_coarse)
            return fail(EMC_E_ARG, "emc_cell_posterior: quat_mapping[%d] = %d out of range", r, quat_mapping[r]);
    // frames in batches, so that the device copy stays small next to the engine's rows
    const int batch = std::max(1, std::min(N, (int)((size_t)64 << 20) / (num_coarse * (int)sizeof(float))));
    DevBuf<int> qm;
    DevBuf<float> dout;
    int rc;
    if ((rc = qm.upload(quat_mapping, Rn, e->stream))) return rc;
    if ((rc = dout.alloc((size_t)batch * num_coarse))) return rc;
    const bool plain = !e->subset_ran;
    const emc::FullRows full{e->num_full > 0 ? e->RF.p : nullptr, e->num_full > 0 ? e->ffull.p : nullptr,
                             e->num_full, e->num_tiles};
    for (int d0 = 0; d0 < N; d0 += batch) {
        const int cnt = std::min(batch, N - d0);
        emc::launch_cell_posterior(e->stream, plain ? e->R.p : e->Rsub.p, plain, N, Rn, e->pair_off.p,
                                   e->pair_tile.p, full, e->slot_rot.p, qm.p, num_coarse, d0, cnt, dout.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out + (size_t)d0 * num_coarse, dout.p, (size_t)cnt * num_coarse * sizeof(float),
                               hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return 0;
}

int emc_cell_topk(emc_engine *e, int num_coarse, const int *quat_mapping, int k, int *idx, float *mass) {
    if (!e || !quat_mapping || !idx || !mass) return fail(EMC_E_ARG, "emc_cell_topk: null argument");
    if (num_coarse <= 0 || num_coarse > 6144)
        return fail(EMC_E_ARG, "emc_cell_topk: num_coarse %d (1 to 6144)", num_coarse);
    if (k < 1 || k > 64 || k > num_coarse) return fail(EMC_E_ARG, "emc_cell_topk: k %d (1 to min(64, num_coarse))", k);
    if (!e->have_results) return fail(EMC_E_STATE, "emc_cell_topk: no iteration has run");
    HIP_TRY(hipSetDevice(e->device));
    const int N = e->num_data, Rn = e->num_rot;
    for (int r = 0; r < Rn; ++r)
        if (quat_mapping[r] < 0 || quat_mapping[r] >= num_coarse)
            return fail(EMC_E_ARG, "emc_cell_topk: quat_mapping[%d] = %d out of range", r, quat_mapping[r]);
    const int batch = std::max(1, std::min(N, (int)((size_t)64 << 20) / (num_coarse * (int)sizeof(float))));
    DevBuf<int> qm, didx;
    DevBuf<float> dout, dmass;
    int rc;
    if ((rc = qm.upload(quat_mapping, Rn, e->stream))) return rc;
    if ((rc = dout.alloc((size_t)batch * num_coarse))) return rc;
    if ((rc = didx.alloc((size_t)batch * k))) return rc;
    if ((rc = dmass.alloc((size_t)batch))) return rc;
    const bool plain = !e->subset_ran;
    const emc::FullRows full{e->num_full > 0 ? e->RF.p : nullptr, e->num_full > 0 ? e->ffull.p : nullptr,
                             e->num_full, e->num_tiles};
    for (int d0 = 0; d0 < N; d0 += batch) {
        const int cnt = std::min(batch, N - d0);
        emc::launch_cell_posterior(e->stream, plain ? e->R.p : e->Rsub.p, plain, N, Rn, e->pair_off.p,
                                   e->pair_tile.p, full, e->slot_rot.p, qm.p, num_coarse, d0, cnt, dout.p);
        emc::launch_prob_topk(e->stream, dout.p, num_coarse, cnt, num_coarse, k, didx.p, dmass.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(idx + (size_t)d0 * k, didx.p, (size_t)cnt * k * sizeof(int), hipMemcpyDeviceToHost,
                               e->stream));
        HIP_TRY(hipMemcpyAsync(mass + d0, dmass.p, (size_t)cnt * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return 0;
}

int emc_set_precision(emc_engine *e, int mode) {
    if (e) ++e->gen, ++e->layout_gen;
    if (e) e->logw_resident = false;
    if (!e) return fail(EMC_E_ARG, "null engine");
    if (mode < 0 || mode > 5) return fail(EMC_E_ARG, "emc_set_precision: mode %d (0 to 5)", mode);
    if (mode != e->precision) {
        e->precision = mode;
        e->mp_built = false;
    }
    return 0;
}

int emc_set_frame_params(emc_engine *e, const double *scale, const double *beta,
                         const uint8_t *blacklist, const int *fcounts, const double *sum_fact,
                         double mean_count, int num_blacklist) {
    if (e) ++e->gen, ++e->layout_gen;
    if (!e || !e->num_data) return fail(EMC_E_STATE, "emc_set_frame_params: frames not set");
    if (!scale || !beta || !blacklist || !fcounts || !sum_fact)
        return fail(EMC_E_ARG, "emc_set_frame_params: bad argument");
    if (num_blacklist < 0 || num_blacklist >= e->num_data)
        return fail(EMC_E_ARG, "emc_set_frame_params: num_blacklist %d out of range", num_blacklist);
    HIP_TRY(hipSetDevice(e->device));
    const size_t N = e->num_data;
    int rc;
    if ((rc = e->scale.upload(scale, N, e->stream))) return rc;
    if ((rc = e->beta.upload(beta, N, e->stream))) return rc;
    if ((rc = e->blacklist.upload(blacklist, N, e->stream))) return rc;
    e->h_blacklist.assign(blacklist, blacklist + N);
    if ((rc = e->fcounts.upload(fcounts, N, e->stream))) return rc;
    if ((rc = e->sum_fact.upload(sum_fact, N, e->stream))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->mean_count = mean_count;
    e->num_blacklist = num_blacklist;
    e->have_params = true;
    return 0;
}

int emc_set_beta(emc_engine *e, const double *beta) {
    if (e) ++e->gen;
    if (!e || !beta || !e->beta.p) return fail(EMC_E_STATE, "emc_set_beta: frame parameters not set");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemcpyAsync(e->beta.p, beta, e->beta.n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int emc_set_scale(emc_engine *e, const double *scale) {
    if (e) ++e->gen;
    if (!e || !scale || !e->scale.p) return fail(EMC_E_STATE, "emc_set_scale: frame parameters not set");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemcpyAsync(e->scale.p, scale, e->scale.n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

// Checks and per-cell inputs of the likelihood judge: the two engines' states, each cell's log
// fine orientation count and log mean fine weight over its centre's weight, and the frames it
// skips (full or blacklisted); leaves the coarse engine's results complete.
static int judge_setup(emc_engine *e, emc_engine *coarse, const char *who) {
    if (!e->refine || !e->subset_ran || !e->have_results)
        return fail(EMC_E_STATE, "%s: needs an iteration on refinement subsets", who);
    if (!coarse->have_results || coarse->subset_ran || coarse->refine)
        return fail(EMC_E_STATE, "%s: the coarse engine needs a plain iteration", who);
    const int N = e->num_data, Nc = e->slot_num_coarse;
    if (coarse->num_data != N || coarse->num_rot != Nc || coarse->num_modes != 1 || coarse->device != e->device)
        return fail(EMC_E_ARG, "%s: the coarse engine must hold the same frames and one orientation "
                               "per coarse cell (%d) on the same device", who, Nc);
    if ((int)e->h_slot_qmap.size() != e->num_rot || (int)coarse->h_quats.size() != 5 * Nc)
        return fail(EMC_E_STATE, "%s: orientation tables not set", who);
    HIP_TRY(hipSetDevice(e->device));
    std::vector<double> cnt(Nc, 0.), wsum(Nc, 0.), logn(Nc), lwr(Nc);
    for (int r = 0; r < e->num_rot; ++r) {
        cnt[e->h_slot_qmap[r]] += 1.;
        wsum[e->h_slot_qmap[r]] += e->h_quats[5 * (size_t)r + 4];
    }
    for (int c = 0; c < Nc; ++c) {
        const double wc = coarse->h_quats[5 * (size_t)c + 4];
        logn[c] = cnt[c] > 0. ? std::log(cnt[c]) : -INFINITY;
        lwr[c] = cnt[c] > 0. && wc > 0. ? std::log(wsum[c] / cnt[c] / wc) : 0.;
    }
    std::vector<uint8_t> skip(N);
    for (int d = 0; d < N; ++d) skip[d] = (uint8_t)(e->h_ffull[d] >= 0 || e->h_blacklist[d] != 0);
    int rc;
    if ((rc = e->judge_logn.upload(logn.data(), Nc, e->stream))) return rc;
    if ((rc = e->judge_lwr.upload(lwr.data(), Nc, e->stream))) return rc;
    if ((rc = e->judge_skip.upload(skip.data(), N, e->stream))) return rc;
    if ((rc = e->judge_out.alloc(N))) return rc;
    // the coarse engine's posterior, maximum and normaliser are complete on its own stream
    HIP_TRY(hipStreamSynchronize(coarse->stream));
    return 0;
}

int emc_judge_outside(emc_engine *e, emc_engine *coarse, double prob_floor, double *outside) {
    if (!e || !coarse || !outside) return fail(EMC_E_ARG, "emc_judge_outside: null argument");
    if (!(prob_floor >= 0.) || prob_floor >= 1.) return fail(EMC_E_ARG, "emc_judge_outside: prob_floor in [0, 1)");
    int rc;
    if ((rc = judge_setup(e, coarse, "emc_judge_outside"))) return rc;
    const int N = e->num_data, Nc = e->slot_num_coarse;
    emc::launch_judge_outside(e->stream, coarse->R.p, N, Nc, coarse->max_exp.p, coarse->p_norm.p, e->max_exp.p,
                              e->p_norm.p, e->beta.p, e->judge_skip.p, e->relbits.p, e->rel_words, e->judge_logn.p,
                              e->judge_lwr.p, prob_floor, e->judge_out.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(outside, e->judge_out.p, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int emc_judge_widen(emc_engine *e, emc_engine *coarse, double outside_max, double target, int max_add,
                    double *outside, int *num_add, long *total) {
    if (!e || !coarse || !outside || !num_add || !total) return fail(EMC_E_ARG, "emc_judge_widen: null argument");
    if (!(outside_max > 0.) || !(outside_max < 1.)) return fail(EMC_E_ARG, "emc_judge_widen: outside_max in (0, 1)");
    if (!(target > 0.) || !(target <= outside_max))
        return fail(EMC_E_ARG, "emc_judge_widen: target in (0, outside_max]");
    if (max_add < 1) return fail(EMC_E_ARG, "emc_judge_widen: max_add %d < 1", max_add);
    if (e->slot_num_coarse > emc::kWidenCells)
        return fail(EMC_E_ARG, "emc_judge_widen: %d coarse cells (at most %d)", e->slot_num_coarse, emc::kWidenCells);
    int rc;
    if ((rc = judge_setup(e, coarse, "emc_judge_widen"))) return rc;
    const int N = e->num_data, Nc = e->slot_num_coarse;
    max_add = std::min(max_add, Nc);
    hipStream_t s = e->stream;
    if ((rc = e->judge_nadd.alloc(N)) || (rc = e->judge_add.alloc((size_t)N * max_add))) return rc;
    emc::launch_judge_widen(s, coarse->R.p, N, Nc, coarse->max_exp.p, coarse->p_norm.p, e->max_exp.p, e->p_norm.p,
                            e->beta.p, e->judge_skip.p, e->relbits.p, e->rel_words, e->judge_logn.p, e->judge_lwr.p,
                            outside_max, target, max_add, e->judge_out.p, e->judge_nadd.p, e->judge_add.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(outside, e->judge_out.p, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(num_add, e->judge_nadd.p, (size_t)N * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // the added cells frame after frame, kept for emc_judge_widen_cells
    std::vector<long> off(N + 1, 0);
    for (int d = 0; d < N; ++d) off[d + 1] = off[d] + std::max(num_add[d], 0);
    *total = off[N];
    e->judge_cells_n = off[N];
    if ((rc = e->judge_off.upload(off.data(), off.size(), s)) ||
        (rc = e->judge_cells.alloc((size_t)std::max(off[N], 1L))))
        return rc;
    emc::launch_widen_compact(s, e->judge_add.p, e->judge_nadd.p, e->judge_off.p, N, max_add, e->judge_cells.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int emc_judge_widen_cells(emc_engine *e, int *cells, long capacity) {
    if (!e || (!cells && e->judge_cells_n > 0)) return fail(EMC_E_ARG, "emc_judge_widen_cells: null argument");
    if (e->judge_cells_n < 0) return fail(EMC_E_STATE, "emc_judge_widen_cells: no emc_judge_widen before");
    if (capacity < e->judge_cells_n)
        return fail(EMC_E_ARG, "emc_judge_widen_cells: capacity %ld < %ld cells", capacity, e->judge_cells_n);
    HIP_TRY(hipSetDevice(e->device));
    if (e->judge_cells_n > 0) {
        HIP_TRY(hipMemcpyAsync(cells, e->judge_cells.p, (size_t)e->judge_cells_n * sizeof(int), hipMemcpyDeviceToHost,
                               e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return 0;
}

int emc_clear_merge(emc_engine *e) {
    if (e) ++e->gen;
    if (!e || !e->size) return fail(EMC_E_STATE, "emc_clear_merge: model not set");
    HIP_TRY(hipSetDevice(e->device));
    const size_t vol = (size_t)e->size * e->size * e->size * e->num_modes;
    emc::launch_fill(e->stream, e->model2.p, vol, 0.);
    emc::launch_fill(e->stream, e->weight.p, vol, 0.);
    HIP_TRY(hipGetLastError());
    return 0;
}

// update_scale_bg (max_scale.c:134-288): per frame, the root in its scale of gradient_d,
// bracketed (stopping once 99% of the frames are bracketed) and then found by Ridder's method.
// The per-frame bookkeeping runs here, as the reference's does; every gradient evaluation is
// gradient_d_kernel over all orientation chunks (views regenerated unless they fit one chunk).
static int update_scale_bg(emc_engine *e, const emc_iter_opts *opts) {
    hipStream_t s = e->stream;
    const int N = e->num_data, P = e->num_pix, C = e->C, Rn = e->num_rot, Rp = e->Rp;
    int rc;
    std::vector<double> scale(N), psum_d(N);
    std::vector<uint8_t> black(N);
    double rescale = 0.;
    HIP_TRY(hipMemcpyAsync(scale.data(), e->scale.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(psum_d.data(), e->psum_d.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(black.data(), e->blacklist.p, N, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&rescale, e->scalars.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if ((rc = e->bg_gpart.alloc((size_t)(C / emc::kRTile) * N))) return rc;
    if ((rc = e->bg_gacc.alloc(N))) return rc;
    if ((rc = e->bg_s.alloc(N))) return rc;
    if ((rc = e->bg_dmask.alloc(N))) return rc;
    const bool one_chunk = C >= Rp;
    bool views_ready = false;
    std::vector<double> tmp(N);
    std::vector<uint8_t> mask(N, 0);
    auto gradient = [&](const std::vector<double> &sv, std::vector<double> &G) -> int {
        int r2;
        if ((r2 = e->bg_dmask.upload(mask.data(), N, s))) return r2;
        if ((r2 = e->bg_s.upload(sv.data(), N, s))) return r2;
        emc::launch_fill(s, e->bg_gacc.p, N, 0.);
        for (int r0 = 0; r0 < Rp; r0 += C) {
            const int Cc = (Rp - r0 < C) ? (Rp - r0) : C;
            // the views of the current model (the tomogram wrote its own into this buffer)
            if (!(one_chunk && views_ready))
                emc::launch_expand(s, e->rots.p, e->model1.p, e->size, e->qvals.p, e->corr.p, e->mask.p, P, r0, Cc,
                                   kPixPerGroup, e->logw.p, e->upart.p, e->ubad.p, nullptr, 1, 1);
            emc::launch_gradient_d(s, e->logw.p, P, e->R.p, N, r0, Cc, Rn, e->lik_off.p, e->lik_pix.p, e->lik_cnt.p,
                                   e->bg_dmask.p, e->bg_s.p, e->bgscale.p, e->bg.p, rescale, e->bg_gpart.p,
                                   e->bg_gacc.p);
        }
        views_ready = true;
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(tmp.data(), e->bg_gacc.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int d = 0; d < N; ++d)
            if (mask[d] == 0) G[d] = tmp[d] - psum_d[d];
        return 0;
    };
    std::vector<double> s_old(N, 0.), s_new(N, 0.), s_mid(N, 0.), s_lat(N, 0.);
    std::vector<double> g_old(N, 0.), g_new(N, 0.), g_mid(N, 0.), g_lat(N, 0.);
    for (int d = 0; d < N; ++d) {
        if (black[d]) {
            mask[d] = 255;
            scale[d] = -1.;
            continue;
        }
        s_new[d] = (opts->iteration > 1 || opts->known_scale) ? 4. * scale[d] : 32.;
    }
    if ((rc = gradient(s_old, g_old))) return rc;
    for (int d = 0; d < N; ++d)
        if ((mask[d] != 255) & (g_old[d] < 1.e-6)) {
            scale[d] = 0.001;  // fewer photons than the background
            mask[d] = 255;
        }
    int num_mask = 0;
    for (int i = 0; i < 5; ++i) {
        if ((rc = gradient(s_new, g_new))) return rc;
        num_mask = 0;
        for (int d = 0; d < N; ++d) {
            if (mask[d] != 0) {
                num_mask++;
            } else if (g_new[d] < 0.) {
                mask[d] = 192;
                num_mask++;
            } else {
                s_new[d] *= 4.;
            }
        }
        if (num_mask > ((double)0.99 * N)) break;
    }
    for (int i = 0; i < 50; ++i) {
        num_mask = 0;
        for (int d = 0; d < N; ++d) {
            if (mask[d] == 255) {
                num_mask++;
            } else if (std::fabs(s_old[d] - s_new[d]) < 1.e-5) {
                mask[d] = 255;
                scale[d] = 0.5 * (s_old[d] + s_new[d]);
                num_mask++;
            } else if (mask[d] == 192) {
                mask[d] = 0;
                s_mid[d] = 0.5 * (s_old[d] + s_new[d]);
            } else {
                s_mid[d] = 0.5 * (s_old[d] + s_new[d]);
            }
        }
        if (num_mask == N) break;
        if ((rc = gradient(s_mid, g_mid))) return rc;
        for (int d = 0; d < N; ++d) {
            if (mask[d] == 255) continue;
            if (std::fabs(g_mid[d]) < 1.e-3) {
                scale[d] = s_mid[d];
                mask[d] = 255;
                num_mask++;
                continue;
            }
            const double sq = std::sqrt(g_mid[d] * g_mid[d] - g_old[d] * g_new[d]);
            if (sq < 1.e-15) {
                s_lat[d] = s_mid[d];
            } else {
                s_lat[d] = s_mid[d] + (s_mid[d] - s_old[d]) * std::copysign(1.0, g_old[d]) * g_mid[d] / sq;
                const double lo = std::fmin(s_old[d], s_new[d]), hi = std::fmax(s_old[d], s_new[d]);
                if (s_lat[d] <= lo || s_lat[d] >= hi) s_lat[d] = s_mid[d];
            }
        }
        if (num_mask == N) break;
        if ((rc = gradient(s_lat, g_lat))) return rc;
        for (int d = 0; d < N; ++d)
            if (mask[d] < 255) {
                if (std::fabs(g_lat[d]) < 1.e-3) {
                    scale[d] = s_lat[d];
                    mask[d] = 255;
                    num_mask++;
                } else if (g_mid[d] * g_lat[d] < 0.) {
                    s_old[d] = s_mid[d];
                    g_old[d] = g_mid[d];
                    s_new[d] = s_lat[d];
                    g_new[d] = g_lat[d];
                } else if (g_old[d] * g_lat[d] < 0.) {
                    s_new[d] = s_lat[d];
                    g_new[d] = g_lat[d];
                } else {
                    s_old[d] = s_lat[d];
                    g_old[d] = g_lat[d];
                }
            }
        if (num_mask == N) break;
    }
    if ((rc = e->scale.upload(scale.data(), N, s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int emc_maximize(emc_engine *e, const emc_iter_opts *opts, emc_iter_stats *stats) {
    int rc;
    if ((rc = check_ready(e))) return rc;
    if (!opts) return fail(EMC_E_ARG, "emc_maximize: null options");
    HIP_TRY(hipSetDevice(e->device));
    if (e->lists_dirty && (rc = build_lists(e))) return rc;
    const bool mp = mp_mode(e);
    const bool g32 = e->precision == 3;
    // background-aware updates (max_prob.c:94, max_update.c:11, max_scale.c:9)
    const bool bgp = e->with_bg && opts->need_scaling;
    if (bgp && (e->precision != 0 || e->num_modes > 1 || e->refine))
        return fail(EMC_E_STATE, "emc_maximize: the background-aware path runs in FP64 with one mode, no refinement");
    if (g32) e->tomo_tile_mode = 2;  // the FP32 posterior feeds the frame-partitioned tomogram
    if (e->estep_only && bgp)
        return fail(EMC_E_STATE, "emc_maximize: estep_only does not run on the background-aware path");
    if (e->mstep_only && !(e->refine && e->subset_mode && e->subset_ready && !mp))
        return fail(EMC_E_STATE, "emc_maximize: mstep_only runs on the refinement-subset path only");
    if (e->refine && e->subset_mode && e->subset_ready && !mp) return maximize_subset(e, opts, stats);
    e->subset_ran = false;
    e->logw_resident = false;  // the plain path's chunks use the log W buffers
    if (mp && e->num_modes > 1)
        return fail(EMC_E_STATE, "emc_maximize: the mixed-precision path takes one mode");
    if (mp) {
        for (uint8_t v : e->h_dense)
            if (!v) return fail(EMC_E_STATE, "emc_maximize: the mixed-precision path needs dense frames");
        if (e->max_count > 2048)
            return fail(EMC_E_STATE, "emc_maximize: photon count %d is not exact in FP16 (limit 2048)", e->max_count);
        if (mp_bf16(e) && e->max_count > 256)
            return fail(EMC_E_STATE, "emc_maximize: photon count %d is not exact in BF16 (limit 256)", e->max_count);
    }
    if ((rc = ensure_work(e))) return rc;
    if (g32 && (rc = ensure_model32(e))) return rc;

    hipStream_t s = e->stream;
    const int N = e->num_data, Rn = e->num_rot, P = e->num_pix, C = e->C, M = e->num_modes;
    const size_t vol = (size_t)e->size * e->size * e->size * M;  // every mode's volume
    const int ngroups = (P + kPixPerGroup - 1) / kPixPerGroup;
    const int use_scale = opts->need_scaling && (opts->iteration > 1 || opts->known_scale);
    const int want_psum_d = opts->need_scaling && opts->update_scale;
    float ms_expand = 0.f, ms_lik = 0.f, ms_tomo = 0.f, ms_reduce = 0.f, ms_merge = 0.f;

    // allocate_memory: zero the merge targets (maximize.c:93-94)
    emc::launch_fill(s, e->model2.p, vol, 0.);
    emc::launch_fill(s, e->weight.p, vol, 0.);

    // mixed path: per-pixel centre of log W and its per-frame photon sum, from this model
    if (mp) {
        HIP_TRY(hipEventRecord(e->ev[0], s));
        emc::launch_center(s, e->rots.p, Rn, e->model1.p, e->size, e->qvals.p, e->corr.p, e->mask.p, P,
                           e->center_t.p);
        emc::launch_center_base(s, e->lik_off.p, e->lik_pix.p, e->lik_cnt.p, e->center_t.p, N,
                                e->center_base.p);
        HIP_TRY(hipEventRecord(e->ev[1], s));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventSynchronize(e->ev[1]));
        float a = 0.f;
        HIP_TRY(hipEventElapsedTime(&a, e->ev[0], e->ev[1]));
        ms_expand += a;
    }

    // expand + likelihood, chunk by chunk over orientations
    for (int r0 = 0; r0 < e->Rp; r0 += C) {
        const int Cc = (e->Rp - r0 < C) ? (e->Rp - r0) : C;
        float a = 0.f, b = 0.f;
        HIP_TRY(hipEventRecord(e->ev[0], s));
        if (mp)
            emc::launch_expand_mp(s, e->rots.p, e->model1.p, e->size, e->qvals.p, e->corr.p, e->mask.p,
                                  e->center_t.p, P, r0, Cc, kPixPerGroup, e->mp_ldp, mp_two(e),
                                  e->lw_hi.p, e->lw_lo.p, e->upart.p, e->ubad.p, mp_bf16(e));
        else if (g32)
            emc::launch_expand(s, e->rots.p, e->model32.p, e->size, e->qvals.p, e->corr.p, e->mask.p,
                               P, r0, Cc, kPixPerGroup, e->logw32.p, e->upart.p, e->ubad.p, nullptr, M);
        else
            emc::launch_expand(s, e->rots.p, e->model1.p, e->size, e->qvals.p, e->corr.p, e->mask.p,
                               P, r0, Cc, kPixPerGroup, e->logw.p, e->upart.p, e->ubad.p, nullptr, M, bgp);
        emc::launch_u_reduce(s, e->upart.p, e->ubad.p, ngroups, Cc, r0, Rn, e->u.p, e->rbad.p);
        HIP_TRY(hipEventRecord(e->ev[1], s));
        if (mp) {
            const size_t ts = (size_t)N * emc::kRTile;
            if (emc::launch_mp_gemm(s, mp_two(e) ? 2 : 1, 0, e->klik.p, e->mp_ldp, e->lw_hi.p,
                                    e->lw_lo.p, e->mp_ldp, (int)e->mp_ldp, (int)e->mp_rows_d,
                                    (int)round_up(Cc, emc::kGemmBN), N, Cc,
                                    e->R.p + (size_t)(r0 / emc::kRTile) * ts, ts, e->center_base.p,
                                    nullptr, 0, 1., mp_bf16(e)))
                return fail(EMC_E_STATE, "likelihood product: shape does not fit the tiles");
        } else if (g32) {
            emc::launch_likelihood(s, e->logw32.p, Cc, P, e->lik_off.p, e->lik_pix.p, e->lik_cnt.p,
                                   e->blacklist.p, N, e->R.p, r0, e->lik_tiles, e->lik_unroll);
        } else if (bgp) {
            emc::launch_likelihood_bg(s, e->logw.p, Cc, P, e->lik_off.p, e->lik_pix.p, e->lik_cnt.p, e->blacklist.p,
                                      N, e->scale.p, e->bgscale.p, e->bg.p, e->R.p, r0);
        } else {
            if (e->lik_part.p)
                emc::launch_likelihood_split(s, e->logw.p, Cc, P, e->lik8_off.p, e->lik_split, e->lik8_pix.p,
                                             e->lik8_cnt.p, e->blacklist.p, N, e->lik_part.p, e->R.p, r0);
            else
                emc::launch_likelihood(s, e->logw.p, Cc, P, e->lik_off.p, e->lik_pix.p, e->lik_cnt.p,
                                       e->blacklist.p, N, e->R.p, r0, e->lik_tiles, e->lik_unroll, e->lik_xcd);
        }
        HIP_TRY(hipEventRecord(e->ev[2], s));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventSynchronize(e->ev[2]));
        HIP_TRY(hipEventElapsedTime(&a, e->ev[0], e->ev[1]));
        HIP_TRY(hipEventElapsedTime(&b, e->ev[1], e->ev[2]));
        ms_expand += a;
        ms_lik += b;
    }
    HIP_TRY(hipEventRecord(e->ev[3], s));
    emc::launch_rescale(s, e->quats.p, e->u.p, Rn, e->mean_count, e->scalars.p, M);
    HIP_TRY(hipEventRecord(e->ev[4], s));
    emc::launch_frame_max(s, e->R.p, Rn, N, e->logwr.p, e->u.p, e->scale.p, use_scale,
                          e->scalars.p, e->blacklist.p, e->any_dense ? e->dense.p : nullptr,
                          e->rbad.p, e->refine ? e->qmap.p : nullptr,
                          e->refine ? e->relbits.p : nullptr, e->rel_words, e->max_exp.p, e->rmax.p);
    HIP_TRY(hipEventRecord(e->ev[5], s));
    emc::launch_frame_norm(s, e->R.p, Rn, N, e->beta.p, e->blacklist.p, e->max_exp.p,
                           e->rmax.p, e->p_norm.p);
    HIP_TRY(hipEventRecord(e->ev[6], s));
    emc::launch_posterior(s, e->R.p, Rn, N, e->logwr.p, e->u.p, e->beta.p, e->sum_fact.p,
                          e->blacklist.p, e->max_exp.p, e->rmax.p, e->p_norm.p, e->scalars.p,
                          want_psum_d, e->likelihood.p, e->info.p, e->psum_d.p, M);
    if (M > 1) {
        if ((rc = e->quat_norm.alloc((size_t)N * M))) return rc;
        emc::launch_occupancy(s, e->R.p, Rn, N, M, e->quat_norm.p);
    }
    emc::launch_psum_r(s, e->R.p, Rn, N, kFramesPerPsumPart, e->scale.p, opts->need_scaling,
                       e->blacklist.p, e->psum_part.p, e->psum_r.p);
    if (bgp) {
        // calc_psum_r's scalemax per orientation, for the background-aware tomogram
        const int nseg = N < 64 ? N : 64;
        if ((rc = e->bg_smax_part.alloc((size_t)nseg * e->Rp))) return rc;
        if ((rc = e->bg_smax.alloc(e->Rp))) return rc;
        if ((rc = e->bg_last.alloc(e->Rp))) return rc;
        emc::launch_scalemax(s, e->R.p, N, e->Rp, e->scale.p, e->bg_smax_part.p, nseg, e->bg_smax.p);
    }
    HIP_TRY(hipEventRecord(e->ev[7], s));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(e->ev[7]));
    float ms_rescale = 0.f, ms_max = 0.f, ms_norm = 0.f, ms_post = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms_rescale, e->ev[3], e->ev[4]));
    HIP_TRY(hipEventElapsedTime(&ms_max, e->ev[4], e->ev[5]));
    HIP_TRY(hipEventElapsedTime(&ms_norm, e->ev[5], e->ev[6]));
    HIP_TRY(hipEventElapsedTime(&ms_post, e->ev[6], e->ev[7]));

    if (e->estep_only) {
        // expectation step only: per-frame statistics of the posterior, model2 and
        // inter_weight left zero, scale factors unchanged
        emc::launch_frame_stats(s, e->likelihood.p, e->info.p, e->blacklist.p, N, N - e->num_blacklist,
                                e->scalars.p);
        HIP_TRY(hipGetLastError());
        double host_scalars[3];
        HIP_TRY(hipMemcpyAsync(host_scalars, e->scalars.p, sizeof host_scalars, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        e->have_results = true;
        if (stats) {
            stats->rescale = host_scalars[0];
            stats->avg_likelihood = host_scalars[1];
            stats->mutual_info = host_scalars[2];
            const float ms[EMC_NUM_STAGES] = {ms_expand, ms_lik, ms_rescale, ms_max, ms_norm, ms_post, 0.f, 0.f, 0.f};
            for (int i = 0; i < EMC_NUM_STAGES; ++i) stats->stage_ms[i] = ms[i];
            stats->num_chunks = (e->Rp + C - 1) / C;
            stats->view_nonzero = 0.;
        }
        return 0;
    }

    // inter_weight does not depend on the data: when every orientation merges (all
    // psum_r > 0) it is the all-orientation sum, computed once and reused
    // brick merge when the view is orientation-major (frame-partitioned tomogram, mixed path)
    const bool bricks = e->merge_mode == 1 && (mp || e->tomo_tile_mode == 2);
    if (bricks && !e->bricks_valid) {
        HIP_TRY(hipEventRecord(e->ev[0], s));
        if ((rc = build_bricks(e))) return rc;
        HIP_TRY(hipEventRecord(e->ev[1], s));
        HIP_TRY(hipEventSynchronize(e->ev[1]));
        float a = 0.f;
        HIP_TRY(hipEventElapsedTime(&a, e->ev[0], e->ev[1]));
        ms_merge += a;
    }
    int unmerged = 0;
    if ((rc = e->d_count.alloc(1))) return rc;
    emc::launch_count_unmerged(s, e->psum_r.p, Rn, e->d_count.p);
    HIP_TRY(hipMemcpyAsync(&unmerged, e->d_count.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    e->last_unmerged = unmerged;
    const bool cached_weight = (unmerged == 0);
    if (cached_weight) {
        HIP_TRY(hipEventRecord(e->ev[0], s));
        if (!e->weight_all_valid) {
            if ((rc = e->weight_all.alloc(vol))) return rc;
            emc::launch_fill(s, e->weight_all.p, vol, 0.);
            if (bricks) {
                // the brick kernel adds into e->weight: accumulate there, then keep a copy
                emc::launch_fill(s, e->weight.p, vol, 0.);
                for (size_t k = 0; k < e->bricks.size(); ++k)
                    if ((rc = merge_chunk(e, (int)k, nullptr, true, false, true))) return rc;
                HIP_TRY(hipMemcpyAsync(e->weight_all.p, e->weight.p, vol * sizeof(double),
                                       hipMemcpyDeviceToDevice, s));
            } else {
                for (int r0 = 0; r0 < e->Rp; r0 += C) {
                    const int Cc = (e->Rp - r0 < C) ? (e->Rp - r0) : C;
                    emc::launch_merge(s, e->rots.p, r0, Cc, Rn, nullptr, e->psum_r.p, 1, e->qvals.p,
                                      e->corr.p, e->mask.p, P, e->size, nullptr, e->weight_all.p, 0, 1, M);
                }
            }
            HIP_TRY(hipGetLastError());
            e->weight_all_valid = true;
        }
        HIP_TRY(hipMemcpyAsync(e->weight.p, e->weight_all.p, vol * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipEventRecord(e->ev[1], s));
        HIP_TRY(hipEventSynchronize(e->ev[1]));
        float a = 0.f;
        HIP_TRY(hipEventElapsedTime(&a, e->ev[0], e->ev[1]));
        ms_merge += a;
    }

    // tomogram + merge, chunk by chunk
    for (int r0 = 0; r0 < e->Rp; r0 += C) {
        const int Cc = (e->Rp - r0 < C) ? (e->Rp - r0) : C;
        float a = 0.f, b = 0.f, c = 0.f;
        HIP_TRY(hipEventRecord(e->ev[0], s));
        const bool parts = !mp && !bgp && e->tomo_tile_mode == 2;
        if (bgp) {
            // update_tomogram_bg: first pass for each orientation's last bracket round, then
            // the solve; the view goes where the merge reads it
            HIP_TRY(hipMemsetAsync(e->bg_last.p + r0, 0, (size_t)Cc * sizeof(int), s));
            for (int pass = 1; pass >= 0; --pass)
                emc::launch_tomogram_bg(s, pass, e->R.p, N, r0, Cc, Rn, e->tom_off.p, e->tom_frame.p, e->tom_cnt.p,
                                        e->scale.p, e->bgscale.p, e->bg.p, e->powder.p, e->mask.p, P,
                                        e->mean_count, e->psum_r.p, e->bg_smax.p, e->bg_last.p, e->logw.p,
                                        bricks ? 1 : 0);
        } else if (mp) {
            // P^T K on matrix cores: P stored as fp16(P * 2^15) where P > 1e-6
            const double pscale = 32768.;
            emc::launch_posterior_fp16(s, e->R.p, N, r0, Cc, e->mp_ldd, pscale, e->p16.p);
            if (emc::launch_mp_gemm(s, 1, 1, e->p16.p, e->mp_ldd, e->ktom.p, nullptr, e->mp_ldd,
                                    (int)e->mp_ldd, (int)round_up(Cc, emc::kGemmBM), (int)e->mp_rows_t, Cc,
                                    P, e->logw.p, (size_t)P, nullptr, e->psum_r.p + r0, Rn - r0,
                                    1. / pscale))
                return fail(EMC_E_STATE, "tomogram product: shape does not fit the tiles");
        } else if (g32) {
            emc::launch_to_f32(s, e->R.p, N, r0, Cc, e->p32.p);
            emc::launch_tomogram_parts_f32(s, e->p32.p, N, Cc, e->tom_off2.p, e->tom_parts, e->tom_frame.p,
                                           e->tom_cnt.p, P, e->tom_partial.p);
        } else if (parts)
            emc::launch_tomogram_parts(s, e->R.p, N, r0, Cc, e->tom_off2.p, e->tom_parts,
                                       e->tom_frame.p, e->tom_cnt.p, e->psum_r.p, Rn, P,
                                       e->tom_partial.p, e->logw.p, e->tomo_width32,
                                       e->tomo_unroll, e->tomo_seq);
        else if (e->tomo_tile_mode == 1)
            emc::launch_tomogram_tiles(s, e->R.p, N, r0, Cc, e->ft_off.p, e->ft_frame.p,
                                       e->fe_off.p, e->fe_code.p, e->psum_r.p, Rn, P, e->logw.p);
        else
            emc::launch_tomogram(s, e->R.p, N, r0, Cc, e->tom_off.p, e->tom_frame.p,
                                 e->tom_cnt.p, e->psum_r.p, Rn, P, e->logw.p, e->tomo_tiles);
        HIP_TRY(hipEventRecord(e->ev[1], s));
        if (parts && bricks)
            emc::launch_tomogram_reduce_rmajor(s, e->tom_partial.p, e->tom_parts, r0, Cc, e->psum_r.p,
                                               Rn, P, e->logw.p);
        else if (parts)
            emc::launch_tomogram_reduce(s, e->tom_partial.p, e->tom_parts, r0, Cc, e->psum_r.p,
                                        Rn, P, e->logw.p);
        HIP_TRY(hipEventRecord(e->ev[2], s));
        if (bricks) {
            if ((rc = merge_chunk(e, r0 / C, e->logw.p, false, true, !cached_weight))) return rc;
        } else {
            emc::launch_merge(s, e->rots.p, r0, Cc, Rn, e->logw.p, e->psum_r.p, 0, e->qvals.p,
                              e->corr.p, e->mask.p, P, e->size, e->model2.p, e->weight.p, 1,
                              cached_weight ? 0 : 1, M);
        }
        HIP_TRY(hipEventRecord(e->ev[8], s));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventSynchronize(e->ev[8]));
        HIP_TRY(hipEventElapsedTime(&a, e->ev[0], e->ev[1]));
        HIP_TRY(hipEventElapsedTime(&b, e->ev[1], e->ev[2]));
        HIP_TRY(hipEventElapsedTime(&c, e->ev[2], e->ev[8]));
        ms_tomo += a;
        ms_reduce += b;
        ms_merge += c;
    }
    // combine_information_mpi statistics and update_scale
    emc::launch_frame_stats(s, e->likelihood.p, e->info.p, e->blacklist.p, N, N - e->num_blacklist,
                            e->scalars.p);
    if (want_psum_d && e->with_bg) {
        if ((rc = update_scale_bg(e, opts))) return rc;
    } else if (want_psum_d) {
        emc::launch_scale(s, e->fcounts.p, e->psum_d.p, e->blacklist.p, N, e->scale.p);
    }
    HIP_TRY(hipGetLastError());
    double host_scalars[3];
    HIP_TRY(hipMemcpyAsync(host_scalars, e->scalars.p, sizeof host_scalars, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    e->have_results = true;
    if (stats) {
        stats->rescale = host_scalars[0];
        stats->avg_likelihood = host_scalars[1];
        stats->mutual_info = host_scalars[2];
        const float ms[EMC_NUM_STAGES] = {ms_expand, ms_lik, ms_rescale, ms_max, ms_norm,
                                          ms_post, ms_tomo, ms_reduce, ms_merge};
        for (int i = 0; i < EMC_NUM_STAGES; ++i) stats->stage_ms[i] = ms[i];
        stats->num_chunks = (e->Rp + C - 1) / C;
        stats->view_nonzero = 0.;
    }
    return 0;
}

int emc_update_model(emc_engine *e, double alpha, double *rms_change) {
    if (e) ++e->gen;
    if (e) e->model32_valid = false, e->logw_resident = false;  // model1 changes below
    if (!e || !e->size) return fail(EMC_E_STATE, "emc_update_model: model not set");
    HIP_TRY(hipSetDevice(e->device));
    int rc;
    if ((rc = e->model_part.alloc(kModelParts))) return rc;
    if ((rc = e->scalars.alloc(8))) return rc;
    emc::launch_update_model(e->stream, e->model1.p, e->model2.p, e->weight.p, e->size, alpha,
                             e->model_part.p, kModelParts, e->scalars.p + 3, e->num_modes);
    HIP_TRY(hipGetLastError());
    double change = 0.;
    HIP_TRY(hipMemcpyAsync(&change, e->scalars.p + 3, sizeof change, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const long vol = e->size * e->size * e->size;
    if (rms_change) *rms_change = std::sqrt(change / e->num_modes / vol);  // recon.pyx:149
    return 0;
}

int emc_normalize_scale(emc_engine *e, double *mean_scale) {
    if (e) ++e->gen;
    if (e) e->model32_valid = false, e->logw_resident = false;  // model1 changes below
    if (!e || !e->have_params || !e->size) return fail(EMC_E_STATE, "emc_normalize_scale: state not set");
    HIP_TRY(hipSetDevice(e->device));
    int rc;
    if ((rc = e->scalars.alloc(8))) return rc;
    emc::launch_normalize_scale(e->stream, e->scale.p, e->blacklist.p, e->num_data, e->model1.p,
                                e->size * e->size * e->size, e->scalars.p + 4);
    HIP_TRY(hipGetLastError());
    double m = 0.;
    HIP_TRY(hipMemcpyAsync(&m, e->scalars.p + 4, sizeof m, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (mean_scale) *mean_scale = m;
    return 0;
}

int emc_slice_gen(emc_engine *e, const double *quat, double *view_out) {
    if (!e || !quat || !view_out) return fail(EMC_E_ARG, "emc_slice_gen: bad argument");
    if (!e->num_pix || !e->size) return fail(EMC_E_STATE, "emc_slice_gen: detector/model not set");
    HIP_TRY(hipSetDevice(e->device));
    double rot[9];
    make_rot(quat, rot);
    int rc;
    if ((rc = e->one_rot.upload(rot, 9, e->stream))) return rc;
    if ((rc = e->one_view.alloc(e->num_pix))) return rc;
    emc::launch_slice_gen(e->stream, e->one_rot.p, e->model1.p, e->size, e->qvals.p, e->corr.p,
                          e->d_perm.p, e->num_pix, e->one_view.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(view_out, e->one_view.p, e->num_pix * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int emc_slice_gen_static(emc_engine *e, const double *quat, double *view_out) {
    if (!e || !quat || !view_out) return fail(EMC_E_ARG, "emc_slice_gen_static: bad argument");
    if (!e->num_pix || !e->size) return fail(EMC_E_STATE, "emc_slice_gen_static: detector/model not set");
    if (e->num_pix > EMC_STATIC_MAX_PIX)
        return fail(EMC_E_ARG, "emc_slice_gen_static: %d pixels exceed the static staging maximum of %d pixels",
                    e->num_pix, EMC_STATIC_MAX_PIX);
    HIP_TRY(hipSetDevice(e->device));
    double *ptr = g_scratch_ptr[e->device];  // set by emc_create on this device
    if (!ptr) return fail(EMC_E_STATE, "emc_slice_gen_static: staging memory of device %d not resolved", e->device);
    double rot[9];
    make_rot(quat, rot);
    // the copy and the kernel are ordered on e->stream, and the stream is drained before rot
    // goes out of scope
    HIP_TRY(hipMemcpyAsync(ptr, rot, sizeof rot, hipMemcpyHostToDevice, e->stream));
    emc::launch_slice_gen(e->stream, ptr, e->model1.p, e->size, e->qvals.p, e->corr.p, e->d_perm.p, e->num_pix,
                          ptr + kStaticViewOffset);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(view_out, ptr + kStaticViewOffset, e->num_pix * sizeof(double), hipMemcpyDeviceToHost,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int emc_slice_merge(emc_engine *e, const double *quat, const double *view) {
    if (!e || !quat || !view) return fail(EMC_E_ARG, "emc_slice_merge: bad argument");
    if (!e->num_pix || !e->size) return fail(EMC_E_STATE, "emc_slice_merge: detector/model not set");
    HIP_TRY(hipSetDevice(e->device));
    double rot[9];
    make_rot(quat, rot);
    int rc;
    if ((rc = e->one_rot.upload(rot, 9, e->stream))) return rc;
    if ((rc = e->one_view.upload(view, e->num_pix, e->stream))) return rc;
    emc::launch_slice_merge(e->stream, e->one_rot.p, e->one_view.p, e->qvals.p, e->corr.p,
                            e->mask.p, e->d_perm.p, e->num_pix, e->size, e->model2.p, e->weight.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

// emc_get_sparse_prob after an iteration on refinement subsets: frames in batches of at most
// 256 MB of dense rows, each scattered from its pair rows (and, for full frames, its RF rows)
// into orientation order and compacted like the dense posterior.
static int sparse_prob_subset(emc_engine *e, int *num_prob, long capacity, int *place, double *prob,
                              long *total) {
    hipStream_t s = e->stream;
    const int N = e->num_data, Rn = e->num_rot;
    const long rowmax = e->sparse_rows > 0 ? (long)e->sparse_rows
                                           : (long)(((size_t)256 << 20) / ((size_t)Rn * sizeof(double)));
    const int nb = (int)std::max(1L, std::min((long)N, rowmax));
    int rc;
    DevBuf<double> dense, d_prob;
    DevBuf<int> cnt, d_place;
    DevBuf<long> d_off;
    if ((rc = dense.alloc((size_t)nb * Rn))) return rc;
    if ((rc = cnt.alloc(nb))) return rc;
    const std::vector<long> &poff = e->h_poff;
    const std::vector<int> &fidx = e->h_fidx;
    auto scatter = [&](int d0, int d1) {
        emc::launch_fill(s, dense.p, (size_t)(d1 - d0) * Rn, 0.);
        emc::launch_pair_scatter(s, e->Rsub.p, e->pair_frame.p, e->pair_tile.p, e->slot_rot.p, poff[d1] - poff[d0],
                                 Rn, dense.p, poff[d0], d0);
        const int f0 = (int)(std::lower_bound(fidx.begin(), fidx.end(), d0) - fidx.begin());
        const int f1 = (int)(std::lower_bound(fidx.begin(), fidx.end(), d1) - fidx.begin());
        if (f1 > f0)
            emc::launch_full_scatter(s, e->RF.p, e->num_full, e->num_tiles, e->fidx.p, e->slot_rot.p, Rn, dense.p,
                                     f0, f1, d0);
    };
    long tot = 0;
    for (int d0 = 0; d0 < N; d0 += nb) {
        const int d1 = std::min(N, d0 + nb);
        scatter(d0, d1);
        emc::launch_sparse_count_rows(s, dense.p, Rn, d1 - d0, cnt.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(num_prob + d0, cnt.p, (size_t)(d1 - d0) * sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int d = d0; d < d1; ++d) tot += num_prob[d];
    }
    if (total) *total = tot;
    if (!place || !prob) return 0;
    if (capacity < tot) return fail(EMC_E_ARG, "emc_get_sparse_prob: capacity %ld < %ld entries", capacity, tot);
    std::vector<long> off;
    long done = 0;
    for (int d0 = 0; d0 < N; d0 += nb) {
        const int d1 = std::min(N, d0 + nb);
        off.assign(d1 - d0, 0);
        long n = 0;
        for (int d = d0; d < d1; ++d) {
            off[d - d0] = n;
            n += num_prob[d];
        }
        if (n > 0) {
            scatter(d0, d1);
            if ((rc = d_off.upload(off.data(), off.size(), s))) return rc;
            if ((rc = d_place.alloc(n))) return rc;
            if ((rc = d_prob.alloc(n))) return rc;
            emc::launch_sparse_fill_rows(s, dense.p, Rn, d1 - d0, d_off.p, d_place.p, d_prob.p);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(place + done, d_place.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(prob + done, d_prob.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        done += n;
    }
    return 0;
}

int emc_get_sparse_prob(emc_engine *e, int *num_prob, long capacity, int *place, double *prob,
                        long *total) {
    if (!e || !num_prob) return fail(EMC_E_ARG, "emc_get_sparse_prob: bad argument");
    if (!e->have_results) return fail(EMC_E_STATE, "emc_get_sparse_prob: no iteration has run");
    HIP_TRY(hipSetDevice(e->device));
    if (e->subset_ran) return sparse_prob_subset(e, num_prob, capacity, place, prob, total);
    hipStream_t s = e->stream;
    const int N = e->num_data, Rn = e->num_rot;
    int rc;
    DevBuf<int> cnt;
    if ((rc = cnt.alloc(N))) return rc;
    emc::launch_sparse_count(s, e->R.p, Rn, N, cnt.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(num_prob, cnt.p, (size_t)N * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    long tot = 0;
    for (int d = 0; d < N; ++d) tot += num_prob[d];
    if (total) *total = tot;
    if (!place || !prob) return 0;
    if (capacity < tot) return fail(EMC_E_ARG, "emc_get_sparse_prob: capacity %ld < %ld entries", capacity, tot);
    // frames in batches of at most kBatch entries (a frame with more goes alone)
    constexpr long kBatch = 32L << 20;
    DevBuf<int> d_place;
    DevBuf<double> d_prob;
    DevBuf<long> d_off;
    std::vector<long> off;
    long done = 0;
    for (int d0 = 0; d0 < N;) {
        int d1 = d0;
        long n = 0;
        off.clear();
        while (d1 < N && (d1 == d0 || n + num_prob[d1] <= kBatch)) {
            off.push_back(n);
            n += num_prob[d1++];
        }
        if (n > 0) {
            if ((rc = d_place.alloc(n))) return rc;
            if ((rc = d_prob.alloc(n))) return rc;
            if ((rc = d_off.upload(off.data(), off.size(), s))) return rc;
            emc::launch_sparse_fill(s, e->R.p, Rn, N, d0, d1 - d0, d_off.p, d_place.p, d_prob.p);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(place + done, d_place.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(prob + done, d_prob.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        done += n;
        d0 = d1;
    }
    return 0;
}

int emc_subset_mass(emc_engine *e, const long *rel_off, const int *rel_quat, double *mass) {
    if (!e || !rel_off || !mass) return fail(EMC_E_ARG, "emc_subset_mass: bad argument");
    if (!e->have_results || e->subset_ran)
        return fail(EMC_E_STATE, "emc_subset_mass: needs the dense posterior of a plain iteration");
    const int N = e->num_data;
    const long total = rel_off[N];
    if (rel_off[0] != 0 || total < 0 || (total > 0 && !rel_quat))
        return fail(EMC_E_ARG, "emc_subset_mass: rel_off must start at 0 and ascend");
    for (int d = 0; d < N; ++d)
        if (rel_off[d + 1] < rel_off[d]) return fail(EMC_E_ARG, "emc_subset_mass: rel_off descends at frame %d", d);
    for (long k = 0; k < total; ++k)
        if (rel_quat[k] < 0 || rel_quat[k] >= e->num_rot)
            return fail(EMC_E_ARG, "emc_subset_mass: orientation %d at entry %ld outside [0, %d)", rel_quat[k], k,
                        e->num_rot);
    HIP_TRY(hipSetDevice(e->device));
    int rc;
    if ((rc = e->mass_off.upload(rel_off, (size_t)N + 1, e->stream))) return rc;
    if (total > 0 && (rc = e->mass_idx.upload(rel_quat, (size_t)total, e->stream))) return rc;
    if ((rc = e->mass_out.alloc((size_t)N))) return rc;
    emc::launch_subset_mass(e->stream, e->R.p, N, e->mass_off.p, total > 0 ? e->mass_idx.p : nullptr, e->mass_out.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(mass, e->mass_out.p, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int emc_get_field(emc_engine *e, int field, void *dst, size_t count) {
    if (!e || !dst) return fail(EMC_E_ARG, "emc_get_field: bad argument");
    HIP_TRY(hipSetDevice(e->device));
    const size_t N = e->num_data, Rn = e->num_rot, vol = (size_t)e->size * e->size * e->size;
    const void *src = nullptr;
    size_t n = 0, elem = sizeof(double);
    bool needs_results = true;
    switch (field) {
        case EMC_FIELD_U: src = e->u.p; n = Rn; break;
        case EMC_FIELD_MAX_EXP: src = e->max_exp.p; n = N; break;
        case EMC_FIELD_P_NORM: src = e->p_norm.p; n = N; break;
        case EMC_FIELD_RMAX: src = e->rmax.p; n = N; elem = sizeof(int); break;
        case EMC_FIELD_LIKELIHOOD: src = e->likelihood.p; n = N; break;
        case EMC_FIELD_INFO: src = e->info.p; n = N; break;
        case EMC_FIELD_PSUM_D: src = e->psum_d.p; n = N; break;
        case EMC_FIELD_PSUM_R: src = e->psum_r.p; n = Rn; break;
        case EMC_FIELD_SCALE: src = e->scale.p; n = N; needs_results = false; break;
        case EMC_FIELD_MODEL1: src = e->model1.p; n = vol * e->num_modes; needs_results = false; break;
        case EMC_FIELD_MODEL2: src = e->model2.p; n = vol * e->num_modes; needs_results = false; break;
        case EMC_FIELD_INTER_WEIGHT: src = e->weight.p; n = vol * e->num_modes; needs_results = false; break;
        case EMC_FIELD_QUAT_NORM:
            if (e->num_modes < 2) return fail(EMC_E_STATE, "emc_get_field: occupancies need num_modes > 1");
            src = e->quat_norm.p;
            n = N * e->num_modes;
            break;
        case EMC_FIELD_PROB: {
            if (!e->have_results) return fail(EMC_E_STATE, "emc_get_field: no iteration has run");
            if (count != N * Rn) return fail(EMC_E_ARG, "emc_get_field: count %zu != %zu", count, N * Rn);
            if (e->subset_ran) {
                DevBuf<double> dense;
                int rc;
                if ((rc = dense.alloc(N * Rn))) return rc;
                HIP_TRY(hipMemsetAsync(dense.p, 0, N * Rn * sizeof(double), e->stream));
                emc::launch_pair_scatter(e->stream, e->Rsub.p, e->pair_frame.p, e->pair_tile.p, e->slot_rot.p,
                                         e->num_pairs, (int)Rn, dense.p);
                if (e->num_full > 0)
                    emc::launch_full_scatter(e->stream, e->RF.p, e->num_full, e->num_tiles, e->fidx.p,
                                             e->slot_rot.p, (int)Rn, dense.p);
                HIP_TRY(hipMemcpyAsync(dst, dense.p, N * Rn * sizeof(double), hipMemcpyDeviceToHost, e->stream));
                HIP_TRY(hipStreamSynchronize(e->stream));
                dense.release();
                return 0;
            }
            // tile-major [Rp/64][N][64] to the caller's [N][num_rot]
            for (size_t j = 0; j * emc::kRTile < Rn; ++j) {
                const size_t w = std::min((size_t)emc::kRTile, Rn - j * emc::kRTile);
                HIP_TRY(hipMemcpy2DAsync((double *)dst + j * emc::kRTile, Rn * sizeof(double),
                                         e->R.p + j * N * emc::kRTile, emc::kRTile * sizeof(double),
                                         w * sizeof(double), N, hipMemcpyDeviceToHost, e->stream));
            }
            HIP_TRY(hipStreamSynchronize(e->stream));
            return 0;
        }
        default: return fail(EMC_E_ARG, "emc_get_field: unknown field %d", field);
    }
    if (needs_results && !e->have_results) return fail(EMC_E_STATE, "emc_get_field: no iteration has run");
    if (!src) return fail(EMC_E_STATE, "emc_get_field: field %d not allocated", field);
    if (count != n) return fail(EMC_E_ARG, "emc_get_field: count %zu != %zu", count, n);
    HIP_TRY(hipMemcpyAsync(dst, src, n * elem, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

}  // extern "C"
