// Internal interface between the kernel translation unit (kernels.hip) and
// the host runtime (runtime.hip) of libfbp.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbp {

// Shape of one chunk, passed by value to every kernel.
struct Dims {
    uint32_t C;        // channels
    uint32_t K;        // samples per channel
    uint32_t T;        // templates
    uint32_t N;        // samples per template and channel
    uint32_t M;        // n_max_shift_inds
    uint32_t S;        // 2M+1 shifts
    uint32_t W;        // ceil(K/N) windows
    uint32_t n_floor;  // floor(K/N): bound used by the flag writes of the reference
    uint32_t Tp;       // T padded to the correlation kernel's template block
    uint32_t reduce_wg;// work-group width of the reference's overlap reduction
    uint64_t items;    // S*S*T overlap combinations per window
    uint64_t nl;       // ceil(items / reduce_wg) reduction groups per window
    uint32_t ov_batch; // screened overlap: windows per Q-table batch (rows of ovQ)
    uint32_t sup_max;  // widest channel support of a template (C until templates are set)
    uint32_t xc_cap;   // screened overlap: candidate slots per batch of the deferred exact
                       // scores (k_overlap_exact), 0 = every window scores in the screen
    uint32_t corr_ctb; // matrix-core correlation: template tiles per block (0 until templates are set)
    uint32_t corr_span;// widest channel span of such a chunk of tiles (0 until templates are set)
    uint32_t xc_wide;  // screened overlap (test switch): eight-wave lean screen and exact scores
                       // from the padded templates in global memory whatever their size
    uint32_t dirty_g;  // channels per bit of a window's dirty mask (ceil(C / 64))
    uint32_t dirty_off;// test switch: recompute every listed (window, template chunk)
};

// Bits of the channel groups (dirty_g channels each) that meet channels [lo, hi).
__host__ __device__ inline uint64_t channel_mask(uint32_t lo, uint32_t hi, uint32_t g)
{
    if (lo >= hi) return 0;
    const uint32_t b0 = lo / g, b1 = (hi - 1) / g;   // b1 <= 63
    const uint64_t upto = b1 >= 63 ? ~0ull : (2ull << b1) - 1;
    return upto & ~((1ull << b0) - 1);
}

// The re-score queue is split into kQueueShards equal shards; a correlation
// block reserves entries in shard blockIdx % kQueueShards, whose counters
// (entries, positions) sit kShardStride words apart, a cache line each, so
// reservations do not serialise on one address.
constexpr uint32_t kQueueShards = 64;
constexpr uint32_t kShardStride = 32;

// Device counters, read back by the host once per pass.
struct Counters {
    uint32_t n_ov;       // overlap windows of the pass
    uint32_t n_new;      // spikes added by the pass
    uint32_t n_next;     // windows of the next pass
    uint32_t n_rescore;  // low-precision path: re-score queue entries (groups of 4 positions)
    uint32_t n_combo;    // screened overlap: combinations re-scored directly
    uint32_t n_rescore_pos;  // low-precision path: positions re-scored in float32
    uint32_t pad[2];
    unsigned long long combos;   // overlap combinations scored exactly since the run started
    unsigned long long screened; // combinations screened in float64 (screened overlap)
    unsigned long long corr_pairs; // (position, template) likelihoods computed since the run
                                   // started (listed blocks whose voltage changed)
};

// Device buffers of one chunk.
struct DeviceState {
    float *V;            // [C][K] voltage, residual in place
    float *V0;           // [C][K] pristine voltage (restored at the start of a run)
    float *Tm;           // [T][C][N] templates
    float *TmT;          // [C][N][Tp] templates, template-fastest, zero padded
    float *bias;         // [Tp]
    float *thr;          // [T]
    float *L;            // [T][K] likelihood
    uint64_t *wdirty;    // [W] channel groups whose voltage changed under the window's positions
                         // since its likelihood was last computed (all set at the start of a run)
    float *best_L;       // [W]
    uint32_t *best_lab;  // [W]
    uint32_t *best_idx;  // [W]
    uint32_t *ov_idx;    // [W]
    uint32_t *ov_lab;    // [W]
    uint8_t *nc;         // [W+2] next-pass flags: bits 0-1 value, bit 2 pruned
    uint8_t *ov;         // [W+2] overlap flags: bit 0 flag, bit 1 gave up
    uint32_t *win;       // [W] window list of the pass
    uint32_t *win_next;  // [W] window list of the next pass
    uint32_t *ov_list;   // [W] overlap windows of the pass
    uint32_t *ov_sorted; // [W] the same, grouped by best label (screened overlap)
    float *part_v;       // [W] best likelihood within each window (argmax partials)
    uint32_t *part_t;    // [W] its template
    uint32_t *part_p;    // [W] its position
    float *ovQ;          // [ov_batch][T][S] screened overlap: Q - B2/2 of a batch of windows
    double *ovVV;        // [ov_batch] squared norm of each such window's voltage (all channels)
    uint32_t *sup_lo;    // [T] first channel on which the template is nonzero
    uint32_t *sup_hi;    // [T] one past the last such channel (sup_lo == sup_hi: all zero)
    uint32_t *tord;      // [CT*16] matrix-core template order: by channel support, then index
    uint32_t *tile_lo;   // [CT] first channel on which a 16-template tile (in tord) is nonzero
    uint32_t *tile_hi;   // [CT] one past its last
    uint8_t *has;        // [W] add decision per entry of the pass list
    uint32_t *cand_idx;  // [W] spike onset per entry of the pass list
    uint32_t *cand_lab;  // [W]
    uint32_t *win_spike; // [W] 1 + position in the pass's spike list, 0 = none
    uint32_t *blk;       // [n_blocks+1] compaction block counts / offsets
    Counters *ctr;       // device counters
    uint32_t *sp_idx;    // [cap] spike onsets, record order
    uint32_t *sp_lab;    // [cap]
    float *clips;        // [n][C*N] adjusted clips (allocated on demand)
    float *tabAB;        // [T][S][T][S] screened overlap: shifted-template cross products
    float *tabABmin;     // [T][S][T]    min over the last shift of tabAB (row bounds of the screen)
    float *tabB2;        // [T][S]       squared norms of shifted, truncated templates
    float *tnorm;        // [T]          template norms (upper bounds)
    float *TmQ;          // screened overlap: split-bf16 A fragments of the matrix-core Q table (uint4 each)
    float *Tpad;         // [T][C][N + 2M] templates at [M, M + N), +0.0 around (overlap scores)
    float *tscale;       // [T]          fp8 screen: per-template power-of-two scale (1 otherwise)
    void *TmB;           // packed bf16/fp16/fp8 template fragments for the matrix cores
    uint64_t *rq;        // re-score queue: (template << 32) | position
    uint32_t *rq_shard;  // per-shard reservation counters of the queue (entries, positions)
    void *rdesc;         // [W * chunks] (queue start, count) per (window, template chunk)
    uint32_t *xc_cand;   // [xc_cap] deferred exact-score candidates of a batch (f << 16 | e)
    uint32_t *xc_win;    // [ov_batch][4] deferred windows: (list position, first candidate, count, 0)
    uint32_t *xc_ctr;    // [2] deferred windows and candidate slots taken in the batch
};

constexpr int kCompactItems = 1024;   // items per block of the compaction kernels

}  // namespace fbp

struct fbp_segment;

namespace fbp {

// Pass-by-pass driving of a segment (runtime.hip), for callers that advance
// many segments from one thread.  segment_begin resets the segment and
// enqueues its first pass on `st` (nullptr: the segment's own stream); once
// segment_pass_event(g) has completed, segment_step reads the pass's counters
// and either enqueues the next pass (*more = 1) or ends the run (*more = 0;
// FBP_OK, FBP_EPASSES or an error, as fbp_segment_run returns).
int segment_begin(fbp_segment *g, hipStream_t st);
hipEvent_t segment_pass_event(const fbp_segment *g);
int segment_step(fbp_segment *g, int *more);
// Device bytes the segment holds now (buffers, spike list, clips).
uint64_t segment_device_bytes(const fbp_segment *g);

// Record the message of a failed C-ABI call for fbp_last_error() on this
// thread (runtime.hip) and return `code`.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Launch wrappers (kernels.hip).  All enqueue on `st` and return hipError_t.
hipError_t launch_fill_u8(uint8_t *p, uint64_t n, uint8_t v, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st);
hipError_t launch_reset_counters(Counters *c, hipStream_t st);
hipError_t launch_iota_u32(uint32_t *p, uint64_t n, hipStream_t st);
hipError_t launch_copy_rows(float *dst, const float *src, uint32_t rows, uint64_t cols,
                            uint64_t src_stride, hipStream_t st);
hipError_t launch_transpose_templates(const Dims &d, const float *Tm, const float *bias_in,
                                      float *TmT, float *bias_out, hipStream_t st);

hipError_t launch_corr_fp32(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st);
uint32_t mfma_kstep(uint32_t precision);
uint32_t mfma_words(uint32_t precision);   // 8-byte words per template fragment
uint32_t mfma_ks(const Dims &d, uint32_t precision);
uint32_t mfma_ct(const Dims &d);
uint32_t mfma_nchunk(const Dims &d);
// tiles per correlation block for templates with these channel supports
uint32_t mfma_ctb_for(uint32_t n_tiles, uint32_t n_chans, uint32_t widest_support);
hipError_t launch_template_norms(const Dims &d, const DeviceState &s, uint32_t precision, hipStream_t st);
hipError_t launch_pack_templates(const Dims &d, const DeviceState &s, uint32_t precision, hipStream_t st);
hipError_t launch_corr_screen(const Dims &d, const DeviceState &s, uint32_t n_win, uint32_t precision,
                              const float *coef, uint32_t rq_cap, uint32_t no_rescore, hipStream_t st);
hipError_t launch_argmax(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st);
hipError_t launch_compact_ov(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st);
hipError_t launch_overlap_direct(const Dims &d, const DeviceState &s, uint32_t max_ov, hipStream_t st);
size_t overlap_screened_lds(const Dims &d);
uint64_t overlap_qtemplate_floats(const Dims &d);
hipError_t launch_pad_templates(const Dims &d, const DeviceState &s, hipStream_t st);
hipError_t launch_overlap_tables(const Dims &d, const DeviceState &s, hipStream_t st);
hipError_t launch_overlap_screened(const Dims &d, const DeviceState &s, uint32_t max_ov, uint32_t debug,
                                  hipStream_t st);
// candidate slots of the deferred exact scores for a batch (0: templates too
// large for the exact-score kernel's LDS, every window scores in the screen)
uint32_t overlap_exact_cap(const Dims &d);
hipError_t launch_check_overlap(const Dims &d, const DeviceState &s, uint32_t max_ov, hipStream_t st);
hipError_t launch_pursuit_add(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st);
hipError_t launch_compact_spikes(const Dims &d, const DeviceState &s, uint32_t n_win,
                                 uint64_t spike_base, hipStream_t st);
hipError_t launch_pursuit_prune(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st);
hipError_t launch_compact_next(const Dims &d, const DeviceState &s, hipStream_t st);
hipError_t launch_residual(const Dims &d, const DeviceState &s, uint64_t spike_base,
                           uint32_t n_new, hipStream_t st);
hipError_t launch_clear_win_spike(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st);
hipError_t launch_get_clips(const float *V, uint64_t row_stride, const int64_t *idx, uint64_t n_events,
                            const uint32_t *chan, uint32_t n_sel, int32_t w0, uint32_t W, float *clips,
                            hipStream_t st);
hipError_t launch_clip_sse(const float *clips, uint64_t n, uint32_t width, const uint32_t *tidx,
                           const float *templates, const uint32_t *prog, uint32_t n_prog, float *sse, hipStream_t st);
template <class F>
hipError_t launch_unit_means(const F *clips, uint32_t width, const uint32_t *rows, const uint64_t *off,
                             uint32_t n_units, F *means, hipStream_t st);
hipError_t launch_overlap_template_residuals(const float *Tm, uint32_t C, uint32_t N, int32_t n_pre, uint32_t S,
                                             const uint32_t *triples, uint64_t n_triples, const uint32_t *prog,
                                             uint32_t n_prog, double *best_res, uint32_t *best_combo,
                                             hipStream_t st);
hipError_t launch_clips(const Dims &d, const DeviceState &s, uint64_t n_spikes, hipStream_t st);
// samples per channel the covariance and template-variance kernels hold in LDS
constexpr uint32_t kMaxCovSamples = 128;
// row slices of the covariance (returns their count R, sets their length);
// scratch: colpart [C][R][N], part [C][R][N][N] doubles
uint32_t channel_covariance_slices(uint64_t n_rows, uint64_t *slice);
hipError_t launch_channel_covariance(const float *clips, uint64_t width, uint64_t n_clips, const uint64_t *rows,
                                     uint64_t n_rows, uint32_t C, uint32_t N, double *colpart, double *part,
                                     double *cov, hipStream_t st);
hipError_t launch_template_variance(const double *Tm, uint32_t T, uint32_t C, uint32_t N, const double *cov,
                                    double *var, hipStream_t st);
// flat cross-correlation and channel-aligned dot product of template pairs
// at lags lag_lo .. lag_lo + n_lags - 1 (separability.hip)
hipError_t launch_template_pair_lags(const double *Tm, uint32_t C, uint32_t N, const uint32_t *pairs,
                                     uint64_t n_pairs, int32_t lag_lo, uint32_t n_lags, double *xcorr,
                                     double *aligned, hipStream_t st);

}  // namespace fbp
