/*
 * fbp.h -- C-ABI of libfbp.so, the MI355X (gfx950) Full Binary Pursuit engine.
 *
 * The library replaces the OpenCL program and the device-side half of the
 * reference's host driver:
 *   spikesorting_fullpursuit/parallel/binary_pursuit_parallel.py:128-845
 *     (binary_pursuit(): device pick, buffers, pass loop, output copies)
 *   spikesorting_fullpursuit/kernels/binary_pursuit.cl:1-1130
 *     (compute_residual, compute_full_likelihood,
 *      compute_template_maximum_likelihood, overlap_recheck_indices,
 *      parse_overlap_recheck_indices, check_overlap_reassignments,
 *      binary_pursuit, get_adjusted_clips)
 * The reference has no FFI of its own: its "plugin" is the Python function
 * binary_pursuit(), which drives pyopencl.  The Python package
 * spikesorting_fullpursuit_amd binds these entry points with ctypes.
 *
 * Conventions follow the reference (binary_pursuit.cl:11-29):
 *   voltage   float32 [C][K]  channel-major, one chunk of K samples
 *   templates float32 [T][C][N] (row t = template label t, 0-based)
 *   bias      float32 [T] = float32(-0.5 * template_SS)  (bpp.py:199)
 *   thr       float32 [T] = neuron_lower_thresholds     (bpp.py:201)
 *   spike indices are template ONSETS within the chunk (uint32)
 *
 * Every call returns 0 on success or a negative FBP_E* code; the message of
 * the last failure on the calling thread is returned by fbp_last_error().
 * No device pointer is returned to the caller except through the *_device
 * entry points, which take caller-owned device memory.
 */
#ifndef FBP_H
#define FBP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FBP_OK = 0,
    FBP_EINVAL = -1,   /* bad argument (bpp.py raises ValueError/KeyError)      */
    FBP_ENOMEM = -2,   /* device memory exhausted                              */
    FBP_ENODEV = -3,   /* no usable GPU (bpp.py:229-230 RuntimeError)          */
    FBP_EHIP = -4,     /* HIP runtime error                                    */
    FBP_EPASSES = -5,  /* pass cap reached before convergence                   */
    FBP_ESTATE = -6,   /* call out of order (e.g. results before run)           */
    FBP_ETIMEDOUT = -7 /* fbp_server_wait: the job has not ended in time        */
};

/* Arithmetic of the likelihood correlation (binary_pursuit.cl:216-300). */
enum {
    FBP_PREC_FP32 = 0,  /* float32 fused multiply-add chain, reference order     */
    FBP_PREC_BF16 = 1,  /* bf16 MFMA screen, float32 re-score of every decision  */
    FBP_PREC_FP16 = 2,  /* fp16 MFMA screen, float32 re-score                    */
    FBP_PREC_FP8 = 3    /* fp8 e4m3 MFMA screen (per-template scale), re-score   */
};

/* Arithmetic of the overlap recheck (binary_pursuit.cl:444-654). */
enum {
    FBP_OVERLAP_DIRECT = 0,   /* every (t, f, s) scored as the reference does      */
    FBP_OVERLAP_SCREENED = 1  /* decomposed float64 screen, direct re-score of the
                                 candidates inside the rounding bound            */
};

typedef struct fbp_config {
    uint32_t n_chans;        /* C (num_neighbor_channels)                       */
    uint32_t n_samples;      /* K: samples per channel of the chunk (> N)       */
    uint32_t n_templates;    /* T (num_templates)                               */
    uint32_t template_len;   /* N (template_length = n_samples_per_chan)        */
    uint32_t max_shift;      /* M = n_max_shift_inds (bpp.py:294-298)          */
    uint32_t do_overlap;     /* sort_info['do_overlap_recheck'] (bpp.py:570)    */
    uint32_t reduce_wg;      /* width of the overlap work-group reduction that
                                breaks exact score ties (power of two; the
                                reference uses the device's work-group size,
                                bpp.py:259; 256 by default)                    */
    uint32_t precision;      /* FBP_PREC_*                                      */
    uint32_t overlap_mode;   /* FBP_OVERLAP_*                                   */
    uint32_t max_passes;     /* 0 = unbounded (the reference has no bound)      */
    uint32_t debug_flags;    /* FBP_DEBUG_*; 0 in normal use                    */
} fbp_config;

/* Test-only switches. */
enum {
    FBP_DEBUG_NO_RESCORE = 1,      /* low-precision paths keep the screened likelihood
                                      (no float32 re-score): checks the error bound */
    FBP_DEBUG_OVERLAP_FULL = 2,    /* screened overlap: find the exact-score
                                      candidates by a second full traversal     */
    FBP_DEBUG_OVERLAP_GLOBAL = 4,  /* screened overlap: exact scores read the
                                      partner template from global memory       */
    FBP_DEBUG_OVERLAP_EXACT = 8,   /* screened overlap: score a single candidate
                                      exactly even when its interval settles the
                                      parse                                      */
    FBP_DEBUG_OVERLAP_ALL_ROWS = 16, /* screened overlap: traverse every row, none
                                      skipped by its upper bound                 */
    FBP_DEBUG_OVERLAP_SMALL_BATCH = 32, /* screened overlap: batches of 37 windows
                                      (Q tables, then the screen, per batch)     */
    FBP_DEBUG_OVERLAP_INBLOCK = 64, /* screened overlap: every exact score in the
                                      screen's block (no deferred exact-score
                                      kernel)                                    */
    FBP_DEBUG_OVERLAP_FEW_SLOTS = 128, /* screened overlap: 64 deferred candidate
                                      slots per batch, so most windows with
                                      candidates score them in the screen       */
    FBP_DEBUG_OVERLAP_WIDE = 256,  /* screened overlap: the eight-wave screen block
                                      and deferred exact scores that read the
                                      templates from global memory (the route of
                                      templates too large for LDS)              */
    FBP_DEBUG_FULL_CORRELATION = 512 /* recompute the likelihood of every listed
                                      window and template, also where no voltage
                                      under it changed since the last pass      */
};

typedef struct fbp_stats {
    uint32_t n_passes;
    uint32_t reserved;
    uint64_t n_spikes;
    uint64_t windows_evaluated;   /* sum over passes of W_p                     */
    uint64_t overlap_windows;     /* sum over passes of O_p                     */
    uint64_t rescored_positions;  /* low-precision paths: float32 re-scores     */
    uint64_t screened_combos;     /* screened overlap: float64 screen evaluations */
    double   ms_total;            /* device time of the last run (HIP events)   */
    double   ms_correlation;      /* summed over passes, HIP events on the run's stream */
    double   ms_overlap;          /* overlap recheck + parse + reassignment      */
    uint64_t corr_launches;       /* correlation kernel launches (one per pass) */
    uint64_t overlap_launches;    /* overlap kernel launches                    */
    uint64_t corr_positions;      /* likelihood positions computed, all passes  */
    uint64_t overlap_combos;      /* (t,f,s) combinations scored with the reference's
                                     float32 chain, all passes                    */
    uint64_t corr_pairs;          /* (position, template) likelihoods computed, all
                                     passes: corr_positions * T less the listed ones
                                     whose template's channels saw no voltage change */
} fbp_stats;

typedef struct fbp_segment fbp_segment;

/* States of a job of the segment server (fbp_server_*). */
enum {
    FBP_JOB_QUEUED = 0,    /* submitted, waiting for a slot and HBM budget, or
                              being created and uploaded                        */
    FBP_JOB_RUNNING = 1,   /* its passes are being driven                       */
    FBP_JOB_DONE = 2,      /* results readable through fbp_server_segment       */
    FBP_JOB_FAILED = 3,
    FBP_JOB_CANCELLED = 4
};

typedef struct fbp_server_stats {
    uint64_t submitted, completed, failed, cancelled;
    uint32_t queued;            /* waiting for admission                        */
    uint32_t running;           /* being created or driven                      */
    uint32_t peak_running;
    uint32_t reserved;
    uint64_t held_bytes;        /* device bytes the admitted jobs count for     */
    uint64_t peak_held_bytes;
    uint64_t budget_bytes;
    uint64_t pass_steps;        /* passes consumed by the driver thread         */
    double   driver_busy_s;     /* driver thread time spent reading counters and
                                   enqueueing passes                            */
    double   uptime_s;
} fbp_server_stats;

typedef struct fbp_job_info {
    int32_t  state;             /* FBP_JOB_*                                    */
    int32_t  rc;                /* run result: FBP_OK, FBP_EPASSES (capped,
                                   results readable) or the failure's code      */
    uint64_t reserve_bytes;     /* bytes reserved at admission (grows to what
                                   the job held when that was more)             */
    uint64_t bytes;             /* device bytes the job holds now               */
    uint32_t passes;
    uint32_t reserved;
    double   t_submit, t_admit, t_start, t_finish;   /* seconds since the server
                                   was created; -1 until reached                */
} fbp_job_info;

typedef struct fbp_server fbp_server;

/* Library version (major*10000 + minor*100 + patch). No reference counterpart. */
int fbp_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *fbp_last_error(void);

/* Number of visible GPUs; FBP_ENODEV when there is none.
 * Replaces the device search of bpp.py:219-230. */
int fbp_device_count(int *count);

/* Total and free bytes of HBM on a device.
 * Replaces GLOBAL_MEM_SIZE, bpp.py:278 (input of the chunk planner). */
int fbp_device_memory(int device, uint64_t *total_bytes, uint64_t *free_bytes);

/* Device-resident state of one chunk: voltage/residual, likelihood buffer,
 * per-window pass state and spike lists (bpp.py:355-423). */
int fbp_segment_create(int device, const fbp_config *cfg, fbp_segment **out);

/* Device bytes fbp_segment_create would allocate for cfg (validates cfg).
 * Plays the role of the reference's per-chunk memory estimate (bpp.py:306-326)
 * for callers that admit several segments against an HBM budget. */
int fbp_segment_bytes(const fbp_config *cfg, uint64_t *bytes);
void fbp_segment_destroy(fbp_segment *seg);

/* Upload templates [T][C][N], bias [T] and thresholds [T] from host memory
 * (bpp.py:355-357, :410).  The library copies them. */
int fbp_segment_set_templates(fbp_segment *seg, const float *templates, const float *bias,
                              const float *thresholds);

/* Upload the chunk voltage from host memory; row_stride is the distance in
 * floats between channel rows (>= K), so a slice of a longer recording needs
 * no host copy (bpp.py:374-377, :386). */
int fbp_segment_set_voltage(fbp_segment *seg, const float *voltage, uint64_t row_stride);

/* Same from caller-owned device memory (device-to-device). */
int fbp_segment_set_voltage_device(fbp_segment *seg, const float *d_voltage, uint64_t row_stride);

/* Run the pass loop to convergence (bpp.py:532-736).  The voltage buffer is
 * turned into the residual in place.  `stream` is a hipStream_t or NULL for
 * the segment's own stream.  Blocks until done. */
int fbp_segment_run(fbp_segment *seg, void *stream);

/* Device bytes the segment holds now: its buffers, the spike list (which
 * grows while it runs) and clips once fetched.  Right after
 * fbp_segment_create it equals fbp_segment_bytes.  No reference counterpart. */
int fbp_segment_device_bytes(const fbp_segment *seg, uint64_t *bytes);

/* Results of the last run. */
int fbp_segment_num_spikes(const fbp_segment *seg, uint64_t *n);
/* Onsets and labels in record order (pass, then ascending window). */
int fbp_segment_get_spikes(fbp_segment *seg, uint32_t *onsets, uint32_t *labels);
/* Adjusted clips [n][C*N] = residual + template (binary_pursuit.cl:1088-1130,
 * bpp.py:771-809). */
int fbp_segment_get_clips(fbp_segment *seg, float *clips);
/* Residual voltage [C][K] (not returned by the reference, bpp.py:822). */
int fbp_segment_get_residual(fbp_segment *seg, float *residual);
/* Likelihood buffer [T][K] as it stands after the run (parity checks). */
int fbp_segment_get_likelihood(fbp_segment *seg, float *likelihood);
int fbp_segment_get_stats(const fbp_segment *seg, fbp_stats *stats);
/* Window list of the pass that would come next when the run was stopped by
 * cfg.max_passes (the reference's window_indices = nonzero(
 * check_window_on_next_pass), binary_pursuit_parallel.py:683-689), ascending;
 * *n = 0 after a run that converged.  For parity checks of the flag rules. */
int fbp_segment_get_window_list(fbp_segment *seg, uint32_t *windows, uint64_t capacity, uint64_t *n);

/* Device address of the segment's input voltage [C][row_stride] (restored
 * before every run), for fbp_get_clips on resident data. */
int fbp_segment_voltage_device(const fbp_segment *seg, const float **d_voltage, uint64_t *row_stride);

/* get_clips (parallel/segment_parallel.py:440-509), as binary pursuit's
 * caller uses it on the input voltage (full_binary_pursuit.py:357-359).
 * d_voltage: device float32 [n_chans][row_stride], n_samples valid per row.
 * event_indices: host int64 [n_events], ascending as the reference requires.
 * channels: host uint32 [n_sel] (the reference's `neighbors`).  The clip of
 * event k is samples [index + w0, index + w1) of each selected channel, in
 * `channels` order (w0, w1 from time_window_to_samples).  With check_valid,
 * events are trimmed from the front while index + w0 < 0 and from the back
 * while index + w1 >= n_samples (segment_parallel.py:462-484); an event left
 * in between whose clip leaves the data is FBP_EINVAL.  valid: host uint8
 * [n_events] (1 = clip returned); clips: host float32
 * [n_valid][n_sel * (w1 - w0)] in event order. */
int fbp_get_clips(int device, const float *d_voltage, uint64_t row_stride, uint32_t n_chans, uint64_t n_samples,
                  const int64_t *event_indices, uint64_t n_events, const uint32_t *channels, uint32_t n_sel,
                  int32_t w0, int32_t w1, int check_valid, float *clips, uint8_t *valid, uint64_t *n_valid);

/* fbp_get_clips on a host voltage float32 [n_chans][row_stride]: the
 * selected channels' rows are uploaded for the gather. */
int fbp_get_clips_host(int device, const float *voltage, uint64_t row_stride, uint32_t n_chans, uint64_t n_samples,
                       const int64_t *event_indices, uint64_t n_events, const uint32_t *channels, uint32_t n_sel,
                       int32_t w0, int32_t w1, int check_valid, float *clips, uint8_t *valid, uint64_t *n_valid);

/* Mean clip of every unit: np.mean(clips[labels == u], axis=0) on float32
 * clips (full_binary_pursuit.py:366-368), bit for bit: rows added in input
 * order in float32, the sum divided by the row count.  clips: host float32
 * [n][width]; labels: host uint32 [n], each < n_units; means: host float32
 * [n_units][width] (NaN for a unit without rows); counts: optional host
 * uint64 [n_units]. */
int fbp_unit_mean_clips(int device, const float *clips, uint64_t n, uint32_t width, const uint32_t *labels,
                        uint32_t n_units, float *means, uint64_t *counts);

/* The same mean in float64, as np.mean computes it for clips of integer or
 * float64 voltage (full_binary_pursuit.py:368 with v_dtype int16, float64):
 * clips host float64 [n][width] (integer clips converted exactly), rows added
 * in input order in float64; means host float64 [n_units][width]. */
int fbp_unit_mean_clips_f64(int device, const double *clips, uint64_t n, uint32_t width, const uint32_t *labels,
                            uint32_t n_units, double *means, uint64_t *counts);

/* Squared distance of every clip to a template, as the duplicate removal
 * after binary pursuit computes it (np.sum((clip - template) ** 2) on float32
 * clips, analyze_spike_timing.py:129-130, with numpy's pairwise summation
 * order: bit-exact).  clips: host float32 [n][width]; template_index: host
 * uint32 [n] (< n_templates); templates: host float32 [n_templates][width];
 * sse: host float32 [n]. */
int fbp_clip_template_sse(int device, const float *clips, uint64_t n, uint32_t width, const uint32_t *template_index,
                          const float *templates, uint32_t n_templates, float *sse);

/* Noise covariance of every channel over a sample of residual clip rows,
 * cov[c] = np.cov(clips[rows, c*N:(c+1)*N], rowvar=False, ddof=0) in float64
 * (full_binary_pursuit.py:245-252; the reference draws the rows with
 * np.random.randint per channel, the caller passes the ones it drew).
 * clips: host float32 [n_clips][n_chans*N]; rows: host uint64 [n_chans][n_rows]
 * (each < n_clips), or NULL for every clip (n_rows = n_clips); cov: host
 * float64 [n_chans][N][N].  N <= 128, n_rows >= 1. */
int fbp_channel_covariance(int device, const float *clips, uint64_t n_clips, uint32_t n_chans, uint32_t N,
                           const uint64_t *rows, uint64_t n_rows, double *cov);

/* fbp_channel_covariance for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Inputs and outputs are the
 * caller's DEVICE memory: d_clips float32 [n_clips][n_chans*N], d_rows uint64
 * [n_chans][n_rows] (each < n_clips; a larger index is not read and makes
 * that channel's covariance NaN) or NULL for every clip, d_cov float64
 * [n_chans][N][N]; results are bit-identical to fbp_channel_covariance.  The
 * call allocates nothing: the partial sums live in a fixed 32 MiB scratch
 * array that every GPU the library runs on holds until the process ends, and
 * requests whose partial sums exceed it (n_chans * R * N * (N + 1) doubles,
 * R = row slices: ceil(n_rows / max(256, ceil(n_rows / 64)))) fail with
 * FBP_EINVAL.  The kernels are enqueued on `stream` (a hipStream_t, NULL for
 * the null stream) and the call returns without waiting.  Every call on one
 * GPU shares its scratch array: calls that overlap on one GPU race. */
int fbp_channel_covariance_noalloc(int device, const float *d_clips, uint64_t n_clips, uint32_t n_chans, uint32_t N,
                                   const uint64_t *d_rows, uint64_t n_rows, double *d_cov, void *stream);

/* Likelihood variance of every template, var[t] = sum_c t_c^T cov_c t_c in
 * float64 (neuron_separability.py:126-135, the input of the binary-pursuit
 * thresholds, :137-142).  templates: host float64 [T][n_chans*N]; cov: host
 * float64 [n_chans][N][N]; var: host float64 [T].  N <= 128. */
int fbp_template_variances(int device, const double *templates, uint32_t n_templates, uint32_t n_chans, uint32_t N,
                           const double *cov, double *var);

/* Template pairs at every lag of [lag_lo, lag_hi]: the alignment step of
 * optimal_align_templates (consolidate.py:9-71) as check_noise_templates
 * runs it for each (good, noisy) template pair (neuron_separability.py:
 * 192-199).  templates: host float64 [n_templates][n_chans*N]; pairs: host
 * uint32 [n_pairs][2] = (a, b) rows.  With L = n_chans*N:
 *   xcorr[p][k]   = sum_i a[i + lag] b[i], 0 <= i, i + lag < L, lag = lag_lo + k
 *                   (np.correlate(a, b, 'full')[lag + L - 1]);
 *   aligned[p][k] = the same sum over the i where i and i + lag fall on one
 *                   channel: np.dot of the two templates after each channel
 *                   is shifted by lag and truncated (consolidate.py:34-54).
 * float64, one fixed summation order per output.  aligned may be NULL.
 * -(L-1) <= lag_lo <= lag_hi <= L-1. */
int fbp_template_pair_lags(int device, const double *templates, uint32_t n_templates, uint32_t n_chans, uint32_t N,
                           const uint32_t *pairs, uint64_t n_pairs, int32_t lag_lo, int32_t lag_hi, double *xcorr,
                           double *aligned);

/* Overlap-template search residuals (find_overlap_templates,
 * c_cython/sort_cython.pyx:333-456, before binary pursuit at
 * full_binary_pursuit.py:257-262).  templates: host float32 [n_templates]
 * [n_chans * n_samples_per_chan].  For each triple (test, n1, n2) of
 * `triples` (host uint32 [n_triples][3]) and every shift pair (s1, s2) in
 * [-n_pre, n_post]^2, the residual np.sum((test - (shift(n1, s1) +
 * shift(n2, s2)))**2) in float64 with numpy's summation order (bit-exact);
 * best_residual / best_combo (host [n_triples]) receive the smallest
 * residual and the first (s1 + n_pre) * (n_pre + n_post + 1) + (s2 + n_pre)
 * holding it (0xffffffff when no residual is below infinity), the choice of
 * the reference's strict < in (s1, s2) loop order. */
int fbp_overlap_template_residuals(int device, const float *templates, uint32_t n_templates, uint32_t n_chans,
                                   uint32_t n_samples_per_chan, int32_t n_pre, int32_t n_post,
                                   const uint32_t *triples, uint64_t n_triples, double *best_residual,
                                   uint32_t *best_combo);

/* Segment server: many segments sorted concurrently on one device.
 * Replaces the serial per-segment loop of the reference's sorter
 * (parallel/spikesorting_parallel.py:1312-1322 calls binary pursuit for one
 * segment at a time, each a blocking call).  Jobs are admitted in submission
 * order when fewer than max_running are staged or running and their
 * reservation (fbp_segment_bytes of cfg plus extra_bytes, e.g. room for
 * clips) fits in hbm_budget next to what the admitted jobs hold (each job
 * counts for its reservation or its current bytes, whichever is larger; a
 * finished job counts until it is released).  One driver thread advances
 * every running job pass by pass on the job's own stream; no thread blocks on
 * a pass.  Results are bitwise those of fbp_segment_run on the same input.
 * hbm_budget 0: 85 % of the device's free memory at creation. */
int fbp_server_create(int device, uint64_t hbm_budget, uint32_t max_running, fbp_server **out);
/* Cancels queued and running jobs, frees every job and stops the threads. */
void fbp_server_destroy(fbp_server *srv);

/* Queue a segment.  templates/bias/thresholds/voltage as for
 * fbp_segment_set_templates and fbp_segment_set_voltage; they are read while
 * the job is staged, so they must stay valid and unchanged until its state is
 * no longer FBP_JOB_QUEUED.  FBP_ENOMEM when the reservation exceeds the
 * whole budget; FBP_EINVAL for a bad cfg. */
int fbp_server_submit(fbp_server *srv, const fbp_config *cfg, const float *templates, const float *bias,
                      const float *thresholds, const float *voltage, uint64_t row_stride, uint64_t extra_bytes,
                      uint64_t *ticket);

/* Current state of a job.  A failed job's message becomes this thread's
 * fbp_last_error(). */
int fbp_server_poll(fbp_server *srv, uint64_t ticket, fbp_job_info *info);

/* Wait until the job has ended or timeout_us has passed (< 0: no limit);
 * FBP_ETIMEDOUT when it has not ended.  info may be NULL. */
int fbp_server_wait(fbp_server *srv, uint64_t ticket, int64_t timeout_us, fbp_job_info *info);

/* Wait until some job has ended that this call has not reported yet, or
 * timeout_us has passed (< 0: no limit): *ticket is that job (each ended job
 * is reported once, in the order jobs ended).  FBP_ETIMEDOUT when none has.
 * A finished job holds its device memory until it is released, so a consumer
 * that takes results in completion order never stalls admission. */
int fbp_server_wait_any(fbp_server *srv, int64_t timeout_us, uint64_t *ticket, fbp_job_info *info);

/* Cancel a job: a queued one at once, a running one at its next pass
 * boundary (fbp_server_wait then reports FBP_JOB_CANCELLED).  FBP_ESTATE when
 * it has already ended. */
int fbp_server_cancel(fbp_server *srv, uint64_t ticket);

/* The finished job's segment, read with fbp_segment_num_spikes /
 * get_spikes / get_clips / get_residual / get_stats from one thread at a
 * time; valid until fbp_server_release.  FBP_ESTATE unless FBP_JOB_DONE. */
int fbp_server_segment(fbp_server *srv, uint64_t ticket, fbp_segment **seg);

/* Forget an ended job and free its device memory. */
int fbp_server_release(fbp_server *srv, uint64_t ticket);

int fbp_server_get_stats(fbp_server *srv, fbp_server_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* FBP_H */
