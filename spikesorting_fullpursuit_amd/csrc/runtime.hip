// runtime.hip -- C-ABI (include/fbp.h) and the device-resident pass loop.
//
// One fbp_segment holds every buffer of one chunk in HBM for the whole call
// and runs the reference's pass loop
// (spikesorting_fullpursuit/parallel/binary_pursuit_parallel.py:532-736) with
// a single 32-byte device-to-host read per pass: window lists and overlap
// lists are compacted on the device, kernels read their counts from device
// memory, so the host never copies flags, lists or spikes between passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fbp.h"
#include "fbp_internal.h"

using namespace fbp;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(FBP_ENOMEM, "%s: out of device memory (%s)", what, hipGetErrorString(e));
    return fail(FBP_EHIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                      \
    do {                                         \
        hipError_t e_ = (expr);                  \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

constexpr uint32_t kTemplateBlock = 8;   // must match the correlation kernel's kCorrTB
constexpr uint64_t kMaxLds = 160 * 1024 - 1024;   // LDS a gfx950 work-group may allocate

}  // namespace

struct fbp_segment {
    int device = 0;
    fbp_config cfg{};
    Dims d{};
    DeviceState s{};
    hipStream_t own_stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_stage[4] = {nullptr, nullptr, nullptr, nullptr};   // correlation / overlap brackets
    Counters *h_ctr = nullptr;   // pinned host copy of the device counters
    uint64_t cap = 0;            // spike capacity of sp_idx / sp_lab
    uint64_t n_spikes = 0;
    uint64_t n_pending = 0;      // window list length of the next pass after a capped run
    uint64_t clips_cap = 0;
    bool have_templates = false, have_voltage = false, ran = false;
    fbp_stats stats{};
    uint32_t rq_cap = 0;         // capacity of the re-score queue
    float coef[4] = {0, 0, 0, 0};  // screen bound coefficients k1..k4
    std::atomic<uint64_t> bytes{0};  // device bytes held (read by the server's admission thread)
    // pass loop state between segment_begin and the last segment_step
    hipStream_t run_stream = nullptr;
    hipEvent_t ev_pass = nullptr;    // recorded after each pass's counter read-back
    uint32_t nw = 0;                 // windows of the pass in flight
    uint64_t total = 0;              // spikes before the pass in flight
    uint32_t passes = 0;
};

namespace {

// Device memory of a segment comes from the device's stream-ordered pool on
// the segment's own stream: creating and destroying segments while others run
// (the serving path) then never synchronises the device, and freed blocks are
// reused by the next segment instead of going back to the driver.
// Once per device and process, whichever thread creates the first segment.
void keep_pool_memory(int device)
{
    static std::once_flag once[64];
    if (device < 0 || device >= 64) return;
    std::call_once(once[device], [device] {
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t keep = ~0ull;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    });
}

template <class T>
hipError_t dalloc(T **p, uint64_t n, hipStream_t st)
{
    return hipMallocAsync((void **)p, (n ? n : 1) * sizeof(T), st);
}

void dfree(void *p, hipStream_t st)
{
    if (p) (void)hipFreeAsync(p, st);
}

uint64_t round256(uint64_t b) { return (b + 255) / 256 * 256; }

// Device bytes of one buffer of n elements as a segment accounts for them.
template <class T>
uint64_t buffer_bytes(uint64_t n) { return round256((n ? n : 1) * sizeof(T)); }

// Allocation and release of a segment's buffers, tracked in g->bytes.
template <class T>
hipError_t salloc(fbp_segment *g, T **p, uint64_t n, hipStream_t st)
{
    const hipError_t e = dalloc(p, n, st);
    if (e == hipSuccess) g->bytes += buffer_bytes<T>(n);
    return e;
}

template <class T>
void sfree(fbp_segment *g, T *p, uint64_t n, hipStream_t st)
{
    if (!p) return;
    dfree(p, st);
    g->bytes -= buffer_bytes<T>(n);
}

uint64_t rescore_queue_cap(const Dims &d)
{
    uint64_t cap = (uint64_t)d.T * d.K / 8;
    if (cap < (1u << 20)) cap = 1u << 20;
    if (cap > (512u << 20)) cap = 512u << 20;
    return cap;
}

// Every buffer a segment allocates when it is created, in one list used both
// to allocate them and to size a segment before creating it:
// visit(pointer, element count) for each.
template <class V>
void for_each_buffer(const fbp_config *cfg, const Dims &d, DeviceState &s, V &&visit)
{
    const uint64_t W = d.W, CK = (uint64_t)d.C * d.K, T = d.T, CT16 = (T + 15) / 16;
    const uint64_t nb = (W + kCompactItems - 1) / kCompactItems;
    visit(s.V, CK);
    visit(s.V0, CK);
    visit(s.Tm, T * d.C * d.N);
    visit(s.TmT, (uint64_t)d.C * d.N * d.Tp);
    visit(s.bias, (uint64_t)d.Tp);
    visit(s.thr, T);
    visit(s.L, T * d.K);
    visit(s.wdirty, W);
    visit(s.best_L, W);
    visit(s.best_lab, W);
    visit(s.best_idx, W);
    visit(s.ov_idx, W);
    visit(s.ov_lab, W);
    visit(s.nc, W + 2);
    visit(s.ov, W + 2);
    visit(s.win, W);
    visit(s.win_next, W);
    visit(s.ov_list, W);
    visit(s.sup_lo, T);
    visit(s.sup_hi, T);
    visit(s.tord, CT16 * 16);
    visit(s.tile_lo, CT16);
    visit(s.tile_hi, CT16);
    visit(s.part_v, W);
    visit(s.part_t, W);
    visit(s.part_p, W);
    visit(s.has, W);
    visit(s.cand_idx, W);
    visit(s.cand_lab, W);
    visit(s.win_spike, W);
    visit(s.blk, nb + 1);
    visit(s.ctr, (uint64_t)1);
    visit(s.tnorm, T);
    visit(s.tscale, T);
    visit(s.Tpad, T * d.C * (d.N + 2 * d.M));
    if (cfg->overlap_mode == FBP_OVERLAP_SCREENED) {
        visit(s.tabAB, T * d.S * T * d.S);
        visit(s.tabABmin, T * d.S * T);
        visit(s.tabB2, T * d.S);
        visit(s.TmQ, overlap_qtemplate_floats(d));
        visit(s.ov_sorted, W);
        visit(s.ovQ, (uint64_t)d.ov_batch * T * d.S);
        visit(s.ovVV, (uint64_t)d.ov_batch);
        if (d.xc_cap) {
            visit(s.xc_cand, (uint64_t)d.xc_cap);
            visit(s.xc_win, (uint64_t)d.ov_batch * 4);
            visit(s.xc_ctr, (uint64_t)2);
        }
    }
    if (cfg->precision != FBP_PREC_FP32) {
        visit(s.rq, rescore_queue_cap(d));
        visit(s.rq_shard, (uint64_t)kQueueShards * kShardStride);
        // the chunking is chosen when the templates are set: room for one per tile
        visit(*(uint64_t **)&s.rdesc, (uint64_t)W * mfma_ct(d));
        visit(*(uint64_t **)&s.TmB, (uint64_t)d.C * mfma_ks(d, cfg->precision) * mfma_ct(d) * 64 *
                                        mfma_words(cfg->precision));
    }
}

// Initial capacity of the spike list, in spikes.
uint64_t initial_spike_cap(const Dims &d) { return 2ull * d.W; }

// Bytes a segment allocates on the device when it is created (each block
// rounded to 256 bytes), including the initial spike list.  The spike list
// may grow while a segment runs and clips are allocated on demand: g->bytes
// tracks what a segment holds.
uint64_t segment_bytes(const fbp_config *cfg, const Dims &d)
{
    DeviceState dummy{};
    uint64_t b = 0;
    for_each_buffer(cfg, d, dummy, [&](auto *&p, uint64_t n) {
        b += buffer_bytes<std::remove_reference_t<decltype(*p)>>(n);
    });
    return b + 2 * buffer_bytes<uint32_t>(initial_spike_cap(d));
}

void free_state(fbp_segment *g)
{
    DeviceState &s = g->s;
    for_each_buffer(&g->cfg, g->d, s, [&](auto *&p, uint64_t) { dfree(p, g->own_stream); });
    dfree(s.sp_idx, g->own_stream);
    dfree(s.sp_lab, g->own_stream);
    dfree(s.clips, g->own_stream);
    if (g->own_stream) (void)hipStreamSynchronize(g->own_stream);
    s = DeviceState{};
    g->bytes = 0;
    g->cap = 0;
    g->clips_cap = 0;
}

int set_device(const fbp_segment *g)
{
    HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
    return FBP_OK;
}

// Grow the spike list to hold `need` entries, keeping the first `keep`.
int ensure_spike_capacity(fbp_segment *g, uint64_t need, uint64_t keep, hipStream_t st)
{
    if (need <= g->cap) return FBP_OK;
    uint64_t cap = g->cap * 2;
    if (cap < need) cap = need;
    uint32_t *ni = nullptr, *nl = nullptr;
    HIP_TRY(salloc(g, &ni, cap, st), "spike buffer");
    hipError_t e = salloc(g, &nl, cap, st);
    if (e != hipSuccess) { sfree(g, ni, cap, st); return hip_fail(e, "spike buffer"); }
    if (keep > 0) {
        HIP_TRY(hipMemcpyAsync(ni, g->s.sp_idx, keep * 4, hipMemcpyDeviceToDevice, st), "grow spikes");
        HIP_TRY(hipMemcpyAsync(nl, g->s.sp_lab, keep * 4, hipMemcpyDeviceToDevice, st), "grow spikes");
    }
    sfree(g, g->s.sp_idx, g->cap, st);
    sfree(g, g->s.sp_lab, g->cap, st);
    g->s.sp_idx = ni;
    g->s.sp_lab = nl;
    g->cap = cap;
    return FBP_OK;
}

// numpy's float64 sum of n contiguous elements as a program for the device:
// blocks of 8192 added in order, each summed pairwise (halves rounded down to
// a multiple of 8, leaves of at most 128 elements).  A leaf op is
// start << 8 | length; 0xffffffff adds the top two stack entries.
void pairwise_leaves(std::vector<uint32_t> &prog, uint64_t start, uint64_t n, int depth, int &max_depth)
{
    if (n <= 128) {
        prog.push_back((uint32_t)(start << 8) | (uint32_t)n);
        max_depth = std::max(max_depth, depth + 1);
        return;
    }
    uint64_t n2 = n / 2;
    n2 -= n2 % 8;
    pairwise_leaves(prog, start, n2, depth, max_depth);
    pairwise_leaves(prog, start + n2, n - n2, depth + 1, max_depth);
    prog.push_back(0xffffffffu);
}

int numpy_sum_program(uint64_t n, std::vector<uint32_t> &prog)
{
    int max_depth = 0;
    for (uint64_t b = 0; b < n; b += 8192) {
        pairwise_leaves(prog, b, std::min<uint64_t>(8192, n - b), b ? 1 : 0, max_depth);
        if (b) prog.push_back(0xffffffffu);
    }
    return max_depth;
}


}  // namespace

int fbp::set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" {

int fbp_version(void) { return 100; }

const char *fbp_last_error(void) { return g_err.c_str(); }

int fbp_device_count(int *count)
{
    if (!count) return fail(FBP_EINVAL, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        *count = 0;
        return fail(FBP_ENODEV, "No GPU available");
    }
    *count = n;
    return FBP_OK;
}

int fbp_device_memory(int device, uint64_t *total_bytes, uint64_t *free_bytes)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(FBP_ENODEV, "No GPU available");
    if (device < 0 || device >= n) return fail(FBP_EINVAL, "device %d out of range (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    if (total_bytes) *total_bytes = tot;
    if (free_bytes) *free_bytes = fr;
    return FBP_OK;
}

static int make_dims(const fbp_config *cfg, Dims &d);

int fbp_segment_bytes(const fbp_config *cfg, uint64_t *bytes)
{
    if (!cfg || !bytes) return fail(FBP_EINVAL, "NULL argument");
    Dims d{};
    if (int rc = make_dims(cfg, d)) return rc;
    *bytes = segment_bytes(cfg, d);
    return FBP_OK;
}

static int make_dims(const fbp_config *cfg, Dims &d)
{
    const uint32_t C = cfg->n_chans, K = cfg->n_samples, T = cfg->n_templates, N = cfg->template_len;
    if (C == 0 || T == 0 || N == 0) return fail(FBP_EINVAL, "channels, templates and template length must be > 0");
    if (K <= N) return fail(FBP_EINVAL, "chunk of %u samples is not longer than the template (%u)", K, N);
    if (cfg->reduce_wg == 0 || (cfg->reduce_wg & (cfg->reduce_wg - 1)) || cfg->reduce_wg > (1u << 20))
        return fail(FBP_EINVAL, "reduce_wg must be a power of two");
    if ((uint64_t)cfg->max_shift >= 4ull * N)
        return fail(FBP_EINVAL, "n_max_shift_inds (%u) must be below 4x the template length", cfg->max_shift);
    if (cfg->precision > FBP_PREC_FP8)
        return fail(FBP_EINVAL, "precision %u is not available in this build", cfg->precision);
    if (cfg->precision != FBP_PREC_FP32 && N > 96)
        return fail(FBP_EINVAL, "the matrix-core screen supports templates of up to 96 samples (N=%u)", N);
    if (cfg->overlap_mode != FBP_OVERLAP_DIRECT && cfg->overlap_mode != FBP_OVERLAP_SCREENED)
        return fail(FBP_EINVAL, "unknown overlap mode %u", cfg->overlap_mode);
    if ((uint64_t)C * (2ull * N + 2ull * cfg->max_shift) * sizeof(float) > kMaxLds)
        return fail(FBP_EINVAL, "C*(2N+2M) = %llu exceeds the overlap kernel's LDS staging",
                    (unsigned long long)C * (2ull * N + 2ull * cfg->max_shift));
    d.C = C; d.K = K; d.T = T; d.N = N; d.M = cfg->max_shift;
    d.S = 2 * cfg->max_shift + 1;
    d.W = (uint32_t)(((uint64_t)K + N - 1) / N);
    d.n_floor = K / N;
    d.Tp = (T + kTemplateBlock - 1) / kTemplateBlock * kTemplateBlock;
    d.reduce_wg = cfg->reduce_wg;
    d.items = (uint64_t)d.S * d.S * T;
    d.nl = (d.items - 1) / cfg->reduce_wg + 1;
    {
        // Q tables of up to 1 GiB of windows per screened-overlap batch (a
        // single batch for most recordings: no launches beyond the windows)
        const uint64_t per = (uint64_t)d.T * d.S * 4;
        uint64_t nb = (1ull << 30) / per;
        if (nb < 1024) nb = 1024;
        if (cfg->debug_flags & FBP_DEBUG_OVERLAP_SMALL_BATCH) nb = 37;
        if (nb > d.W) nb = d.W;
        d.ov_batch = (uint32_t)nb;
    }
    d.xc_cap = (cfg->overlap_mode == FBP_OVERLAP_SCREENED && !(cfg->debug_flags & FBP_DEBUG_OVERLAP_INBLOCK))
                   ? overlap_exact_cap(d) : 0;
    if (d.xc_cap && (cfg->debug_flags & FBP_DEBUG_OVERLAP_FEW_SLOTS)) d.xc_cap = 64;
    d.xc_wide = d.xc_cap && (cfg->debug_flags & FBP_DEBUG_OVERLAP_WIDE) ? 1u : 0u;
    // the screened overlap's LDS grows with the widest template support, known
    // when the templates are set; here it must fit for the narrowest one
    d.sup_max = 1;
    if (cfg->overlap_mode == FBP_OVERLAP_SCREENED && (uint64_t)d.T * d.S > 65535)   // 16-bit (t, s) entries
        return fail(FBP_EINVAL, "screened overlap supports T*S up to 65535 (T*S = %llu)",
                    (unsigned long long)d.T * d.S);
    if (cfg->overlap_mode == FBP_OVERLAP_SCREENED && overlap_screened_lds(d) > kMaxLds)
        return fail(FBP_EINVAL, "screened overlap needs %zu bytes of LDS (T*S too large)", overlap_screened_lds(d));
    d.sup_max = C;
    d.dirty_g = (C + 63) / 64;
    d.dirty_off = (cfg->debug_flags & FBP_DEBUG_FULL_CORRELATION) ? 1u : 0u;
    return FBP_OK;
}

int fbp_segment_create(int device, const fbp_config *cfg, fbp_segment **out)
{
    if (!cfg || !out) return fail(FBP_EINVAL, "NULL argument");
    *out = nullptr;
    Dims dd{};
    if (int rc = make_dims(cfg, dd)) return rc;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(FBP_ENODEV, "No GPU available");
    if (device < 0 || device >= n) return fail(FBP_EINVAL, "device %d out of range", device);
    fbp_segment *g = new fbp_segment();
    g->device = device;
    g->cfg = *cfg;
    g->d = dd;
    const Dims &d = g->d;
    const uint32_t C = d.C, N = d.N;

    int rc = FBP_OK;
    hipError_t e = hipSetDevice(device);
    keep_pool_memory(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->own_stream, hipStreamNonBlocking);
    DeviceState &s = g->s;
    for_each_buffer(cfg, d, s, [&](auto *&p, uint64_t n) {
        if (e == hipSuccess) e = salloc(g, &p, n, g->own_stream);
    });
    if (cfg->precision != FBP_PREC_FP32) {
        g->rq_cap = (uint32_t)rescore_queue_cap(d);
        // Screen bound coefficients (see the correlation screen in kernels.hip):
        // unit roundoff u and flushed-subnormal floor delta of the format.
        const double u = cfg->precision == FBP_PREC_BF16 ? 0x1p-8 : cfg->precision == FBP_PREC_FP16 ? 0x1p-11 : 0x1p-4;
        const double delta = cfg->precision == FBP_PREC_BF16 ? 0x1p-126 : cfg->precision == FBP_PREC_FP16 ? 0x1p-14
                                                                                                        : 0x1p-6;
        const double n = (double)C * mfma_ks(d, cfg->precision) * mfma_kstep(cfg->precision) + 2.0;
        const double gp = n * 0x1p-23 / (1.0 - n * 0x1p-23);
        const double cn = (double)C * N;
        g->coef[0] = (float)((2 * u + u * u + 2 * gp * (1 + u) * (1 + u)) * 1.01);
        g->coef[1] = (float)(delta * (1 + u) * sqrt(cn) * (1 + gp) * 1.01);
        g->coef[2] = (float)(cn * delta * delta * (1 + gp) * 1.01 + 1e-30);
        g->coef[3] = (float)(2 * gp * 1.01);
    }
    if (e == hipSuccess) e = hipEventCreate(&g->ev_start);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_pass, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&g->ev_stop);
    for (int q = 0; q < 4 && e == hipSuccess; q++) e = hipEventCreate(&g->ev_stage[q]);
    if (e == hipSuccess) e = hipHostMalloc((void **)&g->h_ctr, sizeof(Counters), hipHostMallocDefault);
    if (e != hipSuccess) {
        rc = hip_fail(e, "fbp_segment_create");
        fbp_segment_destroy(g);
        return rc;
    }
    rc = ensure_spike_capacity(g, initial_spike_cap(d), 0, g->own_stream);
    if (rc == FBP_OK && hipStreamSynchronize(g->own_stream) != hipSuccess) rc = fail(FBP_EHIP, "allocation failed");
    if (rc) { fbp_segment_destroy(g); return rc; }
    *out = g;
    return FBP_OK;
}

void fbp_segment_destroy(fbp_segment *g)
{
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->own_stream) (void)hipStreamSynchronize(g->own_stream);
    free_state(g);   // stream-ordered frees, then a wait on the segment's stream
    if (g->ev_start) (void)hipEventDestroy(g->ev_start);
    if (g->ev_stop) (void)hipEventDestroy(g->ev_stop);
    if (g->ev_pass) (void)hipEventDestroy(g->ev_pass);
    for (int q = 0; q < 4; q++)
        if (g->ev_stage[q]) (void)hipEventDestroy(g->ev_stage[q]);
    if (g->own_stream) (void)hipStreamDestroy(g->own_stream);
    if (g->h_ctr) (void)hipHostFree(g->h_ctr);
    delete g;
}

int fbp_segment_set_templates(fbp_segment *g, const float *templates, const float *bias, const float *thresholds)
{
    if (!g || !templates || !bias || !thresholds) return fail(FBP_EINVAL, "NULL argument");
    if (int rc = set_device(g)) return rc;
    // channel support of every template: the channels on which it is nonzero
    // (the kernels skip the others, whose terms are exactly zero)
    std::vector<uint32_t> lo(g->d.T), hi(g->d.T);
    uint32_t sup_max = 1;
    for (uint32_t t = 0; t < g->d.T; t++) {
        uint32_t a = g->d.C, b = 0;
        for (uint32_t c = 0; c < g->d.C; c++) {
            const float *row = templates + ((uint64_t)t * g->d.C + c) * g->d.N;
            for (uint32_t j = 0; j < g->d.N; j++)
                if (row[j] != 0.0f) {
                    if (c < a) a = c;
                    b = c + 1;
                    break;
                }
        }
        if (a >= b) a = b = 0;
        lo[t] = a;
        hi[t] = b;
        if (b - a > sup_max) sup_max = b - a;
    }
    g->d.sup_max = sup_max;
    // matrix-core order: templates by first nonzero channel (then index), so a
    // tile of 16 covers few channels; tile ranges are the union of supports
    const uint32_t CT = (g->d.T + 15) / 16;
    std::vector<uint32_t> ord(CT * 16), tlo(CT), thi(CT);
    for (uint32_t t = 0; t < g->d.T; t++) ord[t] = t;
    std::stable_sort(ord.begin(), ord.begin() + g->d.T, [&](uint32_t x, uint32_t y) { return lo[x] < lo[y]; });
    for (uint32_t t = g->d.T; t < CT * 16; t++) ord[t] = g->d.T;
    for (uint32_t ct = 0; ct < CT; ct++) {
        uint32_t a = g->d.C, b = 0;
        for (uint32_t q = ct * 16; q < ct * 16 + 16 && q < g->d.T; q++)
            if (lo[ord[q]] < hi[ord[q]]) {
                a = std::min(a, lo[ord[q]]);
                b = std::max(b, hi[ord[q]]);
            }
        if (a >= b) a = b = 0;
        tlo[ct] = a;
        thi[ct] = b;
    }
    // tiles per correlation block and the widest channel span of such a chunk
    const uint32_t ctb = mfma_ctb_for(CT, g->d.C, sup_max);
    uint32_t span = 0;
    for (uint32_t c0 = 0; c0 < CT; c0 += ctb) {
        uint32_t a = g->d.C, b = 0;
        for (uint32_t ct = c0; ct < CT && ct < c0 + ctb; ct++)
            if (tlo[ct] < thi[ct]) {
                a = std::min(a, tlo[ct]);
                b = std::max(b, thi[ct]);
            }
        if (a < b) span = std::max(span, b - a);
    }
    g->d.corr_ctb = ctb;
    g->d.corr_span = span;
    if (g->cfg.overlap_mode == FBP_OVERLAP_SCREENED && overlap_screened_lds(g->d) > kMaxLds)
        return fail(FBP_EINVAL, "screened overlap needs %zu bytes of LDS (templates span %u channels)",
                    overlap_screened_lds(g->d), sup_max);
    const Dims &d = g->d;
    hipStream_t st = g->own_stream;
    float *tmp_bias = nullptr;
    HIP_TRY(dalloc(&tmp_bias, d.T, st), "bias staging");
    hipError_t e = hipMemcpyAsync(g->s.Tm, templates, (uint64_t)d.T * d.C * d.N * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g->s.sup_lo, lo.data(), (uint64_t)d.T * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g->s.sup_hi, hi.data(), (uint64_t)d.T * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g->s.tord, ord.data(), (uint64_t)CT * 16 * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g->s.tile_lo, tlo.data(), (uint64_t)CT * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g->s.tile_hi, thi.data(), (uint64_t)CT * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(tmp_bias, bias, (uint64_t)d.T * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g->s.thr, thresholds, (uint64_t)d.T * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_transpose_templates(d, g->s.Tm, tmp_bias, g->s.TmT, g->s.bias, st);
    if (e == hipSuccess) e = launch_template_norms(d, g->s, g->cfg.precision, st);
    if (e == hipSuccess) e = launch_pad_templates(d, g->s, st);
    if (e == hipSuccess && g->cfg.overlap_mode == FBP_OVERLAP_SCREENED) e = launch_overlap_tables(d, g->s, st);
    if (e == hipSuccess && g->cfg.precision != FBP_PREC_FP32) e = launch_pack_templates(d, g->s, g->cfg.precision, st);
    dfree(tmp_bias, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "fbp_segment_set_templates");
    g->have_templates = true;
    return FBP_OK;
}

int fbp_segment_set_voltage(fbp_segment *g, const float *voltage, uint64_t row_stride)
{
    if (!g || !voltage) return fail(FBP_EINVAL, "NULL argument");
    const Dims &d = g->d;
    if (row_stride < d.K) return fail(FBP_EINVAL, "row stride %llu below chunk length %u",
                                      (unsigned long long)row_stride, d.K);
    if (int rc = set_device(g)) return rc;
    HIP_TRY(hipMemcpy2DAsync(g->s.V0, (size_t)d.K * 4, voltage, (size_t)row_stride * 4, (size_t)d.K * 4, d.C,
                             hipMemcpyHostToDevice, g->own_stream), "voltage upload");
    HIP_TRY(hipStreamSynchronize(g->own_stream), "voltage upload");
    g->have_voltage = true;
    return FBP_OK;
}

int fbp_segment_set_voltage_device(fbp_segment *g, const float *d_voltage, uint64_t row_stride)
{
    if (!g || !d_voltage) return fail(FBP_EINVAL, "NULL argument");
    const Dims &d = g->d;
    if (row_stride < d.K) return fail(FBP_EINVAL, "row stride below chunk length");
    if (int rc = set_device(g)) return rc;
    HIP_TRY(launch_copy_rows(g->s.V0, d_voltage, d.C, d.K, row_stride, g->own_stream), "voltage copy");
    HIP_TRY(hipStreamSynchronize(g->own_stream), "voltage copy");
    g->have_voltage = true;
    return FBP_OK;
}

}  // extern "C"

namespace {

// Enqueue one pass over the g->nw windows of s.win on g->run_stream
// (bpp.py:532-689), ending with the counter read-back and g->ev_pass.
int issue_pass(fbp_segment *g)
{
    const Dims &d = g->d;
    DeviceState &s = g->s;
    hipStream_t st = g->run_stream;
    const uint32_t nw = g->nw;
    g->passes++;
    if (int rc = ensure_spike_capacity(g, g->total + nw, g->total, st)) return rc;
    HIP_TRY(hipEventRecord(g->ev_stage[0], st), "event");
    if (g->cfg.precision == FBP_PREC_FP32)
        HIP_TRY(launch_corr_fp32(d, s, nw, st), "correlation");               // bpp.py:538-549
    else
        HIP_TRY(launch_corr_screen(d, s, nw, g->cfg.precision, g->coef, g->rq_cap,
                                   g->cfg.debug_flags & FBP_DEBUG_NO_RESCORE, st), "correlation screen");
    HIP_TRY(hipEventRecord(g->ev_stage[1], st), "event");
    HIP_TRY(launch_argmax(d, s, nw, st), "argmax");                            // bpp.py:551-562
    if (g->cfg.do_overlap) {                                                   // bpp.py:565-661
        HIP_TRY(launch_compact_ov(d, s, nw, st), "overlap list");
        HIP_TRY(hipEventRecord(g->ev_stage[2], st), "event");
        if (g->cfg.overlap_mode == FBP_OVERLAP_SCREENED) {
            const uint32_t dbg = g->cfg.debug_flags & (FBP_DEBUG_OVERLAP_FULL | FBP_DEBUG_OVERLAP_GLOBAL |
                                                       FBP_DEBUG_OVERLAP_EXACT | FBP_DEBUG_OVERLAP_ALL_ROWS);
            HIP_TRY(launch_overlap_screened(d, s, nw, dbg, st), "overlap recheck (screened)");
        } else {
            HIP_TRY(launch_overlap_direct(d, s, nw, st), "overlap recheck");
        }
        HIP_TRY(launch_check_overlap(d, s, nw, st), "overlap reassign");
        HIP_TRY(hipEventRecord(g->ev_stage[3], st), "event");
    }
    HIP_TRY(launch_pursuit_add(d, s, nw, st), "pursuit add");                  // bpp.py:663-671
    HIP_TRY(launch_compact_spikes(d, s, nw, g->total, st), "spike list");
    HIP_TRY(launch_pursuit_prune(d, s, nw, st), "pursuit prune");
    HIP_TRY(launch_compact_next(d, s, st), "window list");                     // bpp.py:683-689
    HIP_TRY(hipMemcpyAsync(g->h_ctr, s.ctr, sizeof(Counters), hipMemcpyDeviceToHost, st), "counters");
    HIP_TRY(hipEventRecord(g->ev_pass, st), "event");
    return FBP_OK;
}

// The run is over (converged, capped or failed): final timings and stats.
int finish_run(fbp_segment *g, int rc)
{
    if (rc == FBP_OK || rc == FBP_EPASSES) {
        HIP_TRY(hipEventRecord(g->ev_stop, g->run_stream), "event");
        HIP_TRY(hipEventSynchronize(g->ev_stop), "event");
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, g->ev_start, g->ev_stop);
        g->stats.ms_total = ms;
    }
    g->n_spikes = g->total;
    g->stats.n_passes = g->passes;
    g->stats.n_spikes = g->total;
    g->stats.overlap_combos = g->h_ctr->combos;
    g->stats.screened_combos = g->h_ctr->screened;
    g->stats.corr_pairs = g->h_ctr->corr_pairs;
    // Results stay readable when the pass cap stopped the loop.
    g->ran = (rc == FBP_OK || rc == FBP_EPASSES);
    return rc;
}

}  // namespace

int fbp::segment_begin(fbp_segment *g, hipStream_t stream)
{
    if (!g) return fail(FBP_EINVAL, "NULL segment");
    if (!g->have_templates || !g->have_voltage) return fail(FBP_ESTATE, "templates and voltage must be set first");
    if (int rc = set_device(g)) return rc;
    hipStream_t st = stream ? stream : g->own_stream;
    const Dims &d = g->d;
    DeviceState &s = g->s;
    g->run_stream = st;
    g->ran = false;
    g->n_spikes = 0;
    g->n_pending = 0;
    g->stats = fbp_stats{};
    g->nw = d.W;
    g->total = 0;
    g->passes = 0;
    HIP_TRY(hipEventRecord(g->ev_start, st), "event");
    // Fresh state for the chunk (bpp.py:386-423); L needs no clearing because
    // pass 1 writes every position of it.
    HIP_TRY(hipMemcpyAsync(s.V, s.V0, (uint64_t)d.C * d.K * 4, hipMemcpyDeviceToDevice, st), "voltage restore");
    HIP_TRY(launch_fill_u32((uint32_t *)s.best_L, d.W, 0u, st), "init");
    HIP_TRY(launch_fill_u32(s.best_lab, d.W, 0u, st), "init");
    HIP_TRY(launch_fill_u32(s.best_idx, d.W, 0u, st), "init");
    HIP_TRY(launch_fill_u32(s.ov_idx, d.W, 0u, st), "init");
    HIP_TRY(launch_fill_u32(s.ov_lab, d.W, 0u, st), "init");
    HIP_TRY(launch_fill_u32(s.win_spike, d.W, 0u, st), "init");
    HIP_TRY(launch_fill_u32((uint32_t *)s.wdirty, 2ull * d.W, ~0u, st), "init");   // pass 1 computes everything
    HIP_TRY(launch_fill_u8(s.nc, d.W + 2, 0, st), "init");
    HIP_TRY(launch_fill_u8(s.ov, d.W + 2, 0, st), "init");
    HIP_TRY(launch_iota_u32(s.win, d.W, st), "init");
    HIP_TRY(launch_reset_counters(s.ctr, st), "init");
    return issue_pass(g);
}

hipEvent_t fbp::segment_pass_event(const fbp_segment *g) { return g->ev_pass; }

int fbp::segment_step(fbp_segment *g, int *more)
{
    *more = 0;
    if (int rc = set_device(g)) return rc;
    const Dims &d = g->d;
    DeviceState &s = g->s;
    hipStream_t st = g->run_stream;
    const uint32_t nw = g->nw;
    const uint32_t n_new = g->h_ctr->n_new;
    g->stats.windows_evaluated += nw;
    g->stats.corr_launches++;
    g->stats.corr_positions += (uint64_t)nw * d.N;
    if (g->cfg.precision != FBP_PREC_FP32) g->stats.rescored_positions += g->h_ctr->n_rescore_pos;
    float ms_c = 0.0f, ms_o = 0.0f;
    (void)hipEventElapsedTime(&ms_c, g->ev_stage[0], g->ev_stage[1]);
    g->stats.ms_correlation += ms_c;
    if (g->cfg.do_overlap) {
        g->stats.overlap_windows += g->h_ctr->n_ov;
        g->stats.overlap_launches++;
        (void)hipEventElapsedTime(&ms_o, g->ev_stage[2], g->ev_stage[3]);
        g->stats.ms_overlap += ms_o;
    }
    if (n_new == 0) return finish_run(g, FBP_OK);                              // bpp.py:675-676
    int rc = FBP_OK;
    HIP_TRY(launch_fill_u8(s.ov, d.W + 2, 0, st), "reset");                    // bpp.py:679-680
    HIP_TRY(launch_residual(d, s, g->total, n_new, st), "residual");           // bpp.py:716-731
    HIP_TRY(launch_clear_win_spike(d, s, nw, st), "reset");
    HIP_TRY(launch_fill_u8(s.nc, d.W + 2, 0, st), "reset");                    // bpp.py:699-700
    std::swap(s.win, s.win_next);
    g->total += n_new;
    g->nw = g->h_ctr->n_next;
    if (g->cfg.max_passes && g->passes >= g->cfg.max_passes) {
        g->n_pending = g->nw;   // s.win holds the list the next pass would evaluate
        fail(FBP_EPASSES, "pass cap %u reached", g->passes);
        return finish_run(g, FBP_EPASSES);
    }
    if ((rc = issue_pass(g))) return finish_run(g, rc);
    *more = 1;
    return FBP_OK;
}

uint64_t fbp::segment_device_bytes(const fbp_segment *g) { return g->bytes.load(std::memory_order_relaxed); }

extern "C" {

int fbp_segment_run(fbp_segment *g, void *stream)
{
    int rc = segment_begin(g, (hipStream_t)stream);
    if (rc) return rc;
    for (;;) {
        HIP_TRY(hipEventSynchronize(g->ev_pass), "pass");
        int more = 0;
        rc = segment_step(g, &more);
        if (!more) return rc;
    }
}

int fbp_segment_device_bytes(const fbp_segment *g, uint64_t *bytes)
{
    if (!g || !bytes) return fail(FBP_EINVAL, "NULL argument");
    *bytes = segment_device_bytes(g);
    return FBP_OK;
}

int fbp_segment_num_spikes(const fbp_segment *g, uint64_t *n)
{
    if (!g || !n) return fail(FBP_EINVAL, "NULL argument");
    if (!g->ran) return fail(FBP_ESTATE, "segment has not run");
    *n = g->n_spikes;
    return FBP_OK;
}

int fbp_segment_get_spikes(fbp_segment *g, uint32_t *onsets, uint32_t *labels)
{
    if (!g || !onsets || !labels) return fail(FBP_EINVAL, "NULL argument");
    if (!g->ran) return fail(FBP_ESTATE, "segment has not run");
    if (g->n_spikes == 0) return FBP_OK;
    if (int rc = set_device(g)) return rc;
    HIP_TRY(hipMemcpy(onsets, g->s.sp_idx, g->n_spikes * 4, hipMemcpyDeviceToHost), "spikes");
    HIP_TRY(hipMemcpy(labels, g->s.sp_lab, g->n_spikes * 4, hipMemcpyDeviceToHost), "spikes");
    return FBP_OK;
}

int fbp_segment_get_clips(fbp_segment *g, float *clips)
{
    if (!g || !clips) return fail(FBP_EINVAL, "NULL argument");
    if (!g->ran) return fail(FBP_ESTATE, "segment has not run");
    if (g->n_spikes == 0) return FBP_OK;
    if (int rc = set_device(g)) return rc;
    const uint64_t n = g->n_spikes * g->d.C * g->d.N;
    if (n > g->clips_cap) {
        sfree(g, g->s.clips, g->clips_cap, g->own_stream);
        g->s.clips = nullptr;
        g->clips_cap = 0;
        HIP_TRY(salloc(g, &g->s.clips, n, g->own_stream), "clips");
        g->clips_cap = n;
    }
    HIP_TRY(launch_clips(g->d, g->s, g->n_spikes, g->own_stream), "clips");
    HIP_TRY(hipMemcpyAsync(clips, g->s.clips, n * 4, hipMemcpyDeviceToHost, g->own_stream), "clips");
    HIP_TRY(hipStreamSynchronize(g->own_stream), "clips");
    return FBP_OK;
}

int fbp_segment_get_residual(fbp_segment *g, float *residual)
{
    if (!g || !residual) return fail(FBP_EINVAL, "NULL argument");
    if (!g->ran) return fail(FBP_ESTATE, "segment has not run");
    if (int rc = set_device(g)) return rc;
    HIP_TRY(hipMemcpy(residual, g->s.V, (uint64_t)g->d.C * g->d.K * 4, hipMemcpyDeviceToHost), "residual");
    return FBP_OK;
}

int fbp_segment_get_likelihood(fbp_segment *g, float *likelihood)
{
    if (!g || !likelihood) return fail(FBP_EINVAL, "NULL argument");
    if (!g->ran) return fail(FBP_ESTATE, "segment has not run");
    if (int rc = set_device(g)) return rc;
    HIP_TRY(hipMemcpy(likelihood, g->s.L, (uint64_t)g->d.T * g->d.K * 4, hipMemcpyDeviceToHost), "likelihood");
    return FBP_OK;
}

int fbp_segment_get_window_list(fbp_segment *g, uint32_t *windows, uint64_t capacity, uint64_t *n)
{
    if (!g || !n) return fail(FBP_EINVAL, "NULL argument");
    if (!g->ran) return fail(FBP_ESTATE, "segment has not run");
    *n = g->n_pending;
    if (g->n_pending == 0) return FBP_OK;
    if (!windows || capacity < g->n_pending)
        return fail(FBP_EINVAL, "window list of %llu entries does not fit in %llu",
                    (unsigned long long)g->n_pending, (unsigned long long)capacity);
    if (int rc = set_device(g)) return rc;
    HIP_TRY(hipMemcpy(windows, g->s.win, g->n_pending * 4, hipMemcpyDeviceToHost), "window list");
    return FBP_OK;
}

int fbp_segment_get_stats(const fbp_segment *g, fbp_stats *stats)
{
    if (!g || !stats) return fail(FBP_EINVAL, "NULL argument");
    *stats = g->stats;
    return FBP_OK;
}

int fbp_segment_voltage_device(const fbp_segment *g, const float **d_voltage, uint64_t *row_stride)
{
    if (!g || !d_voltage || !row_stride) return fail(FBP_EINVAL, "NULL argument");
    if (!g->have_voltage) return fail(FBP_ESTATE, "segment has no voltage");
    *d_voltage = g->s.V0;
    *row_stride = g->d.K;
    return FBP_OK;
}

// get_clips from device (d_voltage) or host (h_voltage) float32 rows.
static int get_clips_impl(int device, const float *d_voltage, const float *h_voltage, uint64_t row_stride,
                          uint32_t n_chans, uint64_t n_samples, const int64_t *event_indices, uint64_t n_events,
                          const uint32_t *channels, uint32_t n_sel, int32_t w0, int32_t w1, int check_valid,
                          float *clips, uint8_t *valid, uint64_t *n_valid)
{
    if (!n_valid || (n_events && (!event_indices || !valid)) || (n_sel && !channels))
        return fail(FBP_EINVAL, "NULL argument");
    if (w1 <= w0) return fail(FBP_EINVAL, "empty clip window [%d, %d)", w0, w1);
    if (row_stride < n_samples) return fail(FBP_EINVAL, "row stride %llu below the sample count %llu",
                                            (unsigned long long)row_stride, (unsigned long long)n_samples);
    for (uint32_t j = 0; j < n_sel; j++)
        if (channels[j] >= n_chans) return fail(FBP_EINVAL, "channel %u out of range (%u channels)", channels[j], n_chans);
    *n_valid = 0;
    if (n_events == 0) return FBP_OK;
    // segment_parallel.py:462-484: invalid events are trimmed from the front
    // and the back only; the events between must lie inside the data.
    uint64_t lo = 0, hi = n_events;   // valid events are [lo, hi)
    if (check_valid) {
        while (lo < n_events && event_indices[lo] + w0 < 0) lo++;
        while (hi > lo && event_indices[hi - 1] + w1 >= (int64_t)n_samples) hi--;
    }
    for (uint64_t k = 0; k < n_events; k++) valid[k] = (k >= lo && k < hi) ? 1 : 0;
    for (uint64_t k = lo; k < hi; k++)
        if (event_indices[k] + w0 < 0 || event_indices[k] + w1 > (int64_t)n_samples)
            return fail(FBP_EINVAL, "event %llu (index %lld) has a clip outside the data; event indices must be "
                        "ascending", (unsigned long long)k, (long long)event_indices[k]);
    const uint64_t n = hi - lo;
    *n_valid = n;
    const uint32_t W = (uint32_t)(w1 - w0);
    const uint64_t width = (uint64_t)n_sel * W;
    if (n == 0 || width == 0) return FBP_OK;
    if (!clips) return fail(FBP_EINVAL, "NULL argument");
    if (!d_voltage && !h_voltage) return fail(FBP_EINVAL, "NULL voltage");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
    int64_t *d_idx = nullptr;
    uint32_t *d_chan = nullptr;
    float *d_clips = nullptr, *d_rows = nullptr;
    std::vector<uint32_t> chan(channels, channels + n_sel);
    hipError_t e = hipSuccess;
    if (h_voltage) {   // upload the selected channels' rows; they become rows 0..n_sel-1
        e = dalloc(&d_rows, (uint64_t)n_sel * n_samples, st);
        for (uint32_t j = 0; j < n_sel && e == hipSuccess; j++) {
            e = hipMemcpyAsync(d_rows + (uint64_t)j * n_samples, h_voltage + (uint64_t)channels[j] * row_stride,
                               n_samples * 4, hipMemcpyHostToDevice, st);
            chan[j] = j;
        }
        d_voltage = d_rows;
        row_stride = n_samples;
    }
    if (e == hipSuccess) e = dalloc(&d_idx, n, st);
    if (e == hipSuccess) e = dalloc(&d_chan, n_sel, st);
    if (e == hipSuccess) e = dalloc(&d_clips, n * width, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_idx, event_indices + lo, n * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_chan, chan.data(), (uint64_t)n_sel * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_get_clips(d_voltage, row_stride, d_idx, n, d_chan, n_sel, w0, W, d_clips, st);
    if (e == hipSuccess) e = hipMemcpyAsync(clips, d_clips, n * width * 4, hipMemcpyDeviceToHost, st);
    dfree(d_idx, st);
    dfree(d_chan, st);
    dfree(d_clips, st);
    dfree(d_rows, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (e == hipSuccess) e = es;
    HIP_TRY(e, "get_clips");
    return FBP_OK;
}

int fbp_get_clips(int device, const float *d_voltage, uint64_t row_stride, uint32_t n_chans, uint64_t n_samples,
                  const int64_t *event_indices, uint64_t n_events, const uint32_t *channels, uint32_t n_sel,
                  int32_t w0, int32_t w1, int check_valid, float *clips, uint8_t *valid, uint64_t *n_valid)
{
    if (n_events && !d_voltage) return fail(FBP_EINVAL, "NULL voltage");
    return get_clips_impl(device, d_voltage, nullptr, row_stride, n_chans, n_samples, event_indices, n_events,
                          channels, n_sel, w0, w1, check_valid, clips, valid, n_valid);
}

int fbp_get_clips_host(int device, const float *voltage, uint64_t row_stride, uint32_t n_chans, uint64_t n_samples,
                       const int64_t *event_indices, uint64_t n_events, const uint32_t *channels, uint32_t n_sel,
                       int32_t w0, int32_t w1, int check_valid, float *clips, uint8_t *valid, uint64_t *n_valid)
{
    if (n_events && !voltage) return fail(FBP_EINVAL, "NULL voltage");
    return get_clips_impl(device, nullptr, voltage, row_stride, n_chans, n_samples, event_indices, n_events,
                          channels, n_sel, w0, w1, check_valid, clips, valid, n_valid);
}

int fbp_overlap_template_residuals(int device, const float *templates, uint32_t n_templates, uint32_t n_chans,
                                   uint32_t n_samples_per_chan, int32_t n_pre, int32_t n_post,
                                   const uint32_t *triples, uint64_t n_triples, double *best_residual,
                                   uint32_t *best_combo)
{
    if (n_triples && (!templates || !triples || !best_residual || !best_combo)) return fail(FBP_EINVAL, "NULL argument");
    if (n_pre < 0 || n_post < 0) return fail(FBP_EINVAL, "negative shift range (%d, %d)", n_pre, n_post);
    const uint64_t CN = (uint64_t)n_chans * n_samples_per_chan;
    if (CN == 0 || CN >= (1ull << 24)) return fail(FBP_EINVAL, "template length %llu out of range", (unsigned long long)CN);
    if (3 * CN * sizeof(float) > kMaxLds)
        return fail(FBP_EINVAL, "three templates of %llu samples do not fit in LDS", (unsigned long long)CN);
    for (uint64_t k = 0; k < 3 * n_triples; k++)
        if (triples[k] >= n_templates) return fail(FBP_EINVAL, "template %u out of range", triples[k]);
    const uint64_t S = (uint64_t)n_pre + n_post + 1;
    if (S * S >= (1ull << 32)) return fail(FBP_EINVAL, "shift range too large");
    if (n_triples == 0) return FBP_OK;
    std::vector<uint32_t> prog;
    if (numpy_sum_program(CN, prog) > 16) return fail(FBP_EINVAL, "summation tree too deep");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
    float *d_tm = nullptr;
    uint32_t *d_tri = nullptr, *d_prog = nullptr, *d_combo = nullptr;
    double *d_res = nullptr;
    hipError_t e = dalloc(&d_tm, (uint64_t)n_templates * CN, st);
    if (e == hipSuccess) e = dalloc(&d_tri, 3 * n_triples, st);
    if (e == hipSuccess) e = dalloc(&d_prog, prog.size(), st);
    if (e == hipSuccess) e = dalloc(&d_res, n_triples, st);
    if (e == hipSuccess) e = dalloc(&d_combo, n_triples, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_tm, templates, (uint64_t)n_templates * CN * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tri, triples, 3 * n_triples * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_prog, prog.data(), prog.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_overlap_template_residuals(d_tm, n_chans, n_samples_per_chan, n_pre, (uint32_t)S, d_tri, n_triples,
                                              d_prog, (uint32_t)prog.size(), d_res, d_combo, st);
    if (e == hipSuccess) e = hipMemcpyAsync(best_residual, d_res, n_triples * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(best_combo, d_combo, n_triples * 4, hipMemcpyDeviceToHost, st);
    dfree(d_tm, st);
    dfree(d_tri, st);
    dfree(d_prog, st);
    dfree(d_res, st);
    dfree(d_combo, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (e == hipSuccess) e = es;
    HIP_TRY(e, "overlap template residuals");
    return FBP_OK;
}

int fbp_clip_template_sse(int device, const float *clips, uint64_t n, uint32_t width, const uint32_t *template_index,
                          const float *templates, uint32_t n_templates, float *sse)
{
    if (n && (!clips || !template_index || !templates || !sse)) return fail(FBP_EINVAL, "NULL argument");
    if (width == 0 || width >= (1u << 24)) return fail(FBP_EINVAL, "clip width %u out of range", width);
    for (uint64_t k = 0; k < n; k++)
        if (template_index[k] >= n_templates)
            return fail(FBP_EINVAL, "template index %u out of range (%u templates)", template_index[k], n_templates);
    if (n == 0) return FBP_OK;
    std::vector<uint32_t> prog;
    if (numpy_sum_program(width, prog) > 16) return fail(FBP_EINVAL, "summation tree too deep");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
    float *d_clips = nullptr, *d_tm = nullptr, *d_sse = nullptr;
    uint32_t *d_idx = nullptr, *d_prog = nullptr;
    hipError_t e = dalloc(&d_clips, n * width, st);
    if (e == hipSuccess) e = dalloc(&d_tm, (uint64_t)n_templates * width, st);
    if (e == hipSuccess) e = dalloc(&d_idx, n, st);
    if (e == hipSuccess) e = dalloc(&d_prog, prog.size(), st);
    if (e == hipSuccess) e = dalloc(&d_sse, n, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_clips, clips, n * width * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_tm, templates, (uint64_t)n_templates * width * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_idx, template_index, n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_prog, prog.data(), prog.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_clip_sse(d_clips, n, width, d_idx, d_tm, d_prog, (uint32_t)prog.size(), d_sse, st);
    if (e == hipSuccess) e = hipMemcpyAsync(sse, d_sse, n * 4, hipMemcpyDeviceToHost, st);
    dfree(d_clips, st);
    dfree(d_tm, st);
    dfree(d_idx, st);
    dfree(d_prog, st);
    dfree(d_sse, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (e == hipSuccess) e = es;
    HIP_TRY(e, "clip template sse");
    return FBP_OK;
}

int fbp_channel_covariance(int device, const float *clips, uint64_t n_clips, uint32_t n_chans, uint32_t N,
                           const uint64_t *rows, uint64_t n_rows, double *cov)
{
    if (!clips || !cov) return fail(FBP_EINVAL, "NULL argument");
    if (N == 0 || N > kMaxCovSamples) return fail(FBP_EINVAL, "samples per channel %u out of range (1..%u)", N,
                                                  kMaxCovSamples);
    if (!rows) n_rows = n_clips;
    if (n_rows == 0) return fail(FBP_EINVAL, "no clip rows to compute a covariance from");
    if (rows)
        for (uint64_t k = 0; k < (uint64_t)n_chans * n_rows; k++)
            if (rows[k] >= n_clips)
                return fail(FBP_EINVAL, "row %llu out of range (%llu clips)", (unsigned long long)rows[k],
                            (unsigned long long)n_clips);
    if (n_chans == 0) return FBP_OK;
    const uint64_t width = (uint64_t)n_chans * N;
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
    float *d_clips = nullptr;
    uint64_t *d_rows = nullptr;
    double *d_cov = nullptr, *d_colpart = nullptr, *d_part = nullptr;
    const uint64_t ncov = (uint64_t)n_chans * N * N;
    uint64_t slice = 0;
    const uint32_t R = channel_covariance_slices(n_rows, &slice);
    hipError_t e = dalloc(&d_clips, n_clips * width, st);
    if (e == hipSuccess && rows) e = dalloc(&d_rows, (uint64_t)n_chans * n_rows, st);
    if (e == hipSuccess) e = dalloc(&d_cov, ncov, st);
    if (e == hipSuccess) e = dalloc(&d_colpart, (uint64_t)n_chans * R * N, st);
    if (e == hipSuccess) e = dalloc(&d_part, ncov * R, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_clips, clips, n_clips * width * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && rows)
        e = hipMemcpyAsync(d_rows, rows, (uint64_t)n_chans * n_rows * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_channel_covariance(d_clips, width, n_clips, d_rows, n_rows, n_chans, N, d_colpart, d_part, d_cov,
                                      st);
    if (e == hipSuccess) e = hipMemcpyAsync(cov, d_cov, ncov * 8, hipMemcpyDeviceToHost, st);
    dfree(d_clips, st);
    dfree(d_rows, st);
    dfree(d_cov, st);
    dfree(d_colpart, st);
    dfree(d_part, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (e == hipSuccess) e = es;
    HIP_TRY(e, "channel covariance");
    return FBP_OK;
}

int fbp_template_variances(int device, const double *templates, uint32_t n_templates, uint32_t n_chans, uint32_t N,
                           const double *cov, double *var)
{
    if (n_templates && (!templates || !cov || !var)) return fail(FBP_EINVAL, "NULL argument");
    if (N == 0 || N > kMaxCovSamples) return fail(FBP_EINVAL, "samples per channel %u out of range (1..%u)", N,
                                                  kMaxCovSamples);
    if (n_templates == 0) return FBP_OK;
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
    double *d_tm = nullptr, *d_cov = nullptr, *d_var = nullptr;
    const uint64_t ntm = (uint64_t)n_templates * n_chans * N, ncov = (uint64_t)n_chans * N * N;
    hipError_t e = dalloc(&d_tm, ntm ? ntm : 1, st);
    if (e == hipSuccess) e = dalloc(&d_cov, ncov ? ncov : 1, st);
    if (e == hipSuccess) e = dalloc(&d_var, n_templates, st);
    if (e == hipSuccess && ntm) e = hipMemcpyAsync(d_tm, templates, ntm * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && ncov) e = hipMemcpyAsync(d_cov, cov, ncov * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_template_variance(d_tm, n_templates, n_chans, N, d_cov, d_var, st);
    if (e == hipSuccess) e = hipMemcpyAsync(var, d_var, (uint64_t)n_templates * 8, hipMemcpyDeviceToHost, st);
    dfree(d_tm, st);
    dfree(d_cov, st);
    dfree(d_var, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (e == hipSuccess) e = es;
    HIP_TRY(e, "template variances");
    return FBP_OK;
}

int fbp_template_pair_lags(int device, const double *templates, uint32_t n_templates, uint32_t n_chans, uint32_t N,
                           const uint32_t *pairs, uint64_t n_pairs, int32_t lag_lo, int32_t lag_hi, double *xcorr,
                           double *aligned)
{
    if (n_pairs == 0) return FBP_OK;
    if (!templates || !pairs || !xcorr) return fail(FBP_EINVAL, "NULL argument");
    if (n_chans == 0 || N == 0) return fail(FBP_EINVAL, "empty templates");
    const int64_t L = (int64_t)n_chans * N;
    if (lag_hi < lag_lo || lag_lo <= -L || lag_hi >= L)
        return fail(FBP_EINVAL, "lags %d..%d outside -(L-1)..L-1 for L = %lld", lag_lo, lag_hi, (long long)L);
    if (n_pairs > 0x7fffffffull) return fail(FBP_EINVAL, "too many pairs");
    for (uint64_t p = 0; p < 2 * n_pairs; p++)
        if (pairs[p] >= n_templates) return fail(FBP_EINVAL, "pair entry %u out of range (%u templates)", pairs[p],
                                                 n_templates);
    const uint32_t n_lags = (uint32_t)(lag_hi - lag_lo + 1);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
    double *d_tm = nullptr, *d_x = nullptr, *d_al = nullptr;
    uint32_t *d_pairs = nullptr;
    const uint64_t ntm = (uint64_t)n_templates * L, nout = n_pairs * n_lags;
    hipError_t e = dalloc(&d_tm, ntm, st);
    if (e == hipSuccess) e = dalloc(&d_pairs, 2 * n_pairs, st);
    if (e == hipSuccess) e = dalloc(&d_x, nout, st);
    if (e == hipSuccess && aligned) e = dalloc(&d_al, nout, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tm, templates, ntm * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_pairs, pairs, 2 * n_pairs * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_template_pair_lags(d_tm, n_chans, N, d_pairs, n_pairs, lag_lo, n_lags, d_x, d_al, st);
    if (e == hipSuccess) e = hipMemcpyAsync(xcorr, d_x, nout * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && aligned) e = hipMemcpyAsync(aligned, d_al, nout * 8, hipMemcpyDeviceToHost, st);
    dfree(d_tm, st);
    dfree(d_pairs, st);
    dfree(d_x, st);
    dfree(d_al, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (e == hipSuccess) e = es;
    HIP_TRY(e, "template pair lags");
    return FBP_OK;
}

}  // extern "C"

template <class F>
static int unit_mean_clips_impl(int device, const F *clips, uint64_t n, uint32_t width, const uint32_t *labels,
                                uint32_t n_units, F *means, uint64_t *counts)
{
    if ((n && (!clips || !labels)) || (n_units && width && !means)) return fail(FBP_EINVAL, "NULL argument");
    // rows of each unit in input order (counting sort)
    std::vector<uint64_t> off(n_units + 1, 0);
    for (uint64_t k = 0; k < n; k++) {
        if (labels[k] >= n_units) return fail(FBP_EINVAL, "label %u out of range (%u units)", labels[k], n_units);
        off[labels[k] + 1]++;
    }
    for (uint32_t u = 0; u < n_units; u++) off[u + 1] += off[u];
    if (counts)
        for (uint32_t u = 0; u < n_units; u++) counts[u] = off[u + 1] - off[u];
    if (n_units == 0 || width == 0) return FBP_OK;
    std::vector<uint32_t> rows(n ? n : 1);
    {
        std::vector<uint64_t> pos(off.begin(), off.end() - 1);
        for (uint64_t k = 0; k < n; k++) {
            if (k > 0xffffffffull) return fail(FBP_EINVAL, "more than 2^32 clips");
            rows[pos[labels[k]]++] = (uint32_t)k;
        }
    }
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
    F *d_clips = nullptr, *d_means = nullptr;
    uint32_t *d_rows = nullptr;
    uint64_t *d_off = nullptr;
    hipError_t e = dalloc(&d_clips, n * width, st);
    if (e == hipSuccess) e = dalloc(&d_rows, n, st);
    if (e == hipSuccess) e = dalloc(&d_off, (uint64_t)n_units + 1, st);
    if (e == hipSuccess) e = dalloc(&d_means, (uint64_t)n_units * width, st);
    if (e == hipSuccess && n) e = hipMemcpyAsync(d_clips, clips, n * width * sizeof(F), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n) e = hipMemcpyAsync(d_rows, rows.data(), n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_off, off.data(), ((uint64_t)n_units + 1) * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_unit_means(d_clips, width, d_rows, d_off, n_units, d_means, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(means, d_means, (uint64_t)n_units * width * sizeof(F), hipMemcpyDeviceToHost, st);
    dfree(d_clips, st);
    dfree(d_rows, st);
    dfree(d_off, st);
    dfree(d_means, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (e == hipSuccess) e = es;
    HIP_TRY(e, "unit means");
    return FBP_OK;
}

extern "C" {

int fbp_unit_mean_clips(int device, const float *clips, uint64_t n, uint32_t width, const uint32_t *labels,
                        uint32_t n_units, float *means, uint64_t *counts)
{
    return unit_mean_clips_impl(device, clips, n, width, labels, n_units, means, counts);
}

int fbp_unit_mean_clips_f64(int device, const double *clips, uint64_t n, uint32_t width, const uint32_t *labels,
                            uint32_t n_units, double *means, uint64_t *counts)
{
    return unit_mean_clips_impl(device, clips, n, width, labels, n_units, means, counts);
}

}  // extern "C"
