// The inputs binary pursuit's thresholds are made from: the per-channel noise
// covariance of the residual clips (full_binary_pursuit.py:245-252, np.cov with
// ddof=0 over a sample of clip rows) and each template's likelihood variance
// sum_c t_c^T cov_c t_c (neuron_separability.py:126-135).  Both are small
// float64 reductions; the kernels keep one fixed summation order per output so
// repeated calls agree bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/fbp.h"
#include "fbp_internal.h"

namespace fbp {
namespace {

constexpr uint32_t kCovThreads = 256;
constexpr uint32_t kCovPairs = 8;                           // (i, j) entries per thread
constexpr uint32_t kCovChunk = kCovThreads * kCovPairs;     // entries per block
constexpr uint32_t kCovRows = 32;                           // clip rows staged per step

// cov[c][i][j] = sum_r (x_ri - m_i)(x_rj - m_j) / n over the selected rows r
// (rows[c * n_rows + r], or r itself when rows is null), x_rj = clips[row][c*N + j]
// as float64, m_j the float64 mean of column j.  The rows are cut into R
// slices of `slice` rows so that C * chunks * R blocks share the work: column
// sums per slice, then per (channel, chunk of N*N entries, slice) the centred
// products over the slice's rows, then per entry the slices' sums in slice
// order.  Every sum has one fixed order, so repeated calls agree bit for bit;
// entries with i > j are written from their mirror, so every matrix is
// exactly symmetric.  A row index >= n_clips (possible only when the rows
// come from device memory the host has not checked) is not read: it turns
// its channel's sums into NaN.
__global__ __launch_bounds__(kMaxCovSamples) void k_cov_colsum(const float *__restrict__ clips, uint64_t width,
                                                               uint64_t n_clips, const uint64_t *__restrict__ rows,
                                                               uint64_t n_rows, uint32_t N, uint64_t slice,
                                                               double *__restrict__ part)
{
    const uint32_t c = blockIdx.x, r = blockIdx.y, R = gridDim.y, j = threadIdx.x;
    if (j >= N) return;
    const uint64_t *rc = rows ? rows + (uint64_t)c * n_rows : nullptr;
    const uint64_t b = (uint64_t)r * slice, e = b + slice < n_rows ? b + slice : n_rows;
    const float *col = clips + (uint64_t)c * N + j;
    double s = 0.0;
    for (uint64_t k = b; k < e; k++) {
        const uint64_t row = rc ? rc[k] : k;
        s += row < n_clips ? (double)col[row * width] : __builtin_nan("");
    }
    part[((uint64_t)c * R + r) * N + j] = s;
}

__global__ __launch_bounds__(kCovThreads) void k_cov_products(const float *__restrict__ clips, uint64_t width,
                                                              uint64_t n_clips, const uint64_t *__restrict__ rows,
                                                              uint64_t n_rows,
                                                              uint32_t N, uint64_t slice,
                                                              const double *__restrict__ colpart,
                                                              double *__restrict__ part)
{
    __shared__ double mean[kMaxCovSamples];
    __shared__ double xs[kCovRows][kMaxCovSamples];
    const uint32_t c = blockIdx.x, r = blockIdx.z, R = gridDim.z;
    const uint64_t *rc = rows ? rows + (uint64_t)c * n_rows : nullptr;
    const uint64_t col0 = (uint64_t)c * N;
    for (uint32_t j = threadIdx.x; j < N; j += blockDim.x) {
        double s = 0.0;
        for (uint32_t q = 0; q < R; q++) s += colpart[((uint64_t)c * R + q) * N + j];
        mean[j] = s / (double)n_rows;
    }
    const uint64_t b = (uint64_t)r * slice, e = b + slice < n_rows ? b + slice : n_rows;
    const uint32_t p0 = blockIdx.y * kCovChunk + threadIdx.x;
    double acc[kCovPairs];
#pragma unroll
    for (uint32_t q = 0; q < kCovPairs; q++) acc[q] = 0.0;
    for (uint64_t r0 = b; r0 < e; r0 += kCovRows) {
        const uint32_t nr = e - r0 < kCovRows ? (uint32_t)(e - r0) : kCovRows;
        __syncthreads();   // means ready; the previous tile's reads are done
        for (uint32_t x = threadIdx.x; x < nr * N; x += blockDim.x) {
            const uint32_t rr = x / N, j = x - rr * N;
            const uint64_t row = rc ? rc[r0 + rr] : r0 + rr;
            xs[rr][j] = (row < n_clips ? (double)clips[row * width + col0 + j] : __builtin_nan("")) - mean[j];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kCovPairs; q++) {
            const uint32_t p = p0 + q * kCovThreads;
            if (p >= N * N) break;
            const uint32_t i = p / N, j = p - i * N;
            if (i > j) continue;
            double a = acc[q];
            for (uint32_t rr = 0; rr < nr; rr++) a = __builtin_fma(xs[rr][i], xs[rr][j], a);
            acc[q] = a;
        }
    }
    double *dst = part + ((uint64_t)c * R + r) * N * N;
#pragma unroll
    for (uint32_t q = 0; q < kCovPairs; q++) {
        const uint32_t p = p0 + q * kCovThreads;
        if (p >= N * N) break;
        dst[p] = acc[q];
    }
}

__global__ void k_cov_final(uint32_t C, uint32_t N, uint32_t R, uint64_t n_rows, const double *__restrict__ part,
                            double *__restrict__ cov)
{
    const uint64_t total = (uint64_t)C * N * N;
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total;
         x += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = (uint32_t)(x / ((uint64_t)N * N)), p = (uint32_t)(x - (uint64_t)c * N * N);
        const uint32_t i = p / N, j = p - i * N;
        if (i > j) continue;
        double s = 0.0;
        for (uint32_t q = 0; q < R; q++) s += part[((uint64_t)c * R + q) * N * N + p];
        const double v = s / (double)n_rows;
        cov[(uint64_t)c * N * N + p] = v;
        cov[(uint64_t)c * N * N + (uint64_t)j * N + i] = v;
    }
}

// var[t] = sum over channels c (ascending) of (t_c^T cov_c) t_c, with
// u_j = sum_i t_ci cov_c[i][j] (i ascending) and u . t_c summed over j
// ascending: the evaluation order of (t[None, :] @ cov) @ t[:, None].
__global__ __launch_bounds__(kCovThreads) void k_template_variance(const double *__restrict__ Tm, uint32_t C,
                                                                   uint32_t N, const double *__restrict__ cov,
                                                                   double *__restrict__ var)
{
    __shared__ double tc[kMaxCovSamples];
    __shared__ double u[kMaxCovSamples];
    const uint32_t t = blockIdx.x;
    double total = 0.0;
    for (uint32_t c = 0; c < C; c++) {
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < N; j += blockDim.x) tc[j] = Tm[((uint64_t)t * C + c) * N + j];
        __syncthreads();
        const double *S = cov + (uint64_t)c * N * N;
        for (uint32_t j = threadIdx.x; j < N; j += blockDim.x) {
            double s = 0.0;
            for (uint32_t i = 0; i < N; i++) s = __builtin_fma(tc[i], S[(uint64_t)i * N + j], s);
            u[j] = s;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double q = 0.0;
            for (uint32_t j = 0; j < N; j++) q = __builtin_fma(u[j], tc[j], q);
            total += q;
        }
    }
    if (threadIdx.x == 0) var[t] = total;
}

// Template pairs at a range of lags (consolidate.py:9-71 as called by
// neuron_separability.py:192-196).  With a = templates[pairs[2p]] and
// b = templates[pairs[2p+1]], flat vectors of L = C*N samples, and
// lag = lag_lo + k:
//   xcorr[p][k]   = sum_i a[i + lag] * b[i]  over 0 <= i, i + lag < L
//                   (np.correlate(a, b, 'full')[lag + L - 1]),
//   aligned[p][k] = the same sum over the i whose i and i + lag lie on the
//                   same channel (the dot product of the two templates after
//                   optimal_align_templates has shifted and truncated every
//                   channel by `lag`).
// One wave per (pair, lag): its lanes take i = i0 + lane + 64m in turn, and
// the 64 partial sums are added by a fixed butterfly, so each output has one
// summation order.  Templates are read from global memory (a pair's two
// rows are reused by every lag of the block and stay in L2).
constexpr uint32_t kLagWaves = 4;

__global__ __launch_bounds__(kLagWaves * 64) void k_template_pair_lags(const double *__restrict__ Tm, uint32_t C,
                                                                       uint32_t N, const uint32_t *__restrict__ pairs,
                                                                       int32_t lag_lo, uint32_t n_lags,
                                                                       double *__restrict__ xcorr,
                                                                       double *__restrict__ aligned)
{
    const uint64_t p = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const int64_t L = (int64_t)C * N;
    const double *a = Tm + (uint64_t)pairs[2 * p] * L;
    const double *b = Tm + (uint64_t)pairs[2 * p + 1] * L;
    for (uint32_t k = blockIdx.y * kLagWaves + wave; k < n_lags; k += gridDim.y * kLagWaves) {
        const int64_t lag = (int64_t)lag_lo + k;
        const int64_t i0 = lag < 0 ? -lag : 0, i1 = lag < 0 ? L : L - lag;
        double x = 0.0, al = 0.0;
        for (int64_t i = i0 + lane; i < i1; i += 64) {
            const double prod = a[i + lag] * b[i];
            x += prod;
            if ((i + lag) / N == i / N) al += prod;
        }
        for (int off = 32; off > 0; off >>= 1) {
            x += __shfl_xor(x, off);
            al += __shfl_xor(al, off);
        }
        if (lane == 0) {
            xcorr[p * n_lags + k] = x;
            if (aligned) aligned[p * n_lags + k] = al;
        }
    }
}

}  // namespace

hipError_t launch_template_pair_lags(const double *Tm, uint32_t C, uint32_t N, const uint32_t *pairs,
                                     uint64_t n_pairs, int32_t lag_lo, uint32_t n_lags, double *xcorr,
                                     double *aligned, hipStream_t st)
{
    if (n_pairs == 0 || n_lags == 0) return hipSuccess;
    if (n_pairs > 0x7fffffffull) return hipErrorInvalidValue;
    // about 64 lag waves per block row, at most 1024 blocks along the lags
    uint32_t gy = (n_lags + kLagWaves * 16 - 1) / (kLagWaves * 16);
    if (gy > 1024) gy = 1024;
    hipLaunchKernelGGL(k_template_pair_lags, dim3((uint32_t)n_pairs, gy), dim3(kLagWaves * 64), 0, st, Tm, C, N,
                       pairs, lag_lo, n_lags, xcorr, aligned);
    return hipGetLastError();
}

uint32_t channel_covariance_slices(uint64_t n_rows, uint64_t *slice)
{
    // slices of at least 256 rows, at most 64 of them
    uint64_t len = (n_rows + 63) / 64;
    if (len < 256) len = 256;
    *slice = len;
    return (uint32_t)((n_rows + len - 1) / len);
}

hipError_t launch_channel_covariance(const float *clips, uint64_t width, uint64_t n_clips, const uint64_t *rows,
                                     uint64_t n_rows, uint32_t C, uint32_t N, double *colpart, double *part,
                                     double *cov, hipStream_t st)
{
    if (C == 0 || N == 0 || n_rows == 0) return hipSuccess;
    if (N > kMaxCovSamples) return hipErrorInvalidValue;
    uint64_t slice = 0;
    const uint32_t R = channel_covariance_slices(n_rows, &slice);
    const uint32_t chunks = (N * N + kCovChunk - 1) / kCovChunk;
    hipLaunchKernelGGL(k_cov_colsum, dim3(C, R), dim3(kMaxCovSamples), 0, st, clips, width, n_clips, rows, n_rows, N,
                       slice, colpart);
    hipLaunchKernelGGL(k_cov_products, dim3(C, chunks, R), dim3(kCovThreads), 0, st, clips, width, n_clips, rows,
                       n_rows, N, slice, (const double *)colpart, part);
    const uint64_t total = (uint64_t)C * N * N;
    const uint64_t blocks = (total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096;
    hipLaunchKernelGGL(k_cov_final, dim3((uint32_t)blocks), dim3(256), 0, st, C, N, R, n_rows, (const double *)part,
                       cov);
    return hipGetLastError();
}

hipError_t launch_template_variance(const double *Tm, uint32_t T, uint32_t C, uint32_t N, const double *cov,
                                    double *var, hipStream_t st)
{
    if (T == 0) return hipSuccess;
    if (N > kMaxCovSamples) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_template_variance, dim3(T), dim3(kCovThreads), 0, st, Tm, C, N, cov, var);
    return hipGetLastError();
}

}  // namespace fbp

// ---------------------------------------------------------------------------
// fbp_channel_covariance_noalloc: the covariance entry point for callers that
// cannot allocate device memory (code under graph capture, serving loops).
// Inputs and outputs are the caller's device memory; the per-slice partial
// sums live in `scratch`, a fixed array of this library's code object that
// every GPU holds from the moment the library is loaded on it until the
// process ends.  Its host-side name is not a device address: the address is
// looked up once per device (hipGetSymbolAddress) and only that is used.
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kScratchBytes = 32ull << 20;   // 32 MiB per GPU
constexpr int kMaxScratchDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on `device` (current); looked up on the first
// call for that device.
hipError_t scratch_address(int device, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = (double *)p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" int fbp_channel_covariance_noalloc(int device, const float *d_clips, uint64_t n_clips, uint32_t n_chans,
                                              uint32_t N, const uint64_t *d_rows, uint64_t n_rows, double *d_cov,
                                              void *stream)
{
    using namespace fbp;
    if (!d_clips || !d_cov) return set_error(FBP_EINVAL, "NULL argument");
    if (N == 0 || N > kMaxCovSamples)
        return set_error(FBP_EINVAL, "samples per channel %u out of range (1..%u)", N, kMaxCovSamples);
    if (!d_rows) n_rows = n_clips;
    if (n_rows == 0) return set_error(FBP_EINVAL, "no clip rows to compute a covariance from");
    if (n_chans == 0) return FBP_OK;
    uint64_t slice = 0;
    const uint64_t R = channel_covariance_slices(n_rows, &slice);
    const uint64_t colpart = ((uint64_t)n_chans * R * N + 31) / 32 * 32;   // doubles, 256-byte aligned
    const uint64_t need = (colpart + (uint64_t)n_chans * R * N * N) * sizeof(double);
    if (need > kScratchBytes)
        return set_error(FBP_EINVAL,
                         "covariance of %u channels x %u samples over %llu row slices needs %llu bytes of scratch; "
                         "the allocation-free entry point has %llu (use fbp_channel_covariance)",
                         n_chans, N, (unsigned long long)R, (unsigned long long)need,
                         (unsigned long long)kScratchBytes);
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0) return set_error(FBP_ENODEV, "No GPU available");
    if (device < 0 || device >= n_dev || device >= kMaxScratchDevices)
        return set_error(FBP_EINVAL, "device %d out of range", device);
    hipError_t e = hipSetDevice(device);
    double *base = nullptr;
    if (e == hipSuccess) e = scratch_address(device, &base);
    if (e == hipSuccess)
        e = launch_channel_covariance(d_clips, (uint64_t)n_chans * N, n_clips, d_rows, n_rows, n_chans, N, base,
                                      base + colpart, d_cov, (hipStream_t)stream);
    if (e != hipSuccess) return set_error(FBP_EHIP, "channel covariance (no allocation): %s", hipGetErrorString(e));
    return FBP_OK;
}

