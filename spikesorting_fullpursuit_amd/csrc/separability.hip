// The inputs binary pursuit's thresholds are made from: the per-channel noise
// covariance of the residual clips (full_binary_pursuit.py:245-252, np.cov with
// ddof=0 over a sample of clip rows) and each template's likelihood variance
// sum_c t_c^T cov_c t_c (neuron_separability.py:126-135).  Both are small
// float64 reductions; the kernels keep one fixed summation order per output so
// repeated calls agree bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

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
// exactly symmetric.
__global__ __launch_bounds__(kMaxCovSamples) void k_cov_colsum(const float *__restrict__ clips, uint64_t width,
                                                               const uint64_t *__restrict__ rows, uint64_t n_rows,
                                                               uint32_t N, uint64_t slice, double *__restrict__ part)
{
    const uint32_t c = blockIdx.x, r = blockIdx.y, R = gridDim.y, j = threadIdx.x;
    if (j >= N) return;
    const uint64_t *rc = rows ? rows + (uint64_t)c * n_rows : nullptr;
    const uint64_t b = (uint64_t)r * slice, e = b + slice < n_rows ? b + slice : n_rows;
    const float *col = clips + (uint64_t)c * N + j;
    double s = 0.0;
    for (uint64_t k = b; k < e; k++) s += (double)col[(rc ? rc[k] : k) * width];
    part[((uint64_t)c * R + r) * N + j] = s;
}

__global__ __launch_bounds__(kCovThreads) void k_cov_products(const float *__restrict__ clips, uint64_t width,
                                                              const uint64_t *__restrict__ rows, uint64_t n_rows,
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
            xs[rr][j] = (double)clips[row * width + col0 + j] - mean[j];
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

}  // namespace

uint32_t channel_covariance_slices(uint64_t n_rows, uint64_t *slice)
{
    // slices of at least 256 rows, at most 64 of them
    uint64_t len = (n_rows + 63) / 64;
    if (len < 256) len = 256;
    *slice = len;
    return (uint32_t)((n_rows + len - 1) / len);
}

hipError_t launch_channel_covariance(const float *clips, uint64_t width, const uint64_t *rows, uint64_t n_rows,
                                     uint32_t C, uint32_t N, double *colpart, double *part, double *cov,
                                     hipStream_t st)
{
    if (C == 0 || N == 0 || n_rows == 0) return hipSuccess;
    if (N > kMaxCovSamples) return hipErrorInvalidValue;
    uint64_t slice = 0;
    const uint32_t R = channel_covariance_slices(n_rows, &slice);
    const uint32_t chunks = (N * N + kCovChunk - 1) / kCovChunk;
    hipLaunchKernelGGL(k_cov_colsum, dim3(C, R), dim3(kMaxCovSamples), 0, st, clips, width, rows, n_rows, N, slice,
                       colpart);
    hipLaunchKernelGGL(k_cov_products, dim3(C, chunks, R), dim3(kCovThreads), 0, st, clips, width, rows, n_rows, N,
                       slice, (const double *)colpart, part);
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
