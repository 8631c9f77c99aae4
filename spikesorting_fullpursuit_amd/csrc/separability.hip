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
// (rows[c * n_rows + r], or r itself when rows is null), x_rj = clips[row][c*N + j] as
// float64, m_j the float64 mean of column j.  One block per (channel, chunk of
// N*N entries); entries with i > j are written from their mirror (i <= j), so
// every matrix is exactly symmetric.  Sums run over rows in ascending order.
__global__ __launch_bounds__(kCovThreads) void k_channel_cov(const float *__restrict__ clips, uint64_t width,
                                                             const uint64_t *__restrict__ rows, uint64_t n_rows,
                                                             uint32_t N, double *__restrict__ cov)
{
    __shared__ double mean[kMaxCovSamples];
    __shared__ double xs[kCovRows][kMaxCovSamples];
    const uint32_t c = blockIdx.x;
    const uint64_t col0 = (uint64_t)c * N;
    const uint64_t *rc = rows ? rows + (uint64_t)c * n_rows : nullptr;
    auto row_of = [&](uint64_t r) { return rc ? rc[r] : r; };
    for (uint32_t j = threadIdx.x; j < N; j += blockDim.x) {
        double s = 0.0;
        for (uint64_t r = 0; r < n_rows; r++) s += (double)clips[row_of(r) * width + col0 + j];
        mean[j] = s / (double)n_rows;
    }
    const uint32_t p0 = blockIdx.y * kCovChunk + threadIdx.x;
    double acc[kCovPairs];
#pragma unroll
    for (uint32_t q = 0; q < kCovPairs; q++) acc[q] = 0.0;
    for (uint64_t r0 = 0; r0 < n_rows; r0 += kCovRows) {
        const uint32_t nr = n_rows - r0 < kCovRows ? (uint32_t)(n_rows - r0) : kCovRows;
        __syncthreads();   // means ready; the previous tile's reads are done
        for (uint32_t e = threadIdx.x; e < nr * N; e += blockDim.x) {
            const uint32_t rr = e / N, j = e - rr * N;
            xs[rr][j] = (double)clips[row_of(r0 + rr) * width + col0 + j] - mean[j];
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
    double *dst = cov + (uint64_t)c * N * N;
#pragma unroll
    for (uint32_t q = 0; q < kCovPairs; q++) {
        const uint32_t p = p0 + q * kCovThreads;
        if (p >= N * N) break;
        const uint32_t i = p / N, j = p - i * N;
        if (i > j) continue;
        const double v = acc[q] / (double)n_rows;
        dst[i * N + j] = v;
        dst[j * N + i] = v;
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

hipError_t launch_channel_covariance(const float *clips, uint64_t width, const uint64_t *rows, uint64_t n_rows,
                                     uint32_t C, uint32_t N, double *cov, hipStream_t st)
{
    if (C == 0 || N == 0 || n_rows == 0) return hipSuccess;
    if (N > kMaxCovSamples) return hipErrorInvalidValue;
    const uint32_t chunks = (N * N + kCovChunk - 1) / kCovChunk;
    hipLaunchKernelGGL(k_channel_cov, dim3(C, chunks), dim3(kCovThreads), 0, st, clips, width, rows, n_rows, N, cov);
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
