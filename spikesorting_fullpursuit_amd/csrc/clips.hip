// Clip gathers around the binary pursuit call: get_clips on the input voltage
// (parallel/segment_parallel.py:440-509, called at full_binary_pursuit.py:357-359),
// the per-unit mean clip the caller reduces them to
// (full_binary_pursuit.py:366-368), and the squared distance of every clip to
// its unit's template that the duplicate removal after binary pursuit
// compares (analyze_spike_timing.py:129-130).  All are data movement
// (HBM-bound); the reductions keep numpy's float32 order so they are
// bit-exact.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fbp_internal.h"

namespace fbp {
namespace {

// clips[k][j * W + x] = V[chan[j]][idx[k] + w0 + x] for x < W = w1 - w0.  One
// block per event, one wave per channel at a time: lane x copies sample x
// (and x + 64, ...), so each channel's W samples are read and written
// contiguously.
__global__ __launch_bounds__(256) void k_get_clips(const float *__restrict__ V, uint64_t row_stride,
                                                   const int64_t *__restrict__ idx, uint64_t n_events,
                                                   const uint32_t *__restrict__ chan, uint32_t n_sel, int32_t w0,
                                                   uint32_t W, float *__restrict__ clips)
{
    const uint32_t width = n_sel * W;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (uint64_t k = blockIdx.x; k < n_events; k += gridDim.x) {
        const int64_t start = idx[k] + w0;   // >= 0 and start + W <= n_samples (checked on the host)
        float *dst = clips + k * width;
        for (uint32_t j = wave; j < n_sel; j += nwave) {
            const float *src = V + (uint64_t)chan[j] * row_stride + (uint64_t)start;
            for (uint32_t x = lane; x < W; x += 64) dst[j * W + x] = src[x];
        }
    }
}

// means[u][x] = (((c[r0][x] + c[r1][x]) + c[r2][x]) + ...) / n_u in float32,
// rows r0, r1, ... of unit u in input order (rows[off[u] .. off[u + 1])): the
// order numpy's axis-0 float32 mean adds them.  One thread per (unit,
// column); the row loads of eight steps are issued before their adds.
template <class F>
__global__ __launch_bounds__(256) void k_unit_means(const F *__restrict__ clips, uint32_t width,
                                                    const uint32_t *__restrict__ rows,
                                                    const uint64_t *__restrict__ off, uint32_t n_units,
                                                    F *__restrict__ means)
{
    const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (g >= (uint64_t)n_units * width) return;
    const uint32_t u = (uint32_t)(g / width), x = (uint32_t)(g - (uint64_t)u * width);
    const uint64_t b = off[u], e = off[u + 1];
    if (b == e) {
        means[g] = (F)__builtin_nan("");   // numpy: mean of no rows
        return;
    }
    F acc = clips[(uint64_t)rows[b] * width + x];
    uint64_t r = b + 1;
    for (; r + 8 <= e; r += 8) {
        F v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = clips[(uint64_t)rows[r + q] * width + x];
#pragma unroll
        for (int q = 0; q < 8; q++) acc = acc + v[q];
    }
    for (; r < e; r++) acc = acc + clips[(uint64_t)rows[r] * width + x];
    means[g] = acc / (F)(e - b);
}

// sse[k] = np.sum((clips[k] - templates[tidx[k]]) ** 2) in float32 with
// numpy's pairwise summation order, given as a program (leaf op start << 8 |
// length, 0xffffffff adds the top two stack entries).  One thread per clip.
__global__ __launch_bounds__(256) void k_clip_sse(const float *__restrict__ clips, uint64_t n, uint32_t width,
                                                  const uint32_t *__restrict__ tidx,
                                                  const float *__restrict__ templates,
                                                  const uint32_t *__restrict__ prog, uint32_t n_prog,
                                                  float *__restrict__ sse)
{
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float *c = clips + k * width, *t = templates + (uint64_t)tidx[k] * width;
    auto val = [&](uint32_t e) {
        const float d = c[e] - t[e];
        return d * d;
    };
    float stk[16];
    int sp = 0;
    for (uint32_t p = 0; p < n_prog; p++) {
        const uint32_t op = prog[p];
        if (op == 0xffffffffu) {
            stk[sp - 2] = stk[sp - 2] + stk[sp - 1];
            sp--;
            continue;
        }
        const uint32_t start = op >> 8, len = op & 0xffu;
        float res;
        if (len < 8) {
            res = 0.0f;
            for (uint32_t i = 0; i < len; i++) res += val(start + i);
        } else {
            float r[8];
#pragma unroll
            for (int q = 0; q < 8; q++) r[q] = val(start + q);
            uint32_t i = 8;
            for (; i < len - len % 8; i += 8)
#pragma unroll
                for (int q = 0; q < 8; q++) r[q] += val(start + i + q);
            res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
            for (; i < len; i++) res += val(start + i);
        }
        stk[sp++] = res;
    }
    sse[k] = stk[0];
}

}  // namespace

hipError_t launch_clip_sse(const float *clips, uint64_t n, uint32_t width, const uint32_t *tidx,
                           const float *templates, const uint32_t *prog, uint32_t n_prog, float *sse, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clip_sse, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, clips, n, width, tidx,
                       templates, prog, n_prog, sse);
    return hipGetLastError();
}

hipError_t launch_get_clips(const float *V, uint64_t row_stride, const int64_t *idx, uint64_t n_events,
                            const uint32_t *chan, uint32_t n_sel, int32_t w0, uint32_t W, float *clips,
                            hipStream_t st)
{
    if (n_events == 0) return hipSuccess;
    const uint64_t grid = n_events < 65536 ? n_events : 65536;
    hipLaunchKernelGGL(k_get_clips, dim3((uint32_t)grid), dim3(256), 0, st, V, row_stride, idx, n_events, chan,
                       n_sel, w0, W, clips);
    return hipGetLastError();
}

template <class F>
hipError_t launch_unit_means(const F *clips, uint32_t width, const uint32_t *rows, const uint64_t *off,
                             uint32_t n_units, F *means, hipStream_t st)
{
    const uint64_t total = (uint64_t)n_units * width;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_unit_means<F>, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, clips, width, rows,
                       off, n_units, means);
    return hipGetLastError();
}

template hipError_t launch_unit_means<float>(const float *, uint32_t, const uint32_t *, const uint64_t *, uint32_t,
                                             float *, hipStream_t);
template hipError_t launch_unit_means<double>(const double *, uint32_t, const uint32_t *, const uint64_t *,
                                              uint32_t, double *, hipStream_t);

}  // namespace fbp
