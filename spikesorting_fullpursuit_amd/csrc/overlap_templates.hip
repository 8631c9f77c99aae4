// Residuals of "template = sum of two shifted templates" for the overlap-
// template search before binary pursuit (find_overlap_templates,
// c_cython/sort_cython.pyx:333-456, called at full_binary_pursuit.py:257-262).
//
// For a triple (test, n1, n2) and shifts s1, s2 in [-n_pre, n_post], the
// reference computes np.sum((test - (shift(t1, s1) + shift(t2, s2)))**2) in
// float64 with shift(t, s)[c][j] = t[c][j - s] where 0 <= j - s < N, else 0
// (sort_cython.pyx:340-350, :430-438), and keeps the first smallest residual
// in (s1, s2) loop order with strict < (:440-446).  numpy sums a 1-D float64
// array pairwise (blocks of at most 8192 elements added in order, each block
// split in halves rounded down to a multiple of 8 down to leaves of at most
// 128 elements, a leaf summed with eight interleaved accumulators); the host
// passes that summation tree as a program so every residual is bit-exact.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fbp_internal.h"

namespace fbp {
namespace {

constexpr uint32_t kPwAdd = 0xffffffffu;   // program op: pop two, push their sum
constexpr int kPwDepth = 16;               // stack depth the host program may use

// One block per triple (test, n1, n2): its three templates in LDS, one thread
// per (s1, s2) at a time, a block reduction to (smallest residual, first
// combination).
__global__ __launch_bounds__(256) void k_overlap_template_residuals(
    const float *__restrict__ Tm, uint32_t C, uint32_t N, int32_t n_pre, uint32_t S,
    const uint32_t *__restrict__ triples, uint64_t n_triples, const uint32_t *__restrict__ prog, uint32_t n_prog,
    double *__restrict__ best_res, uint32_t *__restrict__ best_combo)
{
    extern __shared__ float tl[];   // [3][C*N]: test, t1, t2
    __shared__ double r_v[4];
    __shared__ uint32_t r_c[4];
    const uint32_t CN = C * N;
    for (uint64_t k = blockIdx.x; k < n_triples; k += gridDim.x) {
        __syncthreads();
        for (uint32_t q = 0; q < 3; q++) {
            const float *src = Tm + (uint64_t)triples[3 * k + q] * CN;
            for (uint32_t e = threadIdx.x; e < CN; e += blockDim.x) tl[q * CN + e] = src[e];
        }
        __syncthreads();
        const float *tt = tl, *t1 = tl + CN, *t2 = tl + 2 * CN;
        double bv = __builtin_inf();
        uint32_t bc = 0xffffffffu;
        for (uint32_t combo = threadIdx.x; combo < S * S; combo += blockDim.x) {
            const int32_t s1 = (int32_t)(combo / S) - n_pre, s2 = (int32_t)(combo % S) - n_pre;
            // squared difference at flat index e = c*N + j
            auto val = [&](uint32_t c, uint32_t j) {
                const int32_t j1 = (int32_t)j - s1, j2 = (int32_t)j - s2;
                const double a = (j1 >= 0 && j1 < (int32_t)N) ? (double)t1[c * N + j1] : 0.0;
                const double b = (j2 >= 0 && j2 < (int32_t)N) ? (double)t2[c * N + j2] : 0.0;
                const double d = (double)tt[c * N + j] - (a + b);
                return d * d;
            };
            double stk[kPwDepth];
            int sp = 0;
            for (uint32_t p = 0; p < n_prog; p++) {
                const uint32_t op = prog[p];
                if (op == kPwAdd) {
                    stk[sp - 2] = stk[sp - 2] + stk[sp - 1];
                    sp--;
                    continue;
                }
                const uint32_t start = op >> 8, len = op & 0xffu;   // a leaf of at most 128 elements
                uint32_t c = start / N, j = start - c * N;
                auto next = [&]() {
                    const double v = val(c, j);
                    if (++j == N) { j = 0; c++; }
                    return v;
                };
                double res;
                if (len < 8) {
                    res = 0.0;
                    for (uint32_t i = 0; i < len; i++) res += next();
                } else {
                    double r[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) r[q] = next();
                    uint32_t i = 8;
                    for (; i < len - len % 8; i += 8)
#pragma unroll
                        for (int q = 0; q < 8; q++) r[q] += next();
                    res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
                    for (; i < len; i++) res += next();
                }
                stk[sp++] = res;
            }
            const double v = stk[0];
            if (v < bv) { bv = v; bc = combo; }   // combinations ascend per thread: first smallest kept
        }
        // block: smallest residual, then the first combination holding it
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const uint32_t oc = __shfl_xor(bc, off);
            if (ov < bv || (ov == bv && oc < bc)) { bv = ov; bc = oc; }
        }
        if ((threadIdx.x & 63) == 0) { r_v[threadIdx.x >> 6] = bv; r_c[threadIdx.x >> 6] = bc; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint32_t w = 1; w < blockDim.x / 64; w++)
                if (r_v[w] < bv || (r_v[w] == bv && r_c[w] < bc)) { bv = r_v[w]; bc = r_c[w]; }
            best_res[k] = bv;
            best_combo[k] = bc;
        }
    }
}

}  // namespace

hipError_t launch_overlap_template_residuals(const float *Tm, uint32_t C, uint32_t N, int32_t n_pre, uint32_t S,
                                             const uint32_t *triples, uint64_t n_triples, const uint32_t *prog,
                                             uint32_t n_prog, double *best_res, uint32_t *best_combo,
                                             hipStream_t st)
{
    if (n_triples == 0) return hipSuccess;
    const size_t lds = (size_t)3 * C * N * sizeof(float);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_overlap_template_residuals,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const uint64_t grid = n_triples < 65536 ? n_triples : 65536;
    hipLaunchKernelGGL(k_overlap_template_residuals, dim3((uint32_t)grid), dim3(256), lds, st, Tm, C, N, n_pre, S,
                       triples, n_triples, prog, n_prog, best_res, best_combo);
    return hipGetLastError();
}

}  // namespace fbp
