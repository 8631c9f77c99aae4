// kernels.hip -- gfx950 kernels of the Full Binary Pursuit engine.
//
// Each kernel names the reference kernel of
// spikesorting_fullpursuit/kernels/binary_pursuit.cl ("bp.cl") whose results
// it reproduces.  Where the reference races, the kernels implement the outcome
// of one lock-stepped wavefront, applied to the whole chunk:
//   * flags written while deciding adds happen before any pruning read, and
//     pruning reads neighbour flags as they stood after the adds (bp.cl:1009);
//   * check_overlap_reassignments reads the id-2 flag as parsed (bp.cl:852);
//   * spikes of a pass are recorded in ascending window order (bp.cl:1047);
//   * every sample's template subtractions run in ascending (onset, label,
//     record position) order, none lost (bp.cl:35-43, :89-97).
// All float arithmetic that decides an event is IEEE float32 with explicit
// fused multiply-adds in the reference's order; the file is compiled with
// -ffp-contract=off so nothing else is fused.
#include "fbp_internal.h"

namespace fbp {

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_of() { return threadIdx.x & (kWave - 1); }

// ---------------------------------------------------------------------------
// Fills and layout helpers
// ---------------------------------------------------------------------------
__global__ void k_fill_u8(uint8_t *p, uint64_t n, uint8_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_reset_counters(Counters *c)
{
    if (threadIdx.x == 0) *c = Counters{};
}

__global__ void k_iota_u32(uint32_t *p, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

__global__ void k_copy_rows(float *dst, const float *src, uint32_t rows, uint64_t cols, uint64_t stride)
{
    const uint64_t total = (uint64_t)rows * cols;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = e / cols, c = e - r * cols;
        dst[e] = src[r * stride + c];
    }
}

// TmT[(c*N + j)*Tp + t] = Tm[t][c][j]; zero for t >= T (padding).
__global__ void k_transpose_templates(Dims d, const float *__restrict__ Tm, const float *__restrict__ bias_in,
                                      float *__restrict__ TmT, float *__restrict__ bias_out)
{
    const uint64_t total = (uint64_t)d.C * d.N * d.Tp;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t t = (uint32_t)(e % d.Tp);
        const uint64_t cj = e / d.Tp;
        const uint32_t c = (uint32_t)(cj / d.N), j = (uint32_t)(cj % d.N);
        TmT[e] = t < d.T ? Tm[((uint64_t)t * d.C + c) * d.N + j] : 0.0f;
        if (cj == 0) bias_out[t] = t < d.T ? bias_in[t] : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Correlation, float32 (bp.cl:216-300, compute_full_likelihood).
// One 64-lane block per listed window; lane = position in the window.  The
// window's voltage span [wN, wN + 64 + N - 1) is staged in LDS channel block by
// channel block; template values are wave-uniform (scalar loads from the
// transposed [C][N][Tp] copy) and feed CORR_TB accumulators per lane, so one
// LDS read serves CORR_TB fused multiply-adds.  Every accumulator follows the
// reference chain exactly: bias, then c = 0..C-1, j = 0..N-1.
// ---------------------------------------------------------------------------
constexpr int kCorrTB = 8;
constexpr int kCorrLdsFloats = 5120;   // 20 KB per block

__global__ __launch_bounds__(64) void k_corr_fp32(Dims d, const float *__restrict__ V,
                                                  const float *__restrict__ TmT,
                                                  const float *__restrict__ bias,
                                                  const uint32_t *__restrict__ win, float *__restrict__ L,
                                                  uint32_t cb_max, const uint64_t *__restrict__ wdirty)
{
    __shared__ float lds[kCorrLdsFloats];
    const int lane = threadIdx.x;
    const uint32_t w = win[blockIdx.x];
    // no voltage under the window's positions changed since they were
    // computed: L holds these values already
    if (!d.dirty_off && wdirty[w] == 0) return;
    const uint64_t start = (uint64_t)w * d.N;
    const uint64_t stop = min(start + d.N, (uint64_t)d.K);
    const uint32_t span = kWave + d.N - 1;
    const bool staged_once = cb_max >= d.C;
    for (uint64_t base = start; base < stop; base += kWave) {
        const uint64_t i = base + lane;
        const bool active = i < stop;
        const bool valid = active && (i + d.N < d.K);   // bp.cl:231-234
        bool have = false;
        for (uint32_t tb = 0; tb < d.Tp; tb += kCorrTB) {
            float acc[kCorrTB];
#pragma unroll
            for (int r = 0; r < kCorrTB; r++) acc[r] = bias[tb + r];
            for (uint32_t c0 = 0; c0 < d.C; c0 += cb_max) {
                const uint32_t cb = min(cb_max, d.C - c0);
                if (!(staged_once && have)) {
                    __syncthreads();
                    for (uint32_t c = 0; c < cb; c++) {
                        const float *vr = V + (uint64_t)(c0 + c) * d.K;
                        for (uint32_t x = lane; x < span; x += kWave) {
                            const uint64_t g = base + x;
                            lds[c * span + x] = g < d.K ? vr[g] : 0.0f;
                        }
                    }
                    __syncthreads();
                    have = true;
                }
                for (uint32_t c = 0; c < cb; c++) {
                    const float *vrow = lds + c * span + lane;
                    const float *trow = TmT + (uint64_t)(c0 + c) * d.N * d.Tp + tb;
                    for (uint32_t j = 0; j < d.N; j++) {
                        const float v = vrow[j];
                        const float *tj = trow + (uint64_t)j * d.Tp;
#pragma unroll
                        for (int r = 0; r < kCorrTB; r++) acc[r] = __builtin_fmaf(tj[r], v, acc[r]);
                    }
                }
            }
            if (active) {
#pragma unroll
                for (int r = 0; r < kCorrTB; r++)
                    if (tb + r < d.T) L[(uint64_t)(tb + r) * d.K + i] = valid ? acc[r] : 0.0f;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Correlation screen on the matrix cores (bf16 or fp16 inputs, float32
// accumulation) followed by an exact float32 re-score of every position whose
// likelihood could be positive.
//
// The likelihood decides events only through its sign (compared with 0 by the
// argmax, the check flags and the overlap skip rules) and, where it is
// positive, through its exact float32 value (bp.cl:402-417, :572, :733-775).
// The screen computes Lhat = bias + sum over (c, j) of q(T[t,c,j]) q(V[c,i+j])
// with a bound e >= |Lhat - L32| (L32 = the reference's float32 chain):
//   e = k1*|t|*|v_i| + k2*(|t| + |v_i|) + k3 + k4*|bias|
// from the input rounding (unit roundoff u of the format, absolute floor delta
// for flushed subnormals) and the float32 accumulation of both sums (gamma_n
// with a doubled unit roundoff).  Entries with Lhat > -e (or a non-finite
// Lhat) are queued and re-scored with the float32 chain; the others are stored
// as Lhat <= -e <= 0.  Every stored value then takes every branch of the
// reference exactly as the float32 value would.
//
// Matrix layout: rows = positions of the window (16 per tile), columns =
// templates (16 per tile), k = (channel, template sample padded to 16).  The
// A fragment (Hankel of the voltage) is read from four copies of the window's
// converted voltage, shifted by 0..3 samples, so every lane's four samples are
// one aligned 8-byte LDS read.  B fragments (templates) are packed once per
// template set in the exact lane order of v_mfma_f32_16x16x16{bf16,f16}.
// ---------------------------------------------------------------------------
typedef short short4v __attribute__((ext_vector_type(4)));
typedef _Float16 half4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even (inputs are finite)
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ uint16_t to_f16(float x)
{
    const _Float16 h = (_Float16)x;
    return __builtin_bit_cast(uint16_t, h);
}

// Element traits of the three screens: bf16 and fp16 use
// v_mfma_f32_16x16x32{bf16,f16} (k = 32, 8 two-byte elements per lane: four
// aligned 4-byte LDS words, read in pairs, from two shifted voltage copies),
// fp8 e4m3 uses v_mfma_f32_16x16x32_fp8_fp8 (k = 32, 8 one-byte elements per
// lane: two aligned 4-byte words from four copies).  Lane l holds row (A) or
// column (B) l % 16 at k = 8 * (l / 16) + 0..7.
template <int PREC>
struct NarrowTraits {
    static constexpr uint32_t kBytes = PREC == 3 ? 1 : 2;   // bytes per element
    static constexpr uint32_t kK = 32;                        // k per MFMA
    static constexpr uint32_t kPer = 8;                       // elements per lane fragment
    static constexpr uint32_t kWords = kPer * kBytes / 8;     // 8-byte words per fragment
    static constexpr uint32_t kCopies = 4 / kBytes;           // shifted voltage copies in LDS
};

template <int PREC>
struct Frag {
    uint64_t w[NarrowTraits<PREC>::kWords];
};

typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8v __attribute__((ext_vector_type(8)));

// Round a float32 to the screen's element format (round to nearest even).
// fp8 returns the e4m3 byte in the low 8 bits.
template <int PREC>
__device__ __forceinline__ uint32_t to_narrow(float x)
{
    if constexpr (PREC == 1) return to_bf16(x);
    else if constexpr (PREC == 2) return to_f16(x);
    else return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false) & 0xffu;
}

template <int PREC>
__device__ __forceinline__ float4v mfma_narrow(const Frag<PREC> &a, const Frag<PREC> &b, float4v c)
{
    if constexpr (PREC == 1) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8v, a), __builtin_bit_cast(bf16x8v, b),
                                                        c, 0, 0, 0);
    } else if constexpr (PREC == 2) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(halfx8v, a), __builtin_bit_cast(halfx8v, b),
                                                       c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8((long)a.w[0], (long)b.w[0], c, 0, 0, 0);
    }
}

// v_mfma_f32_16x16x16bf16_1k: lane l holds row (A) or column (B) l % 16 at
// k = 4 * (l / 16) + 0..3 (the split-bf16 Q table of the overlap screen).
__device__ __forceinline__ float4v mfma_bf16_k16(uint64_t a, uint64_t b, float4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(short4v, a), __builtin_bit_cast(short4v, b),
                                                    c, 0, 0, 0);
}

// Largest |value| the fp8 screen feeds to the matrix cores unscaled; a window
// holding a larger voltage sample has all its entries re-scored.
constexpr float kFp8Limit = 440.0f;

// TmB[((c*KS + ks)*CT + ct)*64 + lane] = the fragment (kWords 8-byte words) of
// template t = ct*16 + lane%16 at samples j = ks*K + per*(lane/16) + 0..per-1
// (0 when padded), scaled by tscale[t] (a power of two; 1 except for fp8).
template <int PREC>
__global__ void k_pack_templates(Dims d, const float *__restrict__ Tm, const float *__restrict__ tscale,
                                 const uint32_t *__restrict__ tord, Frag<PREC> *__restrict__ TmB, uint32_t KS,
                                 uint32_t CT)
{
    using Tr = NarrowTraits<PREC>;
    const uint64_t total = (uint64_t)d.C * KS * CT * 64;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t lane = (uint32_t)(e % 64);
        const uint64_t r = e / 64;
        const uint32_t ct = (uint32_t)(r % CT);
        const uint32_t ks = (uint32_t)((r / CT) % KS);
        const uint32_t c = (uint32_t)(r / ((uint64_t)CT * KS));
        const uint32_t ts = ct * 16 + (lane & 15);
        const uint32_t t = ts < d.T ? tord[ts] : d.T;   // column ts holds template tord[ts]
        const float sc = t < d.T ? tscale[t] : 1.0f;
        Frag<PREC> v;
        for (uint32_t w = 0; w < Tr::kWords; w++) v.w[w] = 0;
        for (uint32_t q = 0; q < Tr::kPer; q++) {
            const uint32_t j = ks * Tr::kK + Tr::kPer * (lane >> 4) + q;
            const float x = (t < d.T && j < d.N) ? Tm[((uint64_t)t * d.C + c) * d.N + j] * sc : 0.0f;
            const uint32_t bit = 8 * Tr::kBytes * q;
            v.w[bit / 64] |= (uint64_t)to_narrow<PREC>(x) << (bit % 64);
        }
        TmB[e] = v;
    }
}

// Template norms (upper bounds), used by both screens, and the fp8 scales:
// the power of two that brings the template's largest |value| to (112, 224].
__global__ void k_template_norms(Dims d, const float *__restrict__ Tm, float *__restrict__ tnorm,
                                 float *__restrict__ tscale, uint32_t fp8)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.T) return;
    double acc = 0.0;
    float mx = 0.0f;
    for (uint64_t k = 0; k < (uint64_t)d.C * d.N; k++) {
        const float xf = Tm[(uint64_t)t * d.C * d.N + k];
        const double x = xf;
        acc = __builtin_fma(x, x, acc);
        mx = fmaxf(mx, fabsf(xf));
    }
    const double nrm = sqrt(acc) * (1.0 + 1e-12);
    float up = (float)nrm;
    if ((double)up < nrm) up = __uint_as_float(__float_as_uint(up) + 1u);
    tnorm[t] = up;
    float sc = 1.0f;
    if (fp8 && mx > 0.0f) {
        int ex;
        frexpf(224.0f / mx, &ex);   // 224/mx = m * 2^ex with m in [0.5, 1)
        sc = ldexpf(1.0f, ex - 1);  // largest power of two <= 224/mx
    }
    tscale[t] = sc;
}

struct ScreenCoef {
    float k1, k2, k3, k4;
};

__device__ __forceinline__ float exact_likelihood(const Dims &d, const float *__restrict__ V,
                                                  const float *__restrict__ Tm, const float *__restrict__ bias,
                                                  uint32_t c_lo, uint32_t c_hi, uint32_t t, uint64_t i)
{
    // channels outside [c_lo, c_hi) hold a zero template and add exactly 0
    float acc = bias[t];
    for (uint32_t c = c_lo; c < c_hi; c++) {
        const float *tc = Tm + ((uint64_t)t * d.C + c) * d.N;
        const float *vc = V + (uint64_t)c * d.K + i;
        for (uint32_t j = 0; j < d.N; j++) acc = __builtin_fmaf(tc[j], vc[j], acc);
    }
    return acc;
}

constexpr int kMfmaMaxRT = 6;   // row tiles per window: N <= 96

// Register budget: four waves per SIMD for up to 64 positions and 32 template
// columns (measured faster than three, despite a few spilled registers), fewer
// for the larger tiles.
template <int PREC, int RT, int CTB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PREC != 3 && RT <= 4 && CTB <= 2 ? 4 : (RT * CTB > 16 ? 1 : 2)))) void k_corr_mfma(Dims d, const float *__restrict__ V, const float *__restrict__ Tm,
                                                  const Frag<PREC> *__restrict__ TmB, const float *__restrict__ bias,
                                                  const float *__restrict__ tnorm, const float *__restrict__ tscale,
                                                  const uint32_t *__restrict__ win, float *__restrict__ L,
                                                  Counters *ctr, uint64_t *__restrict__ rq, uint32_t rq_cap,
                                                  uint2 *__restrict__ rdesc, uint32_t cb_max, ScreenCoef co,
                                                  uint32_t no_rescore, const uint32_t *__restrict__ tord,
                                                  const uint32_t *__restrict__ tile_lo,
                                                  const uint32_t *__restrict__ tile_hi,
                                                  const uint32_t *__restrict__ sup_lo,
                                                  const uint32_t *__restrict__ sup_hi, uint32_t *rq_shard,
                                                  const uint64_t *__restrict__ wdirty)
{
    using Tr = NarrowTraits<PREC>;
    constexpr uint32_t KS = (RT * 16 + Tr::kK - 1) / Tr::kK;   // k-steps per channel (N padded to RT*16)
    constexpr uint32_t LA = RT * 16 + KS * Tr::kK;            // elements staged per channel and copy
    extern __shared__ uint64_t vb_raw[];
    uint8_t *vb = (uint8_t *)vb_raw;            // [copies][cb][LA] narrow voltage, shifted by 0..copies-1
    __shared__ double s2[kMfmaMaxRT * 16 * 2 + 16];
    __shared__ float vnb[kMfmaMaxRT * 16];
    __shared__ int over;
    const int lane = threadIdx.x;
    const uint32_t CT = (d.T + 15) / 16;
    // one block per (listed window, chunk of CTB template tiles)
    const uint32_t nchunk = (CT + CTB - 1) / CTB;
    const uint32_t wi = blockIdx.x / nchunk, chunk = blockIdx.x - wi * nchunk;
    const uint32_t w = win[wi];
    const uint64_t base = (uint64_t)w * d.N;
    const uint64_t wend = min(base + d.N, (uint64_t)d.K);
    const uint32_t np = (uint32_t)(wend - base);   // positions of this window
    // channels on which the chunk's template tiles are nonzero (templates
    // are ordered by channel support, tiles of 16); elsewhere every B
    // fragment is zero and an MFMA would add exact zeros
    const uint32_t ct0 = chunk * CTB;
    const uint32_t nct = min((uint32_t)CTB, CT - ct0);
    uint32_t tlo[CTB], thi[CTB], clo = d.C, chi = 0;
#pragma unroll
    for (int ct = 0; ct < CTB; ct++) {
        tlo[ct] = ct < (int)nct ? tile_lo[ct0 + ct] : 0u;
        thi[ct] = ct < (int)nct ? tile_hi[ct0 + ct] : 0u;
        if (tlo[ct] < thi[ct]) {
            clo = min(clo, tlo[ct]);
            chi = max(chi, thi[ct]);
        }
    }
    if (chi < clo) chi = clo;
    // no voltage on the chunk's channels changed under the window's positions
    // since they were computed: L holds these values already
    if (!d.dirty_off) {
        const uint64_t cm = channel_mask(clo, chi, d.dirty_g);
        if (cm && !(wdirty[w] & cm)) {
            if (lane == 0) rdesc[(uint64_t)wi * nchunk + chunk] = make_uint2(0u, 0u);
            return;
        }
    }

    // Per-position window norm bound |v_i| over the chunk's channels [clo, chi)
    // (the only ones whose samples enter its products), float64 sums of the
    // float32 samples: s2[x] = sum over those channels of v[c][x]^2,
    // accumulated while staging when they are staged at once, else in a pass
    // of its own.
    const bool staged_once = cb_max >= chi - clo;
    const uint32_t nsq = RT * 16 + d.N;
    if (lane == 0) over = 0;
    if (!staged_once || clo == chi) {
        for (uint32_t x = lane; x < nsq; x += 64) {
            double a = 0.0;
            const uint64_t g = base + x;
            if (g < d.K)
                for (uint32_t c = clo; c < chi; c++) {
                    const double v = V[(uint64_t)c * d.K + g];
                    a = __builtin_fma(v, v, a);
                }
            s2[x] = a;
        }
    }
    auto window_norms = [&]() {
        for (uint32_t p = lane; p < RT * 16; p += 64) {
            double a = 0.0;
            for (uint32_t x = 0; x < d.N; x++) a += s2[p + x];
            const double nv = sqrt(a) * (1.0 + 1e-9);
            float up = (float)nv;
            if ((double)up < nv) up = __uint_as_float(__float_as_uint(up) + 1u);
            vnb[p] = up;
        }
    };
    if (!staged_once || clo == chi) {
        __syncthreads();
        window_norms();
    }
    // Stage channels [c0, c0 + cb): lane owns samples g = lane + 64*slot, and
    // the loads of up to 16 channels of a sample are issued together.  Copy r
    // holds sample base + y + r at y.
    // Samples at or beyond np + N - 1 meet only zero template weights or
    // discarded positions: they are staged as 0 without a load.  Two 64-sample
    // slots are loaded together, up to 16 channels at a time.
    const uint32_t gmax = np + d.N - 1;
    auto stage = [&](uint32_t c0, uint32_t cb) {
        constexpr uint32_t LS = LA + Tr::kCopies;
        auto put = [&](uint32_t g, uint32_t c, float x, double &a) {
            if (PREC == 3 && fabsf(x) > kFp8Limit) over = 1;
            a = __builtin_fma((double)x, (double)x, a);
            const uint32_t qv = to_narrow<PREC>(x);
#pragma unroll
            for (int r = 0; r < (int)Tr::kCopies; r++) {
                const int y = (int)g - r;
                if (y < 0 || y >= (int)LA) continue;
                uint8_t *dst = vb + (((uint64_t)r * cb + c) * LA + y) * Tr::kBytes;
                if constexpr (Tr::kBytes == 2) *(uint16_t *)dst = (uint16_t)qv;
                else *dst = (uint8_t)qv;
            }
        };
        for (uint32_t g0 = lane; g0 < LS; g0 += 128) {
            const uint32_t g1 = g0 + 64;
            const bool l0 = g0 < gmax && base + g0 < d.K, l1 = g1 < gmax && base + g1 < d.K;
            double a0 = 0.0, a1 = 0.0;
            for (uint32_t cc = 0; cc < cb; cc += 16) {
                float x0[16], x1[16];
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const float *vr = V + (uint64_t)(c0 + cc + q) * d.K + base;
                    x0[q] = (cc + q < cb && l0) ? vr[g0] : 0.0f;
                    x1[q] = (cc + q < cb && l1) ? vr[g1] : 0.0f;
                }
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    if (cc + q < cb) {
                        put(g0, cc + q, x0[q], a0);
                        if (g1 < LS) put(g1, cc + q, x1[q], a1);
                    }
                }
            }
            if (staged_once && g0 < nsq) s2[g0] = a0;
            if (staged_once && g1 < nsq) s2[g1] = a1;
        }
    };

    bool have = false;
    {
        float4v acc[RT][CTB];
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int ct = 0; ct < CTB; ct++) acc[rt][ct] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t c0 = clo; c0 < chi; c0 += cb_max) {
            const uint32_t cb = min(cb_max, chi - c0);
            if (!(staged_once && have)) {
                __syncthreads();
                stage(c0, cb);
                __syncthreads();
                if (staged_once) {
                    window_norms();
                    __syncthreads();
                }
                have = true;
            }
            // template fragments of channel c + 1 are fetched during channel c
            Frag<PREC> bn[KS][CTB];
            auto fetch_b = [&](uint32_t c) {
#pragma unroll
                for (uint32_t ks = 0; ks < KS; ks++)
#pragma unroll
                    for (int ct = 0; ct < CTB; ct++)
                        if (ct < (int)nct) bn[ks][ct] = TmB[(((uint64_t)(c0 + c) * KS + ks) * CT + ct0 + ct) * 64 + lane];
                        else
#pragma unroll
                            for (uint32_t w = 0; w < Tr::kWords; w++) bn[ks][ct].w[w] = 0;
            };
            const uint32_t cs = clo > c0 ? clo - c0 : 0u, ce = chi > c0 ? min(cb, chi - c0) : 0u;
            if (cs < ce) fetch_b(cs);
            for (uint32_t c = cs; c < ce; c++) {
                Frag<PREC> b[KS][CTB];
#pragma unroll
                for (uint32_t ks = 0; ks < KS; ks++)
#pragma unroll
                    for (int ct = 0; ct < CTB; ct++) b[ks][ct] = bn[ks][ct];
                if (c + 1 < ce) fetch_b(c + 1);
#pragma unroll
                for (uint32_t ks = 0; ks < KS; ks++) {
#pragma unroll
                    for (int rt = 0; rt < RT; rt++) {
                        const uint32_t o = rt * 16 + ks * Tr::kK + Tr::kPer * (lane >> 4) + (lane & 15);
                        const uint32_t r = o % Tr::kCopies;
                        const uint32_t *src =
                            (const uint32_t *)(vb + (((uint64_t)r * cb + c) * LA + (o - r)) * Tr::kBytes);
                        Frag<PREC> a;
#pragma unroll
                        for (uint32_t w = 0; w < Tr::kWords; w++)
                            a.w[w] = (uint64_t)src[2 * w] | ((uint64_t)src[2 * w + 1] << 32);
#pragma unroll
                        for (int ct = 0; ct < CTB; ct++)
                            if (c0 + c >= tlo[ct] && c0 + c < thi[ct])
                                acc[rt][ct] = mfma_narrow<PREC>(a, b[ks][ct], acc[rt][ct]);
                    }
                }
            }
        }
        const bool force = PREC == 3 && over;
        // per-template constants of this lane's columns
        float cb_bias[CTB], cb_inv[CTB], cb_tn[CTB];
        uint32_t cb_t[CTB];
        bool cb_ok[CTB];
#pragma unroll
        for (int ct = 0; ct < CTB; ct++) {
            const uint32_t ts = (ct0 + ct) * 16 + (lane & 15);
            cb_ok[ct] = ct < (int)nct && ts < d.T;
            const uint32_t t = cb_ok[ct] ? tord[ts] : 0u;   // this lane's column holds template tord[ts]
            cb_t[ct] = t;
            cb_bias[ct] = cb_ok[ct] ? bias[t] : 0.0f;
            cb_inv[ct] = cb_ok[ct] ? 1.0f / tscale[t] : 1.0f;
            cb_tn[ct] = cb_ok[ct] ? tnorm[t] : 0.0f;
        }
        // Epilogue: bias, bound, store; count the entries that need float32.
        uint32_t nflag = 0;
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int ct = 0; ct < CTB; ct++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t p = rt * 16 + 4 * (lane >> 4) + q;
                    const uint32_t t = cb_t[ct];
                    if (!cb_ok[ct] || p >= np) { acc[rt][ct][q] = 0.0f; continue; }
                    const uint64_t i = base + p;
                    if (i + d.N >= d.K) {   // bp.cl:231-234
                        L[(uint64_t)t * d.K + i] = 0.0f;
                        acc[rt][ct][q] = 0.0f;
                        continue;
                    }
                    const float lh = cb_bias[ct] + acc[rt][ct][q] * cb_inv[ct];
                    const float vn = vnb[p];
                    const float e = co.k1 * cb_tn[ct] * vn + co.k2 * (cb_tn[ct] + vn * cb_inv[ct]) + co.k3 * cb_inv[ct]
                                    + co.k4 * fabsf(cb_bias[ct]);
                    L[(uint64_t)t * d.K + i] = lh;
                    const bool fl = force || !(lh <= -e);
                    acc[rt][ct][q] = fl ? 1.0f : 0.0f;   // keep the decision for the queue pass
                }
        if (no_rescore) return;
        // Queue entries are groups: a lane's four consecutive positions of one
        // template, with a mask of those that need float32.
        auto group_mask = [&](int rt, int ct) {
            uint32_t m = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) m |= acc[rt][ct][q] != 0.0f ? 1u << q : 0u;
            return m;
        };
        uint32_t npos = 0;
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int ct = 0; ct < CTB; ct++) {
                const uint32_t m = group_mask(rt, ct);
                nflag += m ? 1u : 0u;
                npos += __popc(m);
            }
        for (int off = 32; off > 0; off >>= 1) npos += __shfl_xor(npos, off);
        // wave-wide reservation in the re-score queue
        uint32_t incl = nflag;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        const uint32_t tot = __shfl(incl, 63);
        uint32_t qbase = 0;
        if (tot) {
            if (lane == 63) {   // one counter pair per shard, each on a cache line of its own
                uint32_t *sh = rq_shard + (blockIdx.x % kQueueShards) * kShardStride;
                qbase = atomicAdd(sh, tot);
                atomicAdd(sh + 1, npos);
            }
            qbase = __shfl(qbase, 63);
        }
        // qbase is within the block's shard [shard * cap, (shard + 1) * cap)
        const uint32_t cap = rq_cap / kQueueShards, sbase = (blockIdx.x % kQueueShards) * cap;
        uint32_t pos = qbase + incl - nflag;
        // this (window, template chunk)'s slice of the queue, for the re-score kernels
        if (lane == 0) {
            const uint32_t kept = qbase >= cap ? 0u : min(tot, cap - qbase);
            rdesc[(uint64_t)wi * nchunk + chunk] = make_uint2(sbase + qbase, kept);
        }
        if (!nflag) return;
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int ct = 0; ct < CTB; ct++) {
                const uint32_t m = group_mask(rt, ct);
                if (!m) continue;
                const uint32_t t = cb_t[ct];
                const uint64_t i0 = base + rt * 16 + 4 * (lane >> 4);
                if (pos < cap) {
                    rq[sbase + pos] = ((uint64_t)m << 48) | ((uint64_t)t << 32) | i0;
                } else {   // shard full
                    for (int q = 0; q < 4; q++)
                        if (m >> q & 1)
                            L[(uint64_t)t * d.K + i0 + q] = exact_likelihood(d, V, Tm, bias, sup_lo[t], sup_hi[t], t, i0 + q);
                }
                pos++;
            }
    }
}

// acc += tp[j * ts] * vp[j] for j = 0..n-1, in order, fused; operands are
// read in batches of 8 ahead of their multiply-adds.
__device__ __forceinline__ float fma_chain(float acc, const float *tp, uint32_t ts, const float *vp, uint32_t n)
{
    uint32_t j = 0;
    for (; j + 8 <= n; j += 8) {
        float a[8], b[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            a[q] = tp[(j + q) * ts];
            b[q] = vp[j + q];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) acc = __builtin_fmaf(a[q], b[q], acc);
    }
    for (; j < n; j++) acc = __builtin_fmaf(tp[j * ts], vp[j], acc);
    return acc;
}

// Four chains of one template at consecutive positions p, p+1, p+2, p+3
// (vp = the voltage at p): each template sample is read once for the four,
// and the voltage slides through registers, one new sample per step.  Every
// chain is the reference's fused multiply-add order.  vp[0, n + 3) must be
// readable.
__device__ __forceinline__ void fma_chain4(float (&acc)[4], const float *tp, uint32_t ts, const float *vp, uint32_t n)
{
    float v0 = vp[0], v1 = vp[1], v2 = vp[2];
    uint32_t j = 0;
    for (; j + 8 <= n; j += 8) {
        float a[8], b[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            a[q] = tp[(j + q) * ts];
            b[q] = vp[j + q + 3];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            acc[0] = __builtin_fmaf(a[q], v0, acc[0]);
            acc[1] = __builtin_fmaf(a[q], v1, acc[1]);
            acc[2] = __builtin_fmaf(a[q], v2, acc[2]);
            acc[3] = __builtin_fmaf(a[q], b[q], acc[3]);
            v0 = v1;
            v1 = v2;
            v2 = b[q];
        }
    }
    for (; j < n; j++) {
        const float a = tp[j * ts], b = vp[j + 3];
        acc[0] = __builtin_fmaf(a, v0, acc[0]);
        acc[1] = __builtin_fmaf(a, v1, acc[1]);
        acc[2] = __builtin_fmaf(a, v2, acc[2]);
        acc[3] = __builtin_fmaf(a, b, acc[3]);
        v0 = v1;
        v1 = v2;
        v2 = b;
    }
}

// acc += tp[j] * vp[j] for j = 0..n-1, in order, fused; operands read from
// global memory eight ahead of their multiply-adds.
__device__ __forceinline__ float fma_chain_contig(float acc, const float *__restrict__ tp,
                                                  const float *__restrict__ vp, uint32_t n)
{
    uint32_t j = 0;
    for (; j + 8 <= n; j += 8) {
        float a[8], b[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            a[q] = tp[j + q];
            b[q] = vp[j + q];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) acc = __builtin_fmaf(a[q], b[q], acc);
    }
    for (; j < n; j++) acc = __builtin_fmaf(tp[j], vp[j], acc);
    return acc;
}

// Float32 re-score of one (window, template chunk)'s queued groups when the
// templates do not fit in LDS: one wave, one queued group per lane (four
// consecutive positions of one template, a mask of those flagged).  The four
// chains run together (fma_chain4: each template sample read once for the
// four, the voltage sliding through registers) over the channels the
// template is nonzero on, template rows and voltage read from global memory
// (L1/L2) eight samples ahead; only the flagged positions are stored.  Each
// chain is the reference's: bias, then c ascending, j = 0..N-1, fused
// multiply-adds; a channel where the template is zero adds exactly 0 (the
// running sum starts at the nonzero bias and is never -0), so it is skipped.
// A group whose fourth chain would read past the recording's last sample
// runs its flagged chains one at a time.
__global__ __launch_bounds__(64) void k_rescore_sup(Dims d, const float *__restrict__ V,
                                                    const float *__restrict__ Tm, const float *__restrict__ bias,
                                                    const uint32_t *__restrict__ sup_lo,
                                                    const uint32_t *__restrict__ sup_hi,
                                                    const uint64_t *__restrict__ rq, const uint2 *__restrict__ rdesc,
                                                    float *L)
{
    const uint2 dsc = rdesc[blockIdx.x];
    const uint32_t cnt = dsc.y;
    for (uint32_t e = threadIdx.x; e < cnt; e += 64) {
        const uint64_t ent = rq[dsc.x + e];
        const uint32_t m = (uint32_t)(ent >> 48) & 15u, t = (uint32_t)(ent >> 32) & 0xffffu;
        const uint64_t i0 = (uint32_t)ent;
        const float b = bias[t];
        const uint32_t lo = sup_lo[t], hi = sup_hi[t];
        const float *tt = Tm + (uint64_t)t * d.C * d.N;
        float *lt = L + (uint64_t)t * d.K + i0;
        if (i0 + d.N + 3 <= d.K) {
            float acc[4] = {b, b, b, b};
            for (uint32_t c = lo; c < hi; c++) fma_chain4(acc, tt + (uint64_t)c * d.N, 1, V + (uint64_t)c * d.K + i0, d.N);
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (m >> q & 1) lt[q] = acc[q];
        } else {
            for (int q = 0; q < 4; q++) {
                if (!(m >> q & 1)) continue;
                float acc = b;
                for (uint32_t c = lo; c < hi; c++)
                    acc = fma_chain_contig(acc, tt + (uint64_t)c * d.N, V + (uint64_t)c * d.K + i0 + q, d.N);
                lt[q] = acc;
            }
        }
    }
}

// Same re-score with every template resident in LDS: one block per CU holds
// the [C][N][Tp] copy for its whole life.  Each wave takes a contiguous slice
// of the queue and walks it in batches of up to 64 entries from at most
// kRescoreBatchWins windows (entries of a window are contiguous in the queue),
// staging those windows' voltage one channel at a time in a region of its own
// (wave-level ordering only; the next channel's samples are fetched while the
// current one runs).  Used when the templates fit in LDS.
constexpr uint32_t kRescoreResWaves = 16;
constexpr uint32_t kRescoreStage = 768;    // voltage floats a wave stages per channel
constexpr uint32_t kRescoreWin = 6;        // windows a wave batch stages together

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(1024) void k_rescore_res(Dims d, const float *__restrict__ V,
                                                      const float *__restrict__ TmT, const float *__restrict__ bias,
                                                      const uint64_t *__restrict__ rq,
                                                      const uint32_t *__restrict__ rq_shard, uint32_t rq_cap, float *L)
{
    extern __shared__ float lds[];   // [C][N][Tp] templates, then [waves][kRescoreStage] voltage rows
    const uint32_t N = d.N, Tp = d.Tp, C = d.C;
    const uint32_t ntm = C * N * Tp;
    const uint32_t span = ((N + 15) / 16) * 16 + N;   // window positions (padded) + template length
    const uint32_t nwb = kRescoreStage / span < kRescoreWin ? kRescoreStage / span : kRescoreWin;   // windows per batch
    if (nwb == 0) return;   // the host launches this kernel only for span <= kRescoreStage
    for (uint32_t b = threadIdx.x; b < ntm; b += 8 * blockDim.x) {   // eight loads in flight per thread
        float v[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) v[q] = b + q * blockDim.x < ntm ? TmT[b + q * blockDim.x] : 0.0f;
#pragma unroll
        for (uint32_t q = 0; q < 8; q++)
            if (b + q * blockDim.x < ntm) lds[b + q * blockDim.x] = v[q];
    }
    __syncthreads();
    const uint32_t lane = lane_of(), nw = blockDim.x >> 6;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));   // scalar per-wave state
    float *vb = lds + ntm + wave * kRescoreStage;
    // the shards' kept entries, concatenated: shard j holds virtual entries
    // [spre[j], spre[j + 1]) at rq[j * cap + (v - spre[j])]
    __shared__ uint32_t spre[kQueueShards + 1];
    const uint32_t cap = rq_cap / kQueueShards;
    if (threadIdx.x < kWave) {
        uint32_t v = threadIdx.x < kQueueShards ? min(rq_shard[threadIdx.x * kShardStride], cap) : 0u;
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(v, off);
            if (lane >= (uint32_t)off) v += y;
        }
        if (threadIdx.x < kQueueShards) spre[threadIdx.x + 1] = v;
        if (threadIdx.x == 0) spre[0] = 0;
    }
    __syncthreads();
    const uint32_t n = spre[kQueueShards];
    const uint32_t gw = blockIdx.x * nw + wave, nwave = gridDim.x * nw;
    const uint32_t per = (n + nwave - 1) / nwave;
    const uint32_t e_end = min(n, (gw + 1) * per);
    for (uint32_t b0 = gw * per; b0 < e_end;) {
        const uint32_t e = b0 + lane;
        const bool in = e < e_end;
        uint32_t sj = 0;   // shard of virtual entry e: the last j with spre[j] <= e
        for (uint32_t step = kQueueShards / 2; step > 0; step >>= 1)
            if (spre[sj + step] <= e) sj += step;
        const uint64_t ent = in ? rq[(uint64_t)sj * cap + (e - spre[sj])] : 0;
        const uint32_t t = (uint32_t)(ent >> 32) & 0xffffu, m = (uint32_t)(ent >> 48) & 15u;
        const uint32_t i0 = (uint32_t)ent;
        const uint32_t wid = i0 / N;
        // windows of the batch: ranks of the first entry of each window
        const uint32_t wprev = __shfl_up(wid, 1);
        const bool first = in && (lane == 0 || wprev != wid);
        const uint64_t fm = __ballot(first);
        const uint32_t rank = __popcll(fm & ((2ull << lane) - 1)) - 1;   // in: lane 0 is a first
        const bool ok = in && rank < nwb;
        const uint64_t okm = __ballot(ok);
        uint32_t wb[kRescoreWin];
        uint32_t nwin = 0;
        {
            uint64_t f = fm;
#pragma unroll
            for (int r = 0; r < (int)kRescoreWin; r++) {
                const int l = f ? __builtin_ctzll(f) : 0;
                wb[r] = __shfl(wid, l) * N;
                if (f && (uint32_t)r < nwb) nwin = r + 1;
                f &= f - 1;
            }
        }
        const uint32_t p = ok ? i0 - wid * N + rank * span : 0u;   // offset in the staged rows
        const float b0v = bias[t];
        float acc[4] = {b0v, b0v, b0v, b0v};
        // staged sample x of a channel: window x / span, offset x % span
        constexpr uint32_t kQ = kRescoreStage / 64;
        uint32_t src[kQ];
        bool has[kQ];
#pragma unroll
        for (uint32_t q = 0; q < kQ; q++) {
            const uint32_t x = lane + 64u * q, r = x / span, o = x - r * span;
            const uint32_t rr = r < kRescoreWin ? r : kRescoreWin - 1;
            has[q] = r < nwin && (uint64_t)wb[rr] + o < d.K;
            src[q] = has[q] ? wb[rr] + o : 0u;
        }
        float rv[kQ];
        auto fetch = [&](uint32_t c) {
            const float *vr = V + (uint64_t)c * d.K;
#pragma unroll
            for (uint32_t q = 0; q < kQ; q++) rv[q] = has[q] ? vr[src[q]] : 0.0f;
        };
        fetch(0);
        for (uint32_t c = 0; c < C; c++) {
            wave_sync();   // every lane is done with channel c - 1
#pragma unroll
            for (uint32_t q = 0; q < kQ; q++) vb[lane + 64u * q] = rv[q];
            wave_sync();
            if (c + 1 < C) fetch(c + 1);
            fma_chain4(acc, lds + (uint64_t)c * N * Tp + t, Tp, vb + p, N);
        }
        if (ok)
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (m >> q & 1) L[(uint64_t)t * d.K + i0 + q] = acc[q];
        b0 += __popcll(okm);
    }
}

static size_t rescore_res_lds(const Dims &d)
{
    return ((uint64_t)d.C * d.N * d.Tp + kRescoreResWaves * kRescoreStage) * sizeof(float);
}

__global__ void k_zero_rescore(Counters *ctr, uint32_t *rq_shard)
{
    if (threadIdx.x == 0) {
        ctr->n_rescore = 0;
        ctr->n_rescore_pos = 0;
    }
    if (threadIdx.x < kQueueShards) {
        rq_shard[threadIdx.x * kShardStride] = 0;
        rq_shard[threadIdx.x * kShardStride + 1] = 0;
    }
}

// Totals of the shards (queued entries kept, positions flagged) for the host.
__global__ void k_sum_rescore(Counters *ctr, const uint32_t *rq_shard, uint32_t rq_cap)
{
    const uint32_t cap = rq_cap / kQueueShards;
    uint32_t e = 0, p = 0;
    if (threadIdx.x < kQueueShards) {
        e = min(rq_shard[threadIdx.x * kShardStride], cap);
        p = rq_shard[threadIdx.x * kShardStride + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off);
        p += __shfl_xor(p, off);
    }
    if (threadIdx.x == 0) {
        ctr->n_rescore = e;
        ctr->n_rescore_pos = p;
    }
}

// ---------------------------------------------------------------------------
// Windowed argmax over all templates (bp.cl:350-441, with the host's T
// sequential launches, bpp.py:551-562, fused).  One wave per listed window.
// The reference scans t ascending, then position ascending, with strict >,
// from a best of 0: the winner is the smallest (t, position) holding the
// largest positive value; with no positive value the label and index stay.
// ---------------------------------------------------------------------------
// Window partials: one wave per listed window reads its own N positions of
// every template once and keeps the best (value, t, position) in the
// reference's scan order (t ascending, then position, strict >, from 0).  A
// window's likelihood changes only when the window is listed, so the partials
// of unlisted windows stay current.
__global__ __launch_bounds__(256) void k_window_partials(Dims d, const float *__restrict__ L,
                                                         const uint32_t *__restrict__ win, uint32_t n_win,
                                                         float *part_v, uint32_t *part_t, uint32_t *part_p)
{
    const uint32_t k = blockIdx.x * 4 + (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = lane_of();
    if (k >= n_win) return;
    const uint32_t w = win[k];
    const uint64_t start_w = (uint64_t)w * d.N;
    const uint64_t stop = min(start_w + d.N, (uint64_t)d.K);
    float bv = 0.0f;
    uint32_t bt = 0xffffffffu, bp = 0xffffffffu;
    for (uint64_t p = start_w + lane; p < stop; p += kWave) {
        // all templates of this position, loads first
        for (uint32_t t0 = 0; t0 < d.T; t0 += 8) {
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = t0 + q < d.T ? L[(uint64_t)(t0 + q) * d.K + p] : 0.0f;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint32_t t = t0 + q;
                if (v[q] > bv || (v[q] == bv && v[q] > 0.0f && t < bt)) { bv = v[q]; bt = t; bp = (uint32_t)p; }
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float ov_ = __shfl_xor(bv, off);
        const uint32_t ot = __shfl_xor(bt, off), op = __shfl_xor(bp, off);
        if (ov_ > bv || (ov_ == bv && (ot < bt || (ot == bt && op < bp)))) { bv = ov_; bt = ot; bp = op; }
    }
    if (lane == 0) {
        part_v[w] = bv;
        part_t[w] = bt;
        part_p[w] = bp;
    }
}

// Windowed argmax over all templates (bp.cl:350-441, with the host's T
// sequential launches, bpp.py:551-562, fused): the best of the partials of
// windows w-1, w, w+1, which cover [wN - N, wN + 2N) exactly.  Ties go to the
// smallest (t, position), as the reference's scan order gives.  With no
// positive value the label and index stay.
__global__ __launch_bounds__(256) void k_argmax(Dims d, const uint32_t *__restrict__ win, uint32_t n_win,
                                                const float *__restrict__ part_v, const uint32_t *__restrict__ part_t,
                                                const uint32_t *__restrict__ part_p, float *best_L,
                                                uint32_t *best_lab, uint32_t *best_idx, uint32_t *ov_idx,
                                                uint32_t *ov_lab, uint8_t *nc, uint8_t *ov)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_win) return;
    const uint32_t w = win[k];
    const uint64_t start_w = (uint64_t)w * d.N, end_w = start_w + d.N;
    if (end_w >= (uint64_t)(d.K - d.N)) return;   // bp.cl:380-383
    float bv = 0.0f;
    uint32_t bt = 0xffffffffu, bp = 0xffffffffu;
    for (uint32_t q = (w > 0 ? w - 1 : 0); q <= w + 1; q++) {
        const float v = part_v[q];
        const uint32_t t = part_t[q], p = part_p[q];
        if (v > bv || (v == bv && v > 0.0f && (t < bt || (t == bt && p < bp)))) { bv = v; bt = t; bp = p; }
    }
    const bool chk = part_v[w] > 0.0f;   // a positive value in the own window (bp.cl:396-400)
    uint32_t lab, idx;
    if (bv > 0.0f) { lab = bt; idx = bp; }
    else { lab = best_lab[w]; idx = best_idx[w]; }
    if (chk) nc[w] = 1;
    ov[w] = (bv > 0.0f && idx >= start_w && idx < end_w) ? 1 : 0;
    best_L[w] = bv;
    best_lab[w] = lab;
    best_idx[w] = idx;
    ov_idx[w] = idx;
    ov_lab[w] = lab;
}

// ---------------------------------------------------------------------------
// Ordered stream compaction: count per block, scan the block counts, scatter.
// Output order equals input order, so the reference's np.nonzero lists
// (bpp.py:569, :686) and a deterministic spike record order come out.
// ---------------------------------------------------------------------------
template <class Pred>
__global__ __launch_bounds__(256) void k_compact_count(Pred pred, uint64_t n, uint32_t *blk)
{
    __shared__ uint32_t part[4];
    const uint64_t base = (uint64_t)blockIdx.x * kCompactItems;
    uint32_t cnt = 0;
    for (int r = 0; r < kCompactItems / 256; r++) {
        const uint64_t i = base + r * 256 + threadIdx.x;
        if (i < n && pred(i)) cnt++;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if (lane_of() == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of nb block counts by one 1024-thread block; total to *out.
__global__ __launch_bounds__(1024) void k_compact_scan(uint32_t *blk, uint32_t nb, uint32_t *total)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? blk[i] : 0;
        uint32_t x = v;
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane_of() >= off) x += y;
        }
        if (lane_of() == kWave - 1) wsum[threadIdx.x >> 6] = x;
        __syncthreads();
        uint32_t wpre = 0;
        for (int q = 0; q < (int)(threadIdx.x >> 6); q++) wpre += wsum[q];
        const uint32_t c = carry;
        if (i < nb) blk[i] = c + wpre + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + wpre + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

template <class Pred, class Emit>
__global__ __launch_bounds__(256) void k_compact_scatter(Pred pred, Emit emit, uint64_t n, const uint32_t *blk)
{
    __shared__ uint32_t wsum[4];
    constexpr int per = kCompactItems / 256;
    const uint64_t base = (uint64_t)blockIdx.x * kCompactItems + (uint64_t)threadIdx.x * per;
    bool f[per];
    uint32_t cnt = 0;
#pragma unroll
    for (int r = 0; r < per; r++) {
        f[r] = (base + r < n) && pred(base + r);
        cnt += f[r] ? 1 : 0;
    }
    uint32_t x = cnt;
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane_of() >= off) x += y;
    }
    if (lane_of() == kWave - 1) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    uint32_t pos = blk[blockIdx.x] + x - cnt;
    for (int q = 0; q < (int)(threadIdx.x >> 6); q++) pos += wsum[q];
#pragma unroll
    for (int r = 0; r < per; r++)
        if (f[r]) emit(base + r, pos++);
}

template <class Pred, class Emit>
hipError_t compact(Pred pred, Emit emit, uint64_t n, uint32_t *blk, uint32_t *total, hipStream_t st)
{
    const uint32_t nb = (uint32_t)((n + kCompactItems - 1) / kCompactItems);
    if (nb > 0) hipLaunchKernelGGL(k_compact_count<Pred>, dim3(nb), dim3(256), 0, st, pred, n, blk);
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(1024), 0, st, blk, nb, total);
    if (nb > 0) hipLaunchKernelGGL((k_compact_scatter<Pred, Emit>), dim3(nb), dim3(256), 0, st, pred, emit, n, blk);
    return hipGetLastError();
}

struct OvPred {
    const uint32_t *win; const uint8_t *ov;
    __device__ bool operator()(uint64_t i) const { return (ov[win[i]] & 1) != 0; }
};
struct OvEmit {
    const uint32_t *win; uint32_t *out;
    __device__ void operator()(uint64_t i, uint32_t pos) const { out[pos] = win[i]; }
};
struct SpikePred {
    const uint8_t *has;
    __device__ bool operator()(uint64_t i) const { return has[i] != 0; }
};
struct SpikeEmit {
    const uint32_t *win; const uint32_t *cand_idx; const uint32_t *cand_lab;
    uint32_t *sp_idx; uint32_t *sp_lab; uint32_t *win_spike;
    __device__ void operator()(uint64_t i, uint32_t pos) const {
        sp_idx[pos] = cand_idx[i];
        sp_lab[pos] = cand_lab[i];
        win_spike[win[i]] = pos + 1;
    }
};
struct NextPred {
    const uint8_t *nc;
    __device__ bool operator()(uint64_t w) const { const uint8_t v = nc[w]; return (v & 3) != 0 && !(v & 4); }
};
struct NextEmit {
    uint32_t *out;
    __device__ void operator()(uint64_t w, uint32_t pos) const { out[pos] = (uint32_t)w; }
};

// ---------------------------------------------------------------------------
// Overlap recheck, direct form (bp.cl:444-654 + parse bp.cl:657-815).
// One 256-thread block per overlap window (grid-stride over the device-side
// count).  Threads take combinations in (t, f, s) order with s fastest so a
// wave's template reads are contiguous; each combination keeps the reference
// work id p = t + T*(f + S*s).  The reference's winner is the largest score;
// among equal scores, the earliest reduction group g = p / WG (parse scans
// groups with strict >), then inside the group the max_local_reduction tree
// (bp.cl:167-186) keeps the lane with the smallest bit-reversed local id.
// That total order is reduced here in any order.  Leftover work items of the
// last group (p >= items) repeat combination p mod items; they can only win
// inside a single-group window, where parse then rejects them (bp.cl:714-718).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t bitrev_bits(uint32_t x, uint32_t bits)
{
    return bits == 0 ? 0u : (__builtin_bitreverse32(x) >> (32 - bits));
}

__device__ __forceinline__ bool key_better(float a, uint64_t ka, float b, uint64_t kb)
{
    return a > b || (a == b && ka < kb);
}

// Winner-order key of one scored combination: (reduction group, bit-reversed
// lane) -- smaller is better among equal scores.  p_out is the reference work id
// that carries the key (a leftover duplicate of a single-group window may
// carry it, which parse then rejects).
__device__ __forceinline__ uint64_t combo_key(const Dims &d, uint32_t wg_bits, uint64_t span, uint32_t t,
                                              uint32_t f, uint32_t s, uint64_t &p_out)
{
    uint64_t p = (uint64_t)t + (uint64_t)d.T * (f + (uint64_t)d.S * s);
    const uint64_t g = p / d.reduce_wg;
    uint32_t br = bitrev_bits((uint32_t)(p % d.reduce_wg), wg_bits);
    for (uint64_t p2 = p + d.items; p2 < span && p2 / d.reduce_wg == g; p2 += d.items) {
        const uint32_t br2 = bitrev_bits((uint32_t)(p2 % d.reduce_wg), wg_bits);
        if (br2 < br) { br = br2; p = p2; }
    }
    p_out = p;
    return (g << 32) | br;
}

// parse_overlap_recheck_indices (bp.cl:690-814) for one window, given the best
// scored combination (score sc, reference work id bp).  Thread 0 only.
// The branches of the parse after the score comparisons (bp.cl:739-814):
// returns whether the window keeps its overlap flag.
__device__ __forceinline__ bool parse_choose(const Dims &d, const float *__restrict__ thr, uint32_t id, uint32_t lab,
                                             uint32_t bidx, uint32_t t, uint32_t af, uint32_t as, float Lm, float Lo,
                                             float *best_L, uint32_t *ov_idx, uint32_t *ov_lab)
{
    if (Lm <= 0.0f || Lo <= 0.0f) return false;                                          // bp.cl:739-744
    if (Lm >= Lo && Lm > thr[lab]) { best_L[id] = Lm; ov_lab[id] = lab; ov_idx[id] = af; return true; }
    if (Lo > Lm && Lo > thr[t]) { best_L[id] = Lo; ov_lab[id] = t; ov_idx[id] = as; return true; }
    if (Lm > thr[lab]) { best_L[id] = Lm; ov_lab[id] = lab; ov_idx[id] = af; return true; }
    ov_idx[id] = bidx;                                                                     // bp.cl:806-814
    ov_lab[id] = lab;
    return false;
}

__device__ __forceinline__ void parse_overlap(const Dims &d, const float *__restrict__ L, const float *__restrict__ thr, uint32_t id,
                              uint32_t lab, uint32_t bidx, float bl, float sc, uint64_t bp, float *best_L,
                              uint32_t *ov_idx, uint32_t *ov_lab, uint8_t *ov)
{
    const uint32_t S = d.S, T = d.T, M = d.M;
    const uint64_t span = d.nl * d.reduce_wg;
    bool keep = sc > bl && sc > 0.0f;                          // bp.cl:707-711
    if (keep && bp % span >= d.items) keep = false;            // bp.cl:714-718
    if (keep) {
        const uint32_t t = (uint32_t)(bp % T);
        const uint32_t f = (uint32_t)((bp / T) % S);
        const uint32_t s = (uint32_t)((bp / ((uint64_t)S * T)) % S);
        const uint32_t af = bidx + f - M, as = bidx + s - M;
        const float Lm = L[(uint64_t)lab * d.K + af];
        const float Lo = L[(uint64_t)t * d.K + as];
        if (sc < Lm || sc < Lo) keep = false;                  // bp.cl:733-738
        else keep = parse_choose(d, thr, id, lab, bidx, t, af, as, Lm, Lo, best_L, ov_idx, ov_lab);
    }
    if (!keep) ov[id] = 0;
}

// The same parse for a winner whose float32 score is only known to lie in
// [lo, hi]: returns false, writing nothing, when a comparison of the score is
// left open by the interval.
__device__ __forceinline__ bool parse_overlap_bounds(const Dims &d, const float *__restrict__ L, const float *__restrict__ thr,
                                     uint32_t id, uint32_t lab, uint32_t bidx, float bl, double lo, double hi,
                                     uint64_t bp, float *best_L, uint32_t *ov_idx, uint32_t *ov_lab, uint8_t *ov)
{
    const uint32_t S = d.S, T = d.T, M = d.M;
    const uint64_t span = d.nl * d.reduce_wg;
    const double g0 = fmax((double)bl, 0.0);                   // keep needs sc > bl and sc > 0
    if (hi <= g0 || bp % span >= d.items) { ov[id] = 0; return true; }
    if (!(lo > g0)) return false;
    const uint32_t t = (uint32_t)(bp % T);
    const uint32_t f = (uint32_t)((bp / T) % S);
    const uint32_t s = (uint32_t)((bp / ((uint64_t)S * T)) % S);
    const uint32_t af = bidx + f - M, as = bidx + s - M;
    const float Lm = L[(uint64_t)lab * d.K + af];
    const float Lo = L[(uint64_t)t * d.K + as];
    const double g1 = fmax((double)Lm, (double)Lo);            // dropped if sc < Lm or sc < Lo
    if (hi < g1) { ov[id] = 0; return true; }
    if (!(lo >= g1)) return false;
    if (!parse_choose(d, thr, id, lab, bidx, t, af, as, Lm, Lo, best_L, ov_idx, ov_lab)) ov[id] = 0;
    return true;
}

// Block-wide reduction of (score, key, work id) in winner order; the result is
// valid in thread 0.
__device__ __forceinline__ void block_key_reduce(float &sc, uint64_t &key, uint64_t &bp, float *r_sc, uint64_t *r_key, uint64_t *r_p)
{
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(sc, off);
        const uint64_t ok = __shfl_xor(key, off), op = __shfl_xor(bp, off);
        if (key_better(o, ok, sc, key)) { sc = o; key = ok; bp = op; }
    }
    if (lane_of() == 0) { r_sc[threadIdx.x >> 6] = sc; r_key[threadIdx.x >> 6] = key; r_p[threadIdx.x >> 6] = bp; }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 1; q < (int)(blockDim.x >> 6); q++)
            if (key_better(r_sc[q], r_key[q], sc, key)) { sc = r_sc[q]; key = r_key[q]; bp = r_p[q]; }
}

// Reference pair score (bp.cl:599-631) from zero-padded template rows:
// a_j = pad_lab[c][j + 2M - f], b_j = pad_t[c][j + 2M - s], where a row holds
// the template at [M, M + N) and +0.0 elsewhere -- the value the reference's
// out-of-range branches use -- so every sum a_j + b_j is the reference's.
__device__ __forceinline__ float overlap_direct_score(const Dims &d, const float *__restrict__ sv,
                                                      const float *__restrict__ amain, const float *__restrict__ bt,
                                                      uint32_t f, uint32_t s)
{
    const uint32_t R = d.N + 2 * d.M;
    float acc = 0.0f;
    for (uint32_t c = 0; c < d.C; c++) {
        const float *a = amain + c * R + 2 * d.M - f;
        const float *b = bt + (uint64_t)c * R + 2 * d.M - s;
        const float *vc = sv + c * d.N;
        float ss = 0.0f;
        // operands are loaded eight samples at a time ahead of the chains
        uint32_t j = 0;
        for (; j + 8 <= d.N; j += 8) {
            float av[8], bv[8], vv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                av[q] = a[j + q];
                bv[q] = b[j + q];
                vv[q] = vc[j + q];
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const float sum = av[q] + bv[q];
                ss = __builtin_fmaf(sum, sum, ss);       // bp.cl:621
                acc = __builtin_fmaf(sum, vv[q], acc);   // bp.cl:622
            }
        }
        for (; j < d.N; j++) {
            const float sum = a[j] + b[j];
            ss = __builtin_fmaf(sum, sum, ss);
            acc = __builtin_fmaf(sum, vc[j], acc);
        }
        if (ss > 0.0f) acc = (float)((double)acc - 0.5 * (double)ss);   // bp.cl:627-630
    }
    return acc;
}

// The same score over the channels where either template is nonzero, with
// the rows outside a template's support read from a zero row in LDS (the
// four-wave screen, whose template rows are in LDS): the
// main template's rows amain[c - lo_m] for c in [lo_m, hi_m), the partner's
// bt[c - lo_t] for c in [lo_t, hi_t), a zero row elsewhere.  A channel where
// both are zero adds exactly 0 to the reference chain (every product is +-0,
// the running sum is never -0 and ss = 0 skips the subtraction), so it is
// skipped.
__device__ __forceinline__ float overlap_support_score_z(const Dims &d, const float *__restrict__ sv,
                                                       const float *amain, uint32_t lo_m, uint32_t hi_m,
                                                       const float *bt, uint32_t lo_t, uint32_t hi_t,
                                                       const float *zero, uint32_t f, uint32_t s)
{
    const uint32_t R = d.N + 2 * d.M;
    const uint32_t c0 = min(lo_m, lo_t), c1 = max(hi_m, hi_t);
    float acc = 0.0f;
    for (uint32_t c = c0; c < c1; c++) {
        const bool ha = c >= lo_m && c < hi_m, hb = c >= lo_t && c < hi_t;
        if (!ha && !hb) continue;
        const float *a = (ha ? amain + (c - lo_m) * R : zero) + 2 * d.M - f;
        const float *b = (hb ? bt + (uint64_t)(c - lo_t) * R : zero) + 2 * d.M - s;
        const float *vc = sv + c * d.N;
        float ss = 0.0f;
        uint32_t j = 0;
        for (; j + 8 <= d.N; j += 8) {
            float av[8], bv[8], vv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                av[q] = a[j + q];
                bv[q] = b[j + q];
                vv[q] = vc[j + q];
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const float sum = av[q] + bv[q];
                ss = __builtin_fmaf(sum, sum, ss);       // bp.cl:621
                acc = __builtin_fmaf(sum, vv[q], acc);   // bp.cl:622
            }
        }
        for (; j < d.N; j++) {
            const float sum = a[j] + b[j];
            ss = __builtin_fmaf(sum, sum, ss);
            acc = __builtin_fmaf(sum, vc[j], acc);
        }
        if (ss > 0.0f) acc = (float)((double)acc - 0.5 * (double)ss);   // bp.cl:627-630
    }
    return acc;
}

// The same score over the channels where either template is nonzero (the
// window's voltage row c at sv + c * vstr): the
// main template's rows amain[c - lo_m] for c in [lo_m, hi_m), the partner's
// bt[c - lo_t] for c in [lo_t, hi_t), +0.0 elsewhere.  A channel where
// both are zero adds exactly 0 to the reference chain (every product is +-0,
// the running sum is never -0 and ss = 0 skips the subtraction), so it is
// skipped.
__device__ __forceinline__ float overlap_support_score(const Dims &d, const float *__restrict__ sv, uint64_t vstr,
                                                       const float *amain, uint32_t lo_m, uint32_t hi_m,
                                                       const float *bt, uint32_t lo_t, uint32_t hi_t,
                                                       uint32_t f, uint32_t s)
{
    const uint32_t R = d.N + 2 * d.M;
    const uint32_t c0 = min(lo_m, lo_t), c1 = max(hi_m, hi_t);
    float acc = 0.0f;
    for (uint32_t c = c0; c < c1; c++) {
        const bool ha = c >= lo_m && c < hi_m, hb = c >= lo_t && c < hi_t;
        if (!ha && !hb) continue;
        const float *a = amain + (uint64_t)(ha ? c - lo_m : 0u) * R + 2 * d.M - f;
        const float *b = bt + (uint64_t)(hb ? c - lo_t : 0u) * R + 2 * d.M - s;
        const float *vc = sv + c * vstr;
        float ss = 0.0f;
        // operands are loaded eight samples at a time ahead of the chains
        uint32_t j = 0;
        for (; j + 8 <= d.N; j += 8) {
            float av[8], bv[8], vv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                av[q] = ha ? a[j + q] : 0.0f;
                bv[q] = hb ? b[j + q] : 0.0f;
                vv[q] = vc[j + q];
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const float sum = av[q] + bv[q];
                ss = __builtin_fmaf(sum, sum, ss);       // bp.cl:621
                acc = __builtin_fmaf(sum, vv[q], acc);   // bp.cl:622
            }
        }
        for (; j < d.N; j++) {
            const float sum = (ha ? a[j] : 0.0f) + (hb ? b[j] : 0.0f);
            ss = __builtin_fmaf(sum, sum, ss);
            acc = __builtin_fmaf(sum, vc[j], acc);
        }
        if (ss > 0.0f) acc = (float)((double)acc - 0.5 * (double)ss);   // bp.cl:627-630
    }
    return acc;
}

// Tpad[t][c][y] = T[t][c][y - M] for y - M in [0, N), +0.0 otherwise (y < N + 2M).
__global__ void k_pad_templates(Dims d, const float *__restrict__ Tm, float *__restrict__ Tpad)
{
    const uint32_t R = d.N + 2 * d.M;
    const uint64_t total = (uint64_t)d.T * d.C * R;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t y = (uint32_t)(e % R);
        const uint64_t tc = e / R;
        Tpad[e] = (y >= d.M && y - d.M < d.N) ? Tm[tc * d.N + (y - d.M)] : 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_overlap_direct(
    Dims d, const float *__restrict__ V, const float *__restrict__ Tpad, const float *__restrict__ L,
    const float *__restrict__ thr, const uint32_t *__restrict__ ov_list, Counters *ctr_rw,
    float *best_L, const uint32_t *best_lab, const uint32_t *best_idx, uint32_t *ov_idx, uint32_t *ov_lab,
    uint8_t *ov)
{
    const Counters *ctr = ctr_rw;
    extern __shared__ float dyn[];
    float *sv = dyn;                       // [C][N] voltage at the best onset
    float *smain = dyn + d.C * d.N;        // [C][N + 2M] padded template of the best label
    const uint32_t R = d.N + 2 * d.M;
    __shared__ float r_sc[4];
    __shared__ uint64_t r_key[4];
    __shared__ uint64_t r_p[4];
    const uint32_t n_ov = ctr->n_ov;
    const uint32_t wg_bits = 31 - __builtin_clz(d.reduce_wg);
    const uint64_t span = d.nl * d.reduce_wg;
    const uint32_t S = d.S, N = d.N, M = d.M;
    for (uint32_t k = blockIdx.x; k < n_ov; k += gridDim.x) {
        const uint32_t id = ov_list[k];
        const uint32_t lab = best_lab[id], bidx = best_idx[id];
        const float bl = best_L[id];
        const float Lb = L[(uint64_t)lab * d.K + bidx];
        const bool wskip = (Lb < thr[lab]) || ((uint64_t)bidx >= (uint64_t)(d.K - d.N));   // bp.cl:518-527
        float sc = -INFINITY;
        uint64_t key = ~0ull, bp = 0;
        uint32_t scored = 0;
        if (!wskip) {
            __syncthreads();
            for (uint32_t e = threadIdx.x; e < d.C * N; e += blockDim.x) {
                const uint32_t c = e / N, j = e % N;
                sv[e] = V[(uint64_t)c * d.K + bidx + j];
            }
            for (uint32_t e = threadIdx.x; e < d.C * R; e += blockDim.x) smain[e] = Tpad[(uint64_t)lab * d.C * R + e];
            __syncthreads();
            const float *Lmain = L + (uint64_t)lab * d.K;
            for (uint64_t q = threadIdx.x; q < d.items; q += blockDim.x) {
                const uint32_t s = (uint32_t)(q % S);
                const uint32_t f = (uint32_t)((q / S) % S);
                const uint32_t t = (uint32_t)(q / ((uint64_t)S * S));
                if ((uint64_t)f + bidx < M || (uint64_t)f + bidx >= (uint64_t)d.K - N + M) continue;   // bp.cl:534
                if ((uint64_t)s + bidx < M || (uint64_t)s + bidx >= (uint64_t)d.K - N + M) continue;   // bp.cl:538
                if ((f > s ? f - s : s - f) > N) continue;                                              // bp.cl:543-556
                const float Lf = Lmain[bidx + f - M];
                const float Ls = L[(uint64_t)t * d.K + bidx + s - M];
                if (Lf <= 0.0f && Ls <= 0.0f) continue;                                                 // bp.cl:572-576
                const float score = overlap_direct_score(d, sv, smain, Tpad + (uint64_t)t * d.C * R, f, s);
                scored++;
                uint64_t p;
                const uint64_t kk = combo_key(d, wg_bits, span, t, f, s, p);
                if (key_better(score, kk, sc, key)) { sc = score; key = kk; bp = p; }
            }
            for (int off = 32; off > 0; off >>= 1) scored += __shfl_xor(scored, off);
            if (lane_of() == 0 && scored) atomicAdd(&ctr_rw->combos, (unsigned long long)scored);
            block_key_reduce(sc, key, bp, r_sc, r_key, r_p);
        }
        if (threadIdx.x == 0) parse_overlap(d, L, thr, id, lab, bidx, bl, sc, bp, best_L, ov_idx, ov_lab, ov);
    }
}

// ---------------------------------------------------------------------------
// Overlap recheck, screened form.  The pair score of bp.cl:599-631 is, in exact
// arithmetic, a sum of terms that separate:
//   score(t,f,s) = Q(lab,f) + Q(t,s) - B2(lab,f)/2 - B2(t,s)/2 - AB(lab,f,t,s)
// with Q(t,s) = <template t shifted by s-M, truncated to the window ; voltage>
// (per window, T*S values), B2(t,s) its squared norm and AB the cross product of
// two shifted, truncated templates (both voltage-free tables built once per
// template set).  The screen evaluates that sum in float64 for every
// combination the reference does not skip, and bounds |screen - float32 score|
// by E = ecoef * ((|t_lab|+|t_t|)*|v| + (|t_lab|+|t_t|)^2), which covers the
// float32 rounding of the reference's fused chain (gamma_{CN+2C+8} with a
// doubled unit roundoff) and the float32 storage of Q and the tables.  Every
// combination whose upper bound reaches the largest lower bound is then scored
// exactly with the reference's float32 chain (overlap_direct_score) and reduced
// in the reference's winner order, so the result equals the direct form's.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_overlap_tables(Dims d, const float *__restrict__ Tm, float *__restrict__ AB,
                                                       float *__restrict__ B2, float *__restrict__ tnorm,
                                                       const uint32_t *__restrict__ sup_lo,
                                                       const uint32_t *__restrict__ sup_hi)
{
    const uint32_t S = d.S, T = d.T, N = d.N, M = d.M, C = d.C;
    const uint64_t nab = (uint64_t)T * S * T * S;
    const uint64_t total = nab + (uint64_t)T * S + T;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        if (e < nab) {
            const uint32_t s = (uint32_t)(e % S);
            const uint32_t t = (uint32_t)((e / S) % T);
            const uint32_t f = (uint32_t)((e / ((uint64_t)S * T)) % S);
            const uint32_t lab = (uint32_t)(e / ((uint64_t)S * T * S));
            const int32_t fa = (int32_t)M - (int32_t)f, sa = (int32_t)M - (int32_t)s;
            const int32_t lo = max(0, max(-fa, -sa));
            const int32_t hi = min((int32_t)N, min((int32_t)N - fa, (int32_t)N - sa));
            double acc = 0.0;   // channels where both templates can be nonzero
            for (uint32_t c = max(sup_lo[lab], sup_lo[t]); c < min(sup_hi[lab], sup_hi[t]); c++) {
                const float *a = Tm + ((uint64_t)lab * C + c) * N;
                const float *b = Tm + ((uint64_t)t * C + c) * N;
                for (int32_t j = lo; j < hi; j++) acc = __builtin_fma((double)a[j + fa], (double)b[j + sa], acc);
            }
            AB[e] = (float)acc;
        } else if (e < nab + (uint64_t)T * S) {
            const uint64_t q = e - nab;
            const uint32_t s = (uint32_t)(q % S), t = (uint32_t)(q / S);
            const int32_t sa = (int32_t)M - (int32_t)s;
            const int32_t lo = max(0, -sa), hi = min((int32_t)N, (int32_t)N - sa);
            double acc = 0.0;
            for (uint32_t c = sup_lo[t]; c < sup_hi[t]; c++) {
                const float *b = Tm + ((uint64_t)t * C + c) * N;
                for (int32_t j = lo; j < hi; j++) acc = __builtin_fma((double)b[j + sa], (double)b[j + sa], acc);
            }
            B2[q] = (float)acc;
        } else {
            const uint32_t t = (uint32_t)(e - nab - (uint64_t)T * S);
            double acc = 0.0;
            for (uint64_t k = 0; k < (uint64_t)C * N; k++) {
                const double x = Tm[(uint64_t)t * C * N + k];
                acc = __builtin_fma(x, x, acc);
            }
            // upper bound of the norm, rounded up to float
            const double nrm = sqrt(acc) * (1.0 + 1e-12);
            float up = (float)nrm;
            if ((double)up < nrm) up = __uint_as_float(__float_as_uint(up) + 1u);
            tnorm[t] = up;
        }
    }
}

// Block-wide staging loop: for x in [0, n) with the block's threads striding,
// U loads are issued before their U stores, so their latencies overlap.
template <int U, class Ld, class St>
__device__ __forceinline__ void batched_stage(uint32_t n, Ld ld, St st)
{
    for (uint32_t b0 = threadIdx.x; b0 < n; b0 += U * blockDim.x) {
        float v[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const uint32_t x = b0 + q * blockDim.x;
            v[q] = x < n ? ld(x) : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            const uint32_t x = b0 + q * blockDim.x;
            if (x < n) st(x, v[q]);
        }
    }
}

// One array of a joint staging: U elements per thread and round.
template <int U, class Ld, class St>
struct StageArr {
    uint32_t n;
    Ld ld;
    St st;
    float v[U];
    __device__ __forceinline__ uint32_t rounds() const { return (n + U * blockDim.x - 1) / (U * blockDim.x); }
    __device__ __forceinline__ void load(uint32_t r)
    {
#pragma unroll
        for (int q = 0; q < U; q++) {
            const uint32_t x = threadIdx.x + (r * U + q) * blockDim.x;
            v[q] = x < n ? ld(x) : 0.0f;
        }
    }
    __device__ __forceinline__ void store(uint32_t r)
    {
#pragma unroll
        for (int q = 0; q < U; q++) {
            const uint32_t x = threadIdx.x + (r * U + q) * blockDim.x;
            if (x < n) st(x, v[q]);
        }
    }
};

template <int U, class Ld, class St>
__device__ __forceinline__ StageArr<U, Ld, St> stage_arr(uint32_t n, Ld ld, St st)
{
    return StageArr<U, Ld, St>{n, ld, st, {}};
}

// Stages several arrays together: each round issues the loads of every array
// before it stores any, so their latencies overlap.
template <class... A>
__device__ __forceinline__ void batched_stage_all(A &...a)
{
    uint32_t nr = 0;
    ((nr = a.rounds() > nr ? a.rounds() : nr), ...);
    for (uint32_t r = 0; r < nr; r++) {
        (a.load(r), ...);
        (a.store(r), ...);
    }
}

// x / d as floor(x * ceil(2^32 / d) / 2^32): exact while x * d < 2^32.
struct FastDiv {
    uint64_t m;
    uint32_t d;
    __device__ explicit FastDiv(uint32_t dv) : m(((1ull << 32) + dv - 1) / dv), d(dv) {}
    __device__ uint32_t div(uint32_t x) const { return (uint32_t)(((uint64_t)x * m) >> 32); }
};

// Calls fn(f, t, s, e = t*S + s, AB[lab][f][t][s], sQ[e], sE[t]) for every
// combination of one window that the reference does not skip (bp.cl:534-576)
// and whose row (main shift f, template t) is listed in krow, in no
// particular order.  Rows r = a*T + t (f = flist[a]) of positive main shifts
// are krow[0, n_kp) and pair with every in-bounds s within N of f; rows of the
// other in-bounds main shifts are krow[TS-1], krow[TS-2], ... (n_kn of them)
// and pair with the s of template t whose likelihood is positive, listed at
// plist[pofs[t], pofs[t] + pcnt[t]) (at most maxcnt).  Each thread loads the
// operands of B combinations (global AB, LDS sQ, sE, spos) before it
// evaluates them, so their latencies overlap.  Row and column indices advance
// incrementally; no division per element.  sE holds the bound per template
// as float64 or, rounded up, as float32 (ET).
template <int B, class ET, class Fn>
__device__ __forceinline__ void for_each_kept(const Dims &d, const FastDiv &divT, const uint8_t *spos,
                                              const uint32_t *flist, const uint16_t *krow, uint32_t n_kp,
                                              uint32_t n_kn, const uint16_t *plist, const uint16_t *pofs,
                                              const uint16_t *pcnt, uint32_t maxcnt, const float *__restrict__ ABl,
                                              const float *sQ, const ET *sE, Fn fn)
{
    const uint32_t S = d.S, N = d.N, T = d.T, TS = d.T * d.S;
    const uint32_t step = blockDim.x;
    {
        const uint32_t total = n_kp * S;
        const uint32_t stK = step / S, stS = step - stK * S;
        uint32_t k = threadIdx.x / S, s = threadIdx.x - k * S;
        for (uint32_t x0 = threadIdx.x; x0 < total; x0 += B * step) {
            float ab[B], qe[B];
            ET Et[B];
            uint32_t fq[B], tq[B], sq[B];
            bool okq[B];
#pragma unroll
            for (int q = 0; q < B; q++) {
                const bool in = x0 + q * step < total;
                const uint32_t r = in ? krow[k] : 0u;
                const uint32_t a = divT.div(r), t = r - a * T, f = flist[a], e = t * S + s;
                const bool ok = in && (spos[e] & 2) && (f > s ? f - s : s - f) <= N;
                ab[q] = ok ? ABl[(uint64_t)f * TS + e] : 0.0f;
                qe[q] = ok ? sQ[e] : 0.0f;
                Et[q] = ok ? sE[t] : (ET)0;
                fq[q] = f;
                tq[q] = t;
                sq[q] = s;
                okq[q] = ok;
                s += stS;
                k += stK;
                if (s >= S) { s -= S; k++; }
            }
#pragma unroll
            for (int q = 0; q < B; q++)
                if (okq[q]) fn(fq[q], tq[q], sq[q], tq[q] * S + sq[q], ab[q], qe[q], Et[q]);
        }
    }
    const uint32_t total = n_kn * maxcnt;
    if (total == 0) return;
    const uint32_t stK = step / maxcnt, stJ = step - stK * maxcnt;
    uint32_t k = threadIdx.x / maxcnt, j = threadIdx.x - k * maxcnt;
    for (uint32_t x0 = threadIdx.x; x0 < total; x0 += B * step) {
        float ab[B], qe[B];
        ET Et[B];
        uint32_t fq[B], eq[B], tq[B];
        bool okq[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const bool in = x0 + q * step < total;
            const uint32_t r = in ? krow[TS - 1 - k] : 0u;
            const uint32_t a = divT.div(r), t = r - a * T, f = flist[a];
            const bool has = in && j < pcnt[t];
            const uint32_t e = has ? plist[pofs[t] + j] : 0u, s = e - t * S;
            const bool ok = has && (f > s ? f - s : s - f) <= N;
            ab[q] = ok ? ABl[(uint64_t)f * TS + e] : 0.0f;
            qe[q] = ok ? sQ[e] : 0.0f;
            Et[q] = ok ? sE[t] : (ET)0;
            fq[q] = f;
            eq[q] = e;
            tq[q] = t;
            okq[q] = ok;
            j += stJ;
            k += stK;
            if (j >= maxcnt) { j -= maxcnt; k++; }
        }
#pragma unroll
        for (int q = 0; q < B; q++)
            if (okq[q]) fn(fq[q], tq[q], eq[q] - tq[q] * S, eq[q], ab[q], qe[q], Et[q]);
    }
}

// Q table on the matrix cores in split bfloat16:
//   Q[t][s] = sum over (c, x < N) of T[t][c][x] * vpad[c][x + s],
// vpad[c][y] = v[c][y - M] for y - M in [0, N), 0 elsewhere.  Every float32
// operand is split into hi = bf16(x) and lo = bf16(x - hi) (both round to
// nearest even; x - hi is exact in float32), and
//   Q ~= sum hi_T*hi_v  +  sum (hi_T*lo_v + lo_T*hi_v)
// on v_mfma_f32_16x16x16bf16_1k (bf16 products are exact in float32), in two
// accumulators.  The omitted terms lo*lo and the representation error of lo
// are below 3.1 * 2^-16 * |T| * |v| per product.
//
// Work: tiles of 16 templates x 16 shifts.  One task = one 16-template row
// tile, up to four shift tiles (each A fragment feeds up to 12 MFMAs), and one
// contiguous slice of the channels when there are too few tiles to occupy the
// block's waves; slice partials are added in slice order afterwards.
constexpr uint32_t kOvWaves = 4;   // waves of a k_overlap_q block

// With at most four tiles, the templates past the last full 16-row tile
// (TR = T % 16 of them, TR <= 8) share one tile as ND shift offsets: row g*TR + i holds template
// TTF*16 + i moved right by 16*g samples, so one 16-column shift tile yields
// shifts 16*(st0 + g) + col of all TR templates over NKQ + ND - 1 K-steps.
// Every output still meets nonzero products in exactly NKQ K-steps (the other
// steps add exact zeros).
struct QPlan {
    uint32_t TT, ST, NKQ, SG, NSG, KS, LPH;
    uint32_t TTF, TR, ND, NTG, NKT;   // full tiles; tail rows, offsets, groups, K-steps
};

__host__ __device__ inline QPlan q_plan(const Dims &d)
{
    QPlan p;
    p.TT = (d.T + 15) / 16;
    p.ST = (d.S + 15) / 16;
    p.NKQ = (d.N + 15) / 16;
    p.SG = p.ST < 4 ? p.ST : 4;
    p.NSG = (p.ST + p.SG - 1) / p.SG;
    const uint32_t rem = d.T % 16;
    if (rem > 0 && rem <= 8 && p.ST >= 2 && p.TT <= 4) {   // where the last tile's padding is a large share
        p.TTF = p.TT - 1;
        p.TR = rem;
        p.ND = 16 / rem < p.ST ? 16 / rem : p.ST;
        p.NTG = (p.ST + p.ND - 1) / p.ND;
        p.NKT = p.NKQ + p.ND - 1;
    } else {
        p.TTF = p.TT;
        p.TR = p.ND = p.NTG = p.NKT = 0;
    }
    const uint32_t base = p.TTF * p.NSG + p.NTG;
    uint32_t ks = base < kOvWaves ? kOvWaves / base : 1;
    p.KS = ks < d.C ? ks : d.C;
    // bf16 elements per channel row of the split voltage: every lane reads the
    // three dwords holding samples y..y+3, y < 16*(NKQ + ST) - 4
    p.LPH = 16 * (p.NKQ + p.ST) + 8;
    return p;
}

__device__ __forceinline__ uint64_t q_bfrag(const uint32_t *row, uint32_t y)
{
    // samples y..y+3 of a bf16 row, sample y in the low half of the low dword
    const uint32_t w = y >> 1, sh = 16u * (y & 1u);
    const uint32_t d0 = row[w], d1 = row[w + 1], d2 = row[w + 2];
    const uint32_t lo = __builtin_amdgcn_alignbit(d1, d0, sh);
    const uint32_t hi = __builtin_amdgcn_alignbit(d2, d1, sh);
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// The same four samples from a row copy in which they start at an even
// sample (two aligned dwords, no shifts).  With QC = 2 the voltage is staged
// twice, the second copy moved left by one sample; a lane's B fragments all
// start at samples of its lane's parity, so it reads the copy of that parity.
__device__ __forceinline__ uint64_t q_bfrag_even(const uint32_t *row, uint32_t y)
{
    const uint32_t w = y >> 1;
    return (uint64_t)row[w] | ((uint64_t)row[w + 1] << 32);
}

template <int QC>
__device__ __forceinline__ uint64_t q_frag(const uint16_t *svc, uint32_t rsel, uint32_t c, uint32_t LPH, uint32_t y)
{
    const uint32_t *row = (const uint32_t *)(svc + c * LPH);
    return QC == 2 ? q_bfrag_even(row, y - rsel) : q_bfrag(row, y);
}

// The tail tile's tasks (see QPlan): A fragments at
// TmQ[(C*NKQ*TT) + (c*NKT + kq)][lane], one shift tile per task.
// Store of a Q-table entry: with B2 (one channel slice), Q - B2/2 straight
// into the window's row of ovQ; else the slice's partial sum.
__device__ __forceinline__ void q_store(float *dst, const float *__restrict__ B2, uint32_t e, float q)
{
    dst[e] = B2 ? (float)((double)q - 0.5 * (double)B2[e]) : q;
}

template <int QC>
__device__ __forceinline__ void q_tail_task(const Dims &d, const QPlan &qp, const uint4 *__restrict__ TmQ,
                                            const uint16_t *svh, const uint16_t *svl, float *dst,
                                            const float *__restrict__ B2, const uint32_t *__restrict__ tord,
                                            uint32_t c0, uint32_t c1, uint32_t tg)
{
    const uint32_t S = d.S, T = d.T;
    const uint32_t lane = lane_of();
    const uint32_t rsel = QC == 2 ? (lane & 1u) : 0u, plane = d.C * qp.LPH;
    const uint16_t *sh = svh + rsel * plane, *sl = svl + rsel * plane;
    const uint32_t st0 = tg * qp.ND;
    const uint32_t nd = qp.ST - st0 < qp.ND ? qp.ST - st0 : qp.ND;   // offsets inside the shift range
    const uint32_t nk = qp.NKQ + nd - 1;
    const uint4 *ap = TmQ + (uint64_t)d.C * qp.NKQ * qp.TT * 64 + lane;
    const uint32_t ybase = 4 * (lane >> 4) + st0 * 16 + (lane & 15);
    // six independent accumulator chains (hi*hi, hi*lo, lo*hi for even and
    // odd K-steps) so consecutive matrix-core operations do not wait on each
    // other; every term still passes through fewer roundings than the chain
    // bound of the screen counts
    const float4v z = float4v{0.0f, 0.0f, 0.0f, 0.0f};
    float4v h0 = z, h1 = z, x0 = z, x1 = z, y0s = z, y1s = z;
    for (uint32_t c = c0; c < c1; c++) {
        uint4 an = ap[(uint64_t)c * qp.NKT * 64];
        for (uint32_t kq = 0; kq < nk; kq++) {
            const uint4 a = an;
            if (kq + 1 < nk) an = ap[((uint64_t)c * qp.NKT + kq + 1) * 64];
            const uint64_t ah = (uint64_t)a.x | ((uint64_t)a.y << 32);
            const uint64_t al = (uint64_t)a.z | ((uint64_t)a.w << 32);
            const uint32_t y0 = kq * 16 + ybase;
            const uint64_t bh = q_frag<QC>(sh, rsel, c, qp.LPH, y0), bl = q_frag<QC>(sl, rsel, c, qp.LPH, y0);
            if (kq & 1) {
                h1 = mfma_bf16_k16(ah, bh, h1);
                x1 = mfma_bf16_k16(ah, bl, x1);
                y1s = mfma_bf16_k16(al, bh, y1s);
            } else {
                h0 = mfma_bf16_k16(ah, bh, h0);
                x0 = mfma_bf16_k16(ah, bl, x0);
                y0s = mfma_bf16_k16(al, bh, y0s);
            }
        }
    }
    float4v acch, accl;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        acch[r] = h0[r] + h1[r];
        accl[r] = (x0[r] + y0s[r]) + (x1[r] + y1s[r]);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint32_t row = 4 * (lane >> 4) + r, g = row / qp.TR, i = row - g * qp.TR;
        const uint32_t ts = qp.TTF * 16 + i, s = (st0 + g) * 16 + (lane & 15);
        if (g < nd && ts < T && s < S) q_store(dst, B2, tord[ts] * S + s, acch[r] + accl[r]);
    }
}

// sQ: the [T][S] table of channel slice 0 (with B2 and a single slice, the
// window's row of ovQ, Q - B2/2 stored directly); sQp: the other slices'.
template <int QC>
__device__ void q_table_split(const Dims &d, const QPlan &qp, const uint4 *__restrict__ TmQ,
                              const uint16_t *svh, const uint16_t *svl, float *sQ, float *sQp,
                              const float *__restrict__ B2, const uint32_t *__restrict__ tord,
                              const uint32_t *__restrict__ tile_lo, const uint32_t *__restrict__ tile_hi,
                              uint32_t rot)
{
    const uint32_t S = d.S, T = d.T, TS = T * S;
    const uint32_t lane = lane_of();
    const uint32_t rsel = QC == 2 ? (lane & 1u) : 0u, plane = d.C * qp.LPH;
    const uint16_t *sh = svh + rsel * plane, *sl = svl + rsel * plane;
    // the wave index through readfirstlane: every task quantity below is then
    // known to be wave-uniform (scalar registers, scalar branches, and load
    // counters the compiler can track across the fragment prefetches)
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), nwave = blockDim.x >> 6;
    const uint32_t ntask = (qp.TTF * qp.NSG + qp.NTG) * qp.KS;
    for (uint32_t it = wave; it < ntask; it += nwave) {
        // with one task per wave, the waves take turns at the (lighter) tail
        // tasks from window to window, so every SIMD gets the same work
        const uint32_t task = ntask == nwave ? (it + rot) % nwave : it;
        const uint32_t ks = task % qp.KS, rest = task / qp.KS;
        if (rest >= qp.TTF * qp.NSG) {
            const uint32_t tl = qp.TT - 1;
            const uint32_t c0 = max(ks * d.C / qp.KS, tile_lo[tl]), c1 = max(c0, min((ks + 1) * d.C / qp.KS, tile_hi[tl]));
            q_tail_task<QC>(d, qp, TmQ, svh, svl, ks == 0 ? sQ : sQp + (uint64_t)(ks - 1) * TS, B2, tord, c0, c1,
                        rest - qp.TTF * qp.NSG);
            continue;
        }
        const uint32_t g = rest % qp.NSG, tt = rest / qp.NSG;
        // the slice's channels on which the tile's templates are nonzero
        const uint32_t c0 = max(ks * d.C / qp.KS, tile_lo[tt]), c1 = max(c0, min((ks + 1) * d.C / qp.KS, tile_hi[tt]));
        const uint32_t st0 = g * qp.SG;
        const uint32_t nst = qp.ST - st0 < qp.SG ? qp.ST - st0 : qp.SG;
        float4v acch[4], accl[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acch[j] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
            accl[j] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
        }
        // steps (c, kq) in c-major order; the A fragment of step st is
        // TmQ[(st*TT + tt)*64 + lane].  Fragments are fetched QD steps ahead
        // (two: with the eight B fragments of a step in flight, four would
        // spill registers).
        constexpr int QD = 2;
        const uint32_t sbeg = c0 * qp.NKQ, send = c1 * qp.NKQ;
        const uint4 *ap = TmQ + (uint64_t)tt * 64 + lane;
        const uint64_t astride = (uint64_t)qp.TT * 64;
        const uint32_t ybase = 4 * (lane >> 4) + st0 * 16 + (lane & 15);
        uint4 an[QD];
#pragma unroll
        for (int q = 0; q < QD; q++) an[q] = sbeg + q < send ? ap[(sbeg + q) * astride] : uint4{0u, 0u, 0u, 0u};
        uint32_t c = c0, kq = 0;
        for (uint32_t s0 = sbeg; s0 < send; s0 += QD) {
            uint4 a[QD];
#pragma unroll
            for (int q = 0; q < QD; q++) a[q] = an[q];
#pragma unroll
            for (int q = 0; q < QD; q++)
                an[q] = s0 + QD + q < send ? ap[(s0 + QD + q) * astride] : uint4{0u, 0u, 0u, 0u};
#pragma unroll
            for (int q = 0; q < QD; q++) {
                if (s0 + q >= send) break;
                const uint64_t ah = (uint64_t)a[q].x | ((uint64_t)a[q].y << 32);
                const uint64_t al = (uint64_t)a[q].z | ((uint64_t)a[q].w << 32);
                const uint32_t y0 = kq * 16 + ybase;
                if (nst == 4) {   // every fragment read first, then the twelve products
                    uint64_t bh[4], bl[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        bh[j] = q_frag<QC>(sh, rsel, c, qp.LPH, y0 + 16 * j);
                        bl[j] = q_frag<QC>(sl, rsel, c, qp.LPH, y0 + 16 * j);
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        acch[j] = mfma_bf16_k16(ah, bh[j], acch[j]);
                        accl[j] = mfma_bf16_k16(ah, bl[j], accl[j]);
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++) accl[j] = mfma_bf16_k16(al, bh[j], accl[j]);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        if ((uint32_t)j < nst) {
                            const uint64_t bh = q_frag<QC>(sh, rsel, c, qp.LPH, y0 + 16 * j),
                                           bl = q_frag<QC>(sl, rsel, c, qp.LPH, y0 + 16 * j);
                            acch[j] = mfma_bf16_k16(ah, bh, acch[j]);
                            accl[j] = mfma_bf16_k16(ah, bl, accl[j]);
                            accl[j] = mfma_bf16_k16(al, bh, accl[j]);
                        }
                    }
                }
                if (++kq == qp.NKQ) { kq = 0; c++; }
            }
        }
        float *dst = ks == 0 ? sQ : sQp + (uint64_t)(ks - 1) * TS;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if ((uint32_t)j >= nst) break;
            const uint32_t scol = (st0 + j) * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint32_t ts = tt * 16 + 4 * (lane >> 4) + r;
                if (ts < T && scol < S) q_store(dst, B2, tord[ts] * S + scol, acch[j][r] + accl[j][r]);
            }
        }
    }
}

// Q - B2/2 of every unskipped overlap window of a batch, on the matrix cores
// (q_table_split), into ovQ[k - kbase][T][S] for k_overlap_screened<true>.
// The window skip rule is the screen kernel's; skipped windows are not written.
template <int QC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_overlap_q(
    Dims d, const float *__restrict__ V, const float *__restrict__ L, const float *__restrict__ thr,
    const uint32_t *__restrict__ ov_list, const Counters *__restrict__ ctr, const uint32_t *__restrict__ best_lab,
    const uint32_t *__restrict__ best_idx, const float *__restrict__ B2, const uint4 *__restrict__ TmQ,
    float *__restrict__ ovQ, uint32_t kbase, const uint32_t *__restrict__ tord, const uint32_t *__restrict__ tile_lo,
    const uint32_t *__restrict__ tile_hi, double *__restrict__ ovVV)
{
    extern __shared__ float dyn[];
    __shared__ double r_q[kOvWaves];
    const uint32_t S = d.S, T = d.T, N = d.N, M = d.M, C = d.C;
    const uint32_t TS = T * S;
    const QPlan qp = q_plan(d);
    const uint32_t LPH = qp.LPH;
    const FastDiv divLPH(LPH);
    // QC copies of the zero-padded voltage, copy r moved left by r samples
    uint16_t *svh = (uint16_t *)dyn;                    // [QC][C][LPH] bf16 hi
    uint16_t *svl = svh + QC * C * LPH;                 // [QC][C][LPH] bf16 lo
    // several channel slices: their partial tables, summed below; one slice:
    // the table goes straight to ovQ (no LDS)
    float *sQp = (float *)(svl + QC * C * LPH);         // [KS-1][T][S] channel-slice partials
    float *sQ = sQp + (qp.KS - 1) * TS;                 // [T][S] (KS > 1)
    const uint32_t kend = min(ctr->n_ov, kbase + d.ov_batch);
    // The zero padding of the staged rows is written once; a window stages
    // only its C*N samples.  The next window's header is loaded in three
    // dependent steps during the current one (no block writes what another
    // block reads).
    for (uint32_t e = threadIdx.x; e < QC * C * LPH; e += blockDim.x) {
        svh[e] = 0;
        svl[e] = 0;
    }
    const uint32_t CN = C * N;
    const FastDiv divN(N);
    uint32_t nid = 0, nlab = 0, nbidx = 0;
    float nLb = 0.0f, nthl = 0.0f;
    int nstep = 0;
    uint32_t kn = kbase + blockIdx.x;
    auto head_step = [&](int upto) {
        if (kn >= kend) return;
        if (nstep < 1 && upto >= 1) { nid = ov_list[kn]; nstep = 1; }
        if (nstep < 2 && upto >= 2) { nlab = best_lab[nid]; nbidx = best_idx[nid]; nstep = 2; }
        if (nstep < 3 && upto >= 3) { nLb = L[(uint64_t)nlab * d.K + nbidx]; nthl = thr[nlab]; nstep = 3; }
    };
    auto put_v = [&](uint32_t e, float x) {
        const uint32_t c = divN.div(e), yr = M + (e - c * N), y = c * LPH + yr;
        const uint32_t h = to_bf16(x);
        const uint16_t l = to_bf16(x - __uint_as_float(h << 16));
        svh[y] = (uint16_t)h;
        svl[y] = l;
        if (QC == 2 && yr > 0) {   // row sample yr at yr - 1 of copy 1
            svh[C * LPH + y - 1] = (uint16_t)h;
            svl[C * LPH + y - 1] = l;
        }
    };
    head_step(3);
    for (uint32_t k = kbase + blockIdx.x; k < kend; k += gridDim.x) {
        const uint32_t bidx = nbidx;
        const bool skip = (nLb < nthl) || ((uint64_t)bidx >= (uint64_t)(d.K - d.N));   // bp.cl:518-527
        kn = k + gridDim.x;
        nstep = 0;
        head_step(1);
        if (skip) {
            head_step(3);
            continue;
        }
        __syncthreads();   // the previous window's LDS reads are done
        double vq = 0.0;   // the window's squared voltage norm, for the screen
        batched_stage<8>(CN, [&](uint32_t e) { const uint32_t c = divN.div(e);
                                               return V[(uint64_t)c * d.K + bidx + (e - c * N)]; },
                         [&](uint32_t e, float x) {
                             put_v(e, x);
                             vq = __builtin_fma((double)x, (double)x, vq);
                         });
        for (int off = 32; off > 0; off >>= 1) vq += __shfl_xor(vq, off);
        if (lane_of() == 0) r_q[threadIdx.x >> 6] = vq;
        head_step(2);
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (uint32_t w = 0; w < (blockDim.x >> 6); w++) t += r_q[w];
            ovVV[k - kbase] = t;
        }
        float *dst = ovQ + (uint64_t)(k - kbase) * TS;
        if (qp.KS == 1) {
            q_table_split<QC>(d, qp, TmQ, svh, svl, dst, sQp, B2, tord, tile_lo, tile_hi, k & 3u);
            head_step(3);
            continue;
        }
        q_table_split<QC>(d, qp, TmQ, svh, svl, sQ, sQp, nullptr, tord, tile_lo, tile_hi, k & 3u);
        __syncthreads();
        head_step(3);
        batched_stage<8>(TS, [&](uint32_t e) { return B2[e]; }, [&](uint32_t e, float b2) {
            float q = sQ[e];
            for (uint32_t j = 1; j < qp.KS; j++) q += sQp[(j - 1) * TS + e];
            dst[e] = (float)((double)q - 0.5 * (double)b2);
        });
    }
}

static size_t overlap_q_lds_qc(const Dims &d, uint32_t qc)
{
    const QPlan qp = q_plan(d);
    return (uint64_t)qc * d.C * qp.LPH * 4 + (qp.KS > 1 ? (uint64_t)qp.KS * d.T * d.S * 4 : 0);
}

// Voltage copies of k_overlap_q: two (aligned fragment reads) where the block
// still fits four times in a CU's LDS, else one.
uint32_t overlap_q_copies(const Dims &d) { return overlap_q_lds_qc(d, 2) <= 40 * 1024 ? 2u : 1u; }

size_t overlap_q_lds(const Dims &d) { return overlap_q_lds_qc(d, overlap_q_copies(d)); }

// TmQ[(c*NKQ + kq)*TT + tt][lane] = {hi, lo} fragments (4 bf16 each, sample
// kq*16 + 4*(lane/16) + i in bits 16i) of template t = tt*16 + lane%16 on
// channel c; 0 when padded.  The A fragments of q_table_split.
__global__ void k_pack_qtemplates(Dims d, const float *__restrict__ Tm, const uint32_t *__restrict__ tord,
                                  uint4 *__restrict__ TmQ)
{
    const uint32_t TT = (d.T + 15) / 16, NKQ = (d.N + 15) / 16;
    const uint64_t total = (uint64_t)d.C * NKQ * TT * 64;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t lane = (uint32_t)(e % 64);
        const uint64_t r = e / 64;
        const uint32_t tt = (uint32_t)(r % TT);
        const uint32_t kq = (uint32_t)((r / TT) % NKQ);
        const uint32_t c = (uint32_t)(r / ((uint64_t)TT * NKQ));
        const uint32_t ts = tt * 16 + (lane & 15);
        const uint32_t t = ts < d.T ? tord[ts] : d.T;   // row ts holds template tord[ts]
        uint32_t h[2] = {0u, 0u}, l[2] = {0u, 0u};
        for (uint32_t i = 0; i < 4; i++) {
            const uint32_t x = kq * 16 + 4 * (lane >> 4) + i;
            const float v = (t < d.T && x < d.N) ? Tm[((uint64_t)t * d.C + c) * d.N + x] : 0.0f;
            const uint32_t vh = to_bf16(v);
            const uint32_t vl = to_bf16(v - __uint_as_float(vh << 16));
            h[i >> 1] |= vh << (16 * (i & 1));
            l[i >> 1] |= vl << (16 * (i & 1));
        }
        TmQ[e] = uint4{h[0], h[1], l[0], l[1]};
    }
    // tail tile: TmQ[total + (c*NKT + kq)][lane], row g*TR + i = template
    // TTF*16 + i moved right by 16*g samples
    const QPlan qp = q_plan(d);
    const uint64_t tail = (uint64_t)d.C * qp.NKT * 64;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < tail;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t lane = (uint32_t)(e % 64);
        const uint64_t r = e / 64;
        const uint32_t kq = (uint32_t)(r % qp.NKT), c = (uint32_t)(r / qp.NKT);
        const uint32_t row = lane & 15, g = row / qp.TR, i = row - g * qp.TR;
        const uint32_t ts = qp.TTF * 16 + i;
        const uint32_t t = (g < qp.ND && ts < d.T) ? tord[ts] : d.T;
        uint32_t h[2] = {0u, 0u}, l[2] = {0u, 0u};
        for (uint32_t q = 0; q < 4; q++) {
            const int32_t x = (int32_t)(kq * 16 + 4 * (lane >> 4) + q) - (int32_t)(16 * g);
            const float v = (t < d.T && x >= 0 && x < (int32_t)d.N) ? Tm[((uint64_t)t * d.C + c) * d.N + x] : 0.0f;
            const uint32_t vh = to_bf16(v);
            const uint32_t vl = to_bf16(v - __uint_as_float(vh << 16));
            h[q >> 1] |= vh << (16 * (q & 1));
            l[q >> 1] |= vl << (16 * (q & 1));
        }
        TmQ[total + e] = uint4{h[0], h[1], l[0], l[1]};
    }
}

// Register budget of four waves per SIMD (measured faster than three despite a
// few spilled registers).  NW waves per block: 4 when the block also scores
// its candidates (main and partner templates staged in LDS); 2 or 8 when the
// exact scores are deferred to k_overlap_exact -- the lean block holds no
// template rows, so twice as many windows are in flight per CU at the same
// wave occupancy (the kernel is bound by the latency of its phases), and
// where the per-window tables leave room for one block per CU (hundreds of
// templates) eight waves share the window's phases.
constexpr uint64_t kXcMaxLds = 156 * 1024;

// LDS floats of the exact-score kernel's padded template rows: M zeros, then
// T*C rows of N samples and M zeros each.  Row entry x of template row r is
// at M + r*(N + M) + x; every x in [-M, N + M) reads the row or the zeros
// around it.
__host__ __device__ inline uint64_t overlap_exact_lds_floats(const Dims &d)
{
    return (uint64_t)d.M + (uint64_t)d.T * d.C * (d.N + d.M);
}

// The exact-score kernel reads the templates from global memory (they do not
// fit in its LDS, or the test switch asks for it).
__host__ __device__ inline bool overlap_exact_global(const Dims &d)
{
    return overlap_exact_lds_floats(d) * 4 > kXcMaxLds || d.xc_wide;
}

template <bool MQ, int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(4))) void k_overlap_screened(
    Dims d, const float *__restrict__ V, const float *__restrict__ Tm, const float *__restrict__ L,
    const float *__restrict__ thr, const uint32_t *__restrict__ ov_list, Counters *ctr_rw, float *best_L,
    const uint32_t *best_lab, const uint32_t *best_idx, uint32_t *ov_idx, uint32_t *ov_lab, uint8_t *ov,
    const float *__restrict__ AB, const float *__restrict__ ABmin, const float *__restrict__ B2,
    const float *__restrict__ tnorm, const float *__restrict__ Tpad, const float *__restrict__ ovQ, uint32_t kbase,
    const uint32_t *__restrict__ sup_lo, const uint32_t *__restrict__ sup_hi, uint32_t debug,
    uint32_t *__restrict__ xc_cand, uint32_t *__restrict__ xc_win, uint32_t *__restrict__ xc_ctr,
    const double *__restrict__ ovVV)
{
    extern __shared__ float dyn[];
    const uint32_t S = d.S, T = d.T, N = d.N, M = d.M, C = d.C;
    const uint32_t TS = T * S, R = N + 2 * M, SM = d.sup_max;
    const QPlan qp = q_plan(d);
    const FastDiv divN(N), divS(S), divT(T);
    double *sE = (double *)dyn;                         // [T] screen bound per template (this window)
    constexpr bool kLean = NW != 4;
    constexpr uint32_t kCand = 4 * NW * 64;             // candidate list: four per thread
    // kQG (the eight-wave lean block on matrix-core Q tables, for hundreds of
    // templates): the window's Q table is read from ovQ and its voltage from
    // V in global memory, so the LDS holds the per-template bookkeeping only
    // and two blocks fit in a CU
    constexpr bool kQG = MQ && NW == 8;
    // kVG (every lean block on matrix-core Q tables): the window's voltage is
    // read from V when a candidate is scored in the block, and its norm comes
    // from the Q-table kernel, so it is not staged
    constexpr bool kVG = MQ && kLean;
    float *sv = (float *)(sE + T);                      // [C][N] voltage at the best onset (not kVG)
    float *smain = sv + (kVG ? 0u : C * N);             // [SM][R] padded best-label template, its support rows
    float *szero = smain + (kLean ? 0u : SM * R);       // [R] zeros (not lean)
    float *sQ = szero + (kLean ? 0u : R);               // [T][S] Q - B2/2 (not kQG)
    float *stn = sQ + (kQG ? 0u : TS);                  // [T] template norms
    uint32_t *flist = (uint32_t *)(stn + T);            // main shifts: positive first, then the rest
    float *sb = (float *)(flist + S);                   // [SM][R] partner template of an exact score (not lean)
    // not lean: kept rows alias the partner template, read before the
    // candidates are written; lean: kept rows, then the candidates
    uint32_t *cand = kLean ? (uint32_t *)sb + (TS + 1) / 2 : (uint32_t *)(sb + SM * R);   // [kCand] (f << 16 | e)
    uint16_t *krow = (uint16_t *)sb;                    // [T*S] rows kept by their bounds
    uint16_t *plist = (uint16_t *)(kLean ? cand + kCand : (uint32_t *)sb + max(SM * R + kCand, (TS + 1) / 2));
                                                        // likelihood, grouped by t
    uint16_t *pofs = plist + TS, *pcnt = pofs + T;      // [T] start and length of each template's group
    uint16_t *tlist = pcnt + T;                         // [T] templates with a positive likelihood, ascending
    uint8_t *spos = (uint8_t *)(tlist + T);             // [T][S] bit1 in bounds, bit0 likelihood > 0
    float *sqmax = (float *)(((uintptr_t)(spos + TS) + 3) & ~(uintptr_t)3);   // [2][T] max over (positive) s of Q'
    float *sEf = sqmax + 2 * T;                         // [T] sE (1 + 2^-20) rounded up to float32
    float *svcs = sEf + T;                              // [C] squared voltage norm per channel (kVG)
    // Sparse channel supports (kVG): the bound of a template pair takes the
    // voltage norm and the operation counts over the channels of their two
    // supports only (every term on another channel is an exact zero, which
    // neither the reference's chain nor the matrix-core sums round).  Used
    // for shift ranges beyond N/2, where partners shifted nearly out of the
    // window leave many combinations within the bound (at N/2 the per-channel
    // norms cost the screen more than the fewer exact scores save).
    const bool sparse = kVG && 2 * SM <= C && 2 * M > N;
    __shared__ double r_ub[NW];
    if (!kLean)
        for (uint32_t x = threadIdx.x; x < R; x += blockDim.x) szero[x] = 0.0f;
    for (uint32_t x = threadIdx.x; x < T; x += blockDim.x) stn[x] = tnorm[x];
    __shared__ uint32_t n_cand, s_settled, s_defer;
    __shared__ uint32_t r_scan[NW];
    __shared__ float r_sc[NW];
    __shared__ uint64_t r_key[NW], r_p[NW];
    __shared__ double r_d[NW][2];
    __shared__ double r_vv[NW];
    __shared__ uint32_t n_ppos, n_fpos, n_fneg, n_kp, n_kn, s_maxcnt, n_tp;
    const Counters *ctr = ctr_rw;
    const uint32_t n_ov = ctr->n_ov;
    const uint32_t wg_bits = 31 - __builtin_clz(d.reduce_wg);
    const uint64_t span = d.nl * d.reduce_wg;
    // E coefficient: float32 chain bound plus float32 storage of Q and tables.
    const double nops = (double)C * N + 2.0 * C + 8.0;
    const double u2 = 0x1p-23;
    const double g = nops * u2 / (1.0 - nops * u2);
    // Split-bf16 matrix-core Q, relative to sum |T||v| <= |T_t| |v|: the hi*hi
    // chain (depth nq: C*NKQ*16 products, the hi + lo add and KS-1 slice
    // adds), the hi*lo + lo*hi chain (depth 2nq, terms <= 2^-7.98 of the
    // products) and the omitted lo*lo terms and lo representation error.
    const double nq = (double)C * qp.NKQ * 16 + qp.KS + 2.0;
    const double gq1 = nq * u2 / (1.0 - nq * u2), gq2 = 2.0 * nq * u2 / (1.0 - 2.0 * nq * u2);
    const double gq = MQ ? gq1 * 1.004 + gq2 * 0.0079 + 3.1 * 0x1p-16 : 0.0;
    // products flushed below the float32 normal range: 2^-126 each, both Q terms
    const double eabs = MQ ? 2.0 * 3.0 * nq * 0x1p-126 : 0.0;
    // + float32 storage of Q - B2/2 and of the AB table (unit roundoff each)
    const double ecoef = (g + 2.0 * gq + 4.0 * u2) * 1.01;
    const uint32_t kend = min(n_ov, kbase + d.ov_batch);
    // Window header (id, label, onset, best likelihood, label support) of the
    // block's next window, loaded in three dependent steps spread over the
    // current window so their latencies hide behind its work.  No block
    // writes another block's window entries.
    uint32_t id = 0, lab = 0, bidx = 0, lo_m = 0, hi_m = 0;
    float bl = 0.0f, Lb = 0.0f, thl = 0.0f;
    uint32_t nid = 0, nlab = 0, nbidx = 0, nlo = 0, nhi = 0;
    float nbl = 0.0f, nLb = 0.0f, nthl = 0.0f;
    int nstep = 0;
    uint32_t kn = kbase + blockIdx.x;   // window whose header is loading
    auto head_step = [&](int upto) {
        if (kn >= kend) return;
        if (nstep < 1 && upto >= 1) { nid = ov_list[kn]; nstep = 1; }
        if (nstep < 2 && upto >= 2) { nlab = best_lab[nid]; nbidx = best_idx[nid]; nbl = best_L[nid]; nstep = 2; }
        if (nstep < 3 && upto >= 3) {
            nLb = L[(uint64_t)nlab * d.K + nbidx];
            nthl = thr[nlab];
            nlo = sup_lo[nlab];
            nhi = sup_hi[nlab];
            nstep = 3;
        }
    };
    head_step(3);
    unsigned long long acc_screened = 0, acc_scored = 0;   // combinations screened / scored exactly
    for (uint32_t k = kbase + blockIdx.x; k < kend; k += gridDim.x) {
        id = nid; lab = nlab; bidx = nbidx; bl = nbl; Lb = nLb; thl = nthl; lo_m = nlo; hi_m = nhi;
        kn = k + gridDim.x;
        nstep = 0;
        head_step(1);
        const bool wskip = (Lb < thl) || ((uint64_t)bidx >= (uint64_t)(d.K - d.N));   // bp.cl:518-527
        float sc = -INFINITY;
        uint64_t key = ~0ull, bp = 0;
        // the window's Q table and voltage rows (stride vstr)
        const float *qT = kQG ? ovQ + (uint64_t)(k - kbase) * TS : sQ;
        const float *vw = kVG ? V + bidx : sv;
        const uint64_t vstr = kVG ? (uint64_t)d.K : (uint64_t)N;
        if (!wskip) {
            __syncthreads();
            if (threadIdx.x == 0) {
                n_ppos = 0;
                n_kp = 0;
                n_kn = 0;
                s_maxcnt = 0;
                n_cand = 0;
                s_settled = 0;
                s_defer = 0;
            }
            // staging (index quotients by FastDiv: every index is below the
            // LDS size); the likelihood signs at every in-bounds (t, s), where
            // the loaded value is -1 if (t, s) is out of bounds
            double vv = 0.0;
            auto a_v = stage_arr<4>(kVG ? 0u : C * N, [&](uint32_t e) { const uint32_t c = divN.div(e);
                                                              return V[(uint64_t)c * d.K + bidx + (e - c * N)]; },
                                    [&](uint32_t e, float x) {
                                        if (!kVG) sv[e] = x;
                                        vv = __builtin_fma((double)x, (double)x, vv);
                                    });
            auto a_main = stage_arr<8>(kLean ? 0u : (hi_m - lo_m) * R, [&](uint32_t e) { return Tpad[((uint64_t)lab * C + lo_m) * R + e]; },
                                       [&](uint32_t e, float x) { smain[e] = x; });
            auto a_q = stage_arr<10>(MQ && !kQG ? TS : 0, [&](uint32_t e) { return ovQ[(uint64_t)(k - kbase) * TS + e]; },
                                     [&](uint32_t e, float q) { sQ[e] = q; });
            auto a_l = stage_arr<10>(TS, [&](uint32_t e) {
                const uint32_t t = divS.div(e), s = e - t * S;
                const bool inb = !((uint64_t)s + bidx < M || (uint64_t)s + bidx >= (uint64_t)d.K - N + M);   // bp.cl:534,538
                return inb ? L[(uint64_t)t * d.K + bidx + s - M] : -1.0f;
            }, [&](uint32_t e, float lv) {
                const uint32_t t = divS.div(e), s = e - t * S;
                const bool inb = !((uint64_t)s + bidx < M || (uint64_t)s + bidx >= (uint64_t)d.K - N + M);
                const bool pos = inb && lv > 0.0f;
                spos[e] = (inb ? 2 : 0) | (pos ? 1 : 0);
                if (!MQ) {
                    const int32_t sa = (int32_t)M - (int32_t)s;
                    const int32_t lo = max(0, -sa), hi = min((int32_t)N, (int32_t)N - sa);
                    double q = 0.0;
                    for (uint32_t c = 0; c < C; c++) {
                        const float *b = Tm + ((uint64_t)t * C + c) * N + sa;
                        const float *vc = sv + c * N;
                        for (int32_t j = lo; j < hi; j++) q = __builtin_fma((double)b[j], (double)vc[j], q);
                    }
                    sQ[e] = (float)q;
                }
            });
            if (MQ) {
                batched_stage_all(a_v, a_main, a_q, a_l);
            } else {   // the float64 Q table reads the staged voltage
                batched_stage_all(a_v, a_main);
                __syncthreads();
                batched_stage_all(a_l);
            }
            for (int off = 32; off > 0; off >>= 1) vv += __shfl_xor(vv, off);
            if (lane_of() == 0) r_vv[threadIdx.x >> 6] = vv;
            if (sparse) {   // a wave per channel row, four rows in flight; float64 squares (exact), rounded up
                const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), l = lane_of();
                for (uint32_t c0 = wv * 4; c0 < C; c0 += NW * 4) {
                    double a[4] = {0.0, 0.0, 0.0, 0.0};
                    for (uint32_t j = l; j < N; j += kWave) {
                        float x[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) x[q] = c0 + q < C ? V[(uint64_t)(c0 + q) * d.K + bidx + j] : 0.0f;
#pragma unroll
                        for (int q = 0; q < 4; q++) a[q] = __builtin_fma((double)x[q], (double)x[q], a[q]);
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        for (int off = 32; off > 0; off >>= 1) a[q] += __shfl_xor(a[q], off);
                        const double sc = a[q] * (1.0 + 0x1p-40);
                        float up = (float)sc;
                        if ((double)up < sc) up = __uint_as_float(__float_as_uint(up) + 1u);
                        if (l == 0 && c0 + q < C) svcs[c0 + q] = up;
                    }
                }
            }
            __syncthreads();
            head_step(2);
            double vsum = 0.0;   // kVG: from the Q-table kernel, which staged the same samples
            if (kVG) vsum = ovVV[k - kbase];
            else
                for (int w = 0; w < NW; w++) vsum += r_vv[w];
            const double vn = sqrt(vsum) * (1.0 + 1e-12);
            const double norm_lab = stn[lab];
            // fold the voltage-free half norms into Q (done by the matrix-core
            // kernel for MQ), and the bound per template
            if (!MQ)
                batched_stage<8>(TS, [&](uint32_t e) { return B2[e]; },
                                 [&](uint32_t e, float b2) { sQ[e] = (float)((double)sQ[e] - 0.5 * (double)b2); });
            for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
                const double nsum = norm_lab + (double)stn[t];
                double e;
                if (sparse) {   // over the union of the two supports
                    const uint32_t lt = sup_lo[t], ht = sup_hi[t];
                    double vs = 0.0;
                    uint32_t u = 0;
                    for (uint32_t c = min(lo_m, lt); c < max(hi_m, ht); c++)
                        if ((c >= lo_m && c < hi_m) || (c >= lt && c < ht)) {
                            vs += (double)svcs[c];
                            u++;
                        }
                    const uint32_t um = max(hi_m - lo_m, ht - lt);
                    const double vnu = sqrt(vs) * (1.0 + 1e-12);
                    const double no = (double)u * N + 2.0 * u + 8.0;
                    const double gu = no * u2 / (1.0 - no * u2);
                    const double nqu = (double)um * qp.NKQ * 16 + qp.KS + 2.0;
                    const double q1 = nqu * u2 / (1.0 - nqu * u2), q2 = 2.0 * nqu * u2 / (1.0 - 2.0 * nqu * u2);
                    const double gqu = MQ ? q1 * 1.004 + q2 * 0.0079 + 3.1 * 0x1p-16 : 0.0;
                    const double eau = MQ ? 2.0 * 3.0 * nqu * 0x1p-126 : 0.0;
                    e = (gu + 2.0 * gqu + 4.0 * u2) * 1.01 * (nsum * vnu + nsum * nsum) + eau;
                } else {
                    e = ecoef * (nsum * vn + nsum * nsum) + eabs;
                }
                sE[t] = e;
                const double eu = e * (1.0 + 0x1p-20) + 0x1p-120;
                float ef = (float)eu;
                if ((double)ef < eu) ef = __uint_as_float(__float_as_uint(ef) + 1u);
                sEf[t] = ef;
            }
            if (!MQ) __syncthreads();   // Q - B2/2 complete
            // Sixteen lanes per template t (four templates per wave), by
            // ballot over s within the group: the group of positive-likelihood
            // (t, s) in plist, ascending s; max over s of Q'(t, s) and over the
            // positive s; for t = lab also the main shifts: the positive ones,
            // then the other in-bounds ones, each ascending.
            {
                const uint32_t l = lane_of(), gl = l & 15u;
                const uint64_t gmask = 0xffffull << (l & 48u);
                const uint64_t below = ((1ull << l) - 1) & gmask;
                for (uint32_t t0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * 4; t0 < T;
                     t0 += NW * 4) {
                    const uint32_t t = t0 + (l >> 4);
                    const bool act = t < T;
                    float m = -INFINITY, mp = -INFINITY;
                    uint32_t cnt = 0, cneg = 0;
                    for (uint32_t s0 = 0; s0 < S; s0 += 16) {
                        const uint32_t s = s0 + gl;
                        const bool in = act && s < S;
                        const uint8_t sp = in ? spos[t * S + s] : (uint8_t)0;
                        const float q = in ? qT[t * S + s] : -INFINITY;
                        m = fmaxf(m, q);
                        if (sp & 1) mp = fmaxf(mp, q);
                        cnt += __popcll(__ballot(sp & 1) & gmask);
                        cneg += __popcll(__ballot((sp & 3) == 2) & gmask);
                    }
                    for (int off = 8; off > 0; off >>= 1) {   // within the group of sixteen lanes
                        m = fmaxf(m, __shfl_xor(m, off));
                        mp = fmaxf(mp, __shfl_xor(mp, off));
                    }
                    uint32_t base = 0;
                    if (gl == 0 && act) {
                        base = atomicAdd(&n_ppos, cnt);
                        atomicMax(&s_maxcnt, cnt);
                        pofs[t] = (uint16_t)base;
                        pcnt[t] = (uint16_t)cnt;
                        sqmax[t] = m;
                        sqmax[T + t] = mp;
                        if (t == lab) {
                            n_fpos = cnt;
                            n_fneg = cneg;
                        }
                    }
                    base = __shfl(base, (int)(l & 48u));
                    uint32_t op = 0, on = 0;
                    for (uint32_t s0 = 0; s0 < S; s0 += 16) {
                        const uint32_t s = s0 + gl;
                        const bool in = act && s < S;
                        const uint8_t sp = in ? spos[t * S + s] : (uint8_t)0;
                        const uint64_t mpos = __ballot(sp & 1) & gmask, mneg = __ballot((sp & 3) == 2) & gmask;
                        if (sp & 1) plist[base + op + __popcll(mpos & below)] = (uint16_t)(t * S + s);
                        if (t == lab) {
                            if (sp & 1) flist[op + __popcll(mpos & below)] = s;
                            if ((sp & 3) == 2) flist[cnt + on + __popcll(mneg & below)] = s;
                        }
                        op += __popcll(mpos);
                        on += __popcll(mneg);
                    }
                }
            }
            __syncthreads();
            if (threadIdx.x < kWave) {   // templates with a positive likelihood, ascending (one wave, by ballot)
                const uint32_t l = threadIdx.x;
                uint32_t nt = 0;
                for (uint32_t t0 = 0; t0 < T; t0 += kWave) {
                    const bool has = t0 + l < T && pcnt[t0 + l] > 0;
                    const uint64_t m = __ballot(has);
                    if (has) tlist[nt + __popcll(m & ((1ull << l) - 1))] = (uint16_t)(t0 + l);
                    nt += __popcll(m);
                }
                if (l == 0) n_tp = nt;
            }
            __syncthreads();
            const float *ABl = AB + (uint64_t)lab * S * TS;
            const float *sQl = qT + lab * S;
            const uint32_t nfp = n_fpos, nfn = n_fneg, maxcnt = s_maxcnt;
            auto screen = [&](uint32_t f, float qe, float ab) {
                return (double)sQl[f] + (double)qe - (double)ab;
            };
            // Row bounds.  For the main shift f of flist row a and template t,
            // every combination (f, t, s) of the row has sh + E <= U(a, t) =
            // Q'(lab, f) + max_s Q'(t, s) - min_s AB(lab, f, t, s) + E_t
            // (the max over s runs over the positive-likelihood s for rows of
            // non-positive main shifts), evaluated in float32 with the margin
            // of the traversal below (so it bounds the float32 upper bounds
            // there too).  Each
            // wave evaluates the row with its largest U; the largest lower bound
            // sh - E found there, lo_est, is at most lo_max.  Rows with U below
            // lo_est hold no combination that reaches lo_max, so neither the
            // bounds nor the candidates change when the traversal skips them.
            // Rows: every (positive main shift, template), then every (other
            // main shift, template with a positive likelihood); rows of
            // templates without one hold no combination.
            const uint32_t ntp = n_tp, nfpr = nfp * T;
            const uint32_t nrows = nfpr + nfn * ntp;
            const FastDiv divTP(ntp ? ntp : 1u);
            const float *mABl = ABmin + (uint64_t)lab * S * T;   // [f][t]
            constexpr int RB = 4;   // rows per thread whose loads are in flight together
            auto row_bounds = [&](auto use) {
                for (uint32_t r0 = threadIdx.x; r0 < nrows; r0 += RB * blockDim.x) {
                    float z[RB];
                    uint32_t aq[RB], tq[RB];
#pragma unroll
                    for (int q = 0; q < RB; q++) {
                        const uint32_t x = r0 + q * blockDim.x;
                        const bool in = x < nrows;
                        uint32_t a, t;
                        if (x < nfpr) {
                            a = divT.div(x);
                            t = x - a * T;
                        } else {
                            const uint32_t y = x - nfpr, b = divTP.div(y);
                            a = nfp + b;
                            t = in ? tlist[y - b * ntp] : 0u;
                        }
                        aq[q] = in ? a : 0u;
                        tq[q] = t;
                        z[q] = in ? mABl[flist[a] * T + t] : 0.0f;
                    }
#pragma unroll
                    for (int q = 0; q < RB; q++) {
                        if (r0 + q * blockDim.x < nrows) {
                            const uint32_t f = flist[aq[q]];
                            const float x = sQl[f], y = sqmax[(aq[q] < nfp ? 0 : T) + tq[q]], zz = z[q];
                            const float m = sEf[tq[q]] + 0x1p-21f * ((fabsf(x) + fabsf(y)) + fabsf(zz));
                            const float u = y == -INFINITY ? -INFINITY   // no combination in the row
                                                           : ((x + y) - zz) + m;
                            use(aq[q] * T + tq[q], aq[q], tq[q], u, q);
                        }
                    }
                }
            };
            // With at most RB rows per thread, the second use of the bounds
            // reads them from registers.
            const bool one_batch = nrows <= RB * blockDim.x;
            float uc[RB];
            uint32_t rc[RB];
            float ub = -INFINITY;
            uint32_t ur = 0;
            row_bounds([&](uint32_t r, uint32_t, uint32_t, float u, int q) {
                if (u > ub) { ub = u; ur = r; }
                uc[q] = u;
                rc[q] = r;
            });
            for (int off = 32; off > 0; off >>= 1) {
                const float ou = __shfl_xor(ub, off);
                const uint32_t orr = __shfl_xor(ur, off);
                if (ou > ub || (ou == ub && orr < ur)) { ub = ou; ur = orr; }
            }
            float lo_est = -INFINITY;
            if (ub > -INFINITY) {   // this wave's best row: float32 lower bounds as in the traversal
                const uint32_t a = ur / T, t = ur - a * T, f = flist[a];
                for (uint32_t s = lane_of(); s < S; s += kWave) {
                    const uint8_t sp = spos[t * S + s];
                    const bool allowed = (a < nfp ? (sp & 2) : (sp & 1)) && (f > s ? f - s : s - f) <= N;
                    if (allowed) {
                        const float x = sQl[f], qe = qT[t * S + s], ab = ABl[(uint64_t)f * TS + t * S + s];
                        const float m = sEf[t] + 0x1p-21f * ((fabsf(x) + fabsf(qe)) + fabsf(ab));
                        lo_est = fmaxf(lo_est, ((x + qe) - ab) - m);
                    }
                }
            }
            for (int off = 32; off > 0; off >>= 1) lo_est = fmaxf(lo_est, __shfl_xor(lo_est, off));
            if (lane_of() == 0) r_ub[threadIdx.x >> 6] = lo_est;
            __syncthreads();
            lo_est = (float)r_ub[0];
            for (int w = 1; w < NW; w++) lo_est = fmaxf(lo_est, (float)r_ub[w]);
            if (debug & 16u) lo_est = -INFINITY;
            // kept rows: positive main shifts from the front, the others from the back
            auto keep_row = [&](uint32_t r, uint32_t a, uint32_t t, float u, int) {
                if (u < lo_est) return;
                if (a < nfp) krow[atomicAdd(&n_kp, 1u)] = (uint16_t)r;
                else if (pcnt[t]) krow[TS - 1 - atomicAdd(&n_kn, 1u)] = (uint16_t)r;
            };
            if (one_batch) {
#pragma unroll
                for (int q = 0; q < RB; q++)
                    if (threadIdx.x + q * blockDim.x < nrows) {
                        const uint32_t a = divT.div(rc[q]);
                        keep_row(rc[q], a, rc[q] - a * T, uc[q], q);
                    }
            } else {
                row_bounds(keep_row);
            }
            __syncthreads();
            const uint32_t nkp = n_kp, nkn = n_kn;
            // pass A: largest lower and upper bounds of the score.  Each thread
            // also keeps its four combinations with the largest upper bounds
            // (kv0 >= kv1 >= kv2 >= kv3, packed f << 16 | e in kc*) and the
            // largest upper bound it did not keep (kdrop).
            // The bounds are evaluated in float32: sh = (x + q) - ab is within
            // 2.001u·A of its exact value (u = 2^-24, A = |x| + |q| + |ab|),
            // and the margin m = Ef + 2^-21·A (Ef >= E(1 + 2^-20), rounded up)
            // covers that, E, and the rounding of A, m and sh ± m, so
            // [sh - m, sh + m] contains [sh_exact - E, sh_exact + E].
            float lo_mf = -INFINITY, hi_mf = -INFINITY, kdrop = -INFINITY;
            float kv0 = -INFINITY, kv1 = -INFINITY, kv2 = -INFINITY, kv3 = -INFINITY;
            uint32_t kc0 = 0, kc1 = 0, kc2 = 0, kc3 = 0;
            uint32_t screened = 0;
            for_each_kept<4>(d, divT, spos, flist, krow, nkp, nkn, plist, pofs, pcnt, maxcnt, ABl, qT, sEf, [&](uint32_t f, uint32_t t, uint32_t s, uint32_t e, float ab, float qe, float Ef) {
                const float x = sQl[f];
                const float sh = (x + qe) - ab;
                const float m = Ef + 0x1p-21f * ((fabsf(x) + fabsf(qe)) + fabsf(ab));
                const float hi = sh + m;
                lo_mf = fmaxf(lo_mf, sh - m);
                hi_mf = fmaxf(hi_mf, hi);
                screened++;
                if (hi > kv3) {
                    kdrop = fmaxf(kdrop, kv3);
                    kv3 = hi;
                    kc3 = (f << 16) | e;
                    if (kv3 > kv2) { const float x2 = kv2; kv2 = kv3; kv3 = x2; const uint32_t y = kc2; kc2 = kc3; kc3 = y; }
                    if (kv2 > kv1) { const float x2 = kv1; kv1 = kv2; kv2 = x2; const uint32_t y = kc1; kc1 = kc2; kc2 = y; }
                    if (kv1 > kv0) { const float x2 = kv0; kv0 = kv1; kv1 = x2; const uint32_t y = kc0; kc0 = kc1; kc1 = y; }
                } else {
                    kdrop = fmaxf(kdrop, hi);
                }
            });
            for (int off = 32; off > 0; off >>= 1) {
                lo_mf = fmaxf(lo_mf, __shfl_xor(lo_mf, off));
                hi_mf = fmaxf(hi_mf, __shfl_xor(hi_mf, off));
                screened += __shfl_xor(screened, off);
            }
            double lo_max, hi_max;
            if (lane_of() == 0) { r_d[threadIdx.x >> 6][0] = lo_mf; r_d[threadIdx.x >> 6][1] = hi_mf; }
            if (lane_of() == 0) acc_screened += screened;
            __syncthreads();
            lo_max = r_d[0][0];
            hi_max = r_d[0][1];
            for (int w = 1; w < NW; w++) {
                lo_max = fmax(lo_max, r_d[w][0]);
                hi_max = fmax(hi_max, r_d[w][1]);
            }
            head_step(3);
            // Reserve `total` slots of the batch's deferred candidates for this
            // window (thread 0; s_defer = first slot + 1, or 0 when they do
            // not fit and the window scores its candidates here).
            auto reserve = [&](uint32_t total) {
                if (threadIdx.x == 0) {
                    const uint32_t off = atomicAdd(&xc_ctr[1], total);
                    if ((uint64_t)off + total <= d.xc_cap) {
                        const uint32_t w = atomicAdd(&xc_ctr[0], 1u);
                        xc_win[4 * w] = k;
                        xc_win[4 * w + 1] = off;
                        xc_win[4 * w + 2] = total;
                        xc_win[4 * w + 3] = 0;
                        s_defer = off + 1;
                    }
                }
                __syncthreads();
                return s_defer != 0;
            };
            // Nothing can beat best_L: parse rejects without any exact score.
            if (hi_max > (double)bl) {
                uint32_t scored = 0;
                // main template rows: staged in LDS (not lean) or read from global memory
                const float *tmain = kLean ? Tpad + ((uint64_t)lab * C + lo_m) * R : smain;
                auto exact = [&](uint32_t f, uint32_t t, uint32_t s, const float *bt) {
                    const float score =
                        kLean ? overlap_support_score(d, vw, vstr, tmain, lo_m, hi_m, bt, sup_lo[t], sup_hi[t], f, s)
                              : overlap_support_score_z(d, sv, tmain, lo_m, hi_m, bt, sup_lo[t], sup_hi[t], szero, f, s);
                    scored++;
                    uint64_t p;
                    const uint64_t kk = combo_key(d, wg_bits, span, t, f, s, p);
                    if (key_better(score, kk, sc, key)) { sc = score; key = kk; bp = p; }
                };
                // Exact scores of every combination with sh + E >= lo_max
                // (bp.cl:599-640).  They are all among the kept ones unless a
                // thread dropped one that reaches lo_max; then the screen is
                // traversed again.
                if (__syncthreads_or(kdrop >= lo_max || (debug & 2u)) == 0) {
                    auto push = [&](double v, uint32_t c) {
                        if (v >= lo_max) cand[atomicAdd(&n_cand, 1u)] = c;
                    };
                    push(kv0, kc0);
                    push(kv1, kc1);
                    push(kv2, kc2);
                    push(kv3, kc3);
                    __syncthreads();
                    const uint32_t nc = n_cand;   // <= 4 * blockDim.x = kCand
                    // A single candidate is the winner (its lower bound is lo_max,
                    // above every other upper bound); its float32 score is needed
                    // only when the interval [sh - E, sh + E] leaves a comparison
                    // of the parse open.
                    if (nc == 1 && !(debug & 8u)) {
                        if (threadIdx.x == 0) {
                            const uint32_t f = cand[0] >> 16, e = cand[0] & 0xffffu;
                            const uint32_t t = e / S, s = e - t * S;
                            const double sh = screen(f, qT[e], ABl[(uint64_t)f * TS + e]);
                            const double E = sE[t];
                            uint64_t p;
                            (void)combo_key(d, wg_bits, span, t, f, s, p);
                            s_settled = parse_overlap_bounds(d, L, thr, id, lab, bidx, bl, sh - E, sh + E, p, best_L,
                                                             ov_idx, ov_lab, ov) ? 1u : 0u;
                        }
                        __syncthreads();
                    }
                    if (s_settled) {
                        // parsed from the interval
                    } else if (d.xc_cap && reserve(nc)) {
                        // exact scores in k_overlap_exact, which also parses
                        for (uint32_t k2 = threadIdx.x; k2 < nc; k2 += blockDim.x)
                            xc_cand[s_defer - 1 + k2] = cand[k2];
                    } else if (!kLean && nc <= 16 && !(debug & 4u)) {
                        // one at a time, the partner template staged in LDS
                        for (uint32_t k2 = 0; k2 < nc; k2++) {
                            const uint32_t f = cand[k2] >> 16, e = cand[k2] & 0xffffu;
                            const uint32_t t = e / S, s = e - t * S;
                            const float *src = Tpad + ((uint64_t)t * C + sup_lo[t]) * R;
                            __syncthreads();
                            batched_stage<8>((sup_hi[t] - sup_lo[t]) * R, [&](uint32_t x) { return src[x]; },
                                             [&](uint32_t x, float v) { sb[x] = v; });
                            __syncthreads();
                            if (threadIdx.x == 0) exact(f, t, s, sb);
                        }
                    } else {
                        for (uint32_t k2 = threadIdx.x; k2 < nc; k2 += blockDim.x) {
                            const uint32_t f = cand[k2] >> 16, e = cand[k2] & 0xffffu;
                            const uint32_t t = e / S, s = e - t * S;
                            exact(f, t, s, Tpad + ((uint64_t)t * C + sup_lo[t]) * R);
                        }
                    }
                } else {
                    // Some thread dropped a combination that reaches lo_max:
                    // count every thread's qualifying combinations and place
                    // them, in each thread's traversal order, after a block-wide
                    // prefix sum of the counts.
                    bool deferred = false;
                    if (d.xc_cap) {
                        // Candidates grouped by main shift f (a counting sort
                        // over the S shifts in the candidate buffer, which this
                        // path does not use otherwise), so that the exact-score
                        // kernel's batches mostly share f; with more shifts than
                        // the buffer holds, each thread's candidates follow a
                        // block-wide prefix sum of the threads' counts.
                        // Templates in global memory: grouped by partner template
                        // instead, so a batch's lanes read few partner rows (each
                        // lane walks its own channels there).
                        const bool by_t = overlap_exact_global(d);
                        const uint32_t KD = by_t ? T : S;   // group keys
                        const bool by_f = KD <= kCand;
                        uint32_t *fpos = cand;   // [KD] counts, then next slot per key
                        if (by_f) {
                            for (uint32_t x = threadIdx.x; x < KD; x += blockDim.x) fpos[x] = 0;
                            __syncthreads();
                        }
                        uint32_t mine = 0;
                        for_each_kept<4>(d, divT, spos, flist, krow, nkp, nkn, plist, pofs, pcnt, maxcnt, ABl, qT, sE, [&](uint32_t f, uint32_t t, uint32_t, uint32_t, float ab, float qe, double E) {
                            if (screen(f, qe, ab) + E >= lo_max) {
                                if (by_f) atomicAdd(&fpos[by_t ? t : f], 1u);
                                else mine++;
                            }
                        });
                        const uint32_t l = lane_of();
                        uint32_t before = 0, total = 0, incl = mine;
                        if (by_f) {
                            __syncthreads();
                            if (threadIdx.x < kWave) {   // exclusive prefix over the keys, one wave
                                uint32_t carry = 0;
                                for (uint32_t x0 = 0; x0 < KD; x0 += kWave) {
                                    const uint32_t v = x0 + l < KD ? fpos[x0 + l] : 0u;
                                    uint32_t inc = v;
                                    for (int off = 1; off < 64; off <<= 1) {
                                        const uint32_t o = __shfl_up(inc, off);
                                        if (l >= (uint32_t)off) inc += o;
                                    }
                                    if (x0 + l < KD) fpos[x0 + l] = carry + inc - v;
                                    carry += __shfl(inc, 63);
                                }
                                if (l == 0) r_scan[0] = carry;
                            }
                            __syncthreads();
                            total = r_scan[0];
                        } else {
                            for (int off = 1; off < 64; off <<= 1) {
                                const uint32_t o = __shfl_up(incl, off);
                                if (l >= (uint32_t)off) incl += o;
                            }
                            if (l == 63) r_scan[threadIdx.x >> 6] = incl;
                            __syncthreads();
                            for (uint32_t w = 0; w < (uint32_t)NW; w++) {
                                if (w < (threadIdx.x >> 6)) before += r_scan[w];
                                total += r_scan[w];
                            }
                        }
                        if (reserve(total)) {
                            const uint32_t first = s_defer - 1;
                            uint32_t slot = first + before + incl - mine;
                            for_each_kept<4>(d, divT, spos, flist, krow, nkp, nkn, plist, pofs, pcnt, maxcnt, ABl, qT, sE, [&](uint32_t f, uint32_t t, uint32_t, uint32_t e, float ab, float qe, double E) {
                                if (screen(f, qe, ab) + E >= lo_max)
                                    xc_cand[by_f ? first + atomicAdd(&fpos[by_t ? t : f], 1u) : slot++] = (f << 16) | e;
                            });
                            deferred = true;
                        }
                    }
                    if (!deferred)
                        for_each_kept<1>(d, divT, spos, flist, krow, nkp, nkn, plist, pofs, pcnt, maxcnt, ABl, qT, sE, [&](uint32_t f, uint32_t t, uint32_t s, uint32_t e, float ab, float qe, double E) {
                            const double sh = screen(f, qe, ab);
                            if (sh + E < lo_max) return;
                            exact(f, t, s, Tpad + ((uint64_t)t * C + sup_lo[t]) * R);
                        });
                }
                for (int off = 32; off > 0; off >>= 1) scored += __shfl_xor(scored, off);
                if (lane_of() == 0) acc_scored += scored;
                if (!s_settled && !s_defer) block_key_reduce(sc, key, bp, r_sc, r_key, r_p);
            }
        }
        if (threadIdx.x == 0 && (wskip || !s_settled) && !(wskip ? 0u : s_defer))
            parse_overlap(d, L, thr, id, lab, bidx, bl, sc, bp, best_L, ov_idx, ov_lab, ov);
        head_step(3);
    }
    // statistics: one update per wave for all of its windows
    if (lane_of() == 0 && acc_screened) atomicAdd(&ctr_rw->screened, acc_screened);
    if (lane_of() == 0 && acc_scored) atomicAdd(&ctr_rw->combos, acc_scored);
}

// Deferred exact scores of the screened overlap (bp.cl:599-640 for the
// candidates the screen could not rule out, then the parse of bp.cl:690-814).
// One wave per deferred window, its lanes one candidate each, 64 at a time.
// The templates sit in LDS as rows of N samples, each followed by M zeros
// (staged once per block) when that fits in kXcMaxLds, else they are read
// from their zero-padded copy in global memory.
// Each lane's chain is the reference's float32 order; the winner is reduced
// in the reference's total order (score, then reduction group and lane).
constexpr uint32_t kXcThreads = 1024;   // sixteen waves per block

// A lane's float32 chains of NC candidates over the channels [cw0, cw1) with
// the templates in padded LDS rows.  Window sample j of a template shifted by
// x is row entry M - x + j, a zero outside the template, so no sample is
// masked; a channel where both templates are zero adds +0.0 sums, which leave
// ss at 0 and acc, never -0, unchanged.  The voltage index is the same in
// every lane, so its samples are scalar loads, shared by the NC chains; UF:
// the lanes share the main shift f, and the main template's samples are
// scalar loads too, from its zero-padded copy in global memory (Tg: rows of
// N + 2M, window sample j at 2M - f + j).
template <bool UF, int NC>
__device__ __forceinline__ void exact_score_pad(const Dims &d, const float *tp, const float *__restrict__ V,
                                                const float *__restrict__ Tg, uint32_t bidx, uint32_t lab,
                                                const uint32_t (&t)[NC], const uint32_t (&f)[NC],
                                                const uint32_t (&s)[NC], uint32_t cw0, uint32_t cw1,
                                                float (&acc)[NC])
{
    const uint32_t N = d.N, C = d.C, P = d.N + d.M, R = d.N + 2 * d.M;
    const uint32_t n8 = N & ~7u;
#pragma unroll
    for (int k = 0; k < NC; k++) acc[k] = 0.0f;
    for (uint32_t c = cw0; c < cw1; c++) {
        // row r's entry x is at M + r*P + x
        const float *ta[NC], *tb[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) {
            ta[k] = UF ? Tg + ((uint64_t)lab * C + c) * R + 2 * d.M - f[k]
                       : tp + ((uint64_t)lab * C + c) * P + 2 * d.M - f[k];
            tb[k] = tp + ((uint64_t)t[k] * C + c) * P + 2 * d.M - s[k];
        }
        const float *vc = V + (uint64_t)c * d.K + bidx;
        float ss[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) ss[k] = 0.0f;
        for (uint32_t j = 0; j < n8; j += 8) {
            float vv[8], av[NC][8], bv[NC][8];
#pragma unroll
            for (uint32_t q = 0; q < 8; q++) {
                vv[q] = vc[j + q];
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    av[k][q] = (UF && k > 0) ? av[0][q] : ta[k][j + q];
                    bv[k][q] = tb[k][j + q];
                }
            }
#pragma unroll
            for (uint32_t q = 0; q < 8; q++) {
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    const float sum = av[k][q] + bv[k][q];
                    ss[k] = __builtin_fmaf(sum, sum, ss[k]);        // bp.cl:621
                    acc[k] = __builtin_fmaf(sum, vv[q], acc[k]);    // bp.cl:622
                }
            }
        }
        for (uint32_t j = n8; j < N; j++) {
            const float v = vc[j];
#pragma unroll
            for (int k = 0; k < NC; k++) {
                const float sum = ta[k][j] + tb[k][j];
                ss[k] = __builtin_fmaf(sum, sum, ss[k]);
                acc[k] = __builtin_fmaf(sum, v, acc[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < NC; k++)
            if (ss[k] > 0.0f) acc[k] = (float)((double)acc[k] - 0.5 * (double)ss[k]);   // bp.cl:627-630
    }
}

// The same score with the templates read from their zero-padded rows in
// global memory (templates too large for LDS; rows of R = N + 2M samples,
// window sample j of a template at shift x is row entry 2M - x + j, +0.0
// outside the template).  Each lane walks its own channels: those of the
// main template's support [lo_m, hi_m) and its partner's [lo_t, hi_t), in
// ascending order (a channel where both are zero leaves the chain unchanged),
// with its own voltage, main and partner rows; an inactive lane walks none.
__device__ __forceinline__ float exact_score_glb(const Dims &d, const float *__restrict__ Tpad,
                                                 const float *__restrict__ V, uint32_t bidx, uint32_t lab,
                                                 uint32_t lo_m, uint32_t hi_m, uint32_t t, uint32_t lo_t,
                                                 uint32_t hi_t, uint32_t f, uint32_t s, bool act)
{
    const uint32_t N = d.N, C = d.C, R = N + 2 * d.M, n8 = N & ~7u;
    uint32_t c = min(lo_m, lo_t);
    const uint32_t cend = act ? max(hi_m, hi_t) : 0u;
    float acc = 0.0f;
    while (c < cend) {
        const float *ta = Tpad + ((uint64_t)lab * C + c) * R + 2 * d.M - f;
        const float *tb = Tpad + ((uint64_t)t * C + c) * R + 2 * d.M - s;
        const float *vc = V + (uint64_t)c * d.K + bidx;
        float ss = 0.0f;
        for (uint32_t j = 0; j < n8; j += 8) {
            float av[8], bv[8], vv[8];
#pragma unroll
            for (uint32_t q = 0; q < 8; q++) {
                av[q] = ta[j + q];
                bv[q] = tb[j + q];
                vv[q] = vc[j + q];
            }
#pragma unroll
            for (uint32_t q = 0; q < 8; q++) {
                const float sum = av[q] + bv[q];
                ss = __builtin_fmaf(sum, sum, ss);      // bp.cl:621
                acc = __builtin_fmaf(sum, vv[q], acc);  // bp.cl:622
            }
        }
        for (uint32_t j = n8; j < N; j++) {
            const float sum = ta[j] + tb[j];
            ss = __builtin_fmaf(sum, sum, ss);
            acc = __builtin_fmaf(sum, vc[j], acc);
        }
        if (ss > 0.0f) acc = (float)((double)acc - 0.5 * (double)ss);   // bp.cl:627-630
        // next channel of either support
        c++;
        if (!(c >= lo_m && c < hi_m) && !(c >= lo_t && c < hi_t)) {
            uint32_t nx = cend;
            if (lo_m > c) nx = min(nx, lo_m);
            if (lo_t > c) nx = min(nx, lo_t);
            c = nx;
        }
    }
    return acc;
}

template <bool GT>
__global__ __launch_bounds__(kXcThreads) __attribute__((amdgpu_waves_per_eu(GT ? 8 : 1))) void k_overlap_exact(
    Dims d, const float *__restrict__ V, const float *__restrict__ Tm, const float *__restrict__ L,
    const float *__restrict__ thr, const uint32_t *__restrict__ list, const uint32_t *__restrict__ xc_win,
    const uint32_t *__restrict__ xc_ctr, const uint32_t *__restrict__ xc_cand, Counters *ctr, float *best_L,
    const uint32_t *__restrict__ best_lab, const uint32_t *__restrict__ best_idx, uint32_t *ov_idx, uint32_t *ov_lab,
    uint8_t *ov, const uint32_t *__restrict__ sup_lo, const uint32_t *__restrict__ sup_hi,
    const float *__restrict__ Tg)
{
    // GT: Tm is the zero-padded [T][C][N + 2M] copy in global memory;
    // otherwise the [T][C][N] templates, staged here as padded rows (Tg:
    // the zero-padded copy, for the main rows shared by a batch)
    extern __shared__ float xs[];
    // every per-window value is wave-uniform (read through readfirstlane, so
    // it lives in scalar registers and the window's loads are scalar)
    auto uni = [](uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); };
    const uint32_t n_win = uni(xc_ctr[0]);
    // blocks with no deferred window leave before staging
    if (blockIdx.x * (blockDim.x >> 6) >= n_win) return;
    if (!GT) {   // rows of N samples, then M zeros; eight loads in flight per thread
        const uint32_t N = d.N, P = d.N + d.M;
        const FastDiv divP(P);
        for (uint32_t x = threadIdx.x; x < d.M; x += blockDim.x) xs[x] = 0.0f;
        batched_stage<8>(d.T * d.C * P, [&](uint32_t x) {
            const uint32_t r = divP.div(x), j = x - r * P;
            return j < N ? Tm[(uint64_t)r * N + j] : 0.0f;
        }, [&](uint32_t x, float v) { xs[d.M + x] = v; });
        __syncthreads();
    }
    const uint32_t lane = lane_of();
    const uint32_t nwave = blockDim.x >> 6;
    const uint32_t wg_bits = 31 - __builtin_clz(d.reduce_wg);
    const uint64_t span = d.nl * d.reduce_wg;
    const FastDiv divS(d.S);
    unsigned long long scored = 0;
    for (uint32_t w = uni(blockIdx.x * nwave + (threadIdx.x >> 6)); w < n_win; w = uni(w + gridDim.x * nwave)) {
        const uint32_t k = uni(xc_win[4 * w]), first = uni(xc_win[4 * w + 1]), count = uni(xc_win[4 * w + 2]);
        const uint32_t id = uni(list[k]), lab = uni(best_lab[id]), bidx = uni(best_idx[id]);
        const float bl = __uint_as_float(uni(__float_as_uint(best_L[id])));
        const uint32_t lo_m = uni(sup_lo[lab]), hi_m = uni(sup_hi[lab]);
        float sc = -INFINITY;
        uint64_t key = ~0ull, bp = 0;
        if (GT) {
            for (uint32_t i0 = 0; i0 < count; i0 += 64) {
                const bool act = i0 + lane < count;
                const uint32_t cnd = act ? xc_cand[first + i0 + lane] : 0u;
                const uint32_t f = cnd >> 16, e = cnd & 0xffffu;
                const uint32_t t = divS.div(e), s = e - t * d.S;
                const uint32_t lo_t = act ? sup_lo[t] : lo_m, hi_t = act ? sup_hi[t] : lo_m;
                const float score = exact_score_glb(d, Tm, V, bidx, lab, lo_m, hi_m, t, lo_t, hi_t, f, s, act);
                if (act) {
                    uint64_t p;
                    const uint64_t kk = combo_key(d, wg_bits, span, t, f, s, p);
                    if (key_better(score, kk, sc, key)) { sc = score; key = kk; bp = p; }
                }
            }
        } else {
            // two candidates per lane (128 per step, two independent chains
            // sharing the voltage and main-row loads); a step whose second
            // half is empty runs one chain
            for (uint32_t i0 = 0; i0 < count; i0 += 128) {
                bool act[2];
                uint32_t f[2], t[2], s[2];
                uint32_t c0 = lo_m, c1 = hi_m;
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    act[k] = i0 + 64 * k + lane < count;
                    const uint32_t cnd = act[k] ? xc_cand[first + i0 + 64 * k + lane] : 0u;
                    f[k] = cnd >> 16;
                    const uint32_t e = cnd & 0xffffu;
                    t[k] = divS.div(e);
                    s[k] = e - t[k] * d.S;
                    if (act[k]) {   // channel range: the union of the lanes' supports
                        c0 = min(c0, sup_lo[t[k]]);
                        c1 = max(c1, sup_hi[t[k]]);
                    }
                }
                if (!act[1]) { f[1] = f[0]; t[1] = t[0]; s[1] = s[0]; }
                for (int off = 32; off > 0; off >>= 1) {
                    c0 = min(c0, (uint32_t)__shfl_xor(c0, off));
                    c1 = max(c1, (uint32_t)__shfl_xor(c1, off));
                }
                const uint32_t cw0 = __builtin_amdgcn_readfirstlane(c0), cw1 = __builtin_amdgcn_readfirstlane(c1);
                const uint32_t f0 = __builtin_amdgcn_readfirstlane(f[0]);
                const bool same_f = __ballot((act[0] && f[0] != f0) || (act[1] && f[1] != f0)) == 0;
                const bool two = __ballot(act[1]) != 0;
                float score[2] = {0.0f, 0.0f};
                if (two) {
                    if (same_f) {
                        const uint32_t fu[2] = {f0, f0};
                        exact_score_pad<true, 2>(d, xs, V, Tg, bidx, lab, t, fu, s, cw0, cw1, score);
                    } else {
                        exact_score_pad<false, 2>(d, xs, V, Tg, bidx, lab, t, f, s, cw0, cw1, score);
                    }
                } else {
                    const uint32_t t1[1] = {t[0]}, f1[1] = {same_f ? f0 : f[0]}, s1[1] = {s[0]};
                    float sc1[1];
                    if (same_f) exact_score_pad<true, 1>(d, xs, V, Tg, bidx, lab, t1, f1, s1, cw0, cw1, sc1);
                    else exact_score_pad<false, 1>(d, xs, V, Tg, bidx, lab, t1, f1, s1, cw0, cw1, sc1);
                    score[0] = sc1[0];
                }
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    if (act[k]) {
                        uint64_t p;
                        const uint64_t kk = combo_key(d, wg_bits, span, t[k], f[k], s[k], p);
                        if (key_better(score[k], kk, sc, key)) { sc = score[k]; key = kk; bp = p; }
                    }
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const float o = __shfl_xor(sc, off);
            const uint64_t ok = __shfl_xor(key, off), op = __shfl_xor(bp, off);
            if (key_better(o, ok, sc, key)) { sc = o; key = ok; bp = op; }
        }
        if (lane == 0) {
            parse_overlap(d, L, thr, id, lab, bidx, bl, sc, bp, best_L, ov_idx, ov_lab, ov);
            scored += count;
        }
    }
    if (lane == 0 && scored) atomicAdd(&ctr->combos, scored);
}

// check_overlap_reassignments (bp.cl:822-875).  ov bit 1 marks a window that
// gave up, so the id-2 reads below see bit 0 as parsed.
__global__ __launch_bounds__(256) void k_check_overlap(Dims d, const uint32_t *__restrict__ ov_list,
                                                       const Counters *__restrict__ ctr, uint8_t *ov, uint8_t *nc,
                                                       uint32_t *best_idx, uint32_t *best_lab,
                                                       const uint32_t *ov_idx, const uint32_t *ov_lab)
{
    const uint32_t n_ov = ctr->n_ov;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n_ov; k += gridDim.x * blockDim.x) {
        const uint32_t id = ov_list[k];
        if (!(ov[id] & 1)) continue;
        const uint64_t start_w = (uint64_t)id * d.N, end_w = start_w + d.N;
        const uint32_t oi = ov_idx[id];
        best_idx[id] = oi;
        best_lab[id] = ov_lab[id];
        if (oi < start_w && id > 1) {
            if (ov[id - 2] & 1) {
                ov[id] = 3;
                best_idx[id] = 0;
                nc[id - 1] = 2;
                nc[id] = 2;
                nc[id + 1] = 2;
            } else {
                nc[id - 2] = 2;
            }
        }
        if (oi >= end_w && id + 2 < d.n_floor) nc[id + 2] = 2;
    }
}

// binary_pursuit, add half (bp.cl:961-999).
__global__ __launch_bounds__(256) void k_pursuit_add(Dims d, const uint32_t *__restrict__ win, uint32_t n_win,
                                                     const float *__restrict__ best_L,
                                                     const uint32_t *__restrict__ best_lab,
                                                     const uint32_t *__restrict__ best_idx,
                                                     const float *__restrict__ thr, const uint8_t *__restrict__ ov,
                                                     uint8_t *nc, uint8_t *has, uint32_t *cand_idx, uint32_t *cand_lab)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_win) return;
    const uint32_t id = win[k];
    const uint64_t start_w = (uint64_t)id * d.N, end_w = start_w + d.N;
    float ml = 0.0f;
    uint32_t lab = 0, idx = 0;
    if (end_w < (uint64_t)(d.K - d.N)) { ml = best_L[id]; lab = best_lab[id]; idx = best_idx[id]; }
    const bool o = ov[id] == 1;
    const bool hs = (ml > thr[lab] || o) && ((idx >= start_w && idx < end_w) || o);
    if (hs) {
        nc[id] = 3;
        if (id > 0) nc[id - 1] = 3;
        if (id + 1 < d.n_floor) nc[id + 1] = 3;
    }
    has[k] = hs ? 1 : 0;
    cand_idx[k] = idx;
    cand_lab[k] = lab;
}

// binary_pursuit, pruning half (bp.cl:1009-1045).  Bit 2 marks a pruned flag so
// neighbours keep reading the value it had after the adds.
__global__ __launch_bounds__(256) void k_pursuit_prune(Dims d, const uint32_t *__restrict__ win, uint32_t n_win,
                                                       const float *__restrict__ best_L,
                                                       const uint32_t *__restrict__ best_lab,
                                                       const uint32_t *__restrict__ best_idx,
                                                       const float *__restrict__ thr,
                                                       const uint8_t *__restrict__ has, uint8_t *nc)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_win) return;
    const uint32_t id = win[k];
    const uint64_t start_w = (uint64_t)id * d.N, end_w = start_w + d.N;
    if (!(end_w < (uint64_t)(d.K - d.N)) || has[k]) return;
    const float ml = best_L[id];
    const uint32_t lab = best_lab[id], idx = best_idx[id];
    const uint8_t orig = nc[id] & 3;
    uint8_t own = orig;
    if (idx < start_w && own == 1 && id > 0 && (nc[id - 1] & 3) == 0) own = 0;
    if (idx >= end_w && own == 1 && id + 1 < d.n_floor && (nc[id + 1] & 3) == 0) own = 0;
    if (ml <= thr[lab] && own < 3) own = 0;
    if (own == 0 && orig != 0) nc[id] = orig | 4;
}

// compute_residual (bp.cl:57-102) without lost updates.  One block per spike of
// the pass; thread (c, j) owns sample x = onset + j if this spike is the first,
// in (onset, label, position) order, of the pass's spikes covering x, and then
// applies all of them in that order.  Spikes are found through the map from
// adding window to spike: a spike added by window w lies in [wN-M, (w+1)N+M).
constexpr int kMaxCover = 32;

__global__ __launch_bounds__(256) void k_residual(Dims d, float *V, const float *__restrict__ Tm,
                                                  const uint32_t *__restrict__ sp_idx,
                                                  const uint32_t *__restrict__ sp_lab, uint32_t n_new,
                                                  const uint32_t *__restrict__ win_spike,
                                                  const uint32_t *__restrict__ sup_lo,
                                                  const uint32_t *__restrict__ sup_hi, uint64_t *wdirty)
{
    // spikes of the pass overlapping this spike's span, in (onset, label, position) order
    __shared__ uint32_t ci[kMaxCover], cl[kMaxCover], cp[kMaxCover];
    __shared__ int ncand;
    for (uint32_t k = blockIdx.x; k < n_new; k += gridDim.x) {
        const uint32_t idx = sp_idx[k], lab = sp_lab[k];
        if (lab >= d.T) continue;   // bp.cl:78-81 (block-uniform)
        __syncthreads();
        if (threadIdx.x == 0) {
            // the likelihood positions whose samples [i, i + N) meet
            // [idx, idx + N), on the channels where the template is nonzero
            // (elsewhere the subtraction leaves every sample's value), must
            // be computed again when listed
            const uint64_t cm = channel_mask(sup_lo[lab], sup_hi[lab], d.dirty_g);
            const uint64_t last = min((uint64_t)idx + d.N - 1, (uint64_t)d.K - 1);
            const uint32_t w0 = idx + 1 > d.N ? (idx + 1 - d.N) / d.N : 0u;
            const uint32_t w1 = min((uint32_t)(last / d.N), d.W - 1);
            for (uint32_t w = w0; w <= w1; w++) atomicOr((unsigned long long *)&wdirty[w], (unsigned long long)cm);
            // windows whose spike can overlap [idx, idx + N): a spike added by
            // window w lies in [wN - M, (w+1)N + M)
            const int64_t lo = (int64_t)idx - 2 * (int64_t)d.N - (int64_t)d.M + 2;
            const int64_t wlo = lo < 0 ? 0 : lo / d.N;
            int64_t whi = ((int64_t)idx + d.N - 1 + d.M) / d.N;
            if (whi > (int64_t)d.W - 1) whi = d.W - 1;
            int n = 0;
            for (int64_t wd = wlo; wd <= whi; wd++) {
                const uint32_t m = win_spike[wd];
                if (!m) continue;
                const uint32_t pos = m - 1;
                const uint32_t i2 = sp_idx[pos], l2 = sp_lab[pos];
                if (l2 >= d.T || (uint64_t)i2 + d.N <= idx || i2 >= (uint64_t)idx + d.N) continue;
                int q = n;
                while (q > 0 && (ci[q - 1] > i2 || (ci[q - 1] == i2 && (cl[q - 1] > l2 || (cl[q - 1] == l2 && cp[q - 1] > pos))))) {
                    ci[q] = ci[q - 1]; cl[q] = cl[q - 1]; cp[q] = cp[q - 1];
                    q--;
                }
                ci[q] = i2; cl[q] = l2; cp[q] = pos;
                if (n < kMaxCover - 1) n++;
            }
            ncand = n;
        }
        __syncthreads();
        const int n = ncand;
        for (uint32_t e = threadIdx.x; e < d.C * d.N; e += blockDim.x) {
            const uint32_t c = e / d.N, j = e % d.N;
            const uint64_t x = (uint64_t)idx + j;
            if (x >= d.K) continue;   // bp.cl:93-96
            // owner = first candidate (in order) covering x
            int first = -1;
            for (int q = 0; q < n; q++)
                if ((uint64_t)ci[q] <= x && x < (uint64_t)ci[q] + d.N) { first = q; break; }
            if (first < 0 || cp[first] != k) continue;
            float *vp = V + (uint64_t)c * d.K + x;
            float v = *vp;
            for (int q = first; q < n; q++)
                if ((uint64_t)ci[q] <= x && x < (uint64_t)ci[q] + d.N)
                    v -= Tm[((uint64_t)cl[q] * d.C + c) * d.N + (uint32_t)(x - ci[q])];
            *vp = v;
        }
    }
}

__global__ void k_clear_win_spike(const uint32_t *__restrict__ win, uint32_t n_win, uint32_t *win_spike)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_win) win_spike[win[k]] = 0;
}

// get_adjusted_clips (bp.cl:1088-1130): clip = residual + template, 0 past K.
__global__ __launch_bounds__(256) void k_clips(Dims d, const float *__restrict__ V, const float *__restrict__ Tm,
                                               const uint32_t *__restrict__ sp_idx,
                                               const uint32_t *__restrict__ sp_lab, uint64_t n,
                                               float *__restrict__ clips)
{
    const uint32_t cn = d.C * d.N;
    for (uint64_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint32_t idx = sp_idx[k], lab = sp_lab[k];
        for (uint32_t e = threadIdx.x; e < cn; e += blockDim.x) {
            const uint32_t c = e / d.N, j = e % d.N;
            float out = 0.0f;
            if (lab < d.T && (uint64_t)idx + j < d.K)
                out = V[(uint64_t)c * d.K + idx + j] + Tm[((uint64_t)lab * d.C + c) * d.N + j];
            clips[k * cn + e] = out;
        }
    }
}

inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 65535u * 4)
{
    uint64_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    return (uint32_t)(g > cap ? cap : g);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------
hipError_t launch_fill_u8(uint8_t *p, uint64_t n, uint8_t v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_reset_counters(Counters *c, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}

hipError_t launch_iota_u32(uint32_t *p, uint64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_copy_rows(float *dst, const float *src, uint32_t rows, uint64_t cols, uint64_t src_stride,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_copy_rows, dim3(grid_for((uint64_t)rows * cols, 256, 16384)), dim3(256), 0, st, dst, src,
                       rows, cols, src_stride);
    return hipGetLastError();
}

hipError_t launch_transpose_templates(const Dims &d, const float *Tm, const float *bias_in, float *TmT,
                                      float *bias_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_transpose_templates, dim3(grid_for((uint64_t)d.C * d.N * d.Tp, 256, 4096)), dim3(256), 0,
                       st, d, Tm, bias_in, TmT, bias_out);
    return hipGetLastError();
}

// The listed windows' likelihood is current: their dirty masks are consumed,
// after counting the (position, template) likelihoods the correlation
// computed (the blocks it did not skip): with ctb template tiles per chunk,
// a chunk is computed when its channel groups meet the mask (or it has
// none); with ctb = 0 (float32 path) a window is computed when its mask is
// not empty.
__global__ __launch_bounds__(256) void k_clear_dirty(Dims d, const uint32_t *__restrict__ win, uint32_t n_win,
                                                     uint64_t *wdirty, const uint32_t *__restrict__ tile_lo,
                                                     const uint32_t *__restrict__ tile_hi, uint32_t ctb,
                                                     Counters *ctr)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long pairs = 0;
    if (k < n_win) {
        const uint32_t w = win[k];
        const uint64_t m = wdirty[w];
        const uint64_t np = min((uint64_t)d.N, (uint64_t)d.K - (uint64_t)w * d.N);
        if (ctb == 0) {
            if (m || d.dirty_off) pairs = np * d.T;
        } else {
            const uint32_t CT = (d.T + 15) / 16;
            for (uint32_t c0 = 0; c0 < CT; c0 += ctb) {
                uint32_t lo = d.C, hi = 0;
                for (uint32_t ct = c0; ct < CT && ct < c0 + ctb; ct++)
                    if (tile_lo[ct] < tile_hi[ct]) {
                        lo = min(lo, tile_lo[ct]);
                        hi = max(hi, tile_hi[ct]);
                    }
                const uint64_t cm = hi > lo ? channel_mask(lo, hi, d.dirty_g) : 0ull;
                if (d.dirty_off || !cm || (m & cm)) pairs += np * (min(d.T, (c0 + ctb) * 16) - c0 * 16);
            }
        }
        wdirty[w] = 0;
    }
    for (int off = 32; off > 0; off >>= 1) pairs += __shfl_xor(pairs, off);
    if (lane_of() == 0 && pairs) atomicAdd(&ctr->corr_pairs, pairs);
}

hipError_t launch_corr_fp32(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st)
{
    if (n_win == 0) return hipSuccess;
    const uint32_t span = kWave + d.N - 1;
    uint32_t cb_max = kCorrLdsFloats / span;
    if (cb_max > d.C) cb_max = d.C;
    if (cb_max == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_corr_fp32, dim3(n_win), dim3(64), 0, st, d, s.V, s.TmT, s.bias, s.win, s.L, cb_max,
                       s.wdirty);
    hipLaunchKernelGGL(k_clear_dirty, dim3((n_win + 255) / 256), dim3(256), 0, st, d, s.win, n_win, s.wdirty,
                       s.tile_lo, s.tile_hi, 0u, s.ctr);
    return hipGetLastError();
}

uint32_t mfma_kstep(uint32_t) { return 32; }
uint32_t mfma_words(uint32_t precision) { return precision == 3 ? 1 : 2; }
// k-steps per channel: the template length padded to whole 16-sample row tiles
uint32_t mfma_ks(const Dims &d, uint32_t precision)
{
    const uint32_t padded = (d.N + 15) / 16 * 16;
    return (padded + mfma_kstep(precision) - 1) / mfma_kstep(precision);
}
uint32_t mfma_ct(const Dims &d) { return (d.T + 15) / 16; }
// Templates that span every channel share one staging of the window: up to
// four tiles per block.  Templates on narrow channel supports (at most half of
// the channels) take two tiles per block: a chunk then spans few channels, so
// the block stages little and more blocks (and waves) fit on a CU.
uint32_t mfma_ctb_for(uint32_t ct, uint32_t n_chans, uint32_t widest_support)
{
    if (ct >= 2 && 2 * widest_support <= n_chans) return 2;
    return ct >= 4 ? 4 : (ct == 3 ? 4 : ct);
}
uint32_t mfma_ctb(const Dims &d) { return d.corr_ctb ? d.corr_ctb : mfma_ctb_for(mfma_ct(d), d.C, d.C); }
uint32_t mfma_nchunk(const Dims &d) { return (mfma_ct(d) + mfma_ctb(d) - 1) / mfma_ctb(d); }

hipError_t launch_template_norms(const Dims &d, const DeviceState &s, uint32_t precision, hipStream_t st)
{
    hipLaunchKernelGGL(k_template_norms, dim3((d.T + 63) / 64), dim3(64), 0, st, d, s.Tm, s.tnorm, s.tscale,
                       precision == 3 ? 1u : 0u);
    return hipGetLastError();
}

template <int PREC>
static hipError_t launch_pack_p(const Dims &d, const DeviceState &s, hipStream_t st)
{
    const uint64_t total = (uint64_t)d.C * mfma_ks(d, PREC) * mfma_ct(d) * 64;
    hipLaunchKernelGGL(k_pack_templates<PREC>, dim3(grid_for(total, 256, 4096)), dim3(256), 0, st, d, s.Tm, s.tscale,
                       s.tord,
                       (Frag<PREC> *)s.TmB, mfma_ks(d, PREC), mfma_ct(d));
    return hipGetLastError();
}

hipError_t launch_pack_templates(const Dims &d, const DeviceState &s, uint32_t precision, hipStream_t st)
{
    if (precision == 1) return launch_pack_p<1>(d, s, st);
    if (precision == 2) return launch_pack_p<2>(d, s, st);
    return launch_pack_p<3>(d, s, st);
}

template <int PREC, int RT, int CTB>
static hipError_t launch_corr_mfma_rt(const Dims &d, const DeviceState &s, uint32_t n_win, const float *coef,
                                      uint32_t rq_cap, uint32_t no_rescore, hipStream_t st)
{
    using Tr = NarrowTraits<PREC>;
    const uint32_t LA = RT * 16 + mfma_ks(d, PREC) * Tr::kK;
    const uint32_t per_channel = Tr::kCopies * LA * Tr::kBytes;
    const uint32_t budget = 32 * 1024;
    uint32_t cb_max = budget / per_channel;
    // no more channels than the widest chunk spans (every chunk then stages once)
    const uint32_t span = d.corr_span ? d.corr_span : d.C;
    if (cb_max > span) cb_max = span;
    if (cb_max > d.C) cb_max = d.C;
    if (cb_max == 0 && span == 0) cb_max = 1;
    if (cb_max == 0) return hipErrorInvalidValue;
    const size_t lds = (size_t)cb_max * per_channel;
    ScreenCoef co{coef[0], coef[1], coef[2], coef[3]};
    hipLaunchKernelGGL((k_corr_mfma<PREC, RT, CTB>), dim3(n_win * mfma_nchunk(d)), dim3(64), lds, st, d, s.V, s.Tm,
                       (const Frag<PREC> *)s.TmB, s.bias, s.tnorm, s.tscale, s.win, s.L, s.ctr, s.rq, rq_cap,
                       (uint2 *)s.rdesc, cb_max, co, no_rescore, s.tord, s.tile_lo, s.tile_hi, s.sup_lo, s.sup_hi,
                       s.rq_shard, s.wdirty);
    hipLaunchKernelGGL(k_clear_dirty, dim3((n_win + 255) / 256), dim3(256), 0, st, d, s.win, n_win, s.wdirty,
                       s.tile_lo, s.tile_hi, (uint32_t)CTB, s.ctr);
    return hipGetLastError();
}

template <int PREC, int RT>
static hipError_t launch_corr_mfma_ct(const Dims &d, const DeviceState &s, uint32_t n_win, const float *coef,
                                      uint32_t rq_cap, uint32_t no_rescore, hipStream_t st)
{
    const uint32_t ctb = mfma_ctb(d);
    if (ctb == 1) return launch_corr_mfma_rt<PREC, RT, 1>(d, s, n_win, coef, rq_cap, no_rescore, st);
    if (ctb == 2) return launch_corr_mfma_rt<PREC, RT, 2>(d, s, n_win, coef, rq_cap, no_rescore, st);
    return launch_corr_mfma_rt<PREC, RT, 4>(d, s, n_win, coef, rq_cap, no_rescore, st);
}

template <int PREC>
static hipError_t launch_corr_mfma_p(const Dims &d, const DeviceState &s, uint32_t n_win, const float *coef,
                                     uint32_t rq_cap, uint32_t no_rescore, hipStream_t st)
{
    switch ((d.N + 15) / 16) {
    case 1: return launch_corr_mfma_ct<PREC, 1>(d, s, n_win, coef, rq_cap, no_rescore, st);
    case 2: return launch_corr_mfma_ct<PREC, 2>(d, s, n_win, coef, rq_cap, no_rescore, st);
    case 3: return launch_corr_mfma_ct<PREC, 3>(d, s, n_win, coef, rq_cap, no_rescore, st);
    case 4: return launch_corr_mfma_ct<PREC, 4>(d, s, n_win, coef, rq_cap, no_rescore, st);
    case 5: return launch_corr_mfma_ct<PREC, 5>(d, s, n_win, coef, rq_cap, no_rescore, st);
    case 6: return launch_corr_mfma_ct<PREC, 6>(d, s, n_win, coef, rq_cap, no_rescore, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_corr_screen(const Dims &d, const DeviceState &s, uint32_t n_win, uint32_t precision,
                              const float *coef, uint32_t rq_cap, uint32_t no_rescore, hipStream_t st)
{
    if (n_win == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_rescore, dim3(1), dim3(64), 0, st, s.ctr, s.rq_shard);
    hipError_t e;
    if (precision == 1) e = launch_corr_mfma_p<1>(d, s, n_win, coef, rq_cap, no_rescore, st);
    else if (precision == 2) e = launch_corr_mfma_p<2>(d, s, n_win, coef, rq_cap, no_rescore, st);
    else e = launch_corr_mfma_p<3>(d, s, n_win, coef, rq_cap, no_rescore, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_rescore, dim3(1), dim3(64), 0, st, s.ctr, s.rq_shard, rq_cap);
    if (!no_rescore) {
        const uint32_t nchunk = mfma_nchunk(d);
        const size_t rlds = rescore_res_lds(d);
        if (rlds <= 150 * 1024 && ((d.N + 15) / 16) * 16 + d.N <= kRescoreStage) {
            hipError_t e2 = hipFuncSetAttribute((const void *)k_rescore_res,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)rlds);
            if (e2 != hipSuccess) return e2;
            int dev = 0, ncu = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
                ncu = 256;
            // every block copies the templates into LDS first: no more blocks
            // than about one wave per listed window
            const uint32_t need = (n_win + kRescoreResWaves - 1) / kRescoreResWaves;
            const uint32_t grid = need < (uint32_t)ncu ? need : (uint32_t)ncu;
            hipLaunchKernelGGL(k_rescore_res, dim3(grid), dim3(64 * kRescoreResWaves), rlds, st, d, s.V, s.TmT,
                               s.bias, s.rq, s.rq_shard, rq_cap, s.L);
            return hipGetLastError();
        }
        hipLaunchKernelGGL(k_rescore_sup, dim3(n_win * nchunk), dim3(64), 0, st, d, s.V, s.Tm, s.bias, s.sup_lo,
                           s.sup_hi, s.rq, (const uint2 *)s.rdesc, s.L);
    }
    return hipGetLastError();
}

hipError_t launch_argmax(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st)
{
    if (n_win == 0) return hipSuccess;
    hipLaunchKernelGGL(k_window_partials, dim3((n_win + 3) / 4), dim3(256), 0, st, d, s.L, s.win, n_win, s.part_v,
                       s.part_t, s.part_p);
    hipLaunchKernelGGL(k_argmax, dim3((n_win + 255) / 256), dim3(256), 0, st, d, s.win, n_win, s.part_v, s.part_t,
                       s.part_p, s.best_L, s.best_lab, s.best_idx, s.ov_idx, s.ov_lab, s.nc, s.ov);
    return hipGetLastError();
}

hipError_t launch_compact_ov(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st)
{
    return compact(OvPred{s.win, s.ov}, OvEmit{s.win, s.ov_list}, n_win, s.blk, &s.ctr->n_ov, st);
}

// Overlap windows grouped by best label (counting sort, one block).  Blocks
// that run at the same time then mostly share a label, so the label's slice of
// the AB table and its padded template stay in each XCD's L2.  The order of
// the windows does not change any result.
constexpr uint32_t kSortMaxLabels = 8192;

__global__ __launch_bounds__(1024) void k_sort_ov_by_label(Dims d, const uint32_t *__restrict__ ov_list,
                                                           const Counters *__restrict__ ctr,
                                                           const uint32_t *__restrict__ best_lab, uint32_t *ov_sorted)
{
    __shared__ uint32_t cnt[kSortMaxLabels];
    const uint32_t n = ctr->n_ov, T = d.T;
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) cnt[t] = 0;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) atomicAdd(&cnt[best_lab[ov_list[k]]], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t t = 0; t < T; t++) {
            const uint32_t c = cnt[t];
            cnt[t] = run;
            run += c;
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        const uint32_t id = ov_list[k];
        ov_sorted[atomicAdd(&cnt[best_lab[id]], 1u)] = id;
    }
}

hipError_t launch_overlap_direct(const Dims &d, const DeviceState &s, uint32_t max_ov, hipStream_t st)
{
    if (max_ov == 0) return hipSuccess;
    const size_t lds = ((uint64_t)d.C * d.N + (uint64_t)d.C * (d.N + 2 * d.M)) * sizeof(float);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_overlap_direct,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const uint32_t grid = max_ov < 4096 ? max_ov : 4096;
    hipLaunchKernelGGL(k_overlap_direct, dim3(grid), dim3(256), lds, st, d, s.V, s.Tpad, s.L, s.thr, s.ov_list, s.ctr,
                       s.best_L, s.best_lab, s.best_idx, s.ov_idx, s.ov_lab, s.ov);
    return hipGetLastError();
}

constexpr uint64_t kLdsPerCu = 160 * 1024;

// The matrix-core Q table is used whenever its kernel's LDS fits.
static bool overlap_uses_mfma_q(const Dims &d) { return overlap_q_lds(d) <= 150 * 1024; }

// Dynamic LDS of k_overlap_screened<MQ, nw> (the layout at the top of the kernel).
static size_t overlap_screened_lds_nw(const Dims &d, int nw)
{
    const bool mq = overlap_uses_mfma_q(d);
    const bool qg = nw == 8 && mq;                       // Q table read from global memory
    const bool vg = nw != 4 && mq;                       // voltage read from global memory
    const uint64_t ts = (uint64_t)d.T * d.S, R = d.N + 2 * d.M;
    const uint64_t krw = (ts + 1) / 2, cand = 4ull * nw * 64;
    const uint64_t partner = (uint64_t)d.sup_max * R;                           // NW 4: staged partner rows
    const uint64_t main_rows = nw != 4 ? 0 : partner + R;                        // main rows + the zero row
    const uint64_t region = nw != 4 ? krw + cand : (partner + cand > krw ? partner + cand : krw);
    return (uint64_t)d.T * 8 + (vg ? 0 : (uint64_t)d.C * d.N * 4) + main_rows * 4 + (qg ? 0 : ts * 4) +
           (uint64_t)d.T * 4 + (uint64_t)d.S * 4 + region * 4 + ts * 2 + (uint64_t)d.T * 6 + ts + 4 +
           (uint64_t)d.T * 12 + (vg ? (uint64_t)d.C * 4 : 0) + 16;
}

// Waves per screen block: 4 when the block scores its candidates; with the
// exact scores deferred (lean block) 2, or 8 when fewer than four two-wave
// blocks fit in a CU's LDS.
static int overlap_screened_waves(const Dims &d)
{
    if (!d.xc_cap) return 4;
    return !d.xc_wide && overlap_screened_lds_nw(d, 2) * 4 <= kLdsPerCu ? 2 : 8;
}

size_t overlap_screened_lds(const Dims &d)
{
    // the block that runs
    const size_t a = overlap_screened_lds_nw(d, overlap_screened_waves(d));
    const size_t b = overlap_uses_mfma_q(d) ? overlap_q_lds(d) : 0;
    return a > b ? a : b;
}

uint64_t overlap_qtemplate_floats(const Dims &d)
{
    // one uint4 per entry: the full tiles' fragments, then the tail tile's
    const QPlan qp = q_plan(d);
    return ((uint64_t)d.C * qp.NKQ * qp.TT + (uint64_t)d.C * qp.NKT) * 64 * 4;
}

hipError_t launch_pad_templates(const Dims &d, const DeviceState &s, hipStream_t st)
{
    const uint64_t total = (uint64_t)d.T * d.C * (d.N + 2 * d.M);
    hipLaunchKernelGGL(k_pad_templates, dim3(grid_for(total, 256, 4096)), dim3(256), 0, st, d, s.Tm, s.Tpad);
    return hipGetLastError();
}

// ABmin[lab][f][t] = min over s of AB[lab][f][t][s] (the row bounds of the screen).
__global__ void k_overlap_abmin(Dims d, const float *__restrict__ AB, float *__restrict__ ABmin)
{
    const uint64_t total = (uint64_t)d.T * d.S * d.T;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const float *row = AB + e * d.S;   // [lab][f][t] rows of S
        float m = INFINITY;
        for (uint32_t s = 0; s < d.S; s++) m = fminf(m, row[s]);
        ABmin[e] = m;
    }
}

hipError_t launch_overlap_tables(const Dims &d, const DeviceState &s, hipStream_t st)
{
    const uint64_t total = (uint64_t)d.T * d.S * d.T * d.S + (uint64_t)d.T * d.S + d.T;
    hipLaunchKernelGGL(k_overlap_tables, dim3(grid_for(total, 256, 65536)), dim3(256), 0, st, d, s.Tm, s.tabAB,
                       s.tabB2, s.tnorm, s.sup_lo, s.sup_hi);
    const uint64_t nmin = (uint64_t)d.T * d.S * d.T;
    hipLaunchKernelGGL(k_overlap_abmin, dim3(grid_for(nmin, 256, 65536)), dim3(256), 0, st, d, s.tabAB, s.tabABmin);
    const uint64_t nq = overlap_qtemplate_floats(d);
    hipLaunchKernelGGL(k_pack_qtemplates, dim3(grid_for(nq / 4, 256, 4096)), dim3(256), 0, st, d, s.Tm, s.tord,
                       (uint4 *)s.TmQ);
    return hipGetLastError();
}

uint32_t overlap_exact_cap(const Dims &d)
{
    // a channel's window samples are held one per lane
    if (d.N > (uint32_t)kWave) return 0;
    uint64_t cap = (uint64_t)d.ov_batch * 256;
    if (cap > (32ull << 20)) cap = 32ull << 20;
    if (cap < 4096) cap = 4096;
    return (uint32_t)cap;
}

hipError_t launch_overlap_screened(const Dims &d, const DeviceState &s, uint32_t max_ov, uint32_t debug,
                                  hipStream_t st)
{
    if (max_ov == 0) return hipSuccess;
    const bool mq = overlap_uses_mfma_q(d);
    const int nw = overlap_screened_waves(d);
    const size_t lds = overlap_screened_lds_nw(d, nw);
    const void *fn = nw == 2   ? (mq ? (const void *)k_overlap_screened<true, 2> : (const void *)k_overlap_screened<false, 2>)
                     : nw == 8 ? (mq ? (const void *)k_overlap_screened<true, 8> : (const void *)k_overlap_screened<false, 8>)
                               : (mq ? (const void *)k_overlap_screened<true, 4> : (const void *)k_overlap_screened<false, 4>);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const uint32_t *list = s.ov_list;
    // grouping by label pays when a label's AB slice is large (measured: a
    // 3 MB slice, 200 templates, gains; a 0.3 MB one, 20 templates, does not
    // repay the sort's launch)
    const uint64_t ab_slice = (uint64_t)d.S * d.T * d.S * 4;
    if (d.T <= kSortMaxLabels && ab_slice > (1ull << 20)) {
        hipLaunchKernelGGL(k_sort_ov_by_label, dim3(1), dim3(1024), 0, st, d, s.ov_list, s.ctr, s.best_lab,
                           s.ov_sorted);
        list = s.ov_sorted;
    }
    const size_t qlds = overlap_q_lds(d);
    if (mq && qlds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(overlap_q_copies(d) == 2 ? (const void *)k_overlap_q<2> : (const void *)k_overlap_q<1>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)qlds);
        if (e != hipSuccess) return e;
    }
    // exact scores: templates staged in LDS when they fit, else read from Tpad
    const size_t xlds_full = (size_t)overlap_exact_lds_floats(d) * 4;
    const bool xg = overlap_exact_global(d);
    const size_t xlds = xg ? 0 : xlds_full;
    if (d.xc_cap && xlds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_overlap_exact<false>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)xlds);
        if (e != hipSuccess) return e;
    }
    // batches of ov_batch windows: Q tables (matrix cores), then the screen
    for (uint32_t kb = 0; kb < max_ov; kb += d.ov_batch) {
        const uint32_t nb = max_ov - kb < d.ov_batch ? max_ov - kb : d.ov_batch;
        // up to 8192 blocks (eight per resident slot at four per CU): windows
        // differ in work, and smaller grids leave a longer tail (measured:
        // 2048 and 4096 slower, 16384 no faster, one block per window slower);
        // the lean two-wave blocks are resident eight per CU, so twice as many
        const uint32_t gmax = nw == 2 ? 16384 : 8192;
        const uint32_t grid = nb < gmax ? nb : gmax;
        if (mq) {
            if (overlap_q_copies(d) == 2)
                hipLaunchKernelGGL(k_overlap_q<2>, dim3(grid), dim3(256), qlds, st, d, s.V, s.L, s.thr, list, s.ctr,
                                   s.best_lab, s.best_idx, s.tabB2, (const uint4 *)s.TmQ, s.ovQ, kb, s.tord, s.tile_lo,
                                   s.tile_hi, s.ovVV);
            else
                hipLaunchKernelGGL(k_overlap_q<1>, dim3(grid), dim3(256), qlds, st, d, s.V, s.L, s.thr, list, s.ctr,
                                   s.best_lab, s.best_idx, s.tabB2, (const uint4 *)s.TmQ, s.ovQ, kb, s.tord, s.tile_lo,
                                   s.tile_hi, s.ovVV);
        }
        if (d.xc_cap) {
            hipError_t e = launch_fill_u32(s.xc_ctr, 2, 0u, st);
            if (e != hipSuccess) return e;
        }
        {
#define FBP_SCREEN_ARGS d, s.V, s.Tm, s.L, s.thr, list, s.ctr, s.best_L, s.best_lab, s.best_idx, s.ov_idx, s.ov_lab, \
                        s.ov, s.tabAB, s.tabABmin, s.tabB2, s.tnorm, s.Tpad, s.ovQ, kb, s.sup_lo, s.sup_hi, debug, \
                        s.xc_cand, s.xc_win, s.xc_ctr, s.ovVV
            if (nw == 2 && mq)
                hipLaunchKernelGGL((k_overlap_screened<true, 2>), dim3(grid), dim3(128), lds, st, FBP_SCREEN_ARGS);
            else if (nw == 2)
                hipLaunchKernelGGL((k_overlap_screened<false, 2>), dim3(grid), dim3(128), lds, st, FBP_SCREEN_ARGS);
            else if (nw == 8 && mq)
                hipLaunchKernelGGL((k_overlap_screened<true, 8>), dim3(grid), dim3(512), lds, st, FBP_SCREEN_ARGS);
            else if (nw == 8)
                hipLaunchKernelGGL((k_overlap_screened<false, 8>), dim3(grid), dim3(512), lds, st, FBP_SCREEN_ARGS);
            else if (mq)
                hipLaunchKernelGGL((k_overlap_screened<true, 4>), dim3(grid), dim3(256), lds, st, FBP_SCREEN_ARGS);
            else
                hipLaunchKernelGGL((k_overlap_screened<false, 4>), dim3(grid), dim3(256), lds, st, FBP_SCREEN_ARGS);
#undef FBP_SCREEN_ARGS
        }
        if (d.xc_cap) {
            // every deferred window of the batch: one wave each, 16 per block,
            // at most two blocks per CU (each stages the templates in LDS
            // before its first window, so a grid sized for every overlap
            // window would mostly stage for nothing: few windows defer)
            uint32_t ncu = 256;
            {
                int dev = 0, n = 0;
                if (hipGetDevice(&dev) == hipSuccess &&
                    hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
                    ncu = (uint32_t)n;
            }
            const uint32_t xmax = 2 * ncu;
            const uint32_t xgrid = (nb + 15) / 16 < xmax ? (nb + 15) / 16 : xmax;
            if (xg)
                hipLaunchKernelGGL(k_overlap_exact<true>, dim3(xgrid), dim3(kXcThreads), 0, st, d, s.V, s.Tpad, s.L,
                                   s.thr, list, s.xc_win, s.xc_ctr, s.xc_cand, s.ctr, s.best_L, s.best_lab,
                                   s.best_idx, s.ov_idx, s.ov_lab, s.ov, s.sup_lo, s.sup_hi, s.Tpad);
            else
                hipLaunchKernelGGL(k_overlap_exact<false>, dim3(xgrid), dim3(kXcThreads), xlds, st, d, s.V, s.Tm,
                                   s.L, s.thr, list, s.xc_win, s.xc_ctr, s.xc_cand, s.ctr, s.best_L, s.best_lab,
                                   s.best_idx, s.ov_idx, s.ov_lab, s.ov, s.sup_lo, s.sup_hi, s.Tpad);
        }
    }
    return hipGetLastError();
}

hipError_t launch_check_overlap(const Dims &d, const DeviceState &s, uint32_t max_ov, hipStream_t st)
{
    if (max_ov == 0) return hipSuccess;
    hipLaunchKernelGGL(k_check_overlap, dim3(grid_for(max_ov, 256, 1024)), dim3(256), 0, st, d, s.ov_list, s.ctr,
                       s.ov, s.nc, s.best_idx, s.best_lab, s.ov_idx, s.ov_lab);
    return hipGetLastError();
}

hipError_t launch_pursuit_add(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st)
{
    if (n_win == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pursuit_add, dim3((n_win + 255) / 256), dim3(256), 0, st, d, s.win, n_win, s.best_L,
                       s.best_lab, s.best_idx, s.thr, s.ov, s.nc, s.has, s.cand_idx, s.cand_lab);
    return hipGetLastError();
}

hipError_t launch_compact_spikes(const Dims &d, const DeviceState &s, uint32_t n_win, uint64_t spike_base,
                                 hipStream_t st)
{
    return compact(SpikePred{s.has},
                   SpikeEmit{s.win, s.cand_idx, s.cand_lab, s.sp_idx + spike_base, s.sp_lab + spike_base, s.win_spike},
                   n_win, s.blk, &s.ctr->n_new, st);
}

hipError_t launch_pursuit_prune(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st)
{
    if (n_win == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pursuit_prune, dim3((n_win + 255) / 256), dim3(256), 0, st, d, s.win, n_win, s.best_L,
                       s.best_lab, s.best_idx, s.thr, s.has, s.nc);
    return hipGetLastError();
}

hipError_t launch_compact_next(const Dims &d, const DeviceState &s, hipStream_t st)
{
    return compact(NextPred{s.nc}, NextEmit{s.win_next}, d.W, s.blk, &s.ctr->n_next, st);
}

hipError_t launch_residual(const Dims &d, const DeviceState &s, uint64_t spike_base, uint32_t n_new, hipStream_t st)
{
    if (n_new == 0) return hipSuccess;
    hipLaunchKernelGGL(k_residual, dim3(grid_for(n_new, 1, 65536)), dim3(256), 0, st, d, s.V, s.Tm,
                       s.sp_idx + spike_base, s.sp_lab + spike_base, n_new, s.win_spike, s.sup_lo, s.sup_hi,
                       s.wdirty);
    return hipGetLastError();
}

hipError_t launch_clear_win_spike(const Dims &d, const DeviceState &s, uint32_t n_win, hipStream_t st)
{
    if (n_win == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_win_spike, dim3((n_win + 255) / 256), dim3(256), 0, st, s.win, n_win, s.win_spike);
    return hipGetLastError();
}

hipError_t launch_clips(const Dims &d, const DeviceState &s, uint64_t n_spikes, hipStream_t st)
{
    if (n_spikes == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clips, dim3(grid_for(n_spikes, 1, 65536)), dim3(256), 0, st, d, s.V, s.Tm, s.sp_idx,
                       s.sp_lab, n_spikes, s.clips);
    return hipGetLastError();
}

}  // namespace fbp
