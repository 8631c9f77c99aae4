/*
 * Address/undefined-behaviour sanitizer run of the CPU oracle
 * (oracle/fbp_oracle.c, test infrastructure): seeded small cases over the
 * shapes the tests use and their edges (one channel, one template, K = N + 1,
 * shift ranges 0 .. N - 1, channels where a template is zero, pass caps),
 * through every entry point, with sanity checks of the outputs.  Built by
 * tests/native/Makefile, run by tests/test_sanitizers.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* the oracle's parameter block and entry points (oracle/fbp_oracle.c) */
typedef struct {
    uint32_t n_chans, n_samples, n_templates, template_len, max_shift, do_overlap, reduce_wg, max_passes;
} fbo_params;
int fbo_likelihood(const fbo_params *p, const float *V, const float *Tm, const float *bias, const uint32_t *windows,
                   uint64_t n_windows, float *L);
int fbo_likelihood_f64(const fbo_params *p, const float *V, const float *Tm, const float *bias_sumsq,
                       const uint32_t *windows, uint64_t n_windows, double *L);
int fbo_residual(const fbo_params *p, float *V, const float *Tm, const uint32_t *idx, const uint32_t *lab,
                 uint64_t n);
int fbo_adjusted_clips(const fbo_params *p, const float *V, const float *Tm, const uint32_t *idx,
                       const uint32_t *lab, uint64_t n, float *clips);
int fbo_pursuit_chunk(const fbo_params *p, float *V, const float *Tm, const float *bias, const float *thr,
                      uint32_t *out_idx, uint32_t *out_lab, uint64_t cap, uint64_t *n_out, uint32_t *n_passes,
                      uint32_t *pass_counts, uint32_t pass_counts_cap, float *L_out, uint32_t *next_windows,
                      uint64_t *n_next);

#define CHECK(c)                                                                     \
    do {                                                                             \
        if (!(c)) {                                                                  \
            fprintf(stderr, "%s:%d: check failed: %s (case %d)\n", __FILE__, __LINE__, #c, g_case); \
            exit(1);                                                                 \
        }                                                                            \
    } while (0)

static int g_case;
static uint64_t g_rng = 88172645463325252ull;

static uint32_t rnd(void)
{
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return (uint32_t)(g_rng >> 11);
}

static float unif(void) { return (float)(rnd() & 0xffffff) / 16777216.0f; }

static void run_case(uint32_t C, uint32_t K, uint32_t T, uint32_t N, uint32_t M, uint32_t do_overlap,
                     uint32_t max_passes)
{
    const uint32_t W = (K + N - 1) / N;
    fbo_params p = {C, K, T, N, M, do_overlap, 256, max_passes};
    float *Tm = calloc((size_t)T * C * N, 4), *bias = malloc(T * 4), *thr = malloc(T * 4);
    float *V = calloc((size_t)C * K, 4), *V0 = malloc((size_t)C * K * 4);
    CHECK(Tm && bias && thr && V && V0);
    for (uint32_t t = 0; t < T; t++) {
        double ss = 0;
        for (uint32_t c = 0; c < C; c++) {
            if (C > 1 && rnd() % 3 == 0) continue;   /* a channel where the template is zero */
            for (uint32_t j = 0; j < N; j++) {
                const float v = 4.0f * (unif() - 0.5f) * expf(-0.5f * ((float)j - N / 2.0f) * ((float)j - N / 2.0f) /
                                                              (1.0f + N / 4.0f));
                Tm[((size_t)t * C + c) * N + j] = v;
                ss += (double)v * v;
            }
        }
        bias[t] = (float)(-0.5 * ss);
        thr[t] = (float)(0.2 * ss);
    }
    for (size_t k = 0; k < (size_t)C * K; k++) V[k] = 0.2f * (unif() - 0.5f);
    const uint32_t n_spk = 1 + rnd() % (2 * W + 1);
    for (uint32_t s = 0; s < n_spk; s++) {
        const uint32_t t = rnd() % T, on = rnd() % K;
        for (uint32_t c = 0; c < C; c++)
            for (uint32_t j = 0; j < N && on + j < K; j++) V[(size_t)c * K + on + j] += Tm[((size_t)t * C + c) * N + j];
    }
    memcpy(V0, V, (size_t)C * K * 4);

    uint64_t cap = 4ull * W + 16, n = 0, n_next = 0;
    uint32_t *idx = malloc(cap * 4), *lab = malloc(cap * 4), *next = malloc(W * 4), passes = 0, counts[64];
    float *L = malloc((size_t)T * K * 4);
    CHECK(idx && lab && next && L);
    int rc = fbo_pursuit_chunk(&p, V, Tm, bias, thr, idx, lab, cap, &n, &passes, counts, 64, L, next, &n_next);
    CHECK(rc == 0 || rc == -3 || rc == -4);   /* ok, spike capacity, pass cap */
    if (rc == -3) {
        /* more spikes than the capacity: run again with room for every pass */
        free(idx);
        free(lab);
        cap = 64ull * W + 64;
        idx = malloc(cap * 4);
        lab = malloc(cap * 4);
        memcpy(V, V0, (size_t)C * K * 4);
        rc = fbo_pursuit_chunk(&p, V, Tm, bias, thr, idx, lab, cap, &n, &passes, counts, 64, L, next, &n_next);
        CHECK(rc == 0 || rc == -4);
    }
    CHECK(n <= cap && passes >= 1);
    CHECK(max_passes == 0 || passes <= max_passes);
    CHECK(rc != 0 || n_next == 0);
    for (uint64_t k = 0; k < n; k++) CHECK(idx[k] < K && lab[k] < T);
    for (uint64_t k = 0; k < n_next; k++) CHECK(next[k] < W && (k == 0 || next[k] > next[k - 1]));

    /* clips of the found spikes on the residual, and the residual rebuilt by
       subtracting them from the input */
    float *clips = malloc((n ? n : 1) * (size_t)C * N * 4);
    CHECK(clips);
    CHECK(fbo_adjusted_clips(&p, V, Tm, idx, lab, n, clips) == 0);
    float *R = malloc((size_t)C * K * 4);
    memcpy(R, V0, (size_t)C * K * 4);
    CHECK(fbo_residual(&p, R, Tm, idx, lab, n) == 0);
    for (size_t k = 0; k < (size_t)C * K; k++) CHECK(fabsf(R[k] - V[k]) < 1e-3f);

    /* likelihood on every window, float32 and float64 */
    uint32_t *wins = malloc(W * 4);
    for (uint32_t w = 0; w < W; w++) wins[w] = w;
    float *L2 = calloc((size_t)T * K, 4);
    double *Ld = calloc((size_t)T * K, 8);
    float *ss = malloc(T * 4);
    for (uint32_t t = 0; t < T; t++) ss[t] = -2.0f * bias[t];
    CHECK(fbo_likelihood(&p, V0, Tm, bias, wins, W, L2) == 0);
    CHECK(fbo_likelihood_f64(&p, V0, Tm, ss, wins, W, Ld) == 0);
    for (size_t k = 0; k < (size_t)T * K; k++) CHECK(fabs((double)L2[k] - Ld[k]) < 1e-2 * (1.0 + fabs(Ld[k])));

    free(Tm); free(bias); free(thr); free(V); free(V0); free(idx); free(lab); free(next); free(L);
    free(clips); free(R); free(wins); free(L2); free(Ld); free(ss);
}

int main(void)
{
    static const uint32_t Ns[] = {1, 2, 7, 16, 41};
    int n_cases = 0;
    for (int rep = 0; rep < 3; rep++)
        for (size_t a = 0; a < sizeof(Ns) / sizeof(Ns[0]); a++) {
            const uint32_t N = Ns[a];
            const uint32_t Ms[] = {0, 1, N / 2, N > 1 ? N - 1 : 0};
            for (int b = 0; b < 4; b++) {
                const uint32_t C = 1 + rnd() % 5, T = 1 + rnd() % 6;
                const uint32_t Ks[] = {N + 1, 3 * N + 5, 40 * N + rnd() % N};
                for (int c = 0; c < 3; c++) {
                    g_case = n_cases++;
                    run_case(C, Ks[c], T, N, Ms[b], (uint32_t)((rep + b) % 2 == 0 || c == 2), (uint32_t)(rnd() % 3));
                }
            }
        }
    printf("oracle checks passed: %d cases\n", n_cases);
    return 0;
}
