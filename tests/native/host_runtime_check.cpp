// Address/undefined-behaviour sanitizer run of libfbp's host code on a
// machine without a GPU: the library's sources with the sanitizers on the
// host side (the device code is built as usual and never run), driven
// through the C-ABI paths that complete without a device -- sizing and validation of segment configurations, argument checks
// of every entry point, the counting and trimming done on the host before any
// upload, and the no-device errors of segment and server creation.  Built by
// tests/native/Makefile, run by tests/test_sanitizers.py.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/fbp.h"

#define CHECK(c)                                                                    \
    do {                                                                            \
        if (!(c)) {                                                                 \
            std::fprintf(stderr, "%s:%d: check failed: %s (last error: %s)\n", __FILE__, __LINE__, #c, \
                         fbp_last_error());                                         \
            std::exit(1);                                                           \
        }                                                                           \
    } while (0)

static fbp_config cfg(uint32_t C, uint32_t K, uint32_t T, uint32_t N, uint32_t M, uint32_t prec, uint32_t mode)
{
    fbp_config c{};
    c.n_chans = C;
    c.n_samples = K;
    c.n_templates = T;
    c.template_len = N;
    c.max_shift = M;
    c.do_overlap = 1;
    c.reduce_wg = 256;
    c.precision = prec;
    c.overlap_mode = mode;
    return c;
}

static void sizing()
{
    uint64_t a = 0, b = 0;
    // every precision and overlap mode over a range of shapes
    const uint32_t shapes[][5] = {{1, 62, 1, 61, 0},       {4, 2400000, 5, 61, 30}, {16, 12000000, 20, 61, 60},
                                  {32, 3600000, 40, 81, 40}, {128, 18000000, 200, 61, 30}, {3, 100, 7, 7, 6}};
    for (auto &s : shapes)
        for (uint32_t prec = 0; prec < 4; prec++)
            for (uint32_t mode = 0; mode < 2; mode++) {
                fbp_config c = cfg(s[0], s[1], s[2], s[3], s[4], prec, mode);
                uint64_t bytes = 0;
                const int rc = fbp_segment_bytes(&c, &bytes);
                if (rc == FBP_OK) {
                    CHECK(bytes >= (uint64_t)2 * s[0] * s[1] * 4 + (uint64_t)s[2] * s[1] * 4);
                    CHECK(bytes % 256 == 0);
                } else {
                    CHECK(rc == FBP_EINVAL && std::strlen(fbp_last_error()) > 0);
                }
            }
    // more samples, more bytes; the screened overlap and the matrix-core
    // screen add their tables
    fbp_config c1 = cfg(8, 100000, 10, 61, 30, FBP_PREC_FP32, FBP_OVERLAP_DIRECT);
    fbp_config c2 = c1;
    c2.n_samples = 200000;
    CHECK(fbp_segment_bytes(&c1, &a) == FBP_OK && fbp_segment_bytes(&c2, &b) == FBP_OK && b > a);
    c2 = c1;
    c2.overlap_mode = FBP_OVERLAP_SCREENED;
    CHECK(fbp_segment_bytes(&c2, &b) == FBP_OK && b > a);
    c2 = c1;
    c2.precision = FBP_PREC_BF16;
    CHECK(fbp_segment_bytes(&c2, &b) == FBP_OK && b > a);
    // invalid configurations name their problem
    fbp_config bad = cfg(4, 61, 5, 61, 30, 0, 0);
    CHECK(fbp_segment_bytes(&bad, &a) == FBP_EINVAL && std::strstr(fbp_last_error(), "not longer"));
    bad = cfg(4, 1000, 5, 61, 300, 0, 0);
    CHECK(fbp_segment_bytes(&bad, &a) == FBP_EINVAL);
    bad = cfg(4, 1000, 5, 61, 30, 0, 0);
    bad.reduce_wg = 100;
    CHECK(fbp_segment_bytes(&bad, &a) == FBP_EINVAL && std::strstr(fbp_last_error(), "power of two"));
    bad = cfg(4, 1000, 5, 61, 30, 9, 0);
    CHECK(fbp_segment_bytes(&bad, &a) == FBP_EINVAL);
    bad = cfg(4, 1000, 5, 101, 30, FBP_PREC_BF16, 0);
    CHECK(fbp_segment_bytes(&bad, &a) == FBP_EINVAL);
    bad = cfg(4096, 1000, 5, 61, 30, 0, 0);
    CHECK(fbp_segment_bytes(&bad, &a) == FBP_EINVAL);
    CHECK(fbp_segment_bytes(nullptr, &a) == FBP_EINVAL);
    std::printf("sizing: ok\n");
}

static void host_side_work()
{
    // get_clips: events trimmed at both ends on the host; no clip left means
    // no upload
    const int64_t ev[] = {-5, -1, 3, 7};
    uint8_t valid[4];
    uint64_t n_valid = 99;
    const uint32_t chans[] = {0, 1};
    float dummy = 0;
    CHECK(fbp_get_clips_host(0, &dummy, 10, 2, 10, ev, 4, chans, 2, 0, 5, 1, nullptr, valid, &n_valid) == FBP_EINVAL);
    const int64_t ev2[] = {-5, -1, 9};
    CHECK(fbp_get_clips_host(0, &dummy, 10, 2, 10, ev2, 3, chans, 2, 0, 5, 1, nullptr, valid, &n_valid) == FBP_OK);
    CHECK(n_valid == 0 && valid[0] == 0 && valid[1] == 0 && valid[2] == 0);
    CHECK(fbp_get_clips_host(0, &dummy, 10, 2, 10, ev2, 3, chans, 2, 5, 5, 1, nullptr, valid, &n_valid) == FBP_EINVAL);
    const uint32_t badch[] = {2};
    CHECK(fbp_get_clips_host(0, &dummy, 10, 2, 10, ev2, 3, badch, 1, 0, 5, 1, nullptr, valid, &n_valid) ==
          FBP_EINVAL);
    CHECK(fbp_get_clips_host(0, &dummy, 5, 2, 10, ev2, 3, chans, 2, 0, 5, 1, nullptr, valid, &n_valid) == FBP_EINVAL);
    CHECK(fbp_get_clips_host(0, &dummy, 10, 2, 10, nullptr, 0, chans, 2, 0, 5, 1, nullptr, nullptr, &n_valid) ==
          FBP_OK);

    // unit means: labels counted on the host (width 0: nothing to upload)
    const uint32_t labels[] = {2, 0, 2, 1, 2, 0};
    uint64_t counts[3] = {0, 0, 0};
    CHECK(fbp_unit_mean_clips(0, &dummy, 6, 0, labels, 3, &dummy, counts) == FBP_OK);
    CHECK(counts[0] == 2 && counts[1] == 1 && counts[2] == 3);
    double dd = 0;
    CHECK(fbp_unit_mean_clips_f64(0, &dd, 6, 0, labels, 3, &dd, counts) == FBP_OK && counts[2] == 3);
    CHECK(fbp_unit_mean_clips(0, &dummy, 6, 4, labels, 2, &dummy, counts) == FBP_EINVAL);

    // overlap template residuals, clip SSE, covariance, variances: argument
    // checks and empty inputs
    std::vector<float> tm(3 * 2 * 5, 1.0f);
    const uint32_t tri[] = {0, 1, 3};
    double best = 0;
    uint32_t combo = 0;
    CHECK(fbp_overlap_template_residuals(0, tm.data(), 3, 2, 5, 2, 2, tri, 1, &best, &combo) == FBP_EINVAL);
    CHECK(fbp_overlap_template_residuals(0, tm.data(), 3, 2, 5, -1, 2, tri, 0, &best, &combo) == FBP_EINVAL);
    CHECK(fbp_overlap_template_residuals(0, tm.data(), 3, 2, 5, 2, 2, tri, 0, &best, &combo) == FBP_OK);
    const uint32_t ti[] = {0, 5};
    float sse[2];
    CHECK(fbp_clip_template_sse(0, tm.data(), 2, 10, ti, tm.data(), 3, sse) == FBP_EINVAL);
    CHECK(fbp_clip_template_sse(0, tm.data(), 0, 10, ti, tm.data(), 3, sse) == FBP_OK);
    CHECK(fbp_clip_template_sse(0, tm.data(), 2, 0, ti, tm.data(), 3, sse) == FBP_EINVAL);
    const uint64_t rows[] = {0, 7};
    double cov = 0;
    CHECK(fbp_channel_covariance(0, tm.data(), 3, 1, 10, rows, 2, &cov) == FBP_EINVAL);
    CHECK(fbp_channel_covariance(0, tm.data(), 3, 1, 0, nullptr, 0, &cov) == FBP_EINVAL);
    CHECK(fbp_channel_covariance(0, tm.data(), 3, 0, 10, nullptr, 0, &cov) == FBP_OK);
    CHECK(fbp_template_variances(0, &dd, 0, 2, 5, &dd, &dd) == FBP_OK);
    CHECK(fbp_template_variances(0, &dd, 1, 2, 0, &dd, &dd) == FBP_EINVAL);
    std::printf("host-side work: ok\n");
}

static void no_device()
{
    int n = -1;
    CHECK(fbp_device_count(&n) == FBP_ENODEV && n == 0);
    uint64_t t = 0, f = 0;
    CHECK(fbp_device_memory(0, &t, &f) == FBP_ENODEV);
    fbp_config c = cfg(4, 1000, 3, 61, 30, FBP_PREC_BF16, FBP_OVERLAP_SCREENED);
    fbp_segment *seg = reinterpret_cast<fbp_segment *>(0x1);
    CHECK(fbp_segment_create(0, &c, &seg) == FBP_ENODEV && seg == nullptr);
    fbp_server *srv = reinterpret_cast<fbp_server *>(0x1);
    CHECK(fbp_server_create(0, 0, 4, &srv) == FBP_ENODEV && srv == nullptr);
    CHECK(fbp_server_create(0, 0, 0, &srv) == FBP_EINVAL);
    CHECK(fbp_segment_run(nullptr, nullptr) == FBP_EINVAL);
    CHECK(fbp_segment_get_spikes(nullptr, nullptr, nullptr) == FBP_EINVAL);
    CHECK(fbp_server_submit(nullptr, &c, nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr) == FBP_EINVAL);
    CHECK(fbp_server_wait(nullptr, 1, 0, nullptr) == FBP_EINVAL);
    fbp_segment_destroy(nullptr);
    fbp_server_destroy(nullptr);
    std::printf("no device: ok\n");
}

int main()
{
    CHECK(fbp_version() > 0);
    sizing();
    host_side_work();
    int n = 0;
    if (fbp_device_count(&n) == FBP_OK)
        std::printf("no device: skipped (%d devices visible)\n", n);
    else
        no_device();
    std::printf("all host runtime checks passed\n");
    return 0;
}
