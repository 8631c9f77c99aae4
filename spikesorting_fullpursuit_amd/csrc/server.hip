// server.hip -- C-ABI of the segment server (include/fbp.h, fbp_server_*):
// the Scheduler of server_core.h over fbp_segment.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/fbp.h"
#include "fbp_internal.h"
#include "server_core.h"

using namespace fbp;

namespace {

struct SegmentRequest {
    fbp_config cfg{};
    const float *templates = nullptr, *bias = nullptr, *thresholds = nullptr, *voltage = nullptr;
    uint64_t row_stride = 0;
};

// Backend of the Scheduler: a job is one fbp_segment driven with
// segment_begin / segment_step, its passes on the segment's own stream.
struct SegmentBackend {
    using Request = SegmentRequest;
    using Job = fbp_segment;
    int device = 0;

    int create(const Request &r, Job **out, std::string *err)
    {
        fbp_segment *g = nullptr;
        int rc = fbp_segment_create(device, &r.cfg, &g);
        if (rc == FBP_OK) rc = fbp_segment_set_templates(g, r.templates, r.bias, r.thresholds);
        if (rc == FBP_OK) rc = fbp_segment_set_voltage(g, r.voltage, r.row_stride);
        if (rc == FBP_OK) rc = segment_begin(g, nullptr);
        if (rc != FBP_OK) {
            *err = fbp_last_error();
            fbp_segment_destroy(g);
            return rc;
        }
        *out = g;
        return FBP_OK;
    }

    int ready(Job *g, std::string *err)
    {
        const hipError_t e = hipEventQuery(segment_pass_event(g));
        if (e == hipSuccess) return 1;
        if (e == hipErrorNotReady) return 0;
        *err = std::string("pass: ") + hipGetErrorString(e);
        return FBP_EHIP;
    }

    int step(Job *g, int *more, std::string *err)
    {
        const int rc = segment_step(g, more);
        if (rc != FBP_OK) *err = fbp_last_error();
        return rc;
    }

    uint64_t bytes(const Job *g) { return segment_device_bytes(g); }

    void destroy(Job *g) { fbp_segment_destroy(g); }
};

void copy_info(const Scheduler<SegmentBackend>::Info &a, fbp_job_info *b)
{
    b->state = a.state;
    b->rc = a.rc;
    b->reserve_bytes = a.reserve;
    b->bytes = a.bytes;
    b->passes = a.passes;
    b->reserved = 0;
    b->t_submit = a.t_submit;
    b->t_admit = a.t_admit;
    b->t_start = a.t_start;
    b->t_finish = a.t_finish;
}

}  // namespace

struct fbp_server {
    SegmentBackend backend;
    Scheduler<SegmentBackend> *sched = nullptr;
};

extern "C" {

int fbp_server_create(int device, uint64_t hbm_budget, uint32_t max_running, fbp_server **out)
{
    if (!out) return set_error(FBP_EINVAL, "NULL argument");
    *out = nullptr;
    if (max_running == 0) return set_error(FBP_EINVAL, "max_running must be at least 1");
    uint64_t total = 0, free_b = 0;
    if (int rc = fbp_device_memory(device, &total, &free_b)) return rc;
    if (hbm_budget == 0) hbm_budget = (uint64_t)(0.85 * (double)free_b);
    fbp_server *srv = new fbp_server();
    srv->backend.device = device;
    srv->sched = new Scheduler<SegmentBackend>(srv->backend, hbm_budget, max_running);
    *out = srv;
    return FBP_OK;
}

void fbp_server_destroy(fbp_server *srv)
{
    if (!srv) return;
    delete srv->sched;
    delete srv;
}

int fbp_server_submit(fbp_server *srv, const fbp_config *cfg, const float *templates, const float *bias,
                      const float *thresholds, const float *voltage, uint64_t row_stride, uint64_t extra_bytes,
                      uint64_t *ticket)
{
    if (!srv || !cfg || !templates || !bias || !thresholds || !voltage || !ticket)
        return set_error(FBP_EINVAL, "NULL argument");
    if (row_stride < cfg->n_samples)
        return set_error(FBP_EINVAL, "row stride %llu below chunk length %u", (unsigned long long)row_stride,
                         cfg->n_samples);
    uint64_t bytes = 0;
    if (int rc = fbp_segment_bytes(cfg, &bytes)) return rc;   // validates cfg
    SegmentRequest r;
    r.cfg = *cfg;
    r.templates = templates;
    r.bias = bias;
    r.thresholds = thresholds;
    r.voltage = voltage;
    r.row_stride = row_stride;
    std::string err;
    const int rc = srv->sched->submit(r, bytes + extra_bytes, ticket, &err);
    if (rc != FBP_OK) return set_error(rc, "%s", err.c_str());
    return FBP_OK;
}

int fbp_server_poll(fbp_server *srv, uint64_t ticket, fbp_job_info *info)
{
    if (!srv || !info) return set_error(FBP_EINVAL, "NULL argument");
    Scheduler<SegmentBackend>::Info a;
    if (srv->sched->info(ticket, &a) != FBP_OK)
        return set_error(FBP_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    copy_info(a, info);
    if (a.state == FBP_JOB_FAILED) set_error(a.rc, "%s", a.err.c_str());
    return FBP_OK;
}

int fbp_server_wait(fbp_server *srv, uint64_t ticket, int64_t timeout_us, fbp_job_info *info)
{
    if (!srv) return set_error(FBP_EINVAL, "NULL argument");
    Scheduler<SegmentBackend>::Info a;
    const int rc = srv->sched->wait(ticket, timeout_us, &a);
    if (rc == FBP_EINVAL) return set_error(FBP_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    if (info) copy_info(a, info);
    if (a.state == FBP_JOB_FAILED) set_error(a.rc, "%s", a.err.c_str());
    if (rc == FBP_ETIMEDOUT) return set_error(FBP_ETIMEDOUT, "job %llu has not ended", (unsigned long long)ticket);
    return FBP_OK;
}

int fbp_server_wait_any(fbp_server *srv, int64_t timeout_us, uint64_t *ticket, fbp_job_info *info)
{
    if (!srv || !ticket) return set_error(FBP_EINVAL, "NULL argument");
    Scheduler<SegmentBackend>::Info a;
    const int rc = srv->sched->wait_any(timeout_us, ticket, &a);
    if (rc == FBP_ETIMEDOUT) return set_error(FBP_ETIMEDOUT, "no job has ended");
    if (info) copy_info(a, info);
    if (a.state == FBP_JOB_FAILED) set_error(a.rc, "%s", a.err.c_str());
    return FBP_OK;
}

int fbp_server_cancel(fbp_server *srv, uint64_t ticket)
{
    if (!srv) return set_error(FBP_EINVAL, "NULL argument");
    const int rc = srv->sched->cancel(ticket);
    if (rc == FBP_EINVAL) return set_error(rc, "unknown ticket %llu", (unsigned long long)ticket);
    if (rc == FBP_ESTATE) return set_error(rc, "job %llu has already ended", (unsigned long long)ticket);
    return FBP_OK;
}

int fbp_server_segment(fbp_server *srv, uint64_t ticket, fbp_segment **seg)
{
    if (!srv || !seg) return set_error(FBP_EINVAL, "NULL argument");
    const int rc = srv->sched->result(ticket, seg);
    if (rc == FBP_EINVAL) return set_error(rc, "unknown ticket %llu", (unsigned long long)ticket);
    if (rc == FBP_ESTATE) return set_error(rc, "job %llu is not done", (unsigned long long)ticket);
    return FBP_OK;
}

int fbp_server_release(fbp_server *srv, uint64_t ticket)
{
    if (!srv) return set_error(FBP_EINVAL, "NULL argument");
    const int rc = srv->sched->release(ticket);
    if (rc == FBP_EINVAL) return set_error(rc, "unknown ticket %llu", (unsigned long long)ticket);
    if (rc == FBP_ESTATE) return set_error(rc, "job %llu has not ended", (unsigned long long)ticket);
    return FBP_OK;
}

int fbp_server_get_stats(fbp_server *srv, fbp_server_stats *stats)
{
    if (!srv || !stats) return set_error(FBP_EINVAL, "NULL argument");
    *stats = srv->sched->stats();
    return FBP_OK;
}

}  // extern "C"
