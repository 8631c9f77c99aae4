// Host-only checks of the segment server's Scheduler (server_core.h) with a
// mock device, built with -fsanitize=thread and with
// -fsanitize=address,undefined by tests/native/Makefile and run by
// tests/test_sanitizers.py.  Every job's "passes" complete after a short
// random delay; the mock counts live jobs and their bytes so the checks can
// see what the scheduler admitted.
#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>

#include "../../spikesorting_fullpursuit_amd/csrc/server_core.h"

using clk = std::chrono::steady_clock;

#define CHECK(cond)                                                                   \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            std::exit(1);                                                             \
        }                                                                             \
    } while (0)

struct MockRequest {
    uint32_t passes = 1;       // passes until convergence
    uint32_t fail_at = 0;      // pass whose step fails (0: none)
    bool fail_create = false;
    bool capped = false;       // the last pass ends with FBP_EPASSES
    uint64_t bytes = 1000;     // held from creation
    uint64_t grow = 0;         // added by every pass
    uint32_t pass_us = 200;    // mean pass duration
};

struct MockJob {
    MockRequest r;
    uint32_t done = 0;
    clk::time_point ready_at;
    std::atomic<uint64_t> bytes{0};
    std::mt19937 rng;
};

struct MockBackend {
    using Request = MockRequest;
    using Job = MockJob;

    std::mutex mu;
    int live = 0, peak_live = 0;
    uint64_t live_bytes = 0, peak_live_bytes = 0;   // bytes at creation of the live jobs
    std::atomic<uint64_t> seed{1};

    void schedule(Job *j)
    {
        std::uniform_int_distribution<uint32_t> d(j->r.pass_us / 2, j->r.pass_us * 3 / 2 + 1);
        j->ready_at = clk::now() + std::chrono::microseconds(d(j->rng));
    }

    int create(const Request &r, Job **out, std::string *err)
    {
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        if (r.fail_create) {
            *err = "mock: create failed";
            return FBP_ENOMEM;
        }
        Job *j = new Job();
        j->r = r;
        j->bytes = r.bytes;
        j->rng.seed((uint32_t)seed.fetch_add(1));
        schedule(j);
        std::lock_guard<std::mutex> g(mu);
        live++;
        peak_live = std::max(peak_live, live);
        live_bytes += r.bytes;
        peak_live_bytes = std::max(peak_live_bytes, live_bytes);
        *out = j;
        return FBP_OK;
    }

    int ready(Job *j, std::string *) { return clk::now() >= j->ready_at ? 1 : 0; }

    int step(Job *j, int *more, std::string *err)
    {
        *more = 0;
        j->done++;
        if (j->r.fail_at && j->done == j->r.fail_at) {
            *err = "mock: pass failed";
            return FBP_EHIP;
        }
        if (j->done >= j->r.passes) return j->r.capped ? FBP_EPASSES : FBP_OK;
        j->bytes += j->r.grow;
        schedule(j);
        *more = 1;
        return FBP_OK;
    }

    uint64_t bytes(const Job *j) { return j->bytes.load(); }

    void destroy(Job *j)
    {
        std::lock_guard<std::mutex> g(mu);
        live--;
        live_bytes -= j->r.bytes;
        delete j;
    }
};

using Sched = fbp::Scheduler<MockBackend>;

static int wait_state(Sched &s, uint64_t t, Sched::Info *info)
{
    const int rc = s.wait(t, -1, info);
    CHECK(rc == FBP_OK);
    return info->state;
}

// Many submitters, a byte budget that admits a few jobs at a time, a slot
// limit, results taken in completion order (wait_any) and released: every
// job completes with its own pass count, admission never exceeds the slots or
// the budget, and admission is in submission order.
static void test_admission()
{
    MockBackend be;
    const uint64_t budget = 5000;
    std::mutex mu;
    std::map<uint64_t, uint32_t> passes;
    {
        Sched s(be, budget, 3);
        std::vector<std::thread> th;
        for (int k = 0; k < 4; k++)
            th.emplace_back([&, k] {
                std::mt19937 rng(k + 11);
                for (int i = 0; i < 40; i++) {
                    MockRequest r;
                    r.passes = 1 + rng() % 6;
                    r.bytes = 500 + rng() % 2000;
                    uint64_t t = 0;
                    std::string err;
                    {
                        // the map entry exists before the job can end
                        std::lock_guard<std::mutex> g(mu);
                        CHECK(s.submit(r, r.bytes, &t, &err) == FBP_OK);
                        passes[t] = r.passes;
                    }
                    if (i % 7 == 3) std::this_thread::sleep_for(std::chrono::microseconds(300));
                }
            });
        std::vector<std::pair<double, uint64_t>> admits;
        double last_finish = 0;
        for (int n = 0; n < 160; n++) {
            uint64_t t = 0;
            Sched::Info info;
            CHECK(s.wait_any(-1, &t, &info) == FBP_OK);
            CHECK(info.state == FBP_JOB_DONE && info.rc == FBP_OK);
            {
                std::lock_guard<std::mutex> g(mu);
                CHECK(info.passes == passes.at(t));
            }
            CHECK(info.t_submit <= info.t_admit && info.t_admit <= info.t_start && info.t_start <= info.t_finish);
            CHECK(info.t_finish >= last_finish);   // completion order
            last_finish = info.t_finish;
            admits.push_back({info.t_admit, t});
            CHECK(s.release(t) == FBP_OK);
        }
        for (auto &t : th) t.join();
        uint64_t t = 0;
        Sched::Info info;
        CHECK(s.wait_any(1000, &t, &info) == FBP_ETIMEDOUT);
        std::sort(admits.begin(), admits.end());
        for (size_t i = 1; i < admits.size(); i++) CHECK(admits[i - 1].second < admits[i].second);
        const fbp_server_stats st = s.stats();
        CHECK(st.submitted == 160 && st.completed == 160 && st.failed == 0 && st.cancelled == 0);
        CHECK(st.peak_running <= 3 && st.peak_held_bytes <= budget);
        CHECK(st.queued == 0 && st.running == 0 && st.held_bytes == 0);
    }
    CHECK(be.live == 0);
    // jobs alive at once: at most 3 staged or running plus finished ones not
    // yet released, all within the budget
    std::printf("peak live %d, bytes %llu\n", be.peak_live, (unsigned long long)be.peak_live_bytes);
    CHECK(be.peak_live_bytes <= budget);
    std::printf("admission: ok (peak %d jobs, %llu bytes)\n", be.peak_live, (unsigned long long)be.peak_live_bytes);
}

// Failures of creation and of a pass reach the caller with their message;
// the failed job's memory is freed at once; capped runs are done.
static void test_errors()
{
    MockBackend be;
    {
        Sched s(be, 1 << 20, 2);
        MockRequest bad_create, bad_pass, capped, big;
        bad_create.fail_create = true;
        bad_pass.passes = 5;
        bad_pass.fail_at = 3;
        capped.passes = 2;
        capped.capped = true;
        big.bytes = 2 << 20;
        uint64_t t1, t2, t3, t4;
        std::string err;
        CHECK(s.submit(bad_create, 1000, &t1, &err) == FBP_OK);
        CHECK(s.submit(bad_pass, 1000, &t2, &err) == FBP_OK);
        CHECK(s.submit(capped, 1000, &t3, &err) == FBP_OK);
        CHECK(s.submit(big, big.bytes, &t4, &err) == FBP_ENOMEM);
        CHECK(err.find("budget") != std::string::npos);
        Sched::Info info;
        CHECK(wait_state(s, t1, &info) == FBP_JOB_FAILED);
        CHECK(info.rc == FBP_ENOMEM && info.err == "mock: create failed");
        CHECK(wait_state(s, t2, &info) == FBP_JOB_FAILED);
        CHECK(info.rc == FBP_EHIP && info.err == "mock: pass failed" && info.passes == 3);
        CHECK(wait_state(s, t3, &info) == FBP_JOB_DONE);
        CHECK(info.rc == FBP_EPASSES);
        MockJob *job = nullptr;
        CHECK(s.result(t2, &job) == FBP_ESTATE);
        CHECK(s.result(t3, &job) == FBP_OK && job && job->done == 2);
        {
            std::lock_guard<std::mutex> g(be.mu);
            CHECK(be.live == 1);   // only the done job holds memory
        }
        CHECK(s.release(t1) == FBP_OK && s.release(t2) == FBP_OK && s.release(t3) == FBP_OK);
        CHECK(s.release(t3) == FBP_EINVAL);
        CHECK(s.info(t3, &info) == FBP_EINVAL);
        const fbp_server_stats st = s.stats();
        CHECK(st.failed == 2 && st.completed == 1);
    }
    CHECK(be.live == 0);
    std::printf("errors: ok\n");
}

// Cancellation of queued and running jobs, wait time-outs, and jobs whose
// bytes grow past their reservation holding back admission.
static void test_cancel_and_growth()
{
    MockBackend be;
    {
        Sched s(be, 10000, 1);
        MockRequest long_job, queued;
        long_job.passes = 1000000;
        long_job.pass_us = 100;
        queued.passes = 1;
        uint64_t tl, tq;
        std::string err;
        CHECK(s.submit(long_job, 1000, &tl, &err) == FBP_OK);
        CHECK(s.submit(queued, 1000, &tq, &err) == FBP_OK);
        Sched::Info info;
        CHECK(s.wait(tl, 2000, &info) == FBP_ETIMEDOUT);
        CHECK(info.state == FBP_JOB_RUNNING || info.state == FBP_JOB_QUEUED);
        CHECK(s.info(tq, &info) == FBP_OK && info.state == FBP_JOB_QUEUED);   // one slot
        CHECK(s.cancel(tq) == FBP_OK);
        CHECK(s.info(tq, &info) == FBP_OK && info.state == FBP_JOB_CANCELLED && info.t_admit < 0);
        CHECK(s.cancel(tl) == FBP_OK);
        CHECK(wait_state(s, tl, &info) == FBP_JOB_CANCELLED);
        CHECK(s.cancel(tl) == FBP_ESTATE);
        CHECK(s.release(tl) == FBP_OK && s.release(tq) == FBP_OK);
        {
            std::lock_guard<std::mutex> g(be.mu);
            CHECK(be.live == 0);
        }
        // a job that grows to 9000 bytes keeps a second 2000-byte job out
        MockRequest grower, other;
        grower.passes = 40;
        grower.bytes = 1000;
        grower.grow = 250;   // 1000 + 39 * 250 = 10750 > budget
        grower.pass_us = 300;
        other.passes = 1;
        other.bytes = 2000;
        uint64_t tg, to;
        CHECK(s.submit(grower, 1000, &tg, &err) == FBP_OK);
        CHECK(wait_state(s, tg, &info) == FBP_JOB_DONE);
        CHECK(info.bytes == 10750);
        CHECK(s.submit(other, 2000, &to, &err) == FBP_OK);
        CHECK(s.wait(to, 20000, &info) == FBP_ETIMEDOUT && info.state == FBP_JOB_QUEUED);
        CHECK(s.release(tg) == FBP_OK);
        CHECK(wait_state(s, to, &info) == FBP_JOB_DONE);
        CHECK(s.release(to) == FBP_OK);
    }
    CHECK(be.live == 0);
    std::printf("cancel and growth: ok\n");
}

// Destroying the scheduler with queued, running and finished jobs frees all.
static void test_shutdown()
{
    MockBackend be;
    {
        Sched s(be, 1 << 30, 4);
        std::string err;
        for (int i = 0; i < 30; i++) {
            MockRequest r;
            r.passes = i % 3 == 0 ? 1 : 100000;
            uint64_t t;
            CHECK(s.submit(r, 100, &t, &err) == FBP_OK);
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(5));
    }
    CHECK(be.live == 0);
    std::printf("shutdown: ok\n");
}

int main()
{
    test_admission();
    test_errors();
    test_cancel_and_growth();
    test_shutdown();
    std::printf("all scheduler checks passed\n");
    return 0;
}
