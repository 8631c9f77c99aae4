// server_core.h -- admission and pass-by-pass driving of many segments on one
// device, independent of HIP.
//
// The reference sorts segments one after another in one process
// (spikesorting_fullpursuit/parallel/spikesorting_parallel.py:1312-1322),
// each through a blocking binary_pursuit() call.  A Scheduler instead keeps
// every admitted segment's state on the device and advances all of them from
// one driver thread: whenever the pass in flight of any segment has
// completed, the driver reads its counters and enqueues its next pass on the
// segment's own stream, so the device always holds one pass of every running
// segment and no host thread blocks on a pass.  A second thread, the stager,
// admits queued requests in arrival order when a running slot is free and the
// request's reservation fits in the byte budget next to what the admitted
// segments hold, then creates the segment and uploads its data while the
// driver keeps the others going.  A job that has finished keeps its device
// memory (its results) until it is released and counts against the budget
// until then; wait_any reports jobs in the order they end, so a consumer that
// releases what it takes keeps admission going.
//
// The device work is behind a Backend (server.hip for HIP; a host-only mock
// in the sanitizer tests), which provides
//   using Request, Job;
//   int  create(const Request &, Job **out, std::string *err);  // allocate,
//        upload and enqueue the first pass; 0 or a negative FBP_E* code
//   int  ready(Job *, std::string *err);   // 1: the pass in flight completed,
//        0: not yet, < 0: error
//   int  step(Job *, int *more, std::string *err);  // consume the completed
//        pass; *more = 1 when the next pass was enqueued; FBP_EPASSES ends a
//        run whose results stay readable
//   uint64_t bytes(const Job *);   // device bytes the job holds now; may be
//        called from any thread while another thread works on the job
//   void destroy(Job *);
// Each Job is used by one thread at a time: the stager while it creates it,
// the driver while it runs, the caller once it is done.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fbp.h"

namespace fbp {

template <class Backend>
class Scheduler {
public:
    using Request = typename Backend::Request;
    using Job = typename Backend::Job;

    struct Info {
        int state = FBP_JOB_QUEUED;
        int rc = FBP_OK;          // result of the run (FBP_EPASSES: capped, results readable)
        uint64_t reserve = 0;     // bytes reserved at admission
        uint64_t bytes = 0;       // bytes the job holds now
        uint32_t passes = 0;      // passes completed
        double t_submit = 0, t_admit = -1, t_start = -1, t_finish = -1;   // seconds since the scheduler started
        std::string err;
    };

    // budget: bytes the admitted jobs may reserve together; max_running:
    // jobs staged or running at once; poll_us: how long the driver sleeps
    // when no pass has completed (new work wakes it at once).
    Scheduler(Backend &be, uint64_t budget, uint32_t max_running, uint32_t poll_us = 20)
        : be_(be), budget_(budget), max_running_(max_running ? max_running : 1), poll_(poll_us),
          t0_(std::chrono::steady_clock::now())
    {
        stager_ = std::thread([this] { stage_loop(); });
        driver_ = std::thread([this] { drive_loop(); });
    }

    Scheduler(const Scheduler &) = delete;
    Scheduler &operator=(const Scheduler &) = delete;

    // Queued jobs are cancelled, running ones stopped at their next pass
    // boundary; every job still held is destroyed.
    ~Scheduler()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            for (Entry *e : queue_) finish_locked(e, FBP_JOB_CANCELLED, FBP_OK, "");
            queue_.clear();
            for (Entry *e : running_) e->cancel = true;
        }
        cv_stage_.notify_all();
        cv_drive_.notify_all();
        stager_.join();
        driver_.join();
        for (auto &kv : jobs_)
            if (kv.second->job) be_.destroy(kv.second->job);
    }

    // Queue a request; *ticket identifies it.  A reservation above the whole
    // budget can never be admitted: FBP_ENOMEM.
    int submit(const Request &req, uint64_t reserve, uint64_t *ticket, std::string *err)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) { *err = "server is shutting down"; return FBP_ESTATE; }
        if (reserve > budget_) {
            *err = "request reserves " + std::to_string(reserve) + " bytes, above the server's budget of " +
                   std::to_string(budget_);
            return FBP_ENOMEM;
        }
        std::unique_ptr<Entry> e(new Entry());
        e->ticket = ++last_ticket_;
        e->req = req;
        e->info.reserve = reserve;
        e->info.t_submit = now();
        queue_.push_back(e.get());
        *ticket = e->ticket;
        jobs_[e->ticket] = std::move(e);
        submitted_++;
        cv_stage_.notify_one();
        return FBP_OK;
    }

    // State and times of a job.
    int info(uint64_t ticket, Info *out)
    {
        std::lock_guard<std::mutex> lk(mu_);
        Entry *e = find_locked(ticket);
        if (!e) return FBP_EINVAL;
        fill_info_locked(e, out);
        return FBP_OK;
    }

    // Wait until the job has ended (done, failed or cancelled) or timeout_us
    // has passed (< 0: no limit).  FBP_ETIMEDOUT when it has not ended.
    int wait(uint64_t ticket, int64_t timeout_us, Info *out)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto ended = [&] {
            Entry *e = find_locked(ticket);
            return !e || e->info.state >= FBP_JOB_DONE;
        };
        if (timeout_us < 0)
            cv_done_.wait(lk, ended);
        else
            cv_done_.wait_until(lk, deadline(timeout_us), ended);
        Entry *e = find_locked(ticket);
        if (!e) return FBP_EINVAL;
        fill_info_locked(e, out);
        return e->info.state >= FBP_JOB_DONE ? FBP_OK : FBP_ETIMEDOUT;
    }

    // Wait until some job has ended that wait_any has not reported yet, or
    // timeout_us has passed (< 0: no limit); jobs are reported once each, in
    // the order they ended.  FBP_ETIMEDOUT when none has.
    int wait_any(int64_t timeout_us, uint64_t *ticket, Info *out)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto any = [&] { return !ended_.empty(); };
        if (timeout_us < 0)
            cv_done_.wait(lk, any);
        else
            cv_done_.wait_until(lk, deadline(timeout_us), any);
        while (!ended_.empty()) {
            const uint64_t t = ended_.front();
            ended_.pop_front();
            Entry *e = find_locked(t);
            if (!e) continue;   // released meanwhile
            *ticket = t;
            fill_info_locked(e, out);
            return FBP_OK;
        }
        return FBP_ETIMEDOUT;
    }

    // Cancel a job: a queued one at once, a staged or running one at its next
    // pass boundary (its device memory is then freed).  A job that has ended
    // is left as it is; FBP_ESTATE tells the caller so.
    int cancel(uint64_t ticket)
    {
        std::lock_guard<std::mutex> lk(mu_);
        Entry *e = find_locked(ticket);
        if (!e) return FBP_EINVAL;
        if (e->info.state >= FBP_JOB_DONE) return FBP_ESTATE;
        if (e->info.state == FBP_JOB_QUEUED && !e->staging) {
            queue_.erase(std::find(queue_.begin(), queue_.end(), e));
            finish_locked(e, FBP_JOB_CANCELLED, FBP_OK, "");
            cv_stage_.notify_one();
            return FBP_OK;
        }
        e->cancel = true;
        cv_drive_.notify_one();
        return FBP_OK;
    }

    // The finished job's device state, for reading results; it stays valid
    // until release().
    int result(uint64_t ticket, Job **job)
    {
        std::lock_guard<std::mutex> lk(mu_);
        Entry *e = find_locked(ticket);
        if (!e) return FBP_EINVAL;
        if (e->info.state != FBP_JOB_DONE) return FBP_ESTATE;
        *job = e->job;
        return FBP_OK;
    }

    // Forget an ended job and free what it holds.
    int release(uint64_t ticket)
    {
        std::unique_ptr<Entry> keep;
        Job *job = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = jobs_.find(ticket);
            if (it == jobs_.end()) return FBP_EINVAL;
            if (it->second->info.state < FBP_JOB_DONE) return FBP_ESTATE;
            keep = std::move(it->second);
            jobs_.erase(it);
            // admission keeps counting the job's bytes until they are freed
            job = detach_locked(keep.get());
        }
        if (job) be_.destroy(job);
        {
            std::lock_guard<std::mutex> lk(mu_);
            held_.erase(ticket);
        }
        cv_stage_.notify_one();
        cv_done_.notify_all();
        return FBP_OK;
    }

    fbp_server_stats stats()
    {
        std::lock_guard<std::mutex> lk(mu_);
        fbp_server_stats s{};
        s.submitted = submitted_;
        s.completed = completed_;
        s.failed = failed_;
        s.cancelled = cancelled_;
        s.queued = (uint32_t)queue_.size();
        s.running = n_running_;
        s.peak_running = peak_running_;
        s.held_bytes = held_bytes_locked();
        s.peak_held_bytes = peak_held_;
        s.budget_bytes = budget_;
        s.pass_steps = pass_steps_;
        s.driver_busy_s = driver_busy_s_;
        s.uptime_s = now();
        return s;
    }

private:
    struct Entry {
        uint64_t ticket = 0;
        Request req{};
        Job *job = nullptr;
        Info info;
        bool staging = false;             // taken by the stager
        std::atomic<bool> cancel{false};
    };

    // Deadline of a timed wait.  On the system clock: condition waits on it
    // use pthread_cond_timedwait, which thread sanitizers follow (waits on
    // the steady clock go through pthread_cond_clockwait, which GCC 11's does
    // not).
    static std::chrono::system_clock::time_point deadline(int64_t us)
    {
        return std::chrono::system_clock::now() + std::chrono::microseconds(us);
    }

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

    Entry *find_locked(uint64_t ticket)
    {
        auto it = jobs_.find(ticket);
        return it == jobs_.end() ? nullptr : it->second.get();
    }

    void fill_info_locked(Entry *e, Info *out)
    {
        *out = e->info;
        out->bytes = e->job ? be_.bytes(e->job) : 0;
    }

    // Bytes the jobs holding device memory count for: each its reservation or
    // what it holds now, whichever is larger (the spike list and clips grow).
    uint64_t held_bytes_locked()
    {
        uint64_t b = 0;
        for (auto &kv : held_) {
            Entry *e = kv.second;
            const uint64_t now_b = e->job ? be_.bytes(e->job) : 0;
            b += std::max(e->info.reserve, now_b);
        }
        return b;
    }

    // Take the job out of its entry before destroying it, so that no other
    // thread reads its byte count meanwhile; the entry's reservation keeps
    // what the job held in the budget until held_ forgets the entry.
    Job *detach_locked(Entry *e)
    {
        Job *job = e->job;
        if (job) e->info.reserve = std::max(e->info.reserve, be_.bytes(job));
        e->job = nullptr;
        return job;
    }

    // A running job has ended: cancelled, failed or done (free_job false:
    // its results stay on the device until release()).
    void retire(Entry *e, int state, int rc, const std::string &err, bool free_job,
                std::chrono::steady_clock::time_point t_a)
    {
        Job *dead = nullptr;
        if (free_job) {
            std::lock_guard<std::mutex> g(mu_);
            dead = detach_locked(e);
        }
        if (dead) be_.destroy(dead);
        std::lock_guard<std::mutex> g(mu_);
        driver_busy_s_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_a).count();
        if (free_job) held_.erase(e->ticket);
        running_.erase(std::find(running_.begin(), running_.end(), e));
        n_running_--;
        finish_locked(e, state, rc, err);
        cv_stage_.notify_one();
    }

    void finish_locked(Entry *e, int state, int rc, const std::string &err)
    {
        e->info.state = state;
        e->info.rc = rc;
        e->info.err = err;
        e->info.t_finish = now();
        ended_.push_back(e->ticket);
        if (state == FBP_JOB_DONE) completed_++;
        else if (state == FBP_JOB_FAILED) failed_++;
        else cancelled_++;
        cv_done_.notify_all();
    }

    void stage_loop()
    {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            if (stop_ && queue_.empty()) return;
            Entry *e = queue_.empty() ? nullptr : queue_.front();
            if (!e || n_running_ >= max_running_ || held_bytes_locked() + e->info.reserve > budget_) {
                cv_stage_.wait(lk);
                continue;
            }
            queue_.pop_front();
            e->staging = true;
            e->info.t_admit = now();
            held_[e->ticket] = e;
            n_running_++;
            peak_running_ = std::max(peak_running_, n_running_);
            peak_held_ = std::max(peak_held_, held_bytes_locked());
            lk.unlock();
            Job *job = nullptr;
            std::string err;
            const int rc = be_.create(e->req, &job, &err);
            const bool cancelled = e->cancel.load() || stop_.load();
            if (rc == FBP_OK && cancelled) {
                be_.destroy(job);
                job = nullptr;
            }
            lk.lock();
            e->staging = false;
            if (rc != FBP_OK || cancelled) {
                held_.erase(e->ticket);
                n_running_--;
                finish_locked(e, rc != FBP_OK ? FBP_JOB_FAILED : FBP_JOB_CANCELLED, rc, err);
                continue;
            }
            e->job = job;
            e->info.state = FBP_JOB_RUNNING;
            e->info.t_start = now();
            running_.push_back(e);
            peak_held_ = std::max(peak_held_, held_bytes_locked());
            cv_drive_.notify_one();
        }
    }

    void drive_loop()
    {
        std::vector<Entry *> snap;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            if (running_.empty()) {
                // the stager has ended once it has nothing left to admit
                if (stop_ && queue_.empty() && n_running_ == 0) return;
                cv_drive_.wait_until(lk, deadline(5000));
                continue;
            }
            snap.assign(running_.begin(), running_.end());
            lk.unlock();
            bool progressed = false;
            for (Entry *e : snap) {
                std::string err;
                const auto t_a = std::chrono::steady_clock::now();
                if (e->cancel.load() || stop_.load()) {
                    retire(e, FBP_JOB_CANCELLED, FBP_OK, "", true, t_a);
                    progressed = true;
                    continue;
                }
                const int r = be_.ready(e->job, &err);
                if (r == 0) continue;
                progressed = true;
                int more = 0;
                const int rc = r < 0 ? r : be_.step(e->job, &more, &err);
                if (r > 0) {
                    std::lock_guard<std::mutex> g(mu_);
                    e->info.passes++;
                    pass_steps_++;
                    peak_held_ = std::max(peak_held_, held_bytes_locked());
                }
                if (rc == FBP_OK && more) {
                    std::lock_guard<std::mutex> g(mu_);
                    driver_busy_s_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_a).count();
                    continue;
                }
                if (rc == FBP_OK || rc == FBP_EPASSES)
                    retire(e, FBP_JOB_DONE, rc, err, false, t_a);
                else
                    retire(e, FBP_JOB_FAILED, rc, err, true, t_a);
            }
            lk.lock();
            if (!progressed) cv_drive_.wait_until(lk, deadline(poll_));
        }
    }

    Backend &be_;
    const uint64_t budget_;
    const uint32_t max_running_;
    const uint32_t poll_;
    const std::chrono::steady_clock::time_point t0_;

    std::mutex mu_;
    std::condition_variable cv_stage_, cv_drive_, cv_done_;
    std::map<uint64_t, std::unique_ptr<Entry>> jobs_;
    std::deque<Entry *> queue_;                 // submitted, not yet admitted
    std::vector<Entry *> running_;              // the driver's jobs
    std::map<uint64_t, Entry *> held_;          // jobs holding device memory
    std::deque<uint64_t> ended_;                // ended jobs not yet reported by wait_any
    uint64_t last_ticket_ = 0;
    uint32_t n_running_ = 0, peak_running_ = 0; // staged or running
    uint64_t peak_held_ = 0;
    uint64_t submitted_ = 0, completed_ = 0, failed_ = 0, cancelled_ = 0, pass_steps_ = 0;
    double driver_busy_s_ = 0;
    std::atomic<bool> stop_{false};

    std::thread stager_, driver_;
};

}  // namespace fbp
