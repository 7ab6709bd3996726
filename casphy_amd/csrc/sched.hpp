// sched.hpp -- the request loop's scheduler: the job queue, admission by a
// device-memory budget without head-of-line blocking, continuous batching of
// small jobs, completion bookkeeping and latency statistics.  Plain C++ (no
// HIP): the device work of a group of jobs is a callback, so the CPU tests run
// the scheduler with a stub callback under ThreadSanitizer.
//
// Contract:
//   * a worker takes the first queued job whose footprint fits the budget next
//     to the running jobs and the resident bytes; a job that fits nowhere runs
//     once nothing else runs; once the oldest job has waited `hol_ms` without
//     fitting, later jobs start only if they leave room for it;
//   * a worker that takes a batchable job also takes every queued batchable
//     job of the same batch class that fits, up to `batch_max_jobs` (when
//     batching is on); with two or more workers, worker 0 takes batchable jobs
//     only, so they never wait behind large ones;
//   * the runner fills the jobs' outputs and returns one status per job (0 or
//     a CASPHY_E* code); status, error text and completion time are written by
//     the scheduler under its lock, before waiters are notified, and nothing
//     touches a job after that (a waiter may free it at once);
//   * wait() hands the finished job back to the caller and forgets it.
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host.hpp"

namespace casphy {

using sched_clock = std::chrono::steady_clock;

struct SchedJob {
    u64 id = 0;
    size_t dev_bytes = 0;    // estimated device footprint while it runs
    bool batchable = false;  // may share one launch with other jobs
    u64 batch_class = 0;     // jobs share a launch only within one class
    double cost_ms = 0;      // estimated run time (shortest-first order)
    int status = 1;          // 1 pending, 0 done, < 0 CASPHY_E* code
    std::string err;
    sched_clock::time_point t_submit, t_done;
    bool claimed = false;    // a wait() is under way (scheduler-owned)
    virtual ~SchedJob() = default;
    // Called under the scheduler's lock when the job completes (success or
    // not), before waiters are notified: drop references to shared inputs.
    virtual void on_done() {}
};

struct SchedStats {
    u64 completed = 0;  // successful jobs
    double p50_ms = 0, p99_ms = 0, jobs_per_s = 0;
    u64 launches = 0, batched_jobs = 0;
};

class Scheduler {
public:
    // group: one or more jobs (more only for a batch of one class); returns one
    // status per job, `err` the message for the failed ones.  Runs on a worker
    // thread without the scheduler's lock.
    using Runner = std::function<std::vector<int>(std::vector<SchedJob *> &group, std::string &err)>;
    // Runs first on every worker thread; false ends that worker.
    using WorkerInit = std::function<bool()>;
    struct Options {
        size_t budget = 0;
        unsigned workers = 1;
        bool batching = true;
        double hol_ms = 500.0;
        size_t batch_max_jobs = 1024;
        // a worker takes the admissible job with the least cost_ms minus
        // time waited (short jobs pass long ones; a job's turn comes at the
        // latest once it has waited its own cost longer than the others);
        // false: oldest first
        bool shortest_first = true;
    };

    Scheduler(const Options &o, WorkerInit init, Runner run);
    ~Scheduler();  // stops the workers (queued jobs are not run) and joins them
    Scheduler(const Scheduler &) = delete;
    Scheduler &operator=(const Scheduler &) = delete;

    // Queues a job; returns its id.
    u64 submit(std::unique_ptr<SchedJob> j);
    // Blocks until job `id` has completed and hands it back; nullptr when the
    // id is unknown (never submitted, or already waited for) or another
    // thread is already waiting for it.
    std::unique_ptr<SchedJob> wait(u64 id);
    // A fresh id from the job id sequence (for other objects that share it).
    u64 new_id();
    // Device bytes held outside the jobs (resident inputs), for admission.
    void add_resident(size_t bytes);
    void remove_resident(size_t bytes);
    SchedStats stats();

private:
    void worker_main(bool small_lane);
    std::vector<SchedJob *> pick(bool small_lane, size_t &bytes);  // with m_ held

    Options opt_;
    WorkerInit init_;
    Runner run_;
    std::mutex m_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<SchedJob *> queue_;
    std::unordered_map<u64, std::unique_ptr<SchedJob>> jobs_;
    size_t running_ = 0, resident_ = 0;
    bool stop_ = false;
    u64 next_id_ = 1;
    std::vector<double> lat_ms_;
    bool have_first_ = false;
    sched_clock::time_point t_first_, t_last_;
    u64 launches_ = 0, batched_jobs_ = 0;
    std::vector<std::thread> workers_;
};

}  // namespace casphy
