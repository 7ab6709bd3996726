// Drives the request loop's scheduler (casphy_amd/csrc/sched.cpp) with a stub
// runner in place of the device work, from many submitting and waiting
// threads; built with ThreadSanitizer by tests/test_sched.py.  Checks:
//   * every job completes once, with its own result or its own failure code;
//   * the running footprint never exceeds the budget, except for a job that
//     fits nowhere running alone;
//   * a group of more than one job holds batchable jobs of one class only, and
//     no group exceeds the batch limit; with batching off every group is one job;
//   * a job larger than the budget still runs (head-of-line reservation), and
//     resident bytes count against the budget;
//   * queued jobs run shortest first (least cost minus time waited), and a
//     long job that has waited longer than its cost runs before a new short
//     one; with shortest_first off, oldest first;
//   * unknown ids and a second waiter on one job are refused;
//   * the statistics count the successful jobs.
// Exit status 0 on success; a message and status 1 on the first failure.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../casphy_amd/csrc/sched.hpp"

using namespace casphy;

namespace {

std::atomic<int> g_fail{0};

void check(bool c, const char *what)
{
    if (!c && g_fail.exchange(1) == 0) std::fprintf(stderr, "sched_check: %s\n", what);
}

struct StubJob : SchedJob {
    u64 input = 0, output = 0;
    bool poison = false;          // the runner fails this job
    std::shared_ptr<int> shared;  // released by on_done
    void on_done() override { shared.reset(); }
};

struct Stub {
    size_t budget = 0, resident = 0, batch_max = 0;
    bool batching = true;
    std::atomic<size_t> running{0};
    std::atomic<u64> groups{0}, jobs{0}, solo_over{0};
    std::mutex mu;

    std::vector<int> run(std::vector<SchedJob *> &g, std::string &err)
    {
        size_t bytes = 0;
        for (SchedJob *j : g) bytes += j->dev_bytes;
        const size_t now = running.fetch_add(bytes) + bytes;
        if (now + resident > budget) {
            // allowed only for a group that runs alone
            check(now == bytes && g.size() == 1, "running footprint over the budget");
            solo_over++;
        }
        check(g.size() <= batch_max, "group larger than the batch limit");
        if (g.size() > 1) {
            check(batching, "a batch formed with batching off");
            for (SchedJob *j : g)
                check(j->batchable && j->batch_class == g[0]->batch_class, "batch mixes classes or large jobs");
        }
        std::vector<int> rc(g.size(), 0);
        bool any_poison = false;
        for (size_t k = 0; k < g.size(); ++k) {
            StubJob *j = static_cast<StubJob *>(g[k]);
            j->output = j->input * 2654435761ull + 17;
            if (j->poison) {
                rc[k] = CASPHY_EARG;
                any_poison = true;
            }
        }
        if (any_poison) err = "poisoned job";
        std::this_thread::sleep_for(std::chrono::microseconds(50 + 10 * (g.size() % 7)));
        running.fetch_sub(bytes);
        groups++;
        jobs += g.size();
        return rc;
    }
};

int run_mix(unsigned workers, bool batching, size_t budget, size_t resident, unsigned threads, unsigned per_thread)
{
    Stub stub;
    stub.budget = budget;
    stub.resident = resident;
    stub.batch_max = 64;
    stub.batching = batching;
    Scheduler::Options o;
    o.budget = budget;
    o.workers = workers;
    o.batching = batching;
    o.hol_ms = 5.0;
    o.batch_max_jobs = stub.batch_max;
    std::atomic<int> inits{0};
    auto sched = std::make_unique<Scheduler>(
        o, [&] { inits++; return true; },
        [&](std::vector<SchedJob *> &g, std::string &err) { return stub.run(g, err); });
    sched->add_resident(resident);
    const u64 new_id = sched->new_id();
    check(sched->wait(new_id) == nullptr, "wait on a non-job id");
    check(sched->wait(1ull << 60) == nullptr, "wait on an unknown id");

    auto shared = std::make_shared<int>(7);
    std::atomic<u64> ok{0}, failed{0};
    std::vector<std::thread> th;
    for (unsigned t = 0; t < threads; ++t) {
        th.emplace_back([&, t] {
            std::mt19937_64 rng(1234 + t);
            std::vector<std::pair<u64, u64>> mine;  // (id, input)
            std::vector<bool> pois;
            for (unsigned i = 0; i < per_thread; ++i) {
                auto j = std::make_unique<StubJob>();
                j->input = rng();
                const unsigned r = (unsigned)(rng() % 100);
                j->batchable = r < 60;
                j->batch_class = rng() % 3;
                // small jobs mostly; some large, a few larger than the budget
                j->dev_bytes = r < 60 ? 1 + rng() % (budget / 64 + 1)
                                      : (r < 97 ? rng() % (budget / 2 + 1) : budget + 1 + rng() % budget);
                j->poison = rng() % 23 == 0;
                j->shared = shared;
                pois.push_back(j->poison);
                const u64 in = j->input;
                mine.emplace_back(sched->submit(std::move(j)), in);
                if (i % 5 == 4) {  // wait for the oldest outstanding one now and then
                    auto [id, input] = mine.front();
                    const bool p = pois.front();
                    mine.erase(mine.begin());
                    pois.erase(pois.begin());
                    auto done = sched->wait(id);
                    check(done != nullptr, "wait returned no job");
                    if (!done) continue;
                    auto *sj = static_cast<StubJob *>(done.get());
                    check(sj->shared == nullptr, "on_done did not run");
                    if (p) {
                        check(sj->status == CASPHY_EARG && sj->err == "poisoned job", "failure not reported");
                        failed++;
                    } else {
                        check(sj->status == 0 && sj->output == input * 2654435761ull + 17, "wrong result");
                        check(sj->t_done >= sj->t_submit, "completion before submission");
                        ok++;
                    }
                }
            }
            for (size_t k = 0; k < mine.size(); ++k) {
                auto done = sched->wait(mine[k].first);
                check(done != nullptr, "wait returned no job");
                if (!done) continue;
                auto *sj = static_cast<StubJob *>(done.get());
                if (pois[k]) {
                    check(sj->status == CASPHY_EARG, "failure not reported");
                    failed++;
                } else {
                    check(sj->status == 0 && sj->output == mine[k].second * 2654435761ull + 17, "wrong result");
                    ok++;
                }
                check(sched->wait(mine[k].first) == nullptr, "a job handed back twice");
            }
        });
    }
    for (auto &t : th) t.join();
    const SchedStats st = sched->stats();
    check(st.completed == ok.load(), "statistics count the successful jobs");
    check(ok + failed == (u64)threads * per_thread, "every job completed once");
    check(stub.jobs.load() == (u64)threads * per_thread, "the runner saw every job once");
    check(st.launches == stub.groups.load(), "launch count");
    check(st.p50_ms <= st.p99_ms, "percentiles ordered");
    sched->remove_resident(resident);
    sched.reset();
    check(inits.load() == (int)workers, "every worker ran its init");
    check(shared.use_count() == 1, "a job kept its shared input");
    std::printf("workers=%u batching=%d jobs=%llu groups=%llu over-budget solo=%llu\n", workers, (int)batching,
                (unsigned long long)stub.jobs.load(), (unsigned long long)stub.groups.load(),
                (unsigned long long)stub.solo_over.load());
    return g_fail.load();
}

// Of two threads waiting for one job that is still running, one is refused at
// once and the other gets the job when it completes.
int second_waiter()
{
    std::atomic<bool> release{false};
    Scheduler::Options o;
    o.budget = 100;
    o.workers = 1;
    Scheduler sched(o, nullptr, [&](std::vector<SchedJob *> &g, std::string &) {
        while (!release.load()) std::this_thread::sleep_for(std::chrono::milliseconds(1));
        return std::vector<int>(g.size(), 0);
    });
    auto j = std::make_unique<StubJob>();
    j->dev_bytes = 10;
    const u64 id = sched.submit(std::move(j));
    std::unique_ptr<SchedJob> got[2];
    std::atomic<int> returned{0};
    std::thread w0([&] { got[0] = sched.wait(id); returned++; });
    std::thread w1([&] { got[1] = sched.wait(id); returned++; });
    for (int k = 0; k < 5000 && returned.load() == 0; ++k) std::this_thread::sleep_for(std::chrono::milliseconds(1));
    check(returned.load() == 1, "one of two waiters returns before the job completes");
    release = true;
    w0.join();
    w1.join();
    check((got[0] == nullptr) != (got[1] == nullptr), "exactly one waiter gets the job");
    const SchedJob *d = got[0] ? got[0].get() : got[1].get();
    check(d && d->status == 0, "the job completed");
    return g_fail.load();
}

// A runner that throws or returns the wrong number of statuses fails the group.
int bad_runner()
{
    Scheduler::Options o;
    o.budget = 100;
    o.workers = 2;
    std::atomic<int> calls{0};
    Scheduler sched(o, nullptr, [&](std::vector<SchedJob *> &g, std::string &) -> std::vector<int> {
        if (calls++ % 2 == 0) throw Error(CASPHY_ENOMEM, "out of device memory");
        return std::vector<int>(g.size() + 1, 0);
    });
    std::vector<u64> ids;
    for (int k = 0; k < 6; ++k) {
        auto j = std::make_unique<StubJob>();
        j->dev_bytes = 10;
        ids.push_back(sched.submit(std::move(j)));
    }
    for (u64 id : ids) {
        auto d = sched.wait(id);
        check(d && (d->status == CASPHY_ENOMEM || d->status == CASPHY_EINTERNAL), "bad runner not reported");
        check(d && !d->err.empty(), "bad runner without a message");
    }
    return g_fail.load();
}

// One worker held by a first job; jobs queued behind it run in the order the
// options ask for once it is released.
std::vector<int> run_order(const std::vector<double> &costs, const std::vector<int> &delay_ms, bool shortest_first,
                           int hold_ms)
{
    std::atomic<bool> release{false};
    std::mutex mu;
    std::vector<int> order;
    Scheduler::Options o;
    o.budget = 1000;
    o.workers = 1;
    o.batching = false;
    o.shortest_first = shortest_first;
    Scheduler sched(o, nullptr, [&](std::vector<SchedJob *> &g, std::string &) {
        const StubJob *j = static_cast<const StubJob *>(g[0]);
        if (j->input == 1000) {
            while (!release.load()) std::this_thread::sleep_for(std::chrono::milliseconds(1));
        } else {
            std::lock_guard<std::mutex> lk(mu);
            order.push_back((int)j->input);
        }
        return std::vector<int>(g.size(), 0);
    });
    auto blocker = std::make_unique<StubJob>();
    blocker->input = 1000;
    blocker->dev_bytes = 1;
    std::vector<u64> ids{sched.submit(std::move(blocker))};
    std::this_thread::sleep_for(std::chrono::milliseconds(20));  // the worker holds the blocker
    const auto t0 = std::chrono::steady_clock::now();
    for (size_t k = 0; k < costs.size(); ++k) {
        std::this_thread::sleep_until(t0 + std::chrono::milliseconds(delay_ms[k]));
        auto j = std::make_unique<StubJob>();
        j->input = k;
        j->dev_bytes = 1;
        j->cost_ms = costs[k];
        ids.push_back(sched.submit(std::move(j)));
    }
    std::this_thread::sleep_until(t0 + std::chrono::milliseconds(hold_ms));
    release = true;
    for (u64 id : ids) check(sched.wait(id) != nullptr, "an ordered job did not complete");
    return order;
}

int ordering()
{
    // all queued within a few milliseconds (cost plus queueing time 60, 21,
    // 42, 3: at least 18 ms apart): shortest first
    check(run_order({60, 20, 40, 0}, {0, 1, 2, 3}, true, 80) == std::vector<int>({3, 1, 2, 0}),
          "queued jobs did not run shortest first");
    check(run_order({60, 20, 40, 0}, {0, 1, 2, 3}, false, 80) == std::vector<int>({0, 1, 2, 3}),
          "oldest-first order not kept with shortest_first off");
    // a 100 ms job queued at 0 ms has waited 300 ms at the release (key -200);
    // a 10 ms job queued at 250 ms has waited 50 ms (key -40): the long one first
    check(run_order({100, 10}, {0, 250}, true, 300) == std::vector<int>({0, 1}),
          "a long job that waited longer than its cost was passed");
    return g_fail.load();
}

}  // namespace

int main()
{
    if (run_mix(1, true, 1000, 0, 6, 150)) return 1;
    if (run_mix(4, true, 1000, 100, 8, 150)) return 1;
    if (run_mix(3, false, 5000, 0, 8, 100)) return 1;
    if (second_waiter()) return 1;
    if (bad_runner()) return 1;
    if (ordering()) return 1;
    std::printf("sched_check ok\n");
    return 0;
}
