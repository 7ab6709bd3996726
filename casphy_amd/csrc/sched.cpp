// sched.cpp -- the request loop's scheduler (see sched.hpp).
#include "sched.hpp"

#include <algorithm>

namespace casphy {

namespace {

double ms_between(sched_clock::time_point a, sched_clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

Scheduler::Scheduler(const Options &o, WorkerInit init, Runner run) : opt_(o), init_(std::move(init)), run_(std::move(run))
{
    CASPHY_REQUIRE(opt_.workers >= 1, CASPHY_EARG, "scheduler: at least one worker");
    try {
        for (unsigned w = 0; w < opt_.workers; ++w)
            workers_.emplace_back([this, lane = opt_.workers >= 2 && w == 0] { worker_main(lane); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_work_.notify_all();
        for (auto &t : workers_) t.join();
        throw;
    }
}

Scheduler::~Scheduler()
{
    {
        std::lock_guard<std::mutex> lk(m_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &t : workers_) t.join();
}

u64 Scheduler::new_id()
{
    std::lock_guard<std::mutex> lk(m_);
    return next_id_++;
}

void Scheduler::add_resident(size_t bytes)
{
    std::lock_guard<std::mutex> lk(m_);
    resident_ += bytes;
}

void Scheduler::remove_resident(size_t bytes)
{
    {
        std::lock_guard<std::mutex> lk(m_);
        resident_ -= std::min(bytes, resident_);
    }
    cv_work_.notify_all();  // room may have freed up for a queued job
}

u64 Scheduler::submit(std::unique_ptr<SchedJob> j)
{
    u64 id;
    j->status = 1;
    j->t_submit = sched_clock::now();
    {
        std::lock_guard<std::mutex> lk(m_);
        if (!have_first_) {
            t_first_ = j->t_submit;
            have_first_ = true;
        }
        id = j->id = next_id_++;
        queue_.push_back(j.get());
        jobs_.emplace(id, std::move(j));
    }
    cv_work_.notify_all();
    return id;
}

std::unique_ptr<SchedJob> Scheduler::wait(u64 id)
{
    std::unique_lock<std::mutex> lk(m_);
    auto it = jobs_.find(id);
    if (it == jobs_.end() || it->second->claimed) return nullptr;
    SchedJob *j = it->second.get();
    j->claimed = true;  // the entry stays until this waiter erases it
    cv_done_.wait(lk, [&] { return j->status != 1; });
    std::unique_ptr<SchedJob> owned = std::move(it->second);
    jobs_.erase(it);
    return owned;
}

SchedStats Scheduler::stats()
{
    std::lock_guard<std::mutex> lk(m_);
    SchedStats st;
    std::vector<double> v = lat_ms_;
    std::sort(v.begin(), v.end());
    st.completed = v.size();
    auto pct = [&](double q) { return v.empty() ? 0.0 : v[std::min(v.size() - 1, (size_t)(q * (v.size() - 1) + 0.5))]; };
    st.p50_ms = pct(0.50);
    st.p99_ms = pct(0.99);
    const double span = have_first_ ? ms_between(t_first_, t_last_) / 1000.0 : 0.0;
    st.jobs_per_s = span > 0 ? v.size() / span : 0.0;
    st.launches = launches_;
    st.batched_jobs = batched_jobs_;
    return st;
}

std::vector<SchedJob *> Scheduler::pick(bool small_lane, size_t &bytes)
{
    std::vector<SchedJob *> group;
    if (queue_.empty()) return group;
    const auto now = sched_clock::now();
    SchedJob *head = queue_.front();
    // `extra`: what the group being formed already holds.  A first job that
    // fits nowhere still runs once nothing else is running.
    auto can_start = [&](size_t need, size_t extra) {
        return resident_ + running_ + extra + need <= opt_.budget || (running_ == 0 && extra == 0);
    };
    // the oldest job, once it has waited hol_ms without fitting, keeps its room
    const bool reserve = !can_start(head->dev_bytes, 0) && ms_between(head->t_submit, now) >= opt_.hol_ms;
    auto admissible = [&](const SchedJob *j, size_t extra) {
        if (j == head || !reserve) return can_start(j->dev_bytes, extra);
        return resident_ + running_ + extra + j->dev_bytes + head->dev_bytes <= opt_.budget;
    };
    auto first = queue_.end();
    double best = 0;
    for (auto it = queue_.begin(); it != queue_.end(); ++it) {
        if (small_lane && !(*it)->batchable) continue;
        if (!admissible(*it, 0)) continue;
        if (!opt_.shortest_first) {
            first = it;
            break;
        }
        const double key = (*it)->cost_ms - ms_between((*it)->t_submit, now);
        if (first == queue_.end() || key < best) {  // ties: the older job
            first = it;
            best = key;
        }
    }
    if (first == queue_.end()) return group;
    SchedJob *h = *first;
    group.push_back(h);
    bytes = h->dev_bytes;
    queue_.erase(first);
    if (h->batchable && opt_.batching) {
        for (auto it = queue_.begin(); it != queue_.end() && group.size() < opt_.batch_max_jobs;) {
            SchedJob *j = *it;
            if (j->batchable && j->batch_class == h->batch_class && admissible(j, bytes)) {
                group.push_back(j);
                bytes += j->dev_bytes;
                it = queue_.erase(it);
            } else {
                ++it;
            }
        }
    }
    return group;
}

void Scheduler::worker_main(bool small_lane)
{
    if (init_ && !init_()) return;
    for (;;) {
        std::vector<SchedJob *> group;
        size_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(m_);
            for (;;) {
                if (stop_) return;
                group = pick(small_lane, bytes);
                if (!group.empty()) break;
                // a reserved head may become admissible by the clock alone
                cv_work_.wait_for(lk, std::chrono::milliseconds(50));
            }
            running_ += bytes;
            ++launches_;
            if (group.size() > 1) batched_jobs_ += group.size();
        }
        std::string err;
        std::vector<int> rc;
        try {
            rc = run_(group, err);
        } catch (const Error &e) {
            rc.assign(group.size(), e.code);
            err = e.what();
        } catch (const std::bad_alloc &) {  // host allocation of the outputs
            rc.assign(group.size(), CASPHY_ENOMEM);
            err = "host allocation failed";
        } catch (const std::exception &e) {
            rc.assign(group.size(), CASPHY_EINTERNAL);
            err = e.what();
        }
        if (rc.size() != group.size()) {  // a runner that miscounted fails the whole group
            rc.assign(group.size(), CASPHY_EINTERNAL);
            err = "scheduler: the runner returned no status for some jobs of the group";
        }
        for (int &c : rc)
            if (c > 0) c = CASPHY_EINTERNAL;  // 1 is "pending": a waiter would never wake
        {
            // a waiter may free a job as soon as it sees its status: everything
            // written into a job is written here, before the notify, and
            // nothing after
            std::lock_guard<std::mutex> lk(m_);
            running_ -= bytes;
            const auto now = sched_clock::now();
            for (size_t k = 0; k < group.size(); ++k) {
                SchedJob *j = group[k];
                j->t_done = now;
                j->on_done();
                if (rc[k] == 0) lat_ms_.push_back(ms_between(j->t_submit, now));
                else j->err = err;
                j->status = rc[k];
            }
            t_last_ = now;
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

}  // namespace casphy
