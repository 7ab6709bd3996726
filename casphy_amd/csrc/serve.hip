// serve.hip -- request loop for tree-build and transfer-distance jobs.
//
// A casphy_server owns a pool of worker threads, each with its own stream on
// the server's device.  Jobs are submitted from any thread (inputs are copied,
// so callers may reuse their buffers) and admitted in FIFO order while their
// estimated device footprint fits the HBM budget.  A free worker takes the
// admitted head of the queue; when the head is a small join it also takes
// every other small join of the same kind and width that has arrived by then
// and runs them as one persistent work-queue launch (continuous batching:
// requests that arrive while others are on the GPU join the next launch on
// the next free worker, and launches of different workers overlap on the
// device).  Large joins and transfer-distance jobs run on their own worker.
// Each job's output equals the corresponding plain entry point's.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace casphy {
namespace {

constexpr u32 BATCH_MAX_CELLS = 2048;  // joins up to this size are batched
constexpr u32 BATCH_MAX_JOBS = 1024;

using clock_t_ = std::chrono::steady_clock;

double ms_between(clock_t_::time_point a, clock_t_::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Job {
    u64 id = 0;
    int kind = 0;  // CASPHY_KIND_BIT / CASPHY_KIND_INDEL joins, CASPHY_JOB_TRANSFER
    u32 n = 0, width = 0;
    int anc = 1;
    std::vector<u32> mat;     // bit join input, or Y then X for transfer
    std::vector<int32_t> imat;
    u32 ny = 0, nx = 0, ncol = 0, ntip = 0;
    std::vector<u32> out;     // join: raw edges; transfer: ymin|ysize|xmin|xsize
    size_t dev_bytes = 0;
    int status = 1;           // 1 pending, 0 done, < 0 error
    std::string err;
    clock_t_::time_point t_submit, t_done;
    bool batchable() const
    {
        return (kind == CASPHY_KIND_BIT || kind == CASPHY_KIND_INDEL) && n <= BATCH_MAX_CELLS;
    }
};

size_t stage_pad_rows(u32 n) { return ((size_t)n + 255) / 256 * 256; }

size_t estimate_bytes(const Job &j)
{
    if (j.kind == CASPHY_KIND_BIT) {
        const size_t ldw = (j.width + 31) / 32 * 32;
        return (size_t)j.n * (ldw * 4 * 5 + j.width * 4 + 96) + (size_t)stage_pad_rows(j.n) * ldw * 4;
    }
    if (j.kind == CASPHY_KIND_INDEL) {
        // one-hot width: per site max(max allele, 1) bits
        size_t bits = 0;
        for (u32 s = 0; s < j.width; ++s) {
            int32_t m = 1;
            for (u32 r = 0; r < j.n; ++r) m = std::max(m, j.imat[(size_t)r * j.width + s]);
            bits += (size_t)m;
        }
        const size_t ldw = ((bits + 31) / 32 + 31) / 32 * 32;
        return (size_t)j.n * (ldw * 4 * 5 + j.width * 4 + 96) + (size_t)stage_pad_rows(j.n) * ldw * 4;
    }
    const size_t ldw = (j.ncol + 31) / 32 * 32;
    return (size_t)(j.ny + j.nx + 512) * ldw * 4 * 4 + (size_t)(j.ny + j.nx) * 32;
}

}  // namespace
}  // namespace casphy

using namespace casphy;

struct casphy_server {
    int device = 0;
    size_t budget = 0;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::unordered_map<u64, std::unique_ptr<Job>> jobs;
    size_t in_use = 0;
    bool stop = false;
    u64 next_id = 1;
    std::vector<std::thread> workers;
    std::vector<double> lat_ms;
    bool have_first = false;
    clock_t_::time_point t_first, t_last;
    u64 launches = 0, batched_jobs = 0;
    bool batching = true;  // CASPHY_SERVE_BATCH=0: one job per launch (queue-only form)

    void worker_main();
    // Runs a group's launch and fills each job's `out`; returns the status of
    // every job (0 or a CASPHY_E* code) and the error message.  Job status,
    // error and completion time are published by worker_main under `m`.
    std::vector<int> run_group(std::vector<Job *> &group, std::string &err);
};

void casphy_server::worker_main()
{
    if (casphy_set_device(device) != CASPHY_OK) return;
    for (;;) {
        std::vector<Job *> group;
        size_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(m);
            for (;;) {
                if (stop) return;
                if (!queue.empty()) {
                    Job *h = queue.front();
                    if (in_use == 0 || in_use + h->dev_bytes <= budget) break;
                }
                cv_work.wait(lk);
            }
            Job *h = queue.front();
            queue.pop_front();
            group.push_back(h);
            bytes = h->dev_bytes;
            if (h->batchable() && batching) {
                for (auto it = queue.begin(); it != queue.end() && group.size() < BATCH_MAX_JOBS;) {
                    Job *j = *it;
                    if (j->batchable() && j->kind == h->kind && j->width == h->width && j->anc == h->anc &&
                        in_use + bytes + j->dev_bytes <= std::max(budget, h->dev_bytes)) {
                        group.push_back(j);
                        bytes += j->dev_bytes;
                        it = queue.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
            in_use += bytes;
            ++launches;
            if (group.size() > 1) batched_jobs += group.size();
        }
        std::string err;
        std::vector<int> rc;
        try {
            rc = run_group(group, err);
        } catch (const std::exception &e) {  // host allocation of the outputs
            rc.assign(group.size(), CASPHY_ENOMEM);
            err = e.what();
        }
        {
            // a waiter may free a job as soon as it sees its status: everything
            // the worker writes into a job is written here, before the notify,
            // and nothing after
            std::lock_guard<std::mutex> lk(m);
            in_use -= bytes;
            const auto now = clock_t_::now();
            for (size_t k = 0; k < group.size(); ++k) {
                Job *j = group[k];
                j->t_done = now;
                if (rc[k] == 0) {
                    lat_ms.push_back(ms_between(j->t_submit, now));
                } else {
                    j->err = err;
                }
                j->status = rc[k];
            }
            t_last = now;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

std::vector<int> casphy_server::run_group(std::vector<Job *> &group, std::string &err)
{
    std::vector<int> rc(group.size(), 0);
    auto fail = [&](int code) {
        err = casphy_last_error();
        std::fill(rc.begin(), rc.end(), code);
        return rc;
    };
    Job *h = group.front();
    if (h->kind == CASPHY_JOB_TRANSFER) {
        std::vector<u32> &o = h->out;
        o.assign((size_t)2 * h->ny + 2 * h->nx, 0);
        const u32 *Y = h->mat.data(), *X = h->mat.data() + (size_t)h->ny * h->ncol;
        const int r = casphy_transfer(Y, h->ny, X, h->nx, h->ncol, h->ntip, o.data(), o.data() + h->ny,
                                      o.data() + 2 * h->ny, o.data() + 2 * h->ny + h->nx);
        return r ? fail(r) : rc;
    }
    for (Job *j : group) j->out.assign((size_t)(2 * j->n - 2) * 2, 0);
    if (group.size() == 1) {
        const int r = h->kind == CASPHY_KIND_BIT
                          ? casphy_join_bit(h->mat.data(), h->n, h->width, nullptr, h->anc, h->out.data(), nullptr,
                                            nullptr)
                          : casphy_join_indel(h->imat.data(), h->n, h->width, nullptr, 0, nullptr, 0, 2.0f,
                                              h->out.data(), nullptr, nullptr);
        return r ? fail(r) : rc;
    }
    std::vector<u32> offs(group.size() + 1, 0);
    for (size_t k = 0; k < group.size(); ++k) offs[k + 1] = offs[k] + group[k]->n;
    std::vector<u32> edges((size_t)(2 * (u64)offs.back() - 2 * group.size()) * 2);
    int r;
    if (h->kind == CASPHY_KIND_BIT) {
        std::vector<u32> all((size_t)offs.back() * h->width);
        for (size_t k = 0; k < group.size(); ++k)
            std::copy(group[k]->mat.begin(), group[k]->mat.end(), all.begin() + (size_t)offs[k] * h->width);
        r = casphy_join_bit_batch(all.data(), offs.data(), (u32)group.size(), h->width, nullptr, h->anc, edges.data());
    } else {
        std::vector<int32_t> all((size_t)offs.back() * h->width);
        for (size_t k = 0; k < group.size(); ++k)
            std::copy(group[k]->imat.begin(), group[k]->imat.end(), all.begin() + (size_t)offs[k] * h->width);
        r = casphy_join_indel_batch(all.data(), offs.data(), (u32)group.size(), h->width, nullptr, 0, nullptr, 0,
                                    2.0f, edges.data());
    }
    if (r) return fail(r);
    size_t pos = 0;
    for (Job *j : group) {
        std::copy(edges.begin() + pos, edges.begin() + pos + j->out.size(), j->out.begin());
        pos += j->out.size();
    }
    return rc;
}

namespace {
int submit(casphy_server *s, std::unique_ptr<Job> j, uint64_t *id)
{
    j->dev_bytes = estimate_bytes(*j);
    j->t_submit = clock_t_::now();
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (!s->have_first) { s->t_first = j->t_submit; s->have_first = true; }
        j->id = s->next_id++;
        *id = j->id;
        s->queue.push_back(j.get());
        s->jobs.emplace(j->id, std::move(j));
    }
    s->cv_work.notify_one();
    return CASPHY_OK;
}
}  // namespace

extern "C" {

int casphy_server_create(int device, uint64_t hbm_budget_bytes, uint32_t workers, casphy_server **out)
{
    return guarded([&] {
        CASPHY_REQUIRE(out && workers >= 1 && workers <= 64, CASPHY_EARG, "casphy_server_create: bad arguments");
        int rc = casphy_set_device(device);
        if (rc) throw Error(rc, casphy_last_error());
        auto s = std::make_unique<casphy_server>();
        s->device = device;
        s->budget = hbm_budget_bytes;
        const char *b = std::getenv("CASPHY_SERVE_BATCH");
        s->batching = !(b && b[0] == '0');
        for (u32 w = 0; w < workers; ++w) s->workers.emplace_back([p = s.get()] { p->worker_main(); });
        *out = s.release();
    });
}

void casphy_server_destroy(casphy_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    delete s;
}

int casphy_server_submit_join(casphy_server *s, int kind, const void *mat, uint32_t n, uint32_t width,
                              int ancestral_type, uint64_t *job)
{
    return guarded([&] {
        CASPHY_REQUIRE(s && mat && job && n >= 2 && width > 0, CASPHY_EARG, "casphy_server_submit_join: bad arguments");
        CASPHY_REQUIRE(kind == CASPHY_KIND_BIT || kind == CASPHY_KIND_INDEL, CASPHY_EARG,
                       "casphy_server_submit_join: kind must be CASPHY_KIND_BIT or CASPHY_KIND_INDEL");
        CASPHY_REQUIRE(ancestral_type == 0 || ancestral_type == 1, CASPHY_EARG,
                       "casphy_server_submit_join: ancestralType must be 0 or 1");
        auto j = std::make_unique<Job>();
        j->kind = kind; j->n = n; j->width = width; j->anc = ancestral_type;
        if (kind == CASPHY_KIND_BIT) {
            const u32 *p = static_cast<const u32 *>(mat);
            j->mat.assign(p, p + (size_t)n * width);
        } else {
            const int32_t *p = static_cast<const int32_t *>(mat);
            j->imat.assign(p, p + (size_t)n * width);
        }
        submit(s, std::move(j), job);
    });
}

int casphy_server_submit_transfer(casphy_server *s, const uint32_t *Y, uint32_t ny, const uint32_t *X, uint32_t nx,
                                  uint32_t ncol, uint32_t ntip, uint64_t *job)
{
    return guarded([&] {
        CASPHY_REQUIRE(s && Y && X && job && ny > 0 && nx > 0 && ncol > 0 && ntip <= 32u * ncol, CASPHY_EARG,
                       "casphy_server_submit_transfer: bad arguments");
        auto j = std::make_unique<Job>();
        j->kind = CASPHY_JOB_TRANSFER; j->ny = ny; j->nx = nx; j->ncol = ncol; j->ntip = ntip;
        j->mat.assign(Y, Y + (size_t)ny * ncol);
        j->mat.insert(j->mat.end(), X, X + (size_t)nx * ncol);
        submit(s, std::move(j), job);
    });
}

int casphy_server_wait(casphy_server *s, uint64_t job, uint32_t *out, uint64_t out_len, double *latency_ms)
{
    return guarded([&] {
        CASPHY_REQUIRE(s, CASPHY_EARG, "casphy_server_wait: bad arguments");
        std::unique_lock<std::mutex> lk(s->m);
        auto it = s->jobs.find(job);
        CASPHY_REQUIRE(it != s->jobs.end(), CASPHY_EARG, "casphy_server_wait: unknown job");
        Job *j = it->second.get();
        s->cv_done.wait(lk, [&] { return j->status != 1; });
        std::unique_ptr<Job> owned = std::move(it->second);
        s->jobs.erase(it);
        if (j->status < 0) throw Error(j->status, j->err);
        CASPHY_REQUIRE(out && out_len >= j->out.size(), CASPHY_EARG,
                       "casphy_server_wait: output buffer needs " + std::to_string(j->out.size()) + " values");
        std::copy(j->out.begin(), j->out.end(), out);
        if (latency_ms) *latency_ms = ms_between(j->t_submit, j->t_done);
    });
}

int casphy_server_stats(casphy_server *s, uint64_t *completed, double *p50_ms, double *p99_ms, double *jobs_per_s,
                        uint64_t *launches)
{
    return guarded([&] {
        CASPHY_REQUIRE(s && completed && p50_ms && p99_ms && jobs_per_s && launches, CASPHY_EARG,
                       "casphy_server_stats: bad arguments");
        std::lock_guard<std::mutex> lk(s->m);
        std::vector<double> v = s->lat_ms;
        std::sort(v.begin(), v.end());
        *completed = v.size();
        auto pct = [&](double q) { return v.empty() ? 0.0 : v[std::min(v.size() - 1, (size_t)(q * (v.size() - 1) + 0.5))]; };
        *p50_ms = pct(0.50);
        *p99_ms = pct(0.99);
        const double span = s->have_first ? ms_between(s->t_first, s->t_last) / 1000.0 : 0.0;
        *jobs_per_s = span > 0 ? v.size() / span : 0.0;
        *launches = s->launches;
    });
}

}  // extern "C"
