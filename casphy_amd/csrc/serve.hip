// serve.hip -- request loop for tree-build and transfer-distance jobs.
//
// A casphy_server owns a pool of worker threads, each with its own stream on
// the server's device, so the launches of different workers overlap on the
// device.  Jobs are submitted from any thread; their inputs are copied, so
// callers may reuse their buffers.
//   * Packed matrices can be made resident once (casphy_server_put_matrix)
//     and joined over any subset of their rows (casphy_server_submit_join_rows):
//     the worker gathers the rows on the device, nothing crosses PCIe per job.
//   * Admission by HBM budget without head-of-line blocking: a worker takes the
//     first queued job whose estimated footprint fits next to the running
//     jobs; once the oldest job has waited `hol_ms` without fitting, later jobs
//     start only if they leave room for it, so it starts as soon as enough
//     running jobs finish.
//   * Continuous batching: a worker that takes a small join also takes every
//     other small join of the same kind, width and rule queued by then and runs
//     them as one persistent work-queue launch.  With two or more workers, one
//     of them takes small joins only, so they never wait behind large ones.
//   * Transfer jobs come as clade bitsets or as trees (edge matrices: the
//     transfer distances are computed from the trees on the device).
// The queue, admission, batching and completion bookkeeping are the host-only
// Scheduler (sched.hpp); this file holds the jobs and their device work.
// Each job's output equals the corresponding plain entry point's.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "context.hpp"
#include "join.hpp"
#include "sched.hpp"

namespace casphy {
namespace {

constexpr u32 BATCH_MAX_CELLS = 2048;  // joins up to this size are batched
constexpr u32 BATCH_MAX_JOBS = 1024;
constexpr int JOB_TREES = 4;           // transfer distances from edge matrices

// A packed matrix kept on the server's device.
struct Resident {
    int kind = CASPHY_KIND_BIT;
    u32 n = 0, width = 0;
    u32 onehot_bits = 0;  // indel: one-hot width of the whole matrix (bounds any row subset's)
    DevBuf d;
};

// out[i][0 .. width) = mat[rows[i]][0 .. width) (32-bit words)
__global__ void k_gather_words(const u32 *mat, u32 width, const u32 *rows, u32 n, u32 *out)
{
    const u64 total = (u64)n * width;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 i = (u32)(x / width), w = (u32)(x - (u64)i * width);
        out[x] = mat[(size_t)rows[i] * width + w];
    }
}

struct Job : SchedJob {
    int kind = 0;  // CASPHY_KIND_BIT / CASPHY_KIND_INDEL joins, CASPHY_JOB_TRANSFER, JOB_TREES
    u32 n = 0, width = 0;
    int anc = 1;
    std::vector<u32> mat;     // bit join input, or Y then X for transfer
    std::vector<int32_t> imat;
    std::shared_ptr<Resident> res;  // join over rows of a resident matrix
    std::vector<u32> rows;
    u32 ny = 0, nx = 0, ncol = 0, ntip = 0;
    // JOB_TREES: reference and replicate edge matrices
    std::vector<int32_t> yanc, ydes, xanc, xdes;
    std::vector<u32> xoff, xnnode;
    u32 ynedge = 0, ynnode = 0, nrep = 0;
    std::vector<u32> out;     // join: raw edges; transfer: ymin|ysize|xmin|xsize
    bool join() const { return kind == CASPHY_KIND_BIT || kind == CASPHY_KIND_INDEL; }
    void on_done() override { res.reset(); }  // the resident matrix may be dropped now
};

size_t stage_pad_rows(u32 n) { return ((size_t)n + 255) / 256 * 256; }

size_t join_bytes(u32 n, size_t ldw, u32 width)
{
    return (size_t)n * (ldw * 4 * 5 + width * 4 * 2 + 96) + stage_pad_rows(n) * ldw * 4;
}

u32 onehot_bits(const int32_t *m, u32 n, u32 width)
{
    size_t bits = 0;
    for (u32 s = 0; s < width; ++s) {
        int32_t mx = 1;
        for (u32 r = 0; r < n; ++r) mx = std::max(mx, m[(size_t)r * width + s]);
        bits += (size_t)mx;
    }
    return (u32)std::min<size_t>(bits, 0xFFFFFFFFu);
}

size_t estimate_bytes(const Job &j)
{
    if (j.kind == CASPHY_KIND_BIT) return join_bytes(j.n, (j.width + 31) / 32 * 32, j.width);
    if (j.kind == CASPHY_KIND_INDEL) {
        const u32 bits = j.res ? j.res->onehot_bits : onehot_bits(j.imat.data(), j.n, j.width);
        return join_bytes(j.n, ((size_t)(bits + 31) / 32 + 31) / 32 * 32, j.width);
    }
    if (j.kind == JOB_TREES) {
        // per tree: the walk programs and positions (about 16 B per node) and the
        // outputs; the clade plans are host memory
        const size_t nodes = (size_t)j.ynedge + j.xanc.size() + j.nrep + 1;
        return nodes * 48 + (size_t)j.nrep * j.ynnode * 8 + (64ull << 20);
    }
    const size_t ldw = (j.ncol + 31) / 32 * 32;
    return (size_t)(j.ny + j.nx + 512) * ldw * 4 * 4 + (size_t)(j.ny + j.nx) * 32;
}

}  // namespace
}  // namespace casphy

using namespace casphy;

struct casphy_server {
    int device = 0;
    std::mutex rm;  // guards `resident`
    std::unordered_map<u64, std::shared_ptr<Resident>> resident;
    std::unique_ptr<Scheduler> sched;
};

namespace {

// Large single joins running on each device (request-loop workers of every
// server in the process).
constexpr u32 LARGE_JOIN_CELLS = 8192;
std::atomic<int> g_large_joins[64];

struct LargeJoin {
    int dev = -1, others = 0;
    explicit LargeJoin(bool large)
    {
        const int d = thread_device();
        if (!large || d < 0 || d >= 64) return;
        dev = d;
        others = g_large_joins[d].fetch_add(1);
    }
    ~LargeJoin()
    {
        if (dev >= 0) g_large_joins[dev].fetch_sub(1);
    }
};

// The device work of one group (one job, or a batch of small joins of one
// class): fills each job's `out`, returns each job's status.
std::vector<int> run_group(std::vector<SchedJob *> &sgroup, std::string &err)
{
    std::vector<Job *> group(sgroup.size());
    for (size_t k = 0; k < sgroup.size(); ++k) group[k] = static_cast<Job *>(sgroup[k]);
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
    if (h->kind == JOB_TREES) {
        const u32 ny = h->ynnode - 1;
        size_t nxr = 0;
        for (u32 b = 0; b < h->nrep; ++b) nxr += h->xnnode[b] - 1;
        std::vector<u32> &o = h->out;
        o.assign((size_t)h->nrep * ny + ny + 2 * nxr, 0);
        const int r = casphy_transfer_trees(h->yanc.data(), h->ydes.data(), h->ynedge, h->ynnode,
                                            h->xanc.empty() ? nullptr : h->xanc.data(),
                                            h->xdes.empty() ? nullptr : h->xdes.data(), h->xoff.data(),
                                            h->xnnode.data(), h->nrep, h->ntip, o.data(),
                                            o.data() + (size_t)h->nrep * ny, o.data() + (size_t)h->nrep * ny + ny,
                                            o.data() + (size_t)h->nrep * ny + ny + nxr);
        return r ? fail(r) : rc;
    }
    // joins: the group's rows assembled on the device (uploaded from the
    // copies taken at submission, or gathered from a resident matrix)
    const int r = guarded([&] {
               hipStream_t s = thread_stream();
               std::vector<u32> offs(group.size() + 1, 0);
               for (size_t k = 0; k < group.size(); ++k) offs[k + 1] = offs[k] + group[k]->n;
               const u32 W = h->width;
               DevBuf raw((size_t)offs.back() * W * 4), didx;
               for (size_t k = 0; k < group.size(); ++k) {
                   Job *j = group[k];
                   u32 *dst = raw.as<u32>() + (size_t)offs[k] * W;
                   if (j->res) {
                       didx.reserve((size_t)j->n * 4);
                       h2d(didx.p, j->rows.data(), (size_t)j->n * 4, s);
                       const u64 words = (u64)j->n * W;
                       hipLaunchKernelGGL(k_gather_words, dim3((u32)std::min<u64>((words + 255) / 256, 4096)),
                                          dim3(256), 0, s, j->res->d.as<u32>(), W, didx.as<u32>(), j->n, dst);
                       CASPHY_HIP(hipGetLastError());
                   } else if (j->kind == CASPHY_KIND_BIT) {
                       h2d(dst, j->mat.data(), (size_t)j->n * W * 4, s);
                   } else {
                       h2d(dst, j->imat.data(), (size_t)j->n * W * 4, s);
                   }
                   j->out.assign((size_t)(2 * j->n - 2) * 2, 0);
               }
               const bool bit = h->kind == CASPHY_KIND_BIT;
               if (group.size() == 1) {
                   // a large join that starts while another runs takes at most 256
                   // workgroups per step, so that the two share the device; alone
                   // it takes the single-join count
                   LargeJoin lj(h->n >= LARGE_JOIN_CELLS);
                   set_join_workgroup_cap(lj.others > 0 ? 256 : 0);
                   if (bit) join_bit_dev(raw.as<u32>(), h->n, W, h->anc, h->out.data(), s);
                   else join_indel_dev(raw.as<int32_t>(), h->n, W, h->out.data(), s);
                   return;
               }
               std::vector<u32> edges((size_t)(2 * (u64)offs.back() - 2 * group.size()) * 2);
               if (bit) join_bit_batch_dev(raw.as<u32>(), offs, W, h->anc, edges.data(), s);
               else join_indel_batch_dev(raw.as<int32_t>(), offs, W, edges.data(), s);
               size_t pos = 0;
               for (Job *j : group) {
                   std::copy(edges.begin() + pos, edges.begin() + pos + j->out.size(), j->out.begin());
                   pos += j->out.size();
               }
           });
    return r == CASPHY_OK ? rc : fail(r);
}

// Estimated device time of a job (ms), for the shortest-first order.  A join
// of n cells takes about 2 n one-launch steps of a few microseconds plus
// rescoring that grows with n^2 and the row width: measured 11 ms at 1,024
// cells, 121-155 ms at 10,000, 462-568 ms at 30,000 and 2.6-2.8 s at 100,000
// (4,096-site binary rows, 200-site indel rows).  A transfer job from trees
// walks its replicates' clades (about 2.6 ms per 65,536-tip replicate); from
// bitsets it scores ny x nx rows.
double estimate_ms(const Job &j)
{
    if (j.join()) {
        // row words: a bit row is `width` words, an indel row one byte per site
        const double n = j.n, words = j.kind == CASPHY_KIND_BIT ? j.width : (j.width + 3) / 4;
        return n * 0.014 + n * n * 1.2e-7 * std::max(1.0, words / 128.0);
    }
    if (j.kind == JOB_TREES) return 1.0 + 4.0e-5 * (double)j.ntip * std::max(1u, j.nrep);
    return 1.0 + 1.0e-9 * (double)j.ny * j.nx * std::max(1u, j.ncol);
}

int submit(casphy_server *s, std::unique_ptr<Job> j, uint64_t *id)
{
    j->dev_bytes = estimate_bytes(*j);
    j->cost_ms = estimate_ms(*j);
    j->batchable = j->join() && j->n <= BATCH_MAX_CELLS;
    j->batch_class = ((u64)(u32)j->kind << 34) | ((u64)(u32)(j->anc & 3) << 32) | j->width;
    *id = s->sched->submit(std::move(j));
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
        Scheduler::Options o;
        o.budget = hbm_budget_bytes;
        o.workers = workers;
        const char *b = std::getenv("CASPHY_SERVE_BATCH");
        o.batching = !(b && b[0] == '0');
        const char *hol = std::getenv("CASPHY_SERVE_HOL_MS");
        if (hol) o.hol_ms = std::atof(hol);
        o.batch_max_jobs = BATCH_MAX_JOBS;
        s->sched = std::make_unique<Scheduler>(
            o, [device] { return casphy_set_device(device) == CASPHY_OK; }, run_group);
        *out = s.release();
    });
}

void casphy_server_destroy(casphy_server *s)
{
    if (!s) return;
    s->sched.reset();  // stops and joins the workers
    delete s;
}

int casphy_server_put_matrix(casphy_server *s, int kind, const void *mat, uint32_t n, uint32_t width,
                             uint64_t *matrix)
{
    return guarded([&] {
        CASPHY_REQUIRE(s && mat && matrix && n > 0 && width > 0, CASPHY_EARG,
                       "casphy_server_put_matrix: bad arguments");
        CASPHY_REQUIRE(kind == CASPHY_KIND_BIT || kind == CASPHY_KIND_INDEL, CASPHY_EARG,
                       "casphy_server_put_matrix: kind must be CASPHY_KIND_BIT or CASPHY_KIND_INDEL");
        auto r = std::make_shared<Resident>();
        r->kind = kind;
        r->n = n;
        r->width = width;
        if (kind == CASPHY_KIND_INDEL) r->onehot_bits = onehot_bits(static_cast<const int32_t *>(mat), n, width);
        const size_t bytes = (size_t)n * width * 4;
        // upload on the server's device from this thread, then restore the thread's device
        const int prev = thread_device();
        int rc = casphy_set_device(s->device);
        if (rc) throw Error(rc, casphy_last_error());
        try {
            hipStream_t st = thread_stream();
            r->d.alloc(bytes);
            h2d(r->d.p, mat, bytes, st);
            CASPHY_HIP(hipStreamSynchronize(st));
        } catch (...) {
            (void)casphy_set_device(prev);
            throw;
        }
        (void)casphy_set_device(prev);
        const u64 id = s->sched->new_id();
        {
            std::lock_guard<std::mutex> lk(s->rm);
            s->resident.emplace(id, std::move(r));
        }
        s->sched->add_resident(bytes);
        *matrix = id;
    });
}

int casphy_server_drop_matrix(casphy_server *s, uint64_t matrix)
{
    return guarded([&] {
        CASPHY_REQUIRE(s, CASPHY_EARG, "casphy_server_drop_matrix: bad arguments");
        size_t bytes;
        {
            std::lock_guard<std::mutex> lk(s->rm);
            auto it = s->resident.find(matrix);
            CASPHY_REQUIRE(it != s->resident.end(), CASPHY_EARG, "casphy_server_drop_matrix: unknown matrix");
            bytes = (size_t)it->second->n * it->second->width * 4;
            s->resident.erase(it);  // queued and running jobs keep their reference
        }
        s->sched->remove_resident(bytes);
    });
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

int casphy_server_submit_join_rows(casphy_server *s, uint64_t matrix, const uint32_t *rows, uint32_t nrows,
                                   int ancestral_type, uint64_t *job)
{
    return guarded([&] {
        CASPHY_REQUIRE(s && rows && job && nrows >= 2, CASPHY_EARG, "casphy_server_submit_join_rows: bad arguments");
        CASPHY_REQUIRE(ancestral_type == 0 || ancestral_type == 1, CASPHY_EARG,
                       "casphy_server_submit_join_rows: ancestralType must be 0 or 1");
        std::shared_ptr<Resident> r;
        {
            std::lock_guard<std::mutex> lk(s->rm);
            auto it = s->resident.find(matrix);
            CASPHY_REQUIRE(it != s->resident.end(), CASPHY_EARG, "casphy_server_submit_join_rows: unknown matrix");
            r = it->second;
        }
        for (u32 i = 0; i < nrows; ++i)
            CASPHY_REQUIRE(rows[i] < r->n, CASPHY_EARG,
                           "casphy_server_submit_join_rows: row " + std::to_string(rows[i]) + " out of range");
        auto j = std::make_unique<Job>();
        j->kind = r->kind; j->n = nrows; j->width = r->width; j->anc = ancestral_type;
        j->res = std::move(r);
        j->rows.assign(rows, rows + nrows);
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

int casphy_server_submit_transfer_trees(casphy_server *s, const int32_t *y_anc, const int32_t *y_des,
                                        uint32_t y_nedge, uint32_t y_nnode, const int32_t *x_anc,
                                        const int32_t *x_des, const uint32_t *x_edge_offsets,
                                        const uint32_t *x_nnode, uint32_t nrep, uint32_t ntip, uint64_t *job)
{
    return guarded([&] {
        CASPHY_REQUIRE(s && y_anc && y_des && x_edge_offsets && x_nnode && job && nrep > 0 && y_nnode >= 2 &&
                           ntip >= 2,
                       CASPHY_EARG, "casphy_server_submit_transfer_trees: bad arguments");
        const u32 xe = x_edge_offsets[nrep];
        CASPHY_REQUIRE(x_edge_offsets[0] == 0 && (xe == 0 || (x_anc && x_des)), CASPHY_EARG,
                       "casphy_server_submit_transfer_trees: bad replicate edges");
        for (u32 b = 0; b < nrep; ++b)
            CASPHY_REQUIRE(x_edge_offsets[b + 1] >= x_edge_offsets[b] && x_nnode[b] >= 1, CASPHY_EARG,
                           "casphy_server_submit_transfer_trees: bad replicate " + std::to_string(b));
        auto j = std::make_unique<Job>();
        j->kind = JOB_TREES; j->ntip = ntip; j->nrep = nrep; j->ynedge = y_nedge; j->ynnode = y_nnode;
        j->yanc.assign(y_anc, y_anc + y_nedge);
        j->ydes.assign(y_des, y_des + y_nedge);
        if (xe) {
            j->xanc.assign(x_anc, x_anc + xe);
            j->xdes.assign(x_des, x_des + xe);
        }
        j->xoff.assign(x_edge_offsets, x_edge_offsets + nrep + 1);
        j->xnnode.assign(x_nnode, x_nnode + nrep);
        submit(s, std::move(j), job);
    });
}

int casphy_server_wait(casphy_server *s, uint64_t job, uint32_t *out, uint64_t out_len, double *latency_ms)
{
    return guarded([&] {
        CASPHY_REQUIRE(s, CASPHY_EARG, "casphy_server_wait: bad arguments");
        std::unique_ptr<SchedJob> owned = s->sched->wait(job);
        CASPHY_REQUIRE(owned, CASPHY_EARG, "casphy_server_wait: unknown job (or another thread waits for it)");
        const Job *j = static_cast<const Job *>(owned.get());
        if (j->status < 0) throw Error(j->status, j->err);
        CASPHY_REQUIRE(out && out_len >= j->out.size(), CASPHY_EARG,
                       "casphy_server_wait: output buffer needs " + std::to_string(j->out.size()) + " values");
        std::copy(j->out.begin(), j->out.end(), out);
        if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
    });
}

int casphy_server_stats(casphy_server *s, uint64_t *completed, double *p50_ms, double *p99_ms, double *jobs_per_s,
                        uint64_t *launches)
{
    return guarded([&] {
        CASPHY_REQUIRE(s && completed && p50_ms && p99_ms && jobs_per_s && launches, CASPHY_EARG,
                       "casphy_server_stats: bad arguments");
        const SchedStats st = s->sched->stats();
        *completed = st.completed;
        *p50_ms = st.p50_ms;
        *p99_ms = st.p99_ms;
        *jobs_per_s = st.jobs_per_s;
        *launches = st.launches;
    });
}

}  // extern "C"
