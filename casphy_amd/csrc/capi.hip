// capi.hip -- C ABI entry points: scoring, radius lists, tryRadius, transfer
// distance, cladewise emission.  Joins live in join.hip.
//
// Every entry point follows the same shape: validate, upload the caller's
// reference-layout buffers once, convert to the device layout with a kernel,
// run the tile kernel with the matching fused epilogue, compact on device and
// copy only the final results back.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "cells.hpp"
#include "context.hpp"
#include "kernels.hpp"
#include "mfma.hpp"
#include "clades.hpp"
#include "fasta.hpp"
#include "newick.hpp"
#include "voteroot.hpp"
#include "nodekeep.hpp"
#include "tipmerge.hpp"
#include "tbe.hpp"
#include "tiles.hpp"

namespace casphy {

namespace {
thread_local std::string g_last_error;
thread_local int g_device = 0;
thread_local hipStream_t g_stream = nullptr;
thread_local int g_stream_device = -1;
}  // namespace

void set_last_error(const std::string &msg) { g_last_error = msg; }

int thread_device() { return g_device; }

hipStream_t thread_stream()
{
    if (!g_stream || g_stream_device != g_device) {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
            throw Error(CASPHY_ENODEV, "no HIP device available");
        CASPHY_REQUIRE(g_device < count, CASPHY_ENODEV, "selected device does not exist");
        CASPHY_HIP(hipSetDevice(g_device));
        CASPHY_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
        g_stream_device = g_device;
        // keep freed stream-ordered memory in the pool instead of returning it
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, g_device) == hipSuccess && pool) {
            uint64_t keep = ~0ull;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    } else {
        CASPHY_HIP(hipSetDevice(g_device));
    }
    return g_stream;
}

namespace {
// Copies of 64 MiB and more between pageable host memory and the device go
// through two pinned 32 MiB buffers: the copy engine fills (or drains) one
// while host threads copy the other out (or in).  The buffers are allocated
// on the first such copy and kept until the process ends; one copy at a time
// uses them.
constexpr size_t STAGED_MIN = 64ull << 20, STAGE_CHUNK = 32ull << 20;
constexpr unsigned STAGE_THREADS = 4;
std::mutex g_stage_mu;
void *g_stage[2] = {nullptr, nullptr};

// Never throws: a part whose thread cannot be started is copied by the
// calling thread, and every started thread is joined, so a staged copy cannot
// unwind (and release the staging buffers) while its DMA is in flight.
void parallel_copy(void *dst, const void *src, size_t bytes) noexcept
{
    const size_t part = (bytes / STAGE_THREADS + 4095) / 4096 * 4096;
    auto copy = [=](unsigned k) {
        const size_t off = std::min(bytes, k * part), len = std::min(bytes - off, part);
        std::memcpy(static_cast<char *>(dst) + off, static_cast<const char *>(src) + off, len);
    };
    std::thread t[STAGE_THREADS - 1];
    for (unsigned k = 1; k < STAGE_THREADS; ++k) {
        try {
            t[k - 1] = std::thread(copy, k);
        } catch (...) {
            copy(k);
        }
    }
    copy(0);
    for (auto &x : t)
        if (x.joinable()) x.join();
}

void staged_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_stage_mu);
    for (auto &b : g_stage)
        if (!b) CASPHY_HIP(hipHostMalloc(&b, STAGE_CHUNK, hipHostMallocPortable));
    hipEvent_t ev[2];
    CASPHY_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    const hipError_t e1 = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    if (e1 != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        CASPHY_HIP(e1);
    }
    const size_t nchunk = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
    auto len = [&](size_t k) { return std::min(STAGE_CHUNK, bytes - k * STAGE_CHUNK); };
    hipError_t err = hipSuccess;
    auto run = [&](hipError_t e) {
        if (err == hipSuccess) err = e;
        return err == hipSuccess;
    };
    if (kind == hipMemcpyDeviceToHost) {
        // chunk k into buffer k % 2; chunk k + 1 is in flight while chunk k is copied out
        auto issue = [&](size_t k) {
            return run(hipMemcpyAsync(g_stage[k & 1], static_cast<const char *>(src) + k * STAGE_CHUNK, len(k),
                                      hipMemcpyDeviceToHost, s)) &&
                   run(hipEventRecord(ev[k & 1], s));
        };
        issue(0);
        for (size_t k = 0; k < nchunk && err == hipSuccess; ++k) {
            if (k + 1 < nchunk && !issue(k + 1)) break;
            if (!run(hipEventSynchronize(ev[k & 1]))) break;
            parallel_copy(static_cast<char *>(dst) + k * STAGE_CHUNK, g_stage[k & 1], len(k));
        }
    } else {
        // buffer k % 2 is refilled once the copy of chunk k - 2 out of it is done
        for (size_t k = 0; k < nchunk && err == hipSuccess; ++k) {
            if (k >= 2 && !run(hipEventSynchronize(ev[k & 1]))) break;
            parallel_copy(g_stage[k & 1], static_cast<const char *>(src) + k * STAGE_CHUNK, len(k));
            if (!run(hipMemcpyAsync(static_cast<char *>(dst) + k * STAGE_CHUNK, g_stage[k & 1], len(k),
                                    hipMemcpyHostToDevice, s)))
                break;
            run(hipEventRecord(ev[k & 1], s));
        }
    }
    run(hipStreamSynchronize(s));
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    CASPHY_HIP(err);
}
}  // namespace

void h2d(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    if (!bytes) return;
    if (bytes >= STAGED_MIN) return staged_copy(dst, src, bytes, hipMemcpyHostToDevice, s);
    CASPHY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    CASPHY_HIP(hipStreamSynchronize(s));
}

void d2h(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    if (!bytes) {
        CASPHY_HIP(hipStreamSynchronize(s));
        return;
    }
    if (bytes >= STAGED_MIN) return staged_copy(dst, src, bytes, hipMemcpyDeviceToHost, s);
    CASPHY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    CASPHY_HIP(hipStreamSynchronize(s));
}

static void check_weight(double v, u32 i)
{
    if (!std::isfinite(v) || v < 0.0 || v != std::floor(v))
        throw Error(CASPHY_EWEIGHT, "weight[" + std::to_string(i) + "] = " + std::to_string(v) +
                                        " is not a non-negative integer");
}

Planes make_planes_bits(const std::vector<double> &w, u32 ldw)
{
    Planes p;
    double sum = 0.0;
    u32 maxw = 0;
    bool unit = true;
    for (u32 i = 0; i < (u32)w.size(); ++i) {
        check_weight(w[i], i);
        sum += w[i];
        maxw = std::max(maxw, (u32)w[i]);
        unit &= w[i] == 1.0;
    }
    if (sum >= 16777216.0)
        throw Error(CASPHY_EWEIGHT, "weights sum to >= 2^24: float accumulation is not exact");
    if (unit) return p;
    std::vector<int> bits;
    for (int b = 23; b >= 0; --b) {
        bool any = false;
        for (double v : w) any |= (((u32)v >> b) & 1u) != 0;
        if (any) bits.push_back(b);
    }
    if (bits.empty()) {  // all weights zero: one empty plane
        bits.push_back(0);
    }
    p.nplanes = (u32)bits.size();
    p.mask.assign((size_t)p.nplanes * ldw, 0u);
    for (u32 q = 0; q < p.nplanes; ++q) {
        for (u32 i = 0; i < (u32)w.size(); ++i)
            if (((u32)w[i] >> bits[q]) & 1u) p.mask[(size_t)q * ldw + (i >> 5)] |= 1u << (i & 31);
        if (q > 0) p.shift[q] = (u32)(bits[q - 1] - bits[q]);
    }
    p.tail = (u32)bits.back();
    return p;
}

Planes make_planes(const double *w, u32 nsite, u32 ldw)
{
    if (!w) return Planes{};
    return make_planes_bits(std::vector<double>(w, w + nsite), ldw);
}

// The int8 matrix-core path serves unit-weight bit scores; CASPHY_MFMA=0
// selects the VALU popcount tiles instead (both are exact).
bool mfma_enabled()
{
    const char *e = std::getenv("CASPHY_MFMA");
    return !(e && e[0] == '0');
}

void apply_planes(const Planes &p, TileArgs &a, DevBuf &buf, hipStream_t s)
{
    a.nplanes = p.nplanes;
    if (!p.nplanes) return;
    buf.reserve(p.mask.size() * 4);
    h2d(buf.p, p.mask.data(), p.mask.size() * 4, s);
    a.pmask = buf.as<u32>();
    for (u32 q = 0; q < p.nplanes; ++q) a.pshift[q] = p.shift[q];
    a.ptail = p.tail;
}

namespace {

// Device copy of a bit matrix in cell-major rows of ldw words.
struct BitMat {
    DevBuf d;
    u32 n = 0, W = 0, ldw = 0;
};

BitMat upload_bit(const u32 *host, u32 W, u32 n, bool wordmajor, hipStream_t s)
{
    BitMat m;
    m.n = n; m.W = W; m.ldw = round_up(std::max<u32>(W, 1), KC_BIT);
    DevBuf raw((size_t)W * n * 4);
    h2d(raw.p, host, (size_t)W * n * 4, s);
    m.d.alloc((size_t)n * m.ldw * 4);
    if (wordmajor) launch_transpose_wordmajor(raw.as<u32>(), W, n, m.d.as<u32>(), m.ldw, s);
    else launch_pad_rows(raw.as<u32>(), W, n, m.d.as<u32>(), m.ldw, s);
    CASPHY_HIP(hipStreamSynchronize(s));
    return m;
}

// DNA in the form chosen for its content.  With A/C/G/T and gaps only, a
// position's code has at most one nucleotide bit, so the low nibbles of the
// byte codes are one-hot bits of (position, base): two sequences share a base
// at a position exactly when their one-hot bits there share a set bit, and
// S(i, j) is the AND-popcount of the nibble rows -- the matrix-core tiles
// score it as bit cells of 4 sites per position (M_AND, unit weights).  The
// VALU form of the same data is 2-bit + mask (M_DNA2, 12 bytes per 32
// positions).  IUPAC ambiguity codes (several bits in a nibble) make the
// share test an OR over bits, which is no dot product: those run as nibbles
// on the VALU tiles (M_NIB).
// CASPHY_DNA_FORM=mfma|valu forces the form of ACGT data; the default is the
// matrix cores when they are enabled (resident cells time both, see
// tune_dna_form).
struct DnaMat {
    DevBuf d;
    u32 n = 0, ldw = 0;
    int metric = M_DNA2;
    DevBuf d2;          // the 2-bit + mask copy kept for the form tuner (M_AND only)
    u32 ldw2 = 0;
};

int dna_form_env()  // 1 = matrix cores, 0 = VALU, -1 = not forced
{
    const char *e = std::getenv("CASPHY_DNA_FORM");
    if (e && !std::strcmp(e, "mfma")) return 1;
    if (e && !std::strcmp(e, "valu")) return 0;
    return -1;
}

void pack_dna2(const u32 *raw, u32 n, u32 nwRaw, DevBuf &d, u32 &ldw, hipStream_t s)
{
    ldw = round_up(std::max<u32>(3 * ceil_div(nwRaw, 8), 1), KC_DNA2);
    d.alloc((size_t)n * ldw * 4);
    launch_dna_pack2(raw, n, nwRaw, d.as<u32>(), ldw, s);
}

// keep_valu: also keep the 2-bit + mask copy when the one-hot form is chosen
DnaMat upload_dna(const u32 *host, u32 n, u32 nwRaw, hipStream_t s, bool keep_valu = false)
{
    DnaMat m;
    m.n = n;
    DevBuf raw((size_t)n * nwRaw * 4), flag(4);
    h2d(raw.p, host, (size_t)n * nwRaw * 4, s);
    launch_fill_u32(flag.as<u32>(), 1, 0, s);
    launch_dna_detect(raw.as<u32>(), (u64)n * nwRaw, flag.as<u32>(), s);
    u32 ambiguous = 0;
    d2h(&ambiguous, flag.p, 4, s);
    const int forced = dna_form_env();
    const bool onehot = !ambiguous && mfma_enabled() && forced != 0;
    if (onehot || ambiguous) {
        m.metric = ambiguous ? M_NIB : M_AND;
        m.ldw = round_up(std::max<u32>(ceil_div(nwRaw, 2), 1), KC_BIT);
        m.d.alloc((size_t)n * m.ldw * 4);
        launch_dna_pack_nib(raw.as<u32>(), n, nwRaw, m.d.as<u32>(), m.ldw, s);
        if (onehot && keep_valu && forced < 0) pack_dna2(raw.as<u32>(), n, nwRaw, m.d2, m.ldw2, s);
    } else {
        m.metric = M_DNA2;
        pack_dna2(raw.as<u32>(), n, nwRaw, m.d, m.ldw, s);
    }
    CASPHY_HIP(hipStreamSynchronize(s));
    return m;
}

}  // namespace

// Stage-major copy of a cell-major matrix for the MFMA tiles.
void make_stage(const u32 *dmat, u32 n, u32 ldw, DevBuf &P, hipStream_t s)
{
    P.reserve((size_t)stage_rows(n) * ldw * 4);
    launch_to_stage_major(dmat, n, ldw, ldw, P.as<u32>(), s);
}

// Radius-mask tiles (threshold `radius`, or *thr_dev when given) between the
// events ev0 / ev1 (may be null); the mask is sized for n rows.
void radius_mask_tiles(int metric, const u32 *dmat, u32 n, u32 ldw, const Planes &planes, int32_t radius,
                       const int32_t *thr_dev, RadiusOut &o, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1,
                       const u32 *stage, const u32 *estage = nullptr)
{
    const u32 nt = ceil_div(n, TILE);
    const u32 mstride = nt * 4;
    o.n = n;
    o.mask.reserve((size_t)nt * TILE * mstride * 4);
    const bool mfma = metric == M_AND && planes.nplanes == 0 && mfma_enabled();
    if (mfma && !stage) {
        make_stage(dmat, n, ldw, o.stage, s);
        stage = o.stage.as<u32>();
    }
    if (ev0) CASPHY_HIP(hipEventRecord(ev0, s));
    if (mfma) {
        MfmaArgs m;
        m.PA = m.PB = stage; m.padA = m.padB = stage_rows(n); m.stages = ldw / 4;
        m.EA = m.EB = estage;
        m.nA = n; m.nB = n;
        m.thr = radius; m.thr_dev = thr_dev; m.mask = o.mask.as<u32>(); m.mstride = mstride;
        launch_mfma_tiles(E_MASK, m, s);
    } else {
        TileArgs a;
        a.A = dmat; a.nA = n; a.B = dmat; a.nB = n;
        a.ldw = ldw; a.kw = ldw; a.ntA = nt; a.ntB = nt;
        a.triangle = 1;
        a.ntiles = (u32)((u64)nt * (nt + 1) / 2);
        a.thr = radius; a.thr_dev = thr_dev; a.mask = o.mask.as<u32>(); a.mstride = mstride;
        apply_planes(planes, a, o.pbuf, s);
        launch_tiles(metric, E_MASK, a, s);
    }
    if (ev1) CASPHY_HIP(hipEventRecord(ev1, s));
}

// Per-row counts and u64 offsets (offsets[n] = total) of the mask of n rows;
// thr_dev: the mask launch's device threshold (INT32_MAX: it scored nothing).
void radius_mask_counts(u32 n, const int32_t *thr_dev, RadiusOut &o, hipStream_t s)
{
    const u32 mstride = ceil_div(n, TILE) * 4;
    o.counts.reserve((size_t)n * 4);
    o.offsets.reserve((size_t)(n + 1) * 8);
    launch_row_popc(o.mask.as<u32>(), n, mstride, mstride, o.counts.as<u32>(), s, thr_dev);
    launch_scan_counts(o.counts.as<u32>(), n, o.offsets.as<u64>(), s);
}

// Radius mask, per-row counts and offsets on the stream; nothing is read back.
void radius_mask(int metric, const u32 *dmat, u32 n, u32 ldw, const Planes &planes, int32_t radius,
                 const int32_t *thr_dev, RadiusOut &o, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1,
                 const u32 *stage, const u32 *estage = nullptr)
{
    radius_mask_tiles(metric, dmat, n, ldw, planes, radius, thr_dev, o, s, ev0, ev1, stage, estage);
    radius_mask_counts(n, thr_dev, o, s);
}

void radius_compute(int metric, const u32 *dmat, u32 n, u32 ldw, const Planes &planes, int32_t radius,
                    RadiusOut &o, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, const u32 *stage,
                    const u32 *estage)
{
    radius_mask(metric, dmat, n, ldw, planes, radius, nullptr, o, s, ev0, ev1, stage, estage);
    const u32 mstride = ceil_div(n, TILE) * 4;
    d2h(&o.total, o.offsets.as<u64>() + n, 8, s);
    o.idx.reserve(std::max<u64>(o.total, 1) * 4);
    launch_compact(o.mask.as<u32>(), n, mstride, mstride, o.offsets.as<u64>(), o.idx.as<u32>(), s);
}

// Largest sum S of R tryRadius counts (each <= n) whose mean passes the
// autoSplit stopping rule log10(max(S / R, 1e-30)) <= bound, evaluated with
// the host's log10 (monotone in S) so the device comparison sums[k] <= smax
// decides exactly as the host rule would; any = 0 when even S = 0 fails.
void autosplit_limit(u32 R, u32 n, double bound, u64 &smax, int &any)
{
    auto pass = [&](u64 S) { return std::log10(std::max((double)S / (double)R, 1e-30)) <= bound; };
    const u64 top = (u64)R * n;
    smax = 0;
    any = pass(0) ? 1 : 0;
    if (!any) return;
    if (pass(top)) {
        smax = top;
        return;
    }
    u64 lo = 0, hi = top;  // pass(lo), !pass(hi)
    while (hi - lo > 1) {
        const u64 mid = lo + (hi - lo) / 2;
        (pass(mid) ? lo : hi) = mid;
    }
    smax = lo;
}

void radius_lists(int metric, const u32 *dmat, u32 n, u32 ldw, const Planes &planes, int32_t radius,
                  u32 *nneighbor, u32 **idx, uint64_t *nidx, hipStream_t s)
{
    RadiusOut o;
    radius_compute(metric, dmat, n, ldw, planes, radius, o, s);
    u32 *host = (u32 *)std::malloc(std::max<u64>(o.total, 1) * 4);
    CASPHY_REQUIRE(host, CASPHY_ENOMEM, "host allocation of the neighbour list failed");
    try {
        d2h(host, o.idx.p, o.total * 4, s);
        d2h(nneighbor, o.counts.p, (size_t)n * 4, s);
    } catch (...) {
        std::free(host);
        throw;
    }
    *idx = host;
    *nidx = o.total;
}

// Few queries over wide unit-weight bit rows take the scan form (one HBM pass);
// CASPHY_SCAN=0 keeps the tiles.
bool use_scan(int metric, const Planes &planes, u32 R, u32 ldw, u32 nradii)
{
    const char *e = std::getenv("CASPHY_SCAN");
    return metric == M_AND && planes.nplanes == 0 && R <= SCAN_Q && nradii <= MAX_RADII_K && scan_fits(ldw) &&
           !(e && e[0] == '0');
}

// r_max - r_min + 1 of ascending radii (the count table's span).
u64 counts_span(const int32_t *radii, u32 nradii)
{
    return (u64)((int64_t)radii[nradii - 1] - radii[0] + 1);
}

// tryRadius: query x cell scores (matrix-core tiles for unit-weight bits,
// VALU tiles otherwise) followed by one counting workgroup per query; on the
// fp4 tiles the counts come straight from the tiles when the radii fit the
// count table.
void try_radii(int metric, const u32 *dmat, u32 n, u32 ldw, const Planes &planes, const int32_t *radii,
               u32 nradii, const u32 *q, u32 R, u32 *counts, hipStream_t s, const u32 *stage, hipEvent_t ev0,
               hipEvent_t ev1, const u32 *estage)
{
    for (u32 k = 0; k < R; ++k)
        CASPHY_REQUIRE(q[k] < n, CASPHY_EARG, "query row " + std::to_string(q[k]) + " out of range");
    for (u32 k = 1; k < nradii; ++k)
        CASPHY_REQUIRE(radii[k - 1] <= radii[k], CASPHY_EARG, "radii must be ascending");
    DevBuf dq((size_t)R * 4), dc((size_t)R * nradii * 4), dr((size_t)nradii * 4), dS, pbuf;
    h2d(dq.p, q, (size_t)R * 4, s);
    h2d(dr.p, radii, (size_t)nradii * 4, s);
    const u32 ntA = ceil_div(R, TILE), ntB = ceil_div(n, TILE);
    if (use_scan(metric, planes, R, ldw, nradii)) {
        // few queries over wide rows: one HBM-bound pass over the cells
        DevBuf qr((size_t)R * ldw * 4);
        for (u32 k = 0; k < R; ++k)
            CASPHY_HIP(hipMemcpyAsync(qr.as<u32>() + (size_t)k * ldw, dmat + (size_t)q[k] * ldw, (size_t)ldw * 4,
                                      hipMemcpyDeviceToDevice, s));
        DevBuf part((size_t)scan_grid(n) * 64 * 4);
        launch_scan_bit(dmat, n, ldw, qr.as<u32>(), R, dr.as<int32_t>(), nradii, dc.as<u32>(), R, nullptr,
                        part.as<u32>(), s, ev0, ev1);
        d2h(counts, dc.p, (size_t)R * nradii * 4, s);
        return;
    }
    if (metric == M_AND && planes.nplanes == 0 && mfma_enabled()) {
        DevBuf own, sq((size_t)stage_rows(R) * ldw * 4);
        if (!stage) {
            make_stage(dmat, n, ldw, own, s);
            stage = own.as<u32>();
        }
        launch_to_stage_major(dmat, R, ldw, ldw, sq.as<u32>(), s, dq.as<u32>());
        MfmaArgs m;
        m.PA = sq.as<u32>(); m.PB = stage; m.padA = stage_rows(R); m.padB = stage_rows(n); m.stages = ldw / 4;
        m.EB = estage;
        m.nA = R; m.nB = n; m.triangle = 0;
        // counts straight from the tiles (no score matrix) when the radii fit
        // the count table
        MfmaArgs f = m;
        f.EA = f.EB = sq.as<u32>();  // stand-ins: only their presence is checked
        const bool fused = fused_counts_ok(f, nradii) && counts_span(radii, nradii) <= COUNT_SPAN_MAX;
        if (fused) {
            const u32 stages = ldw / 4, ntB = ceil_div(n, TILE);
            DevBuf ea(fp4_bytes(m.padA, stages)), eb, tab(COUNT_SPAN_MAX + 64 + 16),
                cpart((size_t)ntB * nradii * m.padA);
            launch_expand_fp4(sq.as<u32>(), m.padA, stages, ea.as<u32>(), s);
            m.EA = ea.as<u32>();
            if (!m.EB) {
                eb.alloc(fp4_bytes(m.padB, stages));
                launch_expand_fp4(stage, m.padB, stages, eb.as<u32>(), s);
                m.EB = eb.as<u32>();
            }
            unsigned char *t = tab.as<unsigned char>();
            m.ctab = t; m.cneed = t + COUNT_SPAN_MAX; m.cpar = reinterpret_cast<const int32_t *>(t + COUNT_SPAN_MAX + 64);
            m.nrad = nradii; m.cpart = cpart.as<unsigned char>();
            launch_count_table(dr.as<int32_t>(), nradii, t, t + COUNT_SPAN_MAX,
                               reinterpret_cast<int32_t *>(t + COUNT_SPAN_MAX + 64), s);
            if (ev0) CASPHY_HIP(hipEventRecord(ev0, s));
            launch_mfma_counts(m, dc.as<u32>(), R, s);
            if (ev1) CASPHY_HIP(hipEventRecord(ev1, s));
            d2h(counts, dc.p, (size_t)R * nradii * 4, s);
            return;
        }
        dS.alloc((size_t)R * n * 4);
        m.S = dS.as<int32_t>(); m.ldS = n;
        if (ev0) CASPHY_HIP(hipEventRecord(ev0, s));
        launch_mfma_tiles(E_SCORES, m, s);
        if (ev1) CASPHY_HIP(hipEventRecord(ev1, s));
        launch_count_radii(dS.as<int32_t>(), R, n, n, dr.as<int32_t>(), nradii, dc.as<u32>(), s);
        d2h(counts, dc.p, (size_t)R * nradii * 4, s);
        return;
    }
    dS.alloc((size_t)R * n * 4);
    TileArgs a;
    a.A = dmat; a.Aidx = dq.as<u32>(); a.nA = R; a.B = dmat; a.nB = n;
    a.ldw = ldw; a.kw = ldw;
    a.ntA = ntA; a.ntB = ntB;
    a.ntiles = ntA * ntB;
    a.S = dS.as<int32_t>(); a.ldS = n;
    apply_planes(planes, a, pbuf, s);
    if (ev0) CASPHY_HIP(hipEventRecord(ev0, s));
    launch_tiles(metric, E_SCORES, a, s);
    if (ev1) CASPHY_HIP(hipEventRecord(ev1, s));
    launch_count_radii(dS.as<int32_t>(), R, n, n, dr.as<int32_t>(), nradii, dc.as<u32>(), s);
    d2h(counts, dc.p, (size_t)R * nradii * 4, s);
}

namespace {
void require(bool c, const char *msg) { CASPHY_REQUIRE(c, CASPHY_EARG, msg); }

}  // namespace
}  // namespace casphy

using namespace casphy;

extern "C" {

const char *casphy_last_error(void) { return g_last_error.c_str(); }

int casphy_version(void) { return 10000; }

void casphy_free(void *p) { std::free(p); }

int casphy_set_device(int device)
{
    return guarded([&] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
            throw Error(CASPHY_ENODEV, "no HIP device available");
        require(device >= 0 && device < count, "device index out of range");
        g_device = device;
    });
}

int casphy_synchronize(void)
{
    return guarded([&] { CASPHY_HIP(hipStreamSynchronize(thread_stream())); });
}

int casphy_mfma_tuning(char *buf, size_t cap, size_t *len)
{
    return guarded([&] {
        require(len && (buf || cap == 0), "casphy_mfma_tuning: bad arguments");
        std::string r = mfma_tuning_report();
        const std::string v = valu_tuning_json();
        if (v.size() > 2) {  // both lists in one
            r.pop_back();
            if (r.size() > 1) r += ",";
            r += v.substr(1);
        }
        *len = r.size();
        if (cap) {
            const size_t k = std::min(cap - 1, r.size());
            std::memcpy(buf, r.data(), k);
            buf[k] = 0;
        }
    });
}

int casphy_score_bit_matrix(const uint32_t *mat, uint32_t n, uint32_t W, const double *weight, int32_t *S)
{
    return guarded([&] {
        require(mat && S && n > 0 && W > 0, "casphy_score_bit_matrix: bad arguments");
        hipStream_t s = thread_stream();
        BitMat m = upload_bit(mat, W, n, false, s);
        Planes planes = make_planes(weight, 32 * W, m.ldw);
        DevBuf dS((size_t)n * n * 4), pbuf, stage;
        const u32 nt = ceil_div(n, TILE);
        if (planes.nplanes == 0 && mfma_enabled()) {
            make_stage(m.d.as<u32>(), n, m.ldw, stage, s);
            MfmaArgs q;
            q.PA = q.PB = stage.as<u32>(); q.padA = q.padB = stage_rows(n); q.stages = m.ldw / 4;
            q.nA = n; q.nB = n;
            q.S = dS.as<int32_t>(); q.ldS = n;
            launch_mfma_tiles(E_SCORES, q, s);
        } else {
            TileArgs a;
            a.A = m.d.as<u32>(); a.nA = n; a.B = a.A; a.nB = n;
            a.ldw = m.ldw; a.kw = m.ldw;
            a.ntA = a.ntB = nt;
            a.triangle = 1;
            a.ntiles = (u32)((u64)nt * (nt + 1) / 2);
            a.S = dS.as<int32_t>(); a.ldS = n;
            apply_planes(planes, a, pbuf, s);
            launch_tiles(M_AND, E_SCORES, a, s);
        }
        d2h(S, dS.p, (size_t)n * n * 4, s);
    });
}

int casphy_radius_bit(const uint32_t *mat_wordmajor, uint32_t W, uint32_t n, const double *weight,
                      int32_t radius, uint32_t *nneighbor, uint32_t **idx, uint64_t *nidx)
{
    return guarded([&] {
        require(mat_wordmajor && nneighbor && idx && nidx && n > 0 && W > 0,
                "casphy_radius_bit: bad arguments");
        hipStream_t s = thread_stream();
        BitMat m = upload_bit(mat_wordmajor, W, n, true, s);
        Planes planes = make_planes(weight, 32 * W, m.ldw);
        radius_lists(M_AND, m.d.as<u32>(), n, m.ldw, planes, radius, nneighbor, idx, nidx, s);
    });
}

int casphy_radius_dna(const uint32_t *mat, uint32_t n, uint32_t nwRaw, int32_t radius,
                      uint32_t *nneighbor, uint32_t **idx, uint64_t *nidx)
{
    return guarded([&] {
        require(mat && nneighbor && idx && nidx && n > 0 && nwRaw > 0, "casphy_radius_dna: bad arguments");
        hipStream_t s = thread_stream();
        DnaMat m = upload_dna(mat, n, nwRaw, s);
        radius_lists(m.metric, m.d.as<u32>(), n, m.ldw, Planes{}, radius, nneighbor, idx, nidx, s);
    });
}

int casphy_try_radii_bit(const uint32_t *mat_wordmajor, uint32_t W, uint32_t n, const double *weight,
                         const int32_t *radii, uint32_t nradii, const uint32_t *q, uint32_t R,
                         uint32_t *counts)
{
    return guarded([&] {
        require(mat_wordmajor && radii && q && counts && n > 0 && W > 0, "casphy_try_radii_bit: bad arguments");
        require(nradii >= 1 && nradii <= MAX_ROUND_RADII, "casphy_try_radii_bit: 1..64 radii");
        if (R == 0) return;
        hipStream_t s = thread_stream();
        BitMat m = upload_bit(mat_wordmajor, W, n, true, s);
        Planes planes = make_planes(weight, 32 * W, m.ldw);
        try_radii(M_AND, m.d.as<u32>(), n, m.ldw, planes, radii, nradii, q, R, counts, s);
    });
}

int casphy_try_radius_bit(const uint32_t *mat_wordmajor, uint32_t W, uint32_t n, const double *weight,
                          int32_t radius, const uint32_t *q, uint32_t R, uint32_t *counts)
{
    return casphy_try_radii_bit(mat_wordmajor, W, n, weight, &radius, 1, q, R, counts);
}

int casphy_try_radii_dna(const uint32_t *mat, uint32_t n, uint32_t nwRaw, const int32_t *radii,
                         uint32_t nradii, const uint32_t *q, uint32_t R, uint32_t *counts)
{
    return guarded([&] {
        require(mat && radii && q && counts && n > 0 && nwRaw > 0, "casphy_try_radii_dna: bad arguments");
        require(nradii >= 1 && nradii <= MAX_ROUND_RADII, "casphy_try_radii_dna: 1..64 radii");
        if (R == 0) return;
        hipStream_t s = thread_stream();
        DnaMat m = upload_dna(mat, n, nwRaw, s);
        try_radii(m.metric, m.d.as<u32>(), n, m.ldw, Planes{}, radii, nradii, q, R, counts, s);
    });
}

int casphy_try_radius_dna(const uint32_t *mat, uint32_t n, uint32_t nwRaw, int32_t radius,
                          const uint32_t *q, uint32_t R, uint32_t *counts)
{
    return casphy_try_radii_dna(mat, n, nwRaw, &radius, 1, q, R, counts);
}

int casphy_transfer(const uint32_t *Y, uint32_t ny, const uint32_t *X, uint32_t nx, uint32_t ncol,
                    uint32_t ntip, uint32_t *ymin, uint32_t *ysize, uint32_t *xmin, uint32_t *xsize)
{
    return guarded([&] {
        require(Y && X && ymin && ysize && xmin && xsize && ncol > 0, "casphy_transfer: bad arguments");
        require(ntip <= 32u * ncol, "casphy_transfer: ntip exceeds 32*ncol");
        hipStream_t s = thread_stream();
        const u32 init = 32u * ncol;
        DevBuf dy, dx, dymin((size_t)std::max(ny, 1u) * 4), dxmin((size_t)std::max(nx, 1u) * 4);
        DevBuf dys((size_t)std::max(ny, 1u) * 4), dxs((size_t)std::max(nx, 1u) * 4);
        BitMat my, mx;
        if (ny) my = upload_bit(Y, ncol, ny, false, s);
        if (nx) mx = upload_bit(X, ncol, nx, false, s);
        launch_fill_u32(dymin.as<u32>(), ny, init, s);
        launch_fill_u32(dxmin.as<u32>(), nx, init, s);
        if (ny && nx && mfma_enabled()) {
            // H = popc(x) + popc(y) - 2 <x, y> with the dot product on the matrix cores
            DevBuf sx, sy, px((size_t)nx * 4), py((size_t)ny * 4);
            make_stage(mx.d.as<u32>(), nx, mx.ldw, sx, s);
            make_stage(my.d.as<u32>(), ny, my.ldw, sy, s);
            launch_row_popc(mx.d.as<u32>(), nx, mx.ldw, mx.ldw, px.as<u32>(), s);
            launch_row_popc(my.d.as<u32>(), ny, my.ldw, my.ldw, py.as<u32>(), s);
            MfmaArgs q;
            q.PA = sx.as<u32>(); q.PB = sy.as<u32>();
            q.padA = stage_rows(nx); q.padB = stage_rows(ny);
            q.stages = mx.ldw / 4; q.nA = nx; q.nB = ny;
            q.triangle = 0;
            q.ntip = ntip; q.popA = px.as<u32>(); q.popB = py.as<u32>();
            q.amin = dxmin.as<u32>(); q.bmin = dymin.as<u32>();
            launch_mfma_tiles(E_TBE, q, s);
            CASPHY_HIP(hipStreamSynchronize(s));
        } else if (ny && nx) {
            TileArgs a;
            a.A = mx.d.as<u32>(); a.nA = nx; a.B = my.d.as<u32>(); a.nB = ny;
            a.ldw = mx.ldw; a.kw = mx.ldw;
            a.ntA = ceil_div(nx, TILE); a.ntB = ceil_div(ny, TILE);
            a.ntiles = a.ntA * a.ntB;
            a.ntip = ntip; a.amin = dxmin.as<u32>(); a.bmin = dymin.as<u32>();
            launch_tiles(M_XOR, E_TBE, a, s);
        }
        if (ny) launch_row_size(my.d.as<u32>(), ny, my.ldw, ntip, dys.as<u32>(), s);
        if (nx) launch_row_size(mx.d.as<u32>(), nx, mx.ldw, ntip, dxs.as<u32>(), s);
        d2h(ymin, dymin.p, (size_t)ny * 4, s);
        d2h(ysize, dys.p, (size_t)ny * 4, s);
        d2h(xmin, dxmin.p, (size_t)nx * 4, s);
        d2h(xsize, dxs.p, (size_t)nx * 4, s);
    });
}

}  // extern "C"

namespace {
// Batched transfer distance on device bit rows: Y rows in my, the replicates'
// rows back to back in mx (replicate b: rows xoffsets[b] .. xoffsets[b+1]),
// minima starting from init = 32 * ncol; results to host buffers.
void transfer_batch_rows(const BitMat &my, const BitMat *mx, const uint32_t *xoffsets, uint32_t nrep, uint32_t ntip,
                         uint32_t init, uint32_t *ymin, uint32_t *ysize, uint32_t *xmin, uint32_t *xsize,
                         hipStream_t s)
{
    const u32 ny = my.n, nx = nrep ? xoffsets[nrep] : 0;
    // every replicate starts on a fresh 128-row tile so that a tile belongs to one replicate
    std::vector<u32> rowidx, rowmap, rep_of_tile;
    for (u32 b = 0; b < nrep; ++b) {
        const u32 cnt = xoffsets[b + 1] - xoffsets[b];
        const u32 padded = ceil_div(cnt, TILE) * TILE;
        for (u32 r = 0; r < padded; ++r) {
            rowidx.push_back(r < cnt ? xoffsets[b] + r : 0xFFFFFFFFu);
            rowmap.push_back(r < cnt ? xoffsets[b] + r : 0xFFFFFFFFu);
        }
        for (u32 t = 0; t < padded / TILE; ++t) rep_of_tile.push_back(b);
    }
    const u32 npad = (u32)rowidx.size();
    DevBuf dymin((size_t)nrep * ny * 4 + 4), dys((size_t)ny * 4), dxmin((size_t)std::max(nx, 1u) * 4),
        dxs((size_t)std::max(nx, 1u) * 4), py((size_t)ny * 4), sy;
    launch_fill_u32(dymin.as<u32>(), (u64)nrep * ny, init, s);
    launch_fill_u32(dxmin.as<u32>(), nx, init, s);
    launch_row_size(my.d.as<u32>(), ny, my.ldw, ntip, dys.as<u32>(), s);
    if (nx) {
        launch_row_size(mx->d.as<u32>(), nx, mx->ldw, ntip, dxs.as<u32>(), s);
        DevBuf didx((size_t)npad * 4), dmap((size_t)npad * 4), drep(rep_of_tile.size() * 4),
            sx((size_t)stage_rows(npad) * mx->ldw * 4), px((size_t)npad * 4);
        h2d(didx.p, rowidx.data(), (size_t)npad * 4, s);
        h2d(dmap.p, rowmap.data(), (size_t)npad * 4, s);
        h2d(drep.p, rep_of_tile.data(), rep_of_tile.size() * 4, s);
        launch_to_stage_major(mx->d.as<u32>(), npad, mx->ldw, mx->ldw, sx.as<u32>(), s, didx.as<u32>());
        make_stage(my.d.as<u32>(), ny, my.ldw, sy, s);
        launch_stage_popc(sx.as<u32>(), npad, stage_rows(npad), mx->ldw, px.as<u32>(), s);  // padding rows: 0
        launch_row_popc(my.d.as<u32>(), ny, my.ldw, my.ldw, py.as<u32>(), s);
        MfmaArgs q;
        q.PA = sx.as<u32>(); q.PB = sy.as<u32>();
        q.padA = stage_rows(npad); q.padB = stage_rows(ny);
        q.stages = mx->ldw / 4; q.nA = npad; q.nB = ny;
        q.triangle = 0;
        q.ntip = ntip; q.popA = px.as<u32>(); q.popB = py.as<u32>();
        q.amin = dxmin.as<u32>(); q.bmin = dymin.as<u32>();
        q.arow_map = dmap.as<u32>(); q.tile_rep = drep.as<u32>(); q.bstride = ny;
        launch_mfma_tiles(E_TBE, q, s);
        d2h(xmin, dxmin.p, (size_t)nx * 4, s);
        d2h(xsize, dxs.p, (size_t)nx * 4, s);
    }
    d2h(ymin, dymin.p, (size_t)nrep * ny * 4, s);
    d2h(ysize, dys.p, (size_t)ny * 4, s);
}

// Replicates in consecutive groups whose rows fit the device budget of the
// bitset pair tiles (about 7 row copies per replicate row: upload, padded,
// stage-major and its fp4 expansion); CASPHY_TBE_BUDGET_MB (default 8,192)
// sets it.  A replicate larger than the budget forms a group of its own.
std::vector<std::pair<u32, u32>> replicate_groups(const u32 *xoffsets, u32 nrep, u32 ldw)
{
    const char *e = std::getenv("CASPHY_TBE_BUDGET_MB");
    const u64 budget = (e ? (u64)std::max(1, std::atoi(e)) : 8192ull) << 20;
    const u64 per_row = 7ull * ldw * 4;
    std::vector<std::pair<u32, u32>> g;
    u32 b0 = 0;
    while (b0 < nrep) {
        u32 b1 = b0 + 1;
        while (b1 < nrep && (u64)(xoffsets[b1 + 1] - xoffsets[b0]) * per_row <= budget) ++b1;
        g.emplace_back(b0, b1);
        b0 = b1;
    }
    return g;
}

// Clade rows of a tree in a BitMat (rows of ldw words, zero padding).
BitMat clade_mat(const CladePlan &p, hipStream_t s)
{
    BitMat m;
    m.n = (u32)p.lo.size(); m.W = p.nbin; m.ldw = round_up(std::max<u32>(p.nbin, 1), KC_BIT);
    m.d.alloc(std::max<size_t>((size_t)m.n * m.ldw * 4, 4));
    build_clade_rows(p, m.d.as<u32>(), m.ldw, s);
    return m;
}
}  // namespace

extern "C" {

int casphy_transfer_batch(const uint32_t *Y, uint32_t ny, const uint32_t *X, const uint32_t *xoffsets, uint32_t nrep,
                          uint32_t ncol, uint32_t ntip, uint32_t *ymin, uint32_t *ysize, uint32_t *xmin,
                          uint32_t *xsize)
{
    return guarded([&] {
        require(Y && X && xoffsets && ymin && ysize && ncol > 0 && ny > 0, "casphy_transfer_batch: bad arguments");
        require(ntip <= 32u * ncol, "casphy_transfer_batch: ntip exceeds 32*ncol");
        require(xoffsets[0] == 0, "casphy_transfer_batch: xoffsets[0] must be 0");
        for (u32 b = 0; b < nrep; ++b)
            require(xoffsets[b + 1] >= xoffsets[b], "casphy_transfer_batch: xoffsets must be ascending");
        const u32 nx = nrep ? xoffsets[nrep] : 0;
        require(nx == 0 || (xmin && xsize), "casphy_transfer_batch: bad arguments");
        hipStream_t s = thread_stream();
        BitMat my = upload_bit(Y, ncol, ny, false, s);
        if (!nrep) {
            transfer_batch_rows(my, nullptr, xoffsets, 0, ntip, 32u * ncol, ymin, ysize, xmin, xsize, s);
            return;
        }
        // replicates stream through the device in groups; Y stays resident
        for (const auto &g : replicate_groups(xoffsets, nrep, my.ldw)) {
            const u32 r0 = xoffsets[g.first], r1 = xoffsets[g.second];
            std::vector<u32> xo(g.second - g.first + 1);
            for (u32 b = g.first; b <= g.second; ++b) xo[b - g.first] = xoffsets[b] - r0;
            u32 *ym = ymin + (size_t)g.first * ny;
            if (r1 > r0) {
                BitMat mx = upload_bit(X + (size_t)r0 * ncol, ncol, r1 - r0, false, s);
                transfer_batch_rows(my, &mx, xo.data(), g.second - g.first, ntip, 32u * ncol, ym, ysize, xmin + r0,
                                    xsize + r0, s);
            } else {
                transfer_batch_rows(my, nullptr, xo.data(), g.second - g.first, ntip, 32u * ncol, ym, ysize, nullptr,
                                    nullptr, s);
            }
        }
    });
}

int casphy_transfer_trees(const int32_t *y_anc, const int32_t *y_des, uint32_t y_nedge, uint32_t y_nnode,
                          const int32_t *x_anc, const int32_t *x_des, const uint32_t *x_edge_offsets,
                          const uint32_t *x_nnode, uint32_t nrep, uint32_t ntip, uint32_t *ymin, uint32_t *ysize,
                          uint32_t *xmin, uint32_t *xsize)
{
    return guarded([&] {
        require(y_anc && y_des && ymin && ysize && (nrep == 0 || (x_anc && x_des && x_edge_offsets && x_nnode)),
                "casphy_transfer_trees: bad arguments");
        require(y_nnode >= 2, "casphy_transfer_trees: the reference tree needs a non-root internal node");
        hipStream_t s = thread_stream();
        // default: the tree walks (no bitsets); CASPHY_TBE=bitsets builds the
        // clade bitsets on the device and scores them on the pair tiles
        const char *form = std::getenv("CASPHY_TBE");
        if (!(form && !std::strcmp(form, "bitsets"))) {
            if (nrep) {
                u64 nx = 0;
                for (u32 b = 0; b < nrep; ++b) nx += x_nnode[b] > 0 ? x_nnode[b] - 1 : 0;
                require(nx == 0 || (xmin && xsize), "casphy_transfer_trees: bad arguments");
            }
            transfer_trees_dp(y_anc, y_des, y_nedge, y_nnode, x_anc, x_des, x_edge_offsets, x_nnode, nrep, ntip,
                              ymin, ysize, xmin, xsize, s);
            return;
        }
        const CladePlan py = plan_clades(y_anc, y_des, y_nedge, y_nnode, ntip);
        std::vector<CladePlan> px(nrep);
        std::vector<u32> xoff(nrep + 1, 0);
        for (u32 b = 0; b < nrep; ++b) {
            require(x_edge_offsets[b + 1] >= x_edge_offsets[b], "casphy_transfer_trees: edge offsets must ascend");
            px[b] = plan_clades(x_anc + x_edge_offsets[b], x_des + x_edge_offsets[b],
                                x_edge_offsets[b + 1] - x_edge_offsets[b], x_nnode[b], ntip);
            xoff[b + 1] = xoff[b] + (u32)px[b].lo.size();
        }
        require(xoff[nrep] == 0 || (xmin && xsize), "casphy_transfer_trees: bad arguments");
        BitMat my = clade_mat(py, s);
        require(my.n > 0, "casphy_transfer_trees: the reference tree has no clade rows");
        if (!nrep) {
            transfer_batch_rows(my, nullptr, xoff.data(), 0, ntip, 32u * py.nbin, ymin, ysize, xmin, xsize, s);
            return;
        }
        // replicates' clade rows are built on the device group by group
        for (const auto &g : replicate_groups(xoff.data(), nrep, my.ldw)) {
            const u32 r0 = xoff[g.first], r1 = xoff[g.second];
            std::vector<u32> xo(g.second - g.first + 1);
            for (u32 b = g.first; b <= g.second; ++b) xo[b - g.first] = xoff[b] - r0;
            u32 *ym = ymin + (size_t)g.first * my.n;
            if (r1 > r0) {
                BitMat mx;
                mx.n = r1 - r0; mx.W = py.nbin; mx.ldw = my.ldw;
                mx.d.alloc((size_t)mx.n * mx.ldw * 4);
                for (u32 b = g.first; b < g.second; ++b)
                    build_clade_rows(px[b], mx.d.as<u32>() + (size_t)(xoff[b] - r0) * mx.ldw, mx.ldw, s);
                transfer_batch_rows(my, &mx, xo.data(), g.second - g.first, ntip, 32u * py.nbin, ym, ysize,
                                    xmin + r0, xsize + r0, s);
            } else {
                transfer_batch_rows(my, nullptr, xo.data(), g.second - g.first, ntip, 32u * py.nbin, ym, ysize,
                                    nullptr, nullptr, s);
            }
        }
    });
}

int casphy_tip_merge(const uint32_t *nneighbor, const uint32_t *idx, uint32_t ntip, double p, int32_t *member)
{
    return guarded([&] {
        require(ntip == 0 || (nneighbor && member), "casphy_tip_merge: bad arguments");
        std::vector<u64> off((size_t)ntip + 1, 0);
        for (u32 i = 0; i < ntip; ++i) {
            CASPHY_REQUIRE(nneighbor[i] <= TIP_MERGE_MAX_LIST, CASPHY_EARG,
                           "casphy_tip_merge: tip " + std::to_string(i) + " has more than 1,048,576 neighbours");
            off[i + 1] = off[i] + nneighbor[i];
        }
        const u64 total = off[ntip];
        require(total == 0 || idx, "casphy_tip_merge: bad arguments");
        for (u32 i = 0; i < ntip; ++i)
            for (u64 k = off[i]; k < off[i + 1]; ++k) {
                CASPHY_REQUIRE(idx[k] < ntip, CASPHY_EARG, "casphy_tip_merge: neighbour id out of range in list " +
                                                               std::to_string(i));
                CASPHY_REQUIRE(k == off[i] || idx[k - 1] < idx[k], CASPHY_EARG,
                               "casphy_tip_merge: list " + std::to_string(i) + " is not strictly ascending");
            }
        if (ntip == 0) return;
        hipStream_t s = thread_stream();
        DevBuf doff(((size_t)ntip + 1) * 8), dids(std::max<u64>(total, 1) * 4);
        h2d(doff.p, off.data(), ((size_t)ntip + 1) * 8, s);
        if (total) h2d(dids.p, idx, total * 4, s);
        tip_merge_device(doff.as<u64>(), dids.as<u32>(), ntip, nullptr, 0, p, member, s);
    });
}

}  // extern "C"

namespace {
int node_keep_entry(bool indel, const char *name, const uint32_t *edge, uint32_t nedge, const uint32_t *state,
                    uint32_t nnode, uint32_t width, const uint32_t *outmat, uint32_t nout, uint32_t cutoff,
                    uint32_t *kept, uint32_t *nkept)
{
    return guarded([&] {
        const std::string fn(name);
        CASPHY_REQUIRE(nkept && (nedge == 0 || edge) && (nnode == 0 || kept), CASPHY_EARG, fn + ": bad arguments");
        CASPHY_REQUIRE(((size_t)nnode * width == 0 || state) && ((size_t)nout * width == 0 || outmat), CASPHY_EARG,
                       fn + ": bad arguments");
        const u64 ntip = (u64)nnode + 1;
        std::vector<unsigned char> seen(nnode, 0);
        for (u32 i = 0; i < nedge; ++i) {
            const u32 a = edge[2 * i], d = edge[2 * i + 1];
            CASPHY_REQUIRE(a < ntip + nnode && d < ntip + nnode, CASPHY_EARG,
                           fn + ": edge " + std::to_string(i) + " names a node beyond the tree");
            if (d >= ntip) {
                CASPHY_REQUIRE(!seen[d - ntip], CASPHY_EARG,
                               fn + ": internal node " + std::to_string(d) + " is the child of two edges");
                seen[d - ntip] = 1;
            }
        }
        *nkept = node_keep(indel, edge, nedge, state, nnode, width, outmat, nout, cutoff, kept, thread_stream());
    });
}
}  // namespace

extern "C" {

int casphy_node_keep_bit(const uint32_t *edge, uint32_t nedge, const uint32_t *node_state, uint32_t nnode,
                         uint32_t nbin, const uint32_t *outmat, uint32_t nout, uint32_t cutoff, uint32_t *kept,
                         uint32_t *nkept)
{
    return node_keep_entry(false, "casphy_node_keep_bit", edge, nedge, node_state, nnode, nbin, outmat, nout, cutoff,
                           kept, nkept);
}

int casphy_node_keep_indel(const uint32_t *edge, uint32_t nedge, const uint32_t *node_state, uint32_t nnode,
                           uint32_t ncol, const uint32_t *outmat, uint32_t nout, uint32_t cutoff, uint32_t *kept,
                           uint32_t *nkept)
{
    return node_keep_entry(true, "casphy_node_keep_indel", edge, nedge, node_state, nnode, ncol, outmat, nout,
                           cutoff, kept, nkept);
}

}  // extern "C"

namespace {
// Checks a root-vote call and returns the sample rows mat[idx[s]], gathered.
std::vector<u32> root_vote_samples(const char *fn, const uint32_t *state, uint32_t nnode, uint32_t nbin,
                                   const uint32_t *edge, uint32_t nedge, const uint32_t *mat, uint64_t nmat,
                                   const uint32_t *idx, uint32_t nsample, const void *root, const void *nshared)
{
    const std::string f(fn);
    CASPHY_REQUIRE(nnode >= 1, CASPHY_EARG, f + ": the tree has no node");
    CASPHY_REQUIRE(((size_t)nnode * nbin == 0 || state) && (nedge == 0 || edge), CASPHY_EARG, f + ": bad arguments");
    CASPHY_REQUIRE(nsample == 0 || (idx && root && nshared && (nbin == 0 || mat)), CASPHY_EARG,
                   f + ": bad arguments");
    std::vector<u32> rows((size_t)nsample * nbin);
    for (u32 k = 0; k < nsample; ++k) {
        CASPHY_REQUIRE(idx[k] < nmat, CASPHY_EARG,
                       f + ": sample " + std::to_string(k) + " names row " + std::to_string(idx[k]) + " of " +
                           std::to_string(nmat));
        if (nbin) std::memcpy(&rows[(size_t)k * nbin], mat + (size_t)idx[k] * nbin, (size_t)nbin * 4);
    }
    return rows;
}
}  // namespace

extern "C" {

int casphy_vote_root(const uint32_t *state, uint32_t nnode, uint32_t nbin, const uint32_t *edge, uint32_t nedge,
                     const uint32_t *mat, uint64_t nmat, const uint32_t *idx, uint32_t nsample, uint32_t *root,
                     uint32_t *nshared)
{
    return guarded([&] {
        const std::vector<u32> rows =
            root_vote_samples("casphy_vote_root", state, nnode, nbin, edge, nedge, mat, nmat, idx, nsample, root,
                              nshared);
        const RootVote plan = plan_vote_voteroot2(edge, nedge, nnode);
        std::vector<u64> keys(nsample);
        vote_root_keys(plan, state, nnode, nbin, rows.data(), nsample, keys.data(), thread_stream());
        for (u32 k = 0; k < nsample; ++k) {  // every node is a candidate: a key is never 0
            root[k] = plan.candidates[0xFFFFFFFFu - (u32)keys[k]];
            nshared[k] = (u32)(keys[k] >> 32);
        }
    });
}

int casphy_find_root(const uint32_t *state, uint32_t nnode, uint32_t nbin, const uint32_t *edge, uint32_t nedge,
                     const uint32_t *mat, uint64_t nmat, const uint32_t *idx, uint32_t nsample, int32_t *root,
                     int32_t *nshared)
{
    return guarded([&] {
        const std::vector<u32> rows =
            root_vote_samples("casphy_find_root", state, nnode, nbin, edge, nedge, mat, nmat, idx, nsample, root,
                              nshared);
        const RootVote plan = plan_vote_findroot(edge, nedge, nnode);
        std::vector<u64> keys(nsample);
        vote_root_keys(plan, state, nnode, nbin, rows.data(), nsample, keys.data(), thread_stream());
        for (u32 k = 0; k < nsample; ++k) {
            if (!keys[k]) {  // no edge: the reference's starting values
                root[k] = 0;
                nshared[k] = -1;
            } else {
                root[k] = (int32_t)plan.candidates[0xFFFFFFFFu - (u32)keys[k]] + 1;
                nshared[k] = (int32_t)(keys[k] >> 32);
            }
        }
    });
}

int casphy_fasta_encode(const uint8_t *seqs, uint32_t nseq, uint64_t len, uint8_t *is_variable, uint32_t *valid,
                        uint32_t *out, uint64_t *width)
{
    return guarded([&] {
        require(width && (nseq == 0 || valid) && (len == 0 || is_variable) &&
                    ((u64)nseq * len == 0 || (seqs && out)),
                "casphy_fasta_encode: bad arguments");
        if (!nseq && len) std::memset(is_variable, 0, len);
        *width = fasta_encode(seqs, nseq, len, is_variable, valid, out, thread_stream());
    });
}

int casphy_newick_read(const char *text, uint64_t len, const char *tip_order, uint64_t tip_order_bytes,
                       uint64_t *consumed, uint32_t *ntip, uint32_t *nnode, int32_t **edge, uint32_t *nedge,
                       char **labels, uint64_t *labels_bytes)
{
    return guarded([&] {
        require((text || !len) && consumed && ntip && nnode && edge && nedge && labels && labels_bytes &&
                    (tip_order || !tip_order_bytes),
                "casphy_newick_read: bad arguments");
        *edge = nullptr;
        *labels = nullptr;
        std::vector<std::string> order;
        if (tip_order) {
            require(tip_order_bytes > 0 && tip_order[tip_order_bytes - 1] == '\0',
                    "casphy_newick_read: the tip order must end with a NUL");
            for (uint64_t b = 0; b < tip_order_bytes;) {
                const std::string l(tip_order + b);
                b += l.size() + 1;
                order.push_back(l);
            }
        }
        u64 pos = 0;
        NewickTree t = read_newick(text, len, pos, tip_order ? &order : nullptr);
        *consumed = pos;
        *ntip = t.ntip;
        *nnode = t.nnode;
        *nedge = (u32)t.anc.size();
        *labels_bytes = 0;
        if (!t.ntip) return;
        int32_t *e = static_cast<int32_t *>(std::malloc(std::max<size_t>(t.anc.size() * 8, 8)));
        require(e != nullptr, "casphy_newick_read: out of host memory");
        std::copy(t.anc.begin(), t.anc.end(), e);  // R's edge matrix: the parent column, then the child column
        std::copy(t.des.begin(), t.des.end(), e + t.anc.size());
        size_t bytes = 0;
        for (auto &l : t.labels) bytes += l.size() + 1;
        char *lb = static_cast<char *>(std::malloc(std::max<size_t>(bytes, 1)));
        if (!lb) {
            std::free(e);
            require(false, "casphy_newick_read: out of host memory");
        }
        size_t at = 0;
        for (auto &l : t.labels) {
            std::memcpy(lb + at, l.c_str(), l.size() + 1);
            at += l.size() + 1;
        }
        *edge = e;
        *labels = lb;
        *labels_bytes = bytes;
    });
}

int casphy_tbe_support(const uint32_t *ymin, uint32_t nrep, uint32_t ny, const uint32_t *ysize, uint32_t ntip,
                       double *support)
{
    return guarded([&] {
        require((ny == 0 || (ysize && support)) && ((u64)nrep * ny == 0 || ymin), "casphy_tbe_support: bad arguments");
        tbe_support(ymin, nrep, ny, ysize, ntip, support);
    });
}

int casphy_clade_bitsets(const int32_t *anc, const int32_t *des, uint32_t nedge, uint32_t nnode, uint32_t ntip,
                         uint32_t *rows, uint32_t *nrows, int32_t *node_size)
{
    return guarded([&] {
        require(nrows && (nedge == 0 || (anc && des)) && (nnode < 2 || (rows && node_size)),
                "casphy_clade_bitsets: bad arguments");
        const CladePlan p = plan_clades(anc, des, nedge, nnode, ntip);
        const u32 nr = (u32)p.lo.size();
        *nrows = nr;
        if (!p.size.empty()) std::memcpy(node_size, p.size.data(), p.size.size() * 4);
        if (!nr) return;
        hipStream_t s = thread_stream();
        DevBuf drows((size_t)nr * p.nbin * 4);
        build_clade_rows(p, drows.as<u32>(), p.nbin, s);
        d2h(rows, drows.p, (size_t)nr * p.nbin * 4, s);
    });
}

int casphy_cladewise(const uint32_t *edge_raw, uint32_t n, uint32_t *edge_out, uint32_t *internal_order)
{
    return guarded([&] {
        require(edge_raw && edge_out && n >= 2, "casphy_cladewise: bad arguments");
        const u32 nedge = 2 * n - 2, nnode = 2 * n - 1;
        // edge rows grouped by parent, ascending row order within a parent
        std::vector<u32> start(nnode + 1, 0), rows(nedge);
        for (u32 e = 0; e < nedge; ++e) {
            const u32 p = edge_raw[2 * e], c = edge_raw[2 * e + 1];
            require(p < nnode && c < nnode, "casphy_cladewise: node id out of range");
            ++start[p + 1];
        }
        for (u32 v = 0; v < nnode; ++v) start[v + 1] += start[v];
        std::vector<u32> fillp(start.begin(), start.end() - 1);
        for (u32 e = 0; e < nedge; ++e) rows[fillp[edge_raw[2 * e]]++] = e;
        std::vector<u32> relabel(nnode, nnode), cursor(start.begin(), start.end() - 1), stack;
        for (u32 i = 0; i < n; ++i) relabel[i] = i;
        stack.reserve(nedge + 1);
        stack.push_back(nedge);  // the root is the last node created
        u32 next = n, out = 0;
        while (!stack.empty()) {
            const u32 top = stack.back();
            if (cursor[top] == start[top + 1]) {
                stack.pop_back();
                continue;
            }
            const u32 e = rows[cursor[top]++];
            const u32 p = edge_raw[2 * e], c = edge_raw[2 * e + 1];
            stack.push_back(c);
            if (relabel[p] == nnode) { internal_order[next - n] = p; relabel[p] = next++; }
            if (relabel[c] == nnode) { internal_order[next - n] = c; relabel[c] = next++; }
            require(out < nedge, "casphy_cladewise: edge list is not a tree");
            edge_out[2 * out] = relabel[p];
            edge_out[2 * out + 1] = relabel[c];
            ++out;
        }
        require(out == nedge, "casphy_cladewise: edge list is not a tree rooted at 2n-2");
    });
}

// ---------------------------------------------------------------- resident cells (struct in cells.hpp)
}  // extern "C"

namespace {

// Form tuner of resident ACGT cells: times the radius-mask tiles of the
// one-hot matrix-core form (after the tile autotuner has picked its variant
// for the shape) and of the 2-bit + mask VALU form over the first
// min(n, 8,192) cells, keeps the faster form and records both times in the
// tuning report.
void tune_dna_form(casphy_cells &h, DnaMat &m, hipStream_t s)
{
    const u32 k = std::min<u32>(h.n, 8192);
    RadiusOut o;
    DevBuf stage, estage;
    make_stage(h.mat.as<u32>(), k, h.ldw, stage, s);
    estage.alloc(fp4_bytes(stage_rows(k), h.ldw / 4));
    launch_expand_fp4(stage.as<u32>(), stage_rows(k), h.ldw / 4, estage.as<u32>(), s);
    hipEvent_t ev[2];
    CASPHY_HIP(hipEventCreate(&ev[0]));
    CASPHY_HIP(hipEventCreate(&ev[1]));
    float ms[2] = {0, 0};
    try {
        for (int form = 0; form < 2; ++form) {
            for (int rep = 0; rep < 2; ++rep) {  // the first run is untimed (variant tuning, code load)
                if (form == 0)
                    radius_mask(M_AND, h.mat.as<u32>(), k, h.ldw, Planes{}, 1, nullptr, o, s, ev[0], ev[1],
                                stage.as<u32>(), estage.as<u32>());
                else
                    radius_mask(M_DNA2, m.d2.as<u32>(), k, m.ldw2, Planes{}, 1, nullptr, o, s, ev[0], ev[1], nullptr);
            }
            CASPHY_HIP(hipEventSynchronize(ev[1]));
            CASPHY_HIP(hipEventElapsedTime(&ms[form], ev[0], ev[1]));
        }
    } catch (...) {
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        throw;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    const bool valu = ms[1] < ms[0];
    add_tuning_note("{\"dna_form\":\"" + std::string(valu ? "valu" : "mfma") + "\",\"cells\":" + std::to_string(k) +
                    ",\"positions\":" + std::to_string(8 * h.ldw) + ",\"ms\":{\"mfma\":" + std::to_string(ms[0]) +
                    ",\"valu\":" + std::to_string(ms[1]) + "}}");
    if (valu) {
        h.mat = std::move(m.d2);
        h.ldw = m.ldw2;
        h.metric = M_DNA2;
    }
    m.d2.release();
}

}  // namespace

extern "C" {

int casphy_cells_create(int kind, const uint32_t *mat, uint32_t n, uint32_t W, int wordmajor,
                        const double *weight, casphy_cells **out)
{
    return guarded([&] {
        require(mat && out && n > 0 && W > 0, "casphy_cells_create: bad arguments");
        require(kind == CASPHY_KIND_BIT || kind == CASPHY_KIND_DNA, "casphy_cells_create: bad kind");
        hipStream_t s = thread_stream();
        auto h = std::make_unique<casphy_cells>();
        h->kind = kind;
        h->n = n;
        h->device = thread_device();
        if (kind == CASPHY_KIND_BIT) {
            BitMat m = upload_bit(mat, W, n, wordmajor != 0, s);
            h->ldw = m.ldw;
            h->mat = std::move(m.d);
            h->metric = M_AND;
            h->planes = make_planes(weight, 32 * W, h->ldw);
        } else {
            require(!weight, "casphy_cells_create: DNA takes unit weights only");
            DnaMat m = upload_dna(mat, n, W, s, true);
            h->ldw = m.ldw;
            h->mat = std::move(m.d);
            h->metric = m.metric;
            if (m.d2.p) tune_dna_form(*h, m, s);
        }
        for (auto &e : h->ev) CASPHY_HIP(hipEventCreate(&e));
        *out = h.release();
    });
}

void casphy_cells_destroy(casphy_cells *h)
{
    if (!h) return;
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    h->drop_graph();
    if (h->pin) (void)hipHostFree(h->pin);
    delete h;
}

int casphy_cells_radius(casphy_cells *h, int32_t radius, uint64_t *nidx)
{
    return guarded([&] {
        require(h && nidx, "casphy_cells_radius: bad arguments");
        hipStream_t s = thread_stream();
        if (h->mfma()) h->staged(s);
        CASPHY_HIP(hipEventRecord(h->ev[2], s));
        h->timed[0] = h->timed[1] = true;
        radius_compute(h->metric, h->mat.as<u32>(), h->n, h->ldw, h->planes, radius, h->out, s, h->ev[0], h->ev[1],
                       h->mfma() ? h->stage.as<u32>() : nullptr, h->mfma() ? h->estage.as<u32>() : nullptr);
        CASPHY_HIP(hipEventRecord(h->ev[3], s));
        CASPHY_HIP(hipEventSynchronize(h->ev[3]));
        *nidx = h->out.total;
    });
}

int casphy_cells_fetch(casphy_cells *h, uint32_t *nneighbor, uint32_t *idx)
{
    return guarded([&] {
        require(h && nneighbor && idx && h->out.n == h->n, "casphy_cells_fetch: no radius result");
        hipStream_t s = thread_stream();
        d2h(nneighbor, h->out.counts.p, (size_t)h->n * 4, s);
        d2h(idx, h->out.idx.p, h->out.total * 4, s);
    });
}

int casphy_cells_tip_merge(casphy_cells *h, double p, int32_t *member)
{
    return guarded([&] {
        require(h && member && h->out.n == h->n, "casphy_cells_tip_merge: no radius result");
        hipStream_t s = thread_stream();
        const u32 n = h->n;
        std::vector<u32> nnb(n);
        d2h(nnb.data(), h->out.counts.p, (size_t)n * 4, s);
        for (u32 i = 0; i < n; ++i)
            CASPHY_REQUIRE(nnb[i] <= TIP_MERGE_MAX_LIST, CASPHY_EARG,
                           "casphy_cells_tip_merge: tip " + std::to_string(i) + " has more than 1,048,576 neighbours");
        // an autoSplit round in which no radius qualified leaves empty lists and
        // an unwritten mask: the dense form then builds its rows from the lists
        tip_merge_device(h->out.offsets.as<u64>(), h->out.idx.as<u32>(), n,
                         h->out.total ? h->out.mask.as<u32>() : nullptr, ceil_div(n, TILE) * 4, p, member, s);
    });
}

int casphy_cells_try_radii(casphy_cells *h, const int32_t *radii, uint32_t nradii, const uint32_t *q,
                           uint32_t R, uint32_t *counts)
{
    return guarded([&] {
        require(h && radii && q && counts, "casphy_cells_try_radii: bad arguments");
        require(nradii >= 1 && nradii <= MAX_ROUND_RADII, "casphy_cells_try_radii: 1..64 radii");
        if (R == 0) return;
        hipStream_t s = thread_stream();
        const bool tiles = h->mfma() && !use_scan(h->metric, h->planes, R, h->ldw, nradii);
        if (tiles) h->staged(s);
        h->timed[0] = true;
        try_radii(h->metric, h->mat.as<u32>(), h->n, h->ldw, h->planes, radii, nradii, q, R, counts, s,
                  tiles ? h->stage.as<u32>() : nullptr, h->ev[0], h->ev[1], tiles ? h->estage.as<u32>() : nullptr);
    });
}

int casphy_cells_autosplit(casphy_cells *h, const int32_t *radii, uint32_t nradii, const uint32_t *q, uint32_t R,
                           double bound, int32_t *radius, uint64_t *nidx, uint32_t *counts)
{
    return guarded([&] {
        require(h && radii && q && radius && nidx && R > 0, "casphy_cells_autosplit: bad arguments");
        require(nradii >= 1 && nradii <= MAX_ROUND_RADII, "casphy_cells_autosplit: 1..64 radii");
        const u32 n = h->n, ldw = h->ldw;
        for (u32 k = 0; k < R; ++k)
            CASPHY_REQUIRE(q[k] < n, CASPHY_EARG, "query row " + std::to_string(q[k]) + " out of range");
        for (u32 k = 1; k < nradii; ++k) require(radii[k - 1] <= radii[k], "radii must be ascending");
        u64 smax;
        int any;
        autosplit_limit(R, n, bound, smax, any);
        hipStream_t s = thread_stream();
        const bool mfma = h->mfma();
        if (mfma) h->staged(s);

        // every buffer the round touches is sized before anything is enqueued, so
        // the enqueued work allocates nothing and can be captured into a graph
        const size_t up = (MAX_ROUND_RADII + (size_t)R) * 4, res_off = (up + 15) / 16 * 16;
        const size_t cnt_bytes = (size_t)R * nradii * 4;
        h->pinned(res_off + 16 + cnt_bytes);
        h->a_in.reserve(up);
        h->a_counts.reserve(cnt_bytes);
        h->a_misc.reserve(16);
        // on the fp4 tiles the counts come straight from the tiles when the radii
        // fit the count table (no score matrix)
        bool fused = false;
        if (mfma) {
            h->a_sq.reserve((size_t)stage_rows(R) * ldw * 4);
            h->a_eq.reserve(fp4_bytes(stage_rows(R), ldw / 4));
            MfmaArgs f;
            f.EA = h->a_eq.as<u32>(); f.EB = h->estage.as<u32>(); f.stages = ldw / 4;
            fused = fused_counts_ok(f, nradii) && counts_span(radii, nradii) <= COUNT_SPAN_MAX;
        }
        if (fused) {
            h->a_tab.reserve(COUNT_SPAN_MAX + 64 + 16);
            h->a_cpart.reserve((size_t)ceil_div(n, TILE) * nradii * stage_rows(R));
        } else {
            h->a_S.reserve((size_t)R * n * 4);
        }
        RadiusOut &o = h->out;
        const u32 nt = ceil_div(n, TILE), mstride = nt * 4;
        o.mask.reserve((size_t)nt * TILE * mstride * 4);
        o.counts.reserve((size_t)n * 4);
        o.offsets.reserve((size_t)(n + 1) * 8);
        const u64 cap = o.idx.bytes / 4;

        int32_t *pr = static_cast<int32_t *>(h->pin);
        for (u32 k = 0; k < MAX_ROUND_RADII; ++k) pr[k] = k < nradii ? radii[k] : 0x7FFFFFFF;
        std::memcpy(pr + MAX_ROUND_RADII, q, (size_t)R * 4);
        char *res = static_cast<char *>(h->pin) + res_off;
        const bool want = counts != nullptr;

        // the inputs come in and the results go out through kernels reading and
        // writing the mapped pinned buffer (no copy engine in the round)
        const int32_t *dr = h->a_in.as<int32_t>();
        const u32 *dq = h->a_in.as<u32>() + MAX_ROUND_RADII;
        int32_t *choice = h->a_misc.as<int32_t>();
        int32_t *dS = h->a_S.as<int32_t>();
        // phase 0: tryRadius counts of the query rows and the device-side choice
        auto phase0 = [&] {
            launch_copy_words(static_cast<const u32 *>(h->pin_dev), h->a_in.as<u32>(), (u32)(up / 4), s);
            if (mfma) {
                launch_to_stage_major(h->mat.as<u32>(), R, ldw, ldw, h->a_sq.as<u32>(), s, dq);
                launch_expand_fp4(h->a_sq.as<u32>(), stage_rows(R), ldw / 4, h->a_eq.as<u32>(), s);
                MfmaArgs m;
                m.PA = h->a_sq.as<u32>(); m.PB = h->stage.as<u32>(); m.padA = stage_rows(R); m.padB = stage_rows(n);
                m.stages = ldw / 4; m.nA = R; m.nB = n; m.triangle = 0;
                m.EA = h->a_eq.as<u32>(); m.EB = h->estage.as<u32>();
                if (fused) {
                    unsigned char *t = h->a_tab.as<unsigned char>();
                    m.ctab = t; m.cneed = t + COUNT_SPAN_MAX;
                    m.cpar = reinterpret_cast<const int32_t *>(t + COUNT_SPAN_MAX + 64);
                    m.nrad = nradii; m.cpart = h->a_cpart.as<unsigned char>();
                    launch_count_table(dr, nradii, t, t + COUNT_SPAN_MAX,
                                       reinterpret_cast<int32_t *>(t + COUNT_SPAN_MAX + 64), s);
                    launch_mfma_counts(m, h->a_counts.as<u32>(), R, s);
                    launch_autosplit_choose(h->a_counts.as<u32>(), R, dr, nradii, smax, any, choice, s);
                    return;
                }
                m.S = dS; m.ldS = n;
                launch_mfma_tiles(E_SCORES, m, s);
            } else {
                TileArgs a;
                a.A = h->mat.as<u32>(); a.Aidx = dq; a.nA = R; a.B = h->mat.as<u32>(); a.nB = n;
                a.ldw = ldw; a.kw = ldw;
                a.ntA = ceil_div(R, TILE); a.ntB = ceil_div(n, TILE);
                a.ntiles = a.ntA * a.ntB;
                a.S = dS; a.ldS = n;
                apply_planes(h->planes, a, h->a_pbuf, s);
                launch_tiles(h->metric, E_SCORES, a, s);
            }
            launch_count_radii(dS, R, n, n, dr, nradii, h->a_counts.as<u32>(), s);
            launch_autosplit_choose(h->a_counts.as<u32>(), R, dr, nradii, smax, any, choice, s);
        };
        // phase 1: counts, offsets and the lists at the chosen radius (nothing
        // qualifies: threshold INT32_MAX, the tiles and the list kernels skip);
        // ids beyond the current capacity are dropped and the compaction reruns
        // once the total is known
        auto phase1 = [&] {
            radius_mask_counts(n, choice, o, s);
            if (cap) launch_compact(o.mask.as<u32>(), n, mstride, mstride, o.offsets.as<u64>(), o.idx.as<u32>(), s, cap);
            launch_publish(choice, o.offsets.as<u64>() + n, h->a_counts.as<u32>(), want ? R * nradii : 0,
                           reinterpret_cast<u32 *>(static_cast<char *>(h->pin_dev) + res_off), s);
        };
        // the radius-mask tiles read their threshold from `choice`: a plain
        // launch between the two phases, timed by the handle's events
        auto mask_tiles = [&] {
            radius_mask_tiles(h->metric, h->mat.as<u32>(), n, ldw, h->planes, 0, choice, o, s, h->ev[0], h->ev[1],
                              mfma ? h->stage.as<u32>() : nullptr, mfma ? h->estage.as<u32>() : nullptr);
        };

        // The two phases are fixed chains of small launches: the second call with
        // the same shapes and buffers captures each into a graph (the first one
        // also runs the autotuner), later calls replay them.
        // CASPHY_AUTOSPLIT_GRAPH=0 keeps plain launches.
        const AutosplitKey key{R,          nradii,     smax,       any,        want,   cap,
                               fused,      o.idx.p,    o.mask.p,   h->a_S.p,   h->a_eq.p, h->a_in.p,
                               h->pin_dev, h->a_tab.p, h->a_cpart.p};
        const char *genv = std::getenv("CASPHY_AUTOSPLIT_GRAPH");
        const bool graphs = !(genv && genv[0] == '0');
        h->timed[0] = h->timed[1] = true;
        auto capture = [&](int k, const std::function<void()> &body) {
            hipGraph_t g = nullptr;
            CASPHY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            CASPHY_HIP(hipStreamEndCapture(s, &g));
            const hipError_t e = hipGraphInstantiate(&h->ag_exec[k], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            CASPHY_HIP(e);
        };
        const bool replay = graphs && h->ag_exec[0] && h->ag_exec[1] && h->ag_key == key;
        if (!replay) {
            h->drop_graph();
            if (graphs && h->ag_seen && h->ag_key == key) {
                capture(0, phase0);
                capture(1, phase1);
            }
        }
        CASPHY_HIP(hipEventRecord(h->ev[2], s));
        if (h->ag_exec[0]) CASPHY_HIP(hipGraphLaunch(h->ag_exec[0], s)); else phase0();
        mask_tiles();
        if (h->ag_exec[1]) CASPHY_HIP(hipGraphLaunch(h->ag_exec[1], s)); else phase1();
        CASPHY_HIP(hipEventRecord(h->ev[3], s));
        h->ag_key = key;
        h->ag_seen = true;
        CASPHY_HIP(hipStreamSynchronize(s));
        int32_t ch[2];
        u64 total;
        std::memcpy(ch, res, 8);
        std::memcpy(&total, res + 8, 8);
        if (want) std::memcpy(counts, res + 16, cnt_bytes);
        if (total > cap) {
            o.idx.reserve(total * 4);  // the captured round wrote into the old buffer: recapture next time
            h->drop_graph();
            h->ag_seen = false;
            launch_compact(o.mask.as<u32>(), n, mstride, mstride, o.offsets.as<u64>(), o.idx.as<u32>(), s, total);
            CASPHY_HIP(hipEventRecord(h->ev[3], s));
            CASPHY_HIP(hipStreamSynchronize(s));
        }
        o.total = total;
        o.n = n;
        *radius = ch[1] < 0 ? -1 : ch[0];
        *nidx = total;
    });
}

int casphy_cells_last_ms(casphy_cells *h, float *tile_ms, float *radius_ms)
{
    return guarded([&] {
        require(h && tile_ms && radius_ms, "casphy_cells_last_ms: bad arguments");
        *tile_ms = *radius_ms = 0.0f;
        if (h->timed[0]) CASPHY_HIP(hipEventElapsedTime(tile_ms, h->ev[0], h->ev[1]));
        if (h->timed[1]) CASPHY_HIP(hipEventElapsedTime(radius_ms, h->ev[2], h->ev[3]));
    });
}

}  // extern "C"
