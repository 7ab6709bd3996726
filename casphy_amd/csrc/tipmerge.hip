// tipmerge.hip -- tipMerge: clusters of tips that share most of their neighbours.
//
// cu/tipMerge.cu intersects, one warp per pair (sharedIdxKernel :25-129), the
// neighbour list of a tip j with the list of each later neighbour t, one tip
// at a time from the host (:307-410), and folds the pairs whose shared count
// exceeds p times either list length into categories (:379-405).  Here every
// linked pair (j, t > j) is counted in one pass:
//   * dense form: the lists as bit rows (bit t of row j set iff t is in the
//     list of j) -- for resident cells the radius mask itself -- so that
//     |N(j) & N(t)| is the AND-popcount of two bit rows, taken on the matrix
//     cores over the upper triangle; the E_PAIRS epilogue stores the count of
//     each linked pair at its place in the pair array (row base + set bits of
//     row j before column t);
//   * sparse form: one workgroup per tip j holds N(j) as a bitmap in LDS (ids
//     in windows of 2^19) and each wave streams the list of one t through it.
// One kernel then keeps, per tip and in list order, the partners that pass
// the reference's test (in double precision, as the reference's u32 * double
// products); the host replays the reference's category bookkeeping over the
// kept pairs in the reference's order.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "context.hpp"
#include "kernels.hpp"
#include "mfma.hpp"
#include "tipmerge.hpp"

namespace casphy {

namespace {

__device__ __forceinline__ u32 wave_sum(u32 x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Row i of B (ldw words, dynamic LDS of ldw words): bits of the ids of list i.
__global__ __launch_bounds__(256) void k_lists_to_rows(const u64 *off, const u32 *ids, u32 ldw, u32 *B)
{
    extern __shared__ u32 row[];
    const u32 i = blockIdx.x;
    for (u32 w = threadIdx.x; w < ldw; w += 256) row[w] = 0u;
    __syncthreads();
    for (u64 k = off[i] + threadIdx.x; k < off[i + 1]; k += 256) {
        const u32 t = ids[k];
        atomicOr(&row[t >> 5], 1u << (t & 31));
    }
    __syncthreads();
    u32 *dst = B + (size_t)i * ldw;
    for (u32 w = threadIdx.x; w < ldw; w += 256) dst[w] = row[w];
}

// One wave per bit row: Q[i][w] = set bits of row i in words < w, first[i] =
// set bits at columns <= i (the list position of the first partner > i),
// upper[i] = set bits at columns > i.
__global__ __launch_bounds__(256) void k_row_prefix(const u32 *B, u32 n, u32 ldw, u32 *Q, u32 *first, u32 *upper)
{
    const u32 lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const u32 *row = B + (size_t)i * ldw;
    u32 *q = Q + (size_t)i * ldw;
    const u32 wi = i >> 5;
    u32 carry = 0, fi = 0;
    for (u32 w0 = 0; w0 < ldw; w0 += 64) {
        const u32 w = w0 + lane;
        const u32 word = w < ldw ? row[w] : 0u;
        const u32 c = (u32)__builtin_popcount(word);
        u32 x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const u32 y = __shfl_up(x, o);
            if (lane >= (u32)o) x += y;
        }
        const u32 ex = carry + x - c;
        if (w < ldw) q[w] = ex;
        if (w == wi) fi = ex + (u32)__builtin_popcount(word & (0xFFFFFFFFu >> (31 - (i & 31))));
        carry += __shfl(x, 63);
    }
    fi = wave_sum(fi);  // held by one lane
    if (lane == 0) {
        first[i] = fi;
        upper[i] = carry - fi;
    }
}

// The same counts from the lists (sparse form): one wave per tip.
__global__ __launch_bounds__(256) void k_list_first(const u64 *off, const u32 *ids, u32 n, u32 *first, u32 *upper)
{
    const u32 lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    u32 c = 0;
    for (u64 k = off[i] + lane; k < off[i + 1]; k += 64) c += ids[k] <= i;
    c = wave_sum(c);
    if (lane == 0) {
        first[i] = c;
        upper[i] = (u32)(off[i + 1] - off[i]) - c;
    }
}

// prow[i] = uoff[i] - first[i] (modulo 2^64): the pair slot of row i's first
// set bit counted from column 0.
__global__ void k_pair_rows(const u64 *uoff, const u32 *first, u32 n, u64 *prow)
{
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) prow[i] = uoff[i] - first[i];
}

// Sparse form: one workgroup per tip j, N(j) as a bitmap of winw words in LDS
// per window of ids; wave u handles the partners u, u + 4, ...
__global__ __launch_bounds__(256) void k_shared_sparse(const u64 *off, const u32 *ids, const u32 *first,
                                                       const u64 *uoff, u32 n, u32 winw, u32 *shared)
{
    extern __shared__ u32 bm[];
    const u32 j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const u64 b = off[j], e = off[j + 1];
    const u32 nj = (u32)(e - b), f = first[j];
    if (f >= nj) return;  // no partner after j (uniform)
    const u32 np = nj - f;
    const u64 o = uoff[j];
    for (u32 w = tid; w < winw; w += 256) bm[w] = 0u;
    __syncthreads();  // every word cleared before any thread sets bits in it
    const u64 span = (u64)winw * 32;
    for (u64 lo = 0; lo < n; lo += span) {
        const u64 hi = lo + span;
        for (u64 k = b + tid; k < e; k += 256) {
            const u32 x = ids[k];
            if (x >= lo && x < hi) atomicOr(&bm[(x - lo) >> 5], 1u << (x & 31));
        }
        __syncthreads();
        for (u32 u = wv; u < np; u += 4) {
            const u32 t = ids[b + f + u];
            u32 c = 0;
            for (u64 k = off[t] + lane; k < off[t + 1]; k += 64) {
                const u32 x = ids[k];
                if (x >= lo && x < hi) c += (bm[(x - lo) >> 5] >> (x & 31)) & 1u;
            }
            c = wave_sum(c);
            if (lane == 0) shared[o + u] = (lo ? shared[o + u] : 0u) + c;
        }
        __syncthreads();
        for (u64 k = b + tid; k < e; k += 256) {
            const u32 x = ids[k];
            if (x >= lo && x < hi) bm[(x - lo) >> 5] = 0u;
        }
        __syncthreads();
    }
}

// One workgroup per tip j: the partners t > j whose shared count s passes
// s > |N(j)| p or s > |N(t)| p (cu/tipMerge.cu:379), in list order, to
// at[uoff[j] ..] with their number in acnt[j].
__global__ __launch_bounds__(256) void k_accept(const u64 *off, const u32 *ids, const u32 *first, const u64 *uoff,
                                                const u32 *shared, double p, u32 *acnt, u32 *at)
{
    __shared__ u32 wsum[4];
    const u32 j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const u64 b = off[j];
    const u32 nj = (u32)(off[j + 1] - b), f = first[j];
    const u32 np = f < nj ? nj - f : 0u;
    const u64 o = uoff[j];
    const double lim = nj * p;
    u32 kept = 0;
    for (u32 u0 = 0; u0 < np; u0 += 256) {
        const u32 u = u0 + tid;
        bool keep = false;
        u32 t = 0;
        if (u < np) {
            t = ids[b + f + u];
            const u32 s = shared[o + u];
            const u32 nt = (u32)(off[t + 1] - off[t]);
            keep = (double)s > lim || (double)s > nt * p;
        }
        const unsigned long long bal = __ballot(keep);
        const u32 before = (u32)__builtin_popcountll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wv] = (u32)__builtin_popcountll(bal);
        __syncthreads();
        u32 base = kept;
        for (u32 k = 0; k < wv; ++k) base += wsum[k];
        if (keep) at[o + base + before] = t;
        kept += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (tid == 0) acnt[j] = kept;
}

// transitionRoot (cu/tipMerge.cu:130-141) with path halving: links only ever
// point at categories that were roots when linked, so halving keeps every
// root and only shortens later walks.
inline int32_t root_of(std::vector<int32_t> &tr, int32_t x)
{
    if (x == -1) return -1;
    while (tr[x] != x && tr[x] != -1) {
        const int32_t up = tr[x];
        if (tr[up] != up && tr[up] != -1) tr[x] = tr[up];
        x = up;
    }
    return x;
}

// The host loop of cu/tipMerge.cu:307-451 over the kept pairs: windows of
// 131,072 tips (CACHEMAX :8); in window [c0, c1) tip j < c1 takes its kept
// partners in [c0, c1) in list order.  The reference also drops a partner t
// whose id, read as a category, leads to j's root (:331); the category count
// never exceeds a partner id there (window 0: at most one category per earlier
// tip, and t > j; window k >= 1: two fresh tips below c1 per category, so at
// most c1 / 2 <= c0 <= t), so that test never drops a pair and is not made.
void replay(u32 n, const std::vector<u64> &uoff, const std::vector<u32> &acnt, const std::vector<u32> &at,
            int32_t *member_out)
{
    constexpr u32 WINDOW = 1024u * 128u;
    std::vector<int32_t> member(n, -1), tr(n, -1);
    std::vector<u32> cur(n, 0);
    int32_t ncat = 0;
    for (u32 c0 = 0; c0 < n; c0 += WINDOW) {
        const u32 c1 = n - c0 < WINDOW ? n : c0 + WINDOW;
        for (u32 j = 0; j < c1; ++j) {
            const u32 *seg = at.data() + uoff[j];
            const u32 na = acnt[j];
            u32 k = cur[j];
            if (k >= na || seg[k] >= c1) continue;
            int32_t r1 = root_of(tr, member[j]);
            for (; k < na && seg[k] < c1; ++k) {
                const u32 t = seg[k];
                const int32_t r2 = root_of(tr, member[t]);
                if (r1 == -1) {
                    if (r2 == -1) {
                        tr[ncat] = ncat;
                        member[j] = ncat;
                        member[t] = ncat;
                        r1 = ncat++;
                    } else {
                        r1 = r2;
                        member[j] = r2;
                    }
                } else if (r2 == -1) {
                    member[t] = r1;
                } else {
                    tr[r1] = r2;
                    member[j] = r2;
                    r1 = r2;
                }
            }
            cur[j] = k;
        }
    }
    for (u32 i = 0; i < n; ++i) member_out[i] = member[i] == -1 ? ncat++ : root_of(tr, member[i]);
}

constexpr u32 DENSE_MAX = 262144;  // bit rows of n^2 / 8 bytes and their fp4 copy stay modest

}  // namespace

void tip_merge_device(const u64 *off, const u32 *ids, u32 n, const u32 *bits, u32 bstride, double p,
                      int32_t *member, hipStream_t s)
{
    if (n == 0) return;
    u64 total = 0;
    d2h(&total, off + n, 8, s);
    // dense (matrix cores) when the lists fill more than 1/140 of the pairs:
    // n^3 / 2 bit products on the matrix cores against the sum over pairs of
    // |N(t)| list reads in the sparse form, both growing as (lists / n^2)^-2
    // relative to each other; 1/140 is where the two forms took the same time
    // on 10,000 cells (CASPHY_TIPMERGE=dense|sparse forces one)
    const char *env = std::getenv("CASPHY_TIPMERGE");
    bool dense = mfma_enabled() && n <= DENSE_MAX && (double)total * 140.0 > (double)n * n;
    if (env && !std::strcmp(env, "dense")) dense = mfma_enabled() && n <= DENSE_MAX;
    if (env && !std::strcmp(env, "sparse")) dense = false;

    DevBuf first((size_t)n * 4), upper((size_t)n * 4), uoff(((size_t)n + 1) * 8);
    DevBuf B, Q, P, prow, shared, acnt((size_t)n * 4), at;
    const u32 ldw = round_up(ceil_div(n, 32), 32);
    if (dense) {
        B.alloc((size_t)n * ldw * 4);
        if (bits) {
            launch_pad_rows(bits, bstride, n, B.as<u32>(), ldw, s);
        } else {
            hipLaunchKernelGGL(k_lists_to_rows, dim3(n), dim3(256), (size_t)ldw * 4, s, off, ids, ldw, B.as<u32>());
            CASPHY_HIP(hipGetLastError());
        }
        Q.alloc((size_t)n * ldw * 4);
        hipLaunchKernelGGL(k_row_prefix, dim3(ceil_div(n, 4)), dim3(256), 0, s, B.as<u32>(), n, ldw, Q.as<u32>(),
                           first.as<u32>(), upper.as<u32>());
    } else {
        hipLaunchKernelGGL(k_list_first, dim3(ceil_div(n, 4)), dim3(256), 0, s, off, ids, n, first.as<u32>(),
                           upper.as<u32>());
    }
    CASPHY_HIP(hipGetLastError());
    launch_scan_counts(upper.as<u32>(), n, uoff.as<u64>(), s);
    u64 npairs = 0;
    d2h(&npairs, uoff.as<u64>() + n, 8, s);
    shared.alloc(std::max<u64>(npairs, 1) * 4);
    at.alloc(std::max<u64>(npairs, 1) * 4);
    if (npairs) {
        if (dense) {
            prow.alloc((size_t)n * 8);
            hipLaunchKernelGGL(k_pair_rows, dim3(std::min<u32>(ceil_div(n, 256), 1024)), dim3(256), 0, s,
                               uoff.as<u64>(), first.as<u32>(), n, prow.as<u64>());
            CASPHY_HIP(hipGetLastError());
            P.alloc((size_t)stage_rows(n) * ldw * 4);
            launch_to_stage_major(B.as<u32>(), n, ldw, ldw, P.as<u32>(), s);
            MfmaArgs m;
            m.PA = m.PB = P.as<u32>();
            m.padA = m.padB = stage_rows(n);
            m.stages = ldw / 4;
            m.nA = m.nB = n;
            m.triangle = 1;
            m.mask = B.as<u32>();
            m.mstride = ldw;
            m.pq = Q.as<u32>();
            m.prow = prow.as<u64>();
            m.pout = shared.as<u32>();
            launch_mfma_tiles(E_PAIRS, m, s);
        } else {
            const u32 winw = std::min<u32>(ceil_div(n, 32), 16384);
            hipLaunchKernelGGL(k_shared_sparse, dim3(n), dim3(256), (size_t)winw * 4, s, off, ids, first.as<u32>(),
                               uoff.as<u64>(), n, winw, shared.as<u32>());
            CASPHY_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_accept, dim3(n), dim3(256), 0, s, off, ids, first.as<u32>(), uoff.as<u64>(),
                           shared.as<u32>(), p, acnt.as<u32>(), at.as<u32>());
        CASPHY_HIP(hipGetLastError());
    } else {
        launch_fill_u32(acnt.as<u32>(), n, 0u, s);
    }
    std::vector<u64> huoff((size_t)n + 1);
    std::vector<u32> hacnt(n), hat(std::max<u64>(npairs, 1));
    d2h(huoff.data(), uoff.p, ((size_t)n + 1) * 8, s);
    d2h(hacnt.data(), acnt.p, (size_t)n * 4, s);
    if (npairs) d2h(hat.data(), at.p, npairs * 4, s);
    replay(n, huoff, hacnt, hat, member);
}

}  // namespace casphy
