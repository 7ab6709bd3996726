// join.hip -- greedy shared-mutation join on device (binary and indel).
//
// State of one tree lives in HBM for the whole join:
//   rows      n slots x (2 planes x ldw words).  Binary: plane 0 = seq
//             (including bistate sites), plane 1 = bistate.  Indel: plane 0 =
//             one-hot (site, allele) bits, plane 1 = per-bit "missing" flags
//             (every bit of a missing site's group is set).  Scores use plane 0.
//   lmax/lidx per slot: best score over later nodes and that node's id
//   slot2node / node2slot, node sizes (indel size rule)
//   edge      raw (parent, child) pairs, nstate = state of each internal node
//   ctl       control block read and written only by the kernels
// A join step is two launches with fixed arguments:
//   k_rescore  (all slots) scores the new node against every live slot and
//              raises lmax where strictly larger, or computes the stale row's
//              (score, smallest node id) maximum for the recompute; each
//              workgroup also leaves the argmax key of its 16 slots;
//   k_select   (one workgroup) finalises a pending recompute, takes the global
//              (score, smallest node id) argmax over those keys, and either
//              requests a recompute of a stale row or performs the join:
//              ancestral update into the first child's slot, bookkeeping, edge
//              record.
// The host only enqueues steps (captured once into a hipGraph and replayed)
// and polls the done flag every chunk, so no per-join round trip exists.
// Semantics follow cu/joinEdgeBitReverse.cu:347-505 / cu/joinEdgeIndel.cu:
// 306-468 (lazy local maxima, (-1, 0) argmax start, strict '>' merge), which
// select the same pair as the CPU twins (src/functions.cpp:1917-1949,
// 3359-3391): the maximal score, ties to the smallest i then smallest j.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "context.hpp"
#include "join.hpp"
#include "kernels.hpp"
#include "mfma.hpp"
#include "tiles.hpp"

namespace casphy {

namespace {

constexpr u32 MODE_NONE = 0, MODE_RESCORE = 1, MODE_RECOMPUTE = 2;

struct Ctl {
    u32 nodeCount;
    u32 mode;
    u32 c1;
    u32 newnode;
    u32 done;
    u32 err;
    u32 steps;
    u32 arrive;  // k_step workgroups that finished their rescoring
    u32 nlive;   // live slots: slots 0 .. nlive-1 hold the live nodes
    unsigned long long key;
};

// rule: 0/1 binary ancestral type, 2 indel
struct JoinArgs {
    u32 *rows;
    int32_t *lmax;
    u32 *lidx;
    u32 *slot2node;
    u32 *node2slot;
    u32 *size;
    u32 *edge;
    u32 *nstate;
    Ctl *ctl;
    unsigned long long *part;  // per-workgroup argmax keys left by the rescoring launch
    u32 nparts;                // = rescoring workgroups, ceil(n / 16)
    u32 n, ldw;
    int rule;
    u32 min_size;
    float min_ratio;
    u32 nplanes;
    const u32 *pmask;  // nplanes x ldw
    u32 pshift[MAX_PLANES];
    u32 ptail;
};

__device__ __forceinline__ void merge_words(int rule, u32 a0, u32 a1, u32 b0, u32 b1, u32 &r0, u32 &r1)
{
    if (rule == 0) {  // calcAncestralBitReverse, src/functions.cpp:3231-3235
        const u32 diff = a0 ^ b0;
        const u32 conflict = diff & (a1 | b1);
        r0 = (a0 | b0) & ~conflict;
        r1 = (diff & ~(a1 ^ b1)) | (a1 & b1);
    } else if (rule == 1) {  // calcAncestralBitReverse2, :3304-3305
        r0 = (a0 & b0) | (a1 & b0) | (b1 & a0);
        r1 = (a0 ^ b0) & ~(a1 | b1);
    } else {  // indel rule of src/functions.cpp:1965-1974 on one-hot planes
        r0 = (a1 & b0) | (~a1 & b1 & a0) | (~a1 & ~b1 & a0 & b0);
        r1 = a1 & b1;
    }
}

// Selection by one workgroup of NT threads (see the file comment); red, sh and
// fin are LDS scratch of the calling kernel.
// PARTS: the argmax runs over nparts per-workgroup maxima (written by the
// rescoring workgroups of the same launch, which leave out a slot under
// recompute) plus the recomputed slot, instead of over every slot.
template <int NT, bool PARTS>
__device__ __forceinline__ void select_phase(const JoinArgs &a, unsigned long long *red, u32 *sh, u32 *fin,
                                             u32 nparts)
{
    Ctl *ctl = a.ctl;
    if (ctl->done) return;
    const u32 tid = threadIdx.x, n = a.n, nnode = 2 * n - 1;
    if (tid == 0) {
        fin[0] = 0xFFFFFFFFu;
        if (ctl->mode == MODE_RECOMPUTE) {
            const unsigned long long key = ctl->key;
            const u32 c = ctl->c1;
            const int32_t v = key ? (int32_t)(key >> 32) : -1;
            const u32 j = key ? 0xFFFFFFFFu - (u32)key : 0xFFFFFFFFu;
            a.lmax[c] = v;
            a.lidx[c] = j;
            fin[0] = c; fin[1] = (u32)v; fin[2] = j;
        }
    }
    __syncthreads();
    const u32 fslot = fin[0];
    // argmax of (lmax, -node id) from the candidate (-1, 0)
    unsigned long long best = 0xFFFFFFFFull;  // value -1 (+1 bias = 0), id 0
    if (PARTS) {
        for (u32 b = tid; b < nparts; b += NT) {
            const unsigned long long k = a.part[b];
            best = k > best ? k : best;
        }
        if (tid == 0 && fslot != 0xFFFFFFFFu) {
            const unsigned long long k = ((unsigned long long)(fin[1] + 1u) << 32) | (0xFFFFFFFFu - a.slot2node[fslot]);
            best = k > best ? k : best;
        }
    } else {
        for (u32 s = tid; s < ctl->nlive; s += NT) {
            const int32_t v = s == fslot ? (int32_t)fin[1] : a.lmax[s];
            const u32 id = a.slot2node[s];
            const unsigned long long k = ((unsigned long long)(u32)(v + 1) << 32) | (0xFFFFFFFFu - id);
            best = k > best ? k : best;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(best, o);
        best = x > best ? x : best;
    }
    if ((tid & 63) == 0) red[tid >> 6] = best;
    __syncthreads();
    if (tid == 0) {
        unsigned long long b = red[0];
        for (int w = 1; w < NT / 64; ++w) b = red[w] > b ? red[w] : b;
        const u32 j1 = 0xFFFFFFFFu - (u32)b;
        const int32_t v = (int32_t)(u32)(b >> 32) - 1;
        u32 c1 = j1 < nnode ? a.node2slot[j1] : n;
        u32 j2 = 0xFFFFFFFFu, c2 = n;
        u32 verdict;
        if (v < 0 || c1 >= n) {
            verdict = 0;  // no live pair: cannot happen with >= 2 live nodes
            ctl->err = 1;
            ctl->done = 1;
            ctl->mode = MODE_NONE;
        } else {
            j2 = c1 == fslot ? fin[2] : a.lidx[c1];
            c2 = j2 < nnode ? a.node2slot[j2] : n;
            if (c2 >= n) {
                verdict = 1;  // partner already joined: recompute this row
                ctl->mode = MODE_RECOMPUTE;
                ctl->c1 = c1;
                ctl->key = 0;
            } else {
                verdict = 2;
            }
        }
        sh[0] = verdict; sh[1] = c1; sh[2] = c2; sh[3] = 0; sh[4] = ctl->nodeCount;
        if (verdict == 2) {
            const u32 nc = sh[4];
            const u32 sa = a.size[j1], sb = a.size[j2];
            u32 choice = 0;  // 0 merge, 1 copy b, 2 keep a (cu/joinEdgeIndel.cu:407-413)
            if (a.rule == 2) {
                if (sa < a.min_size && (float)sa * a.min_ratio < (float)sb) choice = 1;
                else if (sb < a.min_size && (float)sb * a.min_ratio < (float)sa) choice = 2;
            }
            sh[3] = choice;
            const u32 k = nc - n;
            a.edge[4 * k + 0] = nc; a.edge[4 * k + 1] = j1;
            a.edge[4 * k + 2] = nc; a.edge[4 * k + 3] = j2;
            a.node2slot[j1] = n; a.node2slot[j2] = n; a.node2slot[nc] = c1;
            a.slot2node[c1] = nc; a.slot2node[c2] = nnode;
            a.lmax[c1] = -1; a.lmax[c2] = -1;
            a.size[nc] = sa + sb;
            ctl->nodeCount = nc + 1;
            // slot c2 is freed: the last live slot moves into it (after the
            // merge below), so the live slots stay 0 .. nlive-1
            const u32 last = ctl->nlive - 1;
            sh[5] = last;
            ctl->nlive = last;
            ctl->c1 = c1 == last ? c2 : c1;
            ctl->newnode = nc;
            ctl->mode = MODE_RESCORE;
            if (nc + 1 == nnode) {
                ctl->done = 1;
                ctl->mode = MODE_NONE;
            }
        }
        ctl->steps += 1;
    }
    __syncthreads();
    if (sh[0] != 2) return;
    const u32 c1 = sh[1], c2 = sh[2], choice = sh[3], nc = sh[4], last = sh[5];
    const u32 ldw = a.ldw;
    u32 *ra = a.rows + (size_t)c1 * 2 * ldw;
    const u32 *rb = a.rows + (size_t)c2 * 2 * ldw;
    u32 *out = a.nstate + (size_t)(nc - n) * 2 * ldw;
    for (u32 w = tid; w < ldw; w += NT) {
        const u32 a0 = ra[w], a1 = ra[ldw + w], b0 = rb[w], b1 = rb[ldw + w];
        u32 r0, r1;
        if (choice == 1) { r0 = b0; r1 = b1; }
        else if (choice == 2) { r0 = a0; r1 = a1; }
        else merge_words(a.rule, a0, a1, b0, b1, r0, r1);
        ra[w] = r0; ra[ldw + w] = r1;
        out[w] = r0; out[ldw + w] = r1;
    }
    if (c2 == last) return;
    __syncthreads();  // the merged row is complete (it may be the one that moves)
    const u32 *src = a.rows + (size_t)last * 2 * ldw;
    u32 *dst = a.rows + (size_t)c2 * 2 * ldw;
    for (u32 w = tid; w < 2 * ldw; w += NT) dst[w] = src[w];
    if (tid == 0) {
        const u32 node = a.slot2node[last];
        a.slot2node[c2] = node;
        a.node2slot[node] = c2;
        a.lmax[c2] = a.lmax[last];
        a.lidx[c2] = a.lidx[last];
        a.slot2node[last] = 2 * a.n - 1;
        a.lmax[last] = -1;
    }
}

// PARTS = false: full scan over the slots (first selection after the initial
// maxima); PARTS = true: over the maxima k_rescore left per workgroup.
template <bool PARTS>
__global__ __launch_bounds__(1024) void k_select(JoinArgs a)
{
    __shared__ unsigned long long red[16];
    __shared__ u32 sh[6];
    __shared__ u32 fin[3];  // recompute result: slot, value, partner (LDS copy)
    select_phase<1024, PARTS>(a, red, sh, fin, a.nparts);
}

// 16 lanes per target slot, 16 slots per 256-thread workgroup and group of
// slots; the workgroups loop over the groups of live slots (0 .. nlive-1).
// Returns (to every thread) the workgroup's argmax key (lmax + 1, smallest
// node id) over its live slots after the update, leaving out a slot under
// recompute.
__device__ __forceinline__ unsigned long long rescore_phase(const JoinArgs &a, unsigned long long *red)
{
    const Ctl *ctl = a.ctl;
    const u32 mode = ctl->mode;  // uniform over the grid
    const u32 c1 = ctl->c1, newnode = ctl->newnode, nlive = ctl->nlive, nnode = 2 * a.n - 1, ldw = a.ldw;
    const u32 tid = threadIdx.x, l16 = tid & 15;
    const bool active = mode == MODE_RESCORE || mode == MODE_RECOMPUTE;
    const u32 *q = a.rows + (size_t)c1 * 2 * ldw;
    const u32 np = a.nplanes ? a.nplanes : 1;
    unsigned long long pk = 0, rk = 0;
    for (u32 g = blockIdx.x; g * 16 < nlive; g += gridDim.x) {
        const u32 k = g * 16 + (tid >> 4);
        // the live slots are 0 .. nlive-1, so the row loads need not wait for
        // the node id or the current maximum, which load alongside them
        const bool alive = k < nlive;
        const bool live = active && alive && k != c1;
        const u32 node = alive ? a.slot2node[k] : nnode;
        const bool keyed = l16 == 0 && alive && !(mode == MODE_RECOMPUTE && k == c1);
        int32_t v = keyed ? a.lmax[k] : -1;
        u32 S = 0;
        if (live) {
            const u32 *r = a.rows + (size_t)k * 2 * ldw;
            for (u32 p = 0; p < np; ++p) {
                if (p) S <<= a.pshift[p];
                u32 part = 0;
                for (u32 w = 4 * l16; w < ldw; w += 64) {
                    uint4 x = *reinterpret_cast<const uint4 *>(q + w);
                    const uint4 y = *reinterpret_cast<const uint4 *>(r + w);
                    if (a.nplanes) {
                        const uint4 m = *reinterpret_cast<const uint4 *>(a.pmask + (size_t)p * ldw + w);
                        x.x &= m.x; x.y &= m.y; x.z &= m.z; x.w &= m.w;
                    }
                    part += __builtin_popcount(x.x & y.x) + __builtin_popcount(x.y & y.y) +
                            __builtin_popcount(x.z & y.z) + __builtin_popcount(x.w & y.w);
                }
                for (int o = 1; o < 16; o <<= 1) part += __shfl_xor(part, o);
                S += part;
            }
            if (a.nplanes) S <<= a.ptail;
        }
        if (keyed) {
            if (mode == MODE_RESCORE && live && (int32_t)S > v) {
                v = (int32_t)S;
                a.lmax[k] = v;
                a.lidx[k] = newnode;
            }
            const unsigned long long key = ((unsigned long long)(u32)(v + 1) << 32) | (0xFFFFFFFFu - node);
            pk = key > pk ? key : pk;
        }
        if (mode == MODE_RECOMPUTE && live) {
            const unsigned long long key = ((unsigned long long)S << 32) | (0xFFFFFFFFu - node);
            rk = key > rk ? key : rk;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(pk, o), y = __shfl_xor(rk, o);
        pk = x > pk ? x : pk;
        rk = y > rk ? y : rk;
    }
    if ((tid & 63) == 0) { red[tid >> 6] = pk; red[4 + (tid >> 6)] = rk; }
    __syncthreads();
    pk = red[0]; rk = red[4];
    for (int w = 1; w < 4; ++w) {
        pk = red[w] > pk ? red[w] : pk;
        rk = red[4 + w] > rk ? red[4 + w] : rk;
    }
    if (mode == MODE_RECOMPUTE && tid == 0 && rk) atomicMax((unsigned long long *)&((Ctl *)ctl)->key, rk);
    return pk;
}

__global__ __launch_bounds__(256) void k_rescore(JoinArgs a)
{
    __shared__ unsigned long long red[8];
    if (a.ctl->done) return;
    const unsigned long long pk = rescore_phase(a, red);
    if (threadIdx.x == 0) a.part[blockIdx.x] = pk;
}

// One join step in one launch: every workgroup rescores its 16 slots, then the
// last workgroup to arrive (device-scope counter after a release fence) runs
// the selection for the next step.  No workgroup waits on another.
__global__ __launch_bounds__(256) void k_step(JoinArgs a)
{
    __shared__ unsigned long long red[16];
    __shared__ u32 sh[6];
    __shared__ u32 fin[3];
    __shared__ u32 last;
    if (a.ctl->done) return;
    const unsigned long long pk = rescore_phase(a, red);
    if (threadIdx.x == 0) a.part[blockIdx.x] = pk;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&a.ctl->arrive, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x == 0) a.ctl->arrive = 0;
    select_phase<256, true>(a, red, sh, fin, gridDim.x);
}

// keys -> lmax/lidx, maps, sizes, control block.
__global__ void k_join_init(JoinArgs a, const u64 *keys, const u32 *tip_size)
{
    const u32 n = a.n, nnode = 2 * n - 1;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nnode; i += gridDim.x * blockDim.x) {
        if (i < n) {
            const u64 key = keys[i];
            a.lmax[i] = key ? (int32_t)(key >> 32) : -1;
            a.lidx[i] = key ? 0xFFFFFFFFu - (u32)key : 0xFFFFFFFFu;
            a.slot2node[i] = i;
            a.node2slot[i] = i;
            a.size[i] = tip_size ? tip_size[i] : 1u;
        } else {
            a.node2slot[i] = n;
            a.size[i] = 0;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Ctl *c = a.ctl;
        c->nodeCount = n; c->mode = MODE_NONE; c->c1 = 0; c->newnode = 0;
        c->done = n < 2 ? 1u : 0u; c->err = 0; c->steps = 0; c->arrive = 0; c->key = 0; c->nlive = n;
    }
}

// Binary input rows -> slots (plane 1 = 0).
__global__ void k_bit_slots(const u32 *mat, u32 n, u32 W, u32 *rows, u32 ldw)
{
    const u64 total = (u64)n * 2 * ldw;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 r = (u32)(x / (2 * ldw)), w = (u32)(x - (u64)r * 2 * ldw);
        rows[x] = w < W ? mat[(size_t)r * W + w] : 0u;
    }
}

// Indel: per-site maximum allele (states -1 / 0 / k).
__global__ void k_site_max(const int32_t *mat, u32 n, u32 ncol, int32_t *kmax)
{
    const u64 total = (u64)n * ncol;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const int32_t v = mat[x];
        if (v > 0) atomicMax(&kmax[x % ncol], v);
    }
}

// Indel: one-hot encode into the two planes (rows pre-zeroed).
__global__ void k_indel_encode(const int32_t *mat, u32 n, u32 ncol, const u32 *loci, const u32 *width,
                               u32 *rows, u32 ldw)
{
    const u64 total = (u64)n * ncol;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 r = (u32)(x / ncol), s = (u32)(x - (u64)r * ncol);
        const int32_t v = mat[x];
        u32 *row = rows + (size_t)r * 2 * ldw;
        if (v > 0) {
            const u32 b = loci[s] + (u32)v - 1;
            atomicOr(&row[b >> 5], 1u << (b & 31));
        } else if (v < 0) {
            for (u32 b = loci[s]; b < loci[s] + width[s]; ++b) atomicOr(&row[ldw + (b >> 5)], 1u << (b & 31));
        }
    }
}

// Indel: decode planes of `count` rows (stride 2*ldw) back to int32 states.
__global__ void k_indel_decode(const u32 *rows, u32 count, u32 ldw, const u32 *loci, const u32 *width,
                               u32 ncol, int32_t *out)
{
    const u64 total = (u64)count * ncol;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 r = (u32)(x / ncol), s = (u32)(x - (u64)r * ncol);
        const u32 *row = rows + (size_t)r * 2 * ldw;
        const u32 b0 = loci[s];
        int32_t v = 0;
        if ((row[ldw + (b0 >> 5)] >> (b0 & 31)) & 1u) {
            v = -1;
        } else {
            for (u32 t = 0; t < width[s]; ++t) {
                const u32 b = b0 + t;
                if ((row[b >> 5] >> (b & 31)) & 1u) { v = (int32_t)t + 1; break; }
            }
        }
        out[x] = v;
    }
}


// ---------------------------------------------------------------------------
// Batched form: one persistent workgroup per tree, jobs taken from an atomic
// work queue, the whole join (initial local maxima, argmax, stale recompute,
// ancestral merge, rescoring) inside the workgroup with LDS broadcasts and
// barriers only.  Job k owns rows [off_k, off_k + n_k) of every per-slot array;
// its node arrays start at 2*off_k - k, its edges at 2*(2*off_k - 2*k), its
// internal states at (off_k - k) rows.
struct BatchArgs {
    u32 *rows;
    int32_t *lmax;
    u32 *lidx, *slot2node, *node2slot, *size, *edge, *nstate;
    const u32 *offsets;
    const u32 *tip_size;  // sum n_k or nullptr
    u32 njobs, ldw;
    int rule;
    u32 min_size;
    float min_ratio;
    u32 *next_job;
    u32 *status;          // per job: 1 ok, 2 no live pair
    u32 nplanes;
    const u32 *pmask;
    u32 pshift[MAX_PLANES];
    u32 ptail;
};

// Score of rows q and r (ldw words each, plane 0), by 16 cooperating lanes;
// every lane of the group returns the total.
__device__ __forceinline__ u32 group16_score(const BatchArgs &b, const u32 *q, const u32 *r, u32 l16)
{
    const u32 np = b.nplanes ? b.nplanes : 1;
    u32 S = 0;
    for (u32 p = 0; p < np; ++p) {
        if (p) S <<= b.pshift[p];
        u32 part = 0;
        for (u32 w = 4 * l16; w < b.ldw; w += 64) {
            uint4 x = *reinterpret_cast<const uint4 *>(q + w);
            const uint4 y = *reinterpret_cast<const uint4 *>(r + w);
            if (b.nplanes) {
                const uint4 m = *reinterpret_cast<const uint4 *>(b.pmask + (size_t)p * b.ldw + w);
                x.x &= m.x; x.y &= m.y; x.z &= m.z; x.w &= m.w;
            }
            part += __builtin_popcount(x.x & y.x) + __builtin_popcount(x.y & y.y) +
                    __builtin_popcount(x.z & y.z) + __builtin_popcount(x.w & y.w);
        }
        for (int o = 1; o < 16; o <<= 1) part += __shfl_xor(part, o);
        S += part;
    }
    if (b.nplanes) S <<= b.ptail;
    return S;
}

// Workgroup-wide max of a u64 key (1024 threads).
__device__ __forceinline__ unsigned long long wg_max(unsigned long long k, unsigned long long *red)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(k, o);
        k = x > k ? x : k;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = k;
    __syncthreads();
    unsigned long long m = red[0];
    for (int w = 1; w < 16; ++w) m = red[w] > m ? red[w] : m;
    return m;
}

__global__ __launch_bounds__(1024) void k_join_batch(BatchArgs b)
{
    __shared__ unsigned long long red[16];
    __shared__ u32 sh[8];
    const u32 tid = threadIdx.x, l16 = tid & 15, grp = tid >> 4;  // 64 groups of 16 lanes
    const u32 ldw = b.ldw;
    for (;;) {
        __syncthreads();
        if (tid == 0) sh[0] = atomicAdd(b.next_job, 1u);
        __syncthreads();
        const u32 job = sh[0];
        if (job >= b.njobs) return;
        const u32 off = b.offsets[job], n = b.offsets[job + 1] - off, nnode = 2 * n - 1;
        u32 *rows = b.rows + (size_t)off * 2 * ldw;
        int32_t *lmax = b.lmax + off;
        u32 *lidx = b.lidx + off, *s2n = b.slot2node + off;
        u32 *n2s = b.node2slot + (2 * (size_t)off - job);
        u32 *size = b.size + (2 * (size_t)off - job);
        u32 *edge = b.edge + 2 * (2 * (size_t)off - 2 * (size_t)job);
        u32 *nst = b.nstate + ((size_t)off - job) * 2 * ldw;

        // maps and sizes
        for (u32 i = tid; i < nnode; i += 1024) {
            if (i < n) {
                s2n[i] = i;
                n2s[i] = i;
                size[i] = b.tip_size ? b.tip_size[off + i] : 1u;
            } else {
                n2s[i] = n;
            }
        }
        // initial local maxima: group g scores rows i = g, g+64, ... against all j > i
        for (u32 i = grp; i < n; i += 64) {
            unsigned long long best = 0;
            const u32 *qi = rows + (size_t)i * 2 * ldw;
            for (u32 j = i + 1; j < n; ++j) {
                const u32 S = group16_score(b, qi, rows + (size_t)j * 2 * ldw, l16);
                const unsigned long long key = ((unsigned long long)S << 32) | (0xFFFFFFFFu - j);
                best = key > best ? key : best;
            }
            if (l16 == 0) {
                lmax[i] = best ? (int32_t)(best >> 32) : -1;
                lidx[i] = best ? 0xFFFFFFFFu - (u32)best : 0xFFFFFFFFu;
            }
        }
        __syncthreads();
        u32 nodeCount = n;
        u32 status = 1;
        while (nodeCount < nnode) {
            // argmax of (lmax, -node id) from (-1, 0)
            unsigned long long best = 0xFFFFFFFFull;
            for (u32 s = tid; s < n; s += 1024) {
                const unsigned long long k =
                    ((unsigned long long)(u32)(lmax[s] + 1) << 32) | (0xFFFFFFFFu - s2n[s]);
                best = k > best ? k : best;
            }
            best = wg_max(best, red);
            if (tid == 0) {
                const u32 j1 = 0xFFFFFFFFu - (u32)best;
                const int32_t v = (int32_t)(u32)(best >> 32) - 1;
                const u32 c1 = j1 < nnode ? n2s[j1] : n;
                u32 verdict = 0, j2 = 0xFFFFFFFFu, c2 = n;
                if (v >= 0 && c1 < n) {
                    j2 = lidx[c1];
                    c2 = j2 < nnode ? n2s[j2] : n;
                    verdict = c2 < n ? 2 : 1;
                }
                sh[1] = verdict; sh[2] = c1; sh[3] = c2; sh[4] = j1; sh[5] = j2;
            }
            __syncthreads();
            const u32 verdict = sh[1], c1 = sh[2], c2 = sh[3], j1 = sh[4], j2 = sh[5];
            if (verdict == 0) { status = 2; break; }
            const u32 *q = rows + (size_t)c1 * 2 * ldw;
            if (verdict == 1) {
                // recompute row c1 over every live slot
                unsigned long long kb = 0;
                for (u32 k = grp; k < n; k += 64) {
                    const u32 id = s2n[k];
                    if (k == c1 || id >= nnode) continue;
                    const u32 S = group16_score(b, q, rows + (size_t)k * 2 * ldw, l16);
                    const unsigned long long key = ((unsigned long long)S << 32) | (0xFFFFFFFFu - id);
                    kb = key > kb ? key : kb;
                }
                kb = wg_max(kb, red);
                if (tid == 0) {
                    lmax[c1] = kb ? (int32_t)(kb >> 32) : -1;
                    lidx[c1] = kb ? 0xFFFFFFFFu - (u32)kb : 0xFFFFFFFFu;
                }
                __syncthreads();
                continue;
            }
            // join: size rule, merge into slot c1, record
            u32 choice = 0;
            if (b.rule == 2) {
                const u32 sa = size[j1], sb = size[j2];
                if (sa < b.min_size && (float)sa * b.min_ratio < (float)sb) choice = 1;
                else if (sb < b.min_size && (float)sb * b.min_ratio < (float)sa) choice = 2;
            }
            u32 *ra = rows + (size_t)c1 * 2 * ldw;
            const u32 *rb = rows + (size_t)c2 * 2 * ldw;
            u32 *out = nst + (size_t)(nodeCount - n) * 2 * ldw;
            for (u32 w = tid; w < ldw; w += 1024) {
                const u32 a0 = ra[w], a1 = ra[ldw + w], b0 = rb[w], b1 = rb[ldw + w];
                u32 r0, r1;
                if (choice == 1) { r0 = b0; r1 = b1; }
                else if (choice == 2) { r0 = a0; r1 = a1; }
                else merge_words(b.rule, a0, a1, b0, b1, r0, r1);
                ra[w] = r0; ra[ldw + w] = r1;
                out[w] = r0; out[ldw + w] = r1;
            }
            if (tid == 0) {
                const u32 k = nodeCount - n;
                edge[4 * k + 0] = nodeCount; edge[4 * k + 1] = j1;
                edge[4 * k + 2] = nodeCount; edge[4 * k + 3] = j2;
                n2s[j1] = n; n2s[j2] = n; n2s[nodeCount] = c1;
                s2n[c1] = nodeCount; s2n[c2] = nnode;
                lmax[c1] = -1; lmax[c2] = -1;
                size[nodeCount] = size[j1] + size[j2];
            }
            __syncthreads();
            // rescore the new node against every live slot (strict '>')
            for (u32 k = grp; k < n; k += 64) {
                if (k == c1 || s2n[k] >= nnode) continue;
                const u32 S = group16_score(b, q, rows + (size_t)k * 2 * ldw, l16);
                if (l16 == 0 && (int32_t)S > lmax[k]) {
                    lmax[k] = (int32_t)S;
                    lidx[k] = nodeCount;
                }
            }
            ++nodeCount;
            __syncthreads();
        }
        if (tid == 0) b.status[job] = status;
    }
}

u32 grid_for(u64 items) { return (u32)std::min<u64>((items + 255) / 256, 65535ull * 8); }

// Host-side layout of one-hot indel encoding.
struct IndelLayout {
    std::vector<u32> loci, width;
    std::vector<int32_t> kmax;
    u32 nbits = 0, ldw = 0;
    DevBuf dloci, dwidth;
};

IndelLayout indel_layout(const int32_t *dmat, u32 n, u32 ncol, hipStream_t s)
{
    IndelLayout L;
    DevBuf dk((size_t)ncol * 4);
    launch_fill_u32(dk.as<u32>(), ncol, 0, s);
    hipLaunchKernelGGL(k_site_max, dim3(grid_for((u64)n * ncol)), dim3(256), 0, s, dmat, n, ncol,
                       dk.as<int32_t>());
    CASPHY_HIP(hipGetLastError());
    L.kmax.resize(ncol);
    d2h(L.kmax.data(), dk.p, (size_t)ncol * 4, s);
    L.loci.resize(ncol);
    L.width.resize(ncol);
    u64 off = 0;
    for (u32 c = 0; c < ncol; ++c) {
        L.loci[c] = (u32)off;
        L.width[c] = std::max<u32>((u32)L.kmax[c], 1u);
        off += L.width[c];
    }
    CASPHY_REQUIRE(off < (1ull << 31), CASPHY_EARG, "indel one-hot encoding too wide");
    L.nbits = (u32)off;
    L.ldw = round_up(ceil_div(L.nbits, 32), KC_BIT);
    L.dloci.alloc((size_t)ncol * 4);
    L.dwidth.alloc((size_t)ncol * 4);
    h2d(L.dloci.p, L.loci.data(), (size_t)ncol * 4, s);
    h2d(L.dwidth.p, L.width.data(), (size_t)ncol * 4, s);
    return L;
}

// Per-bit weights of the one-hot encoding: bit loci[s]+a-1 has weight[a-1][s].
Planes indel_planes(const IndelLayout &L, const double *weight, u32 nstate, u32 ncol)
{
    if (!weight) return Planes{};
    std::vector<double> wb((size_t)L.ldw * 32, 0.0);
    for (u32 s = 0; s < ncol; ++s)
        for (int32_t a = 1; a <= L.kmax[s]; ++a) {
            CASPHY_REQUIRE((u32)a <= nstate, CASPHY_EARG,
                           "allele " + std::to_string(a) + " at site " + std::to_string(s) +
                               " exceeds the weight matrix rows (nState)");
            wb[L.loci[s] + a - 1] = weight[(size_t)(a - 1) * ncol + s];
        }
    // validate every supplied weight even if its allele does not occur
    for (u64 i = 0; i < (u64)nstate * ncol; ++i)
        if (!std::isfinite(weight[i]) || weight[i] < 0 || weight[i] != std::floor(weight[i]))
            throw Error(CASPHY_EWEIGHT, "indel weights must be non-negative integers");
    return make_planes_bits(wb, L.ldw);
}

struct Tree {
    u32 n = 0, ldw = 0;
    int rule = 1;
    DevBuf rows, lmax, lidx, s2n, n2s, size, edge, nstate, ctl, keys, pmask, part;
    JoinArgs args{};
};

void tree_alloc(Tree &t, u32 n, u32 ldw, int rule)
{
    t.n = n; t.ldw = ldw; t.rule = rule;
    const u32 nnode = 2 * n - 1;
    t.rows.alloc((size_t)n * 2 * ldw * 4);
    t.lmax.alloc((size_t)n * 4);
    t.lidx.alloc((size_t)n * 4);
    t.s2n.alloc((size_t)n * 4);
    t.n2s.alloc((size_t)nnode * 4);
    t.size.alloc((size_t)nnode * 4);
    t.edge.alloc((size_t)(2 * n - 2) * 2 * 4);
    t.nstate.alloc((size_t)(n - 1) * 2 * ldw * 4);
    t.ctl.alloc(sizeof(Ctl));
    t.keys.alloc((size_t)n * 8);
    t.part.alloc((size_t)ceil_div(n, 16) * 8);
}

// Initial local maxima (all-pairs upper triangle with the LOCALMAX epilogue)
// and the join loop.
void tree_run(Tree &t, const Planes &planes, const u32 *dtip_size, u32 min_size, float min_ratio,
              hipStream_t s)
{
    const u32 n = t.n, ldw = t.ldw;
    JoinArgs &a = t.args;
    a.rows = t.rows.as<u32>(); a.lmax = t.lmax.as<int32_t>(); a.lidx = t.lidx.as<u32>();
    a.slot2node = t.s2n.as<u32>(); a.node2slot = t.n2s.as<u32>(); a.size = t.size.as<u32>();
    a.edge = t.edge.as<u32>(); a.nstate = t.nstate.as<u32>(); a.ctl = t.ctl.as<Ctl>();
    a.part = t.part.as<unsigned long long>();
    a.n = n; a.ldw = ldw; a.rule = t.rule; a.min_size = min_size; a.min_ratio = min_ratio;
    a.nplanes = planes.nplanes;
    if (planes.nplanes) {
        t.pmask.alloc(planes.mask.size() * 4);
        h2d(t.pmask.p, planes.mask.data(), planes.mask.size() * 4, s);
        a.pmask = t.pmask.as<u32>();
        for (u32 q = 0; q < planes.nplanes; ++q) a.pshift[q] = planes.shift[q];
        a.ptail = planes.tail;
    }

    launch_fill_u64(t.keys.as<u64>(), n, 0, s);
    if (planes.nplanes == 0 && mfma_enabled()) {
        DevBuf stage((size_t)stage_rows(n) * ldw * 4);
        launch_to_stage_major(a.rows, n, 2 * ldw, ldw, stage.as<u32>(), s);
        MfmaArgs q;
        q.PA = q.PB = stage.as<u32>(); q.padA = q.padB = stage_rows(n); q.stages = ldw / 4;
        q.nA = n; q.nB = n;
        q.keys = t.keys.as<u64>();
        launch_mfma_tiles(E_LOCALMAX, q, s);
        CASPHY_HIP(hipStreamSynchronize(s));
    } else {
    TileArgs ta;
    ta.A = a.rows; ta.nA = n; ta.B = a.rows; ta.nB = n;
    ta.ldw = 2 * ldw; ta.kw = ldw;
    ta.ntA = ta.ntB = ceil_div(n, TILE);
    ta.triangle = 1;
    ta.ntiles = (u32)((u64)ta.ntA * (ta.ntA + 1) / 2);
    ta.keys = t.keys.as<u64>();
    ta.nplanes = planes.nplanes;
    ta.pmask = a.pmask;
    for (u32 q = 0; q < planes.nplanes; ++q) ta.pshift[q] = planes.shift[q];
    ta.ptail = planes.tail;
    launch_tiles(M_AND, E_LOCALMAX, ta, s);
    }
    hipLaunchKernelGGL(k_join_init, dim3(grid_for(2 * (u64)n)), dim3(256), 0, s, a, t.keys.as<u64>(), dtip_size);
    CASPHY_HIP(hipGetLastError());

    // Join loop: CH steps per captured graph, poll the done flag per chunk.
    const u32 CH = 64;
    // one group of 16 slots per workgroup up to 2,048 workgroups (8 per CU),
    // then the workgroups loop over the live slots
    // (CASPHY_JOIN_RESCORE_WG lowers the cap, so small joins take the loop too)
    const char *cap_env = std::getenv("CASPHY_JOIN_RESCORE_WG");
    const u32 cap = cap_env ? std::max(1, std::atoi(cap_env)) : 2048;
    const u32 rescore_grid = std::min<u32>(ceil_div(n, 16), cap);
    a.nparts = rescore_grid;
    const char *env = std::getenv("CASPHY_JOIN_GRAPH");
    const bool use_graph = !(env && env[0] == '0');
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    // First selection by a full scan, then per step a rescoring launch that
    // leaves per-workgroup maxima and a one-workgroup selection over them.
    // CASPHY_JOIN_FUSED=1 runs both in one launch (the last workgroup to
    // arrive selects); its per-workgroup device-scope release fences cost more
    // than the second launch on this part, so it is not the default.
    const char *fenv = std::getenv("CASPHY_JOIN_FUSED");
    const bool fused = fenv && fenv[0] == '1';
    hipLaunchKernelGGL(k_select<false>, dim3(1), dim3(1024), 0, s, a);
    CASPHY_HIP(hipGetLastError());
    auto enqueue_chunk = [&]() {
        for (u32 i = 0; i < CH; ++i) {
            if (fused) {
                hipLaunchKernelGGL(k_step, dim3(rescore_grid), dim3(256), 0, s, a);
            } else {
                hipLaunchKernelGGL(k_rescore, dim3(rescore_grid), dim3(256), 0, s, a);
                hipLaunchKernelGGL(k_select<true>, dim3(1), dim3(1024), 0, s, a);
            }
        }
    };
    if (use_graph) {
        CASPHY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_chunk();
        CASPHY_HIP(hipStreamEndCapture(s, &graph));
        CASPHY_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    Ctl h{};
    const u64 max_steps = 64ull * n + 4096;
    u64 launched = 0;
    try {
        for (;;) {
            if (use_graph) CASPHY_HIP(hipGraphLaunch(exec, s));
            else enqueue_chunk();
            CASPHY_HIP(hipGetLastError());
            launched += CH;
            // poll every few chunks; the steps after completion are no-ops
            if ((launched / CH) % 4 == 0 || launched >= n) {
                d2h(&h, a.ctl, sizeof(Ctl), s);
                if (h.done) break;
                CASPHY_REQUIRE(launched < max_steps, CASPHY_EINTERNAL, "join did not converge");
            }
        }
    } catch (...) {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    if (exec) CASPHY_HIP(hipGraphExecDestroy(exec));
    if (graph) CASPHY_HIP(hipGraphDestroy(graph));
    CASPHY_REQUIRE(!h.err, CASPHY_EINTERNAL, "join found no live pair");
    const char *st = std::getenv("CASPHY_JOIN_STATS");
    if (st && st[0] == '1')
        std::fprintf(stderr, "join stats: n=%u selections=%u launched=%llu\n", n, h.steps,
                     (unsigned long long)launched);
}


// Runs the persistent batch kernel on encoded rows (sum n_k slots, 2 planes).
void batch_run(DevBuf &rows, const std::vector<u32> &offs, u32 ldw, int rule, const Planes &planes,
               const u32 *dtip, u32 min_size, float min_ratio, u32 *edge_host, hipStream_t s)
{
    const u32 njobs = (u32)offs.size() - 1;
    const u64 total = offs.back();
    const u64 nodes = 2 * total - njobs, edges = 2 * total - 2 * (u64)njobs, internal = total - njobs;
    DevBuf lmax(total * 4), lidx(total * 4), s2n(total * 4), n2s(nodes * 4), size(nodes * 4),
        dedge(std::max<u64>(edges, 1) * 8), nst(std::max<u64>(internal, 1) * 2 * ldw * 4),
        doff(offs.size() * 4), ctr(4), status((size_t)njobs * 4), pm;
    h2d(doff.p, offs.data(), offs.size() * 4, s);
    launch_fill_u32(ctr.as<u32>(), 1, 0, s);
    launch_fill_u32(status.as<u32>(), njobs, 0, s);
    BatchArgs b{};
    b.rows = rows.as<u32>(); b.lmax = lmax.as<int32_t>(); b.lidx = lidx.as<u32>();
    b.slot2node = s2n.as<u32>(); b.node2slot = n2s.as<u32>(); b.size = size.as<u32>();
    b.edge = dedge.as<u32>(); b.nstate = nst.as<u32>(); b.offsets = doff.as<u32>();
    b.tip_size = dtip; b.njobs = njobs; b.ldw = ldw; b.rule = rule;
    b.min_size = min_size; b.min_ratio = min_ratio; b.next_job = ctr.as<u32>(); b.status = status.as<u32>();
    b.nplanes = planes.nplanes;
    if (planes.nplanes) {
        pm.alloc(planes.mask.size() * 4);
        h2d(pm.p, planes.mask.data(), planes.mask.size() * 4, s);
        b.pmask = pm.as<u32>();
        for (u32 q = 0; q < planes.nplanes; ++q) b.pshift[q] = planes.shift[q];
        b.ptail = planes.tail;
    }
    int dev = 0, cus = 0;
    CASPHY_HIP(hipGetDevice(&dev));
    CASPHY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const u32 grid = std::min<u32>(njobs, (u32)std::max(cus, 1) * 2);
    hipLaunchKernelGGL(k_join_batch, dim3(grid), dim3(1024), 0, s, b);
    CASPHY_HIP(hipGetLastError());
    std::vector<u32> st(njobs);
    d2h(st.data(), status.p, (size_t)njobs * 4, s);
    for (u32 k = 0; k < njobs; ++k)
        CASPHY_REQUIRE(st[k] == 1, CASPHY_EINTERNAL, "batched join " + std::to_string(k) + " did not complete");
    d2h(edge_host, dedge.p, edges * 8, s);
}

void require(bool c, const char *msg) { CASPHY_REQUIRE(c, CASPHY_EARG, msg); }

}  // namespace

void join_bit_dev(const u32 *raw, u32 n, u32 W, int ancestral_type, u32 *edge_host, hipStream_t s)
{
    const u32 ldw = round_up(W, KC_BIT);
    Tree t;
    tree_alloc(t, n, ldw, ancestral_type);
    hipLaunchKernelGGL(k_bit_slots, dim3(grid_for((u64)n * 2 * ldw)), dim3(256), 0, s, raw, n, W, t.rows.as<u32>(),
                       ldw);
    CASPHY_HIP(hipGetLastError());
    tree_run(t, Planes{}, nullptr, 0, 2.0f, s);
    d2h(edge_host, t.edge.p, (size_t)(2 * n - 2) * 2 * 4, s);
}

void join_indel_dev(const int32_t *raw, u32 n, u32 ncol, u32 *edge_host, hipStream_t s)
{
    IndelLayout L = indel_layout(raw, n, ncol, s);
    Tree t;
    tree_alloc(t, n, L.ldw, 2);
    launch_fill_u32(t.rows.as<u32>(), (u64)n * 2 * L.ldw, 0, s);
    hipLaunchKernelGGL(k_indel_encode, dim3(grid_for((u64)n * ncol)), dim3(256), 0, s, raw, n, ncol,
                       L.dloci.as<u32>(), L.dwidth.as<u32>(), t.rows.as<u32>(), L.ldw);
    CASPHY_HIP(hipGetLastError());
    tree_run(t, Planes{}, nullptr, 0, 2.0f, s);
    d2h(edge_host, t.edge.p, (size_t)(2 * n - 2) * 2 * 4, s);
}

void join_bit_batch_dev(const u32 *raw, const std::vector<u32> &offs, u32 W, int ancestral_type, u32 *edge_host,
                        hipStream_t s)
{
    const u32 ldw = round_up(W, KC_BIT), total = offs.back();
    DevBuf rows((size_t)total * 2 * ldw * 4);
    hipLaunchKernelGGL(k_bit_slots, dim3(grid_for((u64)total * 2 * ldw)), dim3(256), 0, s, raw, total, W,
                       rows.as<u32>(), ldw);
    CASPHY_HIP(hipGetLastError());
    batch_run(rows, offs, ldw, ancestral_type, Planes{}, nullptr, 0, 2.0f, edge_host, s);
}

void join_indel_batch_dev(const int32_t *raw, const std::vector<u32> &offs, u32 ncol, u32 *edge_host,
                          hipStream_t s)
{
    const u32 total = offs.back();
    // one layout for every job: per-site allele groups sized by the maximum over all jobs
    IndelLayout L = indel_layout(raw, total, ncol, s);
    DevBuf rows((size_t)total * 2 * L.ldw * 4);
    launch_fill_u32(rows.as<u32>(), (u64)total * 2 * L.ldw, 0, s);
    hipLaunchKernelGGL(k_indel_encode, dim3(grid_for((u64)total * ncol)), dim3(256), 0, s, raw, total, ncol,
                       L.dloci.as<u32>(), L.dwidth.as<u32>(), rows.as<u32>(), L.ldw);
    CASPHY_HIP(hipGetLastError());
    batch_run(rows, offs, L.ldw, 2, Planes{}, nullptr, 0, 2.0f, edge_host, s);
}

}  // namespace casphy

using namespace casphy;

extern "C" {

int casphy_join_bit(const uint32_t *mat, uint32_t n, uint32_t W, const double *weight, int ancestral_type,
                    uint32_t *edge, uint32_t *node_state, uint32_t *node_bistate)
{
    return guarded([&] {
        require(mat && edge && W > 0, "casphy_join_bit: bad arguments");
        require(n >= 2, "casphy_join_bit: at least two cells are needed");
        require(ancestral_type == 0 || ancestral_type == 1, "casphy_join_bit: ancestralType must be 0 or 1");
        hipStream_t s = thread_stream();
        const u32 ldw = round_up(W, KC_BIT);
        Planes planes = make_planes(weight, 32 * W, ldw);
        Tree t;
        tree_alloc(t, n, ldw, ancestral_type);
        DevBuf raw((size_t)n * W * 4);
        h2d(raw.p, mat, (size_t)n * W * 4, s);
        hipLaunchKernelGGL(k_bit_slots, dim3(grid_for((u64)n * 2 * ldw)), dim3(256), 0, s, raw.as<u32>(), n, W,
                           t.rows.as<u32>(), ldw);
        CASPHY_HIP(hipGetLastError());
        tree_run(t, planes, nullptr, 0, 2.0f, s);
        d2h(edge, t.edge.p, (size_t)(2 * n - 2) * 2 * 4, s);
        if (node_state || node_bistate) {
            std::vector<u32> ns((size_t)(n - 1) * 2 * ldw);
            d2h(ns.data(), t.nstate.p, ns.size() * 4, s);
            for (u32 v = 0; v < 2 * n - 1; ++v)
                for (u32 w = 0; w < W; ++w) {
                    u32 sq, bs;
                    if (v < n) { sq = mat[(size_t)v * W + w]; bs = 0; }
                    else { sq = ns[(size_t)(v - n) * 2 * ldw + w]; bs = ns[(size_t)(v - n) * 2 * ldw + ldw + w]; }
                    if (node_state) node_state[(size_t)v * W + w] = sq & ~bs;
                    if (node_bistate) node_bistate[(size_t)v * W + w] = bs;
                }
        }
    });
}

int casphy_join_indel(const int32_t *mat, uint32_t n, uint32_t ncol, const double *weight, uint32_t nstate,
                      const uint32_t *tip_size, uint32_t min_size, float min_ratio, uint32_t *edge,
                      int32_t *node_state, uint32_t *branch)
{
    return guarded([&] {
        require(mat && edge && ncol > 0, "casphy_join_indel: bad arguments");
        require(n >= 2, "casphy_join_indel: at least two cells are needed");
        require(!weight || nstate > 0, "casphy_join_indel: weights need nState > 0");
        hipStream_t s = thread_stream();
        DevBuf dmat((size_t)n * ncol * 4);
        h2d(dmat.p, mat, (size_t)n * ncol * 4, s);
        IndelLayout L = indel_layout(dmat.as<int32_t>(), n, ncol, s);
        Planes planes = indel_planes(L, weight, nstate, ncol);
        Tree t;
        tree_alloc(t, n, L.ldw, 2);
        launch_fill_u32(t.rows.as<u32>(), (u64)n * 2 * L.ldw, 0, s);
        hipLaunchKernelGGL(k_indel_encode, dim3(grid_for((u64)n * ncol)), dim3(256), 0, s, dmat.as<int32_t>(), n,
                           ncol, L.dloci.as<u32>(), L.dwidth.as<u32>(), t.rows.as<u32>(), L.ldw);
        CASPHY_HIP(hipGetLastError());
        DevBuf dts;
        if (tip_size) {
            dts.alloc((size_t)n * 4);
            h2d(dts.p, tip_size, (size_t)n * 4, s);
        }
        tree_run(t, planes, tip_size ? dts.as<u32>() : nullptr, min_size, min_ratio, s);
        d2h(edge, t.edge.p, (size_t)(2 * n - 2) * 2 * 4, s);
        if (node_state || branch) {
            DevBuf dec((size_t)(n - 1) * ncol * 4);
            hipLaunchKernelGGL(k_indel_decode, dim3(grid_for((u64)(n - 1) * ncol)), dim3(256), 0, s,
                               t.nstate.as<u32>(), n - 1, L.ldw, L.dloci.as<u32>(), L.dwidth.as<u32>(), ncol,
                               dec.as<int32_t>());
            CASPHY_HIP(hipGetLastError());
            std::vector<int32_t> ns((size_t)(2 * n - 1) * ncol);
            std::copy(mat, mat + (size_t)n * ncol, ns.begin());
            d2h(ns.data() + (size_t)n * ncol, dec.p, (size_t)(n - 1) * ncol * 4, s);
            if (node_state) std::copy(ns.begin(), ns.end(), node_state);
            if (branch)
                for (u32 e = 0; e < 2 * n - 2; ++e) {
                    const int32_t *p = ns.data() + (size_t)edge[2 * e] * ncol;
                    const int32_t *c = ns.data() + (size_t)edge[2 * e + 1] * ncol;
                    u32 cnt = 0;
                    for (u32 k = 0; k < ncol; ++k) cnt += (p[k] == 0 && c[k] > 0);
                    branch[e] = cnt;
                }
        }
    });
}

int casphy_score_indel_matrix(const int32_t *mat, uint32_t n, uint32_t ncol, const double *weight,
                              uint32_t nstate, int32_t *S)
{
    return guarded([&] {
        require(mat && S && n > 0 && ncol > 0, "casphy_score_indel_matrix: bad arguments");
        require(!weight || nstate > 0, "casphy_score_indel_matrix: weights need nState > 0");
        hipStream_t s = thread_stream();
        DevBuf dmat((size_t)n * ncol * 4);
        h2d(dmat.p, mat, (size_t)n * ncol * 4, s);
        IndelLayout L = indel_layout(dmat.as<int32_t>(), n, ncol, s);
        Planes planes = indel_planes(L, weight, nstate, ncol);
        DevBuf rows((size_t)n * 2 * L.ldw * 4), dS((size_t)n * n * 4), pbuf;
        launch_fill_u32(rows.as<u32>(), (u64)n * 2 * L.ldw, 0, s);
        hipLaunchKernelGGL(k_indel_encode, dim3(grid_for((u64)n * ncol)), dim3(256), 0, s, dmat.as<int32_t>(), n,
                           ncol, L.dloci.as<u32>(), L.dwidth.as<u32>(), rows.as<u32>(), L.ldw);
        CASPHY_HIP(hipGetLastError());
        TileArgs a;
        a.A = rows.as<u32>(); a.nA = n; a.B = a.A; a.nB = n;
        a.ldw = 2 * L.ldw; a.kw = L.ldw;
        a.ntA = a.ntB = ceil_div(n, TILE);
        a.triangle = 1;
        a.ntiles = (u32)((u64)a.ntA * (a.ntA + 1) / 2);
        a.S = dS.as<int32_t>(); a.ldS = n;
        apply_planes(planes, a, pbuf, s);
        launch_tiles(M_AND, E_SCORES, a, s);
        d2h(S, dS.p, (size_t)n * n * 4, s);
    });
}

int casphy_join_bit_batch(const uint32_t *mat, const uint32_t *offsets, uint32_t njobs, uint32_t W,
                          const double *weight, int ancestral_type, uint32_t *edge)
{
    return guarded([&] {
        require(mat && offsets && edge && W > 0, "casphy_join_bit_batch: bad arguments");
        require(ancestral_type == 0 || ancestral_type == 1, "casphy_join_bit_batch: ancestralType must be 0 or 1");
        if (njobs == 0) return;
        std::vector<u32> offs(offsets, offsets + njobs + 1);
        for (u32 k = 0; k < njobs; ++k)
            require(offs[k + 1] >= offs[k] + 2, "casphy_join_bit_batch: every job needs >= 2 cells");
        require(offs[0] == 0, "casphy_join_bit_batch: offsets[0] must be 0");
        hipStream_t s = thread_stream();
        const u32 ldw = round_up(W, KC_BIT);
        const u32 total = offs.back();
        Planes planes = make_planes(weight, 32 * W, ldw);
        DevBuf raw((size_t)total * W * 4), rows((size_t)total * 2 * ldw * 4);
        h2d(raw.p, mat, (size_t)total * W * 4, s);
        hipLaunchKernelGGL(k_bit_slots, dim3(grid_for((u64)total * 2 * ldw)), dim3(256), 0, s, raw.as<u32>(), total,
                           W, rows.as<u32>(), ldw);
        CASPHY_HIP(hipGetLastError());
        batch_run(rows, offs, ldw, ancestral_type, planes, nullptr, 0, 2.0f, edge, s);
    });
}

int casphy_join_indel_batch(const int32_t *mat, const uint32_t *offsets, uint32_t njobs, uint32_t ncol,
                            const double *weight, uint32_t nstate, const uint32_t *tip_size, uint32_t min_size,
                            float min_ratio, uint32_t *edge)
{
    return guarded([&] {
        require(mat && offsets && edge && ncol > 0, "casphy_join_indel_batch: bad arguments");
        require(!weight || nstate > 0, "casphy_join_indel_batch: weights need nState > 0");
        if (njobs == 0) return;
        std::vector<u32> offs(offsets, offsets + njobs + 1);
        require(offs[0] == 0, "casphy_join_indel_batch: offsets[0] must be 0");
        for (u32 k = 0; k < njobs; ++k)
            require(offs[k + 1] >= offs[k] + 2, "casphy_join_indel_batch: every job needs >= 2 cells");
        hipStream_t s = thread_stream();
        const u32 total = offs.back();
        DevBuf dmat((size_t)total * ncol * 4);
        h2d(dmat.p, mat, (size_t)total * ncol * 4, s);
        // one layout for every job: per-site allele groups sized by the maximum over all jobs
        IndelLayout L = indel_layout(dmat.as<int32_t>(), total, ncol, s);
        Planes planes = indel_planes(L, weight, nstate, ncol);
        DevBuf rows((size_t)total * 2 * L.ldw * 4), dts;
        launch_fill_u32(rows.as<u32>(), (u64)total * 2 * L.ldw, 0, s);
        hipLaunchKernelGGL(k_indel_encode, dim3(grid_for((u64)total * ncol)), dim3(256), 0, s, dmat.as<int32_t>(),
                           total, ncol, L.dloci.as<u32>(), L.dwidth.as<u32>(), rows.as<u32>(), L.ldw);
        CASPHY_HIP(hipGetLastError());
        if (tip_size) {
            dts.alloc((size_t)total * 4);
            h2d(dts.p, tip_size, (size_t)total * 4, s);
        }
        batch_run(rows, offs, L.ldw, 2, planes, tip_size ? dts.as<u32>() : nullptr, min_size, min_ratio, edge, s);
    });
}

}  // extern "C"
