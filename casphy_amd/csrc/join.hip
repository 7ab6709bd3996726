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
#include <cstring>
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
// Per-site maximum allele: a thread owns one column over the rows y, y +
// gridDim.y, ... (coalesced across the threads of a row) and issues one
// atomicMax at the end (one per site and workgroup row, not one per entry:
// with a few hundred sites the entries' atomics all met on those addresses).
__global__ void k_site_max(const int32_t *mat, u32 n, u32 ncol, int32_t *kmax)
{
    const u32 c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    int32_t m = 0;
    for (u32 r = blockIdx.y; r < n; r += gridDim.y) m = max(m, mat[(size_t)r * ncol + c]);
    if (m > 0) atomicMax(&kmax[c], m);
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
// One-launch join step (single tree, the default form).  Every workgroup of a
// step repeats the selection from the previous step's per-workgroup results,
// so no workgroup waits for another and no device-scope fence is needed inside
// a launch: the per-slot state a step reads (parity p) and the state it writes
// (parity 1 - p) are different buffers, a new node's row goes to a fresh row
// of `nstate` (tip rows are never written), and the few node -> slot changes
// of a step travel in the step's control block: workgroup 0 of the next step
// writes them into node2slot while every reader overrides node2slot with them.
// Each workgroup also leaves, for its best slot, the partner node and that
// node's slot, so the next selection needs no dependent lookups.
// Semantics are those of k_select / k_rescore (and of the CPU twins): the
// maximal (score, smallest node id) pair, stale partners recomputed over every
// live slot, strict '>' on rescoring, live slots kept compact (the last live
// slot moves into the second child's slot).
constexpr u32 FS_INIT = 0, FS_JOIN = 1, FS_RECOMPUTE = 2;
constexpr u32 FS_LPS = 8;             // lanes per slot
constexpr u32 FS_SPP = 256 / FS_LPS;  // slots per workgroup pass

struct FCtl {
    u32 nodeCount, nlive;
    u32 mode;      // what the writing step did (FS_*)
    u32 c, cnode;  // FS_RECOMPUTE: the slot recomputed and its node
    u32 done, err, steps;
    u32 nupd;      // node -> slot changes of the writing step
    u32 upd_node[4], upd_slot[4];
};

struct FMeta {  // a workgroup's best slot, its partner node and that node's slot
    u32 slot, j2, c2, pad;
};

// Row formats of the one-launch step: FMT_PLANES, two bit planes of ldw words
// (binary: state and bistate; indel: one-hot alleles and missing flags);
// FMT_CODES (indel, unit weights), one byte per site of ldw words: 0 unmutated,
// 1..254 the allele, 255 missing.  A node's indel state holds at most one
// allele per site (a tip one; a merge keeps the common allele or takes the
// other child's where one is missing), so the byte is the one-hot group.
constexpr int FMT_PLANES = 0, FMT_CODES = 1;
constexpr u32 CODE_MISSING = 0xFFu;

struct FArgs {
    const u32 *rows;  // tip rows: n x rstride words
    u32 *nstate;      // internal node rows: (n - 1) x rstride words
    int32_t *lmax[2];
    u32 *lidx[2], *s2n[2];
    unsigned long long *key[2];   // per workgroup: best (lmax + 1, ~node) over its slots
    FMeta *meta[2];
    unsigned long long *rkey[2];  // recompute: per workgroup best (score, ~node)
    u32 *rslot[2];                // ... and that node's slot
    u32 *node2slot, *size, *edge;
    FCtl *ctl[2];
    u32 n, ldw, nwg;
    u32 rstride;  // words per row: 2 ldw (planes) or ldw (codes)
    int rule;
    u32 min_size;
    float min_ratio;
    u32 nplanes;
    const u32 *pmask;
    u32 pshift[MAX_PLANES];
    u32 ptail;
};

__device__ __forceinline__ const u32 *frow(const FArgs &a, u32 node)
{
    return node < a.n ? a.rows + (size_t)node * a.rstride : a.nstate + (size_t)(node - a.n) * a.rstride;
}

// Per byte of a word: 0x80 where the byte is zero, else 0 (no carries across bytes).
__device__ __forceinline__ u32 zero_bytes(u32 t)
{
    return ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t | 0x7F7F7F7Fu);
}

// Sites of a code word where a and b hold the same allele (neither
// unmutated nor missing).
__device__ __forceinline__ u32 code_matches(u32 x, u32 y)
{
    return (u32)__builtin_popcount(zero_bytes(x ^ y) & ~zero_bytes(x) & ~zero_bytes(~x));
}

// Merge of two code words (the indel rule of src/functions.cpp:1965-1974 on
// one-hot groups): a missing site takes the other child's state, two present
// sites keep their common allele or become unmutated.
__device__ __forceinline__ u32 code_merge(u32 x, u32 y)
{
    const u32 mx = (zero_bytes(~x) >> 7) * 0xFFu, my = (zero_bytes(~y) >> 7) * 0xFFu;
    const u32 eq = (zero_bytes(x ^ y) >> 7) * 0xFFu;
    return (mx & y) | (~mx & my & x) | (~mx & ~my & x & eq);
}

// Sum over each aligned group of LPS lanes (4 .. 64), in every lane of it:
// quad swaps, the half-row mirror, a row rotation, then cross-row shuffles.
template <u32 LPS>
__device__ __forceinline__ u32 sum_lanes(u32 v)
{
    v += __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xF, 0xF, false);                            // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xF, 0xF, false);                            // quad_perm [2,3,0,1]
    if constexpr (LPS >= 8) v += __builtin_amdgcn_update_dpp(0u, v, 0x141, 0xF, 0xF, false);   // row_half_mirror
    if constexpr (LPS >= 16) v += __builtin_amdgcn_update_dpp(0u, v, 0x128, 0xF, 0xF, false);  // row_ror:8
    if constexpr (LPS >= 32) v += __shfl_xor(v, 16);
    if constexpr (LPS >= 64) v += __shfl_xor(v, 32);
    return v;
}

// Score of the LDS row q against the global row r (plane 0, weight planes) by
// the LPS lanes of a slot (lane l reads words 4 l + 4 LPS i); every lane of
// the group returns the total.
template <u32 LPS, int FMT>
__device__ __forceinline__ u32 fscore(const FArgs &a, const u32 *q, const u32 *r, u32 l)
{
    if constexpr (FMT == FMT_CODES) {
        u32 part = 0;
#pragma unroll 4
        for (u32 w = 4 * l; w < a.ldw; w += 4 * LPS) {
            const uint4 x = *reinterpret_cast<const uint4 *>(q + w);
            const uint4 y = *reinterpret_cast<const uint4 *>(r + w);
            part += code_matches(x.x, y.x) + code_matches(x.y, y.y) + code_matches(x.z, y.z) +
                    code_matches(x.w, y.w);
        }
        return sum_lanes<LPS>(part);
    }
    const u32 np = a.nplanes ? a.nplanes : 1;
    u32 S = 0;
    for (u32 p = 0; p < np; ++p) {
        if (p) S <<= a.pshift[p];
        u32 part = 0;
#pragma unroll 4
        for (u32 w = 4 * l; w < a.ldw; w += 4 * LPS) {
            uint4 x = *reinterpret_cast<const uint4 *>(q + w);
            const uint4 y = *reinterpret_cast<const uint4 *>(r + w);
            if (a.nplanes) {
                const uint4 m = *reinterpret_cast<const uint4 *>(a.pmask + (size_t)p * a.ldw + w);
                x.x &= m.x; x.y &= m.y; x.z &= m.z; x.w &= m.w;
            }
            part += __builtin_popcount(x.x & y.x) + __builtin_popcount(x.y & y.y) +
                    __builtin_popcount(x.z & y.z) + __builtin_popcount(x.w & y.w);
        }
        S += sum_lanes<LPS>(part);
    }
    if (a.nplanes) S <<= a.ptail;
    return S;
}

// Maximum of a u64 over the wave (every lane active): a prefix maximum by DPP
// row shifts and row broadcasts (zero filled), read from lane 63.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long k)
{
#define CASPHY_MAX_STEP(ctrl, rmask)                                                                 \
    {                                                                                                \
        const u32 lo_ = __builtin_amdgcn_update_dpp(0u, (u32)k, ctrl, rmask, 0xF, false);            \
        const u32 hi_ = __builtin_amdgcn_update_dpp(0u, (u32)(k >> 32), ctrl, rmask, 0xF, false);    \
        const unsigned long long o_ = ((unsigned long long)hi_ << 32) | lo_;                         \
        k = o_ > k ? o_ : k;                                                                         \
    }
    CASPHY_MAX_STEP(0x111, 0xF)  // row_shr:1
    CASPHY_MAX_STEP(0x112, 0xF)  // row_shr:2
    CASPHY_MAX_STEP(0x114, 0xF)  // row_shr:4
    CASPHY_MAX_STEP(0x118, 0xF)  // row_shr:8
    CASPHY_MAX_STEP(0x142, 0xA)  // row_bcast:15 -> rows 1, 3
    CASPHY_MAX_STEP(0x143, 0xC)  // row_bcast:31 -> rows 2, 3
#undef CASPHY_MAX_STEP
    // readlane returns int: through u32, so the low word is not sign-extended
    return ((unsigned long long)(u32)__builtin_amdgcn_readlane((u32)(k >> 32), 63) << 32) |
           (unsigned long long)(u32)__builtin_amdgcn_readlane((u32)k, 63);
}

// Workgroup (256 threads) argmax of a u64 key carrying three u32 payloads; the
// result lands in every thread.  Keys are distinct or carry equal payloads.
struct FRed {
    unsigned long long k[4];
    u32 x[4], y[4], z[4];
};
__device__ __forceinline__ void fmax3(unsigned long long &k, u32 &x, u32 &y, u32 &z, FRed &r)
{
    const unsigned long long wm = wave_max_u64(k);
    const u32 src = (u32)__builtin_ctzll(__ballot(k == wm));
    const u32 wx = __builtin_amdgcn_readlane(x, src), wy = __builtin_amdgcn_readlane(y, src),
              wz = __builtin_amdgcn_readlane(z, src);
    __syncthreads();  // the previous use of r is over
    if ((threadIdx.x & 63) == 0) {
        const u32 w = threadIdx.x >> 6;
        r.k[w] = wm; r.x[w] = wx; r.y[w] = wy; r.z[w] = wz;
    }
    __syncthreads();
    k = r.k[0]; x = r.x[0]; y = r.y[0]; z = r.z[0];
    for (int w = 1; w < 4; ++w)
        if (r.k[w] > k) { k = r.k[w]; x = r.x[w]; y = r.y[w]; z = r.z[w]; }
}

// Initial per-workgroup keys and partners (state in parity 0, written by
// k_join_init) and the first control block.
__global__ __launch_bounds__(256) void k_fstep_init(FArgs a)
{
    __shared__ FRed red;
    const u32 tid = threadIdx.x, n = a.n;
    unsigned long long bk = 0;
    u32 bs = n, bj = 0xFFFFFFFFu, bc = n;
    for (u32 base = blockIdx.x * FS_SPP; base < n; base += a.nwg * FS_SPP) {
        const u32 k = base + (tid / FS_LPS);
        if (k < n && tid % FS_LPS == 0) {
            const unsigned long long key = ((unsigned long long)(u32)(a.lmax[0][k] + 1) << 32) |
                                           (0xFFFFFFFFu - a.s2n[0][k]);
            if (key > bk) {
                bk = key; bs = k; bj = a.lidx[0][k];
                bc = bj < 2 * n - 1 ? a.node2slot[bj] : n;
            }
        }
    }
    fmax3(bk, bs, bj, bc, red);
    if (tid == 0) {
        a.key[0][blockIdx.x] = bk;
        a.meta[0][blockIdx.x] = FMeta{bs, bj, bc, 0};
        if (blockIdx.x == 0) {
            FCtl c{};
            c.nodeCount = n; c.nlive = n; c.mode = FS_INIT; c.c = n; c.cnode = 0;
            c.done = n < 2 ? 1u : 0u;
            *a.ctl[0] = c;
        }
    }
}

// One step; LPS lanes score a slot.
template <u32 LPS, int FMT>
__global__ __launch_bounds__(256) void k_fstep(FArgs a, u32 p)
{
    extern __shared__ __attribute__((aligned(16))) u32 fq[];  // the step's query row: plane 0 (and 1 for a join)
    __shared__ FRed red;
    constexpr u32 SPP = 256 / LPS;  // slots per workgroup pass
    const u32 tid = threadIdx.x, R = p, W = 1 - p;
    const u32 n = a.n, ldw = a.ldw, NONE = 0xFFFFFFFFu, nnode = 2 * n - 1;
    const u32 l = tid % LPS;
    // ---- selection, the same in every workgroup: each workgroup's best key
    // comes with its slot, partner and partner slot
    unsigned long long best = 0xFFFFFFFFull;  // value -1, node 0: no pair
    u32 c1 = n, j2 = NONE, c2 = n;
    for (u32 b = tid; b < a.nwg; b += 256) {
        const unsigned long long k = a.key[R][b];
        const FMeta m = a.meta[R][b];
        if (k > best) { best = k; c1 = m.slot; j2 = m.j2; c2 = m.c2; }
    }
    const FCtl C = *a.ctl[R];
    if (C.done) {
        if (blockIdx.x == 0 && tid == 0) *a.ctl[W] = C;
        return;
    }
    fmax3(best, c1, j2, c2, red);
    const u32 fslot = C.mode == FS_RECOMPUTE ? C.c : n;  // the slot whose maximum the last step recomputed
    unsigned long long fk = 0;
    u32 fks = n, d0 = 0, d1 = 0;
    if (fslot < n) {
        for (u32 b = tid; b < a.nwg; b += 256) {
            const unsigned long long k = a.rkey[R][b];
            const u32 rs = a.rslot[R][b];
            if (k > fk) { fk = k; fks = rs; }
        }
        fmax3(fk, fks, d0, d1, red);
    }
    const int32_t fv = fk ? (int32_t)(u32)(fk >> 32) : -1;
    const u32 fj = fk ? 0xFFFFFFFFu - (u32)fk : NONE;
    if (fslot < n) {
        const unsigned long long kf = ((unsigned long long)(u32)(fv + 1) << 32) | (0xFFFFFFFFu - C.cnode);
        if (kf > best) {
            best = kf;
            c1 = fslot; j2 = fj; c2 = fk ? fks : n;
        }
    }
    const int32_t v = (int32_t)(u32)(best >> 32) - 1;
    const u32 j1 = 0xFFFFFFFFu - (u32)best;
    const u32 verdict = (v < 0 || c1 >= n) ? 0u : (c2 < n ? FS_JOIN : FS_RECOMPUTE);
    if (verdict == 0) {  // no live pair: cannot happen with >= 2 live nodes
        if (blockIdx.x == 0 && tid == 0) {
            FCtl c = C;
            c.done = 1; c.err = 1; c.nupd = 0;
            *a.ctl[W] = c;
        }
        return;
    }
    const int32_t *lmR = a.lmax[R];
    const u32 *liR = a.lidx[R], *snR = a.s2n[R];
    int32_t *lmW = a.lmax[W];
    u32 *liW = a.lidx[W], *snW = a.s2n[W];
    const u32 last = C.nlive - 1;
    unsigned long long bk = 0, rbk = 0;
    u32 bs = n, bj = NONE, bc = n, rbs = n;
    // slot of a partner node once the changes of this step (cn/cs) and of the
    // previous one (C) are in, else node2slot
    auto slot_of = [&](u32 node, const u32 *cn, const u32 *cs, u32 ncur) {
        if (node >= nnode) return n;
        for (u32 i = 0; i < ncur; ++i)
            if (cn[i] == node) return cs[i];
        for (u32 i = 0; i < C.nupd; ++i)
            if (C.upd_node[i] == node) return C.upd_slot[i];
        return a.node2slot[node];
    };
    if (verdict == FS_JOIN) {
        const u32 nc = C.nodeCount, nl = last;
        const u32 s_nc = c1 == last ? c2 : c1;
        const u32 mv = (c2 != last && c1 != last) ? snR[last] : NONE;  // node moving into slot c2
        u32 cn[4] = {j1, j2, nc, mv}, cs[4] = {n, n, s_nc, c2};
        const u32 ncur = mv != NONE ? 4u : 3u;
        // merged row of the new node (cu/joinEdgeIndel.cu:407-413 size rule)
        u32 choice = 0;
        if (a.rule == 2) {
            const u32 sa = a.size[j1], sb = a.size[j2];
            if (sa < a.min_size && (float)sa * a.min_ratio < (float)sb) choice = 1;
            else if (sb < a.min_size && (float)sb * a.min_ratio < (float)sa) choice = 2;
        }
        const u32 *ra = frow(a, j1), *rb = frow(a, j2);
        if constexpr (FMT == FMT_CODES) {
            for (u32 w = tid; w < ldw; w += 256) {
                const u32 x = ra[w], y = rb[w];
                fq[w] = choice == 1 ? y : choice == 2 ? x : code_merge(x, y);
            }
        } else {
            for (u32 w = tid; w < ldw; w += 256) {
                const u32 a0 = ra[w], a1 = ra[ldw + w], b0 = rb[w], b1 = rb[ldw + w];
                u32 r0, r1;
                if (choice == 1) { r0 = b0; r1 = b1; }
                else if (choice == 2) { r0 = a0; r1 = a1; }
                else merge_words(a.rule, a0, a1, b0, b1, r0, r1);
                fq[w] = r0;
                fq[ldw + w] = r1;
            }
        }
        __syncthreads();
        if (blockIdx.x == 0) {
            u32 *out = a.nstate + (size_t)(nc - n) * a.rstride;
            for (u32 w = tid; w < a.rstride; w += 256) out[w] = fq[w];
            if (tid == 0) {
                const u32 e = nc - n;
                a.edge[4 * e + 0] = nc; a.edge[4 * e + 1] = j1;
                a.edge[4 * e + 2] = nc; a.edge[4 * e + 3] = j2;
                a.size[nc] = a.size[j1] + a.size[j2];
                for (u32 i = 0; i < C.nupd; ++i) a.node2slot[C.upd_node[i]] = C.upd_slot[i];
                FCtl c{};
                c.nodeCount = nc + 1; c.nlive = nl; c.mode = FS_JOIN; c.c = n; c.cnode = 0;
                c.done = nc + 1 == 2 * n - 1 ? 1u : 0u; c.err = 0; c.steps = C.steps + 1; c.nupd = ncur;
                for (u32 i = 0; i < 4; ++i) { c.upd_node[i] = cn[i]; c.upd_slot[i] = cs[i]; }
                *a.ctl[W] = c;
            }
        }
        // the new node against every live slot; the state moves to parity W
        for (u32 base = blockIdx.x * SPP; base < nl; base += a.nwg * SPP) {
            const u32 k = base + tid / LPS;
            if (k >= nl) continue;  // the whole slot group
            const u32 o = (k == c2 && c2 != last) ? last : k;  // old slot of the node now in k
            u32 node, li;
            int32_t lm;
            if (o == c1) {
                node = nc; lm = -1; li = NONE;
            } else {
                node = snR[o];
                if (o == fslot) { lm = fv; li = fj; }
                else { lm = lmR[o]; li = liR[o]; }
                const u32 S = fscore<LPS, FMT>(a, fq, frow(a, node), l);
                if ((int32_t)S > lm) { lm = (int32_t)S; li = nc; }
            }
            if (l == 0) { lmW[k] = lm; liW[k] = li; snW[k] = node; }
            const unsigned long long key = ((unsigned long long)(u32)(lm + 1) << 32) | (0xFFFFFFFFu - node);
            if (key > bk) {
                bk = key; bs = k; bj = li;
                bc = slot_of(li, cn, cs, ncur);
            }
        }
        fmax3(bk, bs, bj, bc, red);
        if (tid == 0) {
            a.key[W][blockIdx.x] = bk;
            a.meta[W][blockIdx.x] = FMeta{bs, bj, bc, 0};
        }
        return;
    }
    // ---- recompute: the maximum of slot c1 (node j1) over every other live slot
    const u32 *rq = frow(a, j1);
    for (u32 w = tid; w < ldw; w += 256) fq[w] = rq[w];
    __syncthreads();
    if (blockIdx.x == 0 && tid == 0) {
        for (u32 i = 0; i < C.nupd; ++i) a.node2slot[C.upd_node[i]] = C.upd_slot[i];
        FCtl c{};
        c.nodeCount = C.nodeCount; c.nlive = C.nlive; c.mode = FS_RECOMPUTE; c.c = c1; c.cnode = j1;
        c.done = 0; c.err = 0; c.steps = C.steps + 1; c.nupd = 0;
        *a.ctl[W] = c;
    }
    for (u32 base = blockIdx.x * SPP; base < C.nlive; base += a.nwg * SPP) {
        const u32 k = base + tid / LPS;
        if (k >= C.nlive) continue;
        const u32 node = snR[k];
        int32_t lm;
        u32 li;
        if (k == fslot) { lm = fv; li = fj; }
        else { lm = lmR[k]; li = liR[k]; }
        if (l == 0) { lmW[k] = lm; liW[k] = li; snW[k] = node; }
        if (k == c1) continue;
        const u32 S = fscore<LPS, FMT>(a, fq, frow(a, node), l);
        const unsigned long long r = ((unsigned long long)S << 32) | (0xFFFFFFFFu - node);
        if (r > rbk) { rbk = r; rbs = k; }
        const unsigned long long key = ((unsigned long long)(u32)(lm + 1) << 32) | (0xFFFFFFFFu - node);
        if (key > bk) {
            bk = key; bs = k; bj = li;
            bc = slot_of(li, nullptr, nullptr, 0);
        }
    }
    u32 e0 = 0, e1 = 0;
    fmax3(rbk, rbs, e0, e1, red);
    fmax3(bk, bs, bj, bc, red);
    if (tid == 0) {
        a.rkey[W][blockIdx.x] = rbk;
        a.rslot[W][blockIdx.x] = rbs;
        a.key[W][blockIdx.x] = bk;
        a.meta[W][blockIdx.x] = FMeta{bs, bj, bc, 0};
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
    u32 rstride;          // words per row: 2 ldw (planes) or ldw (byte codes)
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

// Scores of row q against NR rows r[] (ldw words each, plane 0), by LPS
// cooperating lanes; every lane of the group gets the totals.  A lane loads
// its (up to) four 16-byte chunks of every row before it counts, so 4 NR loads
// are in flight (rows of more than 16 LPS words take several rounds).
template <int FMT, u32 LPS, u32 NR>
__device__ __forceinline__ void group_scores(const BatchArgs &b, const u32 *q, const u32 *const (&r)[NR], u32 l,
                                             u32 (&S)[NR])
{
    const u32 np = FMT == FMT_CODES ? 1u : (b.nplanes ? b.nplanes : 1);
#pragma unroll
    for (u32 k = 0; k < NR; ++k) S[k] = 0;
    for (u32 p = 0; p < np; ++p) {
        u32 part[NR];
#pragma unroll
        for (u32 k = 0; k < NR; ++k) part[k] = 0;
        for (u32 w0 = 4 * l; w0 < b.ldw; w0 += 16 * LPS) {
            uint4 x[4], y[NR][4];
#pragma unroll
            for (u32 i = 0; i < 4; ++i) {
                const u32 w = w0 + 4 * LPS * i;
                const bool in = w < b.ldw;
                x[i] = in ? *reinterpret_cast<const uint4 *>(q + w) : make_uint4(0, 0, 0, 0);
#pragma unroll
                for (u32 k = 0; k < NR; ++k)
                    y[k][i] = in ? *reinterpret_cast<const uint4 *>(r[k] + w) : make_uint4(0, 0, 0, 0);
                if (FMT == FMT_PLANES && b.nplanes && in) {
                    const uint4 m = *reinterpret_cast<const uint4 *>(b.pmask + (size_t)p * b.ldw + w);
                    x[i].x &= m.x; x[i].y &= m.y; x[i].z &= m.z; x[i].w &= m.w;
                }
            }
#pragma unroll
            for (u32 i = 0; i < 4; ++i)
#pragma unroll
                for (u32 k = 0; k < NR; ++k) {
                    if constexpr (FMT == FMT_CODES)
                        part[k] += code_matches(x[i].x, y[k][i].x) + code_matches(x[i].y, y[k][i].y) +
                                   code_matches(x[i].z, y[k][i].z) + code_matches(x[i].w, y[k][i].w);
                    else
                        part[k] += __builtin_popcount(x[i].x & y[k][i].x) + __builtin_popcount(x[i].y & y[k][i].y) +
                                   __builtin_popcount(x[i].z & y[k][i].z) + __builtin_popcount(x[i].w & y[k][i].w);
                }
        }
#pragma unroll
        for (u32 k = 0; k < NR; ++k) {
            if (FMT == FMT_PLANES && b.nplanes && p) S[k] <<= b.pshift[p];
            S[k] += sum_lanes<LPS>(part[k]);
        }
    }
    if (FMT == FMT_PLANES && b.nplanes)
#pragma unroll
        for (u32 k = 0; k < NR; ++k) S[k] <<= b.ptail;
}

// Workgroup-wide max of a u64 key (1024 threads).
__device__ __forceinline__ unsigned long long wg_max(unsigned long long k, unsigned long long *red)
{
    k = wave_max_u64(k);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = k;
    __syncthreads();
    unsigned long long m = red[0];
    for (int w = 1; w < 16; ++w) m = red[w] > m ? red[w] : m;
    return m;
}

// LPS lanes score a pair of rows (G = 1024 / LPS groups); the row loops are
// unrolled so that every group keeps several rows' loads in flight.
template <int FMT, u32 LPS>
__global__ __launch_bounds__(1024) void k_join_batch(BatchArgs b)
{
    __shared__ unsigned long long red[16];
    __shared__ u32 sh[8];
    constexpr u32 G = 1024 / LPS;
    const u32 tid = threadIdx.x, l = tid % LPS, grp = tid / LPS;
    const u32 ldw = b.ldw;
    for (;;) {
        __syncthreads();
        if (tid == 0) sh[0] = atomicAdd(b.next_job, 1u);
        __syncthreads();
        const u32 job = sh[0];
        if (job >= b.njobs) return;
        const u32 off = b.offsets[job], n = b.offsets[job + 1] - off, nnode = 2 * n - 1;
        const u32 rs = b.rstride;
        u32 *rows = b.rows + (size_t)off * rs;
        int32_t *lmax = b.lmax + off;
        u32 *lidx = b.lidx + off, *s2n = b.slot2node + off;
        u32 *n2s = b.node2slot + (2 * (size_t)off - job);
        u32 *size = b.size + (2 * (size_t)off - job);
        u32 *edge = b.edge + 2 * (2 * (size_t)off - 2 * (size_t)job);
        u32 *nst = b.nstate + ((size_t)off - job) * rs;

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
        // initial local maxima: group g scores rows i = g, g+G, ... against all j > i
        for (u32 i = grp; i < n; i += G) {
            unsigned long long best = 0;
            const u32 *qi = rows + (size_t)i * rs;
            for (u32 j = i + 1; j < n; j += 2) {  // two rows at a time (the second may be past n)
                const u32 j1 = j + 1 < n ? j + 1 : j;
                const u32 *rr[2] = {rows + (size_t)j * rs, rows + (size_t)j1 * rs};
                u32 S[2];
                group_scores<FMT, LPS, 2>(b, qi, rr, l, S);
                const unsigned long long k0 = ((unsigned long long)S[0] << 32) | (0xFFFFFFFFu - j);
                const unsigned long long k1 = ((unsigned long long)S[1] << 32) | (0xFFFFFFFFu - j1);
                best = k0 > best ? k0 : best;
                best = k1 > best ? k1 : best;
            }
            if (l == 0) {
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
            const u32 *q = rows + (size_t)c1 * rs;
            if (verdict == 1) {
                // recompute row c1 over every live slot
                unsigned long long kb = 0;
                for (u32 k = grp; k < n; k += 2 * G) {  // slots k and k + G together
                    const u32 k1 = k + G;
                    const u32 id0 = s2n[k], id1 = k1 < n ? s2n[k1] : nnode;
                    const bool v0 = k != c1 && id0 < nnode, v1 = k1 != c1 && id1 < nnode;
                    if (!v0 && !v1) continue;
                    const u32 *rr[2] = {rows + (size_t)(v0 ? k : k1) * rs, rows + (size_t)(v1 ? k1 : k) * rs};
                    u32 S[2];
                    group_scores<FMT, LPS, 2>(b, q, rr, l, S);
                    const unsigned long long k0 = v0 ? ((unsigned long long)S[0] << 32) | (0xFFFFFFFFu - id0) : 0;
                    const unsigned long long kk = v1 ? ((unsigned long long)S[1] << 32) | (0xFFFFFFFFu - id1) : 0;
                    kb = k0 > kb ? k0 : kb;
                    kb = kk > kb ? kk : kb;
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
            u32 *ra = rows + (size_t)c1 * rs;
            const u32 *rb = rows + (size_t)c2 * rs;
            u32 *out = nst + (size_t)(nodeCount - n) * rs;
            for (u32 w = tid; w < ldw; w += 1024) {
                if constexpr (FMT == FMT_CODES) {
                    const u32 x = ra[w], y = rb[w];
                    const u32 r = choice == 1 ? y : choice == 2 ? x : code_merge(x, y);
                    ra[w] = r;
                    out[w] = r;
                } else {
                    const u32 a0 = ra[w], a1 = ra[ldw + w], b0 = rb[w], b1 = rb[ldw + w];
                    u32 r0, r1;
                    if (choice == 1) { r0 = b0; r1 = b1; }
                    else if (choice == 2) { r0 = a0; r1 = a1; }
                    else merge_words(b.rule, a0, a1, b0, b1, r0, r1);
                    ra[w] = r0; ra[ldw + w] = r1;
                    out[w] = r0; out[ldw + w] = r1;
                }
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
            for (u32 k = grp; k < n; k += 2 * G) {  // slots k and k + G together
                const u32 k1 = k + G;
                const bool v0 = k != c1 && s2n[k] < nnode, v1 = k1 < n && k1 != c1 && s2n[k1] < nnode;
                if (!v0 && !v1) continue;
                const u32 *rr[2] = {rows + (size_t)(v0 ? k : k1) * rs, rows + (size_t)(v1 ? k1 : k) * rs};
                u32 S[2];
                group_scores<FMT, LPS, 2>(b, q, rr, l, S);
                if (l == 0 && v0 && (int32_t)S[0] > lmax[k]) {
                    lmax[k] = (int32_t)S[0];
                    lidx[k] = nodeCount;
                }
                if (l == 0 && v1 && (int32_t)S[1] > lmax[k1]) {
                    lmax[k1] = (int32_t)S[1];
                    lidx[k1] = nodeCount;
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
    hipLaunchKernelGGL(k_site_max, dim3(ceil_div(ncol, 256u), std::max<u32>(1, std::min<u32>(n, 1024))), dim3(256), 0, s, dmat, n, ncol,
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

// Indel states -> byte codes (0 unmutated, 1..254 allele, 255 missing), 4
// sites a word, rows of cw words (zero past ncol).
__global__ void k_indel_codes(const int32_t *mat, u32 n, u32 ncol, u32 *out, u32 cw)
{
    const u64 total = (u64)n * cw;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 r = (u32)(x / cw), w = (u32)(x - (u64)r * cw);
        u32 word = 0;
        for (u32 b = 0; b < 4; ++b) {
            const u32 s = 4 * w + b;
            if (s < ncol) {
                const int32_t v = mat[(size_t)r * ncol + s];
                const u32 c = v > 0 ? (u32)v : v < 0 ? CODE_MISSING : 0u;
                word |= c << (8 * b);
            }
        }
        out[x] = word;
    }
}

// Replicate r of a column resample: rows r * n .. r * n + n - 1 of out are the
// n rows of mat at the columns cols[r * ncol ..] (a bootstrap replicate).
__global__ void k_gather_cols(const int32_t *mat, u32 n, u32 ncol, const u32 *cols, u32 nrep, int32_t *out)
{
    const u64 total = (u64)nrep * n * ncol;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u64 row = x / ncol;
        const u32 c = (u32)(x - row * ncol), r = (u32)(row / n), i = (u32)(row - (u64)r * n);
        out[x] = mat[(size_t)i * ncol + cols[(size_t)r * ncol + c]];
    }
}

// Byte codes of `count` rows back to states (-1 missing).
__global__ void k_codes_decode(const u32 *codes, u32 count, u32 cw, u32 ncol, int32_t *out)
{
    const u64 total = (u64)count * ncol;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 r = (u32)(x / ncol), s = (u32)(x - (u64)r * ncol);
        const u32 c = (codes[(size_t)r * cw + s / 4] >> (8 * (s % 4))) & 0xFFu;
        out[x] = c == CODE_MISSING ? -1 : (int32_t)c;
    }
}

struct Tree {
    u32 n = 0, ldw = 0;
    int rule = 1;
    DevBuf rows, lmax, lidx, s2n, n2s, size, edge, nstate, ctl, keys, pmask, part;
    // indel joins on byte codes (FMT_CODES): the source states, their width,
    // and the code rows of the tips and of the internal nodes
    const int32_t *code_src = nullptr;
    u32 ncol = 0, cw = 0;
    bool codes = false;
    DevBuf ctips, cnodes;
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

// The join loop in one launch per step (k_fstep), after k_join_init has put
// the initial maxima into parity 0 of the per-slot state.

thread_local u32 g_join_wg_cap = 0;

// Workgroups of the one-launch step for n slots (slots_per_wg per pass).
// Every workgroup repeats the selection over all workgroups' results of the
// previous step, so past a few hundred workgroups that redundant work outweighs
// the rescoring parallelism.  Measured single-join times (ms, binary / indel
// codes): 10,000 cells 126 / 164 at 313 workgroups, 121 / 155 at 256; 30,000
// cells 562 / 709 at 938, 464 / 568 at 256, 462 / 581 at 384; 100,000 cells
// 2,835 / 2,994 at 1,024, 2,595 / 2,764 at 512, 3,101 / 3,033 at 256.  Four
// concurrent 30,000-cell joins: 2.42 joins/s at 938, 4.95 at 256, 4.29 at 512.
// CASPHY_JOIN_WG overrides both rules.
u32 fstep_workgroups(u32 n, u32 slots_per_wg)
{
    const char *wenv = std::getenv("CASPHY_JOIN_WG");
    u32 cap = 256 + n / 400;
    if (g_join_wg_cap) cap = std::min(cap, g_join_wg_cap);
    if (wenv) cap = (u32)std::max(1, std::atoi(wenv));
    return std::max<u32>(1, std::min<u32>({ceil_div(n, slots_per_wg), cap, 1024u}));
}

void fstep_run(Tree &t, const Planes &planes, u32 min_size, float min_ratio, hipStream_t s)
{
    const u32 n = t.n;
    if (t.codes) {  // indel rows as byte codes: one word of 4 sites, 16-byte multiples
        t.cw = round_up(ceil_div(t.ncol, 4), 4u);
        t.ctips.alloc((size_t)n * t.cw * 4);
        t.cnodes.alloc((size_t)std::max(n - 1, 1u) * t.cw * 4);
        hipLaunchKernelGGL(k_indel_codes, dim3(grid_for((u64)n * t.cw)), dim3(256), 0, s, t.code_src, n, t.ncol,
                           t.ctips.as<u32>(), t.cw);
        CASPHY_HIP(hipGetLastError());
    }
    const u32 ldw = t.codes ? t.cw : t.ldw;
    // 8 lanes per slot, 32 slots per workgroup pass: measured faster than 16,
    // 32 or 64 lanes per slot for every row width (more slots in flight beat
    // fewer serial 16-byte loads per lane)
    const u32 lps = 8;
    const u32 nwg = fstep_workgroups(n, 256 / lps);
    DevBuf lm1((size_t)n * 4), li1((size_t)n * 4), sn1((size_t)n * 4), keys(2 * (size_t)nwg * 8),
        meta(2 * (size_t)nwg * sizeof(FMeta)), rkeys(2 * (size_t)nwg * 8), rslots(2 * (size_t)nwg * 4),
        ctl(2 * sizeof(FCtl));
    FArgs f{};
    if (t.codes) {
        f.rows = t.ctips.as<u32>(); f.nstate = t.cnodes.as<u32>(); f.rstride = ldw;
    } else {
        f.rows = t.rows.as<u32>(); f.nstate = t.nstate.as<u32>(); f.rstride = 2 * ldw;
    }
    f.lmax[0] = t.lmax.as<int32_t>(); f.lmax[1] = lm1.as<int32_t>();
    f.lidx[0] = t.lidx.as<u32>(); f.lidx[1] = li1.as<u32>();
    f.s2n[0] = t.s2n.as<u32>(); f.s2n[1] = sn1.as<u32>();
    for (int q = 0; q < 2; ++q) {
        f.key[q] = keys.as<unsigned long long>() + (size_t)q * nwg;
        f.meta[q] = reinterpret_cast<FMeta *>(meta.p) + (size_t)q * nwg;
        f.rkey[q] = rkeys.as<unsigned long long>() + (size_t)q * nwg;
        f.rslot[q] = rslots.as<u32>() + (size_t)q * nwg;
        f.ctl[q] = reinterpret_cast<FCtl *>(ctl.p) + q;
    }
    f.node2slot = t.n2s.as<u32>(); f.size = t.size.as<u32>(); f.edge = t.edge.as<u32>();
    f.n = n; f.ldw = ldw; f.nwg = nwg;
    f.rule = t.rule; f.min_size = min_size; f.min_ratio = min_ratio;
    f.nplanes = planes.nplanes;
    f.pmask = t.args.pmask;
    for (u32 q = 0; q < planes.nplanes; ++q) f.pshift[q] = planes.shift[q];
    f.ptail = planes.tail;
    hipLaunchKernelGGL(k_fstep_init, dim3(nwg), dim3(256), 0, s, f);
    CASPHY_HIP(hipGetLastError());
    const u32 CH = 64;  // steps per captured graph: even, so every chunk starts on parity 0
    const size_t lds = (size_t)f.rstride * 4;
    auto step = [&](u32 par) {
        if (t.codes) hipLaunchKernelGGL((k_fstep<8, FMT_CODES>), dim3(nwg), dim3(256), lds, s, f, par);
        else hipLaunchKernelGGL((k_fstep<8, FMT_PLANES>), dim3(nwg), dim3(256), lds, s, f, par);
    };
    auto enqueue_chunk = [&]() {
        for (u32 i = 0; i < CH; ++i) step(i & 1u);
    };
    const char *genv = std::getenv("CASPHY_JOIN_GRAPH");
    const bool use_graph = !(genv && genv[0] == '0');
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    if (use_graph) {
        CASPHY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_chunk();
        CASPHY_HIP(hipStreamEndCapture(s, &graph));
        CASPHY_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    FCtl h{};
    const u64 max_steps = 64ull * n + 4096;
    u64 launched = 0;
    try {
        for (;;) {
            if (use_graph) CASPHY_HIP(hipGraphLaunch(exec, s));
            else enqueue_chunk();
            CASPHY_HIP(hipGetLastError());
            launched += CH;
            // poll every few chunks (the chunk's last step wrote parity 0); the
            // steps after completion only carry the control block over
            if ((launched / CH) % 4 == 0 || launched >= n) {
                d2h(&h, f.ctl[0], sizeof(FCtl), s);
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
        std::fprintf(stderr, "join stats: n=%u selections=%u launched=%llu (one launch per step, %u workgroups%s)\n",
                     n, h.steps, (unsigned long long)launched, nwg, t.codes ? ", byte codes" : "");
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
    // One launch per step (k_fstep) unless CASPHY_JOIN_STEP=two asks for the
    // rescoring + selection pair.  Measured per join on MI355X (ms, one-launch
    // vs two-launch step): binary rows of 128 words 11.5 / 12.7 at 1,024 cells,
    // 575 / 654 at 30,000, 3,163 / 3,291 at 100,000; one-hot indel rows of 384
    // words 216 / 228 at 10,000 but 1,208 / 1,085 at 30,000 and 7,481 / 6,502
    // at 100,000, where the 2,048 rescoring workgroups of the two-launch step
    // keep more rows in flight.  So wide rows of large joins take two launches.
    const char *senv = std::getenv("CASPHY_JOIN_STEP");
    const bool force_one = senv && !std::strcmp(senv, "one");
    // Indel joins of unit weights whose alleles fit a byte take byte codes
    // (t.code_src set by the caller): rows of 4 sites a word instead of the
    // one-hot planes, so the step reads a fraction of the bytes (200 sites:
    // 52 words against 352-608 one-hot words a plane).
    // (CASPHY_JOIN_CODES=0 keeps the one-hot planes)
    const char *cenv = std::getenv("CASPHY_JOIN_CODES");
    t.codes = t.code_src && t.rule == 2 && planes.nplanes == 0 && !(cenv && cenv[0] == '0') &&
              !(senv && !std::strcmp(senv, "two")) && !fused;
    const bool two = !t.codes && ((senv && !std::strcmp(senv, "two")) || fused || (!force_one && ldw > 256 && n > 16384));
    if (t.codes || (!two && (size_t)2 * ldw * 4 <= 64 * 1024)) return fstep_run(t, planes, min_size, min_ratio, s);
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
// codes: the rows are indel byte codes of ldw words (FMT_CODES), else two planes.
void batch_run(DevBuf &rows, const std::vector<u32> &offs, u32 ldw, int rule, const Planes &planes,
               const u32 *dtip, u32 min_size, float min_ratio, u32 *edge_host, hipStream_t s, bool codes = false)
{
    const u32 njobs = (u32)offs.size() - 1;
    const u64 total = offs.back();
    const u64 nodes = 2 * total - njobs, edges = 2 * total - 2 * (u64)njobs, internal = total - njobs;
    DevBuf lmax(total * 4), lidx(total * 4), s2n(total * 4), n2s(nodes * 4), size(nodes * 4),
        dedge(std::max<u64>(edges, 1) * 8), nst(std::max<u64>(internal, 1) * (codes ? 1 : 2) * ldw * 4),
        doff(offs.size() * 4), ctr(4), status((size_t)njobs * 4), pm;
    h2d(doff.p, offs.data(), offs.size() * 4, s);
    launch_fill_u32(ctr.as<u32>(), 1, 0, s);
    launch_fill_u32(status.as<u32>(), njobs, 0, s);
    BatchArgs b{};
    b.rows = rows.as<u32>(); b.lmax = lmax.as<int32_t>(); b.lidx = lidx.as<u32>();
    b.slot2node = s2n.as<u32>(); b.node2slot = n2s.as<u32>(); b.size = size.as<u32>();
    b.edge = dedge.as<u32>(); b.nstate = nst.as<u32>(); b.offsets = doff.as<u32>();
    b.tip_size = dtip; b.njobs = njobs; b.ldw = ldw; b.rstride = codes ? ldw : 2 * ldw; b.rule = rule;
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
    // lanes per pair: the fewest whose 16-byte loads of a row are at most four a lane
    const u32 lps = ldw <= 64 ? 4u : ldw <= 128 ? 8u : ldw <= 256 ? 16u : ldw <= 512 ? 32u : 64u;
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(grid), dim3(1024), 0, s, b); };
    const char *st_env = std::getenv("CASPHY_JOIN_STATS");
    const bool stats = st_env && st_env[0] == '1';
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (stats) {
        CASPHY_HIP(hipEventCreate(&ev[0]));
        CASPHY_HIP(hipEventCreate(&ev[1]));
        CASPHY_HIP(hipEventRecord(ev[0], s));
    }
    if (codes) {
        switch (lps) {
        case 4: go(k_join_batch<FMT_CODES, 4>); break;
        case 8: go(k_join_batch<FMT_CODES, 8>); break;
        case 16: go(k_join_batch<FMT_CODES, 16>); break;
        case 32: go(k_join_batch<FMT_CODES, 32>); break;
        default: go(k_join_batch<FMT_CODES, 64>); break;
        }
    } else {
        switch (lps) {
        case 4: go(k_join_batch<FMT_PLANES, 4>); break;
        case 8: go(k_join_batch<FMT_PLANES, 8>); break;
        case 16: go(k_join_batch<FMT_PLANES, 16>); break;
        case 32: go(k_join_batch<FMT_PLANES, 32>); break;
        default: go(k_join_batch<FMT_PLANES, 64>); break;
        }
    }
    CASPHY_HIP(hipGetLastError());
    if (stats) {
        CASPHY_HIP(hipEventRecord(ev[1], s));
        CASPHY_HIP(hipEventSynchronize(ev[1]));
        float ms = 0;
        CASPHY_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        std::fprintf(stderr, "join stats: batch of %u joins, %u workgroups of 1024, %u lanes per pair%s: %.3f ms\n",
                     njobs, grid, lps, codes ? ", byte codes" : "", ms);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    std::vector<u32> st(njobs);
    d2h(st.data(), status.p, (size_t)njobs * 4, s);
    for (u32 k = 0; k < njobs; ++k)
        CASPHY_REQUIRE(st[k] == 1, CASPHY_EINTERNAL, "batched join " + std::to_string(k) + " did not complete");
    d2h(edge_host, dedge.p, edges * 8, s);
}

void require(bool c, const char *msg) { CASPHY_REQUIRE(c, CASPHY_EARG, msg); }


}  // namespace

void set_join_workgroup_cap(u32 cap) { g_join_wg_cap = cap; }

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

// Byte codes need every allele below the missing code.
bool codes_fit(const IndelLayout &L)
{
    for (int32_t k : L.kmax)
        if (k >= (int32_t)CODE_MISSING) return false;
    return true;
}

// CASPHY_JOIN_CODES=0 keeps indel joins on the one-hot planes.
bool no_codes_env()
{
    const char *e = std::getenv("CASPHY_JOIN_CODES");
    return e && e[0] == '0';
}

// No weights or all-one weights (what R passes by default): unit weights.
bool unit_weights(const double *weight, u32 nstate, u32 ncol)
{
    if (!weight) return true;
    for (u64 i = 0; i < (u64)nstate * ncol; ++i)
        if (weight[i] != 1.0) return false;
    return true;
}

// Batched indel joins over jobs whose rows are already on the device (dmat:
// offs.back() x ncol allele codes): byte codes when the weights are unit and
// every site has at most 254 alleles, else one-hot planes.
static void indel_batch_dev(DevBuf &dmat, const std::vector<u32> &offs, u32 ncol, const double *weight, u32 nstate,
                     const u32 *tip_size, u32 min_size, float min_ratio, u32 *edge, hipStream_t s)
{
    const u32 total = offs.back();
    // one layout for every job: per-site allele groups sized by the maximum over all jobs
    IndelLayout L = indel_layout(dmat.as<int32_t>(), total, ncol, s);
    Planes planes = indel_planes(L, weight, nstate, ncol);
    DevBuf dts;
    if (tip_size) {
        dts.alloc((size_t)total * 4);
        h2d(dts.p, tip_size, (size_t)total * 4, s);
    }
    if (unit_weights(weight, nstate, ncol) && codes_fit(L) && !no_codes_env()) {  // byte codes
        const u32 cw = round_up(ceil_div(ncol, 4), 4u);
        DevBuf crows((size_t)total * cw * 4);
        hipLaunchKernelGGL(k_indel_codes, dim3(grid_for((u64)total * cw)), dim3(256), 0, s, dmat.as<int32_t>(), total,
                           ncol, crows.as<u32>(), cw);
        CASPHY_HIP(hipGetLastError());
        batch_run(crows, offs, cw, 2, Planes{}, tip_size ? dts.as<u32>() : nullptr, min_size, min_ratio, edge, s,
                  true);
        return;
    }
    DevBuf rows((size_t)total * 2 * L.ldw * 4);
    launch_fill_u32(rows.as<u32>(), (u64)total * 2 * L.ldw, 0, s);
    hipLaunchKernelGGL(k_indel_encode, dim3(grid_for((u64)total * ncol)), dim3(256), 0, s, dmat.as<int32_t>(), total,
                       ncol, L.dloci.as<u32>(), L.dwidth.as<u32>(), rows.as<u32>(), L.ldw);
    CASPHY_HIP(hipGetLastError());
    batch_run(rows, offs, L.ldw, 2, planes, tip_size ? dts.as<u32>() : nullptr, min_size, min_ratio, edge, s);
}

void join_indel_dev(const int32_t *raw, u32 n, u32 ncol, u32 *edge_host, hipStream_t s)
{
    IndelLayout L = indel_layout(raw, n, ncol, s);
    Tree t;
    tree_alloc(t, n, L.ldw, 2);
    if (codes_fit(L)) { t.code_src = raw; t.ncol = ncol; }
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
    if (codes_fit(L) && !no_codes_env()) {  // byte codes
        const u32 cw = round_up(ceil_div(ncol, 4), 4u);
        DevBuf rows((size_t)total * cw * 4);
        hipLaunchKernelGGL(k_indel_codes, dim3(grid_for((u64)total * cw)), dim3(256), 0, s, raw, total, ncol,
                           rows.as<u32>(), cw);
        CASPHY_HIP(hipGetLastError());
        batch_run(rows, offs, cw, 2, Planes{}, nullptr, 0, 2.0f, edge_host, s, true);
        return;
    }
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
        if (unit_weights(weight, nstate, ncol)) planes = Planes{};  // all-one weights are unit weights
        Tree t;
        tree_alloc(t, n, L.ldw, 2);
        if (codes_fit(L)) { t.code_src = dmat.as<int32_t>(); t.ncol = ncol; }
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
            if (t.codes)
                hipLaunchKernelGGL(k_codes_decode, dim3(grid_for((u64)(n - 1) * ncol)), dim3(256), 0, s,
                                   t.cnodes.as<u32>(), n - 1, t.cw, ncol, dec.as<int32_t>());
            else
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
        indel_batch_dev(dmat, offs, ncol, weight, nstate, tip_size, min_size, min_ratio, edge, s);
    });
}

int casphy_join_indel_bootstrap(const int32_t *mat_cellmajor, uint32_t n, uint32_t ncol, const uint32_t *cols,
                                uint32_t nrep, uint32_t *edge)
{
    return guarded([&] {
        require(mat_cellmajor && cols && edge && n >= 2 && ncol > 0, "casphy_join_indel_bootstrap: bad arguments");
        if (nrep == 0) return;
        CASPHY_REQUIRE((u64)nrep * n < (1ull << 31), CASPHY_EARG, "casphy_join_indel_bootstrap: too many rows");
        for (u64 k = 0; k < (u64)nrep * ncol; ++k)
            CASPHY_REQUIRE(cols[k] < ncol, CASPHY_EARG,
                           "casphy_join_indel_bootstrap: column index " + std::to_string(cols[k]) + " >= ncol");
        hipStream_t s = thread_stream();
        DevBuf base((size_t)n * ncol * 4), dcols((size_t)nrep * ncol * 4), dmat((size_t)nrep * n * ncol * 4);
        h2d(base.p, mat_cellmajor, (size_t)n * ncol * 4, s);
        h2d(dcols.p, cols, (size_t)nrep * ncol * 4, s);
        hipLaunchKernelGGL(k_gather_cols, dim3(grid_for((u64)nrep * n * ncol)), dim3(256), 0, s, base.as<int32_t>(), n,
                           ncol, dcols.as<u32>(), nrep, dmat.as<int32_t>());
        CASPHY_HIP(hipGetLastError());
        std::vector<u32> offs(nrep + 1);
        for (u32 r = 0; r <= nrep; ++r) offs[r] = r * n;
        indel_batch_dev(dmat, offs, ncol, nullptr, 0, nullptr, 0, 2.0f, edge, s);
    });
}

}  // extern "C"
