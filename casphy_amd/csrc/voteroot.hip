// voteroot.hip -- root votes of outgroup samples (voteRoot2, findRoot).
//
// For every outgroup sample the reference counts, against every node's edge
// state (the OR of the node's and its parent's byte codes), the positions
// whose codes share a nucleotide bit, and keeps the node of the largest count
// (cu/voteRoot2.cu:34-169; src/functions.cpp:4170-4410).  A position's
// "share" test is an OR over bits, not a sum, so it is not an AND-popcount as
// it stands.  It becomes one per code: for the codes c that occur among the
// samples at position p, the sample side sets feature (p, c) when its code is
// c and the node side when its edge state shares a bit with c, so
//     [edge & sample has a base bit] = sum over c of [sample = c] [edge & c != 0]
// exactly, one term at most being non-zero.  Typical columns carry 2-5 codes,
// so the K dimension is a few times the positions and the counts run on the
// matrix cores (the bit-row MFMA tiles); a fused epilogue keeps, per sample,
// the largest (count, smallest candidate rank) key.
#include <algorithm>

#include "context.hpp"
#include "kernels.hpp"
#include "mfma.hpp"
#include "voteroot.hpp"

namespace casphy {

namespace {

// pres[p] |= 1 << (low nibble of the code at p) over the rows: which codes
// occur at each position (bit 0: a code without a base bit).  Lanes take
// words, the 4 waves of a workgroup interleave the rows of its slice of
// PRES_ROWS rows (four loads in flight), and the waves' masks meet in LDS
// before one atomic per position and workgroup.
constexpr u32 PRES_ROWS = 128;
__global__ __launch_bounds__(256) void k_code_presence(const u32 *rows, u32 n, u32 nbin, u32 *pres)
{
    __shared__ uint4 sm[4][64];
    const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u32 w = blockIdx.x * 64 + lane;
    const u32 r0 = blockIdx.y * PRES_ROWS, r1 = min(n, r0 + PRES_ROWS);
    u32 m[4] = {0, 0, 0, 0};
    if (w < nbin) {
        for (u32 r = r0 + wave; r < r1; r += 16) {
            u32 x[4];
#pragma unroll
            for (u32 k = 0; k < 4; ++k) x[k] = r + 4 * k < r1 ? rows[(size_t)(r + 4 * k) * nbin + w] : 0u;
#pragma unroll
            for (u32 k = 0; k < 4; ++k)
#pragma unroll
                for (u32 b = 0; b < 4; ++b) m[b] |= 1u << ((x[k] >> (8 * b)) & 15u);
        }
    }
    sm[wave][lane] = make_uint4(m[0], m[1], m[2], m[3]);
    __syncthreads();
    if (wave == 0 && w < nbin) {
        uint4 t = sm[0][lane];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            const uint4 u = sm[k][lane];
            t.x |= u.x; t.y |= u.y; t.z |= u.z; t.w |= u.w;
        }
        atomicOr(&pres[4 * w], t.x);
        atomicOr(&pres[4 * w + 1], t.y);
        atomicOr(&pres[4 * w + 2], t.z);
        atomicOr(&pres[4 * w + 3], t.w);
    }
}

// esT[w][r] = word w of the edge state of candidate cand[r] (its row | its
// parent's row, zero without a parent edge; zero for rows ncand .. pad), the
// word-major copy the operand packer reads.  32 x 32 tiles through LDS: reads
// along the words of a row, writes along the rows of a word.
__global__ __launch_bounds__(256) void k_edge_cols(const u32 *state, u32 nbin, const u32 *parent, const u32 *cand,
                                                   u32 ncand, u32 pad, u32 *esT)
{
    __shared__ u32 t[32][33];
    const u32 w0 = blockIdx.x * 32, r0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (u32 y = ty; y < 32; y += 8) {
        const u32 r = r0 + y, w = w0 + tx;
        u32 v = 0;
        if (r < ncand && w < nbin) {
            const u32 c = cand[r], pv = parent[c];
            if (pv != 0xFFFFFFFFu) v = state[(size_t)c * nbin + w] | state[(size_t)pv * nbin + w];
        }
        t[y][tx] = v;
    }
    __syncthreads();
    for (u32 y = ty; y < 32; y += 8) {
        const u32 w = w0 + y, r = r0 + tx;
        if (w < nbin && r < pad) esT[(size_t)w * pad + r] = t[tx][y];
    }
}

// The stage-major operand of the matrix-core tiles, written directly: stage
// s of rows 64 rb .. 64 rb + 63 is one wave (lane = row), P[s][r] = the bits
// of features 128 s .. 128 s + 127 of row r (feat[f] = p << 5 | c, sorted by
// p, padded to whole stages with c = 16, which no nibble matches):
// [byte at p & c != 0] for edge-state rows (EQ false), [low nibble at p == c]
// for sample rows.  The rows come word-major (T[w][r], pad rows, zero past
// the last), so a wave's load of one word is one contiguous 256-byte piece;
// the feature walk is wave-uniform (scalar loads of feat) and loads a word
// only when it reaches a new one.
template <bool EQ>
__global__ __launch_bounds__(256) void k_vote_stage(const u32 *__restrict__ T, u32 pad, const u32 *__restrict__ feat,
                                                    u32 stages, u32 *__restrict__ P)
{
    const u32 lane = threadIdx.x & 63, rblocks = pad / 64;
    const u32 items = rblocks * stages;  // below 2^31 (checked by the launcher): `it` cannot wrap
    // the item loop runs on scalar registers: every wave-level value below is uniform
    const u32 first = __builtin_amdgcn_readfirstlane((blockIdx.x * 256 + threadIdx.x) >> 6);
    for (u32 it = first; it < items; it += gridDim.x * 4) {
        const u32 st = it / rblocks, rb = it - st * rblocks, r = rb * 64 + lane;
        const u32 *fs = feat + (size_t)st * 128;
        u32 cw = 0xFFFFFFFFu, word = 0;
#pragma unroll 1
        for (u32 q = 0; q < 4; ++q) {
            u32 bits = 0;
#pragma unroll
            for (u32 b = 0; b < 32; ++b) {
                const u32 e = fs[32 * q + b], p = e >> 5, c = e & 31u, wi = p >> 2;
                if (wi != cw) {
                    cw = wi;
                    word = T[(size_t)wi * pad + r];
                }
                const u32 x = (word >> (8 * (p & 3u))) & 15u;
                bits |= (u32)(EQ ? x == c : (x & c) != 0u) << b;
            }
            P[((size_t)st * pad + r) * 4 + q] = bits;
        }
    }
}

// Workgroups for one wave per item.
u32 pack_grid(u64 waves) { return (u32)std::min<u64>(std::max<u64>(ceil_div(waves, (u64)4), 1), 65536); }

// Parent per node from 1-based (parent, child) pairs, checked.
std::vector<u32> parents(const u32 *edge, u32 nedge, u32 nnode)
{
    std::vector<u32> par(nnode, 0xFFFFFFFFu);
    for (u32 e = 0; e < nedge; ++e) {
        const u32 a = edge[2 * e], d = edge[2 * e + 1];
        CASPHY_REQUIRE(a >= 1 && a <= nnode && d >= 1 && d <= nnode, CASPHY_EARG,
                       "root vote: edge " + std::to_string(e) + " names a node outside 1.." + std::to_string(nnode));
        CASPHY_REQUIRE(par[d - 1] == 0xFFFFFFFFu, CASPHY_EARG,
                       "root vote: node " + std::to_string(d) + " is the child of two edges");
        par[d - 1] = a - 1;
    }
    return par;
}

}  // namespace

// cu/voteRoot2.cu:125-133,140-164: nodes in blocks of 8 reduced by halving
// strides (4, 2, 1) where the lower slot keeps ties, then lane L (of 32)
// scanning blocks L, L + 32, ... and keeping the first strictly larger one
// (from count 0, node 0), then the 32 lanes reduced by halving offsets (16 ..
// 1), the lower lane keeping ties.  A halving-stride tournament prefers even
// slots, then slots 0 mod 4, ...: among equal values the winner is the first
// in bit-reversed slot order.  A node of the largest count M > 0 therefore
// wins by (bit-reversed block mod 32, block / 32, bit-reversed node mod 8),
// and with M = 0 node 0 wins -- which is also first in that order.
RootVote plan_vote_voteroot2(const u32 *edge, u32 nedge, u32 nnode)
{
    RootVote p;
    p.parent = parents(edge, nedge, nnode);
    const u32 nblk = ceil_div(nnode, 8u);
    const auto rev = [](u32 x, int bits) {
        u32 r = 0;
        for (int k = 0; k < bits; ++k) r |= ((x >> k) & 1u) << (bits - 1 - k);
        return r;
    };
    p.candidates.reserve(nnode);
    for (u32 l = 0; l < 32; ++l)
        for (u32 b = rev(l, 5); b < nblk; b += 32)
            for (u32 y = 0; y < 8; ++y) {
                const u32 v = 8 * b + rev(y, 3);
                if (v < nnode) p.candidates.push_back(v);
            }
    return p;
}

// src/functions.cpp:4385-4398: the edge children in edge order, the first of
// the largest count kept.
RootVote plan_vote_findroot(const u32 *edge, u32 nedge, u32 nnode)
{
    RootVote p;
    p.parent = parents(edge, nedge, nnode);
    p.candidates.resize(nedge);
    for (u32 e = 0; e < nedge; ++e) p.candidates[e] = edge[2 * e + 1] - 1;
    return p;
}

void vote_root_keys(const RootVote &plan, const u32 *state, u32 nnode, u32 nbin, const u32 *samples, u32 nsample,
                    u64 *keys, hipStream_t s)
{
    const u32 ncand = (u32)plan.candidates.size();
    if (!nsample) return;
    if (!ncand) {
        std::fill(keys, keys + nsample, 0ull);
        return;
    }
    const u64 P = 4ull * nbin;
    DevBuf dsamp(std::max<size_t>((size_t)nsample * nbin * 4, 4)), dpres(std::max<size_t>(P * 4, 4));
    std::vector<u32> pres(P);
    if (P) {
        h2d(dsamp.p, samples, (size_t)nsample * nbin * 4, s);
        launch_fill_u32(dpres.as<u32>(), P, 0u, s);
        const dim3 grid(ceil_div(nbin, 64u), std::min(ceil_div(nsample, PRES_ROWS), 65535u));
        hipLaunchKernelGGL(k_code_presence, grid, dim3(256), 0, s, dsamp.as<u32>(), nsample, nbin, dpres.as<u32>());
        CASPHY_HIP(hipGetLastError());
        d2h(pres.data(), dpres.p, P * 4, s);
    }
    std::vector<u32> feat;
    CASPHY_REQUIRE(P < (1ull << 27), CASPHY_EARG, "root vote: rows beyond 2^25 words");
    for (u64 p = 0; p < P; ++p)
        for (u32 c = 1; c < 16; ++c)
            if ((pres[p] >> c) & 1u) feat.push_back((u32)(p << 5) | c);
    CASPHY_REQUIRE(feat.size() < (1ull << 31), CASPHY_EARG, "root vote: too many (position, code) features");
    const u32 K = (u32)feat.size();
    if (!K) {  // no sample has a base: every count is 0, rank 0 wins
        std::fill(keys, keys + nsample, 0xFFFFFFFFull);
        return;
    }
    const u32 ldw = round_up(ceil_div(K, 32u), (u32)KC_BIT);
    CASPHY_REQUIRE((u64)stage_rows(std::max(ncand, nsample)) / 64 * (ldw / 4) < (1ull << 31), CASPHY_EARG,
                   "root vote: rows x features beyond 2^44 bits");
    CASPHY_REQUIRE(std::max(ncand, nsample) <= 2000000u, CASPHY_EARG, "root vote: at most 2,000,000 nodes or samples");
    const u32 stages = ldw / 4, pa = stage_rows(ncand), pb = stage_rows(nsample);
    // whole stages: padding features (last position, code 16) set no bit
    feat.resize((size_t)stages * 128, (u32)((P - 1) << 5) | 16u);
    DevBuf dstate((size_t)nnode * nbin * 4), dpar((size_t)nnode * 4), dcand((size_t)ncand * 4),
        dfeat(feat.size() * 4), esT((size_t)pa * nbin * 4), smT((size_t)pb * nbin * 4), sa((size_t)pa * ldw * 4),
        sb((size_t)pb * ldw * 4), dkeys((size_t)nsample * 8);
    h2d(dstate.p, state, (size_t)nnode * nbin * 4, s);
    h2d(dpar.p, plan.parent.data(), (size_t)nnode * 4, s);
    h2d(dcand.p, plan.candidates.data(), (size_t)ncand * 4, s);
    h2d(dfeat.p, feat.data(), feat.size() * 4, s);
    // word-major rows, padded with zero: edge states [nbin][pa], samples [nbin][pb]
    hipLaunchKernelGGL(k_edge_cols, dim3(ceil_div(nbin, 32u), pa / 32), dim3(256), 0, s, dstate.as<u32>(), nbin,
                       dpar.as<u32>(), dcand.as<u32>(), ncand, pa, esT.as<u32>());
    launch_transpose_wordmajor(dsamp.as<u32>(), nsample, nbin, smT.as<u32>(), pb, s);
    hipLaunchKernelGGL(k_vote_stage<false>, dim3(pack_grid((u64)pa / 64 * stages)), dim3(256), 0, s, esT.as<u32>(),
                       pa, dfeat.as<u32>(), stages, sa.as<u32>());
    hipLaunchKernelGGL(k_vote_stage<true>, dim3(pack_grid((u64)pb / 64 * stages)), dim3(256), 0, s, smT.as<u32>(),
                       pb, dfeat.as<u32>(), stages, sb.as<u32>());
    CASPHY_HIP(hipGetLastError());
    launch_fill_u64(dkeys.as<u64>(), nsample, 0ull, s);
    MfmaArgs q;
    q.PA = sa.as<u32>(); q.PB = sb.as<u32>();
    q.padA = stage_rows(ncand); q.padB = stage_rows(nsample);
    q.stages = ldw / 4; q.nA = ncand; q.nB = nsample;
    q.triangle = 0;
    q.keys = dkeys.as<u64>();
    launch_mfma_tiles(E_ARGMAX, q, s);
    d2h(keys, dkeys.p, (size_t)nsample * 8, s);
}

}  // namespace casphy
