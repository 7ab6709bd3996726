// groups.hip -- autoSplit's radius search for many groups of one resident cell
// matrix at once.  R splits a layer group by group: run.split
// (R/CasPhy_functions.R:2896-2937) calls autoSplit (:1964-2082) on
// mat[inIdx, ] for each group, and each call runs tryRadius at rising radii,
// radiusNeighbor at the first radius whose mean count passes
// log10(mean) <= nNeighborBound - log10(ntip), then tipMerge.  Groups are of
// very uneven size, so one launch per group leaves most of the device idle on
// the small ones.  Here, on the matrix cores:
//   * the groups' rows are gathered into one stage-major operand (each group
//     padded to 128 rows) and their query rows into another;
//   * one launch of score tiles covers every group's (query block, row block)
//     pairs, group g's scores in its own R_g x n_g matrix; one counting pass
//     per 16 radii covers every query (rows of different lengths);
//   * one workgroup per group applies the stopping rule against the group's own
//     limit;
//   * one launch of radius-mask tiles covers every group's upper triangle,
//     each group at its own chosen radius (a group with none scores nothing);
//   * one row-popcount launch, one scan and one compaction cover all groups'
//     rows; ids are local to their group.
// The host synchronises once to size the lists.  Cells on the VALU tiles
// (weighted bits, IUPAC DNA, the 2-bit DNA form) run the groups one by one
// through the single-matrix path; the results are laid out the same way.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "cells.hpp"
#include "tipmerge.hpp"

namespace casphy {
namespace {

void require(bool c, const std::string &msg) { CASPHY_REQUIRE(c, CASPHY_EARG, msg); }

// out[i][0 .. ldw) = mat[rows[i]][0 .. ldw)
__global__ void k_gather_rows(const u32 *mat, u32 ldw, const u32 *rows, u32 n, u32 *out)
{
    const u64 total = (u64)n * ldw;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 i = (u32)(x / ldw), w = (u32)(x - (u64)i * ldw);
        out[x] = mat[(size_t)rows[i] * ldw + w];
    }
}

// out[i] = in[i] - sub for i <= n
__global__ void k_rebase(const u64 *in, u32 n, u64 sub, u64 *out)
{
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) out[i] = in[i] - sub;
}

// Appends a value of type T to a byte image at an offset aligned for it.
template <class T> size_t place(std::vector<unsigned char> &img, const std::vector<T> &v)
{
    const size_t off = (img.size() + 15) / 16 * 16;
    img.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(img.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

// Upper triangle of a tg x tg tile grid in 8 x 8 super-tiles (the order the
// XCD-contiguous launch slots share operand rows in), offset by t0.
void triangle_tiles(u32 tg, u32 t0, u32 g, std::vector<uint4> &out)
{
    const u32 S = 8, ns = ceil_div(tg, S);
    for (u32 SJ = 0; SJ < ns; ++SJ)
        for (u32 SI = 0; SI <= SJ; ++SI)
            for (u32 l = 0; l < S * S; ++l) {
                const u32 I = SI * S + l / S, J = SJ * S + l % S;
                if (I <= J && J < tg) out.push_back(make_uint4(t0 + I, t0 + J, g, 0));
            }
}

// The groups of one search, host side.
struct Plan {
    u32 G = 0, Nt = 0, Rt = 0;
    std::vector<u32> n, R, base, qbase, cpos, qpos, ms, ld;
    std::vector<u64> moff, soff;
    std::vector<u64> smax;
    std::vector<int32_t> any;
    u64 mask_words = 0, scores = 0;
    u32 Tc = 0, Tq = 0;
};

Plan make_plan(const casphy_cells &h, const u32 *goff, const u32 *rows, u32 G, const u32 *qoff, const u32 *q,
               double nb)
{
    Plan p;
    p.G = G;
    require(goff[0] == 0 && qoff[0] == 0, "casphy_cells_split_groups: goff[0] and qoff[0] must be 0");
    for (u32 g = 0; g < G; ++g) {
        require(goff[g + 1] > goff[g], "casphy_cells_split_groups: group " + std::to_string(g) + " is empty");
        require(qoff[g + 1] > qoff[g], "casphy_cells_split_groups: group " + std::to_string(g) + " has no query");
        const u32 ng = goff[g + 1] - goff[g], Rg = qoff[g + 1] - qoff[g];
        p.n.push_back(ng);
        p.R.push_back(Rg);
        p.base.push_back(goff[g]);
        p.qbase.push_back(qoff[g]);
        p.cpos.push_back(p.Tc);
        p.qpos.push_back(p.Tq);
        p.Tc += round_up(ng, TILE);
        p.Tq += round_up(Rg, TILE);
        p.ms.push_back(ceil_div(ng, TILE) * 4);
        p.moff.push_back(p.mask_words);
        p.mask_words += (u64)ng * p.ms.back();
        p.ld.push_back(round_up(ng, 4));  // 16-byte aligned score rows
        p.soff.push_back(p.scores);
        p.scores += (u64)Rg * p.ld.back();
        u64 smax;
        int any;
        autosplit_limit(Rg, ng, nb - std::log10((double)ng), smax, any);
        p.smax.push_back(smax);
        p.any.push_back(any);
    }
    p.Nt = goff[G];
    p.Rt = qoff[G];
    for (u32 i = 0; i < p.Nt; ++i)
        require(rows[i] < h.n, "casphy_cells_split_groups: row " + std::to_string(rows[i]) + " out of range");
    for (u32 g = 0; g < G; ++g)
        for (u32 k = 0; k < p.R[g]; ++k)
            require(q[p.qbase[g] + k] < p.n[g], "casphy_cells_split_groups: query " +
                                                    std::to_string(q[p.qbase[g] + k]) + " of group " +
                                                    std::to_string(g) + " out of range");
    return p;
}

void finish(casphy_cells &h, const Plan &p, const std::vector<u64> &off, const int32_t *choice, int32_t *radius,
            uint64_t *nidx)
{
    GroupOut &o = h.groups;
    o.ngroups = p.G;
    o.base = p.base;
    o.n = p.n;
    o.ld = p.ms;
    o.moff = p.moff;
    o.start.assign(p.G, 0);
    o.total.assign(p.G, 0);
    o.radius.assign(p.G, -1);
    for (u32 g = 0; g < p.G; ++g) {
        o.start[g] = off[p.base[g]];
        o.total[g] = off[p.base[g] + p.n[g]] - off[p.base[g]];
        o.radius[g] = choice[2 * g + 1] < 0 ? -1 : choice[2 * g];
        radius[g] = o.radius[g];
        nidx[g] = o.total[g];
    }
}

// Matrix-core path: every phase one launch over all groups.
void search_mfma(casphy_cells &h, const Plan &p, const u32 *rows, const int32_t *radii, u32 nradii, const u32 *q,
                 int32_t *radius, uint64_t *nidx, hipStream_t s)
{
    const u32 G = p.G, ldw = h.ldw, stages = ldw / 4;
    // host images of every index table, uploaded in one copy
    std::vector<u32> crow(p.Tc, 0xFFFFFFFFu), qrow(p.Tq, 0xFFFFFFFFu);
    std::vector<uint4> glS, glM;
    std::vector<GroupDesc> dS(G), dM(G);
    std::vector<u64> qsoff(p.Rt), rmoff(p.Nt);
    std::vector<u32> qlen(p.Rt), qoffs(G + 1), rwords(p.Nt), rgroup(p.Nt);
    std::vector<int32_t> rad(MAX_ROUND_RADII, 0x7FFFFFFF);
    for (u32 k = 0; k < nradii; ++k) rad[k] = radii[k];
    for (u32 g = 0; g < G; ++g) {
        const u32 ng = p.n[g], Rg = p.R[g];
        for (u32 i = 0; i < ng; ++i) {
            crow[p.cpos[g] + i] = rows[p.base[g] + i];
            rmoff[p.base[g] + i] = p.moff[g] + (u64)i * p.ms[g];
            rwords[p.base[g] + i] = p.ms[g];
            rgroup[p.base[g] + i] = g;
        }
        for (u32 k = 0; k < Rg; ++k) {
            qrow[p.qpos[g] + k] = rows[p.base[g] + q[p.qbase[g] + k]];
            qsoff[p.qbase[g] + k] = p.soff[g] + (u64)k * p.ld[g];
            qlen[p.qbase[g] + k] = ng;
        }
        qoffs[g] = p.qbase[g];
        const u32 tq = ceil_div(Rg, TILE), tc = ceil_div(ng, TILE);
        for (u32 I = 0; I < tq; ++I)
            for (u32 J = 0; J < tc; ++J) glS.push_back(make_uint4(p.qpos[g] / TILE + I, p.cpos[g] / TILE + J, g, 0));
        triangle_tiles(tc, p.cpos[g] / TILE, g, glM);
        dS[g].a0 = p.qpos[g]; dS[g].b0 = p.cpos[g]; dS[g].nA = Rg; dS[g].nB = ng;
        dS[g].off = p.soff[g]; dS[g].ld = p.ld[g];
        dM[g].a0 = dM[g].b0 = p.cpos[g]; dM[g].nA = dM[g].nB = ng;
        dM[g].off = p.moff[g]; dM[g].ld = p.ms[g];
    }
    qoffs[G] = p.Rt;
    std::vector<unsigned char> img;
    const size_t o_crow = place(img, crow), o_qrow = place(img, qrow), o_glS = place(img, glS),
                 o_glM = place(img, glM), o_dS = place(img, dS), o_dM = place(img, dM), o_qsoff = place(img, qsoff),
                 o_qlen = place(img, qlen), o_qoffs = place(img, qoffs), o_smax = place(img, p.smax),
                 o_any = place(img, p.any), o_rad = place(img, rad), o_rmoff = place(img, rmoff),
                 o_rwords = place(img, rwords), o_rgroup = place(img, rgroup);
    DevBuf up(img.size());
    h2d(up.p, img.data(), img.size(), s);
    auto at = [&](size_t off) { return static_cast<unsigned char *>(up.p) + off; };

    const u32 padC = stage_rows(p.Tc), padQ = stage_rows(p.Tq);
    DevBuf PC((size_t)padC * ldw * 4), EC(fp4_bytes(padC, stages)), PQ((size_t)padQ * ldw * 4),
        EQ(fp4_bytes(padQ, stages));
    DevBuf S(std::max<u64>(p.scores, 1) * 4), counts((size_t)nradii * p.Rt * 4), choice((size_t)G * 8);
    GroupOut &o = h.groups;
    o.mask.reserve(std::max<u64>(p.mask_words, 1) * 4);
    o.counts.reserve((size_t)p.Nt * 4);
    o.offsets.reserve(((size_t)p.Nt + 1) * 8);

    launch_to_stage_major(h.mat.as<u32>(), p.Tc, ldw, ldw, PC.as<u32>(), s, reinterpret_cast<const u32 *>(at(o_crow)),
                          h.n);
    launch_expand_fp4(PC.as<u32>(), padC, stages, EC.as<u32>(), s);
    launch_to_stage_major(h.mat.as<u32>(), p.Tq, ldw, ldw, PQ.as<u32>(), s, reinterpret_cast<const u32 *>(at(o_qrow)),
                          h.n);
    launch_expand_fp4(PQ.as<u32>(), padQ, stages, EQ.as<u32>(), s);

    MfmaArgs m;
    m.PA = PQ.as<u32>(); m.PB = PC.as<u32>(); m.EA = EQ.as<u32>(); m.EB = EC.as<u32>();
    m.padA = padQ; m.padB = padC; m.stages = stages; m.triangle = 0;
    m.S = S.as<int32_t>();
    m.glist = reinterpret_cast<const uint4 *>(at(o_glS));
    m.gdesc = reinterpret_cast<const GroupDesc *>(at(o_dS));
    CASPHY_HIP(hipEventRecord(h.ev[2], s));
    launch_mfma_groups(E_SCORES, m, (u32)glS.size(), s);
    const int32_t *drad = reinterpret_cast<const int32_t *>(at(o_rad));
    launch_count_radii_rows(S.as<int32_t>(), p.Rt, reinterpret_cast<const u64 *>(at(o_qsoff)),
                            reinterpret_cast<const u32 *>(at(o_qlen)), drad, nradii, counts.as<u32>(), s);
    launch_autosplit_choose_groups(counts.as<u32>(), p.Rt, reinterpret_cast<const u32 *>(at(o_qoffs)), G, drad, nradii,
                                   reinterpret_cast<const u64 *>(at(o_smax)),
                                   reinterpret_cast<const int32_t *>(at(o_any)), choice.as<int32_t>(), s);

    MfmaArgs mm;
    mm.PA = mm.PB = PC.as<u32>(); mm.EA = mm.EB = EC.as<u32>();
    mm.padA = mm.padB = padC; mm.stages = stages; mm.triangle = 1;
    mm.mask = o.mask.as<u32>();
    mm.glist = reinterpret_cast<const uint4 *>(at(o_glM));
    mm.gdesc = reinterpret_cast<const GroupDesc *>(at(o_dM));
    mm.gthr = choice.as<int32_t>();
    CASPHY_HIP(hipEventRecord(h.ev[0], s));
    launch_mfma_groups(E_MASK, mm, (u32)glM.size(), s);
    CASPHY_HIP(hipEventRecord(h.ev[1], s));
    RowMap map;
    map.moff = reinterpret_cast<const u64 *>(at(o_rmoff));
    map.words = reinterpret_cast<const u32 *>(at(o_rwords));
    map.group = reinterpret_cast<const u32 *>(at(o_rgroup));
    map.gthr = choice.as<int32_t>();
    launch_row_popc_map(o.mask.as<u32>(), p.Nt, map, o.counts.as<u32>(), s);
    launch_scan_counts(o.counts.as<u32>(), p.Nt, o.offsets.as<u64>(), s);
    std::vector<u64> off((size_t)p.Nt + 1);
    std::vector<int32_t> ch((size_t)2 * G);
    d2h(off.data(), o.offsets.p, off.size() * 8, s);
    d2h(ch.data(), choice.p, ch.size() * 4, s);
    const u64 total = off[p.Nt];
    o.idx.reserve(std::max<u64>(total, 1) * 4);
    launch_compact_map(o.mask.as<u32>(), p.Nt, map, o.offsets.as<u64>(), o.idx.as<u32>(), s);
    CASPHY_HIP(hipEventRecord(h.ev[3], s));
    CASPHY_HIP(hipStreamSynchronize(s));
    h.timed[0] = h.timed[1] = true;
    finish(h, p, off, ch.data(), radius, nidx);
}

// VALU path: the single-matrix search on each group's gathered rows in turn.
void search_valu(casphy_cells &h, const Plan &p, const u32 *rows, const int32_t *radii, u32 nradii, const u32 *q,
                 int32_t *radius, uint64_t *nidx, hipStream_t s)
{
    const u32 G = p.G, ldw = h.ldw;
    GroupOut &o = h.groups;
    o.mask.reserve(std::max<u64>(p.mask_words, 1) * 4);
    o.counts.reserve((size_t)p.Nt * 4);
    o.offsets.reserve(((size_t)p.Nt + 1) * 8);
    std::vector<int32_t> ch((size_t)2 * G, 0);
    std::vector<RadiusOut> res(G);
    u64 total = 0;
    CASPHY_HIP(hipEventRecord(h.ev[2], s));
    for (u32 g = 0; g < G; ++g) {
        const u32 ng = p.n[g], Rg = p.R[g];
        DevBuf dr((size_t)ng * 4), sub((size_t)ng * ldw * 4);
        h2d(dr.p, rows + p.base[g], (size_t)ng * 4, s);
        const u64 words = (u64)ng * ldw;
        hipLaunchKernelGGL(k_gather_rows, dim3((u32)std::min<u64>(ceil_div(words, (u64)256), 4096)), dim3(256), 0, s,
                           h.mat.as<u32>(), ldw, dr.as<u32>(), ng, sub.as<u32>());
        CASPHY_HIP(hipGetLastError());
        std::vector<u32> cnt((size_t)nradii * Rg);
        try_radii(h.metric, sub.as<u32>(), ng, ldw, h.planes, radii, nradii, q + p.qbase[g], Rg, cnt.data(), s);
        ch[2 * g] = 0x7FFFFFFF;
        ch[2 * g + 1] = -1;
        if (p.any[g])
            for (u32 k = 0; k < nradii; ++k) {
                u64 sum = 0;
                for (u32 j = 0; j < Rg; ++j) sum += cnt[(size_t)k * Rg + j];
                if (sum <= p.smax[g]) {
                    ch[2 * g] = radii[k];
                    ch[2 * g + 1] = (int32_t)k;
                    break;
                }
            }
        if (ch[2 * g + 1] < 0) {
            launch_fill_u32(o.counts.as<u32>() + p.base[g], ng, 0u, s);
            continue;
        }
        RadiusOut &r = res[g];
        radius_compute(h.metric, sub.as<u32>(), ng, ldw, h.planes, ch[2 * g], r, s);
        CASPHY_HIP(hipMemcpyAsync(o.counts.as<u32>() + p.base[g], r.counts.p, (size_t)ng * 4, hipMemcpyDeviceToDevice,
                                  s));
        CASPHY_HIP(hipMemcpyAsync(o.mask.as<u32>() + p.moff[g], r.mask.p, (size_t)ng * p.ms[g] * 4,
                                  hipMemcpyDeviceToDevice, s));
        total += r.total;
    }
    launch_scan_counts(o.counts.as<u32>(), p.Nt, o.offsets.as<u64>(), s);
    o.idx.reserve(std::max<u64>(total, 1) * 4);
    std::vector<u64> off((size_t)p.Nt + 1);
    d2h(off.data(), o.offsets.p, off.size() * 8, s);
    for (u32 g = 0; g < G; ++g)
        if (res[g].total)
            CASPHY_HIP(hipMemcpyAsync(o.idx.as<u32>() + off[p.base[g]], res[g].idx.p, res[g].total * 4,
                                      hipMemcpyDeviceToDevice, s));
    CASPHY_HIP(hipEventRecord(h.ev[3], s));
    CASPHY_HIP(hipStreamSynchronize(s));
    h.timed[0] = false;
    h.timed[1] = true;
    finish(h, p, off, ch.data(), radius, nidx);
}

}  // namespace
}  // namespace casphy

using namespace casphy;

extern "C" {

int casphy_cells_split_groups(casphy_cells *h, const uint32_t *goff, const uint32_t *rows, uint32_t ngroups,
                              const int32_t *radii, uint32_t nradii, const uint32_t *qoff, const uint32_t *q,
                              double nneighbor_bound, int32_t *radius, uint64_t *nidx)
{
    return guarded([&] {
        require(h && goff && rows && radii && qoff && q && radius && nidx && ngroups > 0,
                "casphy_cells_split_groups: bad arguments");
        require(nradii >= 1 && nradii <= MAX_ROUND_RADII, "casphy_cells_split_groups: 1..64 radii");
        for (u32 k = 1; k < nradii; ++k) require(radii[k - 1] <= radii[k], "radii must be ascending");
        const Plan p = make_plan(*h, goff, rows, ngroups, qoff, q, nneighbor_bound);
        hipStream_t s = thread_stream();
        h->groups.ngroups = 0;  // no valid result until this call completes
        if (h->mfma()) search_mfma(*h, p, rows, radii, nradii, q, radius, nidx, s);
        else search_valu(*h, p, rows, radii, nradii, q, radius, nidx, s);
    });
}

int casphy_cells_group_fetch(casphy_cells *h, uint32_t g, uint32_t *nneighbor, uint32_t *idx)
{
    return guarded([&] {
        require(h && nneighbor && idx, "casphy_cells_group_fetch: bad arguments");
        const GroupOut &o = h->groups;
        require(g < o.ngroups, "casphy_cells_group_fetch: no such group in the last group search");
        hipStream_t s = thread_stream();
        d2h(nneighbor, o.counts.as<u32>() + o.base[g], (size_t)o.n[g] * 4, s);
        if (o.total[g]) d2h(idx, o.idx.as<u32>() + o.start[g], o.total[g] * 4, s);
    });
}

int casphy_cells_group_tip_merge(casphy_cells *h, uint32_t g, double p, int32_t *member)
{
    return guarded([&] {
        require(h && member, "casphy_cells_group_tip_merge: bad arguments");
        const GroupOut &o = h->groups;
        require(g < o.ngroups, "casphy_cells_group_tip_merge: no such group in the last group search");
        hipStream_t s = thread_stream();
        const u32 n = o.n[g];
        std::vector<u32> nnb(n);
        d2h(nnb.data(), o.counts.as<u32>() + o.base[g], (size_t)n * 4, s);
        for (u32 i = 0; i < n; ++i)
            CASPHY_REQUIRE(nnb[i] <= TIP_MERGE_MAX_LIST, CASPHY_EARG,
                           "casphy_cells_group_tip_merge: tip " + std::to_string(i) +
                               " has more than 1,048,576 neighbours");
        DevBuf off(((size_t)n + 1) * 8);
        hipLaunchKernelGGL(k_rebase, dim3(std::min<u32>(ceil_div(n + 1, 256u), 1024u)), dim3(256), 0, s,
                           o.offsets.as<u64>() + o.base[g], n, o.start[g], off.as<u64>());
        CASPHY_HIP(hipGetLastError());
        // a group with empty lists has an unwritten mask: the dense form then
        // builds its rows from the lists
        tip_merge_device(off.as<u64>(), o.idx.as<u32>() + o.start[g], n,
                         o.total[g] ? o.mask.as<u32>() + o.moff[g] : nullptr, o.ld[g], p, member, s);
    });
}

}  // extern "C"
