// cells.hpp -- a cell matrix resident in HBM (casphy_cells) and the radius /
// tryRadius helpers that run on it, shared by the resident entry points
// (capi.hip) and the group-mode search (groups.hip).
#pragma once

#include <cstring>

#include "context.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

namespace casphy {

// Radius lists of a device matrix: mask tiles -> counts -> scan -> compaction,
// all kept on the device.
struct RadiusOut {
    DevBuf mask, counts, offsets, idx, pbuf, stage;
    u64 total = 0;
    u32 n = 0;
};

// Stage-major copy of a cell-major matrix for the MFMA tiles.
void make_stage(const u32 *dmat, u32 n, u32 ldw, DevBuf &P, hipStream_t s);
// Radius lists of the n rows of dmat at `radius`, left on the device in o
// (synchronises s to size the list).
void radius_compute(int metric, const u32 *dmat, u32 n, u32 ldw, const Planes &planes, int32_t radius,
                    RadiusOut &o, hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                    const u32 *stage = nullptr, const u32 *estage = nullptr);
// tryRadius counts (host array, counts[k*R + q]) of the R query rows q of dmat.
void try_radii(int metric, const u32 *dmat, u32 n, u32 ldw, const Planes &planes, const int32_t *radii,
               u32 nradii, const u32 *q, u32 R, u32 *counts, hipStream_t s, const u32 *stage = nullptr,
               hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, const u32 *estage = nullptr);
// Largest sum of R tryRadius counts (each <= n) that passes the autoSplit rule
// log10(max(S / R, 1e-30)) <= bound, by the host's log10; any = 0 when even 0 fails.
void autosplit_limit(u32 R, u32 n, double bound, u64 &smax, int &any);
bool use_scan(int metric, const Planes &planes, u32 R, u32 ldw, u32 nradii);

// Key of a captured autoSplit round: the shapes and buffers it was captured with.
struct AutosplitKey {
    u32 R = 0, nradii = 0;
    u64 smax = 0;
    int any = -1;
    bool want = false;
    u64 cap = 0;
    bool fused = false;
    const void *idx = nullptr, *mask = nullptr, *S = nullptr, *eq = nullptr, *in = nullptr, *pin = nullptr;
    const void *tab = nullptr, *cpart = nullptr;
    bool operator==(const AutosplitKey &o) const
    {
        return R == o.R && nradii == o.nradii && smax == o.smax && any == o.any && want == o.want && cap == o.cap &&
               fused == o.fused && idx == o.idx && mask == o.mask && S == o.S && eq == o.eq && in == o.in &&
               pin == o.pin && tab == o.tab && cpart == o.cpart;
    }
};

// Results of a group-mode search (casphy_cells_split_groups): the lists of
// every group back to back, group g's rows at rows base[g] .. base[g] + n[g]
// of counts / offsets (offsets over all groups), its ids local to the group,
// its mask at word moff[g] with ld[g] words per row.
struct GroupOut {
    u32 ngroups = 0;
    std::vector<u32> base, n, ld;
    std::vector<u64> moff, start, total;  // start[g] = offsets[base[g]]
    std::vector<int32_t> radius;
    DevBuf mask, counts, offsets, idx;
};

}  // namespace casphy

struct casphy_cells {
    int kind = CASPHY_KIND_BIT;
    int metric = casphy::M_AND;
    casphy::u32 n = 0, ldw = 0;
    casphy::DevBuf mat;
    casphy::Planes planes;
    casphy::DevBuf stage;   // stage-major copy for the MFMA tiles (bit kind, unit weights)
    casphy::DevBuf estage;  // its fp4 expansion (for the fp4 tile variant)
    bool mfma() const { return metric == casphy::M_AND && planes.nplanes == 0 && casphy::mfma_enabled(); }
    // stage-major copy and its fp4 expansion, made on first use
    void staged(hipStream_t s)
    {
        using namespace casphy;
        if (!stage.p) make_stage(mat.as<u32>(), n, ldw, stage, s);
        if (!estage.p) {
            estage.alloc(fp4_bytes(stage_rows(n), ldw / 4));
            launch_expand_fp4(stage.as<u32>(), stage_rows(n), ldw / 4, estage.as<u32>(), s);
        }
    }
    casphy::RadiusOut out;
    casphy::GroupOut groups;  // the last group-mode search
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timed[2] = {false, false};  // ev[0..1] / ev[2..3] recorded by some call
    int device = 0;
    // autoSplit round: uploads (radii | query rows), query stage copy, scores,
    // counts and the choice; pinned host staging for one upload and
    // the read-back of (choice, total)
    casphy::DevBuf a_in, a_sq, a_eq, a_S, a_counts, a_misc, a_pbuf;
    casphy::DevBuf a_tab, a_cpart;  // fused counts: count table, per-tile byte counts
    // captured autoSplit round and the shapes / buffers it was captured with
    // (the round's two small-kernel phases; the mask launch between them is a
    // plain launch bracketed by the handle's events)
    hipGraphExec_t ag_exec[2] = {nullptr, nullptr};
    casphy::AutosplitKey ag_key{};
    bool ag_seen = false;
    void drop_graph()
    {
        for (auto &g : ag_exec) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }
    void *pin = nullptr;      // pinned host staging, mapped into the device's address space
    void *pin_dev = nullptr;  // the device's address of it
    size_t pin_bytes = 0;
    void pinned(size_t b)
    {
        if (b <= pin_bytes) return;
        if (pin) CASPHY_HIP(hipHostFree(pin));
        pin = pin_dev = nullptr;
        pin_bytes = 0;
        CASPHY_HIP(hipHostMalloc(&pin, b, hipHostMallocMapped));
        CASPHY_HIP(hipHostGetDevicePointer(&pin_dev, pin, 0));
        pin_bytes = b;
    }
};
